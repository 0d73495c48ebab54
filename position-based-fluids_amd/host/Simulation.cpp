// Simulation.cpp -- C++ drop-in for the reference's Simulation class over the
// pbf C ABI (see Simulation.hpp).
#include "Simulation.hpp"

#include <cmath>
#include <cstring>

namespace pbf_host {

pbf_params Params{};

void LoadParameters(const std::string& text) {
    const int rc = pbf_load_parameters(text.c_str(), &Params);
    if (rc < 0) throw std::runtime_error("LoadParameters: invalid input");
}

Simulation::Simulation() { pbf_default_options(&mOptions); }

Simulation::Simulation(const pbf_options& options) : mOptions(options) {}

Simulation::~Simulation() {
    if (mCtx) pbf_destroy(mCtx);
}

void Simulation::Check(int rc, const char* what) const {
    if (rc != PBF_OK) throw std::runtime_error(std::string(what) + ": " + pbf_last_error(mCtx));
}

void Simulation::InitBuffers() {
    if (mCtx) {
        pbf_destroy(mCtx);
        mCtx = nullptr;
    }
    pbf_params p = Params;
    pbf_params_finalize(&p);
    const int rc = pbf_create(&p, &mOptions, &mCtx);
    if (rc != PBF_OK) throw std::runtime_error(std::string("InitBuffers: ") + pbf_last_error(nullptr));
    CreateParticles();
}

void Simulation::InitCells() {
    if (!mCtx) throw std::runtime_error("InitCells before InitBuffers");
    // pbf_create leaves every cell empty and every friends count at zero.
}

bool Simulation::InitKernels() { return mCtx != nullptr; }

// src/Simulation.cpp:71-101: a cube of ceil(N^(1/3)) particles per axis at
// spacing h * setupSpacing, offset (1 - axis*d)/2 in x and z and 0.3 in y,
// with the reference's index-to-axis assignment (y fastest, then x, then z).
void Simulation::CreateParticles() {
    const uint32_t n = Params.particleCount;
    std::vector<float> x(4 * (size_t)n, 0.0f);
    const int per_axis = (int)std::ceil(std::pow((double)n, 1.0 / 3.0));
    const float d = Params.h * Params.setupSpacing;
    const float off_x = (1.0f - per_axis * d) / 2.0f;
    const float off_y = 0.3f;
    const float off_z = (1.0f - per_axis * d) / 2.0f;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t ix = ((uint32_t)(i / std::pow(per_axis, 1)) % per_axis);
        const uint32_t iy = ((uint32_t)(i / std::pow(per_axis, 0)) % per_axis);
        const uint32_t iz = ((uint32_t)(i / std::pow(per_axis, 2)) % per_axis);
        x[4 * (size_t)i + 0] = off_x + ix * d;
        x[4 * (size_t)i + 1] = off_y + iy * d;
        x[4 * (size_t)i + 2] = off_z + iz * d;
        x[4 * (size_t)i + 3] = 0.0f;
    }
    Check(pbf_set_state(mCtx, x.data(), nullptr, n), "CreateParticles");
}

void Simulation::SetState(const std::vector<float>& x4, const std::vector<float>& v4) {
    if (!mCtx) throw std::runtime_error("SetState before InitBuffers");
    const uint32_t n = (uint32_t)(x4.size() / 4);
    Check(pbf_set_state(mCtx, x4.data(), v4.empty() ? nullptr : v4.data(), n), "SetState");
}

void Simulation::Step() {
    if (!mCtx) throw std::runtime_error("Step before InitBuffers");
    Check(pbf_step(mCtx, bPauseSim ? 1u : 0u), "Step");
    if (!bPauseSim) {
        const uint32_t t = mSharedPingBufferID;
        mSharedPingBufferID = mSharedPongBufferID;
        mSharedPongBufferID = t;
    }
}

void Simulation::CopyRenderBuffers(std::vector<float>* positions4, std::vector<float>* texture4) const {
    const size_t n = Params.particleCount;
    if (positions4) positions4->resize(4 * n);
    if (texture4) {
        // 2048-wide texture rows (src/Runner.cpp:60), zero padded
        const size_t rows = (n + 2047) / 2048;
        texture4->assign(4 * 2048 * rows, 0.0f);
    }
    Check(pbf_get_state(mCtx, positions4 ? positions4->data() : nullptr, nullptr,
                        texture4 ? texture4->data() : nullptr, nullptr, nullptr, nullptr),
          "CopyRenderBuffers");
}

void Simulation::GetState(std::vector<float>* x4, std::vector<float>* v4) const {
    const size_t n = Params.particleCount;
    if (x4) x4->resize(4 * n);
    if (v4) v4->resize(4 * n);
    Check(pbf_get_state(mCtx, x4 ? x4->data() : nullptr, v4 ? v4->data() : nullptr, nullptr, nullptr, nullptr,
                        nullptr),
          "GetState");
}

void Simulation::UpdateTimings() {
    int count = 0;
    Check(pbf_timings(mCtx, nullptr, 0, &count), "UpdateTimings");
    std::vector<pbf_timing> t((size_t)count);
    Check(pbf_timings(mCtx, t.data(), count, &count), "UpdateTimings");
    PerfData.Trackers.clear();
    for (const pbf_timing& q : t) PerfData.Trackers.push_back({q.name, q.ema_ms, q.last_ms});
}

}  // namespace pbf_host
