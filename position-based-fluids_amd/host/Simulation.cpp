// Simulation.cpp -- C++ drop-in for the reference's Simulation class over the
// pbf C ABI (see Simulation.hpp).
#include "Simulation.hpp"

#include <fcntl.h>
#include <sys/file.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <cstring>

namespace pbf_host {

pbf_params Params{};

void LoadParameters(const std::string& text) {
    const int rc = pbf_load_parameters(text.c_str(), &Params);
    if (rc < 0) throw std::runtime_error("LoadParameters: invalid input");
}

Simulation::Simulation() {
    // the reference kernels apply no box walls, only the force plates
    // (compute_delta.cl:163-182)
    pbf_default_options(&mOptions);
    mOptions.walls = 0;
}

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

bool Simulation::InitKernels() {
    if (!mCtx) return false;
    pbf_params p = Params;
    pbf_params_finalize(&p);
    return pbf_update_params(mCtx, 0, &p) == PBF_OK;
}

void Simulation::ReadLog(std::vector<pbf_log_msg>* out, uint64_t* lost) {
    if (!mCtx) throw std::runtime_error("ReadLog before InitBuffers");
    out->clear();
    uint64_t total_lost = 0;
    for (;;) {
        pbf_log_msg buf[256];
        uint32_t count = 0;
        uint64_t l = 0;
        Check(pbf_read_log(mCtx, buf, 256, &count, &l), "ReadLog");
        total_lost += l;
        out->insert(out->end(), buf, buf + count);
        if (count < 256) break;
    }
    if (lost) *lost = total_lost;
}

bool DetectResourceChanges(std::list<std::pair<std::string, time_t>>& fileList) {
    bool changed = false;
    for (auto& f : fileList) {
        struct stat st;
        if (stat(f.first.c_str(), &st) != 0) {
            std::fprintf(stderr, "Unable to locate %s\n", f.first.c_str());
            continue;
        }
        if (f.second != st.st_mtime) {
            changed = true;
            break;
        }
    }
    if (!changed) return false;
    // a writer may still hold a file: report the change once all can be locked
    for (auto& f : fileList) {
        const int fd = open(f.first.c_str(), O_RDWR);
        const int locked = fd < 0 ? -1 : flock(fd, LOCK_EX | LOCK_NB);
        if (locked == 0) flock(fd, LOCK_UN);
        if (fd >= 0) close(fd);
        if (locked != 0) {
            std::fprintf(stderr, "File is exclusivly open [%s]\n", f.first.c_str());
            return false;
        }
    }
    for (auto& f : fileList) {
        struct stat st;
        if (stat(f.first.c_str(), &st) == 0) f.second = st.st_mtime;
    }
    return true;
}

// The initial block of src/Simulation.cpp:71-101: a cube with
// a = ceil(N^(1/3)) particles per side at spacing d = h * setupSpacing,
// filled y first, then x, then z (particle i sits at lattice point
// (floor(i/a) mod a, i mod a, floor(i/a^2) mod a)), centred on x = z = 1/2
// and lifted to y = 0.3; v = 0.
void Simulation::CreateParticles() {
    const uint32_t n = Params.particleCount;
    std::vector<float> x(4 * (size_t)n, 0.0f);
    const uint32_t a = (uint32_t)std::ceil(std::pow((double)n, 1.0 / 3.0));
    const float d = Params.h * Params.setupSpacing;
    const float side = (float)a * d;
    const float corner_xz = (1.0f - side) / 2.0f;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t col = i / a;
        const float lattice[3] = {(float)(col % a), (float)(i % a), (float)((col / a) % a)};
        const float corner[3] = {corner_xz, 0.3f, corner_xz};
        for (int k = 0; k < 3; ++k) x[4 * (size_t)i + k] = corner[k] + lattice[k] * d;
    }
    Check(pbf_set_state(mCtx, x.data(), nullptr, n), "CreateParticles");
}

void Simulation::LoadForceMasks(const std::vector<uint32_t>& texels, uint32_t width, uint32_t height,
                                const std::vector<pbf_quad>& quads, float edge_base) {
    if (!mCtx) throw std::runtime_error("LoadForceMasks before InitBuffers");
    if ((size_t)width * height != texels.size()) throw std::runtime_error("LoadForceMasks: texels != width x height");
    Check(pbf_set_collision_quads(mCtx, 0, quads.data(), (uint32_t)quads.size(), texels.data(), width, height,
                                  edge_base),
          "LoadForceMasks");
}

std::vector<pbf_quad> Simulation::ForcePlates() {
    // corner, E0, E1, mask rows (first, count) of each plate, in call order
    static const float t[14][11] = {
        {-165.8f, -106.4f, 28.32f, -15.39f, 41.82f, 0.0f, 52.28f, 19.23f, 0.0f, 0, 409},
        {-158.5f, -126.3f, -37.11f, -7.638f, 20.76f, 66.1f, 52.05f, 19.15f, 0.0f, 409, 643},
        {-21.79f, -23.62f, -34.55f, 89.42f, 39.44f, 0.0f, -27.26f, 61.81f, 0.0f, 1052, 740},
        {-39.96f, -11.55f, 29.95f, -8.877f, 49.84f, 0.0f, 103.3f, 18.4f, 0.0f, 1792, 247},
        {-180.5f, -64.4f, -37.25f, -0.0009671f, -0.0003558f, 66.38f, 22.87f, -62.17f, 0.0f, 2039, 512},
        {35.5f, 2.278f, -35.25f, 0.0f, 0.0f, 65.91f, 21.74f, 36.07f, 0.0f, 2551, 801},
        {-31.54f, 1.768f, -8.499f, 0.0f, 0.0f, 15.86f, -15.69f, -32.17f, 0.0f, 3352, 226},
        {-67.22f, -41.65f, 7.364f, 11.45f, 0.0f, 0.0f, 0.0f, 0.0f, -15.86f, 3578, 369},
        {-65.91f, -42.02f, 7.364f, 0.0f, 0.0f, -15.87f, -7.275f, 4.603f, 0.0f, 3947, 943},
        {-9.05f, -4.783f, -8.792f, -47.98f, -37.3f, 0.0f, 0.0f, 0.0f, 16.45f, 4890, 1891},
        {-53.98f, -25.37f, 6.656f, 27.86f, 34.59f, 0.0f, 17.42f, -14.03f, 0.0f, 6781, 1016},
        {-48.37f, -30.96f, -7.792f, 21.75f, -10.61f, 0.0f, 17.92f, 36.76f, 0.0f, 7797, 303},
        {-48.28f, 38.39f, -35.25f, 0.0f, 0.0f, 65.91f, 20.93f, -47.46f, 0.0f, 8100, 650},
        {36.56f, 3.007f, 30.66f, 0.0f, 0.0f, -65.91f, -64.89f, -11.56f, 0.0f, 8750, 512},
    };
    std::vector<pbf_quad> out(14);
    for (int k = 0; k < 14; ++k) {
        for (int a = 0; a < 3; ++a) {
            out[k].origin[a] = t[k][a];
            out[k].e0[a] = t[k][3 + a];
            out[k].e1[a] = t[k][6 + a];
        }
        out[k].mask_y = (int32_t)t[k][9];
        out[k].mask_h = (int32_t)t[k][10];
    }
    return out;
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

void Simulation::CopyRenderBuffers(float* positions4, float* texture4) const {
    Check(pbf_get_state(mCtx, positions4, nullptr, texture4, nullptr, nullptr, nullptr), "CopyRenderBuffers");
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
