// pbf_run -- headless frame loop over the C++ Simulation drop-in, the
// Runner::run sequence (src/Runner.cpp:43-132) without the renderer: when
// the scenario file changes (DetectResourceChanges; always on the first
// frame) LoadParameters, then InitBuffers if the particle count changed or
// ResetSimOnChange is set, InitCells, InitKernels; then subSteps x Step per
// frame.  Prints one line of state statistics per reported frame and one
// line per reload.
//
//   pbf_run <scenario.par> <frames> [report_every] [ramp] [--watch] [--fps N] [--log] [--render]
//
// ramp:    LoadForceMasks installs one collision quad, a ramp over the floor
//          of the box (y = 0.4h at xMin up to 0.1 of the height at xMax),
//          solid everywhere.
// --watch: reload the scenario whenever it changes (the reference always
//          watches); without it the file is read once.
// --fps N: pace the loop to N frames per second, as a renderer would.
// --log:   print the device event log after each reported frame.
// --render: stage the render buffers (CopyRenderBuffers, both forms) after
//          each reported frame and print whether the VBO equals the state's
//          positions and the two forms agree.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <list>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "Simulation.hpp"

namespace {

std::string read_file(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

void install_ramp(pbf_host::Simulation& sim) {
    const pbf_params& P = pbf_host::Params;
    pbf_quad q{};
    q.origin[0] = P.xMin;
    q.origin[1] = P.yMin + 0.4f * P.h;
    q.origin[2] = P.zMin;
    q.e0[2] = P.zMax - P.zMin;
    q.e1[0] = P.xMax - P.xMin;
    q.e1[1] = 0.1f * (P.yMax - P.yMin);
    q.mask_h = 1;
    sim.LoadForceMasks({1u}, 1, 1, {q}, 0.0f);
}

const char* event_name(uint32_t code) {
    switch (code) {
        case PBF_LOG_FRIENDS_OVERFLOW: return "friends_overflow";
        case PBF_LOG_NAN: return "nan";
        case PBF_LOG_SORT_TIMEOUT: return "sort_timeout";
        case PBF_LOG_INTERNAL: return "internal";
        default: return "unknown";
    }
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr,
                     "usage: %s <scenario.par> <frames> [report_every] [ramp] [--watch] [--fps N] [--log] [--render]\n",
                     argv[0]);
        return 2;
    }
    const std::string path = argv[1];
    const int frames = std::atoi(argv[2]);
    int every = 10;
    bool ramp = false, watch = false, log = false, render = false;
    double fps = 0.0;
    for (int a = 3; a < argc; ++a) {
        const std::string s = argv[a];
        if (s == "ramp")
            ramp = true;
        else if (s == "--watch")
            watch = true;
        else if (s == "--log")
            log = true;
        else if (s == "--render")
            render = true;
        else if (s == "--fps" && a + 1 < argc)
            fps = std::atof(argv[++a]);
        else
            every = std::max(1, std::atoi(argv[a]));
    }
    try {
        pbf_host::Simulation sim;
        std::list<std::pair<std::string, time_t>> tracked = {{path, 0}};
        uint32_t prev_count = 0;
        bool kernels_ok = false;
        std::vector<float> x, v;
        auto next = std::chrono::steady_clock::now();
        for (int frame = 1; frame <= frames; ++frame) {
            // src/Runner.cpp:43-77 (the first call sees the file as changed)
            if ((frame == 1 || watch) && pbf_host::DetectResourceChanges(tracked)) {
                pbf_host::LoadParameters(read_file(path));
                const bool reset = frame == 1 || pbf_host::Params.particleCount != prev_count ||
                                   pbf_host::Params.resetSimOnChange;
                if (reset) {
                    prev_count = pbf_host::Params.particleCount;
                    sim.InitBuffers();
                }
                sim.InitCells();
                if (ramp) install_ramp(sim);
                kernels_ok = sim.InitKernels();
                if (frame > 1) std::printf("reload frame %d reset %d ok %d\n", frame, reset ? 1 : 0, kernels_ok ? 1 : 0);
                std::fflush(stdout);
            }
            if (!kernels_ok) continue;
            for (uint32_t s = 0; s < pbf_host::Params.subSteps; ++s) sim.Step();
            if (frame % every == 0 || frame == frames) {
                sim.GetState(&x, &v);
                double ke = 0.0, ymin = 1e30, ymax = -1e30;
                bool finite = true;
                const size_t n = x.size() / 4;
                for (size_t i = 0; i < n; ++i) {
                    const double vx = v[4 * i], vy = v[4 * i + 1], vz = v[4 * i + 2];
                    ke += 0.5 * (vx * vx + vy * vy + vz * vz);
                    ymin = std::fmin(ymin, x[4 * i + 1]);
                    ymax = std::fmax(ymax, x[4 * i + 1]);
                    finite = finite && std::isfinite(x[4 * i]) && std::isfinite(x[4 * i + 1]) && std::isfinite(x[4 * i + 2]);
                }
                std::printf("frame %d n %zu ke %.6e ymin %.5f ymax %.5f finite %d\n", frame, n, ke, ymin, ymax,
                            finite ? 1 : 0);
                if (render) {
                    std::vector<float> vbo, tex;
                    sim.CopyRenderBuffers(&vbo, &tex);
                    std::vector<float> vbo2(vbo.size(), -1.0f), tex2(tex.size(), 0.0f);
                    sim.CopyRenderBuffers(vbo2.data(), tex2.data());
                    const bool same_vbo = vbo.size() == x.size() &&
                                          std::memcmp(vbo.data(), x.data(), 4 * x.size()) == 0;
                    const bool same_forms = vbo == vbo2 && tex == tex2;
                    std::printf("render frame %d vbo_is_x %d forms_agree %d rows %zu\n", frame, same_vbo ? 1 : 0,
                                same_forms ? 1 : 0, tex.size() / (4 * 2048));
                }
                if (log) {
                    std::vector<pbf_log_msg> msgs;
                    uint64_t lost = 0;
                    sim.ReadLog(&msgs, &lost);
                    for (const pbf_log_msg& m : msgs)
                        std::printf("event %s %u %u %u\n", event_name(m.code), m.a, m.b, m.c);
                    if (lost) std::printf("event lost %llu\n", (unsigned long long)lost);
                }
                std::fflush(stdout);
            }
            if (fps > 0.0) {
                next += std::chrono::microseconds((long long)(1e6 / fps));
                std::this_thread::sleep_until(next);
            }
        }
        sim.UpdateTimings();
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
