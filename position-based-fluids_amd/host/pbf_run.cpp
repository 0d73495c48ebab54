// pbf_run -- headless frame loop over the C++ Simulation drop-in, the
// Runner::run sequence (src/Runner.cpp:43-132) without the renderer:
// LoadParameters -> InitBuffers -> InitCells -> InitKernels -> subSteps x
// Step per frame.  Prints one line of state statistics per reported frame.
// With "ramp", LoadForceMasks installs one collision quad: a ramp over the
// floor of the box (y = 0.4h at xMin up to 0.1 of the height at xMax),
// solid everywhere.
//
//   pbf_run <scenario.par> <frames> [report_every] [ramp]
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "Simulation.hpp"

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <scenario.par> <frames> [report_every] [ramp]\n", argv[0]);
        return 2;
    }
    std::ifstream f(argv[1]);
    if (!f) {
        std::fprintf(stderr, "cannot open %s\n", argv[1]);
        return 2;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    const int frames = std::atoi(argv[2]);
    const int every = argc > 3 ? std::atoi(argv[3]) : 10;
    try {
        pbf_host::LoadParameters(ss.str());
        pbf_host::Simulation sim;
        sim.InitBuffers();
        sim.InitCells();
        if (!sim.InitKernels()) return 1;
        if (argc > 4 && std::string(argv[4]) == "ramp") {
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
        std::vector<float> x, v;
        for (int frame = 1; frame <= frames; ++frame) {
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
            }
        }
        sim.UpdateTimings();
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
