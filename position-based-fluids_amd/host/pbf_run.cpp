// pbf_run -- headless frame loop over the C++ Simulation drop-in, the
// Runner::run sequence (src/Runner.cpp:43-132) without the renderer: when
// the scenario file changes (DetectResourceChanges; always on the first
// frame) LoadParameters, then InitBuffers if the particle count changed or
// ResetSimOnChange is set, InitCells, InitKernels; then subSteps x Step per
// frame.  Prints one line of state statistics per reported frame and one
// line per reload.
//
//   pbf_run <scenario.par> <frames> [report_every] [ramp] [--walls] [--plates MASK] [--pause F,F,...]
//           [--dump DIR] [--watch] [--fps N] [--log] [--render]
//
// ramp:    LoadForceMasks installs one collision quad, a ramp over the floor
//          of the box (y = 0.4h at xMin up to 0.1 of the height at xMax),
//          solid everywhere.
// --walls: the six box planes at x/y/zMin/Max (the reference kernels have
//          none; the drop-in Simulation leaves them off by default).
// --plates MASK: LoadForceMasks with the reference scene's 14 force plates
//          (edge offset 1.0) and MASK, a raw file of one byte per texel,
//          512 texels per row (the red channel of the reference's mask).
// --pause F,F,...: frames run with bPauseSim set.
// --dump DIR: write raw little-endian arrays to DIR: init.x (float4 per
//          particle after InitBuffers), and for every reported frame F
//          F.vbo and F.tex (CopyRenderBuffers: positions with w = |v|, the
//          2048-wide x* texture with w = density), F.ids (ping and pong
//          buffer ids, uint32) and F.keys, F.perm, F.cells, F.friends
//          (pbf_get_structure, friends in the reference layout).
// --watch: reload the scenario whenever it changes (the reference always
//          watches); without it the file is read once.
// --fps N: pace the loop to N frames per second, as a renderer would.
// --log:   print the device event log after each reported frame.
// --render: stage the render buffers (CopyRenderBuffers, both forms) after
//          each reported frame and print whether the VBO equals the state's
//          positions and the two forms agree.
#include <algorithm>
#include <set>
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

void write_raw(const std::string& path, const void* data, size_t bytes) {
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + path);
    const size_t w = bytes ? std::fwrite(data, 1, bytes, f) : 0;
    std::fclose(f);
    if (w != bytes) throw std::runtime_error("short write to " + path);
}

std::vector<uint32_t> read_mask(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::vector<char> bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    if (bytes.empty() || bytes.size() % 512) throw std::runtime_error(path + ": not a whole number of 512-texel rows");
    std::vector<uint32_t> texels(bytes.size());
    for (size_t k = 0; k < bytes.size(); ++k) texels[k] = (uint8_t)bytes[k];
    return texels;
}

void dump_frame(const std::string& dir, int frame, pbf_host::Simulation& sim) {
    const std::string base = dir + "/" + std::to_string(frame);
    std::vector<float> vbo, tex;
    sim.CopyRenderBuffers(&vbo, &tex);
    write_raw(base + ".vbo", vbo.data(), 4 * vbo.size());
    write_raw(base + ".tex", tex.data(), 4 * tex.size());
    const uint32_t ids[2] = {sim.mSharedPingBufferID, sim.mSharedPongBufferID};
    write_raw(base + ".ids", ids, sizeof ids);
    const size_t n = pbf_host::Params.particleCount;
    std::vector<uint32_t> keys(n), perm(n), cells(2 * (size_t)pbf_host::Params.gridBufSize),
        friends(n * pbf_host::Params.friendsCircles * (1 + pbf_host::Params.particlesPerCircle));
    const int rc = pbf_get_structure(sim.context(), keys.data(), perm.data(), cells.data(), friends.data());
    if (rc != PBF_OK) throw std::runtime_error(std::string("pbf_get_structure: ") + pbf_last_error(sim.context()));
    write_raw(base + ".keys", keys.data(), 4 * n);
    write_raw(base + ".perm", perm.data(), 4 * n);
    write_raw(base + ".cells", cells.data(), 4 * cells.size());
    write_raw(base + ".friends", friends.data(), 4 * friends.size());
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
                     "usage: %s <scenario.par> <frames> [report_every] [ramp] [--walls] [--plates MASK] "
                     "[--pause F,F,...] [--dump DIR] [--watch] [--fps N] [--log] [--render]\n",
                     argv[0]);
        return 2;
    }
    const std::string path = argv[1];
    const int frames = std::atoi(argv[2]);
    int every = 10;
    bool ramp = false, watch = false, log = false, render = false, walls = false;
    std::string plates, dump;
    std::set<int> paused;
    double fps = 0.0;
    for (int a = 3; a < argc; ++a) {
        const std::string s = argv[a];
        if (s == "ramp")
            ramp = true;
        else if (s == "--walls")
            walls = true;
        else if (s == "--plates" && a + 1 < argc)
            plates = argv[++a];
        else if (s == "--dump" && a + 1 < argc)
            dump = argv[++a];
        else if (s == "--pause" && a + 1 < argc) {
            std::stringstream list(argv[++a]);
            for (std::string f; std::getline(list, f, ',');) paused.insert(std::atoi(f.c_str()));
        }
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
        pbf_options opt;
        pbf_default_options(&opt);
        opt.walls = walls ? 1u : 0u;
        pbf_host::Simulation sim(opt);
        const std::vector<uint32_t> mask = plates.empty() ? std::vector<uint32_t>() : read_mask(plates);
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
                    // the renderer's two position buffers (src/Runner.cpp:58-59)
                    sim.mSharedPingBufferID = 1;
                    sim.mSharedPongBufferID = 2;
                    if (!dump.empty()) {
                        std::vector<float> x0, v0;
                        sim.GetState(&x0, &v0);
                        write_raw(dump + "/init.x", x0.data(), 4 * x0.size());
                    }
                }
                sim.InitCells();
                if (ramp) install_ramp(sim);
                if (!mask.empty()) sim.LoadForceMasks(mask, 512, (uint32_t)(mask.size() / 512));
                kernels_ok = sim.InitKernels();
                if (frame > 1) std::printf("reload frame %d reset %d ok %d\n", frame, reset ? 1 : 0, kernels_ok ? 1 : 0);
                std::fflush(stdout);
            }
            if (!kernels_ok) continue;
            sim.bPauseSim = paused.count(frame) != 0;
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
                if (!dump.empty()) dump_frame(dump, frame, sim);
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
