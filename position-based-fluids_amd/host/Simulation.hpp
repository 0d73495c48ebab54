// Simulation.hpp -- C++ drop-in for the reference's Simulation class
// (src/Simulation.hpp:28-151) on top of the pbf C ABI.
//
// Same public surface for the caller (Runner::run, src/Runner.cpp:43-132):
// InitBuffers / InitCells / InitKernels / Step, the bPauseSim /
// bReadFriendsList / fWavePos flags, PerfData trackers, and the ping/pong
// position-buffer swap the renderer draws from.  OpenCL objects are gone:
// the constructor takes no cl::Context, buffers live in HBM, and the GL
// hand-off is a host staging copy (CopyRenderBuffers) until a GL context is
// available to register the buffers with HIP.
#pragma once

#include <cstdint>
#include <ctime>
#include <list>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pbf/pbf.h"

namespace pbf_host {

// Macro used for the end of cell list (src/Simulation.hpp:22)
static const int END_OF_CELL_LIST = -1;

// A global parameter object and loader (src/ParamUtils.hpp:10-13).
extern pbf_params Params;
void LoadParameters(const std::string& text);

// src/Resources.cpp:166-238: true when a tracked file's modification time
// differs from the stored one and every tracked file can be locked
// exclusively (no writer still holds it); the stored times are then
// updated.  A file tracked with time 0 reports a change on the first call,
// which is how the reference's frame loop does its initial load.
bool DetectResourceChanges(std::list<std::pair<std::string, time_t>>& fileList);

// OCLPerfMon::Trackers equivalent (src/OCLPerfMon.h).
struct PerfTracker {
    std::string eventName;
    double total_time = 0.0;  // EMA, ms
    double last_time = 0.0;   // ms
};
struct PerfMon {
    std::vector<PerfTracker> Trackers;
};

class Simulation {
public:
    // The reference takes the CL context and device; the HIP context is the
    // current device of the calling thread.  As in the reference kernels no
    // box walls are applied (only the force plates of LoadForceMasks);
    // options with walls = 1 add the six planes at x/y/zMin/Max.
    Simulation();
    explicit Simulation(const pbf_options& options);
    ~Simulation();
    Simulation(const Simulation&) = delete;
    Simulation& operator=(const Simulation&) = delete;

    // src/Simulation.cpp:206-244: allocates buffers and places the initial
    // particle block (CreateParticles, :71-101).
    void InitBuffers();
    // src/Simulation.cpp:246-255: cells to END_OF_CELL_LIST, friends to 0.
    void InitCells();
    // src/Simulation.cpp:129-204: the reference compiles its kernels with
    // the current Params baked in.  The HIP code objects are built ahead of
    // time, so this applies Params to the live context instead
    // (pbf_update_params); false when the context cannot take them (a changed
    // particleCount, gridBufSize or simIterations needs InitBuffers first).
    bool InitKernels();
    // src/Simulation.cpp:554-621
    void Step();

    // src/Simulation.cpp:257-281: the force-plate mask (the red channel of
    // Scene_fp_mask.png as width x height texels, row-major; the caller
    // decodes the image, as SOIL did there) and the quads that sample it
    // (by default the reference scene's plates, compute_delta.cl:169-182,
    // with its edge offset 1.0).  Call after InitBuffers.
    void LoadForceMasks(const std::vector<uint32_t>& texels, uint32_t width, uint32_t height,
                        const std::vector<pbf_quad>& quads = ForcePlates(), float edge_base = 1.0f);
    // The reference scene's 14 force plates (compute_delta.cl:169-182).
    static std::vector<pbf_quad> ForcePlates();

    // Replace the particle state (float4 x and v, particle-major).
    void SetState(const std::vector<float>& x4, const std::vector<float>& v4);
    // Host staging of the two render inputs: the positions VBO (w = speed)
    // and the 2048-wide RGBA32F texture contents (x*, w = density).
    void CopyRenderBuffers(std::vector<float>* positions4, std::vector<float>* texture4) const;
    // The same into caller-owned memory: positions4 holds 4 * particleCount
    // floats, texture4 the 2048-wide rows (4 * 2048 * ceil(n / 2048) floats;
    // the padding is left as it is).  Page-locked destinations (hipHostMalloc,
    // or a renderer's pinned upload buffers) copy at about 3.5x the rate of
    // pageable ones.
    void CopyRenderBuffers(float* positions4, float* texture4) const;
    void GetState(std::vector<float>* x4, std::vector<float>* v4) const;
    // Device event log since the last read (pbf_read_log), the role of
    // OCL_Logger::CycleExecute (src/OCL_Logger.cpp:80-116); *lost counts the
    // messages the ring overwrote before they were read.
    void ReadLog(std::vector<pbf_log_msg>* out, uint64_t* lost = nullptr);
    // Collect per-stage timings into PerfData (OCLPerfMon::UpdateTimings).
    void UpdateTimings();

    pbf_ctx* context() const { return mCtx; }

    // Rendering state (src/Simulation.hpp:147-150)
    bool bPauseSim = false;
    bool bReadFriendsList = false;
    bool bDumpParticlesData = false;
    float fWavePos = 0.0f;

    // The renderer draws mSharedPingBufferID; Step() swaps the pair after a
    // sort as the reference does (src/Simulation.cpp:535).
    uint32_t mSharedPingBufferID = 0;
    uint32_t mSharedPongBufferID = 0;
    uint32_t mSharedParticlesPos = 0;

    PerfMon PerfData;

private:
    void CreateParticles();
    void Check(int rc, const char* what) const;

    pbf_options mOptions{};
    pbf_ctx* mCtx = nullptr;
};

}  // namespace pbf_host
