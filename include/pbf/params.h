/*
 * params.h -- simulation parameter block, byte-compatible with the
 * reference's `struct Parameters` (src/Parameters.hpp:5-56, copied to the
 * device bytewise at src/Simulation.cpp:220,243).  Field names, including
 * the reference's spellings, are kept so that code which fills a
 * `Parameters` fills a `pbf_params` the same way.
 */
#ifndef PBF_PARAMS_H
#define PBF_PARAMS_H

#include <stddef.h>
#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pbf_params {
    /* Runner related */
    int32_t resetSimOnChange;      /*   0 */
    /* Scene related */
    uint32_t particleCount;        /*   4 */
    float xMin;                    /*   8  box walls (used by the wall planes) */
    float xMax;                    /*  12 */
    float yMin;                    /*  16 */
    float yMax;                    /*  20 */
    float zMin;                    /*  24 */
    float zMax;                    /*  28 */
    float waveGenAmp;              /*  32 */
    float waveGenFreq;             /*  36 */
    float waveGenDuty;             /*  40 */
    /* Simulation consts */
    float timeStep;                /*  44 */
    uint32_t simIterations;        /*  48 */
    uint32_t subSteps;             /*  52 */
    float h;                       /*  56 */
    float restDensity;             /*  60 */
    float epsilon;                 /*  64 */
    float garvity;                 /*  68 */
    float vorticityFactor;         /*  72 */
    float viscosityFactor;         /*  76 */
    float surfaceTenstionK;        /*  80 */
    float surfaceTenstionDist;     /*  84 */
    /* Grid and friends list */
    uint32_t friendsCircles;       /*  88  must be 5  */
    uint32_t particlesPerCircle;   /*  92  must be 50 */
    uint32_t gridBufSize;          /*  96 */
    /* Setup related */
    float setupSpacing;            /* 100 */
    /* Sorting */
    uint32_t segmentSize;          /* 104 */
    uint32_t sortIterations;       /* 108 */
    /* Rendering related */
    float particleRenderSize;      /* 112 */
    /* Computed fields */
    float h_2;                     /* 116  h*h, see pbf_params_finalize */
    /* Kernel setup related */
    bool EnableCachedBuffers;      /* 120 */
} pbf_params;

#ifdef __cplusplus
static_assert(sizeof(pbf_params) == 124, "pbf_params must match the reference Parameters layout");
static_assert(offsetof(pbf_params, timeStep) == 44, "layout");
static_assert(offsetof(pbf_params, h) == 56, "layout");
static_assert(offsetof(pbf_params, gridBufSize) == 96, "layout");
static_assert(offsetof(pbf_params, h_2) == 116, "layout");
static_assert(offsetof(pbf_params, EnableCachedBuffers) == 120, "layout");
#else
_Static_assert(sizeof(pbf_params) == 124, "pbf_params must match the reference Parameters layout");
_Static_assert(offsetof(pbf_params, h_2) == 116, "layout");
#endif

#ifdef __cplusplus
}
#endif
#endif
