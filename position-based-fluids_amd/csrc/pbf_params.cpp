// pbf_params.cpp -- the reference's .par loader (src/ParamUtils.cpp:24-100)
// behind the C ABI: '#' starts a comment, keys are case-insensitive, one
// "key value" pair per line, unknown keys are reported and skipped.
#include <algorithm>
#include <cctype>
#include <sstream>
#include <string>

#include "../../include/pbf/pbf.h"

extern "C" int pbf_load_parameters(const char* text, pbf_params* p) {
    if (!text || !p) return PBF_E_INVALID;
    std::istringstream ss{std::string(text)};
    std::string line, key;
    int unknown = 0;
    while (std::getline(ss, line)) {
        // ParamUtils.cpp:35-39: a stream that is no longer good() (a last
        // line without a trailing newline) ends the parse.
        if (!ss.good()) break;
        line.erase(std::find(line.begin(), line.end(), '#'), line.end());
        if (line.empty()) continue;
        std::transform(line.begin(), line.end(), line.begin(), [](unsigned char ch) { return (char)std::tolower(ch); });
        std::istringstream ls(line);
        if (!(ls >> key)) continue;
        if (key == "resetsimonchange") ls >> p->resetSimOnChange;
        else if (key == "particlecount") ls >> p->particleCount;
        else if (key == "xmin") ls >> p->xMin;
        else if (key == "xmax") ls >> p->xMax;
        else if (key == "ymin") ls >> p->yMin;
        else if (key == "ymax") ls >> p->yMax;
        else if (key == "zmin") ls >> p->zMin;
        else if (key == "zmax") ls >> p->zMax;
        else if (key == "wavegenamp") ls >> p->waveGenAmp;
        else if (key == "wavegenfreq") ls >> p->waveGenFreq;
        else if (key == "wavegenduty") ls >> p->waveGenDuty;
        else if (key == "timestep") ls >> p->timeStep;
        else if (key == "simiterations") ls >> p->simIterations;
        else if (key == "substeps") ls >> p->subSteps;
        else if (key == "smoothlen") ls >> p->h;
        else if (key == "gridbuffersize") ls >> p->gridBufSize;
        else if (key == "restdensity") ls >> p->restDensity;
        else if (key == "epsilon") ls >> p->epsilon;
        else if (key == "garvity") ls >> p->garvity;
        else if (key == "vorticityfactor") ls >> p->vorticityFactor;
        else if (key == "viscosityfactor") ls >> p->viscosityFactor;
        else if (key == "surfacetenstionk") ls >> p->surfaceTenstionK;
        else if (key == "surfacetenstiondist") ls >> p->surfaceTenstionDist;
        else if (key == "setupspacing") ls >> p->setupSpacing;
        else if (key == "segmentsize") ls >> p->segmentSize;
        else if (key == "sortiterations") ls >> p->sortIterations;
        else if (key == "particlerendersize") ls >> p->particleRenderSize;
        else if (key == "enablecachedbuffers") ls >> p->EnableCachedBuffers;
        else ++unknown;
    }
    pbf_params_finalize(p);
    return unknown;
}
