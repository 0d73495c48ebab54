// pbf_params.cpp -- the scenario (.par) format of the reference's loader
// (src/ParamUtils.cpp:24-100) behind the C ABI: '#' starts a comment, keys
// are case-insensitive, one "key value" pair per line, unknown keys are
// counted and skipped, and a last line without a newline is not read.
#include <algorithm>
#include <cctype>
#include <cstddef>
#include <sstream>
#include <string>

#include "../../include/pbf/pbf.h"

namespace {

// One scenario key: its lower-case name, the field it sets and how the value
// is read (stream extraction into the field's own type, so a value that does
// not parse behaves as it does in the reference).
struct ParamKey {
    const char* name;
    size_t offset;
    void (*read)(std::istream&, unsigned char*);
};

template <typename T>
void read_field(std::istream& in, unsigned char* field) {
    in >> *reinterpret_cast<T*>(field);
}

#define PBF_KEY(NAME, FIELD)                                                                       \
    ParamKey {                                                                                     \
        NAME, offsetof(pbf_params, FIELD), &read_field<decltype(pbf_params::FIELD)>                \
    }

// The keys of src/ParamUtils.cpp:41-95 (the friends-circle sizes, h^2 and the
// kernel factors are fixed or computed, not read).
const ParamKey kKeys[] = {
    PBF_KEY("resetsimonchange", resetSimOnChange), PBF_KEY("particlecount", particleCount),
    PBF_KEY("xmin", xMin), PBF_KEY("xmax", xMax), PBF_KEY("ymin", yMin), PBF_KEY("ymax", yMax),
    PBF_KEY("zmin", zMin), PBF_KEY("zmax", zMax), PBF_KEY("wavegenamp", waveGenAmp),
    PBF_KEY("wavegenfreq", waveGenFreq), PBF_KEY("wavegenduty", waveGenDuty), PBF_KEY("timestep", timeStep),
    PBF_KEY("simiterations", simIterations), PBF_KEY("substeps", subSteps), PBF_KEY("smoothlen", h),
    PBF_KEY("gridbuffersize", gridBufSize), PBF_KEY("restdensity", restDensity), PBF_KEY("epsilon", epsilon),
    PBF_KEY("garvity", garvity), PBF_KEY("vorticityfactor", vorticityFactor),
    PBF_KEY("viscosityfactor", viscosityFactor), PBF_KEY("surfacetenstionk", surfaceTenstionK),
    PBF_KEY("surfacetenstiondist", surfaceTenstionDist), PBF_KEY("setupspacing", setupSpacing),
    PBF_KEY("segmentsize", segmentSize), PBF_KEY("sortiterations", sortIterations),
    PBF_KEY("particlerendersize", particleRenderSize), PBF_KEY("enablecachedbuffers", EnableCachedBuffers),
};

#undef PBF_KEY

const ParamKey* find_key(const std::string& name) {
    for (const ParamKey& k : kKeys)
        if (name == k.name) return &k;
    return nullptr;
}

}  // namespace

extern "C" int pbf_load_parameters(const char* text, pbf_params* p) {
    if (!text || !p) return PBF_E_INVALID;
    std::istringstream doc{std::string(text)};
    unsigned char* const base = reinterpret_cast<unsigned char*>(p);
    int unknown = 0;
    for (std::string line; std::getline(doc, line);) {
        // the format's last-line rule: a line that ends the stream without a
        // newline (the stream is no longer good()) ends the parse unread
        if (!doc.good()) break;
        line.erase(std::find(line.begin(), line.end(), '#'), line.end());
        for (char& ch : line) ch = (char)std::tolower((unsigned char)ch);
        std::istringstream fields(line);
        std::string name;
        if (!(fields >> name)) continue;   // blank or comment-only
        if (const ParamKey* k = find_key(name))
            k->read(fields, base + k->offset);
        else
            ++unknown;
    }
    pbf_params_finalize(p);
    return unknown;
}
