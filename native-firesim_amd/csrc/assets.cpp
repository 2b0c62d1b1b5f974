// assets.cpp - reference scene placement and frame statistics (host only).
//
// fs_scene_emitter_params reads the reference's scene contracts the way its
// host does (loadSceneEmitterParams src/main.cpp:1397-1476): keys are found by
// their first occurrence anywhere in the file text and the number after the
// following colon is taken (jsonNumberForKey :1247-1264); arrays are the
// bracketed text after the key with every number in it (jsonArrayForKey
// :1148-1169, parseJsonFloats :1214-1228). Matching that lookup, rather than
// parsing JSON properly, keeps the chosen values identical when a key appears
// more than once (emitter-mask.json has heightBandMeters in two objects).
#include <algorithm>
#include <array>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/firesim.h"

namespace {

std::string read_text(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return {};
    std::ostringstream b;
    b << in.rdbuf();
    return b.str();
}

bool number_for_key(const std::string& text, const char* key, float& out) {
    const std::string needle = std::string("\"") + key + "\"";
    const size_t at = text.find(needle);
    if (at == std::string::npos) return false;
    const size_t colon = text.find(':', at + needle.size());
    if (colon == std::string::npos) return false;
    const char* start = text.c_str() + colon + 1;
    char* end = nullptr;
    const float v = std::strtof(start, &end);
    if (end == start) return false;
    out = v;
    return true;
}

// Text from the first opening delimiter after the key to its matching close.
std::string bracketed_after_key(const std::string& text, const char* key, char open, char close) {
    const std::string needle = std::string("\"") + key + "\"";
    const size_t at = text.find(needle);
    if (at == std::string::npos) return {};
    const size_t start = text.find(open, at);
    if (start == std::string::npos) return {};
    int depth = 0;
    bool inString = false, escaped = false;
    for (size_t i = start; i < text.size(); ++i) {
        const char c = text[i];
        if (open == '{') {  // objects skip string contents
            if (escaped) { escaped = false; continue; }
            if (c == '\\') { escaped = inString; continue; }
            if (c == '"') { inString = !inString; continue; }
            if (inString) continue;
        }
        if (c == open) ++depth;
        else if (c == close && --depth == 0) return text.substr(start, i - start + 1);
    }
    return {};
}

std::vector<float> numbers_in(const std::string& text) {
    std::vector<float> v;
    const char* p = text.c_str();
    while (*p) {
        char* end = nullptr;
        const float x = std::strtof(p, &end);
        if (end != p) {
            v.push_back(x);
            p = end;
        } else {
            ++p;
        }
    }
    return v;
}

std::array<float, 3> triplet(const std::string& text, const char* key, std::array<float, 3> fallback) {
    const std::vector<float> v = numbers_in(bracketed_after_key(text, key, '[', ']'));
    return v.size() >= 3 ? std::array<float, 3>{v[0], v[1], v[2]} : fallback;
}

const char* asset_id(int scene) {
    switch (scene) {
        case 0: return "room";
        case 1: return "campfire";
        case 2: return "gas-burner-aver1";
        default: return "";
    }
}

struct Placement {
    float centerX = 0.0f, centerZ = 0.0f, heightNorm = 0.0f, heightBandNorm = 0.06f, radius = 0.48f;
    int burners = 0;
    float bx[4] = {}, by[4] = {}, bz[4] = {};
};

Placement load(int scene, const std::string& root) {
    Placement p;
    const std::string dir = root + "/" + asset_id(scene);
    const std::string contract = read_text(dir + "/scene.json");
    if (!contract.empty()) {
        const std::string emitter = bracketed_after_key(contract, "emitter", '{', '}');
        const std::string& src = emitter.empty() ? contract : emitter;
        const std::array<float, 3> c = triplet(src, "centerMeters", {0.0f, 0.02f, 0.0f});
        p.centerX = c[0];
        p.centerZ = c[2];
        p.heightNorm = (c[1] - 0.02f) / 2.03f;
        float v = 0.0f;
        if (number_for_key(src, "radiusMeters", v)) p.radius = v;
        if (number_for_key(src, "heightBandMeters", v)) p.heightBandNorm = v / 2.03f;
    } else if (scene == 1) {  // built-in placements (:1416-1424)
        p.radius = 0.32f;
        p.heightNorm = (0.045f - 0.02f) / 2.03f;
        p.heightBandNorm = 0.20f / 2.03f;
    } else if (scene == 2) {
        p.radius = 0.24f;
        p.heightNorm = (0.138f - 0.02f) / 2.03f;
        p.heightBandNorm = 0.105f / 2.03f;
    }
    if (scene == 0) return p;
    const std::string mask = read_text(dir + "/emitter-mask.json");
    if (mask.empty()) return p;
    const std::array<float, 3> shift = triplet(contract, "translationMeters", {0.0f, scene == 2 ? -0.24f : 0.0f, 0.0f});
    float v = 0.0f;
    if (number_for_key(mask, "centerXMeters", v)) p.centerX = v;
    if (number_for_key(mask, "centerZMeters", v)) p.centerZ = v;
    if (number_for_key(mask, "centerYMeters", v)) p.heightNorm = (v - 0.02f) / 2.03f;
    else if (number_for_key(mask, "worldHeightMeters", v)) p.heightNorm = (v - 0.02f) / 2.03f;
    if (number_for_key(mask, "heightBandMeters", v)) p.heightBandNorm = v / 2.03f;
    if (number_for_key(mask, "radiusMeters", v)) {
        p.radius = v;
    } else if (number_for_key(mask, "radiusFraction", v)) {
        float w = 0.0f, d = 0.0f;
        if (number_for_key(mask, "width", w) && number_for_key(mask, "depth", d)) p.radius = v * std::max(w, d);
    }
    const std::vector<float> centres = numbers_in(bracketed_after_key(mask, "burnerCentersMeters", '[', ']'));
    p.burners = static_cast<int>(std::min<size_t>(4, centres.size() / 3));
    for (int i = 0; i < p.burners; ++i) {
        p.bx[i] = centres[3 * i + 0] + shift[0];
        p.by[i] = centres[3 * i + 1] + shift[1];
        p.bz[i] = centres[3 * i + 2] + shift[2];
    }
    if (scene == 2 && p.burners > 0) p.heightNorm = (p.by[0] - 0.02f) / 2.03f;
    return p;
}

}  // namespace

extern "C" {

int fs_scene_emitter_params(int32_t scene, const char* assetsDir, fs_params* params) {
    if (!params || scene < 0 || scene > 2) return FS_ERR_INVALID_ARGUMENT;
    const Placement p = load(scene, assetsDir ? assetsDir : "assets/fire-scenes");
    params->emitter_center_x = p.centerX;
    params->emitter_center_z = p.centerZ;
    params->emitter_height_norm = p.heightNorm;
    params->emitter_height_band_norm = p.heightBandNorm;
    params->emitter_radius = p.radius;
    params->burner_center_count = p.burners;
    for (int i = 0; i < 4; ++i) {
        params->burner_center_x[i] = p.bx[i];
        params->burner_center_y[i] = p.by[i];
        params->burner_center_z[i] = p.bz[i];
    }
    return FS_OK;
}

int fs_image_stats_of(const uint32_t* px, int64_t n, fs_image_stats* out) {
    if (!out || (n > 0 && !px) || n < 0) return FS_ERR_INVALID_ARGUMENT;
    fs_image_stats s{};
    if (n == 0) {
        *out = s;
        return FS_OK;
    }
    double lumaSum = 0.0, satSum = 0.0;
    int64_t warm = 0, tan = 0, black = 0, white = 0, bright = 0;
    for (int64_t i = 0; i < n; ++i) {
        const float r = static_cast<float>((px[i] >> 16) & 0xff) / 255.0f;
        const float g = static_cast<float>((px[i] >> 8) & 0xff) / 255.0f;
        const float b = static_cast<float>(px[i] & 0xff) / 255.0f;
        const float luma = r * 0.2126f + g * 0.7152f + b * 0.0722f;
        const float hi = std::max(r, std::max(g, b)), lo = std::min(r, std::min(g, b));
        const float sat = hi > 0.0001f ? (hi - lo) / hi : 0.0f;
        const bool isWarm = r > g * 1.05f && r > b * 1.55f && luma > 0.20f;
        lumaSum += luma;
        satSum += sat;
        s.max_luma = std::max(s.max_luma, luma);
        if (luma > 0.68f) ++bright;
        if (isWarm) ++warm;
        if (r > g * 1.03f && g > b * 1.35f && sat < 0.58f && luma > 0.20f && luma < 0.68f) ++tan;
        if (luma < 0.16f && sat < 0.42f) ++black;
        if (isWarm && luma > 0.82f && sat < 0.40f) ++white;
    }
    const double cnt = static_cast<double>(n);
    s.mean_luma = static_cast<float>(lumaSum / cnt);
    s.mean_saturation = static_cast<float>(satSum / cnt);
    s.warm_fire_fraction = static_cast<float>(warm / cnt);
    s.tan_smoke_fraction = static_cast<float>(tan / cnt);
    s.black_smoke_fraction = static_cast<float>(black / cnt);
    s.white_core_fraction = static_cast<float>(white / cnt);
    s.bright_fraction = static_cast<float>(bright / cnt);
    s.bright_pixels = static_cast<int32_t>(std::min<int64_t>(bright, INT32_MAX));
    *out = s;
    return FS_OK;
}

}  // extern "C"
