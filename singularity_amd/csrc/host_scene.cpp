// singularity_amd/csrc/host_scene.cpp — host-side scene ingestion for the C ABI:
// scene-file parsing, Scene/SceneConfig defaults, the orbital camera and the
// 160-byte uniform packing.  Pure C++20, no GPU calls.

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "../../include/sing_abi.h"

static_assert(sizeof(SingUniforms) == 160, "SingUniforms must match the 160-byte uniforms.h layout");
static_assert(offsetof(SingUniforms, mass_M) == 64 && offsetof(SingUniforms, width) == 112 &&
                  offsetof(SingUniforms, exposure) == 144 && offsetof(SingUniforms, disc_turbulence) == 156,
              "SingUniforms field offsets");

namespace {

// Backend defaults of the CUDA path (cuda_backend.cu:35-36,67-68).
constexpr std::uint32_t kDefaultMaxSteps = 2500;
constexpr float kDefaultHStep = 0.4f;
constexpr float kDefaultEscapeM = 200.0f;
constexpr float kDefaultHorizonRs = 1.02f;

void trim(std::string& s) {
    auto sp = [](unsigned char c) { return std::isspace(c) != 0; };
    while (!s.empty() && sp((unsigned char)s.back())) s.pop_back();
    std::size_t i = 0;
    while (i < s.size() && sp((unsigned char)s[i])) ++i;
    s.erase(0, i);
}

bool only_space_after(const char* end) {
    for (; end && *end; ++end)
        if (!std::isspace((unsigned char)*end)) return false;
    return true;
}

bool to_float(const std::string& v, float& out) {
    if (v.empty()) return false;
    char* end = nullptr;
    const float x = std::strtof(v.c_str(), &end);
    if (end == v.c_str() || !only_space_after(end)) return false;
    out = x;
    return true;
}

bool to_int(const std::string& v, std::int32_t& out) {
    if (v.empty()) return false;
    char* end = nullptr;
    const long x = std::strtol(v.c_str(), &end, 10);
    if (end == v.c_str() || !only_space_after(end)) return false;
    out = std::int32_t(x);
    return true;
}

bool to_bool(const std::string& v, std::uint32_t& out) {
    std::string l;
    for (char c : v) l.push_back(char(std::tolower((unsigned char)c)));
    if (l == "true" || l == "on" || l == "yes" || l == "1") {
        out = 1;
        return true;
    }
    if (l == "false" || l == "off" || l == "no" || l == "0") {
        out = 0;
        return true;
    }
    return false;
}

void copy_err(char* buf, std::size_t n, const std::string& s) {
    if (!buf || n == 0) return;
    std::snprintf(buf, n, "%s", s.c_str());
}

int parse_stream(std::istream& in, SingSceneConfig* out, int* line_out, char* errbuf, std::size_t errlen) {
    std::string line;
    int no = 0;
    while (std::getline(in, line)) {
        ++no;
        const std::size_t hash = line.find('#');
        if (hash != std::string::npos) line.erase(hash);
        trim(line);
        if (line.empty()) continue;
        const std::size_t eq = line.find('=');
        if (eq == std::string::npos) {
            if (line_out) *line_out = no;
            copy_err(errbuf, errlen, line);
            return SING_ERR_PARSE;
        }
        std::string key = line.substr(0, eq), v = line.substr(eq + 1);
        trim(key);
        trim(v);
        if (key.empty() || v.empty()) {
            if (line_out) *line_out = no;
            copy_err(errbuf, errlen, line);
            return SING_ERR_PARSE;
        }
        bool ok;
        if (key == "mass") ok = to_float(v, out->mass);
        else if (key == "spin") ok = to_float(v, out->spin);
        else if (key == "disc_r_inner") ok = to_float(v, out->disc_r_inner);
        else if (key == "disc_r_outer") ok = to_float(v, out->disc_r_outer);
        else if (key == "disc_doppler_on") ok = to_bool(v, out->disc_doppler_on);
        else if (key == "disc_redshift_on") ok = to_bool(v, out->disc_redshift_on);
        else if (key == "camera_fov_deg") ok = to_float(v, out->camera_fov_deg);
        else if (key == "camera_distance") ok = to_float(v, out->camera_distance);
        else if (key == "camera_elevation") ok = to_float(v, out->camera_elevation);
        else if (key == "h_step") ok = to_float(v, out->h_step);
        else if (key == "max_steps") ok = to_int(v, out->max_steps);
        else {
            if (line_out) *line_out = no;
            copy_err(errbuf, errlen, key);
            return SING_ERR_KEY;
        }
        if (!ok) {
            if (line_out) *line_out = no;
            copy_err(errbuf, errlen, line);
            return SING_ERR_PARSE;
        }
    }
    if (line_out) *line_out = no;
    return SING_OK;
}

}  // namespace

extern "C" {

void sing_scene_defaults(SingScene* s) {
    if (!s) return;
    std::memset(s, 0, sizeof(*s));
    s->metric = SING_METRIC_KERR;
    s->mass = 1.0f;
    s->spin_a_over_M = 0.6f;
    s->disc_inner_M = 3.83f;
    s->disc_outer_M = 25.0f;
    s->doppler_on = 1;
    s->redshift_on = 1;
    s->disc_on = 1;
    s->supersample = 1;
    s->exposure = 1.8f;
    s->bloom_threshold = 0.8f;
    s->bloom_strength = 0.55f;
    s->disc_turbulence = 0.5f;
    s->disc_peak_T_K = 4500.0f;
}

void sing_scene_config_defaults(SingSceneConfig* c) {
    if (!c) return;
    c->mass = 1.0f;
    c->spin = 0.0f;
    c->disc_r_inner = -1.0f;
    c->disc_r_outer = 20.0f;
    c->disc_doppler_on = 1;
    c->disc_redshift_on = 1;
    c->camera_fov_deg = 60.0f;
    c->camera_distance = 30.0f;
    c->camera_elevation = 0.0f;
    c->h_step = 0.1f;
    c->max_steps = 5000;
}

int sing_parse_scene_config(const char* text, SingSceneConfig* out, int* line, char* errbuf, size_t errlen) {
    if (!text || !out) return SING_ERR_ARG;
    std::istringstream in{std::string(text)};
    return parse_stream(in, out, line, errbuf, errlen);
}

int sing_load_scene_config(const char* path, SingSceneConfig* out, int* line, char* errbuf, size_t errlen) {
    if (!path || !out) return SING_ERR_ARG;
    std::ifstream in(path);
    if (!in) {
        if (line) *line = 0;
        copy_err(errbuf, errlen, path);
        return SING_ERR_IO;
    }
    return parse_stream(in, out, line, errbuf, errlen);
}

void sing_scene_from_config(const SingSceneConfig* cfg, SingScene* s) {
    if (!cfg || !s) return;
    s->mass = cfg->mass;
    s->spin_a_over_M = cfg->spin;
    s->disc_inner_M = cfg->disc_r_inner;
    s->disc_outer_M = cfg->disc_r_outer;
    s->doppler_on = cfg->disc_doppler_on;
    s->redshift_on = cfg->disc_redshift_on;
    s->metric = (cfg->spin > 0.0f) ? SING_METRIC_KERR : SING_METRIC_SCHWARZSCHILD;
}

void sing_orbital_camera(float distance_M, float elevation_rad, float azimuth_rad, float fov_y_deg, SingCamera* c) {
    if (!c) return;
    const float ce = std::cos(elevation_rad), se = std::sin(elevation_rad);
    const float ca = std::cos(azimuth_rad), sa = std::sin(azimuth_rad);
    c->position[0] = distance_M * ce * ca;
    c->position[1] = distance_M * ce * sa;
    c->position[2] = distance_M * se;
    const float inv = 1.0f / distance_M;
    const float fwd[3] = {-c->position[0] * inv, -c->position[1] * inv, -c->position[2] * inv};
    float rt[3] = {fwd[1], -fwd[0], 0.0f};
    const float rl = std::sqrt(rt[0] * rt[0] + rt[1] * rt[1] + rt[2] * rt[2]);
    const float ir = (rl > 1e-6f) ? 1.0f / rl : 1.0f;
    for (float& v : rt) v *= ir;
    const float up[3] = {fwd[1] * rt[2] - fwd[2] * rt[1], fwd[2] * rt[0] - fwd[0] * rt[2],
                         fwd[0] * rt[1] - fwd[1] * rt[0]};
    for (int i = 0; i < 3; ++i) {
        c->basis[i] = rt[i];
        c->basis[3 + i] = up[i];
        c->basis[6 + i] = -fwd[i];
    }
    c->fov_y_radians = fov_y_deg * (3.14159265358979323846f / 180.0f);
}

void sing_pack_uniforms(const SingScene* s, const SingCamera* cam, uint32_t w, uint32_t h, float time_sec,
                        SingUniforms* o) {
    if (!s || !cam || !o) return;
    std::memset(o, 0, sizeof(*o));
    o->cam_pos = {cam->position[0], cam->position[1], cam->position[2], 0.0f};
    o->cam_right = {cam->basis[0], cam->basis[1], cam->basis[2], 0.0f};
    o->cam_up = {cam->basis[3], cam->basis[4], cam->basis[5], 0.0f};
    o->cam_fwd = {-cam->basis[6], -cam->basis[7], -cam->basis[8], 0.0f};
    o->tan_half_fov = std::tan(0.5f * cam->fov_y_radians);
    o->aspect = float(w) / float(h);
    o->mass_M = s->mass;
    o->spin_a = s->spin_a_over_M * s->mass;
    o->rs = 2.0f * s->mass;
    o->disc_r_inner = s->disc_inner_M * s->mass;
    o->disc_r_outer = s->disc_outer_M * s->mass;
    o->disc_peak_T = s->disc_peak_T_K;
    o->h_step = s->h_step > 0.0f ? s->h_step : kDefaultHStep;
    o->escape_r = (s->escape_r_M > 0.0f ? s->escape_r_M : kDefaultEscapeM) * s->mass;
    o->horizon_cut = (s->horizon_cut_rs > 0.0f ? s->horizon_cut_rs : kDefaultHorizonRs) * o->rs;
    o->time_sec = time_sec;
    o->width = w;
    o->height = h;
    o->metric_type = s->metric == SING_METRIC_KERR ? SING_METRIC_KERR : SING_METRIC_SCHWARZSCHILD;
    std::uint32_t flags = SING_FLAG_STARFIELD_ON;
    if (s->disc_on) flags |= SING_FLAG_DISC_ON;
    if (s->doppler_on) flags |= SING_FLAG_DOPPLER_ON;
    if (s->redshift_on) flags |= SING_FLAG_REDSHIFT_ON;
    o->flags = flags;
    o->max_steps = s->max_steps != 0 ? s->max_steps : kDefaultMaxSteps;
    o->frame_index = 0;
    o->supersample = s->supersample == 0 ? 1u : s->supersample;
    o->pad_b = 0;
    o->exposure = s->exposure;
    o->bloom_threshold = s->bloom_threshold;
    o->bloom_strength = s->bloom_strength;
    o->disc_turbulence = s->disc_turbulence;
}

double sing_kerr_radius(int which, double M, double a) {
    const double d = M * M - a * a;
    const double sq = std::sqrt(d > 0.0 ? d : 0.0);
    switch (which) {
        case 0: return M + sq;
        case 1: return M - sq;
        case 2:
        case 3: {
            const double x = a / M, x2 = x * x;
            const double z1 = 1.0 + std::cbrt(1.0 - x2) * (std::cbrt(1.0 + x) + std::cbrt(1.0 - x));
            const double z2 = std::sqrt(3.0 * x2 + z1 * z1);
            const double rad = (3.0 - z1) * (3.0 + z1 + 2.0 * z2);
            const double root = std::sqrt(rad > 0.0 ? rad : 0.0);
            return M * (3.0 + z2 + (which == 2 ? -root : root));
        }
        case 4: return 2.0 * M * (1.0 + std::cos((2.0 / 3.0) * std::acos(-a / M)));
        case 5: return 2.0 * M * (1.0 + std::cos((2.0 / 3.0) * std::acos(a / M)));
    }
    return 0.0;
}

}  // extern "C"
