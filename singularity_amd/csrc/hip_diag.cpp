// singularity_amd/csrc/hip_diag.cpp — hip_diag, the drop-in for the CPU
// driver's diagnostic modes (cli/main.cpp:43-104,1269-1621,1625-1834):
//
//   hip_diag --mode kerr-geometry [--mass M] [--spin A] [--output PATH]
//   hip_diag --mode photon-orbit  [--mass M] [--spin A] [--r-init R] [--orbits N] [--h-step H] [--output PATH]
//   hip_diag --mode benchmark     [--metric schw|kerr] [--spin A] [--n-rays N] [--max-steps N] [--b-range MIN MAX]
//
// Same flags, defaults, files, stdout lines and exit codes (0 ok, 2 usage or
// argument error, 1 runtime); the orbit and the ray fan are integrated on the
// GPU (libsing_hip.so), the geometry scalars are host closed forms.  The
// driver's image modes (2d-toy, kerr-2d-toy, disc-preview, cpu-render,
// kerr-cpu-render) are not provided: hip_cli renders.

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <string_view>
#include <vector>

#include "../../include/sing_abi.h"

namespace {

void print_help() {
    std::puts(
        "hip_diag — the CPU driver's diagnostic modes on the MI355X\n"
        "\n"
        "Usage:\n"
        "  hip_diag --help\n"
        "  hip_diag --mode kerr-geometry [options]\n"
        "  hip_diag --mode photon-orbit  [options]\n"
        "  hip_diag --mode benchmark     [options]\n"
        "\n"
        "Modes:\n"
        "  kerr-geometry    Closed-form Kerr radii as JSON: both horizons,\n"
        "                   ergosurface at pole and equator, prograde and\n"
        "                   retrograde photon orbits and ISCOs.\n"
        "  photon-orbit     March one circular null orbit (r_init, by default\n"
        "                   3 M, or the prograde Kerr photon orbit) on the GPU\n"
        "                   and write its t, r, theta, phi samples as CSV.\n"
        "  benchmark        Steps/s of the Schwarzschild or Kerr RK4 stepper\n"
        "                   over a fixed fan of equatorial rays, one GPU lane\n"
        "                   per ray; prints one JSON line.\n"
        "\n"
        "Options:\n"
        "  --output PATH        where the JSON / CSV goes (each mode has a default)\n"
        "  --b-range MIN MAX    (benchmark) span of impact parameters, M (default -10 10)\n"
        "  --n-rays N           (benchmark) rays in the fan (default 100)\n"
        "  --max-steps N        (benchmark) step limit of each ray (default 5000)\n"
        "  --spin VALUE         spin a/M, 0 to 1 (default 0)\n"
        "  --mass VALUE         mass M, geometrized units (default 1)\n"
        "  --r-init VALUE       (photon-orbit) starting radius, M (default 3)\n"
        "  --orbits VALUE       (photon-orbit) periods to march (default 1)\n"
        "  --h-step VALUE       (photon-orbit) step size, M (default 0.01)\n"
        "  --metric NAME        (benchmark) schw (default) or kerr\n"
        "  --scene PATH         key = value scene file; this tool reads its\n"
        "                       mass, spin, max_steps and h_step\n"
        "  --device N           HIP device (default 0)\n");
}

int fail_usage(const char* msg) {
    std::fprintf(stderr, "hip_diag: %s\n", msg);
    return 2;
}

int run_kerr_geometry(const std::string& path, float mass, float spin) {
    if (!(mass > 0.0f)) return fail_usage("--mass must be positive");
    if (spin < 0.0f || spin > mass) {
        std::fprintf(stderr, "hip_diag: --spin %.6f out of range [0, M=%.6f]\n", spin, mass);
        return 2;
    }
    SingKerrGeometry g;
    if (sing_kerr_geometry(mass, spin, &g) != SING_OK) {
        std::fprintf(stderr, "hip_diag: %s\n", sing_last_error(nullptr));
        return 2;
    }
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) {
        std::fprintf(stderr, "hip_diag: failed to open '%s'\n", path.c_str());
        return 1;
    }
    // cli/main.cpp:1444-1471: one flat object, %.9g (a float round-trips)
    std::fprintf(f,
                 "{\n  \"mass\": %.9g,\n  \"spin\": %.9g,\n  \"spin_over_mass\": %.9g,\n"
                 "  \"outer_horizon\": %.9g,\n  \"inner_horizon\": %.9g,\n"
                 "  \"ergosphere_polar\": %.9g,\n  \"ergosphere_equatorial\": %.9g,\n"
                 "  \"photon_sphere_prograde\": %.9g,\n  \"photon_sphere_retrograde\": %.9g,\n"
                 "  \"isco_prograde\": %.9g,\n  \"isco_retrograde\": %.9g\n}\n",
                 g.mass, g.spin, g.spin_over_mass, g.outer_horizon, g.inner_horizon, g.ergosphere_polar,
                 g.ergosphere_equatorial, g.photon_sphere_prograde, g.photon_sphere_retrograde, g.isco_prograde,
                 g.isco_retrograde);
    std::fclose(f);
    std::printf("wrote %s (Kerr geometry at a/M = %.6g)\n", path.c_str(), spin / mass);
    return 0;
}

SingContext* open_device(int device) {
    SingContext* ctx = sing_create(1, 1, device);
    if (!ctx) std::fprintf(stderr, "hip_diag: device initialize failed: %s\n", sing_last_error(nullptr));
    return ctx;
}

int run_photon_orbit(const std::string& path, float mass, float spin, float r_init, float orbits, float h, int dev) {
    // argument errors (cli/main.cpp:1282-1305) are reported before a device is opened
    if (!(mass > 0.0f)) return fail_usage("--mass must be positive");
    if (spin < 0.0f || spin > mass) return fail_usage("--spin must be in [0, M]");
    SingKerrGeometry g;
    sing_kerr_geometry(mass, spin, &g);
    const float r_plus = spin > 0.0f ? g.outer_horizon : 2.0f * mass;
    if (r_init <= r_plus * 1.001f) {
        std::fprintf(stderr, "hip_diag: --r-init %.4g must be > r_+ = %.4g\n", r_init, r_plus);
        return 2;
    }
    if (orbits <= 0.0f) return fail_usage("--orbits must be positive");
    if (h <= 0.0f) return fail_usage("--h-step must be positive");
    SingContext* ctx = open_device(dev);
    if (!ctx) return 1;
    std::uint32_t bound = 0;
    int rc = sing_photon_orbit(ctx, mass, spin, r_init, orbits, h, nullptr, 0, &bound, nullptr);
    if (bound == 0) {  // an argument error, not the sizing query
        std::fprintf(stderr, "hip_diag: %s\n", sing_last_error(ctx));
        sing_destroy(ctx);
        return rc == SING_ERR_ARG ? 2 : 1;
    }
    std::vector<SingOrbitRow> rows(bound);
    std::uint32_t n = 0;
    rc = sing_photon_orbit(ctx, mass, spin, r_init, orbits, h, rows.data(), bound, &n, nullptr);
    if (rc != SING_OK) {
        std::fprintf(stderr, "hip_diag: photon orbit failed: %s\n", sing_last_error(ctx));
        sing_destroy(ctx);
        return rc == SING_ERR_ARG ? 2 : 1;
    }
    sing_destroy(ctx);
    FILE* fh = std::fopen(path.c_str(), "wb");
    if (!fh) {
        std::fprintf(stderr, "hip_diag: failed to open '%s'\n", path.c_str());
        return 1;
    }
    std::fprintf(fh, "step,t,r,theta,phi\n");
    for (std::uint32_t i = 0; i < n; ++i)
        std::fprintf(fh, "%d,%.9g,%.9g,%.9g,%.9g\n", rows[i].step, rows[i].t, rows[i].r, rows[i].theta,
                     rows[i].phi);
    std::fclose(fh);
    if (spin <= 0.0f)
        std::printf("wrote %s (Schwarzschild photon orbit at r = %.4g M, %.2f orbits)\n", path.c_str(), r_init / mass,
                    orbits);
    else
        std::printf("wrote %s (Kerr prograde photon orbit at r = %.4g M, a/M = %.3g, %.2f orbits)\n", path.c_str(),
                    r_init / mass, spin / mass, orbits);
    return 0;
}

int run_benchmark(const std::string& metric, float spin, int n_rays, int max_steps, float b_min, float b_max,
                  int dev) {
    if (metric != "schw" && metric != "kerr") return fail_usage("--metric must be 'schw' or 'kerr'");
    const bool kerr = metric == "kerr";
    if (kerr && (spin < 0.0f || spin >= 1.0f)) return fail_usage("--spin must be in [0, 1) for kerr benchmark");
    SingContext* ctx = open_device(dev);
    if (!ctx) return 1;
    SingFanOpts o{};
    o.metric = kerr ? SING_METRIC_KERR : SING_METRIC_SCHWARZSCHILD;
    o.spin = spin;
    o.n_rays = std::uint32_t(n_rays);
    o.max_steps = std::uint32_t(max_steps);
    o.b_min = b_min;
    o.b_max = b_max;
    SingFanStats st{};
    const int rc = sing_benchmark_fan(ctx, &o, nullptr, &st);
    if (rc != SING_OK) {
        std::fprintf(stderr, "hip_diag: benchmark failed: %s\n", sing_last_error(ctx));
        sing_destroy(ctx);
        return rc == SING_ERR_ARG ? 2 : 1;
    }
    sing_destroy(ctx);
    // cli/main.cpp:1598-1619; total_ms is the integration kernel on the GPU
    const double ms = st.kernel_ms;
    const double steps_per_s = ms > 0.0 ? double(st.total_steps) / (ms * 1e-3) : 0.0;
    const double mean_step_ns = st.total_steps > 0 ? ms * 1e6 / double(st.total_steps) : 0.0;
    std::printf("{\"mode\":\"benchmark\",\"metric\":\"%s\",\"spin\":%.6f,\"n_rays\":%d,\"max_steps\":%d,"
                "\"total_steps\":%lld,\"total_ms\":%.3f,\"steps_per_s\":%.1f,\"mean_step_ns\":%.1f,"
                "\"device\":\"%s\",\"setup_ms\":%.3f}\n",
                metric.c_str(), kerr ? spin : 0.0f, n_rays, max_steps, (long long)st.total_steps, ms, steps_per_s,
                mean_step_ns, sing_backend_name(), st.setup_ms);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    std::string mode, output_path, metric = "schw";
    float spin = 0.0f, mass = 1.0f, r_init = -1.0f, orbits = 1.0f, h_step = 0.01f;
    float b_min = -10.0f, b_max = 10.0f;
    int n_rays = 100, max_steps = 5000, device = 0;

    for (int i = 1; i < argc; ++i) {
        const std::string_view a = argv[i];
        if (a == "--help" || a == "-h") {
            print_help();
            return 0;
        }
        if (a == "--mode" && i + 1 < argc) mode = argv[++i];
        else if (a == "--output" && i + 1 < argc) output_path = argv[++i];
        else if (a == "--spin" && i + 1 < argc) spin = std::strtof(argv[++i], nullptr);
        else if (a == "--mass" && i + 1 < argc) mass = std::strtof(argv[++i], nullptr);
        else if (a == "--b-range" && i + 2 < argc) {
            b_min = std::strtof(argv[++i], nullptr);
            b_max = std::strtof(argv[++i], nullptr);
            if (b_min >= b_max) return fail_usage("--b-range MIN MAX requires MIN < MAX");
        } else if (a == "--n-rays" && i + 1 < argc) {
            n_rays = int(std::strtol(argv[++i], nullptr, 10));
            if (n_rays < 2) return fail_usage("--n-rays must be >= 2");
        } else if (a == "--max-steps" && i + 1 < argc) {
            max_steps = int(std::strtol(argv[++i], nullptr, 10));
            if (max_steps < 100) return fail_usage("--max-steps must be >= 100");
        } else if (a == "--r-init" && i + 1 < argc) r_init = std::strtof(argv[++i], nullptr);
        else if (a == "--orbits" && i + 1 < argc) orbits = std::strtof(argv[++i], nullptr);
        else if (a == "--h-step" && i + 1 < argc) h_step = std::strtof(argv[++i], nullptr);
        else if (a == "--metric" && i + 1 < argc) metric = argv[++i];
        else if (a == "--device" && i + 1 < argc) device = int(std::strtol(argv[++i], nullptr, 10));
        else if (a == "--scene" && i + 1 < argc) {
            // cli/main.cpp:1713-1745: a scene file sets the defaults, later flags override
            SingSceneConfig cfg;
            sing_scene_config_defaults(&cfg);
            int line = 0;
            char err[512] = {0};
            const int rc = sing_load_scene_config(argv[++i], &cfg, &line, err, sizeof(err));
            if (rc != SING_OK) {
                const char* why = rc == SING_ERR_IO ? "file-not-found" : rc == SING_ERR_KEY ? "unknown-key" : "parse-error";
                std::fprintf(stderr, "hip_diag: --scene load failed (%s, line %d): %s\n", why, line, err);
                return 2;
            }
            mass = cfg.mass;
            spin = cfg.spin;
            max_steps = cfg.max_steps;
            h_step = cfg.h_step;
        } else {
            std::fprintf(stderr, "hip_diag: unknown argument: %.*s\n", int(a.size()), a.data());
            print_help();
            return 2;
        }
    }

    if (mode.empty()) {
        print_help();
        return 0;
    }
    if (mode == "kerr-geometry")
        return run_kerr_geometry(output_path.empty() ? "kerr_geometry.json" : output_path, mass, spin);
    if (mode == "photon-orbit") {
        // cli/main.cpp:1780-1787: default r_init is the photon sphere (3 M, or prograde Kerr)
        if (r_init < 0.0f) {
            if (spin <= 0.0f) {
                r_init = 3.0f * mass;
            } else {
                // out-of-range mass or spin is reported by run_photon_orbit
                SingKerrGeometry g;
                if (sing_kerr_geometry(mass, spin, &g) == SING_OK) r_init = g.photon_sphere_prograde;
            }
        }
        return run_photon_orbit(output_path.empty() ? "photon_orbit.csv" : output_path, mass, spin, r_init, orbits,
                                h_step, device);
    }
    if (mode == "benchmark") return run_benchmark(metric, spin, n_rays, max_steps, b_min, b_max, device);
    if (mode == "2d-toy" || mode == "kerr-2d-toy" || mode == "disc-preview" || mode == "cpu-render" ||
        mode == "kerr-cpu-render") {
        std::fprintf(stderr, "hip_diag: mode '%s' is not provided here (images are rendered by hip_cli)\n",
                     mode.c_str());
        return 2;
    }
    std::fprintf(stderr, "hip_diag: unknown mode '%s'\n", mode.c_str());
    return 2;
}
