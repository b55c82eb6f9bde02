// singularity_amd/csrc/diag.hip — the CPU driver's diagnostics
// (cli/main.cpp:1269-1621) behind the C ABI: the Kerr geometry scalars (host
// closed forms), the circular photon orbit trail and the benchmark ray fan.
// The orbit and the fan are set up on the host in float exactly as the driver
// does, so step counts, periods and row indices come out the same; the
// integration runs on the device with the renderer's RK4 steppers
// (physics.hpp).  The Kerr RHS here keeps the reference's grouping
// (KerrKDirect): rays of the fan that graze the horizon are chaotic under
// fixed steps, and their step counts follow the rounding of that grouping.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/sing_abi.h"
#include "context.hpp"
#include "physics.hpp"
#include "scene_const.hpp"

namespace sing {
namespace {

constexpr float kPiF = 3.14159265358979323846f;

// core/include/physics/kerr.hpp:36-108 in float (the driver prints these).
float kerr_r_plus_f(float M, float a) {
    const float d = M * M - a * a;
    return M + std::sqrt(d > 0.0f ? d : 0.0f);
}
float kerr_r_minus_f(float M, float a) {
    const float d = M * M - a * a;
    return M - std::sqrt(d > 0.0f ? d : 0.0f);
}
float kerr_ergo_f(float M, float a, float theta) {
    const float c = std::cos(theta);
    const float d = M * M - a * a * c * c;
    return M + std::sqrt(d > 0.0f ? d : 0.0f);
}
float kerr_photon_f(float M, float a, bool prograde) {
    const float x = prograde ? -a / M : a / M;
    return 2.0f * M * (1.0f + std::cos((2.0f / 3.0f) * std::acos(x)));
}
float kerr_isco_f(float M, float a, bool prograde) {
    const float x = a / M;
    const float x2 = x * x;
    const float z1 = 1.0f + std::cbrt(1.0f - x2) * (std::cbrt(1.0f + x) + std::cbrt(1.0f - x));
    const float z2 = std::sqrt(3.0f * x2 + z1 * z1);
    const float rad = (3.0f - z1) * (3.0f + z1 + 2.0f * z2);
    const float root = std::sqrt(rad > 0.0f ? rad : 0.0f);
    return M * (3.0f + z2 + (prograde ? -root : root));
}

// d(phi)/d(lambda) of the Hamiltonian form (kerr_hamilton.h:104-118) at one
// state, in the driver's float expression order: it fixes the orbit's period
// and therefore its step count and row indices.
float kerr_dphi_f(float r, float theta, float E, float L, float a, float M) {
    const float st = std::sin(theta);
    const float s2 = st * st;
    const float safe = (s2 < 1e-10f) ? 1e-10f : s2;
    const float ct = std::cos(theta);
    const float c2 = ct * ct;
    const float r2 = r * r;
    const float a2 = a * a;
    const float Sigma = r2 + a2 * c2;
    const float Delta = r2 - 2.0f * M * r + a2;
    const float SD = Sigma * Delta;
    return (2.0f * M * r * a * E * s2 + (Delta - a2 * s2) * L) / (SD * safe);
}

SceneConst<float> metric_const(float M, float a) {
    SceneConst<float> sc{};
    sc.M = M;
    sc.a = a;
    sc.a2 = a * a;
    sc.twoM = 2.0f * M;
    sc.rs = 2.0f * M;
    return sc;
}

struct OrbitPlan {
    float r0, th0, v0, v1;  // Schwarzschild: u^t, u^phi; Kerr: E, L_z
    float h, cut;
    int total, every;
    bool kerr;
};

// One lane integrates the whole orbit (it is a single ray); rows are written
// as the driver writes its CSV lines.  The host sized `rows` for the largest
// count the plan can produce (1 + total/every + 1).
template <bool KERR>
__global__ __launch_bounds__(64) void orbit_kernel(OrbitPlan p, SceneConst<float> sc, SingOrbitRow* __restrict__ rows,
                                                   std::uint32_t cap, std::uint32_t* __restrict__ n_rows) {
    if (threadIdx.x != 0) return;
    float r = p.r0, th = p.th0, ph = 0.0f, t = 0.0f;
    float ut = p.v0, ur = 0.0f, uth = 0.0f, uph = p.v1;  // Schwarzschild velocities
    float pr = 0.0f, pth = 0.0f;                         // Kerr momenta
    KerrKDirect<float> k{};
    if (KERR) k = kerr_k_direct(sc, p.v0, p.v1);
    std::uint32_t n = 0;
    auto emit = [&](int step) {
        if (n < cap) {
            SingOrbitRow row;
            row.step = step;
            row.t = t;
            row.r = r;
            row.theta = th;
            row.phi = ph;
            rows[n] = row;
        }
        ++n;
    };
    emit(0);
    for (int step = 1; step <= p.total; ++step) {
        if (KERR) kerr_rk4<float, true>(r, th, ph, pr, pth, p.h, sc, k, &t);
        else schw_rk4<float, true>(r, th, ph, ut, ur, uth, uph, p.h, sc.rs, &t);
        if (!__builtin_isfinite(r) || r < p.cut) {
            emit(-1);
            break;
        }
        if (step % p.every == 0 || step == p.total) emit(step);
    }
    *n_rows = n;
}

// Initial state of one fan ray: Schwarzschild (r, phi, u^t, u^r, u^phi),
// Kerr (r, phi, p_r, E, L_z); r = NaN marks a Kerr ray the driver skips.
struct FanRay {
    float r, ph, v0, v1, v2, pad0, pad1, pad2;
};

template <bool KERR>
__global__ __launch_bounds__(256) void fan_kernel(const FanRay* __restrict__ rays, SceneConst<float> sc, float h,
                                                  float cut, std::uint32_t n, std::uint32_t max_steps,
                                                  std::uint32_t* __restrict__ steps) {
    const std::uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const FanRay q = rays[i];
    float r = q.r, th = 0.5f * kPiF, ph = q.ph;
    std::uint32_t cnt = 0;
    if (KERR) {
        float pr = q.v0, pth = 0.0f;
        const KerrKDirect<float> k = kerr_k_direct(sc, q.v1, q.v2);
        for (std::uint32_t s = 0; s < max_steps; ++s) {
            if (!__builtin_isfinite(r) || !__builtin_isfinite(ph)) break;
            if (r < cut) break;
            kerr_rk4(r, th, ph, pr, pth, h, sc, k);
            ++cnt;
        }
    } else {
        float ut = q.v0, ur = q.v1, uth = 0.0f, uph = q.v2;
        for (std::uint32_t s = 0; s < max_steps; ++s) {
            if (!__builtin_isfinite(r) || !__builtin_isfinite(ph)) break;
            if (r < cut) break;
            schw_rk4(r, th, ph, ut, ur, uth, uph, h, sc.rs);
            ++cnt;
        }
    }
    steps[i] = cnt;
}

}  // namespace
}  // namespace sing

using sing::kPiF;

extern "C" {

int sing_kerr_geometry(float M, float a, SingKerrGeometry* out) {
    char buf[160];
    if (!out) {
        sing_set_global_error("sing_kerr_geometry: null output");
        return SING_ERR_ARG;
    }
    if (!(M > 0.0f)) {
        sing_set_global_error("mass must be positive");
        return SING_ERR_ARG;
    }
    if (!(a >= 0.0f && a <= M)) {
        std::snprintf(buf, sizeof(buf), "spin %.6f out of range [0, M=%.6f]", a, M);
        sing_set_global_error(buf);
        return SING_ERR_ARG;
    }
    out->mass = M;
    out->spin = a;
    out->spin_over_mass = a / M;
    out->outer_horizon = sing::kerr_r_plus_f(M, a);
    out->inner_horizon = sing::kerr_r_minus_f(M, a);
    out->ergosphere_polar = sing::kerr_ergo_f(M, a, 0.0f);
    out->ergosphere_equatorial = sing::kerr_ergo_f(M, a, 0.5f * kPiF);
    out->photon_sphere_prograde = sing::kerr_photon_f(M, a, true);
    out->photon_sphere_retrograde = sing::kerr_photon_f(M, a, false);
    out->isco_prograde = sing::kerr_isco_f(M, a, true);
    out->isco_retrograde = sing::kerr_isco_f(M, a, false);
    return SING_OK;
}

int sing_photon_orbit(SingContext* c, float M, float a, float r_init, float orbits, float h, SingOrbitRow* rows,
                      std::uint32_t cap, std::uint32_t* n_rows, double* kernel_ms) {
    if (!c) return SING_ERR_ARG;
    c->complete_pending();
    if (!n_rows) return c->fail(SING_ERR_ARG, "sing_photon_orbit: null n_rows");
    *n_rows = 0;
    // cli/main.cpp:1282-1305: argument checks in the driver's order
    if (!(M > 0.0f)) return c->fail(SING_ERR_ARG, "mass must be positive");
    if (!(a >= 0.0f && a <= M)) return c->fail(SING_ERR_ARG, "spin must be in [0, M]");
    const float rs = 2.0f * M;
    const float r_plus = a > 0.0f ? sing::kerr_r_plus_f(M, a) : rs;
    if (!(r_init > r_plus * 1.001f))
        return c->fail(SING_ERR_ARG, "r_init %.4g must be > r_+ = %.4g", r_init, r_plus);
    if (!(orbits > 0.0f)) return c->fail(SING_ERR_ARG, "orbits must be positive");
    if (!(h > 0.0f)) return c->fail(SING_ERR_ARG, "h_step must be positive");

    sing::OrbitPlan p{};
    p.r0 = r_init;
    p.th0 = 0.5f * kPiF;
    p.h = h;
    p.cut = r_plus * 1.001f;
    p.kerr = a > 0.0f;
    float period;
    if (!p.kerr) {  // cli/main.cpp:1308-1325
        const float f = 1.0f - rs / r_init;
        p.v0 = 1.0f;
        p.v1 = std::sqrt(f) / r_init;
        period = 2.0f * kPiF / p.v1;
    } else {  // cli/main.cpp:1352-1390
        const float Delta = r_init * r_init - 2.0f * M * r_init + a * a;
        if (!(Delta > 0.0f))
            return c->fail(SING_ERR_ARG, "Delta(%.4g, a=%.4g) <= 0: r_init inside the horizon", r_init, a);
        const float b_pro = (r_init * std::sqrt(Delta) - 2.0f * M * a) / (r_init - 2.0f * M);
        p.v0 = 1.0f;
        p.v1 = b_pro;
        const float dphi = sing::kerr_dphi_f(r_init, p.th0, 1.0f, b_pro, a, M);
        if (!std::isfinite(dphi) || dphi <= 0.0f)
            return c->fail(SING_ERR_ARG, "bad Kerr init (dphi/dlambda = %.3e)", dphi);
        period = 2.0f * kPiF / dphi;
    }
    const float total_f = orbits * period / h;
    // one lane integrates the orbit (about a microsecond per step): keep a call under a few minutes
    if (!(total_f <= 1.0e8f)) return c->fail(SING_ERR_ARG, "orbit needs %.3g steps (limit 1e8)", total_f);
    p.total = int(total_f);
    p.every = std::max(1, int(period / (h * 200.0f)));
    const std::uint64_t bound = 2u + std::uint64_t(std::max(p.total, 0)) / std::uint64_t(p.every);
    if (bound > (1u << 26)) return c->fail(SING_ERR_ARG, "orbit produces %llu rows (limit 2^26)",
                                           (unsigned long long)bound);
    if (!rows || cap < bound) {
        *n_rows = std::uint32_t(bound);
        return c->fail(SING_ERR_ARG, "sing_photon_orbit: rows buffer holds %u rows, the orbit can produce %llu", cap,
                       (unsigned long long)bound);
    }

    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return c->hip_fail(e, "hipSetDevice");
    int rc = c->ensure_events(2);
    if (rc != SING_OK) return rc;
    SingOrbitRow* d_rows = nullptr;
    std::uint32_t* d_n = nullptr;
    if ((e = hipMalloc(&d_rows, bound * sizeof(SingOrbitRow))) != hipSuccess) return c->hip_fail(e, "hipMalloc(rows)");
    if ((e = hipMalloc(&d_n, sizeof(std::uint32_t))) != hipSuccess) {
        hipFree(d_rows);
        return c->hip_fail(e, "hipMalloc(count)");
    }
    const sing::SceneConst<float> sc = sing::metric_const(M, a);
    std::uint32_t n = 0;
    hipEventRecord(c->events[0], c->stream);
    if (p.kerr)
        hipLaunchKernelGGL(sing::orbit_kernel<true>, dim3(1), dim3(64), 0, c->stream, p, sc, d_rows,
                           std::uint32_t(bound), d_n);
    else
        hipLaunchKernelGGL(sing::orbit_kernel<false>, dim3(1), dim3(64), 0, c->stream, p, sc, d_rows,
                           std::uint32_t(bound), d_n);
    e = hipGetLastError();
    hipEventRecord(c->events[1], c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&n, d_n, sizeof(n), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess && n > bound) {
        hipFree(d_rows);
        hipFree(d_n);
        return c->fail(SING_ERR_STATE, "sing_photon_orbit: kernel reported %u rows above the bound %llu", n,
                       (unsigned long long)bound);
    }
    if (e == hipSuccess) e = hipMemcpy(rows, d_rows, std::size_t(n) * sizeof(SingOrbitRow), hipMemcpyDeviceToHost);
    hipFree(d_rows);
    hipFree(d_n);
    if (e != hipSuccess) return c->hip_fail(e, "photon orbit");
    if (kernel_ms) {
        float ms = 0.0f;
        hipEventElapsedTime(&ms, c->events[0], c->events[1]);
        *kernel_ms = ms;
    }
    *n_rows = n;
    return SING_OK;
}

int sing_benchmark_fan(SingContext* c, const SingFanOpts* o, std::uint32_t* steps_per_ray, SingFanStats* stats) {
    if (!c) return SING_ERR_ARG;
    c->complete_pending();
    if (!o) return c->fail(SING_ERR_ARG, "sing_benchmark_fan: null options");
    const bool kerr = o->metric == SING_METRIC_KERR;
    if (!kerr && o->metric != SING_METRIC_SCHWARZSCHILD)
        return c->fail(SING_ERR_ARG, "sing_benchmark_fan: unknown metric %u", o->metric);
    if (kerr && !(o->spin >= 0.0f && o->spin < 1.0f))
        return c->fail(SING_ERR_ARG, "spin must be in [0, 1) for kerr benchmark");
    if (o->n_rays < 1 || o->n_rays > (1u << 28)) return c->fail(SING_ERR_ARG, "n_rays must be in [1, 2^28]");
    if (!(o->b_min < o->b_max)) return c->fail(SING_ERR_ARG, "b_min must be < b_max");

    // cli/main.cpp:1499-1577: the fan's initial states, in the driver's float order
    const auto t0 = std::chrono::steady_clock::now();
    constexpr float M = 1.0f, rs = 2.0f * M, h = 0.1f;
    const float a = kerr ? o->spin * M : 0.0f;
    const float b_min = o->b_min * M, b_max = o->b_max * M;
    const float x0 = std::max(50.0f * M, 2.0f * std::max(std::fabs(b_min), std::fabs(b_max)));
    const std::uint32_t n = o->n_rays;
    std::vector<sing::FanRay> rays(n);
    for (std::uint32_t i = 0; i < n; ++i) {
        const float b = b_min + (b_max - b_min) * (float(i) + 0.5f) / float(n);
        const float r0 = std::sqrt(x0 * x0 + b * b);
        const float phi0 = std::atan2(b, x0);
        const float ur = -std::cos(phi0);
        const float uphi = std::sin(phi0) / r0;
        sing::FanRay q{};
        q.r = r0;
        q.ph = phi0;
        if (!kerr) {
            const float f = 1.0f - rs / r0;
            q.v0 = std::sqrt(((ur * ur) / f + r0 * r0 * uphi * uphi) / f);
            q.v1 = ur;
            q.v2 = uphi;
        } else {
            const float Delta = r0 * r0 - 2.0f * M * r0 + a * a;
            const float g_tt = -(1.0f - 2.0f * M / r0);
            const float g_tp = -2.0f * M * a / r0;
            const float g_rr = r0 * r0 / Delta;
            const float g_pp = ((r0 * r0 + a * a) * (r0 * r0 + a * a) - a * a * Delta) / (r0 * r0);
            const float qa = g_tt, qb = 2.0f * g_tp * uphi, qc = g_rr * ur * ur + g_pp * uphi * uphi;
            const float disc = qb * qb - 4.0f * qa * qc;
            if (disc < 0.0f) {
                q.r = std::nanf("");  // the driver skips this ray
            } else {
                const float ut = (-qb - std::sqrt(disc)) / (2.0f * qa);
                // p_r = Sigma u^r / Delta at theta = pi/2 (kerr_hamilton.h:192-200)
                const float ct = std::cos(0.5f * kPiF);
                const float Sig = r0 * r0 + a * a * ct * ct;
                const float Del = r0 * r0 - 2.0f * M * r0 + a * a;
                q.v0 = Sig * ur / Del;
                q.v1 = -(g_tt * ut + g_tp * uphi);
                q.v2 = g_tp * ut + g_pp * uphi;
            }
        }
        rays[i] = q;
    }
    const float cut = kerr ? 1.02f * sing::kerr_r_plus_f(M, a) : 1.01f * rs;

    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return c->hip_fail(e, "hipSetDevice");
    int rc = c->ensure_events(2);
    if (rc != SING_OK) return rc;
    sing::FanRay* d_rays = nullptr;
    std::uint32_t* d_steps = nullptr;
    if ((e = hipMalloc(&d_rays, std::size_t(n) * sizeof(sing::FanRay))) != hipSuccess)
        return c->hip_fail(e, "hipMalloc(fan)");
    if ((e = hipMalloc(&d_steps, std::size_t(n) * sizeof(std::uint32_t))) != hipSuccess) {
        hipFree(d_rays);
        return c->hip_fail(e, "hipMalloc(fan steps)");
    }
    e = hipMemcpyAsync(d_rays, rays.data(), std::size_t(n) * sizeof(sing::FanRay), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    const double setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    const sing::SceneConst<float> sc = sing::metric_const(M, a);
    const dim3 grid((n + 255u) / 256u), block(256);
    if (e == hipSuccess) {
        hipEventRecord(c->events[0], c->stream);
        if (kerr)
            hipLaunchKernelGGL(sing::fan_kernel<true>, grid, block, 0, c->stream, d_rays, sc, h, cut, n, o->max_steps,
                               d_steps);
        else
            hipLaunchKernelGGL(sing::fan_kernel<false>, grid, block, 0, c->stream, d_rays, sc, h, cut, n,
                               o->max_steps, d_steps);
        e = hipGetLastError();
        hipEventRecord(c->events[1], c->stream);
    }
    std::vector<std::uint32_t> counts(n);
    if (e == hipSuccess)
        e = hipMemcpyAsync(counts.data(), d_steps, std::size_t(n) * sizeof(std::uint32_t), hipMemcpyDeviceToHost,
                           c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    hipFree(d_rays);
    hipFree(d_steps);
    if (e != hipSuccess) return c->hip_fail(e, "benchmark fan");
    std::uint64_t total = 0;
    for (std::uint32_t v : counts) total += v;
    if (steps_per_ray) std::memcpy(steps_per_ray, counts.data(), std::size_t(n) * sizeof(std::uint32_t));
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        float ms = 0.0f;
        hipEventElapsedTime(&ms, c->events[0], c->events[1]);
        stats->total_steps = total;
        stats->kernel_ms = ms;
        stats->setup_ms = setup_ms;
    }
    return SING_OK;
}

}  // extern "C"
