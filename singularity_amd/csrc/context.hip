// singularity_amd/csrc/context.hip — the C ABI's render context: device
// buffers, the frame pipeline (trace -> [FP64 pass] -> shade per sample chunk,
// then the optional post passes), stats and capture.  Replaces CudaBackend
// (render/cuda/cuda_backend.cu:105-166): the context owns every device
// allocation, render is synchronous, capture copies to caller memory, errors
// are returned as status codes with a per-context message.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sing_abi.h"
#include "context.hpp"
#include "post.hpp"
#include "scene_const.hpp"
#include "trace.hpp"

namespace {

// The last error of the calling thread (as errno): contexts may be driven
// from several host threads at once, and a pointer handed out by
// sing_last_error(NULL) must not change under its reader.
thread_local std::string g_last_error;

// Device memory budget for one chunk of termination records (16 B per ray).
constexpr std::size_t kRecordBudget = std::size_t(2) << 30;

}  // namespace

void sing_set_global_error(const std::string& m) { g_last_error = m; }

int SingContext::fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    err = buf;
    sing_set_global_error(err);
    return code;
}

int SingContext::hip_fail(hipError_t e, const char* where) {
    return fail(SING_ERR_HIP, "%s failed: %s", where, hipGetErrorString(e));
}

void SingContext::free_frame() {
    hipFree(d_rgba);
    hipFree(d_hdr);
    hipFree(d_hdr16);
    hipFree(d_cls);
    hipFree(d_steps);
    hipFree(d_acc);
    d_rgba = nullptr;
    d_hdr = nullptr;
    d_hdr16 = nullptr;
    d_cls = nullptr;
    d_steps = nullptr;
    d_acc = nullptr;
    outputs_valid = 0;
}

int SingContext::alloc_frame(std::uint32_t w, std::uint32_t h) {
    const std::size_t n = std::size_t(w) * h;
    hipError_t e;
    if ((e = hipMalloc(&d_rgba, n * 4)) != hipSuccess || (e = hipMalloc(&d_hdr, n * 12)) != hipSuccess ||
        (e = hipMalloc(&d_hdr16, n * 8)) != hipSuccess || (e = hipMalloc(&d_cls, n)) != hipSuccess ||
        (e = hipMalloc(&d_steps, n * 4)) != hipSuccess ||
        (e = hipMalloc(&d_acc, n * sizeof(sing::PixAcc))) != hipSuccess) {
        free_frame();
        return fail(SING_ERR_NOMEM, "hipMalloc of %ux%u frame buffers failed: %s", w, h, hipGetErrorString(e));
    }
    width = w;
    height = h;
    return SING_OK;
}

int SingContext::ensure_records(std::size_t rays) {
    if (rays <= rec_capacity) return SING_OK;
    hipFree(d_rec);
    hipFree(d_band);
    d_rec = nullptr;
    d_band = nullptr;
    rec_capacity = 0;
    hipError_t e;
    if ((e = hipMalloc(&d_rec, rays * sizeof(sing::Rec))) != hipSuccess ||
        (e = hipMalloc(&d_band, rays * sizeof(std::uint32_t))) != hipSuccess) {
        hipFree(d_rec);
        d_rec = nullptr;
        return fail(SING_ERR_NOMEM, "hipMalloc of %zu ray records failed: %s", rays, hipGetErrorString(e));
    }
    rec_capacity = rays;
    return SING_OK;
}

int SingContext::ensure_items(std::size_t items) {
    if (items <= item_capacity) return SING_OK;
    hipFree(d_init);
    hipFree(d_item_rec);
    d_init = nullptr;
    d_item_rec = nullptr;
    item_capacity = 0;
    hipError_t e;
    if ((e = hipMalloc(&d_init, items * sizeof(float4))) != hipSuccess ||
        (e = hipMalloc(&d_item_rec, items * sizeof(std::uint32_t))) != hipSuccess) {
        hipFree(d_init);
        d_init = nullptr;
        return fail(SING_ERR_NOMEM, "hipMalloc of %zu ray set-up entries failed: %s", items, hipGetErrorString(e));
    }
    item_capacity = items;
    return SING_OK;
}

int SingContext::ensure_events(std::size_t n) {
    while (events.size() < n) {
        hipEvent_t e;
        const hipError_t r = hipEventCreate(&e);
        if (r != hipSuccess) return hip_fail(r, "hipEventCreate");
        events.push_back(e);
    }
    return SING_OK;
}

int SingContext::ensure_join_events(std::size_t n) {
    while (join_events.size() < n) {
        hipEvent_t e;
        const hipError_t r = hipEventCreateWithFlags(&e, hipEventDisableTiming);
        if (r != hipSuccess) return hip_fail(r, "hipEventCreateWithFlags");
        join_events.push_back(e);
    }
    return SING_OK;
}

std::uint32_t SingContext::chunk_samples(std::uint32_t spp) const {
    const std::size_t P = std::size_t(width) * height;
    std::size_t S = kRecordBudget / (P * sizeof(sing::Rec));
    S = std::max<std::size_t>(S, 1);
    S = std::min<std::size_t>(S, spp);
    S = std::min<std::size_t>(S, (std::size_t(1) << 31) / P);  // record indices stay in u32
    return std::uint32_t(std::max<std::size_t>(S, 1));
}

sing::LaunchShape SingContext::launch_shape(const SingRenderOpts& o) const { return sing::launch_shape_for(num_cu, o); }

sing::LaunchShape sing::launch_shape_for(int num_cu, const SingRenderOpts& o) {
    sing::LaunchShape sh{};
    // 24 waves per CU (the trace kernels' occupancy limit) from a launch-shape
    // sweep of the 4K Kerr frame; step batch 0 = each pass's own default
    // (launch_trace_chunk).  A step batch is a multiple of 8 (rounded up): the
    // step loop's sin/cos phase relies on it.
    sh.step_batch = o.step_batch ? std::uint32_t(std::min<std::uint64_t>((std::uint64_t(o.step_batch) + 7u) & ~7ull,
                                                                         0xFFFFFFF8ull))
                                 : 0u;
    const std::uint32_t wpc = o.waves_per_cu ? o.waves_per_cu : 24u;
    sh.blocks = std::uint32_t(num_cu) * ((wpc + 3u) / 4u);
    sh.blocks64 = std::uint32_t(num_cu) * ((std::max(wpc / 2u, 4u) + 3u) / 4u);
    return sh;
}

namespace sing {

bool render_args_ok(const SingUniforms& u, const SingRenderOpts& o, const char* who, char* msg, std::size_t len) {
    char why[96] = "";
    if (o.precision > SING_PREC_MIXED) std::snprintf(why, sizeof(why), "unknown precision %u", o.precision);
    else if (o.post > SING_POST_BLOOM) std::snprintf(why, sizeof(why), "unknown post pass %u", o.post);
    else if (o.sampler > SING_SAMPLER_GRID) std::snprintf(why, sizeof(why), "unknown sampler %u", o.sampler);
    else if (u.metric_type > SING_METRIC_KERR) std::snprintf(why, sizeof(why), "unknown metric %u", u.metric_type);
    else if (!(o.adapt_tol >= 0.0f && o.adapt_tol <= 0.1f)) std::snprintf(why, sizeof(why), "adapt_tol must be in [0, 0.1]");
    else if (o.adapt_tol > 0.0f && u.metric_type != SING_METRIC_KERR)
        std::snprintf(why, sizeof(why), "error-controlled stepping (adapt_tol) is Kerr-only");
    else if (o.refine & ~(SING_REFINE_CROSS64 | SING_REFINE_REPROJECT))
        std::snprintf(why, sizeof(why), "unknown refine bits 0x%x", o.refine);
    else if (o.refine && (u.metric_type != SING_METRIC_KERR || o.adapt_tol > 0.0f))
        std::snprintf(why, sizeof(why), "refine options apply to the fixed-step Kerr march only");
    else if (o.post == SING_POST_BLOOM && (u.width < 2 || u.height < 2))
        std::snprintf(why, sizeof(why), "bloom needs at least 2x2 pixels");
    if (!why[0]) return true;
    std::snprintf(msg, len, "%s: %s", who, why);
    return false;
}

}  // namespace sing

extern "C" {

int sing_abi_version(void) { return SING_ABI_VERSION; }
const char* sing_backend_name(void) { return "HIP-gfx950"; }

const char* sing_last_error(const SingContext* ctx) {
    if (ctx) return ctx->err.c_str();
    return g_last_error.c_str();
}

SingContext* sing_create(uint32_t width, uint32_t height, int device) {
    if (width < 1 || height < 1 || width > 32768 || height > 32768) {
        sing_set_global_error("sing_create: width/height must be in [1, 32768]");
        return nullptr;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        sing_set_global_error("sing_create: no HIP device available");
        return nullptr;
    }
    if (device < 0 || device >= ndev) {
        sing_set_global_error("sing_create: device index out of range");
        return nullptr;
    }
    auto* c = new SingContext();
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&c->d_ctl, 2 * sizeof(sing::QueueCtl));  // one per chunk parity
    if (e == hipSuccess) e = hipMalloc(&c->d_halton, (1024 + sing::kGridTableLen) * sizeof(float2));
    if (e == hipSuccess) e = hipMalloc(&c->d_sd, sizeof(sing::SceneConst<double>));
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->h_ctl, 2 * sizeof(sing::QueueCtl), hipHostMallocDefault);
    if (e != hipSuccess) {
        sing_set_global_error(std::string("sing_create: HIP setup failed: ") + hipGetErrorString(e));
        sing_destroy(c);
        return nullptr;
    }
    // Halton(2,3) offsets of samples 1..1024 (geodesic_kernel.cu:491-495),
    // followed by the grid sampler's tables for ss = 1..4.
    std::vector<float2> tab(1024 + sing::kGridTableLen);
    for (std::uint32_t s = 0; s < 1024; ++s)
        tab[s] = make_float2(sing::halton_f(s + 1u, 2u), sing::halton_f(s + 1u, 3u));
    for (std::uint32_t ss = 1; ss <= 4u; ++ss)
        for (std::uint32_t s = 0; s < ss * ss; ++s) {
            float ox, oy;
            sing::grid_offset_f(s, ss, ox, oy);
            tab[1024 + sing::grid_table_base(ss) + s] = make_float2(ox, oy);
        }
    e = hipMemcpy(c->d_halton, tab.data(), tab.size() * sizeof(float2), hipMemcpyHostToDevice);
    if (e != hipSuccess || c->alloc_frame(width, height) != SING_OK || c->ensure_events(8) != SING_OK) {
        sing_set_global_error(e != hipSuccess ? std::string("sing_create: halton upload failed") : c->err);
        sing_destroy(c);
        return nullptr;
    }
    return c;
}

void sing_destroy(SingContext* c) {
    if (!c) return;
    hipSetDevice(c->device);
    c->complete_pending();
    if (c->stream) hipStreamSynchronize(c->stream);
    c->free_frame();
    hipFree(c->d_ctl);
    hipFree(c->d_halton);
    hipFree(c->d_sd);
    hipFree(c->d_rec);
    hipFree(c->d_band);
    hipFree(c->d_init);
    hipFree(c->d_item_rec);
    if (c->h_ctl) hipHostFree(c->h_ctl);
    c->free_diff();
    c->free_batch();
    for (auto& e : c->events) hipEventDestroy(e);
    for (auto& e : c->join_events) hipEventDestroy(e);
    if (c->stream2) {
        hipStreamSynchronize(c->stream2);
        hipStreamDestroy(c->stream2);
    }
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

int sing_resize(SingContext* c, uint32_t w, uint32_t h) {
    if (!c) return SING_ERR_ARG;
    if (w < 1 || h < 1 || w > 32768 || h > 32768) return c->fail(SING_ERR_ARG, "sing_resize: bad size %ux%u", w, h);
    c->complete_pending();
    if (w == c->width && h == c->height) return SING_OK;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    c->free_frame();
    return c->alloc_frame(w, h);
}

int sing_render_async(SingContext* c, const SingUniforms* u, const SingRenderOpts* opts_in) {
    if (!c) return SING_ERR_ARG;
    c->complete_pending();
    c->stats_unclaimed = false;
    if (!u) return c->fail(SING_ERR_ARG, "sing_render: null uniforms");
    if (u->width != c->width || u->height != c->height)
        return c->fail(SING_ERR_ARG, "sing_render: uniforms size %ux%u != context %ux%u", u->width, u->height,
                       c->width, c->height);
    SingRenderOpts o{};
    if (opts_in) o = *opts_in;
    char msg[160];
    if (!sing::render_args_ok(*u, o, "sing_render", msg, sizeof(msg))) return c->fail(SING_ERR_ARG, "%s", msg);
    c->outputs_valid = 0;  // the outputs are overwritten from here on
    std::uint32_t outs = o.outputs ? o.outputs : SING_OUT_RGBA8;
    if (o.post == SING_POST_BLOOM) outs |= SING_OUT_HDR16 | SING_OUT_RGBA8;
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return c->hip_fail(e, "hipSetDevice");

    const sing::SceneConst<float> sf = sing::make_scene_const<float>(*u, o);
    const sing::SceneConst<double> sd = sing::make_scene_const<double>(*u, o);
    sing::TraceOutputs to{};
    // With the bloom pass the RGBA8 image is the composite, not the trace's own.
    to.rgba = (outs & SING_OUT_RGBA8) && o.post != SING_POST_BLOOM ? c->d_rgba : nullptr;
    to.hdr = (outs & SING_OUT_HDR) ? c->d_hdr : nullptr;
    to.hdr16 = (outs & SING_OUT_HDR16) ? c->d_hdr16 : nullptr;
    to.cls = (outs & SING_OUT_CLASS) ? c->d_cls : nullptr;
    to.steps = (outs & SING_OUT_STEPS) ? c->d_steps : nullptr;

    const std::uint32_t spp = sf.spp;
    const std::uint32_t S = o.chunk_samples ? std::min(c->chunk_samples(spp), o.chunk_samples) : c->chunk_samples(spp);
    const std::uint32_t n_chunks = (spp + S - 1u) / S;
    const std::size_t P = std::size_t(c->width) * c->height;
    // Overlapped chunks (overlap_chunks, mixed form, two or more): chunk k's tail — the FP64
    // list pass, then its shade — runs on the second stream while the context
    // stream goes on with chunk k+1's set-up and FP32 pass, so the list pass's
    // long rays and the shade share the GPU with FP32 work instead of leaving
    // lanes idle.  The chunk buffers (records, FP64 list, set-up arrays, queue
    // counters) are doubled and used by chunk parity; chunk k+2 waits for
    // chunk k's shade before it reuses them.  Shades stay in chunk order on
    // one stream and every ray's arithmetic is unchanged: the frame is
    // bit-identical to the serial order.
    // On request only, and for the mixed form only: a plain FP32 frame gains
    // nothing (0.0 % on the 4K headline) and an all-FP64 frame loses 3.7 % when
    // its shades run beside the next chunk's FP64 pass; the mixed form gained
    // 1.2 % on the headline while its FP64 list ran in record order, and with
    // the list's longest-first order the serial schedule is as fast there and
    // 1.6 % faster at a = 0.998 (overlapped, -0 % and +2 %).
    const bool overlap = n_chunks >= 2 && o.overlap_chunks && o.precision == SING_PREC_MIXED;
    const std::uint32_t nbuf = overlap ? 2u : 1u;
    hipStream_t sa = c->stream, sb = overlap ? c->stream2 : c->stream;
    int rc = c->ensure_records(std::size_t(nbuf) * S * P);
    if (rc != SING_OK) return rc;
    // per chunk: start, front (set-up, FP32 pass, audit) done, shade done,
    // shade start, and the FP32/FP64 trace kernels' own brackets; then the
    // post pass's two and the frame's end
    if ((rc = c->ensure_events(8u * n_chunks + 4u)) != SING_OK) return rc;
    if (overlap && (rc = c->ensure_join_events(2u * n_chunks)) != SING_OK) return rc;

    sing::Chunk ch{};
    ch.P = std::uint32_t(P);
    ch.tiles_x = (c->width + 7u) / 8u;
    ch.n_tiles = ch.tiles_x * ((c->height + 7u) / 8u);
    const std::size_t items = std::size_t(S) * ch.n_tiles * 64u;
    if (o.precision != SING_PREC_FP64 && (rc = c->ensure_items(nbuf * items)) != SING_OK) return rc;
    const sing::LaunchShape shape = c->launch_shape(o);
    const float2* offsets = c->d_halton;
    if (sf.grid) {
        std::uint32_t ss = 1;
        while (ss * ss < spp) ++ss;
        offsets = c->d_halton + 1024 + sing::grid_table_base(ss);
    }

    if ((e = hipMemsetAsync(c->d_ctl, 0, nbuf * sizeof(sing::QueueCtl), sa)) != hipSuccess)
        return c->hip_fail(e, "hipMemsetAsync(queue)");
    c->h_sd = sd;
    if ((e = hipMemcpyAsync(c->d_sd, &c->h_sd, sizeof(sd), hipMemcpyHostToDevice, sa)) != hipSuccess)
        return c->hip_fail(e, "hipMemcpyAsync(scene constants)");
    ch.sd = c->d_sd;
    hipEvent_t* ev = c->events.data();
    hipEvent_t* join = overlap ? c->join_events.data() : nullptr;  // per chunk: front done, shade done
    // a failure after the first launch returns only after both streams have
    // drained, so no queued pass of this frame still uses the chunk buffers or
    // the queue counters when the next call resets them
    auto bail = [&](hipError_t err, const char* where) {
        if (overlap) (void)hipStreamSynchronize(sb);
        (void)hipStreamSynchronize(sa);
        return c->hip_fail(err, where);
    };
    for (std::uint32_t k = 0; k < n_chunks; ++k) {
        const std::uint32_t b = k % nbuf;
        ch.s0 = k * S;
        ch.S = std::min(S, spp - ch.s0);
        ch.rec = c->d_rec + b * S * P;
        ch.band = c->d_band + b * S * P;
        ch.init = o.precision != SING_PREC_FP64 ? c->d_init + b * items : nullptr;
        ch.item_rec = o.precision != SING_PREC_FP64 ? c->d_item_rec + b * items : nullptr;
        sing::QueueCtl* ctl = c->d_ctl + b;
        if (overlap && k >= 2 && (e = hipStreamWaitEvent(sa, join[2 * (k - 2) + 1], 0)) != hipSuccess)
            return bail(e, "hipStreamWaitEvent(chunk buffers)");
        // the work queues restart per chunk; the frame counters keep accumulating
        if (k >= nbuf && (e = hipMemsetAsync(ctl, 0, offsetof(sing::QueueCtl, steps), sa)) != hipSuccess)
            return bail(e, "hipMemsetAsync(queue)");
        hipEventRecord(ev[8 * k + 0], sa);
        const bool mixed = o.precision == SING_PREC_MIXED;
        if ((e = sing::launch_trace_chunk(sf, sd, offsets, ch, ctl, o.precision, shape, sa, ev + 8 * k + 4,
                                          mixed ? sb : nullptr, overlap ? join[2 * k] : nullptr)) != hipSuccess)
            return bail(e, "trace kernel launch");
        hipEventRecord(ev[8 * k + 1], sa);
        if (overlap && !mixed) {
            if ((e = hipEventRecord(join[2 * k], sa)) != hipSuccess ||
                (e = hipStreamWaitEvent(sb, join[2 * k], 0)) != hipSuccess)
                return bail(e, "cross-stream join");
        }
        hipEventRecord(ev[8 * k + 3], sb);
        if ((e = sing::launch_shade_chunk(sf, ch, c->d_acc, k == 0, k + 1 == n_chunks, to, sb)) != hipSuccess)
            return bail(e, "shade kernel launch");
        hipEventRecord(ev[8 * k + 2], sb);
        if (overlap && (e = hipEventRecord(join[2 * k + 1], sb)) != hipSuccess)
            return bail(e, "hipEventRecord(shade done)");
    }
    if (overlap && (e = hipStreamWaitEvent(sa, join[2 * (n_chunks - 1) + 1], 0)) != hipSuccess)
        return bail(e, "hipStreamWaitEvent(last shade)");
    hipEvent_t post0 = ev[8 * n_chunks], post1 = ev[8 * n_chunks + 1], fin = ev[8 * n_chunks + 2];
    if (o.post == SING_POST_BLOOM) {
        hipEventRecord(post0, sa);
        e = sing::launch_bloom(c->d_hdr16, c->width, c->height, u->bloom_threshold, u->bloom_strength, u->exposure,
                               c->d_rgba, sa);
        if (e != hipSuccess) return bail(e, "bloom launch");
        hipEventRecord(post1, sa);
    }
    hipEventRecord(fin, sa);
    if ((e = hipMemcpyAsync(c->h_ctl, c->d_ctl, nbuf * sizeof(sing::QueueCtl), hipMemcpyDeviceToHost, sa)) !=
        hipSuccess)
        return bail(e, "hipMemcpyAsync(stats)");
    c->pend = SingContext::PendingFrame{n_chunks, nbuf, outs, o.precision, o.post, overlap, std::uint64_t(P) * spp};
    c->pending = true;
    return SING_OK;
}

int SingContext::complete_pending() {
    if (!pending) return SING_OK;
    pending = false;
    stats_unclaimed = true;
    std::memset(&last_stats, 0, sizeof(last_stats));
    hipSetDevice(device);
    const hipError_t e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return last_rc = hip_fail(e, "hipStreamSynchronize");
    const PendingFrame& f = pend;
    sing::QueueCtl ctl = h_ctl[0];
    if (f.nbuf == 2u) {
        ctl.steps += h_ctl[1].steps;
        ctl.rays += h_ctl[1].rays;
        ctl.rays64 += h_ctl[1].rays64;
        ctl.rejected += h_ctl[1].rejected;
        ctl.steps32 += h_ctl[1].steps32;
        ctl.steps64 += h_ctl[1].steps64;
    }
    hipEvent_t* ev = events.data();
    const std::uint32_t n_chunks = f.n_chunks;
    double trace_ms = 0.0, shade_ms = 0.0, post_ms = 0.0, t32_ms = 0.0, t64_ms = 0.0;
    for (std::uint32_t k = 0; k < n_chunks; ++k) {
        float a = 0.0f, b = 0.0f, t32 = 0.0f, t64 = 0.0f;
        hipEventElapsedTime(&a, ev[8 * k + 0], ev[8 * k + 1]);
        hipEventElapsedTime(&b, ev[8 * k + 3], ev[8 * k + 2]);
        hipEventElapsedTime(&t32, ev[8 * k + 4], ev[8 * k + 5]);
        hipEventElapsedTime(&t64, ev[8 * k + 6], ev[8 * k + 7]);
        // the mixed form's list pass sits inside [start, front done] unless it
        // ran on the second stream
        trace_ms += a + ((f.overlap && f.precision == SING_PREC_MIXED) ? t64 : 0.0f);
        shade_ms += b;
        t32_ms += t32;
        t64_ms += t64;
    }
    float total = 0.0f;
    hipEventElapsedTime(&total, ev[0], ev[8 * n_chunks + 2]);
    if (f.post == SING_POST_BLOOM) {
        float p = 0.0f;
        hipEventElapsedTime(&p, ev[8 * n_chunks], ev[8 * n_chunks + 1]);
        post_ms = p;
    }
    SingRenderStats& st = last_stats;
    st.kernel_ms = total;
    st.trace_ms = trace_ms;
    st.shade_ms = shade_ms;
    st.post_ms = post_ms;
    st.rays = ctl.rays;
    st.steps = ctl.steps;
    st.fp64_rays = ctl.rays64;
    st.chunks = n_chunks;
    st.rejected = ctl.rejected;
    st.trace32_ms = t32_ms;
    st.trace64_ms = t64_ms;
    st.steps32 = ctl.steps32;
    st.steps64 = ctl.steps64;
    if (ctl.rays != f.expect_rays)
        return last_rc = fail(SING_ERR_STATE, "sing_render: traced %llu rays, expected %llu",
                              (unsigned long long)ctl.rays, (unsigned long long)f.expect_rays);
    outputs_valid = f.outs;
    return last_rc = SING_OK;
}

int sing_wait(SingContext* c, SingRenderStats* stats) {
    if (!c) return SING_ERR_ARG;
    if (!c->pending && !c->stats_unclaimed) return c->fail(SING_ERR_STATE, "sing_wait: no frame was submitted");
    const int rc = c->pending ? c->complete_pending() : c->last_rc;
    c->stats_unclaimed = false;
    if (stats) *stats = c->last_stats;
    return rc;
}

int sing_render(SingContext* c, const SingUniforms* u, const SingRenderOpts* opts, SingRenderStats* stats) {
    const int rc = sing_render_async(c, u, opts);
    if (rc != SING_OK) return rc;
    return sing_wait(c, stats);
}

static int capture(SingContext* c, std::uint32_t bit, const void* src, void* dst, std::size_t need, std::size_t bytes,
                   const char* what) {
    if (!c) return SING_ERR_ARG;
    // a pending frame that fails reports its own error here (and to sing_wait)
    if (const int rc = c->complete_pending(); rc != SING_OK) return rc;
    if (!dst) return c->fail(SING_ERR_ARG, "capture %s: null destination", what);
    if (bytes < need) return c->fail(SING_ERR_ARG, "capture %s: buffer %zu bytes < %zu", what, bytes, need);
    if (!(c->outputs_valid & bit)) return c->fail(SING_ERR_STATE, "capture %s: not rendered in the last frame", what);
    hipSetDevice(c->device);
    const hipError_t e = hipMemcpy(dst, src, need, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return c->hip_fail(e, "hipMemcpy D->H");
    return SING_OK;
}

int sing_capture_rgba8(SingContext* c, uint8_t* dst, size_t bytes) {
    return c ? capture(c, SING_OUT_RGBA8, c->d_rgba, dst, std::size_t(c->width) * c->height * 4, bytes, "rgba8")
             : SING_ERR_ARG;
}
int sing_capture_hdr(SingContext* c, float* dst, size_t bytes) {
    return c ? capture(c, SING_OUT_HDR, c->d_hdr, dst, std::size_t(c->width) * c->height * 12, bytes, "hdr")
             : SING_ERR_ARG;
}
int sing_capture_class(SingContext* c, uint8_t* dst, size_t bytes) {
    return c ? capture(c, SING_OUT_CLASS, c->d_cls, dst, std::size_t(c->width) * c->height, bytes, "class")
             : SING_ERR_ARG;
}
int sing_capture_steps(SingContext* c, uint32_t* dst, size_t bytes) {
    return c ? capture(c, SING_OUT_STEPS, c->d_steps, dst, std::size_t(c->width) * c->height * 4, bytes, "steps")
             : SING_ERR_ARG;
}
int sing_capture_hdr16(SingContext* c, uint16_t* dst, size_t bytes) {
    return c ? capture(c, SING_OUT_HDR16, c->d_hdr16, dst, std::size_t(c->width) * c->height * 8, bytes, "hdr16")
             : SING_ERR_ARG;
}

int sing_post_bloom(SingContext* c, const SingUniforms* u, uint32_t repeat, SingRenderStats* stats) {
    if (!c) return SING_ERR_ARG;
    if (const int rc = c->complete_pending(); rc != SING_OK) return rc;
    if (!u) return c->fail(SING_ERR_ARG, "sing_post_bloom: null uniforms");
    if (repeat < 1u || repeat > 1000u) return c->fail(SING_ERR_ARG, "sing_post_bloom: repeat must be in [1, 1000]");
    if (c->width < 2 || c->height < 2) return c->fail(SING_ERR_ARG, "sing_post_bloom: bloom needs at least 2x2 pixels");
    if (!(c->outputs_valid & SING_OUT_HDR16))
        return c->fail(SING_ERR_STATE, "sing_post_bloom: the last frame has no half-float HDR (SING_OUT_HDR16 or bloom)");
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return c->hip_fail(e, "hipSetDevice");
    if (const int rc = c->ensure_events(2); rc != SING_OK) return rc;
    hipEvent_t p0 = c->events[0], p1 = c->events[1];
    hipEventRecord(p0, c->stream);
    for (std::uint32_t i = 0; i < repeat; ++i) {
        e = sing::launch_bloom(c->d_hdr16, c->width, c->height, u->bloom_threshold, u->bloom_strength, u->exposure,
                               c->d_rgba, c->stream);
        if (e != hipSuccess) return c->hip_fail(e, "bloom launch");
    }
    hipEventRecord(p1, c->stream);
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return c->hip_fail(e, "hipStreamSynchronize");
    c->outputs_valid |= SING_OUT_RGBA8;
    if (stats) {
        float ms = 0.0f;
        hipEventElapsedTime(&ms, p0, p1);
        std::memset(stats, 0, sizeof(*stats));
        stats->kernel_ms = ms;
        stats->post_ms = double(ms) / repeat;
    }
    return SING_OK;
}

}  // extern "C"
