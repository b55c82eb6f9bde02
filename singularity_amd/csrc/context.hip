// singularity_amd/csrc/context.hip — the C ABI's render context: device
// buffers, the launch of the trace and post-pass kernels, stats and capture.
// Replaces CudaBackend (render/cuda/cuda_backend.cu:105-166): the context owns
// every device allocation, render is synchronous, capture copies to caller
// memory, errors are returned as status codes with a per-context message.

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/sing_abi.h"
#include "post.hpp"
#include "scene_const.hpp"
#include "trace.hpp"

namespace {

std::mutex g_err_mu;
std::string g_last_error;

void set_global_error(const std::string& m) {
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_last_error = m;
}

}  // namespace

struct SingContext {
    int device = 0;
    int num_cu = 256;
    std::uint32_t width = 0, height = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    uchar4* d_rgba = nullptr;
    float* d_hdr = nullptr;
    uint2* d_hdr16 = nullptr;
    std::uint8_t* d_cls = nullptr;
    std::uint32_t* d_steps = nullptr;
    float* d_bloom_a = nullptr;  // half-res ping-pong (RGBA16F stored as uint2)
    float* d_bloom_b = nullptr;
    sing::QueueCtl* d_ctl = nullptr;
    float2* d_halton = nullptr;
    std::uint32_t outputs_valid = 0;  // SING_OUT_* of the last frame
    std::string err;

    int fail(int code, const char* fmt, ...) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        std::vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        err = buf;
        set_global_error(err);
        return code;
    }
    int hip_fail(hipError_t e, const char* where) {
        return fail(SING_ERR_HIP, "%s failed: %s", where, hipGetErrorString(e));
    }
    void free_frame() {
        hipFree(d_rgba); hipFree(d_hdr); hipFree(d_hdr16); hipFree(d_cls); hipFree(d_steps);
        hipFree(d_bloom_a); hipFree(d_bloom_b);
        d_rgba = nullptr; d_hdr = nullptr; d_hdr16 = nullptr; d_cls = nullptr; d_steps = nullptr;
        d_bloom_a = nullptr; d_bloom_b = nullptr;
        outputs_valid = 0;
    }
    int alloc_frame(std::uint32_t w, std::uint32_t h) {
        const std::size_t n = std::size_t(w) * h;
        const std::size_t nb = std::size_t(w / 2 > 0 ? w / 2 : 1) * (h / 2 > 0 ? h / 2 : 1);
        hipError_t e;
        if ((e = hipMalloc(&d_rgba, n * 4)) != hipSuccess ||
            (e = hipMalloc(&d_hdr, n * 12)) != hipSuccess ||
            (e = hipMalloc(&d_hdr16, n * 8)) != hipSuccess ||
            (e = hipMalloc(&d_cls, n)) != hipSuccess ||
            (e = hipMalloc(&d_steps, n * 4)) != hipSuccess ||
            (e = hipMalloc(&d_bloom_a, nb * 8)) != hipSuccess ||
            (e = hipMalloc(&d_bloom_b, nb * 8)) != hipSuccess) {
            free_frame();
            return fail(SING_ERR_NOMEM, "hipMalloc of %ux%u frame buffers failed: %s", w, h, hipGetErrorString(e));
        }
        width = w;
        height = h;
        return SING_OK;
    }
};

extern "C" {

int sing_abi_version(void) { return SING_ABI_VERSION; }
const char* sing_backend_name(void) { return "HIP-gfx950"; }

const char* sing_last_error(const SingContext* ctx) {
    if (ctx) return ctx->err.c_str();
    std::lock_guard<std::mutex> lk(g_err_mu);
    return g_last_error.c_str();
}

SingContext* sing_create(uint32_t width, uint32_t height, int device) {
    if (width < 1 || height < 1 || width > 32768 || height > 32768) {
        set_global_error("sing_create: width/height must be in [1, 32768]");
        return nullptr;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_global_error("sing_create: no HIP device available");
        return nullptr;
    }
    if (device < 0 || device >= ndev) {
        set_global_error("sing_create: device index out of range");
        return nullptr;
    }
    auto* c = new SingContext();
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    for (int i = 0; i < 4 && e == hipSuccess; ++i) e = hipEventCreate(&c->ev[i]);
    if (e == hipSuccess) e = hipMalloc(&c->d_ctl, sizeof(sing::QueueCtl));
    if (e == hipSuccess) e = hipMalloc(&c->d_halton, 1024 * sizeof(float2));
    if (e != hipSuccess) {
        set_global_error(std::string("sing_create: HIP setup failed: ") + hipGetErrorString(e));
        sing_destroy(c);
        return nullptr;
    }
    // Halton(2,3) offsets of samples 1..1024 (geodesic_kernel.cu:491-495).
    float2 tab[1024];
    for (std::uint32_t s = 0; s < 1024; ++s) tab[s] = make_float2(sing::halton_f(s + 1u, 2u), sing::halton_f(s + 1u, 3u));
    e = hipMemcpy(c->d_halton, tab, sizeof(tab), hipMemcpyHostToDevice);
    if (e != hipSuccess || c->alloc_frame(width, height) != SING_OK) {
        set_global_error(e != hipSuccess ? std::string("sing_create: halton upload failed") : c->err);
        sing_destroy(c);
        return nullptr;
    }
    return c;
}

void sing_destroy(SingContext* c) {
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    c->free_frame();
    hipFree(c->d_ctl);
    hipFree(c->d_halton);
    for (auto& e : c->ev)
        if (e) hipEventDestroy(e);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

int sing_resize(SingContext* c, uint32_t w, uint32_t h) {
    if (!c) return SING_ERR_ARG;
    if (w < 1 || h < 1 || w > 32768 || h > 32768) return c->fail(SING_ERR_ARG, "sing_resize: bad size %ux%u", w, h);
    if (w == c->width && h == c->height) return SING_OK;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    c->free_frame();
    return c->alloc_frame(w, h);
}

int sing_render(SingContext* c, const SingUniforms* u, const SingRenderOpts* opts_in, SingRenderStats* stats) {
    if (!c) return SING_ERR_ARG;
    if (!u) return c->fail(SING_ERR_ARG, "sing_render: null uniforms");
    if (u->width != c->width || u->height != c->height)
        return c->fail(SING_ERR_ARG, "sing_render: uniforms size %ux%u != context %ux%u", u->width, u->height,
                       c->width, c->height);
    SingRenderOpts o{};
    if (opts_in) o = *opts_in;
    if (o.precision > SING_PREC_MIXED) return c->fail(SING_ERR_ARG, "sing_render: unknown precision %u", o.precision);
    if (o.post > SING_POST_BLOOM) return c->fail(SING_ERR_ARG, "sing_render: unknown post pass %u", o.post);
    if (u->metric_type > SING_METRIC_KERR) return c->fail(SING_ERR_ARG, "sing_render: unknown metric %u", u->metric_type);
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

    sing::LaunchShape shape{};
    shape.precision = o.precision;
    shape.step_batch = o.step_batch ? o.step_batch : 8u;
    const std::uint32_t waves_per_cu = o.waves_per_cu ? o.waves_per_cu : 16u;
    shape.blocks = std::uint32_t(c->num_cu) * ((waves_per_cu + 3u) / 4u);

    if ((e = hipMemsetAsync(c->d_ctl, 0, sizeof(sing::QueueCtl), c->stream)) != hipSuccess)
        return c->hip_fail(e, "hipMemsetAsync(queue)");
    hipEventRecord(c->ev[0], c->stream);
    if ((e = sing::launch_trace(sf, sd, c->d_halton, to, c->d_ctl, shape, c->stream)) != hipSuccess)
        return c->hip_fail(e, "trace kernel launch");
    hipEventRecord(c->ev[1], c->stream);
    if (o.post == SING_POST_BLOOM) {
        hipEventRecord(c->ev[2], c->stream);
        e = sing::launch_bloom(c->d_hdr16, c->width, c->height, u->bloom_threshold, u->bloom_strength, u->exposure,
                               reinterpret_cast<uint2*>(c->d_bloom_a), reinterpret_cast<uint2*>(c->d_bloom_b),
                               c->d_rgba, c->stream);
        if (e != hipSuccess) return c->hip_fail(e, "bloom launch");
        hipEventRecord(c->ev[3], c->stream);
    }
    sing::QueueCtl ctl{};
    if ((e = hipMemcpyAsync(&ctl, c->d_ctl, sizeof(ctl), hipMemcpyDeviceToHost, c->stream)) != hipSuccess)
        return c->hip_fail(e, "hipMemcpyAsync(stats)");
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return c->hip_fail(e, "hipStreamSynchronize");
    c->outputs_valid = outs;
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        float ms = 0.0f;
        hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
        stats->kernel_ms = ms;
        if (o.post == SING_POST_BLOOM) {
            hipEventElapsedTime(&ms, c->ev[2], c->ev[3]);
            stats->post_ms = ms;
        }
        stats->rays = ctl.rays;
        stats->steps = ctl.steps;
        stats->fp64_rays = ctl.rays64;
    }
    const std::uint64_t expect = std::uint64_t(c->width) * c->height * sf.spp;
    if (ctl.rays != expect)
        return c->fail(SING_ERR_STATE, "sing_render: traced %llu rays, expected %llu",
                       (unsigned long long)ctl.rays, (unsigned long long)expect);
    return SING_OK;
}

static int capture(SingContext* c, std::uint32_t bit, const void* src, void* dst, std::size_t need, std::size_t bytes,
                   const char* what) {
    if (!c) return SING_ERR_ARG;
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

}  // extern "C"
