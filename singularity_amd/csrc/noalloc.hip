// singularity_amd/csrc/noalloc.hip — sing_render_noalloc: the frame pipeline of
// sing_render (trace -> [FP64 pass] -> shade per sample chunk, optional bloom)
// for callers that cannot allocate device memory: code under HIP graph
// capture, serving loops, harnesses that forbid hipMalloc.  The caller passes
// the uniform block and device output buffers; every working buffer (queue
// counters, FP64 scene constants, sample-offset tables, per-pixel
// accumulators, ray records, the FP64 list, the set-up arrays, the bloom
// pass's half-float HDR input) is carved from `scratch`, one file-scope device array of
// kScratchBytes per GPU that lives until the process ends.
//
// Host code never passes `scratch` itself to a kernel or copy: in host code the
// name is a host-side stand-in.  The first call on a device looks its device
// address up with hipGetSymbolAddress (per device: each GPU holds its own copy
// at its own address) and uploads the sample-offset tables; every later call
// on that device uses the stored address.
//
// All calls on one GPU share its array: calls that overlap in time on the same
// GPU (two host threads, or two streams) race and corrupt each other's frames.
// Each call clears what must start at zero (queue counters, first-chunk
// accumulators) itself, so it does not depend on what the last call left.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "../../include/sing_abi.h"
#include "context.hpp"
#include "post.hpp"
#include "scene_const.hpp"
#include "trace.hpp"

namespace {

constexpr std::size_t kScratchBytes = std::size_t(256) << 20;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr std::size_t round256(std::size_t n) { return (n + 255u) & ~std::size_t(255); }

// Fixed head of the array: queue counters, the FP64 scene constants, then the
// Halton table (1024 float2) followed by the grid sampler's tables.
constexpr std::size_t kCtlOff = 0;
constexpr std::size_t kSdOff = round256(sizeof(sing::QueueCtl));
constexpr std::size_t kTabOff = kSdOff + round256(sizeof(sing::SceneConst<double>));
constexpr std::size_t kTabLen = 1024u + sing::kGridTableLen;
constexpr std::size_t kHeadBytes = kTabOff + round256(kTabLen * sizeof(float2));

struct DeviceScratch {
    unsigned char* ptr = nullptr;  // device address of this GPU's `scratch`
    int num_cu = 0;
};
std::mutex g_setup_mu;
DeviceScratch g_dev[kMaxDevices];

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...) {
    char buf[384];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    sing_set_global_error(buf);
    return code;
}

// Once per device, with that device current: the array's device address, the
// CU count and the sample-offset tables (the same host-built values
// sing_create uploads).
int setup_device(int device, DeviceScratch& out) {
    std::lock_guard<std::mutex> lk(g_setup_mu);
    DeviceScratch& d = g_dev[device];
    if (!d.ptr) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(SING_ERR_HIP, "sing_render_noalloc: hipGetSymbolAddress(scratch) failed on device %d: %s",
                        device, hipGetErrorString(e));
        int cus = 0;
        e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
        if (e != hipSuccess)
            return fail(SING_ERR_HIP, "sing_render_noalloc: hipDeviceGetAttribute failed: %s", hipGetErrorString(e));
        float2 tab[kTabLen];
        for (std::uint32_t s = 0; s < 1024u; ++s)
            tab[s] = make_float2(sing::halton_f(s + 1u, 2u), sing::halton_f(s + 1u, 3u));
        for (std::uint32_t ss = 1; ss <= 4u; ++ss)
            for (std::uint32_t s = 0; s < ss * ss; ++s) {
                float ox, oy;
                sing::grid_offset_f(s, ss, ox, oy);
                tab[1024u + sing::grid_table_base(ss) + s] = make_float2(ox, oy);
            }
        e = hipMemcpy(static_cast<unsigned char*>(p) + kTabOff, tab, sizeof(tab), hipMemcpyHostToDevice);
        if (e != hipSuccess)
            return fail(SING_ERR_HIP, "sing_render_noalloc: sample table upload failed: %s", hipGetErrorString(e));
        d.num_cu = cus;
        d.ptr = static_cast<unsigned char*>(p);
    }
    out = d;
    return SING_OK;
}

// Writes the frame's FP64 scene constants into the array (a kernel argument is
// copied at launch, so the caller's host memory is free at once) and clears
// the queue counters: all of them (frame start) or only the work queues
// (between chunks, the frame totals keep accumulating).
__global__ void noalloc_reset_kernel(sing::SceneConst<double> sd, sing::SceneConst<double>* __restrict__ sd_dst,
                                     std::uint64_t* __restrict__ ctl, std::uint32_t words, std::uint32_t write_sd) {
    const std::uint32_t t = threadIdx.x;
    if (t < words) ctl[t] = 0ull;
    if (write_sd && t == 0) *sd_dst = sd;
}

}  // namespace

extern "C" int sing_render_noalloc(int device, const SingUniforms* u, const SingRenderOpts* opts_in,
                                   const SingDeviceOutputs* out, void* stream_in, SingRenderStats* stats) {
    if (!u || !out) return fail(SING_ERR_ARG, "sing_render_noalloc: null uniforms or outputs");
    SingRenderOpts o{};
    if (opts_in) o = *opts_in;
    char msg[160];
    if (!sing::render_args_ok(*u, o, "sing_render_noalloc", msg, sizeof(msg))) return fail(SING_ERR_ARG, "%s", msg);
    if (u->width < 1 || u->height < 1 || u->width > 32768 || u->height > 32768)
        return fail(SING_ERR_ARG, "sing_render_noalloc: width/height must be in [1, 32768]");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(SING_ERR_HIP, "sing_render_noalloc: no HIP device available");
    if (device < 0 || device >= ndev || device >= kMaxDevices)
        return fail(SING_ERR_ARG, "sing_render_noalloc: device index %d out of range", device);
    std::uint32_t outs = o.outputs ? o.outputs : SING_OUT_RGBA8;
    if (o.post == SING_POST_BLOOM) outs |= SING_OUT_RGBA8;
    if (((outs & SING_OUT_RGBA8) && !out->rgba) || ((outs & SING_OUT_HDR) && !out->hdr) ||
        ((outs & SING_OUT_HDR16) && !out->hdr16) || ((outs & SING_OUT_CLASS) && !out->cls) ||
        ((outs & SING_OUT_STEPS) && !out->steps))
        return fail(SING_ERR_ARG, "sing_render_noalloc: an output selected in opts->outputs has no buffer");
    auto misaligned = [](const void* p, std::uintptr_t a) { return (reinterpret_cast<std::uintptr_t>(p) & (a - 1u)) != 0; };
    if (misaligned(out->rgba, 4) || misaligned(out->hdr, 4) || misaligned(out->hdr16, 16) ||
        misaligned(out->steps, 4))
        return fail(SING_ERR_ARG, "sing_render_noalloc: output buffers must be aligned to rgba 4, hdr 4, hdr16 16, "
                                  "steps 4 bytes");

    const sing::SceneConst<float> sf = sing::make_scene_const<float>(*u, o);
    const sing::SceneConst<double> sd = sing::make_scene_const<double>(*u, o);
    const std::uint32_t W = u->width, H = u->height, spp = sf.spp;
    const std::size_t P = std::size_t(W) * H;
    const bool bloom = o.post == SING_POST_BLOOM;
    const bool hdr16_in_scratch = bloom && !(outs & SING_OUT_HDR16);

    sing::Chunk ch{};
    ch.P = std::uint32_t(P);
    ch.tiles_x = (W + 7u) / 8u;
    ch.n_tiles = ch.tiles_x * ((H + 7u) / 8u);
    const std::size_t items_per_sample = std::size_t(ch.n_tiles) * 64u;

    // layout: head | accumulators | [hdr16] | records | [FP64 list] | [set-up arrays]
    std::size_t fixed = kHeadBytes + round256(P * sizeof(sing::PixAcc));
    if (hdr16_in_scratch) fixed += round256(P * 8u);
    const bool mixed = o.precision == SING_PREC_MIXED, fp32_items = o.precision != SING_PREC_FP64;
    const std::size_t per_sample = P * sizeof(sing::Rec) + (mixed ? P * 4u : 0u) +
                                   (fp32_items ? items_per_sample * (sizeof(float4) + 4u) : 0u);
    const std::size_t slack = 4u * 256u;  // rounding of the per-chunk arrays
    if (fixed + slack + per_sample > kScratchBytes)
        return fail(SING_ERR_ARG,
                    "sing_render_noalloc: a %ux%u frame needs %zu MiB of scratch for one sample per chunk; the "
                    "no-alloc entry point has %zu MiB (use sing_render)",
                    W, H, (fixed + slack + per_sample + (std::size_t(1) << 20) - 1) >> 20, kScratchBytes >> 20);
    std::size_t S = (kScratchBytes - fixed - slack) / per_sample;
    S = std::min<std::size_t>(S, spp);
    S = std::min<std::size_t>(S, (std::size_t(1) << 31) / P);  // record indices stay in u32
    if (o.chunk_samples) S = std::min<std::size_t>(S, o.chunk_samples);
    S = std::max<std::size_t>(S, 1);

    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return fail(SING_ERR_HIP, "sing_render_noalloc: hipSetDevice: %s", hipGetErrorString(e));
    DeviceScratch ds;
    int rc = setup_device(device, ds);
    if (rc != SING_OK) return rc;
    const hipStream_t stream = static_cast<hipStream_t>(stream_in);

    unsigned char* base = ds.ptr;
    std::size_t off = kHeadBytes;
    auto take = [&](std::size_t bytes) {
        unsigned char* p = base + off;
        off += round256(bytes);
        return p;
    };
    sing::PixAcc* acc = reinterpret_cast<sing::PixAcc*>(take(P * sizeof(sing::PixAcc)));
    uint2* hdr16 = reinterpret_cast<uint2*>(out->hdr16);
    if (hdr16_in_scratch) hdr16 = reinterpret_cast<uint2*>(take(P * 8u));
    ch.rec = reinterpret_cast<sing::Rec*>(take(S * P * sizeof(sing::Rec)));
    ch.band = mixed ? reinterpret_cast<std::uint32_t*>(take(S * P * 4u)) : nullptr;
    ch.init = fp32_items ? reinterpret_cast<float4*>(take(S * items_per_sample * sizeof(float4))) : nullptr;
    ch.item_rec = fp32_items ? reinterpret_cast<std::uint32_t*>(take(S * items_per_sample * 4u)) : nullptr;
    if (off > kScratchBytes)
        return fail(SING_ERR_STATE, "sing_render_noalloc: scratch layout overflow (%zu > %zu bytes)", off,
                    kScratchBytes);
    sing::QueueCtl* ctl = reinterpret_cast<sing::QueueCtl*>(base + kCtlOff);
    sing::SceneConst<double>* d_sd = reinterpret_cast<sing::SceneConst<double>*>(base + kSdOff);
    const float2* table = reinterpret_cast<const float2*>(base + kTabOff);
    const float2* offsets = table;
    if (sf.grid) {
        std::uint32_t ss = 1;
        while (ss * ss < spp) ++ss;
        offsets = table + 1024 + sing::grid_table_base(ss);
    }
    ch.sd = d_sd;

    sing::TraceOutputs to{};
    to.rgba = (outs & SING_OUT_RGBA8) && !bloom ? reinterpret_cast<uchar4*>(out->rgba) : nullptr;
    to.hdr = (outs & SING_OUT_HDR) ? out->hdr : nullptr;
    to.hdr16 = (outs & SING_OUT_HDR16) || bloom ? hdr16 : nullptr;
    to.cls = (outs & SING_OUT_CLASS) ? out->cls : nullptr;
    to.steps = (outs & SING_OUT_STEPS) ? out->steps : nullptr;
    const sing::LaunchShape shape = sing::launch_shape_for(ds.num_cu, o);

    constexpr std::uint32_t kCtlWords = sizeof(sing::QueueCtl) / 8u;
    constexpr std::uint32_t kQueueWords = offsetof(sing::QueueCtl, steps) / 8u;
    static_assert(sizeof(sing::QueueCtl) % 8u == 0 && kCtlWords <= 64u, "queue counters are cleared by one wave");
    hipLaunchKernelGGL(noalloc_reset_kernel, dim3(1), dim3(64), 0, stream, sd, d_sd,
                       reinterpret_cast<std::uint64_t*>(ctl), kCtlWords, 1u);
    if ((e = hipGetLastError()) != hipSuccess)
        return fail(SING_ERR_HIP, "sing_render_noalloc: reset launch: %s", hipGetErrorString(e));
    const std::uint32_t n_chunks = std::uint32_t((spp + S - 1u) / S);
    for (std::uint32_t k = 0; k < n_chunks; ++k) {
        ch.s0 = k * std::uint32_t(S);
        ch.S = std::min(std::uint32_t(S), spp - ch.s0);
        if (k > 0) {
            hipLaunchKernelGGL(noalloc_reset_kernel, dim3(1), dim3(64), 0, stream, sd, d_sd,
                               reinterpret_cast<std::uint64_t*>(ctl), kQueueWords, 0u);
            if ((e = hipGetLastError()) != hipSuccess)
                return fail(SING_ERR_HIP, "sing_render_noalloc: reset launch: %s", hipGetErrorString(e));
        }
        if ((e = sing::launch_trace_chunk(sf, sd, offsets, ch, ctl, o.precision, shape, stream)) != hipSuccess)
            return fail(SING_ERR_HIP, "sing_render_noalloc: trace launch: %s", hipGetErrorString(e));
        if ((e = sing::launch_shade_chunk(sf, ch, acc, k == 0, k + 1 == n_chunks, to, stream)) != hipSuccess)
            return fail(SING_ERR_HIP, "sing_render_noalloc: shade launch: %s", hipGetErrorString(e));
    }
    if (bloom) {
        e = sing::launch_bloom(hdr16, W, H, u->bloom_threshold, u->bloom_strength, u->exposure,
                               reinterpret_cast<uchar4*>(out->rgba), stream);
        if (e != hipSuccess) return fail(SING_ERR_HIP, "sing_render_noalloc: bloom launch: %s", hipGetErrorString(e));
    }
    sing::QueueCtl hc{};
    if ((e = hipMemcpyAsync(&hc, ctl, sizeof(hc), hipMemcpyDeviceToHost, stream)) != hipSuccess ||
        (e = hipStreamSynchronize(stream)) != hipSuccess)
        return fail(SING_ERR_HIP, "sing_render_noalloc: %s", hipGetErrorString(e));
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->rays = hc.rays;
        stats->steps = hc.steps;
        stats->fp64_rays = hc.rays64;
        stats->chunks = n_chunks;
        stats->rejected = hc.rejected;
        stats->steps32 = hc.steps32;
        stats->steps64 = hc.steps64;
    }
    const std::uint64_t expect = std::uint64_t(P) * spp;
    if (hc.rays != expect)
        return fail(SING_ERR_STATE, "sing_render_noalloc: traced %llu rays, expected %llu",
                    (unsigned long long)hc.rays, (unsigned long long)expect);
    return SING_OK;
}
