"""sing_render_noalloc (the render entry point that allocates no device memory)
against sing_render: the same frames, bit for bit, on two calls in a row (the
second smaller), through the C ABI with caller-owned device buffers.

The caller's buffers come from hipMalloc of the HIP runtime the library itself
links (found in the process's mappings), not from a second runtime such as the
one bundled with a Python framework: two HIP runtimes in one process do not
share a device context."""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import singularity_amd as sa
from helpers import scene_uniforms

pytestmark = pytest.mark.gpu


def _hip() -> C.CDLL:
    sa.lib()  # loads the library and, with it, its HIP runtime
    with open("/proc/self/maps") as f:
        paths = {ln.split()[-1] for ln in f if "libamdhip64.so" in ln}
    linked = [p for p in paths if "/torch/" not in p]
    assert linked, paths
    h = C.CDLL(sorted(linked)[0])
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMalloc.restype = C.c_int
    h.hipFree.argtypes = [C.c_void_p]
    h.hipFree.restype = C.c_int
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    h.hipMemset.restype = C.c_int
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemcpy.restype = C.c_int
    h.hipDeviceSynchronize.restype = C.c_int
    return h


SHAPES = {"rgba": (4, np.uint8), "hdr": (3, np.float32), "hdr16": (4, np.uint16), "cls": (1, np.uint8),
          "steps": (1, np.uint32)}
BITS = {"rgba": sa.OUT_RGBA8, "hdr": sa.OUT_HDR, "hdr16": sa.OUT_HDR16, "cls": sa.OUT_CLASS, "steps": sa.OUT_STEPS}


class DeviceBuffers:
    """Caller-owned device outputs (hipMalloc here, in the test: the entry point allocates nothing)."""

    def __init__(self, W, H, outputs):
        self.h = _hip()
        self.W, self.H = W, H
        self.ptrs = {}
        self.d = sa.DeviceOutputs()
        for k, (ch, dt) in SHAPES.items():
            if k != "rgba" and not outputs & BITS[k]:
                continue
            n = W * H * ch * np.dtype(dt).itemsize
            p = C.c_void_p()
            assert self.h.hipMalloc(C.byref(p), n) == 0
            assert self.h.hipMemset(p, 0, n) == 0
            self.ptrs[k] = (p, n, ch, dt)
            setattr(self.d, k, p.value)

    def read(self):
        assert self.h.hipDeviceSynchronize() == 0
        out = {k: None for k in SHAPES}
        for k, (p, n, ch, dt) in self.ptrs.items():
            a = np.empty(n // np.dtype(dt).itemsize, dt)
            assert self.h.hipMemcpy(a.ctypes.data, p, n, 2) == 0  # hipMemcpyDeviceToHost
            out[k] = a
        return out

    def free(self):
        for p, *_ in self.ptrs.values():
            self.h.hipFree(p)
        self.ptrs = {}


def _noalloc(u, opts):
    bufs = DeviceBuffers(u.width, u.height, opts.outputs)
    try:
        st = sa.RenderStats()
        rc = sa.lib().sing_render_noalloc(0, C.byref(u), C.byref(opts), C.byref(bufs.d), None, C.byref(st))
        assert rc == 0, sa.lib().sing_last_error(None).decode()
        return bufs.read(), st
    finally:
        bufs.free()


def _reference(u, opts):
    be = sa.HipBackend(u.width, u.height)
    st = be.render(u, opts)
    f = be.capture(opts.outputs | sa.OUT_RGBA8)
    be.shutdown()
    return f, st


def _compare(host, f, st_a, st_b, outputs):
    assert np.array_equal(host["rgba"].reshape(f.rgba.shape), f.rgba)
    if outputs & sa.OUT_HDR:
        assert np.array_equal(host["hdr"].reshape(f.hdr.shape), f.hdr)
    if outputs & sa.OUT_HDR16:
        assert np.array_equal(host["hdr16"], f.hdr16.view(np.uint16).reshape(-1))
    if outputs & sa.OUT_CLASS:
        assert np.array_equal(host["cls"].reshape(f.cls.shape), f.cls)
    if outputs & sa.OUT_STEPS:
        assert np.array_equal(host["steps"].reshape(f.steps.shape), f.steps)
    assert (st_a.rays, st_a.steps, st_a.fp64_rays) == (st_b.rays, st_b.steps, st_b.fp64_rays)


def test_noalloc_equals_sing_render_on_two_calls():
    # call 1: 1080p, 4 spp, mixed precision: the scratch array holds two samples per
    # chunk here, so this also checks that chunking leaves the frame unchanged
    outs1 = sa.OUT_RGBA8 | sa.OUT_HDR | sa.OUT_HDR16 | sa.OUT_CLASS | sa.OUT_STEPS
    u1 = scene_uniforms(metric=sa.METRIC_KERR, spin=0.94, r_in=1.7, r_out=30.0, width=1920, height=1080, spp=4)
    o1 = sa.RenderOpts(precision=sa.PREC_MIXED, outputs=outs1)
    print("noalloc call 1: Kerr a=0.94 1920x1080 4 spp mixed", flush=True)
    h1, s1 = _noalloc(u1, o1)
    assert s1.chunks >= 2
    f1, r1 = _reference(u1, o1)
    _compare(h1, f1, s1, r1, outs1)
    # call 2, smaller: 200x120, 3 spp, FP32 with the bloom post pass (its HDR16 input
    # and ping-pong buffers come from the scratch array)
    outs2 = sa.OUT_RGBA8 | sa.OUT_CLASS
    u2 = scene_uniforms(metric=sa.METRIC_SCHWARZSCHILD, spin=0.0, r_in=6.0, r_out=20.0, width=200, height=120,
                        spp=3)
    o2 = sa.RenderOpts(precision=sa.PREC_FP32, outputs=outs2, post=sa.POST_BLOOM)
    print("noalloc call 2: Schwarzschild 200x120 3 spp fp32 + bloom", flush=True)
    h2, s2 = _noalloc(u2, o2)
    f2, r2 = _reference(u2, o2)
    _compare(h2, f2, s2, r2, outs2)
    # call 3, smaller again: FP64 with a grid sampler
    outs3 = sa.OUT_RGBA8 | sa.OUT_HDR
    u3 = scene_uniforms(metric=sa.METRIC_KERR, spin=0.5, width=64, height=48, spp=4)
    o3 = sa.RenderOpts(precision=sa.PREC_FP64, outputs=outs3, sampler=sa.SAMPLER_GRID)
    print("noalloc call 3: Kerr a=0.5 64x48 grid 2x2 fp64", flush=True)
    h3, s3 = _noalloc(u3, o3)
    f3, r3 = _reference(u3, o3)
    _compare(h3, f3, s3, r3, outs3)
    # call 4: mixed with both fixed-step refinements (their own trace kernels)
    outs4 = sa.OUT_RGBA8 | sa.OUT_HDR | sa.OUT_CLASS
    u4 = scene_uniforms(metric=sa.METRIC_KERR, spin=0.9, r_in=2.3, width=48, height=32, spp=2)
    o4 = sa.RenderOpts(precision=sa.PREC_MIXED, outputs=outs4, refine=sa.REFINE_CROSS64 | sa.REFINE_REPROJECT)
    print("noalloc call 4: Kerr a=0.9 48x32 2 spp mixed, refine cross64+reproject", flush=True)
    h4, s4 = _noalloc(u4, o4)
    f4, r4 = _reference(u4, o4)
    _compare(h4, f4, s4, r4, outs4)


def test_noalloc_rejects_frames_beyond_its_scratch():
    L = sa.lib()
    u = scene_uniforms(width=3840, height=2160, spp=1)
    o = sa.RenderOpts(precision=sa.PREC_MIXED)
    bufs = DeviceBuffers(3840, 2160, sa.OUT_RGBA8)
    try:
        rc = L.sing_render_noalloc(0, C.byref(u), C.byref(o), C.byref(bufs.d), None, None)
        assert rc == 1
        assert b"scratch" in L.sing_last_error(None)
        d = sa.DeviceOutputs()  # a selected output without a buffer
        assert L.sing_render_noalloc(0, C.byref(scene_uniforms(width=64, height=64)), C.byref(o), C.byref(d), None,
                                     None) == 1
    finally:
        bufs.free()
