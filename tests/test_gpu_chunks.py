"""Frames of several sample chunks: the mixed form's overlapped schedule
(SingRenderOpts.overlap_chunks = 1: a chunk's FP64 list pass and shade on a
second stream beside the next chunk's set-up and FP32 pass, chunk buffers used
by parity) against the default serial one, bit for bit; FP32 and FP64 frames
(always serial) give the same bits either way.

A chunk holds at most 2 GiB of 16-byte ray records, so 1024x768 at 400 spp is
three chunks (170, 170 and 60 samples): the third reuses the first one's
buffers and has to wait for its shade."""

from __future__ import annotations

import numpy as np
import pytest

import singularity_amd as sa
from helpers import scene_uniforms

pytestmark = pytest.mark.gpu

OUTS = sa.OUT_RGBA8 | sa.OUT_HDR | sa.OUT_CLASS | sa.OUT_STEPS | sa.OUT_HDR16


@pytest.fixture(scope="module")
def be():
    b = sa.HipBackend(1024, 768)
    yield b
    b.shutdown()


CASES = {
    "fp32": dict(precision=sa.PREC_FP32),
    "fp64": dict(precision=sa.PREC_FP64),
    "mixed": dict(precision=sa.PREC_MIXED),
    "mixed_adaptive_bloom": dict(precision=sa.PREC_MIXED, adapt_tol=1e-6, post=sa.POST_BLOOM, turb_seed=7),
    "mixed_cross64": dict(precision=sa.PREC_MIXED, refine=sa.REFINE_CROSS64),
}


@pytest.mark.parametrize("name", list(CASES))
def test_overlapped_chunks_equal_serial_chunks(be, name):
    u = scene_uniforms(metric=sa.METRIC_KERR, spin=0.94, r_in=1.7, r_out=30.0, width=1024, height=768, spp=400)
    kw = dict(CASES[name])
    prec = kw.pop("precision")
    print(f"chunks {name}: Kerr a=0.94 1024x768 400 spp, overlapped", flush=True)
    a = be.render_outputs(u, OUTS, prec, overlap_chunks=1, **kw)
    print(f"chunks {name}: serial", flush=True)
    b = be.render_outputs(u, OUTS, prec, **kw)
    assert a.stats.chunks == 3 and b.stats.chunks == 3
    n = 1024 * 768 * 400
    assert a.stats.rays == b.stats.rays == n
    assert (a.stats.steps, a.stats.fp64_rays, a.stats.steps32, a.stats.steps64, a.stats.rejected) == (
        b.stats.steps, b.stats.fp64_rays, b.stats.steps32, b.stats.steps64, b.stats.rejected)
    assert np.array_equal(a.rgba, b.rgba)
    assert np.array_equal(a.hdr.view(np.uint32), b.hdr.view(np.uint32))
    assert np.array_equal(a.hdr16.view(np.uint16), b.hdr16.view(np.uint16))
    assert np.array_equal(a.cls, b.cls)
    assert np.array_equal(a.steps, b.steps)
    # the per-pixel step totals add up to the frame's
    assert int(a.steps.astype(np.uint64).sum()) == a.stats.steps
    assert a.stats.kernel_ms > 0.0 and a.stats.trace32_ms + a.stats.trace64_ms > 0.0


@pytest.mark.parametrize("prec", [sa.PREC_FP32, sa.PREC_MIXED])
def test_chunk_samples_option_does_not_change_the_frame(prec):
    """SingRenderOpts.chunk_samples caps the samples per chunk (less record memory,
    more chunks): 40 spp in chunks of 7 (six chunks, the last of 5) equals the
    one-chunk frame bit for bit."""
    u = scene_uniforms(metric=sa.METRIC_KERR, spin=0.94, r_in=1.7, r_out=30.0, width=256, height=192, spp=40)
    b = sa.HipBackend(256, 192)
    try:
        print(f"chunk_samples 7 vs default, prec={prec}", flush=True)
        a = b.render_outputs(u, OUTS, prec)
        c = b.render_outputs(u, OUTS, prec, chunk_samples=7)
        d = b.render_outputs(u, OUTS, prec, chunk_samples=7, overlap_chunks=1)
        assert np.array_equal(d.rgba, c.rgba) and np.array_equal(d.hdr.view(np.uint32), c.hdr.view(np.uint32))
        assert a.stats.chunks == 1 and c.stats.chunks == 6
        assert (a.stats.rays, a.stats.steps, a.stats.fp64_rays) == (c.stats.rays, c.stats.steps, c.stats.fp64_rays)
        assert np.array_equal(a.rgba, c.rgba)
        assert np.array_equal(a.hdr.view(np.uint32), c.hdr.view(np.uint32))
        assert np.array_equal(a.steps, c.steps)
    finally:
        b.shutdown()


def test_grid_sampler_across_chunks_equals_one_chunk():
    """The grid sampler's tone map of the mean radiance happens in the last chunk's
    shade: 16 grid samples in chunks of 5 (four chunks, mixed, overlapped) equal the
    one-chunk frame."""
    u = scene_uniforms(metric=sa.METRIC_KERR, spin=0.94, r_in=1.7, r_out=30.0, width=200, height=150, spp=16)
    b = sa.HipBackend(200, 150)
    try:
        print("grid sampler, chunk_samples 5 vs default", flush=True)
        a = b.render_outputs(u, OUTS, sa.PREC_MIXED, sampler=sa.SAMPLER_GRID)
        c = b.render_outputs(u, OUTS, sa.PREC_MIXED, sampler=sa.SAMPLER_GRID, chunk_samples=5, overlap_chunks=1)
        assert a.stats.chunks == 1 and c.stats.chunks == 4
        assert np.array_equal(a.rgba, c.rgba)
        assert np.array_equal(a.hdr.view(np.uint32), c.hdr.view(np.uint32))
    finally:
        b.shutdown()
