"""sing_render_async / sing_wait through the C ABI: the same frame as
sing_render, bit for bit; the host is free while the GPU renders; any other
call on the context completes a pending frame first; argument errors are
returned by the submission and leave nothing pending."""

from __future__ import annotations

import time

import numpy as np
import pytest

import singularity_amd as sa
from helpers import scene_uniforms

pytestmark = pytest.mark.gpu

OUTS = sa.OUT_RGBA8 | sa.OUT_HDR | sa.OUT_CLASS | sa.OUT_STEPS


def _u(spp=4, w=512, h=384, spin=0.94):
    return scene_uniforms(metric=sa.METRIC_KERR, spin=spin, r_in=1.7, r_out=30.0, width=w, height=h, spp=spp)


def _frames_equal(a, b):
    assert np.array_equal(a.rgba, b.rgba)
    assert np.array_equal(a.hdr.view(np.uint32), b.hdr.view(np.uint32))
    assert np.array_equal(a.cls, b.cls)
    assert np.array_equal(a.steps, b.steps)


@pytest.fixture(scope="module")
def be():
    b = sa.HipBackend(512, 384)
    yield b
    b.shutdown()


@pytest.mark.parametrize("prec", [sa.PREC_FP32, sa.PREC_MIXED])
def test_async_then_wait_equals_render(be, prec):
    u = _u()
    o = sa.RenderOpts(precision=prec, outputs=OUTS)
    print(f"async vs sync prec={prec}", flush=True)
    st_sync = be.render(u, o)
    ref = be.capture(OUTS)
    be.render_async(u, o)
    st = be.wait()
    f = be.capture(OUTS)
    _frames_equal(f, ref)
    assert (st.rays, st.steps, st.fp64_rays, st.steps32, st.steps64) == (
        st_sync.rays, st_sync.steps, st_sync.fp64_rays, st_sync.steps32, st_sync.steps64)
    assert st.rays == 512 * 384 * 4 and st.kernel_ms > 0.0


def test_capture_completes_a_pending_frame_and_wait_still_reports_it(be):
    u = _u(spp=8)
    o = sa.RenderOpts(precision=sa.PREC_MIXED, outputs=OUTS)
    be.render(u, o)
    ref = be.capture(OUTS)
    u2 = _u(spp=2, spin=0.5)
    be.render(u2, o)  # something else in the outputs first
    print("capture while pending", flush=True)
    be.render_async(u, o)
    f = be.capture(OUTS)  # waits for the frame
    _frames_equal(f, ref)
    st = be.wait()  # the completed frame's stats are still there
    assert st.rays == 512 * 384 * 8
    with pytest.raises(sa.SingError) as e:  # and only once
        be.wait()
    assert e.value.code == 3


def test_wait_without_a_submission_is_a_state_error(be):
    be.render(_u(spp=1), sa.RenderOpts(outputs=OUTS))
    with pytest.raises(sa.SingError) as e:
        be.wait()
    assert e.value.code == 3


def test_second_submission_completes_the_first(be):
    o = sa.RenderOpts(precision=sa.PREC_FP32, outputs=OUTS)
    ua, ub = _u(spp=3), _u(spp=5, spin=0.7)
    be.render(ub, o)
    ref_b = be.capture(OUTS)
    print("two submissions", flush=True)
    be.render_async(ua, o)
    be.render_async(ub, o)
    st = be.wait()
    assert st.rays == 512 * 384 * 5
    _frames_equal(be.capture(OUTS), ref_b)


def test_bad_submission_returns_the_error_and_leaves_nothing_pending(be):
    u = _u()
    with pytest.raises(sa.SingError) as e:
        be.render_async(u, sa.RenderOpts(precision=7))
    assert e.value.code == 1
    with pytest.raises(sa.SingError):
        be.wait()
    # validation fails before anything is overwritten: the previous frame
    # stays capturable
    be.render(u, sa.RenderOpts(outputs=OUTS))
    with pytest.raises(sa.SingError):
        be.render_async(u, sa.RenderOpts(precision=7))
    assert be.capture(sa.OUT_RGBA8).rgba.shape == (384, 512, 4)


def test_host_is_free_while_the_gpu_renders():
    W, H = 1920, 1080
    u = scene_uniforms(metric=sa.METRIC_KERR, spin=0.94, r_in=1.7, r_out=30.0, width=W, height=H, spp=64)
    b = sa.HipBackend(W, H)
    try:
        o = sa.RenderOpts(precision=sa.PREC_MIXED, outputs=sa.OUT_RGBA8)
        st = b.render(u, o)  # warm: buffers allocated
        print(f"1080p 64 spp: {st.kernel_ms:.1f} ms", flush=True)
        t0 = time.perf_counter()
        b.render_async(u, o)
        submit_ms = (time.perf_counter() - t0) * 1e3
        st2 = b.wait()
        print(f"submission returned after {submit_ms:.2f} ms of a {st2.kernel_ms:.1f} ms frame", flush=True)
        assert submit_ms < 0.5 * st2.kernel_ms
        assert st2.rays == W * H * 64
    finally:
        b.shutdown()


def test_contexts_driven_from_two_host_threads_at_once():
    """Two contexts, each rendered by its own host thread at the same time (ctypes
    releases the GIL during the calls): every frame equals the same frame rendered
    alone, and a bad call on one thread leaves its error on that thread only."""
    import threading

    scenes = {
        "kerr": _u(spp=16, w=640, h=480),
        "schw": scene_uniforms(metric=sa.METRIC_SCHWARZSCHILD, spin=0.0, r_in=6.0, r_out=20.0, width=480, height=320,
                               spp=8),
    }
    precs = {"kerr": sa.PREC_MIXED, "schw": sa.PREC_FP32}
    refs = {}
    for k, u in scenes.items():
        b = sa.HipBackend(u.width, u.height)
        b.render(u, sa.RenderOpts(precision=precs[k], outputs=OUTS))
        refs[k] = b.capture(OUTS)
        b.shutdown()
    results, errors = {}, {}

    def worker(k):
        try:
            u = scenes[k]
            b = sa.HipBackend(u.width, u.height)
            frames = []
            for _ in range(3):
                b.render(u, sa.RenderOpts(precision=precs[k], outputs=OUTS))
                frames.append(b.capture(OUTS))
            # an error on this thread: the thread's own last error names it
            assert sa.lib().sing_create(0, 1, 0) is None
            errors[k] = sa.lib().sing_last_error(None)
            b.shutdown()
            results[k] = frames
        except Exception as exc:  # reported by the main thread
            results[k] = exc

    print("two host threads, two contexts", flush=True)
    threads = [threading.Thread(target=worker, args=(k,)) for k in scenes]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    for k in scenes:
        assert not isinstance(results.get(k), Exception), results.get(k)
        assert len(results[k]) == 3
        for f in results[k]:
            _frames_equal(f, refs[k])
        assert b"width/height" in errors[k]
