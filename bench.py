"""bench.py — headline benchmark of the offline null-geodesic renderer on one MI355X.

Workload (the reference's published headline, README.md:39,130; docs/PHASE8_CUDA.md:84-87):
scenes/kerr_a094.scene (Kerr a/M = 0.94, disc [1.7 M, 30 M], Doppler + redshift) at
cuda_cli defaults (camera 30 M, elevation 0.15 rad, FOV 60 deg), 3840x2160, 256 spp
Halton(2,3), h = 0.4 M, <= 2500 steps, escape 200 M.  The reference renders it in
162 s on an RTX 3090 (13.1 Mrays/s, FP32).

One step = one full frame in the mixed FP32/FP64 form (photon-band rays in FP64).
Warmup frames are untimed; the K timed frames are bracketed by device-synchronous
calls (sing_render returns after the frame completes).  After them, one plain-FP32
frame and one all-FP64 frame of the same scene give the precision ratios and the
per-pixel radiance error of mixed and plain FP32 against FP64.

Prints ONE JSON line.  The CPU baseline is the repo's C++ restatement of the
reference path compiled with the reference's Release flags (-O3 -ffast-math),
row-parallel over the host cores, timed on a seeded pixel subset.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

import singularity_amd as sa
from singularity_amd import flops as FL

METRIC = "Mrays/s Kerr 4K (≤1e-3 err vs FP64), mixed vs plain FP32; grads/s"
PUBLISHED_MRAYS = 13.1  # 4K Kerr a/M=0.94, 256 spp, RTX 3090 (README.md:39,130)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def headline_uniforms(width: int, height: int, spp: int) -> sa.Uniforms:
    # scenes/kerr_a094.scene: spin 0.94, disc_r_inner 1.7, disc_r_outer 30, doppler/redshift on
    cfg = sa.parse_scene_config("mass = 1.0\nspin = 0.94\ndisc_r_inner = 1.7\ndisc_r_outer = 30.0\n"
                                "disc_doppler_on = 1\ndisc_redshift_on = 1\n")
    scene = sa.scene_from_config(cfg)
    scene.supersample = spp
    cam = sa.orbital_camera(30.0, 0.15, 0.0, 60.0)
    return sa.pack_uniforms(scene, cam, width, height, 0.0)


def err_stats(a: np.ndarray, ref: np.ndarray, ca: np.ndarray, cref: np.ndarray) -> dict:
    """Per-pixel radiance error of a frame against the all-FP64 GPU frame of the same
    scene (the same discrete algorithm in double on the device).  The FP64 host
    oracle checks of these scenes are the GPU tests (tests/test_gpu_headline.py,
    tests/test_gpu_config3.py)."""
    a = a.reshape(-1, 3).astype(np.float64)
    b = ref.reshape(-1, 3).astype(np.float64)
    rel = np.abs(a - b).max(-1) / np.maximum(np.abs(b).max(-1), 1e-6)
    return {
        "frac_pixels_rel_le_1e-3": float(np.mean(rel <= 1e-3)),
        "median_rel": float(np.median(rel)),
        "p99_rel": float(np.quantile(rel, 0.99)),
        "p999_rel": float(np.quantile(rel, 0.999)),
        "image_rel_l1": float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-30)),
        "class_agreement": float(np.mean(ca.ravel() == cref.ravel())),
    }


def cpu_baseline(u: sa.Uniforms, rows_per_thread: int = 2, seed: int = 1234) -> dict:
    """The C++ restatement in FP32 with the reference Release build's flags (-O3
    -ffast-math), scheduled like the reference CPU renderer (cli/main.cpp:739-740,
    861-872): one worker per hardware thread pulling whole image rows from an
    atomic counter.  Workers are capped at the 16 cores a one-GPU share of the host
    provides; the machine's own hardware_concurrency is reported beside it.  The
    sample is rows_per_thread seeded whole rows of the same frame per worker (a
    4K row at 256 spp is about 10^6 rays; rows differ in cost by 2-3x, so with
    one row per worker the slowest row would set the time)."""
    import oracle  # CPU baseline leg only

    hw = os.cpu_count() or 1
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = hw
    threads = max(1, min(avail, 16))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), model)
    except OSError:
        pass
    rows = np.random.default_rng(seed).permutation(u.height)[: rows_per_thread * threads]
    t0 = time.perf_counter()
    _, steps = oracle.render_rows(u, rows, threads=threads, lib=oracle.fast())
    el = time.perf_counter() - t0
    rays = rows.size * u.width * max(1, u.supersample)
    return {
        "value": round(rays / el / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
        "host_hardware_concurrency": hw, "cpu_model": model,
        "rk4_steps_per_s": round(steps / el, 1),
        "sample": f"{rows.size} seeded whole rows x {u.width} px x {max(1, u.supersample)} spp of the same frame "
                  f"({rays} rays, {el:.1f} s); FP32 C++ restatement, -O3 -ffast-math, {threads} std::threads "
                  f"pulling rows from an atomic counter (host reports {hw} hardware threads)",
    }


def kerr_uniforms(W, H, spp, spin, r_out, dist, incl_deg, fov_deg):
    """Kerr scene with r_in = prograde ISCO(a), Doppler + redshift, orbital camera."""
    s = sa.scene_defaults()
    s.metric = sa.METRIC_KERR
    s.spin_a_over_M = spin
    s.disc_inner_M = sa.isco_prograde(spin)
    s.disc_outer_M = r_out
    s.supersample = spp
    cam = sa.orbital_camera(dist, np.radians(90.0 - incl_deg), 0.0, fov_deg)
    return sa.pack_uniforms(s, cam, W, H, 0.0)


def c2_leg(frames: int = 3, seed: int = 20261203) -> dict:
    """Kerr a=0.94 1920x1080 16 spp, error-controlled RK4 (tol 1e-6), seeded disc
    turbulence, half-float HDR -> bloom (0.8 / 0.55, 7-tap) -> exposure -> ACES -> sRGB,
    mixed FP32/FP64; against an all-FP64 frame of the same algorithm."""
    W, H, spp = 1920, 1080, 16
    u = kerr_uniforms(W, H, spp, 0.94, 25.0, 40.0, 80.0, 30.0)
    be = sa.HipBackend(W, H)
    outs = sa.OUT_RGBA8 | sa.OUT_HDR | sa.OUT_CLASS

    def opts(prec, tol=1e-6):
        return sa.RenderOpts(precision=prec, outputs=outs, post=sa.POST_BLOOM, adapt_tol=tol, turb_seed=seed)

    be.render(u, opts(sa.PREC_MIXED))
    st = []
    t0 = time.perf_counter()
    for _ in range(frames):
        st.append(be.render(u, opts(sa.PREC_MIXED)))
    wall = (time.perf_counter() - t0) / frames
    fm = be.capture(outs)
    s32 = be.render(u, opts(sa.PREC_FP32))
    f32 = be.capture(outs)
    s64 = be.render(u, opts(sa.PREC_FP64))
    f64 = be.capture(outs)
    sfix = be.render(u, opts(sa.PREC_MIXED, 0.0))
    be.shutdown()
    rays = W * H * spp
    post_ms = float(np.mean([s.post_ms for s in st]))
    kern = float(np.mean([s.kernel_ms for s in st]))
    res = {
        "workload": f"Kerr a=0.94 {W}x{H} {spp} spp, camera 40 M incl 80 deg FOV 30, disc [ISCO, 25 M], "
                    f"adaptive RK4 tol 1e-6, turbulence seed {seed}, bloom post, mixed FP32/FP64",
        "mrays_s": round(rays / wall / 1e6, 2), "kernel_ms": round(kern, 2),
        "trace_ms": round(float(np.mean([s.trace_ms for s in st])), 2),
        "accepted_steps": int(st[-1].steps), "rejected_steps": int(st[-1].rejected),
        "rk4_evals_per_s": round(3 * (st[-1].steps + st[-1].rejected) / (kern * 1e-3), 1),
        "fp64_rays": int(st[-1].fp64_rays),
        "post_ms": round(post_ms, 4),
        "post_hbm_gbs": round(FL.POST_BYTES_PER_PIXEL * W * H / (post_ms * 1e-3) / 1e9, 1),
        "post_bytes_per_pixel": FL.POST_BYTES_PER_PIXEL,
        "plain_fp32": {"kernel_ms": round(s32.kernel_ms, 2), "error_vs_gpu_fp64": err_stats(f32.hdr, f64.hdr, f32.cls, f64.cls),
                       "rgba_max_lsb_vs_gpu_fp64": int(np.abs(f32.rgba.astype(int) - f64.rgba.astype(int)).max())},
        "fp64": {"kernel_ms": round(s64.kernel_ms, 2)},
        "mixed_error_vs_gpu_fp64": err_stats(fm.hdr, f64.hdr, fm.cls, f64.cls),
        "mixed_rgba_max_lsb_vs_gpu_fp64": int(np.abs(fm.rgba.astype(int) - f64.rgba.astype(int)).max()),
        "fixed_step_mixed": {"kernel_ms": round(sfix.kernel_ms, 2), "steps": int(sfix.steps)},
    }
    log(f"c2: {res['mrays_s']} Mrays/s, post {res['post_hbm_gbs']} GB/s")
    return res


def post_leg(reps: int = 5, launches: int = 50) -> dict:
    """The bloom post pass (extract, 7-tap H and V blurs, composite, ACES, sRGB) on
    1-spp FP32 frames of the 1080p Kerr scene at 1920x1080 and 3840x2160.  post_ms:
    the mean launch time of `launches` back-to-back passes over the frame's
    half-float HDR (sing_post_bloom, HIP events around all of them, best of reps);
    HBM GB/s in the four-pass algorithm's bytes (32 B per full-res pixel).
    in_frame_post_ms: the pass timed alone inside a frame (one launch between two
    events, launch latency included)."""
    res = {}
    for W, H in ((1920, 1080), (3840, 2160)):
        u = kerr_uniforms(W, H, 1, 0.94, 25.0, 40.0, 80.0, 30.0)
        be = sa.HipBackend(W, H)
        o = sa.RenderOpts(precision=sa.PREC_FP32, outputs=sa.OUT_RGBA8, post=sa.POST_BLOOM)
        be.render(u, o)
        in_frame = min(be.render(u, o).post_ms for _ in range(reps))
        ms = min(be.post_bloom(u, launches).post_ms for _ in range(reps))
        be.shutdown()
        gbs = FL.POST_BYTES_PER_PIXEL * W * H / (ms * 1e-3) / 1e9
        res[f"{W}x{H}"] = {"post_ms": round(ms, 4), "hbm_gbs": round(gbs, 1),
                           "frac_of_hbm_peak": round(gbs / FL.PEAK_HBM_GBS, 3), "launches": launches,
                           "in_frame_post_ms": round(in_frame, 4)}
        log(f"post {W}x{H}: {res[f'{W}x{H}']}")
    return res


def refine_leg(be: sa.HipBackend, u: sa.Uniforms, base_mixed_ms: float) -> dict:
    """The headline frame with the precision refinements of the fixed-step march
    (SING_REFINE_CROSS64: disc crossing at the FP64 cubic dense-output root;
    SING_REFINE_REPROJECT: null-constraint re-projection after every step): the
    mixed frame's time against the same frame without them, and its error
    against the all-FP64 GPU frame under the same rule."""
    outs = sa.OUT_HDR | sa.OUT_CLASS
    res = {}
    for name, rf in (("cross64", sa.REFINE_CROSS64), ("reproject", sa.REFINE_REPROJECT),
                     ("cross64+reproject", sa.REFINE_CROSS64 | sa.REFINE_REPROJECT)):
        stm = be.render(u, sa.RenderOpts(precision=sa.PREC_MIXED, outputs=outs, refine=rf))
        fm = be.capture(outs)
        st64 = be.render(u, sa.RenderOpts(precision=sa.PREC_FP64, outputs=outs, refine=rf))
        f64 = be.capture(outs)
        res[name] = {"mixed_kernel_ms": round(stm.kernel_ms, 2),
                     "mixed_trace32_ms": round(stm.trace32_ms, 2), "mixed_trace64_ms": round(stm.trace64_ms, 2),
                     "mixed_over_unrefined_mixed_time": round(stm.kernel_ms / base_mixed_ms, 3),
                     "fp64_kernel_ms": round(st64.kernel_ms, 2),
                     "fp64_over_mixed_time": round(st64.kernel_ms / stm.kernel_ms, 3),
                     "mixed_error_vs_gpu_fp64_same_rule": err_stats(fm.hdr, f64.hdr, fm.cls, f64.cls)}
        log(f"refine {name}: {res[name]}")
    return res


def c3_leg() -> dict:
    """Kerr a=0.998 3840x2160 256 spp, incl 60 deg, disc [ISCO, 30 M], cuda_cli camera
    defaults (30 M, FOV 60), fixed h = 0.4: plain FP32, mixed and all-FP64 frames."""
    W, H, spp = 3840, 2160, 256
    u = kerr_uniforms(W, H, spp, 0.998, 30.0, 30.0, 60.0, 60.0)
    be = sa.HipBackend(W, H)
    outs = sa.OUT_HDR | sa.OUT_CLASS | sa.OUT_RGBA8
    res = {"workload": f"Kerr a=0.998 {W}x{H} {spp} spp, incl 60 deg, disc [ISCO, 30 M], camera 30 M FOV 60, h 0.4"}
    frames = {}
    for name, prec in (("mixed", sa.PREC_MIXED), ("fp32", sa.PREC_FP32), ("fp64", sa.PREC_FP64)):
        st = be.render(u, sa.RenderOpts(precision=prec, outputs=outs))
        frames[name] = be.capture(outs)
        res[name] = {"kernel_ms": round(st.kernel_ms, 1), "mrays_s": round(W * H * spp / st.kernel_ms / 1e3, 2),
                     "steps": int(st.steps), "fp64_rays": int(st.fp64_rays)}
        log(f"c3 {name}: {st.kernel_ms:.0f} ms")
    be.shutdown()
    f64 = frames["fp64"]
    for name in ("mixed", "fp32"):
        f = frames[name]
        res[name]["error_vs_gpu_fp64"] = err_stats(f.hdr, f64.hdr, f.cls, f64.cls)
    res["mixed_over_plain_time"] = round(res["mixed"]["kernel_ms"] / res["fp32"]["kernel_ms"], 3)
    res["fp64_over_mixed_time"] = round(res["fp64"]["kernel_ms"] / res["mixed"]["kernel_ms"], 3)
    return res


def batch_scenes(seed: int = 20261203, n: int = 64) -> list:
    """64 uneven Kerr scenes from one recorded seed (numpy PCG64): resolution from
    {128^2, 256^2, 512^2, 1024^2, 1920x1080}, a/M ~ U[0, 0.998], inclination ~
    U[5, 89] deg, FOV ~ U[20, 60] deg, spp ~ U{1..64}, r_out ~ U[12, 40] M,
    r_in = ISCO_pro(a), camera 30 M."""
    rng = np.random.default_rng(seed)
    sizes = [(128, 128), (256, 256), (512, 512), (1024, 1024), (1920, 1080)]
    out = []
    for _ in range(n):
        W, H = sizes[int(rng.integers(0, len(sizes)))]
        a = float(rng.uniform(0.0, 0.998))
        incl = float(rng.uniform(5.0, 89.0))
        fov = float(rng.uniform(20.0, 60.0))
        spp = int(rng.integers(1, 65))
        r_out = float(rng.uniform(12.0, 40.0))
        s = sa.scene_defaults()
        s.metric = sa.METRIC_KERR if a > 0.0 else sa.METRIC_SCHWARZSCHILD
        s.spin_a_over_M = a
        s.disc_inner_M = sa.isco_prograde(a)
        s.disc_outer_M = r_out
        s.supersample = spp
        cam = sa.orbital_camera(30.0, np.radians(90.0 - incl), 0.0, fov)
        out.append(sa.pack_uniforms(s, cam, W, H, 0.0))
    return out


def c4_leg(seed: int = 20261203, reps: int = 2) -> dict:
    """64 uneven scenes in one persistent batched launch (recorded as a hipGraph and
    replayed), mixed precision, launch shape autotuned once; against rendering the
    same scenes one after another through sing_render."""
    scenes = batch_scenes(seed)
    rays = sum(u.width * u.height * u.supersample for u in scenes)
    be = sa.HipBackend(128, 128)
    opts = sa.BatchOpts(precision=sa.PREC_MIXED, use_graph=1, autotune=1)
    _, st = sa.render_batch(be, scenes, opts)  # tunes, records the graph
    log(f"c4: tuned step_batch={st.step_batch} waves_per_cu={st.waves_per_cu} rays_per_refill={st.rays_per_refill} "
        f"queue_chunk={st.queue_chunk}, {st.kernel_ms:.0f} ms")
    opts.autotune = 0
    times, walls = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        _, st = sa.render_batch(be, scenes, opts)
        walls.append(time.perf_counter() - t0)
        times.append(st.kernel_ms)
    # the scenes one after another, at the single-scene default launch shape
    # and at the batch's tuned one; the faster total is the comparison
    seq_shapes = {"default": (0, 0), "batch_tuned": (st.step_batch, st.waves_per_cu)}
    seq_ms = {}
    for name, (sb, wpc) in seq_shapes.items():
        seq_ms[name] = 0.0
        for u in scenes:
            be.resize(u.width, u.height)
            seq_ms[name] += be.render(u, sa.RenderOpts(precision=sa.PREC_MIXED, outputs=sa.OUT_RGBA8,
                                                       step_batch=sb, waves_per_cu=wpc)).kernel_ms
    seq = min(seq_ms.values())
    be.shutdown()
    kern = float(np.mean(times))
    res = {
        "workload": f"64 Kerr scenes (seed {seed}): res in {{128^2..1920x1080}}, a~U[0,0.998], incl~U[5,89] deg, "
                    f"FOV~U[20,60], spp~U{{1..64}}, r_out~U[12,40], r_in=ISCO, mixed precision, {rays} rays",
        "rays": rays, "mrays_s": round(rays / (kern * 1e-3) / 1e6, 2), "kernel_ms": round(kern, 1),
        "wall_ms_incl_output_copies": round(float(np.mean(walls)) * 1e3, 1), "copy_ms": round(st.copy_ms, 1),
        "steps": int(st.steps), "fp64_rays": int(st.fp64_rays), "groups": int(st.groups),
        "graph_replayed": int(st.graph_replayed), "step_batch": int(st.step_batch),
        "waves_per_cu": int(st.waves_per_cu), "rays_per_refill": int(st.rays_per_refill),
        "queue_chunk": int(st.queue_chunk),
        "sequential_single_scene_ms": round(seq, 1), "batch_speedup_vs_sequential": round(seq / kern, 3),
        "sequential_ms_by_shape": {k: round(v, 1) for k, v in seq_ms.items()},
    }
    log(f"c4: {res['mrays_s']} Mrays/s, batch {kern:.0f} ms vs sequential {seq:.0f} ms")
    return res


def adjoint_leg(evals: int = 50, W: int = 512) -> dict:
    """grads/s of the WxW (512x512) inverse render: one evaluation = forward with
    checkpoints + discrete adjoint + reduction of dL/d(a/M, incl, r_in, r_out, Ts),
    at the optimisation's start point against a target rendered at the goal.
    Also the longest ray's RK4 steps per precision: each launch lasts as long as
    its longest ray, so forward_ms / longest steps is the per-step time on the
    critical path.  batched_8: eight independent evaluations (the central-
    difference stencil of the four free parameters) per sing_diff_loss_grad_batch
    call."""
    be = sa.HipBackend(W, W)
    base = sa.diff_base_uniforms(W, W)
    cam = sa.DiffCamera(50.0, 0.0, 50.0)

    def opts(prec):
        o = sa.DiffOpts()
        o.precision, o.soft_edge, o.r_in_isco = prec, 1.0, 1
        return o

    goal = sa.DiffParams.from_array([0.7, np.radians(70.0), 0.0, 20.0, 1.0])
    start = sa.DiffParams.from_array([0.5, np.radians(60.0), 0.0, 15.0, 0.8])
    tgt, _ = sa.diff_forward(be, base, goal, cam, opts(sa.PREC_FP64))
    res = {"workload": "Kerr 512x512 1 spp, L2 over HDR RGB, start (0.5, 60 deg, 15 M, 0.8) vs target "
                       "(0.7, 70 deg, 20 M, 1.0), r_in = ISCO(a), soft annulus 1 M, checkpoint every 64 steps"}
    for name, prec in (("mixed", sa.PREC_MIXED), ("fp32", sa.PREC_FP32), ("fp64", sa.PREC_FP64)):
        for _ in range(3):
            sa.diff_loss_grad(be, base, start, cam, tgt, None, opts(prec))
        fwd, bwd = [], []
        t0 = time.perf_counter()
        for _ in range(evals):
            _, g, _, st = sa.diff_loss_grad(be, base, start, cam, tgt, None, opts(prec))
            fwd.append(st.forward_ms)
            bwd.append(st.backward_ms)
        wall = time.perf_counter() - t0
        _, meta = sa.diff_capture_pixels(be)
        steps, on64 = meta[..., 0], meta[..., 3] == 1
        # independent evaluations in one call (sing_diff_loss_grad_batch): the
        # central-difference stencil of the four free parameters, 8 points
        x0 = start.as_array()
        pts = [x0 + sgn * h * np.eye(5)[j] for j, h in ((0, 1e-3), (1, 1e-3), (3, 1e-2), (4, 1e-3))
               for sgn in (-1.0, 1.0)]
        sa.diff_loss_grad_batch(be, base, pts, cam, tgt, None, opts(prec))
        tb = time.perf_counter()
        for _ in range(max(2, evals // 8)):
            _, _, _, bst = sa.diff_loss_grad_batch(be, base, pts, cam, tgt, None, opts(prec))
        batched = len(pts) * max(2, evals // 8) / (time.perf_counter() - tb)
        res[name] = {"grads_per_s": round(evals / wall, 1), "forward_ms": round(float(np.mean(fwd)), 3),
                     "adjoint_ms": round(float(np.mean(bwd)), 3), "fp64_rays": int(st.fp64_rays),
                     "forward_steps": int(st.steps),
                     "forward_steps_per_s": round(int(st.steps) / (float(np.mean(fwd)) * 1e-3), 1),
                     "longest_ray_steps_fp32": int(steps[~on64].max()) if (~on64).any() else 0,
                     "longest_ray_steps_fp64": int(steps[on64].max()) if on64.any() else 0,
                     "batched_8_grads_per_s": round(batched, 1),
                     "batched_8_ms_per_call": round(bst.forward_ms + bst.backward_ms, 3),
                     "batched_8_forward_steps_per_s": round(int(bst.steps) / (bst.forward_ms * 1e-3), 1)}
        log(f"adjoint {name}: {res[name]}")
    be.shutdown()
    return res


def fan_leg(n_rays: int = 1 << 20, max_steps: int = 5000) -> dict:
    """The CPU driver's benchmark mode (cli/main.cpp:1492-1621) on the GPU: the
    equatorial ray fan over b in [-10, 10] M with h = 0.1, one lane per ray,
    at 2^20 rays (the driver's default is 100 rays on one CPU thread).  Reports
    RK4 steps/s of the integration kernel (HIP events)."""
    be = sa.HipBackend(1, 1)
    res = {"workload": f"ray fan b in [-10, 10] M, {n_rays} rays, max {max_steps} steps, h = 0.1 M"}
    for name, metric, spin in (("schw", sa.METRIC_SCHWARZSCHILD, 0.0), ("kerr_a09", sa.METRIC_KERR, 0.9)):
        sa.benchmark_fan(be, metric, spin, 4096, 200)  # warm-up
        best = None
        for _ in range(3):
            total, _, st = sa.benchmark_fan(be, metric, spin, n_rays, max_steps)
            best = st.kernel_ms if best is None else min(best, st.kernel_ms)
        res[name] = {"total_steps": total, "kernel_ms": round(best, 3),
                     "steps_per_s": round(total / (best * 1e-3), 1)}
        log(f"fan {name}: {res[name]}")
    be.shutdown()
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--res", default="3840x2160")
    ap.add_argument("--spp", type=int, default=256)
    ap.add_argument("--cpu-rows-per-thread", type=int, default=2)
    ap.add_argument("--no-compare", action="store_true", help="skip the plain-FP32 and FP64 comparison frames")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the grads/s leg (512x512 inverse render)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the refinement, 1080p adaptive+bloom, a=0.998 4K three-precision, post, "
                         "batch and integrator-fan legs")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU; the renderer has no multi-GPU path "
              "(rays shard by pixel tiles with no exchange inside the trace)", file=sys.stderr)
        return 2
    W, H = (int(v) for v in args.res.lower().split("x"))
    u = headline_uniforms(W, H, args.spp)
    # The adjoint leg runs first: once another context in the process has
    # rendered and released its streams, the runtime adds ~0.6 ms of host
    # time to each two-stream (mixed) evaluation for a while; its GPU time and
    # the single-stream forms are unchanged.
    adjoint = None if args.no_adjoint else adjoint_leg()
    be = sa.HipBackend(W, H)
    outs = sa.OUT_RGBA8 | sa.OUT_HDR | sa.OUT_CLASS
    mixed = sa.RenderOpts(precision=sa.PREC_MIXED, outputs=outs)
    rays_per_frame = W * H * args.spp

    for i in range(args.warmup):
        st = be.render(u, mixed)
        log(f"warmup {i}: {st.kernel_ms:.0f} ms")
    stats = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        stats.append(be.render(u, mixed))
        log(f"step {i}: {stats[-1].kernel_ms:.0f} ms")
    wall = time.perf_counter() - t0
    f_mixed = be.capture(outs)
    value = rays_per_frame * args.steps / wall / 1e6
    trace_ms = float(np.mean([s.trace_ms for s in stats]))
    kern_ms = float(np.mean([s.kernel_ms for s in stats]))
    steps_total = stats[-1].steps
    chunks = max(1, int(stats[-1].chunks))
    # the dominant kernel is the FP32 trace pass (trace_kernel<1, 1>), launched once
    # per sample chunk: its algorithmic flops per launch over its average launch time
    t32 = float(np.mean([s.trace32_ms for s in stats]))
    t64 = float(np.mean([s.trace64_ms for s in stats]))
    steps32, steps64 = int(stats[-1].steps32), int(stats[-1].steps64)
    achieved = FL.KERR_STEP_FLOPS * steps32 / (t32 * 1e-3) / 1e12
    achieved64 = FL.KERR_STEP_FLOPS * steps64 / (t64 * 1e-3) / 1e12 if t64 > 0 else 0.0
    # VALU issue of each pass per launch: its steps x instructions per step (SQ
    # counters) over its launch time (HIP events), in wave64 instructions per CU
    # per clock
    issue32 = FL.valu_issue_per_cu_clk(steps32 / chunks, t32 / chunks * 1e-3, FL.KERR_STEP_VALU_INSTS,
                                       FL.KERR_STEP_LANES)
    issue64 = (FL.valu_issue_per_cu_clk(steps64 / chunks, t64 / chunks * 1e-3, FL.KERR_STEP_VALU_INSTS_FP64,
                                        FL.KERR_STEP_LANES_FP64) if t64 > 0 else 0.0)

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": round(value / PUBLISHED_MRAYS, 3) if (W, H, args.spp) == (3840, 2160, 256) else None,
        "dtype": "fp32+fp64", "data": "synthetic",
        "config": {
            "workload": f"kerr_a094 scene {W}x{H} {args.spp} spp mixed FP32/FP64",
            "scene": "Kerr a/M=0.94, disc [1.7 M, 30 M], Doppler+redshift, 4500 K (scenes/kerr_a094.scene)",
            "camera": "r=30 M, elevation 0.15 rad, azimuth 0, FOV-y 60 deg",
            "integrator": "Hamiltonian RK4 h=0.4 M with pole damping, <=2500 steps, escape 200 M, horizon 1.02 r+",
            "resolution": f"{W}x{H}", "spp": args.spp, "precision": "mixed (photon-band rays FP64)",
        },
        "roofline": {
            "bound": "valu", "achieved": round(achieved, 3), "peak": FL.PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FL.PEAK_FP32_TFLOPS, 4),
            "traffic": round(FL.TRACE_HBM_BYTES_PER_RAY * rays_per_frame / chunks),
            "kernel": "trace_kernel<1, 1> (FP32 trace pass, one launch per sample chunk)",
            "flops_per_step": FL.KERR_STEP_FLOPS, "divides_per_step": FL.KERR_STEP_DIVS,
            "sincos_per_step": FL.KERR_STEP_TRIG, "launches_per_frame": chunks,
            "steps_per_launch": round(steps32 / chunks), "launch_ms": round(t32 / chunks, 3),
            "valu_insts_per_step": FL.KERR_STEP_VALU_INSTS,
            "valu_issue_per_cu_clk": round(issue32, 3),
            "valu_issue_frac": round(issue32 / FL.VALU_ISSUE_PEAK_PER_CU_CLK, 3),
        },
        "roofline_fp64": {
            "bound": "valu", "achieved": round(achieved64, 3), "peak": FL.PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved64 / FL.PEAK_FP64_TFLOPS, 4),
            "kernel": "trace_kernel<1, 3> (FP64 pass over the rays routed to it)",
            "flops_per_step": FL.KERR_STEP_FLOPS, "steps_per_launch": round(steps64 / chunks),
            "launch_ms": round(t64 / chunks, 3),
            "valu_insts_per_step": FL.KERR_STEP_VALU_INSTS_FP64,
            "valu_issue_per_cu_clk": round(issue64, 3),
            "valu_issue_frac": round(issue64 / FL.VALU_ISSUE_PEAK_FP64_PER_CU_CLK, 3),
        },
        "kernel_ms_per_frame": round(kern_ms, 2), "trace_ms_per_frame": round(trace_ms, 2),
        "shade_ms_per_frame": round(float(np.mean([s.shade_ms for s in stats])), 2),
        "rk4_steps_per_s": round(steps_total / (kern_ms * 1e-3), 1),
        "rk4_steps_per_s_fp32_pass": round(steps32 / (t32 * 1e-3), 1),
        "fp64_rays_per_frame": int(stats[-1].fp64_rays),
        "chunks_per_frame": int(stats[-1].chunks),
    }
    if not args.no_compare:
        st32 = be.render(u, sa.RenderOpts(precision=sa.PREC_FP32, outputs=outs))
        f32 = be.capture(outs)
        log(f"plain fp32: {st32.kernel_ms:.0f} ms")
        st64 = be.render(u, sa.RenderOpts(precision=sa.PREC_FP64, outputs=outs))
        f64 = be.capture(outs)
        log(f"fp64: {st64.kernel_ms:.0f} ms")
        out["plain_fp32"] = {"mrays_s": round(rays_per_frame / st32.kernel_ms / 1e3, 3),
                             "kernel_ms": round(st32.kernel_ms, 2),
                             "error_vs_gpu_fp64": err_stats(f32.hdr, f64.hdr, f32.cls, f64.cls)}
        out["fp64"] = {"mrays_s": round(rays_per_frame / st64.kernel_ms / 1e3, 3), "kernel_ms": round(st64.kernel_ms, 2)}
        out["mixed_error_vs_gpu_fp64"] = err_stats(f_mixed.hdr, f64.hdr, f_mixed.cls, f64.cls)
        out["mixed_over_plain_time"] = round(kern_ms / st32.kernel_ms, 3)
        out["fp64_over_mixed_time"] = round(st64.kernel_ms / kern_ms, 3)
    if not args.no_extra:
        out["refine"] = refine_leg(be, u, kern_ms)
    be.shutdown()
    if not args.no_extra:
        out["adaptive_bloom_1080p"] = c2_leg()
        out["post_passes"] = post_leg()
        out["kerr_a0998_4k"] = c3_leg()
        out["batch_64_scenes"] = c4_leg()
        out["integrator_fan"] = fan_leg()
    if adjoint is not None:
        out["adjoint"] = adjoint
        out["grads_per_s"] = adjoint["mixed"]["grads_per_s"]
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(u, args.cpu_rows_per_thread)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
