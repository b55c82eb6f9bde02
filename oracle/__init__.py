"""oracle — CPU restatement of the reference renderer (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker or the timed CPU baseline; the product
(singularity_amd) never does.

Libraries (built by oracle/Makefile):
    liboracle.so       strict IEEE build (-O2 -fno-fast-math -ffp-contract=off)
    liboracle_fast.so  reference Release flags (-O3 -ffast-math): CPU baseline
    _ref/libref.so     the reference's own physics headers, built from
                       /root/reference where that tree exists
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_LIBS: dict[str, C.CDLL] = {}


class Extras(C.Structure):
    _fields_ = [("soft_edge", C.c_double), ("band_phase1", C.c_double), ("band_phase2", C.c_double),
                ("adapt_tol", C.c_double), ("peak_T_scale", C.c_double), ("jitter_x", C.c_double),
                ("jitter_y", C.c_double), ("grid_sampler", C.c_double), ("reproject", C.c_double),
                ("cross_refine", C.c_double)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def _load(name: str) -> C.CDLL:
    if name not in _LIBS:
        path = HERE / name
        if not path.exists():
            build()
        L = C.CDLL(str(path))
        if name != "_ref/libref.so":
            _type_oracle(L)
        _LIBS[name] = L
    return _LIBS[name]


def strict() -> C.CDLL:
    return _load("liboracle.so")


def fast() -> C.CDLL:
    return _load("liboracle_fast.so")


def ref_available() -> bool:
    return (HERE / "_ref" / "libref.so").exists()


def ref() -> C.CDLL:
    return _load("_ref/libref.so")


def _type_oracle(L: C.CDLL) -> None:
    P = C.c_void_p
    L.orc_render.restype = C.c_int
    L.orc_render.argtypes = [P, C.c_int, P, P, C.c_int64, C.c_int, P, P, P, P, P, P]
    L.orc_render_rows.restype = C.c_longlong
    L.orc_render_rows.argtypes = [P, C.c_int, P, C.c_int64, C.c_int, P]
    L.orc_trace_ray.restype = C.c_int
    L.orc_trace_ray.argtypes = [P, C.c_int, P, C.c_uint32, C.c_uint32, C.c_double, C.c_double, P]
    L.orc_halton.restype = C.c_float
    L.orc_halton.argtypes = [C.c_uint32, C.c_uint32]
    L.orc_kerr_closed_form.restype = C.c_double
    L.orc_kerr_closed_form.argtypes = [C.c_int, C.c_double, C.c_double]
    L.orc_kerr_closed_form_f32.restype = C.c_float
    L.orc_kerr_closed_form_f32.argtypes = [C.c_int, C.c_float, C.c_float]
    L.orc_disc_temperature_f32.restype = C.c_float
    L.orc_disc_temperature_f32.argtypes = [C.c_float, C.c_float, C.c_float]
    L.orc_disc_temperature_f64.restype = C.c_double
    L.orc_disc_temperature_f64.argtypes = [C.c_double, C.c_double, C.c_double]
    L.orc_band_mask_f32.restype = C.c_float
    L.orc_band_mask_f32.argtypes = [C.c_float] * 7
    L.orc_blackbody_f32.argtypes = [C.c_float, P]
    L.orc_blackbody_f64.argtypes = [C.c_double, P]
    L.orc_starfield_f32.argtypes = [C.c_float, C.c_float, C.c_float, P]
    L.orc_schw_shift_f32.argtypes = [C.c_float] * 5 + [C.c_uint, P]
    L.orc_kerr_shift_f32.argtypes = [C.c_float] * 5 + [C.c_uint, P]
    L.orc_kerr_shift_f64.argtypes = [C.c_double] * 5 + [C.c_uint, P]
    L.orc_disc_flux_f32.restype = C.c_float
    L.orc_disc_flux_f32.argtypes = [C.c_float, C.c_float]
    L.orc_disc_emit_f32.argtypes = [P] + [C.c_float] * 5 + [P]
    L.orc_mix32.restype = C.c_uint32
    L.orc_mix32.argtypes = [C.c_uint32, C.c_uint32]
    L.orc_aces_f32.argtypes = [P, P]
    L.orc_linear_to_srgb_f32.restype = C.c_float
    L.orc_linear_to_srgb_f32.argtypes = [C.c_float]
    L.orc_to_u8_f32.restype = C.c_uint8
    L.orc_to_u8_f32.argtypes = [C.c_float]
    L.orc_camera_ray_f32.argtypes = [P, C.c_uint32, C.c_uint32, C.c_float, C.c_float, P]
    L.orc_kerr_rhs_f32.argtypes = [P, P, P]
    L.orc_kerr_rhs_f64.argtypes = [P, P, P]
    L.orc_kerr_two_H_f32.restype = C.c_float
    L.orc_kerr_two_H_f32.argtypes = [P, P]
    L.orc_kerr_two_H_f64.restype = C.c_double
    L.orc_kerr_two_H_f64.argtypes = [P, P]
    L.orc_kerr_reproject_f32.argtypes = [P, P, P]
    L.orc_kerr_reproject_f64.argtypes = [P, P, P]
    L.orc_kerr_refined_crossing_f32.argtypes = [P, P, C.c_float, P, P]
    L.orc_kerr_refined_crossing_f64.argtypes = [P, P, C.c_double, P, P]
    L.orc_kerr_rk4_f32.argtypes = [P, C.c_float, P, C.c_int, P]
    L.orc_kerr_rk4_f64.argtypes = [P, C.c_double, P, C.c_int, P]
    L.orc_schw_rhs_f32.argtypes = [P, C.c_float, P]
    L.orc_schw_rhs_f64.argtypes = [P, C.c_double, P]
    L.orc_schw_rk4_f32.argtypes = [P, C.c_float, C.c_float, C.c_int, P]
    L.orc_schw_rk4_f64.argtypes = [P, C.c_double, C.c_double, C.c_int, P]
    L.orc_crossing_f32.argtypes = [C.c_float, C.c_float, P, P]
    L.orc_to_half.argtypes = [P, C.c_int64]
    L.orc_bloom_extract.argtypes = [P, C.c_int, C.c_int, C.c_float, P, C.c_int]
    L.orc_bloom_blur.argtypes = [P, C.c_int, C.c_int, C.c_int, P, C.c_int]
    L.orc_bloom_composite.argtypes = [P, P, C.c_int, C.c_int, C.c_float, C.c_float, P]
    L.orc_count_kerr_step.argtypes = [P]
    L.orc_count_schw_step.argtypes = [P]
    L.orc_make_uniforms.argtypes = [P, P, C.c_double, C.c_double, C.c_double, P]
    L.orc_loss_grad_dual.restype = C.c_double
    L.orc_loss_grad_dual.argtypes = [P, P, P, C.c_double, C.c_double, C.c_double, P, C.c_int64, P, C.c_int, P, P]
    L.orc_loss_f64.restype = C.c_double
    L.orc_kerr_adaptive_step.restype = C.c_double
    L.orc_kerr_adaptive_step.argtypes = [P, P, C.c_double, C.c_double, P, P]
    L.orc_kerr_rk4_substeps.argtypes = [P, P, C.c_double, C.c_int, P]
    L.orc_loss_f64.argtypes = [P, P, P, C.c_double, C.c_double, C.c_double, P, C.c_int64, P, C.c_int, P]
    L.orc_kerr_geometry.restype = C.c_int
    L.orc_kerr_geometry.argtypes = [C.c_int, C.c_double, C.c_double, P]
    L.orc_photon_orbit.restype = C.c_int
    L.orc_photon_orbit.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, P, C.c_int, P]
    L.orc_bench_fan.restype = C.c_longlong
    L.orc_bench_fan.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_double, C.c_double, P]


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data


def _ubuf(u) -> C.Array:
    b = bytes(u)
    assert len(b) == 160, "uniform block must be 160 bytes"
    return C.create_string_buffer(b, 160)


def render(u, precision: int = 32, pixels: np.ndarray | None = None, threads: int = 0,
           extras: Extras | None = None, lib: C.CDLL | None = None) -> dict:
    """Render all pixels (or the listed linear pixel indices) with the oracle.

    Returns hdr (n,3) f64, ldr (n,3) f64, rgba (n,4) u8, cls (n,) u8, steps (n,) u32,
    drift (n,3) f64 (max |invariant(end) - invariant(start)| over the pixel's samples).
    """
    L = lib or strict()
    ub = _ubuf(u)
    width = int.from_bytes(bytes(u)[112:116], "little")
    height = int.from_bytes(bytes(u)[116:120], "little")
    if pixels is None:
        n = width * height
        px = None
    else:
        px = np.ascontiguousarray(pixels, dtype=np.int64)
        n = int(px.size)
    out = {
        "hdr": np.zeros((n, 3), np.float64), "ldr": np.zeros((n, 3), np.float64),
        "rgba": np.zeros((n, 4), np.uint8), "cls": np.zeros(n, np.uint8),
        "steps": np.zeros(n, np.uint32), "drift": np.zeros((n, 3), np.float64),
    }
    ex = C.byref(extras) if extras is not None else None
    rc = L.orc_render(ub, precision, ex, _ptr(px), n, threads or (os.cpu_count() or 1), _ptr(out["hdr"]),
                      _ptr(out["ldr"]), _ptr(out["rgba"]), _ptr(out["cls"]), _ptr(out["steps"]), _ptr(out["drift"]))
    assert rc == 0
    return out


def render_rows(u, rows, precision: int = 32, threads: int = 0, lib: C.CDLL | None = None) -> tuple[np.ndarray, int]:
    """Render whole rows with workers pulling the next row from an atomic counter
    (the reference CPU renderer's schedule); returns (rgba (n, W, 4), RK4 steps)."""
    L = lib or strict()
    width = int.from_bytes(bytes(u)[112:116], "little")
    r = np.ascontiguousarray(rows, dtype=np.int32)
    rgba = np.zeros((r.size, width, 4), np.uint8)
    steps = L.orc_render_rows(_ubuf(u), precision, _ptr(r), r.size, threads or (os.cpu_count() or 1), _ptr(rgba))
    assert steps >= 0
    return rgba, int(steps)


def boundary_band(u, pixels: np.ndarray | None = None, eps: float = 1e-4, threads: int = 0,
                  extras: Extras | None = None) -> np.ndarray:
    """Pixels whose FP64 termination class changes when the sample moves by
    +-eps pixel in x or y: the chaotic boundary band (photon ring, disc edges,
    shadow edge) where FP32 and FP64 may legitimately disagree."""
    base = render(u, 64, pixels, threads, extras=extras)["cls"]
    band = np.zeros(base.shape, bool)
    for jx, jy in ((eps, 0.0), (-eps, 0.0), (0.0, eps), (0.0, -eps)):
        ex = Extras()
        if extras is not None:
            C.memmove(C.addressof(ex), C.addressof(extras), C.sizeof(Extras))
        ex.jitter_x, ex.jitter_y = jx, jy
        band |= render(u, 64, pixels, threads, extras=ex)["cls"] != base
    return band


def splitmix64(seed: int, n: int) -> list[int]:
    """First n outputs of SplitMix64 with state = seed (Steele, Lea & Flood 2014)."""
    m = (1 << 64) - 1
    z, out = seed & m, []
    for _ in range(n):
        z = (z + 0x9E3779B97F4A7C15) & m
        x = z
        x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & m
        x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & m
        out.append(x ^ (x >> 31))
    return out


def turbulence_extras(seed: int, adapt_tol: float = 0.0) -> Extras:
    """Extras for a seeded disc-band pattern: phases = top 53 bits of the first two
    SplitMix64 outputs, scaled to [0, 2 pi)."""
    ex = Extras()
    if seed:
        a, b = splitmix64(seed, 2)
        ex.band_phase1 = (a >> 11) * (1.0 / 9007199254740992.0) * 6.283185307179586476925286766559
        ex.band_phase2 = (b >> 11) * (1.0 / 9007199254740992.0) * 6.283185307179586476925286766559
    ex.adapt_tol = adapt_tol
    return ex


def trace_ray(u, gx: int, gy: int, ox: float = 0.5, oy: float = 0.5, precision: int = 32,
              extras: Extras | None = None) -> dict:
    o = np.zeros(19, np.float64)
    ex = C.byref(extras) if extras is not None else None
    strict().orc_trace_ray(_ubuf(u), precision, ex, gx, gy, ox, oy, _ptr(o))
    return {"rgb": o[0:3].copy(), "cls": int(o[3]), "steps": int(o[4]), "r_cross": o[5],
            "inv0": o[6:9].copy(), "inv1": o[9:12].copy(), "E": o[12], "Lz": o[13], "r_min": o[14],
            "th_min": o[15], "esc_theta": o[16], "esc_phi": o[17], "th_max": o[18]}


def count_step(kerr: bool = True) -> dict:
    c = np.zeros(6, np.float64)
    (strict().orc_count_kerr_step if kerr else strict().orc_count_schw_step)(_ptr(c))
    return dict(zip(["add", "mul", "div", "sqrt", "trig", "other"], c.astype(int).tolist()))


def kerr_radius(which: int, M: float, a: float) -> float:
    return float(strict().orc_kerr_closed_form(which, M, a))


KERR_GEOMETRY_KEYS = ("mass", "spin", "spin_over_mass", "outer_horizon", "inner_horizon", "ergosphere_polar",
                      "ergosphere_equatorial", "photon_sphere_prograde", "photon_sphere_retrograde", "isco_prograde",
                      "isco_retrograde")


def kerr_geometry(mass: float, spin: float, precision: int = 32) -> dict | None:
    """The eleven Kerr scalars of the CPU driver's kerr-geometry mode (None: argument error)."""
    o = np.zeros(11, np.float64)
    if strict().orc_kerr_geometry(precision, mass, spin, _ptr(o)) != 0:
        return None
    return dict(zip(KERR_GEOMETRY_KEYS, o.tolist()))


def photon_orbit(mass: float, spin: float, r_init: float, orbits: float = 1.0, h: float = 0.01,
                 precision: int = 32) -> np.ndarray | None:
    """Rows (step, t, r, theta, phi) of the CPU driver's photon-orbit mode (None: argument error)."""
    n = C.c_int(0)
    strict().orc_photon_orbit(precision, mass, spin, r_init, orbits, h, None, 0, C.byref(n))
    rows = np.zeros((max(n.value, 1), 5), np.float64)
    rc = strict().orc_photon_orbit(precision, mass, spin, r_init, orbits, h, _ptr(rows), n.value, C.byref(n))
    return None if rc != 0 else rows[: n.value]


def bench_fan(kerr: bool, spin: float, n_rays: int, max_steps: int, b_min: float = -10.0, b_max: float = 10.0,
              precision: int = 32) -> tuple[int, np.ndarray]:
    """Total and per-ray RK4 step counts of the CPU driver's benchmark ray fan."""
    per = np.zeros(n_rays, np.uint32)
    total = strict().orc_bench_fan(precision, int(kerr), spin, n_rays, max_steps, b_min, b_max, _ptr(per))
    return int(total), per


def bloom_pipeline(hdr_rgb: np.ndarray, threshold: float, strength: float, exposure: float,
                   half_store: bool = True) -> dict:
    """Post passes on an (H, W, 3) float32 HDR image; returns all intermediates."""
    L = strict()
    H, W, _ = hdr_rgb.shape
    h = np.ascontiguousarray(hdr_rgb, dtype=np.float32).copy()
    if half_store:
        L.orc_to_half(_ptr(h), h.size)
    Wb, Hb = W // 2, H // 2
    ext = np.zeros((Hb, Wb, 3), np.float32)
    bh = np.zeros_like(ext)
    bv = np.zeros_like(ext)
    L.orc_bloom_extract(_ptr(h), W, H, threshold, _ptr(ext), int(half_store))
    L.orc_bloom_blur(_ptr(ext), Wb, Hb, 1, _ptr(bh), int(half_store))
    L.orc_bloom_blur(_ptr(bh), Wb, Hb, 0, _ptr(bv), int(half_store))
    rgba = np.zeros((H, W, 4), np.uint8)
    L.orc_bloom_composite(_ptr(h), _ptr(bv), W, H, strength, exposure, _ptr(rgba))
    return {"hdr": h, "extract": ext, "blur_h": bh, "blur_v": bv, "rgba": rgba}
