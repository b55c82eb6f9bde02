"""Error study of the fixed-step refinements (oracle Extras.cross_refine and
Extras.reproject, the rules behind SING_REFINE_CROSS64 / SING_REFINE_REPROJECT)
on the two 4K Kerr scenes, one pixel-centre ray per seeded pixel.  Test-side
measurement: it runs the oracle only.

    python -m oracle.refine_study [n_fp32_vs_fp64] [n_converged]

Prints two JSON tables:
  * "fp32_vs_fp64": FP32 restatement against FP64 under the same rule, over the
    rays the mixed form keeps in FP32 (rays a numpy restatement of the
    renderer's pre-march router flags, and grazing rays, excluded);
  * "vs_converged": each rule and precision against a converged FP64 trace
    (h/16, refined crossing), relative radiance error over disc pixels.
"""

from __future__ import annotations

import copy
import json
import math
import sys

import numpy as np

from . import Extras, render, trace_ray

ROOT_SCENES = {
    # scenes/kerr_a094.scene (the headline) and the a = 0.998 4K scene
    "a094": dict(spin=0.94, r_in=1.7, r_out=30.0, elev=0.15, fov=60.0),
    "a0998": dict(spin=0.998, r_in=None, r_out=30.0, elev=math.radians(30.0), fov=60.0),
}


def scene(name: str):
    import singularity_amd as sa

    p = ROOT_SCENES[name]
    s = sa.scene_defaults()
    s.metric = sa.METRIC_KERR
    s.spin_a_over_M = p["spin"]
    s.disc_inner_M = p["r_in"] if p["r_in"] is not None else sa.isco_prograde(p["spin"])
    s.disc_outer_M = p["r_out"]
    s.supersample = 1
    return sa.pack_uniforms(s, sa.orbital_camera(30.0, p["elev"], 0.0, p["fov"]), 3840, 2160, 0.0), p["spin"]


def routed_to_fp64(xi, eta, a, M, r_plus, band=0.03):
    """numpy restatement of the renderer's pre-march router (photon-orbit margin,
    ingoing rays reaching 2 r+, vortical and near-polar rays); camera rays of
    these scenes are ingoing."""
    C = (xi - a) ** 2 + eta
    P = 0.5 * (a * a - xi * xi - eta)
    Qc = 0.5 * M * C
    if P < 0 and -(4 * P ** 3 + 27 * Qc * Qc) > 0:
        arg = min(1.0, max(-1.0, (3 * Qc / (2 * P)) * math.sqrt(-3 / P)))
        rs = 2 * math.sqrt(-P / 3) * math.cos(math.acos(arg) / 3)
        if rs > r_plus:
            r2 = rs * rs
            m = abs((r2 * r2 + (a * a - xi * xi - eta) * r2 + 2 * M * C * rs - a * a * eta) / (r2 * r2))
            if m < band:
                return True
    r1 = 2 * r_plus
    r2 = r1 * r1
    if r2 * r2 + (a * a - xi * xi - eta) * r2 + 2 * M * C * r1 - a * a * eta >= 0:
        return True
    if eta < 0:
        return True
    B = eta + xi * xi - a * a
    umax = 2 * eta / (B + math.sqrt(B * B + 4 * a * a * eta)) if eta > 0 else 0.0
    return 1 - umax < 8.9973e-4


def rel(x, b):
    return np.abs(x - b).max(-1) / np.maximum(np.abs(b).max(-1), 1e-6)


def ex_for(rp: int, cr: int) -> Extras:
    e = Extras()
    e.reproject, e.cross_refine = rp, cr
    return e


def fp32_vs_fp64(name: str, n: int, seed: int = 7) -> dict:
    u, spin = scene(name)
    rng = np.random.default_rng(seed)
    px = np.sort(rng.choice(3840 * 2160, n, replace=False)).astype(np.int64)
    keep = np.ones(n, bool)
    r_plus = 1.0 + math.sqrt(max(1.0 - spin * spin, 0.0))
    for i, p in enumerate(px):
        gx, gy = int(p % 3840), int(p // 3840)
        r64 = trace_ray(u, gx, gy, 0.5, 0.5, precision=64)
        r32 = trace_ray(u, gx, gy, 0.5, 0.5, precision=32)
        E = r64["E"]
        keep[i] = not (routed_to_fp64(r64["Lz"] / E, r64["inv0"][1] / E ** 2, spin, 1.0, r_plus)
                       or r32["th_min"] < 2e-6)
    out = {"rays": n, "kept_in_fp32": int(keep.sum())}
    ref64 = render(u, 64, px)
    for rp, cr, label in ((0, 0, "reference"), (0, 1, "cross64"), (1, 0, "reproject"), (1, 1, "both")):
        a = render(u, 32, px, extras=ex_for(rp, cr))
        b = render(u, 64, px, extras=ex_for(rp, cr))
        r = rel(a["hdr"], b["hdr"])[keep]
        row = {"p99": float(np.quantile(r, 0.99)), "p999": float(np.quantile(r, 0.999)),
               "within_1e-3": float(np.mean(r <= 1e-3)),
               "drift_2H_median_fp32": float(np.median(a["drift"][keep, 0]))}
        if label != "reference":
            rr = rel(b["hdr"], ref64["hdr"])
            row["fp64_rule_vs_reference_fp64_within_1e-3"] = float(np.mean(rr <= 1e-3))
            row["fp64_rule_vs_reference_fp64_p99"] = float(np.quantile(rr, 0.99))
        out[label] = row
    return out


def vs_converged(name: str, n: int, seed: int = 7) -> dict:
    u, _ = scene(name)
    rng = np.random.default_rng(seed)
    px = np.sort(rng.choice(3840 * 2160, n, replace=False)).astype(np.int64)
    fine = copy.copy(u)
    fine.h_step = u.h_step / 16
    fine.max_steps = u.max_steps * 16
    truth = render(fine, 64, px, extras=ex_for(0, 1))
    disc = truth["cls"] == 2
    out = {"pixels": n, "disc_pixels": int(disc.sum())}
    for prec in (64, 32):
        for rp, cr, label in ((0, 0, "reference"), (0, 1, "cross64"), (1, 0, "reproject"), (1, 1, "both")):
            o = render(u, prec, px, extras=ex_for(rp, cr))
            ok = disc & (o["cls"] == truth["cls"])
            r = rel(o["hdr"], truth["hdr"])[ok]
            out[f"{label} fp{prec}"] = {"median": float(np.median(r)), "p90": float(np.quantile(r, 0.9)),
                                        "p99": float(np.quantile(r, 0.99))}
    return out


def main() -> int:
    n1 = int(sys.argv[1]) if len(sys.argv) > 1 else 40000
    n2 = int(sys.argv[2]) if len(sys.argv) > 2 else 20000
    res = {"fp32_vs_fp64": {"a094": fp32_vs_fp64("a094", n1)},
           "vs_converged": {k: vs_converged(k, n2) for k in ROOT_SCENES}}
    print(json.dumps(res, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
