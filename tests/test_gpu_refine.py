"""SING_REFINE_CROSS64 and SING_REFINE_REPROJECT (include/sing_abi.h) on the GPU
against the oracle's same rules (oracle/oracle.hpp: kerr_refined_crossing,
kerr_reproject), on the headline scene (Kerr a/M = 0.94, disc [1.7 M, 30 M],
camera 30 M / 0.15 rad / FOV 60, h = 0.4) at 1920x1080, 1 spp, on a seeded 1 %
pixel subset (numpy PCG64, seed 20261203; 20,736 pixels).

Bars (tolerances stated here):
  FP64 GPU   class agreement >= 99.99 %, HDR within 1e-5 relative on >= 99.9 %
  mixed GPU  outside the oracle's boundary band (same rule): class >= 99.9 %,
             HDR within 1e-3 on >= 99.9 %
  CROSS64    the mixed frame's disc pixels against a converged oracle trace
             (h/16, same rule): median relative error <= 1e-5 and at most
             1/100 of the linear rule's
"""

from __future__ import annotations

import copy

import numpy as np
import pytest

import singularity_amd as sa
from helpers import rel_err

pytestmark = pytest.mark.gpu

W, H = 1920, 1080
SEED = 20261203


def headline() -> sa.Uniforms:
    cfg = sa.parse_scene_config("mass = 1.0\nspin = 0.94\ndisc_r_inner = 1.7\ndisc_r_outer = 30.0\n"
                                "disc_doppler_on = 1\ndisc_redshift_on = 1\n")
    return sa.pack_uniforms(sa.scene_from_config(cfg), sa.orbital_camera(30.0, 0.15, 0.0, 60.0), W, H, 0.0)


def extras_for(orc, refine: int):
    ex = orc.Extras()
    ex.cross_refine = 1 if refine & sa.REFINE_CROSS64 else 0
    ex.reproject = 1 if refine & sa.REFINE_REPROJECT else 0
    return ex


@pytest.fixture(scope="module")
def be():
    b = sa.HipBackend(W, H)
    yield b
    b.shutdown()


@pytest.fixture(scope="module")
def subset():
    rng = np.random.default_rng(SEED)
    return np.sort(rng.choice(W * H, size=W * H // 100, replace=False))


@pytest.mark.parametrize("refine", [sa.REFINE_CROSS64, sa.REFINE_REPROJECT, sa.REFINE_CROSS64 | sa.REFINE_REPROJECT],
                         ids=["cross64", "reproject", "both"])
def test_refine_vs_fp64_oracle(be, orc, subset, refine):
    u = headline()
    ex = extras_for(orc, refine)
    print(f"oracle fp64 refine={refine} on {subset.size} pixels + boundary band", flush=True)
    o = orc.render(u, 64, pixels=subset, extras=ex)
    band = orc.boundary_band(u, pixels=subset, extras=ex)
    res = {}
    for name, prec in (("mixed", sa.PREC_MIXED), ("fp64", sa.PREC_FP64)):
        print(f"render {W}x{H} 1 spp refine={refine} {name}", flush=True)
        f = be.render_outputs(u, sa.OUT_HDR | sa.OUT_CLASS, prec, refine=refine)
        hdr = f.hdr.reshape(-1, 3)[subset]
        cls = f.cls.ravel()[subset]
        rel = rel_err(hdr, o["hdr"])
        res[name] = {
            "class": float((cls == o["cls"]).mean()),
            "class_outside_band": float((cls[~band] == o["cls"][~band]).mean()),
            "rel_le_1e-3_outside_band": float((rel[~band] <= 1e-3).mean()),
            "rel_le_1e-5": float((rel <= 1e-5).mean()),
        }
        print(name, res[name], flush=True)
    assert res["fp64"]["class"] >= 0.9999
    assert res["fp64"]["rel_le_1e-5"] >= 0.999
    assert res["mixed"]["class_outside_band"] >= 0.999
    assert res["mixed"]["rel_le_1e-3_outside_band"] >= 0.999


def test_cross64_tracks_the_converged_trace(be, orc, subset):
    u = headline()
    px = subset[::4]
    fine = copy.copy(u)
    fine.h_step = u.h_step / 16
    fine.max_steps = u.max_steps * 16
    ex = extras_for(orc, sa.REFINE_CROSS64)
    print(f"oracle fp64 h/16 on {px.size} pixels", flush=True)
    truth = orc.render(fine, 64, pixels=px, extras=ex)
    print("render mixed, linear and refined crossings", flush=True)
    lin = be.render_outputs(u, sa.OUT_HDR | sa.OUT_CLASS, sa.PREC_MIXED)
    ref = be.render_outputs(u, sa.OUT_HDR | sa.OUT_CLASS, sa.PREC_MIXED, refine=sa.REFINE_CROSS64)
    disc = (truth["cls"] == sa.CLASS_DISC) & (lin.cls.ravel()[px] == sa.CLASS_DISC) & \
           (ref.cls.ravel()[px] == sa.CLASS_DISC)
    assert disc.sum() > 1000
    e_lin = float(np.median(rel_err(lin.hdr.reshape(-1, 3)[px][disc], truth["hdr"][disc])))
    e_ref = float(np.median(rel_err(ref.hdr.reshape(-1, 3)[px][disc], truth["hdr"][disc])))
    print(f"disc pixels {int(disc.sum())}: median rel err linear {e_lin:.3g}, refined {e_ref:.3g}", flush=True)
    assert e_ref <= 1e-5 and e_ref <= e_lin / 100.0


def test_refine_argument_checks(be):
    u = headline()
    with pytest.raises(sa.SingError):
        be.render(u, sa.RenderOpts(refine=4))
    with pytest.raises(sa.SingError):
        be.render(u, sa.RenderOpts(refine=sa.REFINE_CROSS64, adapt_tol=1e-6))
    s = sa.scene_defaults()
    s.metric = sa.METRIC_SCHWARZSCHILD
    us = sa.pack_uniforms(s, sa.orbital_camera(30.0, 0.15, 0.0, 60.0), 64, 64, 0.0)
    small = sa.HipBackend(64, 64)
    try:
        with pytest.raises(sa.SingError):
            small.render(us, sa.RenderOpts(refine=sa.REFINE_REPROJECT))
    finally:
        small.shutdown()
