"""Scene builders and comparison metrics shared by the tests.

Uniform blocks are built with the product's host functions (pack_uniforms,
orbital_camera); the same 160 bytes feed the HIP path and the oracle.
"""

from __future__ import annotations

import math
from pathlib import Path

import numpy as np

import singularity_amd as sa

GOLDEN = Path(__file__).resolve().parent / "golden"


def scene_uniforms(metric: int = sa.METRIC_KERR, spin: float = 0.6, r_in: float = 3.83, r_out: float = 25.0,
                   width: int = 256, height: int = 256, spp: int = 1, dist: float = 30.0, elev: float = 0.15,
                   az: float = 0.0, fov: float = 60.0, h: float = 0.0, max_steps: int = 0, escape_M: float = 0.0,
                   horizon_rs: float = 0.0, peak_T: float = 4500.0, time_sec: float = 0.0,
                   disc: bool = True, doppler: bool = True, redshift: bool = True) -> sa.Uniforms:
    s = sa.scene_defaults()
    s.metric = metric
    s.spin_a_over_M = spin
    s.disc_inner_M = r_in
    s.disc_outer_M = r_out
    s.supersample = spp
    s.h_step = h
    s.max_steps = max_steps
    s.escape_r_M = escape_M
    s.horizon_cut_rs = horizon_rs
    s.disc_peak_T_K = peak_T
    s.disc_on = int(disc)
    s.doppler_on = int(doppler)
    s.redshift_on = int(redshift)
    return sa.pack_uniforms(s, sa.orbital_camera(dist, elev, az, fov), width, height, time_sec)


def golden_scene(name: str, res: int = 256) -> sa.Uniforms:
    """The scenes of verification/golden/cuda/*.png (verification/test_cuda_goldens.py:36-40).

    schw.png is the Schwarzschild scene with the disc at the a=0 ISCO (6 M):
    its golden README row says Schwarzschild a/M = 0, and the oracle reproduces
    it at pHash distance 2 in that configuration."""
    if name == "schw.png":
        return scene_uniforms(metric=sa.METRIC_SCHWARZSCHILD, spin=0.0, r_in=6.0, r_out=25.0, width=res, height=res)
    cfg = sa.load_scene_config(GOLDEN / ("kerr_a05.scene" if name == "kerr_a05.png" else "kerr_a094.scene"))
    s = sa.scene_from_config(cfg)
    s.supersample = 1
    return sa.pack_uniforms(s, sa.orbital_camera(30.0, 0.15, 0.0, 60.0), res, res, 0.0)


def config1(width: int = 256, height: int = 256) -> sa.Uniforms:
    """Schwarzschild 256x256, 1 spp, camera r=30 inclination 85 deg FOV 40, disc [6, 20],
    h=0.05, <=4000 steps, escape 60 M, horizon 1.01 r_h."""
    return scene_uniforms(metric=sa.METRIC_SCHWARZSCHILD, spin=0.0, r_in=6.0, r_out=20.0, width=width,
                          height=height, elev=math.radians(5.0), fov=40.0, h=0.05, max_steps=4000, escape_M=60.0,
                          horizon_rs=1.01)


def config2(width: int = 1920, height: int = 1080, spp: int = 16) -> sa.Uniforms:
    """Kerr a=0.94, camera r=40 inclination 80 deg FOV 30, disc [ISCO, 25]."""
    return scene_uniforms(metric=sa.METRIC_KERR, spin=0.94, r_in=sa.isco_prograde(0.94), r_out=25.0, width=width,
                          height=height, spp=spp, dist=40.0, elev=math.radians(10.0), fov=30.0)


def config3(width: int = 3840, height: int = 2160, spp: int = 256) -> sa.Uniforms:
    """Kerr a=0.998, inclination 60 deg, disc [ISCO, 30], cuda_cli camera defaults (30 M, FOV 60)."""
    return scene_uniforms(metric=sa.METRIC_KERR, spin=0.998, r_in=sa.isco_prograde(0.998), r_out=30.0, width=width,
                          height=height, spp=spp, dist=30.0, elev=math.radians(30.0), fov=60.0)


def headline(width: int = 3840, height: int = 2160, spp: int = 256) -> sa.Uniforms:
    """scenes/kerr_a094.scene at cuda_cli defaults: the reference's published 4K/256-spp render."""
    cfg = sa.load_scene_config(GOLDEN / "kerr_a094.scene")
    s = sa.scene_from_config(cfg)
    s.supersample = spp
    return sa.pack_uniforms(s, sa.orbital_camera(30.0, 0.15, 0.0, 60.0), width, height, 0.0)


def phash(rgb: np.ndarray) -> np.ndarray:
    """64-bit DCT perceptual hash (the imagehash.phash algorithm, hash_size 8)."""
    import scipy.fftpack as F
    from PIL import Image

    im = Image.fromarray(np.ascontiguousarray(rgb[..., :3])).convert("L").resize((32, 32), Image.Resampling.LANCZOS)
    p = np.asarray(im, dtype=np.float64)
    d = F.dct(F.dct(p, axis=0), axis=1)[:8, :8]
    return (d > np.median(d)).flatten()


def phash_distance(a: np.ndarray, b: np.ndarray) -> int:
    return int((phash(a) != phash(b)).sum())


def load_png(path: Path) -> np.ndarray:
    from PIL import Image

    return np.asarray(Image.open(path).convert("RGBA"))


def rel_err(a: np.ndarray, b: np.ndarray, floor: float = 1e-6) -> np.ndarray:
    """Per-pixel max-channel relative error of RGB radiance."""
    a = a.reshape(-1, 3).astype(np.float64)
    b = b.reshape(-1, 3).astype(np.float64)
    return np.abs(a - b).max(-1) / np.maximum(np.abs(b).max(-1), floor)
