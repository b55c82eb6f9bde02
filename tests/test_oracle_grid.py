"""The oracle's grid sampler (SING_SAMPLER_GRID), the realtime backends'
sub-pixel scheme (render/metal/shaders/geodesic_kernel.metal:416-433 with the
tone map of blit.metal:54-68), pinned on the CPU:

* the pixel radiance is the mean of single rays traced at
  ((sx + 0.5) / ss, (sy + 0.5) / ss), ss = clamp(supersample, 1, 4), checked
  against orc_trace_ray at those offsets;
* the LDR value is ACES of the mean radiance times the exposure (Narkowicz's
  published curve, restated here), not the mean of per-sample ACES;
* one sample per pixel is the pixel centre, the same ray as the Halton
  sampler's spp = 1;
* supersample above 16 is clamped to the 4 x 4 grid.
"""

from __future__ import annotations

import numpy as np

import singularity_amd as sa
from helpers import scene_uniforms


def aces_narkowicz(x: np.ndarray) -> np.ndarray:
    return np.clip((x * (2.51 * x + 0.03)) / (x * (2.43 * x + 0.59) + 0.14), 0.0, 1.0)


def grid(orc):
    ex = orc.Extras()
    ex.grid_sampler = 1.0
    return ex


def test_grid_pixel_is_mean_of_grid_rays(orc):
    u = scene_uniforms(spin=0.94, r_in=sa.isco_prograde(0.94), r_out=25.0, width=10, height=7, spp=3, dist=40.0,
                       elev=0.2, fov=35.0)
    out = orc.render(u, 64, extras=grid(orc))
    assert len(set(out["cls"].tolist())) >= 2  # the frame holds more than one kind of ray
    for pix in (0, 13, 34, 41, 69):
        gx, gy = pix % u.width, pix // u.width
        rgb = np.zeros(3)
        for sy in range(3):
            for sx in range(3):
                ox = np.float32((np.float32(sx) + np.float32(0.5)) * (np.float32(1.0) / np.float32(3.0)))
                oy = np.float32((np.float32(sy) + np.float32(0.5)) * (np.float32(1.0) / np.float32(3.0)))
                rgb += orc.trace_ray(u, gx, gy, float(ox), float(oy), 64)["rgb"]
        hdr = rgb * (1.0 / 9.0)
        np.testing.assert_allclose(out["hdr"][pix], hdr, rtol=1e-14, atol=1e-300)
        np.testing.assert_allclose(out["ldr"][pix], aces_narkowicz(hdr * u.exposure), rtol=1e-12, atol=1e-15)
    # the Halton sampler places its samples elsewhere
    halton = orc.render(u, 64)
    assert np.abs(halton["hdr"] - out["hdr"]).max() > 1e-3


def test_grid_single_sample_is_pixel_centre(orc):
    u = scene_uniforms(spin=0.5, r_in=4.0, r_out=30.0, width=9, height=5, spp=1)
    g = orc.render(u, 64, extras=grid(orc))
    h = orc.render(u, 64)
    for k in ("hdr", "ldr", "rgba", "cls", "steps"):
        np.testing.assert_array_equal(g[k], h[k])


def test_grid_size_is_clamped_supersample(orc):
    """ss = clamp(supersample, 1, 4): supersample 4, 5 and 40 all give the 4 x 4
    grid; supersample 2 gives the 2 x 2 grid."""
    def frame(spp):
        u = scene_uniforms(spin=0.3, r_in=5.0, r_out=15.0, width=6, height=4, spp=spp)
        return u, orc.render(u, 64, extras=grid(orc))

    _, g4 = frame(4)
    for spp in (5, 40):
        _, g = frame(spp)
        np.testing.assert_array_equal(g["hdr"], g4["hdr"])
        np.testing.assert_array_equal(g["steps"], g4["steps"])
    u2, g2 = frame(2)
    for pix in (0, 9, 23):
        gx, gy = pix % u2.width, pix // u2.width
        rays = [orc.trace_ray(u2, gx, gy, (sx + 0.5) / 2, (sy + 0.5) / 2, 64) for sy in range(2) for sx in range(2)]
        assert g2["steps"][pix] == sum(r["steps"] for r in rays)
        np.testing.assert_allclose(g2["hdr"][pix], sum(r["rgb"] for r in rays) / 4, rtol=1e-14, atol=1e-300)
