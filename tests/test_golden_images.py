"""End-to-end structure against the reference's only end-to-end fixtures:
verification/golden/cuda/{schw,kerr_a05,kerr_a094}.png (copied as data to
tests/golden/cuda/), 256x256, 1 spp, camera (30 M, 0.15 rad, 60 deg).
The reference gates them at pHash distance <= 4 (verification/test_cuda_goldens.py:32,72-80).
The PNGs predate the reference's current shading defaults (4500 K disc, dark
sky), so they pin image structure, not radiance.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import GOLDEN, golden_scene, load_png, phash_distance

NAMES = ["schw.png", "kerr_a05.png", "kerr_a094.png"]


@pytest.mark.parametrize("name", NAMES)
def test_oracle_render_matches_cuda_golden_structure(orc, name):
    u = golden_scene(name)
    o = orc.render(u, 32)
    img = o["rgba"].reshape(256, 256, 4)
    gold = load_png(GOLDEN / "cuda" / name)
    assert phash_distance(img, gold) <= 4


def test_goldens_distinguish_scenes(orc):
    """The hash is discriminative: each render is far from the other scenes' goldens."""
    imgs = {n: orc.render(golden_scene(n), 32)["rgba"].reshape(256, 256, 4) for n in NAMES}
    assert phash_distance(imgs["schw.png"], load_png(GOLDEN / "cuda" / "kerr_a094.png")) > 4
    assert np.any(imgs["kerr_a05.png"] != imgs["kerr_a094.png"])


# ---- the reference's Vulkan goldens (verification/golden/vulkan/*.png, copied as
# data to tests/golden/vulkan/): the realtime pipeline (bloom + ACES blit) at
# camera (30 M, 0.15 rad), 256x256, 1 spp; the reference gates the offline CUDA
# renderer against them at pHash <= 20 (verification/test_backend_equivalence.py:207-256).
def vulkan_scene(name):
    import singularity_amd as sa

    if name == "schw.png":  # Schwarzschild, Scene defaults otherwise (README of the golden set)
        s = sa.scene_defaults()
        s.metric = sa.METRIC_SCHWARZSCHILD
        s.spin_a_over_M = 0.0
    else:
        s = sa.scene_from_config(sa.load_scene_config(GOLDEN / name.replace(".png", ".scene")))
    return sa.pack_uniforms(s, sa.orbital_camera(30.0, 0.15, 0.0, 60.0), 256, 256, 0.0)


@pytest.mark.parametrize("name", NAMES)
def test_oracle_render_matches_vulkan_golden_structure(orc, name):
    img = orc.render(vulkan_scene(name), 32)["rgba"].reshape(256, 256, 4)
    assert phash_distance(img, load_png(GOLDEN / "vulkan" / name)) <= 20


def test_vulkan_goldens_distinguish_metric(orc):
    img = orc.render(vulkan_scene("schw.png"), 32)["rgba"].reshape(256, 256, 4)
    assert phash_distance(img, load_png(GOLDEN / "vulkan" / "kerr_a094.png")) > 20
