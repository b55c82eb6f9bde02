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
