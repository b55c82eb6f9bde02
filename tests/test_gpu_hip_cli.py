"""hip_cli end to end on the GPU: the PNG equals the library's RGBA8 frame for
the same scene file and camera (both are the same code path), the orbit
sequence writes one frame per azimuth with cuda_cli's "wrote" lines and its
1/30-s clock, and the HDR (PFM) and class (PGM) side outputs match the
library's HDR and class outputs."""

from __future__ import annotations

import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import singularity_amd as sa
from helpers import GOLDEN, load_png

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
CLI = ROOT / "singularity_amd" / "hip_cli"
W, H = 160, 90


def run(*args, cwd=None):
    print("hip_cli", *args, flush=True)
    return subprocess.run([str(CLI), *map(str, args)], capture_output=True, text=True, timeout=120, cwd=cwd)


def read_pfm(path: Path) -> np.ndarray:
    data = path.read_bytes()
    head, rest = data.split(b"\n", 3)[:3], data.split(b"\n", 3)[3]
    assert head[0] == b"PF"
    w, h = map(int, head[1].split())
    assert float(head[2]) < 0  # little endian
    return np.frombuffer(rest, "<f4").reshape(h, w, 3)[::-1]


def read_pgm(path: Path) -> np.ndarray:
    data = path.read_bytes()
    m = re.match(rb"P5\n(\d+) (\d+)\n255\n", data)
    w, h = int(m.group(1)), int(m.group(2))
    return np.frombuffer(data[m.end():], np.uint8).reshape(h, w)


def library_frame(scene_file, az=0.0, t=0.0, spp=1, prec=sa.PREC_FP32, **kw):
    s = sa.scene_from_config(sa.load_scene_config(scene_file))
    s.supersample = spp
    u = sa.pack_uniforms(s, sa.orbital_camera(30.0, 0.15, az, 60.0), W, H, t)
    be = sa.HipBackend(W, H)
    try:
        return be.render_outputs(u, sa.OUT_RGBA8 | sa.OUT_HDR | sa.OUT_CLASS, prec, **kw)
    finally:
        be.shutdown()


def test_png_equals_library_frame(tmp_path):
    scene = GOLDEN / "kerr_a094.scene"
    out = tmp_path / "a.png"
    r = run("--scene", scene, "--output", out, "--res", f"{W}x{H}", "--samples-per-pixel", 4,
            "--hdr-output", tmp_path / "a.pfm", "--class-output", tmp_path / "a.pgm")
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == f"wrote {out} ({W}x{H}, HIP)"
    f = library_frame(scene, spp=4)
    np.testing.assert_array_equal(load_png(out), f.rgba)
    np.testing.assert_array_equal(read_pfm(tmp_path / "a.pfm"), f.hdr)
    np.testing.assert_array_equal(read_pgm(tmp_path / "a.pgm"), f.cls)


def test_refine_png_equals_library_frame(tmp_path):
    scene = GOLDEN / "kerr_a094.scene"
    out = tmp_path / "r.png"
    r = run("--scene", scene, "--output", out, "--res", f"{W}x{H}", "--precision", "mixed",
            "--refine", "cross64,reproject", "--hdr-output", tmp_path / "r.pfm")
    assert r.returncode == 0, r.stderr
    f = library_frame(scene, prec=sa.PREC_MIXED, refine=sa.REFINE_CROSS64 | sa.REFINE_REPROJECT)
    np.testing.assert_array_equal(load_png(out), f.rgba)
    np.testing.assert_array_equal(read_pfm(tmp_path / "r.pfm"), f.hdr)
    assert not np.array_equal(f.hdr, library_frame(scene, prec=sa.PREC_MIXED).hdr)


def test_grid_sampler_png_equals_library_frame(tmp_path):
    scene = GOLDEN / "kerr_a094.scene"
    out = tmp_path / "g.png"
    r = run("--scene", scene, "--output", out, "--res", f"{W}x{H}", "--samples-per-pixel", 3, "--sampler", "grid",
            "--hdr-output", tmp_path / "g.pfm")
    assert r.returncode == 0, r.stderr
    f = library_frame(scene, spp=3, sampler=sa.SAMPLER_GRID)
    np.testing.assert_array_equal(load_png(out), f.rgba)
    np.testing.assert_array_equal(read_pfm(tmp_path / "g.pfm"), f.hdr)
    assert not np.array_equal(f.hdr, library_frame(scene, spp=3).hdr)


def test_orbit_sequence(tmp_path):
    scene = GOLDEN / "kerr_a05.scene"
    r = run("--scene", scene, "--res", f"{W}x{H}", "--frames", 3, "--output-pattern", tmp_path / "f_%02d.png")
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    assert lines == [f"wrote {tmp_path / f'f_{k:02d}.png'} ({W}x{H}, frame {k + 1}/3, HIP)" for k in range(3)]
    frames = [load_png(tmp_path / f"f_{k:02d}.png") for k in range(3)]
    # frame k: azimuth 2 pi k / 3, virtual clock advanced by 1/30 s per frame
    # (cuda_backend.cu:149-151); both computed in float as the CLI does
    t = np.float32(0.0)
    for k in range(3):
        az = np.float32(6.28318530717958647693) * np.float32(k) / np.float32(3)
        f = library_frame(scene, az=float(az), t=float(t))
        np.testing.assert_array_equal(frames[k], f.rgba)
        t = np.float32(t + np.float32(1.0) / np.float32(30.0))
    assert not np.array_equal(frames[0], frames[1])


def test_precision_and_extra_flags(tmp_path):
    scene = GOLDEN / "kerr_a094.scene"
    for extra in (("--precision", "mixed"), ("--precision", "fp64"), ("--bloom",), ("--adapt-tol", "1e-6"),
                  ("--turb-seed", "5", "--stats")):
        out = tmp_path / "x.png"
        r = run("--scene", scene, "--output", out, "--res", f"{W}x{H}", *extra)
        assert r.returncode == 0, (extra, r.stderr)
        img = load_png(out)
        assert img.shape == (H, W, 4) and np.all(img[..., 3] == 255) and img[..., :3].max() > 0
