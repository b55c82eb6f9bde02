"""hip_cli, the drop-in for singularity_cuda_cli (cuda_cli/main.cpp:29-64,113-341):
flags, usage errors and exit codes (2 = usage or scene error, 1 = runtime).
These paths end before any device call, so they run on the CPU; the rendering
paths are in test_gpu_hip_cli.py."""

from __future__ import annotations

import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CLI = ROOT / "singularity_amd" / "hip_cli"


def run(*args, cwd=None):
    if not CLI.exists():
        pytest.fail(f"{CLI} is not built (python -m singularity_amd.build)")
    return subprocess.run([str(CLI), *map(str, args)], capture_output=True, text=True, timeout=60, cwd=cwd)


def test_help_lists_the_reference_flags():
    r = run("--help")
    assert r.returncode == 0
    for flag in ("--scene", "--output", "--res", "--samples-per-pixel", "--camera-distance", "--camera-elevation",
                 "--camera-azimuth", "--fov-y-deg", "--frames", "--output-pattern", "--encode-mp4", "--mp4-fps",
                 "--precision", "--hdr-output", "--class-output", "--bloom", "--adapt-tol", "--turb-seed",
                 "--sampler", "--refine"):
        assert flag in r.stdout, flag


@pytest.mark.parametrize("args,needle", [
    (("--res", "10"), "--res expects WxH"),
    (("--res", "0x5"), "--res too small"),
    (("--samples-per-pixel", "0"), "--samples-per-pixel must be in [1, 1024]"),
    (("--samples-per-pixel", "1025"), "--samples-per-pixel must be in [1, 1024]"),
    (("--frames", "0"), "--frames must be in [1, 100000]"),
    (("--mp4-fps", "0"), "--mp4-fps must be in [1, 240]"),
    (("--precision", "fp16"), "--precision must be fp32, fp64 or mixed"),
    (("--sampler", "stratified"), "--sampler must be halton or grid"),
    (("--refine", "cross64,newton"), "--refine takes cross64 and/or reproject"),
    (("--bogus",), "unknown argument: --bogus"),
    (("--frames", "3"), "--frames > 1 requires --output-pattern"),
    (("--frames", "3", "--output-pattern", "f.png"), "must contain a printf placeholder"),
    (("--encode-mp4", "x.mp4"), "--encode-mp4 requires --frames > 1"),
    (("--scene", "/nonexistent/scene.txt"), "scene load failed (file-not-found"),
])
def test_usage_errors_exit_2(args, needle):
    r = run(*args)
    assert r.returncode == 2, (r.returncode, r.stderr)
    assert needle in r.stderr, r.stderr


def test_scene_file_errors_exit_2(tmp_path):
    bad = tmp_path / "bad.scene"
    bad.write_text("spin = 0.5\nwarp_factor = 9\n")
    r = run("--scene", bad)
    assert r.returncode == 2 and "unknown-key" in r.stderr and "line 2" in r.stderr
    bad.write_text("spin = fast\n")
    r = run("--scene", bad)
    assert r.returncode == 2 and "parse-error" in r.stderr
