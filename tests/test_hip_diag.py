"""hip_diag, the drop-in for the CPU driver's diagnostic modes
(cli/main.cpp:43-104,1269-1621,1625-1834): flags, usage errors and exit codes
(2 = usage or argument error), and the kerr-geometry JSON, which is host
closed forms and must be byte-identical to the driver's file.  These paths end
before any device call, so they run on the CPU; the photon-orbit and
benchmark modes run on the GPU (test_gpu_diag.py)."""

from __future__ import annotations

import json
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
DIAG = ROOT / "singularity_amd" / "hip_diag"
GOLD = Path(__file__).resolve().parent / "golden"
REF_CLI = ROOT / "oracle" / "_ref" / "singularity_cli"


def run(*args, cwd=None):
    if not DIAG.exists():
        pytest.fail(f"{DIAG} is not built (python -m singularity_amd.build)")
    return subprocess.run([str(DIAG), *map(str, args)], capture_output=True, text=True, timeout=60, cwd=cwd)


def test_help_lists_modes_and_flags():
    r = run("--help")
    assert r.returncode == 0
    for word in ("kerr-geometry", "photon-orbit", "benchmark", "--output", "--b-range", "--n-rays", "--max-steps",
                 "--spin", "--mass", "--r-init", "--orbits", "--h-step", "--metric", "--scene"):
        assert word in r.stdout, word
    assert run().returncode == 0  # no mode: help, exit 0 (cli/main.cpp:1754-1757)


@pytest.mark.parametrize("args,needle", [
    (("--mode", "benchmark", "--n-rays", "1"), "--n-rays must be >= 2"),
    (("--mode", "benchmark", "--max-steps", "99"), "--max-steps must be >= 100"),
    (("--mode", "benchmark", "--b-range", "3", "1"), "--b-range MIN MAX requires MIN < MAX"),
    (("--mode", "benchmark", "--metric", "minkowski"), "--metric must be 'schw' or 'kerr'"),
    (("--mode", "benchmark", "--metric", "kerr", "--spin", "1.0"), "--spin must be in [0, 1) for kerr benchmark"),
    (("--mode", "photon-orbit", "--r-init", "2"), "--r-init 2 must be > r_+ = 2"),
    (("--mode", "photon-orbit", "--orbits", "0"), "--orbits must be positive"),
    (("--mode", "photon-orbit", "--h-step", "-1"), "--h-step must be positive"),
    (("--mode", "photon-orbit", "--mass", "0"), "--mass must be positive"),
    (("--mode", "photon-orbit", "--spin", "1.5"), "--spin"),
    (("--mode", "kerr-geometry", "--spin", "1.5"), "--spin 1.500000 out of range [0, M=1.000000]"),
    (("--mode", "kerr-geometry", "--mass", "-1"), "--mass must be positive"),
    (("--mode", "warp-drive",), "unknown mode 'warp-drive'"),
    (("--mode", "cpu-render",), "is not provided here"),
    (("--bogus",), "unknown argument: --bogus"),
    (("--scene", "/nonexistent/x.scene", "--mode", "benchmark"), "--scene load failed (file-not-found"),
])
def test_usage_errors_exit_2(args, needle):
    r = run(*args)
    assert r.returncode == 2, (r.returncode, r.stderr)
    assert needle in r.stderr, r.stderr


def test_kerr_geometry_json_equals_reference(tmp_path):
    for case in json.loads((GOLD / "cli" / "kerr_geometry.json").read_text()):
        m, s = case["args"]["mass"], case["args"]["spin"]
        out = tmp_path / "g.json"
        r = run("--mode", "kerr-geometry", "--mass", m, "--spin", s, "--output", out)
        assert r.returncode == 0, r.stderr
        assert r.stdout.strip() == f"wrote {out} (Kerr geometry at a/M = {s / m:.6g})"
        assert json.loads(out.read_text()) == case["json"], case["args"]


def test_kerr_geometry_default_path_and_scene(tmp_path):
    r = run("--mode", "kerr-geometry", "--scene", GOLD / "kerr_a094.scene", cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    data = json.loads((tmp_path / "kerr_geometry.json").read_text())
    assert abs(data["spin"] - 0.94) < 1e-7


@pytest.mark.skipif(not REF_CLI.exists(), reason="oracle/_ref/singularity_cli is not built (no reference tree)")
@pytest.mark.parametrize("mass,spin", [(1.0, 0.0), (1.0, 0.37), (1.0, 0.9999), (3.5, 2.0), (0.5, 0.5)])
def test_kerr_geometry_file_is_byte_identical_to_live_driver(tmp_path, mass, spin):
    a, b = tmp_path / "a.json", tmp_path / "b.json"
    assert run("--mode", "kerr-geometry", "--mass", mass, "--spin", spin, "--output", a).returncode == 0
    subprocess.run([str(REF_CLI), "--mode", "kerr-geometry", "--mass", str(mass), "--spin", str(spin), "--output",
                    str(b)], check=True, capture_output=True, timeout=60)
    assert a.read_bytes() == b.read_bytes()
