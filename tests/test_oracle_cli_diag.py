"""The oracle's restatement of the CPU driver's diagnostics (cli/main.cpp:
1269-1621) pinned against what the reference driver itself wrote
(tests/golden/cli/, made by tests/golden/make_cli_golden.py from
oracle/_ref/singularity_cli), against the live reference driver where it is
built, and against mathematics that does not go through the closed forms:

* kerr-geometry: every scalar bit-equal to the driver's JSON; horizons are
  roots of Delta, photon orbits roots of Bardeen's circular-photon condition
  r^(3/2) - 3 M r^(1/2) +- 2 a M^(1/2) = 0, ISCOs roots of the BPT polynomial
  r^2 - 6 M r +- 8 a sqrt(M r) - 3 a^2 = 0 (upper sign prograde; in double);
* photon-orbit: every CSV row bit-equal (float), including the -1 sentinel of
  captured orbits; the reference harness's closure, equatorial and
  monotone-phi bars (verification/test_photon_sphere.py) hold; the FP64 orbit
  stays on the sphere far longer than the float one;
* benchmark: total step counts equal the driver's; rays inside the critical
  impact parameter 3 sqrt(3) M are captured before max_steps, rays outside
  run to it.
"""

from __future__ import annotations

import csv
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

GOLD = Path(__file__).resolve().parent / "golden" / "cli"
REF_CLI = Path(__file__).resolve().parents[1] / "oracle" / "_ref" / "singularity_cli"


def read_csv(path: Path) -> np.ndarray:
    with path.open(newline="") as fh:
        rows = [[float(row[k]) for k in ("step", "t", "r", "theta", "phi")] for row in csv.DictReader(fh)]
    return np.array(rows)


def f32(x):
    return np.asarray(x, np.float64).astype(np.float32)


def orbit_args(args: list[str]) -> dict:
    kv = dict(zip(args[::2], args[1::2]))
    return {"spin": float(kv.get("--spin", 0.0)), "r_init": float(kv["--r-init"]) if "--r-init" in kv else None,
            "orbits": float(kv["--orbits"]), "h": float(kv["--h-step"])}


def default_r_init(orc, spin: float) -> float:
    return 3.0 if spin <= 0.0 else float(orc.strict().orc_kerr_closed_form_f32(4, 1.0, spin))


# ---- kerr-geometry -----------------------------------------------------------------
def test_geometry_equals_reference_json(orc):
    for case in json.loads((GOLD / "kerr_geometry.json").read_text()):
        o = orc.kerr_geometry(case["args"]["mass"], case["args"]["spin"])
        for k, v in case["json"].items():
            assert f32(o[k]) == f32(v), (case["args"], k)


@pytest.mark.parametrize("spin", [0.0, 0.3, 0.5, 0.9, 0.94, 0.998])
def test_geometry_roots_of_independent_equations(orc, spin):
    g = orc.kerr_geometry(1.0, spin, precision=64)
    M, a = 1.0, spin
    for r in (g["outer_horizon"], g["inner_horizon"]):
        assert abs(r * r - 2 * M * r + a * a) < 1e-12
    for r, s in ((g["photon_sphere_prograde"], 1.0), (g["photon_sphere_retrograde"], -1.0)):
        assert abs(r ** 1.5 - 3 * M * r ** 0.5 + s * 2 * a * M ** 0.5) < 1e-9
    for r, s in ((g["isco_prograde"], 1.0), (g["isco_retrograde"], -1.0)):
        assert abs(r * r - 6 * M * r + s * 8 * a * np.sqrt(M * r) - 3 * a * a) < 1e-8 * r * r
    assert g["ergosphere_polar"] == g["outer_horizon"] and abs(g["ergosphere_equatorial"] - 2.0) < 1e-12


def test_geometry_argument_errors(orc):
    assert orc.kerr_geometry(0.0, 0.0) is None
    assert orc.kerr_geometry(1.0, 1.5) is None
    assert orc.kerr_geometry(1.0, -0.1) is None


# ---- photon-orbit ------------------------------------------------------------------
def orbit_cases():
    return json.loads((GOLD / "orbit_args.json").read_text())


@pytest.mark.parametrize("name", sorted(json.loads((GOLD / "orbit_args.json").read_text())))
def test_orbit_equals_reference_csv(orc, name):
    args = orbit_args(orbit_cases()[name])
    r0 = args["r_init"] if args["r_init"] is not None else default_r_init(orc, args["spin"])
    rows = orc.photon_orbit(1.0, args["spin"], r0, args["orbits"], args["h"])
    gold = read_csv(GOLD / name)
    assert rows.shape == gold.shape, name
    np.testing.assert_array_equal(rows[:, 0], gold[:, 0])
    np.testing.assert_array_equal(f32(rows[:, 1:]), f32(gold[:, 1:]))


def test_orbit_reference_harness_bars(orc):
    """verification/test_photon_sphere.py: closure 0.5 % after one period,
    theta within 1e-5 of pi/2, phi increasing by 2 pi (4 % slack); 3 periods at
    h = 0.005 close within 1 %; Kerr a in {0.3, 0.5, 0.7, 0.9} at h = 0.001
    close within 0.5 % with theta within 5e-4."""
    rows = orc.photon_orbit(1.0, 0.0, 3.0, 1.0, 0.01)
    assert np.all(rows[:, 0] >= 0)
    assert abs(rows[-1, 2] - rows[0, 2]) / rows[0, 2] <= 0.005
    assert np.max(np.abs(rows[:, 3] - np.pi / 2)) < 1e-5
    assert np.all(np.diff(rows[:, 4]) > 0) and abs(rows[-1, 4] - rows[0, 4] - 2 * np.pi) < 0.04 * np.pi
    rows = orc.photon_orbit(1.0, 0.0, 3.0, 3.0, 0.005)
    assert np.all(rows[:, 0] >= 0) and abs(rows[-1, 2] - rows[0, 2]) / rows[0, 2] < 0.01
    for a in (0.3, 0.5, 0.7, 0.9):
        rows = orc.photon_orbit(1.0, a, default_r_init(orc, a), 1.0, 0.001)
        assert np.all(rows[:, 0] >= 0)
        assert abs(rows[-1, 2] - rows[0, 2]) / rows[0, 2] <= 0.005
        assert np.max(np.abs(rows[:, 3] - np.pi / 2)) < 5e-4
        assert np.all(np.diff(rows[:, 4]) > 0)


def test_orbit_fp64_stays_on_sphere_and_period_is_exact(orc):
    """In double the Schwarzschild orbit closes to 1e-9 after one period; the
    coordinate time of one period is 2 pi 3 sqrt(3) M (the photon sphere's
    angular frequency is 1 / (3 sqrt(3) M))."""
    rows = orc.photon_orbit(1.0, 0.0, 3.0, 1.0, 0.01, precision=64)
    assert np.max(np.abs(rows[:, 2] - 3.0)) < 1e-9
    # phi(t) = t / (3 sqrt(3)) on the sphere
    np.testing.assert_allclose(rows[:, 4], rows[:, 1] / (3.0 * np.sqrt(3.0)), rtol=1e-9, atol=1e-12)


def test_orbit_argument_errors(orc):
    assert orc.photon_orbit(1.0, 0.0, 2.0) is None          # inside 1.001 r_s
    assert orc.photon_orbit(1.0, 0.0, 3.0, orbits=0.0) is None
    assert orc.photon_orbit(1.0, 0.0, 3.0, h=0.0) is None
    assert orc.photon_orbit(1.0, 1.2, 3.0) is None          # a > M
    assert orc.photon_orbit(1.0, 0.6, 2.0) is None          # r = 2M: L_z diverges


# ---- benchmark ---------------------------------------------------------------------
def bench_args(args: list[str]) -> dict:
    kv, i = {}, 0
    while i < len(args):
        if args[i] == "--b-range":
            kv["b_min"], kv["b_max"] = float(args[i + 1]), float(args[i + 2])
            i += 3
        else:
            kv[args[i]] = args[i + 1]
            i += 2
    return {"kerr": kv.get("--metric") == "kerr", "spin": float(kv.get("--spin", 0.0)),
            "n_rays": int(kv.get("--n-rays", 100)), "max_steps": int(kv.get("--max-steps", 5000)),
            "b_min": kv.get("b_min", -10.0), "b_max": kv.get("b_max", 10.0)}


def test_bench_totals_equal_reference(orc):
    for case in json.loads((GOLD / "benchmark.json").read_text()):
        a = bench_args(case["args"])
        total, per = orc.bench_fan(a["kerr"], a["spin"], a["n_rays"], a["max_steps"], a["b_min"], a["b_max"])
        assert total == case["json"]["total_steps"], case["args"]
        assert int(per.sum()) == total


def test_bench_capture_follows_critical_impact_parameter(orc):
    n, steps = 200, 3000
    total, per = orc.bench_fan(False, 0.0, n, steps, -10.0, 10.0)
    b = -10.0 + 20.0 * (np.arange(n) + 0.5) / n
    bc = 3.0 * np.sqrt(3.0)
    far = np.abs(np.abs(b) - bc) > 0.1
    assert np.all(per[far & (np.abs(b) < bc)] < steps)
    assert np.all(per[far & (np.abs(b) > bc)] == steps)


# ---- the live reference driver (built from /root/reference where present) --------
@pytest.mark.skipif(not REF_CLI.exists(), reason="oracle/_ref/singularity_cli is not built (no reference tree)")
def test_live_reference_driver_more_cases(orc, tmp_path):
    rng = np.random.default_rng(11)
    for _ in range(4):
        spin = float(np.round(rng.uniform(0.0, 0.99), 3))
        n, steps = int(rng.integers(2, 40)), int(rng.integers(100, 1500))
        lo = float(np.round(rng.uniform(-9, 4), 2))
        hi = float(np.round(lo + rng.uniform(0.5, 8), 2))
        for metric in ("schw", "kerr"):
            r = subprocess.run([str(REF_CLI), "--mode", "benchmark", "--metric", metric, "--spin", str(spin),
                                "--n-rays", str(n), "--max-steps", str(steps), "--b-range", str(lo), str(hi)],
                               capture_output=True, text=True, timeout=60)
            ref_total = json.loads(r.stdout)["total_steps"]
            total, _ = orc.bench_fan(metric == "kerr", spin, n, steps, lo, hi)
            assert total == ref_total, (metric, spin, n, steps, lo, hi)
        out = tmp_path / "o.csv"
        subprocess.run([str(REF_CLI), "--mode", "photon-orbit", "--spin", str(spin), "--orbits", "0.5",
                        "--h-step", "0.004", "--output", str(out)], check=True, capture_output=True, timeout=60)
        rows = orc.photon_orbit(1.0, spin, default_r_init(orc, spin) if spin > 0 else 3.0, 0.5, 0.004)
        gold = read_csv(out)
        assert rows.shape == gold.shape
        np.testing.assert_array_equal(f32(rows[:, 1:]), f32(gold[:, 1:]))
