"""Generate tests/golden/cli/ from the reference's CPU driver.

Runs only where /root/reference exists: it builds oracle/_ref/singularity_cli
(the reference's cli/main.cpp with its header-only core and vendored stb,
compiled where they lie, see oracle/Makefile) and records what its
diagnostic modes print or write:

  * kerr-geometry JSON at the reference's spin grid {0, 0.5, 0.94, 0.998}
    (mass 1) and at mass 2, spin 1.2 (kerr_geometry.json: one object per case);
  * photon-orbit CSVs: Schwarzschild r = 3 M for 1 orbit (h 0.01) and 3 orbits
    (h 0.005), Kerr prograde orbits at a/M in {0.3, 0.5, 0.7, 0.9} (h 0.001),
    orbits started off the photon sphere that leave it outwards (r = 3.2 M) or
    fall in and end with the -1 sentinel row (Schwarzschild r = 2.9 M, Kerr
    a/M = 0.6 at r = 2.1 M) (orbit_*.csv, the driver's own files);
  * benchmark total_steps for Schwarzschild and Kerr fans (benchmark.json; the
    timing fields are dropped, the step counts are deterministic).

Usage: python tests/golden/make_cli_golden.py
"""

from __future__ import annotations

import json
import subprocess
import sys
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
OUT = HERE / "cli"
CLI = ROOT / "oracle" / "_ref" / "singularity_cli"

GEOMETRY_CASES = [(1.0, 0.0), (1.0, 0.5), (1.0, 0.94), (1.0, 0.998), (2.0, 1.2)]
ORBIT_CASES = {
    "orbit_schw_1.csv": ["--r-init", "3", "--orbits", "1", "--h-step", "0.01"],
    "orbit_schw_3.csv": ["--r-init", "3", "--orbits", "3", "--h-step", "0.005"],
    "orbit_kerr_a03.csv": ["--spin", "0.3", "--orbits", "1", "--h-step", "0.001"],
    "orbit_kerr_a05.csv": ["--spin", "0.5", "--orbits", "1", "--h-step", "0.001"],
    "orbit_kerr_a07.csv": ["--spin", "0.7", "--orbits", "1", "--h-step", "0.001"],
    "orbit_kerr_a09.csv": ["--spin", "0.9", "--orbits", "1", "--h-step", "0.001"],
    "orbit_schw_off_sphere.csv": ["--r-init", "3.2", "--orbits", "2", "--h-step", "0.01"],
    "orbit_schw_capture.csv": ["--r-init", "2.9", "--orbits", "2", "--h-step", "0.01"],
    "orbit_kerr_capture.csv": ["--spin", "0.6", "--r-init", "2.1", "--orbits", "2", "--h-step", "0.002"],
}
BENCH_CASES = [
    ["--metric", "schw", "--n-rays", "8", "--max-steps", "500"],
    ["--metric", "kerr", "--spin", "0.9", "--n-rays", "8", "--max-steps", "500"],
    ["--metric", "schw"],
    ["--metric", "kerr", "--spin", "0.9"],
    ["--metric", "kerr", "--spin", "0.5", "--n-rays", "64", "--max-steps", "2000", "--b-range", "-3", "7"],
    ["--metric", "schw", "--n-rays", "33", "--max-steps", "1000", "--b-range", "-6", "-4"],
]


def run(*args: str, cwd: Path) -> str:
    r = subprocess.run([str(CLI), *args], capture_output=True, text=True, cwd=cwd, timeout=120)
    if r.returncode != 0:
        raise RuntimeError(f"{args}: {r.stderr}")
    return r.stdout


def main() -> None:
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "ref"], check=True)
    if not CLI.exists():
        sys.exit("the reference tree is absent: nothing to generate")
    OUT.mkdir(exist_ok=True)
    with tempfile.TemporaryDirectory() as tmp:
        t = Path(tmp)
        geometry = []
        for mass, spin in GEOMETRY_CASES:
            run("--mode", "kerr-geometry", "--mass", str(mass), "--spin", str(spin), "--output", "g.json", cwd=t)
            geometry.append({"args": {"mass": mass, "spin": spin}, "json": json.loads((t / "g.json").read_text())})
        (OUT / "kerr_geometry.json").write_text(json.dumps(geometry, indent=1) + "\n")
        for name, args in ORBIT_CASES.items():
            line = run("--mode", "photon-orbit", *args, "--output", name, cwd=t)
            (OUT / name).write_bytes((t / name).read_bytes())
            print(line.strip())
        bench = []
        for args in BENCH_CASES:
            payload = json.loads(run("--mode", "benchmark", *args, cwd=t).strip())
            for k in ("total_ms", "steps_per_s", "mean_step_ns"):
                payload.pop(k)
            bench.append({"args": args, "json": payload})
        (OUT / "benchmark.json").write_text(json.dumps(bench, indent=1) + "\n")
        (OUT / "orbit_args.json").write_text(json.dumps(ORBIT_CASES, indent=1) + "\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
