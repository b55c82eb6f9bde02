"""bench.py's command-line contract on the host: one GPU only, and the
constants its roofline line is built from."""
import os
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def _bench(args, env_extra=None):
    env = dict(os.environ)
    env.update(env_extra or {})
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True, text=True,
                          timeout=300, cwd=str(ROOT), env=env)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = _bench(["--gpus", "2"])
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""  # no JSON line for a refused run


def test_distributed_launch_is_refused_with_a_message():
    r = _bench([], {"WORLD_SIZE": "2", "RANK": "0", "LOCAL_RANK": "0"})
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_roofline_constants():
    import singularity_amd.flops as fl

    # the per-step flops are the op-counted reference step's adds and multiplies; divides and trig apart
    assert fl.KERR_STEP_FLOPS == fl.KERR_STEP_OPS["add"] + fl.KERR_STEP_OPS["mul"] == 490
    assert (fl.KERR_STEP_DIVS, fl.KERR_STEP_TRIG) == (53, 9)
    # measured HBM bytes per traced ray: 16-byte init entry + 4-byte record index
    # read, 16-byte record written, plus deferral markers and list entries
    assert 36.0 < fl.TRACE_HBM_BYTES_PER_RAY < 42.0
    assert fl.PEAK_HBM_GBS == 8000.0
