"""Host side of the C ABI: the library loads, exports every symbol that
include/*.h declares, and the scene ingestion (scene files, Scene
defaults, orbital camera, uniform packing) behaves like the reference's
(core/include/scene/scene_config.hpp, render_backend.hpp:46-99,
cuda_cli/main.cpp:68-109,193-218, cuda_backend.cu:45-101).  No GPU calls."""

from __future__ import annotations

import ctypes as C
import math
import re
from pathlib import Path

import numpy as np
import pytest

import singularity_amd as sa
from helpers import GOLDEN

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols() -> list[str]:
    names = set()
    for h in sorted((ROOT / "include").glob("*.h")):
        names |= set(re.findall(r"\b(sing_[a-z0-9_]+)\s*\(", h.read_text()))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = sa.lib()
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert L.sing_abi_version() == 3
    assert L.sing_backend_name() == b"HIP-gfx950"


def test_create_fails_cleanly_without_device_or_bad_size():
    L = sa.lib()
    assert not L.sing_create(0, 10, 0)
    assert b"width/height" in L.sing_last_error(None)
    # null-context calls return SING_ERR_ARG instead of crashing
    assert L.sing_render(None, None, None, None) == 1
    assert L.sing_render_async(None, None, None) == 1
    assert L.sing_wait(None, None) == 1
    assert L.sing_capture_rgba8(None, None, 0) == 1
    assert L.sing_resize(None, 4, 4) == 1
    assert L.sing_render_batch(None, None, 0, None, None) == 1
    assert L.sing_diff_forward(None, None, None, None, None, None, None) == 1


def test_python_mirror_types_every_declared_symbol():
    # every declared entry point has a ctypes signature (or is a host diagnostic typed by its test)
    typed = set(sa.exported_symbols())
    diag = set(re.findall(r"\b(sing_[a-z0-9_]+)\s*\(", (ROOT / "include" / "sing_diag.h").read_text()))
    assert set(declared_symbols()) - diag <= typed, set(declared_symbols()) - diag - typed


def test_struct_sizes_match_header_layout():
    # sizes the C compiler gives the ABI structs (include/sing_abi.h), mirrored by ctypes
    assert C.sizeof(sa.Uniforms) == 160
    assert C.sizeof(sa.RenderOpts) == 64
    assert C.sizeof(sa.RenderStats) == 104
    assert C.sizeof(sa.DiffOpts) == 40 and C.sizeof(sa.DiffStats) == 64
    assert C.sizeof(sa.DiffParams) == 40 and C.sizeof(sa.DiffCamera) == 24
    assert C.sizeof(sa.BatchItem) == 32 and C.sizeof(sa.BatchOpts) == 48 and C.sizeof(sa.BatchStats) == 72


def test_scene_defaults_match_reference():
    s = sa.scene_defaults()
    assert s.metric == sa.METRIC_KERR
    assert (s.mass, s.spin_a_over_M, s.disc_outer_M) == (1.0, pytest.approx(0.6), 25.0)
    assert s.disc_inner_M == pytest.approx(3.83)
    assert (s.exposure, s.bloom_threshold, s.bloom_strength, s.disc_turbulence) == (
        pytest.approx(1.8), pytest.approx(0.8), pytest.approx(0.55), 0.5)
    assert s.disc_peak_T_K == 4500.0 and s.doppler_on and s.redshift_on and s.disc_on
    c = sa.scene_config_defaults()
    assert (c.mass, c.spin, c.disc_r_inner, c.disc_r_outer, c.max_steps) == (1.0, 0.0, -1.0, 20.0, 5000)
    assert c.h_step == pytest.approx(0.1) and c.camera_fov_deg == 60.0 and c.camera_distance == 30.0


def test_scene_file_parsing_known_answers():
    c = sa.load_scene_config(GOLDEN / "kerr_a094.scene")
    assert c.spin == pytest.approx(0.94) and c.disc_r_inner == pytest.approx(1.7) and c.disc_r_outer == 30.0
    c = sa.parse_scene_config("# comment\nmass = 2.0   # trailing\n\nspin=0.5\ndisc_doppler_on = off\nmax_steps = 77\n")
    assert (c.mass, c.spin, c.disc_doppler_on, c.max_steps) == (2.0, 0.5, 0, 77)
    for word, val in [("true", 1), ("ON", 1), ("yes", 1), ("1", 1), ("False", 0), ("no", 0), ("0", 0)]:
        assert sa.parse_scene_config(f"disc_redshift_on = {word}").disc_redshift_on == val


@pytest.mark.parametrize("text,code", [
    ("spin 0.5", 6), ("spin =", 6), ("= 0.5", 6), ("spin = abc", 6), ("spin = 0.5x", 6),
    ("disc_doppler_on = maybe", 6), ("max_steps = 1.5", 6), ("colour = red", 7),
])
def test_scene_file_errors(text, code):
    with pytest.raises(sa.SingError) as e:
        sa.parse_scene_config("mass = 1\n" + text)
    assert e.value.code == code
    assert "line 2" in str(e.value)


def test_missing_scene_file_is_io_error(tmp_path):
    with pytest.raises(sa.SingError) as e:
        sa.load_scene_config(tmp_path / "nope.scene")
    assert e.value.code == 5


def test_scene_from_config_picks_metric_by_spin():
    s = sa.scene_from_config(sa.parse_scene_config("spin = 0.0"))
    assert s.metric == sa.METRIC_SCHWARZSCHILD
    s = sa.scene_from_config(sa.parse_scene_config("spin = 0.3\ndisc_r_inner = 5\ndisc_r_outer = 9"))
    assert s.metric == sa.METRIC_KERR and s.disc_inner_M == 5.0 and s.disc_outer_M == 9.0
    assert s.exposure == pytest.approx(1.8)  # untouched Scene defaults survive


def test_orbital_camera_is_orthonormal_and_looks_at_origin():
    cam = sa.orbital_camera(30.0, 0.15, 0.7, 60.0)
    pos = np.array(cam.position[:])
    B = np.array(cam.basis[:]).reshape(3, 3)
    assert np.linalg.norm(pos) == pytest.approx(30.0, rel=1e-6)
    np.testing.assert_allclose(B @ B.T, np.eye(3), atol=1e-6)
    np.testing.assert_allclose(-B[2], -pos / np.linalg.norm(pos), atol=1e-6)  # row 2 = -forward
    assert pos[2] == pytest.approx(30.0 * math.sin(0.15), rel=1e-6)
    assert cam.fov_y_radians == pytest.approx(math.radians(60.0), rel=1e-6)


def test_pack_uniforms_reference_semantics(orc):
    s = sa.scene_defaults()
    cam = sa.orbital_camera(30.0, 0.15, 0.0, 60.0)
    u = sa.pack_uniforms(s, cam, 640, 360, 0.25)
    assert u.width == 640 and u.height == 360 and u.aspect == pytest.approx(640 / 360)
    assert u.tan_half_fov == pytest.approx(math.tan(math.radians(30.0)), rel=1e-6)
    assert u.h_step == pytest.approx(0.4) and u.max_steps == 2500 and u.escape_r == 200.0
    assert u.horizon_cut == pytest.approx(1.02 * 2.0) and u.rs == 2.0
    assert u.spin_a == pytest.approx(0.6) and u.metric_type == sa.METRIC_KERR and u.time_sec == 0.25
    assert u.flags == sa.FLAG_STARFIELD | sa.FLAG_DISC | sa.FLAG_DOPPLER | sa.FLAG_REDSHIFT
    assert u.supersample == 1 and u.pad_b == 0 and u.frame_index == 0
    fwd = np.array([u.cam_fwd.x, u.cam_fwd.y, u.cam_fwd.z])
    np.testing.assert_allclose(fwd, -np.array(cam.position[:]) / 30.0, atol=1e-6)
    # the oracle's independent double-precision camera agrees to float rounding
    base = C.create_string_buffer(bytes(u), 160)
    out = C.create_string_buffer(160)
    params = np.array([0.6, math.pi / 2 - 0.15, 3.83, 25.0, 1.0])
    orc.strict().orc_make_uniforms(base, params.ctypes.data, 30.0, 0.0, 60.0, out)
    v = sa.Uniforms.from_buffer_copy(out.raw)
    for f in ("cam_pos", "cam_fwd", "cam_right", "cam_up"):
        a, b = getattr(u, f), getattr(v, f)
        np.testing.assert_allclose([a.x, a.y, a.z], [b.x, b.y, b.z], atol=2e-6)
    assert v.tan_half_fov == pytest.approx(u.tan_half_fov, rel=1e-6)
    # overrides
    s.max_steps, s.h_step, s.escape_r_M, s.horizon_cut_rs, s.supersample = 4000, 0.05, 60.0, 1.01, 0
    s.disc_on = 0
    u = sa.pack_uniforms(s, cam, 8, 8, 0.0)
    assert (u.max_steps, u.escape_r, u.supersample) == (4000, 60.0, 1)
    assert u.h_step == pytest.approx(0.05) and u.horizon_cut == pytest.approx(2.02)
    assert not (u.flags & sa.FLAG_DISC)
