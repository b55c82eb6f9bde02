"""singularity_amd — MI355X-native offline null-geodesic renderer.

Python view of the C ABI in ``include/sing_abi.h`` (libsing_hip.so, built
in-tree by ``singularity_amd.build``).  ``HipBackend`` mirrors the reference's
``RenderBackend`` (render/include/render_backend.hpp:107-132): initialize,
shutdown, resize, render_frame, capture_frame, name.  The host helpers mirror
``cuda_cli``'s scene ingestion (cuda_cli/main.cpp:68-109,193-218) and
``pack_uniforms_cuda`` (render/cuda/cuda_backend.cu:45-101).

There is no CPU fallback: every render call goes through the HIP library, and
loading fails loudly when it has not been built.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libsing_hip.so"

# --- constants (include/sing_abi.h) ------------------------------------------
FLAG_DOPPLER = 1 << 0
FLAG_REDSHIFT = 1 << 1
FLAG_DISC = 1 << 2
FLAG_STARFIELD = 1 << 3
FLAG_ADAPTIVE = 1 << 4
METRIC_SCHWARZSCHILD = 0
METRIC_KERR = 1
CLASS_HORIZON, CLASS_ESCAPE, CLASS_DISC, CLASS_BUDGET, CLASS_INITFAIL = range(5)
PREC_FP32, PREC_FP64, PREC_MIXED = 0, 1, 2
OUT_RGBA8, OUT_HDR, OUT_CLASS, OUT_STEPS, OUT_HDR16 = 1, 2, 4, 8, 16
POST_NONE, POST_BLOOM = 0, 1
SAMPLER_HALTON, SAMPLER_GRID = 0, 1
REFINE_CROSS64, REFINE_REPROJECT = 1, 2
SING_OK = 0
ERR_NAMES = {1: "ARG", 2: "HIP", 3: "STATE", 4: "NOMEM", 5: "IO", 6: "PARSE", 7: "KEY"}


class Vec4(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("w", C.c_float)]


class Uniforms(C.Structure):
    """The 160-byte block of shared_shader/uniforms.h:54-95."""

    _fields_ = [
        ("cam_pos", Vec4), ("cam_fwd", Vec4), ("cam_right", Vec4), ("cam_up", Vec4),
        ("mass_M", C.c_float), ("spin_a", C.c_float), ("rs", C.c_float), ("tan_half_fov", C.c_float),
        ("disc_r_inner", C.c_float), ("disc_r_outer", C.c_float), ("disc_peak_T", C.c_float), ("aspect", C.c_float),
        ("h_step", C.c_float), ("escape_r", C.c_float), ("horizon_cut", C.c_float), ("time_sec", C.c_float),
        ("width", C.c_uint32), ("height", C.c_uint32), ("metric_type", C.c_uint32), ("flags", C.c_uint32),
        ("max_steps", C.c_uint32), ("frame_index", C.c_uint32), ("supersample", C.c_uint32), ("pad_b", C.c_uint32),
        ("exposure", C.c_float), ("bloom_threshold", C.c_float), ("bloom_strength", C.c_float),
        ("disc_turbulence", C.c_float),
    ]

    def to_bytes(self) -> bytes:
        return bytes(self)

    def copy(self) -> "Uniforms":
        return Uniforms.from_buffer_copy(bytes(self))


class RenderOpts(C.Structure):
    _fields_ = [
        ("precision", C.c_uint32), ("outputs", C.c_uint32), ("post", C.c_uint32), ("turb_seed", C.c_uint32),
        ("adapt_tol", C.c_float), ("band_width", C.c_float), ("peak_T_scale", C.c_float),
        ("step_batch", C.c_uint32), ("waves_per_cu", C.c_uint32), ("sampler", C.c_uint32),
        ("refine", C.c_uint32), ("overlap_chunks", C.c_uint32), ("chunk_samples", C.c_uint32),
        ("reserved", C.c_uint32 * 3),
    ]


class RenderStats(C.Structure):
    _fields_ = [
        ("kernel_ms", C.c_double), ("trace_ms", C.c_double), ("shade_ms", C.c_double), ("post_ms", C.c_double),
        ("rays", C.c_uint64), ("steps", C.c_uint64), ("fp64_rays", C.c_uint64), ("chunks", C.c_uint64),
        ("rejected", C.c_uint64), ("trace32_ms", C.c_double), ("trace64_ms", C.c_double), ("steps32", C.c_uint64),
        ("steps64", C.c_uint64),
    ]


class DeviceOutputs(C.Structure):
    """Device output pointers of sing_render_noalloc (include/sing_abi.h)."""

    _fields_ = [("rgba", C.c_void_p), ("hdr", C.c_void_p), ("hdr16", C.c_void_p), ("cls", C.c_void_p),
                ("steps", C.c_void_p)]


class Scene(C.Structure):
    """render_backend.hpp:46-82 Scene (the fields the offline path uses)."""

    _fields_ = [
        ("metric", C.c_uint32), ("mass", C.c_float), ("spin_a_over_M", C.c_float), ("disc_inner_M", C.c_float),
        ("disc_outer_M", C.c_float), ("doppler_on", C.c_uint32), ("redshift_on", C.c_uint32),
        ("disc_on", C.c_uint32), ("supersample", C.c_uint32), ("exposure", C.c_float),
        ("bloom_threshold", C.c_float), ("bloom_strength", C.c_float), ("disc_turbulence", C.c_float),
        ("disc_peak_T_K", C.c_float), ("max_steps", C.c_uint32), ("h_step", C.c_float),
        ("escape_r_M", C.c_float), ("horizon_cut_rs", C.c_float),
    ]


class Camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("basis", C.c_float * 9), ("fov_y_radians", C.c_float)]


class SceneConfig(C.Structure):
    _fields_ = [
        ("mass", C.c_float), ("spin", C.c_float), ("disc_r_inner", C.c_float), ("disc_r_outer", C.c_float),
        ("disc_doppler_on", C.c_uint32), ("disc_redshift_on", C.c_uint32), ("camera_fov_deg", C.c_float),
        ("camera_distance", C.c_float), ("camera_elevation", C.c_float), ("h_step", C.c_float),
        ("max_steps", C.c_int32),
    ]


class SingError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{ERR_NAMES.get(code, code)}] {msg}")
        self.code = code


_LIB: C.CDLL | None = None

# name -> (restype, argtypes)
_SIGS = {
    "sing_abi_version": (C.c_int, []),
    "sing_backend_name": (C.c_char_p, []),
    "sing_last_error": (C.c_char_p, [C.c_void_p]),
    "sing_create": (C.c_void_p, [C.c_uint32, C.c_uint32, C.c_int]),
    "sing_destroy": (None, [C.c_void_p]),
    "sing_resize": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32]),
    "sing_render": (C.c_int, [C.c_void_p, C.POINTER(Uniforms), C.POINTER(RenderOpts), C.POINTER(RenderStats)]),
    "sing_render_async": (C.c_int, [C.c_void_p, C.POINTER(Uniforms), C.POINTER(RenderOpts)]),
    "sing_wait": (C.c_int, [C.c_void_p, C.POINTER(RenderStats)]),
    "sing_render_noalloc": (C.c_int, [C.c_int, C.POINTER(Uniforms), C.POINTER(RenderOpts), C.POINTER(DeviceOutputs),
                                      C.c_void_p, C.POINTER(RenderStats)]),
    "sing_capture_rgba8": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "sing_capture_hdr": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "sing_capture_class": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "sing_capture_steps": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "sing_capture_hdr16": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "sing_post_bloom": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "sing_scene_defaults": (None, [C.POINTER(Scene)]),
    "sing_scene_config_defaults": (None, [C.POINTER(SceneConfig)]),
    "sing_parse_scene_config": (
        C.c_int, [C.c_char_p, C.POINTER(SceneConfig), C.POINTER(C.c_int), C.c_char_p, C.c_size_t]),
    "sing_load_scene_config": (
        C.c_int, [C.c_char_p, C.POINTER(SceneConfig), C.POINTER(C.c_int), C.c_char_p, C.c_size_t]),
    "sing_scene_from_config": (None, [C.POINTER(SceneConfig), C.POINTER(Scene)]),
    "sing_orbital_camera": (None, [C.c_float, C.c_float, C.c_float, C.c_float, C.POINTER(Camera)]),
    "sing_pack_uniforms": (
        None, [C.POINTER(Scene), C.POINTER(Camera), C.c_uint32, C.c_uint32, C.c_float, C.POINTER(Uniforms)]),
    "sing_kerr_radius": (C.c_double, [C.c_int, C.c_double, C.c_double]),
}


def lib() -> C.CDLL:
    """Load libsing_hip.so (built in-tree); raises when it is missing."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is not built; run `python -m singularity_amd.build` (no CPU fallback exists)"
            )
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def exported_symbols() -> list[str]:
    return sorted(_SIGS)


# --- host-side scene ingestion ----------------------------------------------
def scene_defaults() -> Scene:
    s = Scene()
    lib().sing_scene_defaults(C.byref(s))
    return s


def scene_config_defaults() -> SceneConfig:
    c = SceneConfig()
    lib().sing_scene_config_defaults(C.byref(c))
    return c


def parse_scene_config(text: str) -> SceneConfig:
    c = scene_config_defaults()
    line = C.c_int(0)
    buf = C.create_string_buffer(512)
    rc = lib().sing_parse_scene_config(text.encode(), C.byref(c), C.byref(line), buf, 512)
    if rc != SING_OK:
        raise SingError(rc, f"scene parse failed (line {line.value}): {buf.value.decode()}")
    return c


def load_scene_config(path: str | Path) -> SceneConfig:
    c = scene_config_defaults()
    line = C.c_int(0)
    buf = C.create_string_buffer(512)
    rc = lib().sing_load_scene_config(str(path).encode(), C.byref(c), C.byref(line), buf, 512)
    if rc != SING_OK:
        raise SingError(rc, f"scene load failed (line {line.value}): {buf.value.decode()}")
    return c


def scene_from_config(cfg: SceneConfig, base: Scene | None = None) -> Scene:
    s = base if base is not None else scene_defaults()
    lib().sing_scene_from_config(C.byref(cfg), C.byref(s))
    return s


def orbital_camera(distance: float = 30.0, elevation: float = 0.15, azimuth: float = 0.0,
                   fov_y_deg: float = 60.0) -> Camera:
    cam = Camera()
    lib().sing_orbital_camera(distance, elevation, azimuth, fov_y_deg, C.byref(cam))
    return cam


def pack_uniforms(scene: Scene, cam: Camera, width: int, height: int, time_sec: float = 0.0) -> Uniforms:
    u = Uniforms()
    lib().sing_pack_uniforms(C.byref(scene), C.byref(cam), width, height, time_sec, C.byref(u))
    return u


def kerr_radius(which: int, mass: float, spin_a: float) -> float:
    return float(lib().sing_kerr_radius(which, mass, spin_a))


def isco_prograde(a_over_M: float, mass: float = 1.0) -> float:
    return kerr_radius(2, mass, a_over_M * mass)


@dataclass
class Frame:
    rgba: np.ndarray | None = None
    hdr: np.ndarray | None = None
    cls: np.ndarray | None = None
    steps: np.ndarray | None = None
    hdr16: np.ndarray | None = None
    stats: RenderStats | None = None


class HipBackend:
    """RenderBackend mirror over the C ABI (render/include/render_backend.hpp:107-132)."""

    def __init__(self, width: int = 1280, height: int = 720, device: int = 0):
        self._ctx = None
        self.width = self.height = 0
        self.frame_time_sec = 0.0
        self.initialize(width, height, device)

    def name(self) -> str:
        return lib().sing_backend_name().decode()

    def initialize(self, width: int, height: int, device: int = 0) -> bool:
        self.shutdown()
        ctx = lib().sing_create(width, height, device)
        if not ctx:
            raise SingError(2, lib().sing_last_error(None).decode())
        self._ctx = ctx
        self.width, self.height = width, height
        self.frame_time_sec = 0.0
        return True

    def shutdown(self) -> None:
        if self._ctx:
            lib().sing_destroy(self._ctx)
        self._ctx = None

    def __del__(self):
        try:
            self.shutdown()
        except Exception:
            pass

    def resize(self, width: int, height: int) -> None:
        self._check(lib().sing_resize(self._ctx, width, height))
        self.width, self.height = width, height

    def _check(self, rc: int) -> None:
        if rc != SING_OK:
            raise SingError(rc, lib().sing_last_error(self._ctx).decode())

    def render(self, u: Uniforms, opts: RenderOpts | None = None) -> RenderStats:
        """One synchronous frame from a packed uniform block."""
        st = RenderStats()
        o = opts if opts is not None else RenderOpts()
        self._check(lib().sing_render(self._ctx, C.byref(u), C.byref(o), C.byref(st)))
        return st

    def render_async(self, u: Uniforms, opts: RenderOpts | None = None) -> None:
        """sing_render_async: enqueue the frame and return; wait() completes it (as does
        any other call on this backend)."""
        o = opts if opts is not None else RenderOpts()
        self._check(lib().sing_render_async(self._ctx, C.byref(u), C.byref(o)))

    def wait(self) -> RenderStats:
        """sing_wait: the stats of the frame submitted by render_async."""
        st = RenderStats()
        self._check(lib().sing_wait(self._ctx, C.byref(st)))
        return st

    def render_frame(self, scene: Scene, camera: Camera, opts: RenderOpts | None = None) -> RenderStats:
        """CudaBackend::render_frame semantics: pack, render, advance the 1/30 s virtual clock."""
        u = pack_uniforms(scene, camera, self.width, self.height, self.frame_time_sec)
        st = self.render(u, opts)
        self.frame_time_sec += np.float32(1.0) / np.float32(30.0)
        self.frame_time_sec = float(np.float32(self.frame_time_sec))
        return st

    def post_bloom(self, u: Uniforms, repeat: int = 1) -> RenderStats:
        """sing_post_bloom: the bloom pass again over the last frame's half-float HDR with
        u's exposure and bloom settings (no re-trace); post_ms = mean per launch."""
        st = RenderStats()
        self._check(lib().sing_post_bloom(self._ctx, C.byref(u), int(repeat), C.byref(st)))
        return st

    def capture(self, outputs: int) -> Frame:
        n = self.width * self.height
        f = Frame()
        if outputs & OUT_RGBA8:
            f.rgba = np.empty((self.height, self.width, 4), np.uint8)
            self._check(lib().sing_capture_rgba8(self._ctx, f.rgba.ctypes.data, f.rgba.nbytes))
        if outputs & OUT_HDR:
            f.hdr = np.empty((self.height, self.width, 3), np.float32)
            self._check(lib().sing_capture_hdr(self._ctx, f.hdr.ctypes.data, f.hdr.nbytes))
        if outputs & OUT_CLASS:
            f.cls = np.empty((self.height, self.width), np.uint8)
            self._check(lib().sing_capture_class(self._ctx, f.cls.ctypes.data, f.cls.nbytes))
        if outputs & OUT_STEPS:
            f.steps = np.empty((self.height, self.width), np.uint32)
            self._check(lib().sing_capture_steps(self._ctx, f.steps.ctypes.data, f.steps.nbytes))
        if outputs & OUT_HDR16:
            f.hdr16 = np.empty((self.height, self.width, 4), np.float16)
            self._check(lib().sing_capture_hdr16(self._ctx, f.hdr16.ctypes.data, f.hdr16.nbytes))
        assert n == self.width * self.height
        return f

    def capture_frame(self) -> np.ndarray:
        return self.capture(OUT_RGBA8).rgba

    def render_outputs(self, u: Uniforms, outputs: int, precision: int = PREC_FP32, **kw) -> Frame:
        o = RenderOpts()
        o.precision = precision
        o.outputs = outputs
        for k, v in kw.items():
            setattr(o, k, v)
        st = self.render(u, o)
        f = self.capture(outputs | (OUT_RGBA8 if o.post else 0))
        f.stats = st
        return f


# --- differentiable render (include/sing_abi.h: sing_diff_*) -------------------
class DiffParams(C.Structure):
    _fields_ = [("spin", C.c_double), ("incl", C.c_double), ("r_in", C.c_double), ("r_out", C.c_double),
                ("t_scale", C.c_double)]

    def as_array(self) -> np.ndarray:
        return np.array([self.spin, self.incl, self.r_in, self.r_out, self.t_scale])

    @staticmethod
    def from_array(a) -> "DiffParams":
        return DiffParams(*[float(v) for v in a])


class DiffCamera(C.Structure):
    _fields_ = [("distance", C.c_double), ("azimuth", C.c_double), ("fov_y_deg", C.c_double)]


class DiffOpts(C.Structure):
    _fields_ = [("precision", C.c_uint32), ("checkpoint_every", C.c_uint32), ("soft_edge", C.c_double),
                ("r_in_isco", C.c_uint32), ("target_per_eval", C.c_uint32), ("reserved", C.c_uint32 * 4)]


class DiffStats(C.Structure):
    _fields_ = [("forward_ms", C.c_double), ("backward_ms", C.c_double), ("rays", C.c_uint64),
                ("steps", C.c_uint64), ("fp64_rays", C.c_uint64), ("reserved", C.c_uint64 * 3)]


_SIGS.update({
    "sing_diff_forward": (C.c_int, [C.c_void_p, C.POINTER(Uniforms), C.POINTER(DiffParams), C.POINTER(DiffCamera),
                                    C.POINTER(DiffOpts), C.c_void_p, C.POINTER(DiffStats)]),
    "sing_diff_loss_grad": (C.c_int, [C.c_void_p, C.POINTER(Uniforms), C.POINTER(DiffParams),
                                      C.POINTER(DiffCamera), C.POINTER(DiffOpts), C.c_void_p, C.c_void_p,
                                      C.POINTER(C.c_double), C.c_void_p, C.c_void_p, C.POINTER(DiffStats)]),
    "sing_diff_loss_grad_batch": (C.c_int, [C.c_void_p, C.POINTER(Uniforms), C.c_uint32, C.c_void_p, C.c_void_p,
                                            C.c_uint32, C.POINTER(DiffOpts), C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_void_p, C.POINTER(DiffStats)]),
    "sing_diff_vjp": (C.c_int, [C.c_void_p, C.POINTER(Uniforms), C.POINTER(DiffParams), C.POINTER(DiffCamera),
                                C.POINTER(DiffOpts), C.c_void_p, C.POINTER(C.c_double), C.c_void_p, C.c_void_p,
                                C.POINTER(DiffStats)]),
})
DIFF_MAX_BATCH = 32


_SIGS["sing_diff_capture_pixels"] = (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p])


def diff_capture_pixels(be: "HipBackend") -> tuple[np.ndarray, np.ndarray]:
    """Per-pixel gradient (H, W, 5) and (steps, crossings, termination, precision) (H, W, 4)
    of the last differentiable call."""
    g = np.empty((be.height, be.width, 5), np.float64)
    m = np.empty((be.height, be.width, 4), np.uint32)
    be._check(lib().sing_diff_capture_pixels(be._ctx, g.ctypes.data, m.ctypes.data))
    return g, m


def diff_base_uniforms(width: int, height: int, starfield: bool = False, peak_T: float = 4500.0,
                       h_step: float = 0.4, max_steps: int = 2500, escape_M: float = 100.0) -> Uniforms:
    """Base block of the differentiable Kerr scene: resolution, integrator and shading settings."""
    s = scene_defaults()
    s.metric = METRIC_KERR
    s.disc_peak_T_K = peak_T
    s.h_step = h_step
    s.max_steps = max_steps
    s.escape_r_M = escape_M
    u = pack_uniforms(s, orbital_camera(), width, height, 0.0)
    if not starfield:
        u.flags &= ~FLAG_STARFIELD
    return u


def _diff_call(be: "HipBackend", fn: str, base: Uniforms, params: DiffParams, cam: DiffCamera,
               opts: DiffOpts | None, *rest):
    L = lib()
    f = getattr(L, fn)
    o = opts if opts is not None else DiffOpts()
    st = DiffStats()
    be._check(f(be._ctx, C.byref(base), C.byref(params), C.byref(cam), C.byref(o), *rest, C.byref(st)))
    return st


def diff_forward(be: "HipBackend", base: Uniforms, params: DiffParams, cam: DiffCamera,
                 opts: DiffOpts | None = None) -> tuple[np.ndarray, DiffStats]:
    hdr = np.empty((be.height, be.width, 3), np.float32)
    st = _diff_call(be, "sing_diff_forward", base, params, cam, opts, hdr.ctypes.data)
    return hdr, st


def diff_loss_grad(be: "HipBackend", base: Uniforms, params: DiffParams, cam: DiffCamera, target: np.ndarray,
                   weights: np.ndarray | None = None, opts: DiffOpts | None = None):
    """Loss = sum w |I - target|^2 and d loss / d(a/M, incl, r_in/M, r_out/M, T-scale)."""
    t = np.ascontiguousarray(target, dtype=np.float32)
    if t.shape != (be.height, be.width, 3):
        raise ValueError(f"target shape {t.shape} != {(be.height, be.width, 3)}")
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
    if w is not None and w.shape != (be.height, be.width):
        raise ValueError(f"weights shape {w.shape} != {(be.height, be.width)}")
    loss = C.c_double(0.0)
    grad = np.zeros(5, np.float64)
    hdr = np.empty((be.height, be.width, 3), np.float32)
    st = _diff_call(be, "sing_diff_loss_grad", base, params, cam, opts, t.ctypes.data,
                    None if w is None else w.ctypes.data, C.byref(loss), grad.ctypes.data, hdr.ctypes.data)
    return loss.value, grad, hdr, st


def diff_vjp(be: "HipBackend", base: Uniforms, params: DiffParams, cam: DiffCamera, dl_dhdr: np.ndarray,
             opts: DiffOpts | None = None):
    """sing_diff_vjp: the backward pass for a caller's own loss.  dl_dhdr (H, W, 3) =
    dLoss/dI; returns (sum dl_dhdr . dI/dparams (5,), the pairing sum dl_dhdr . I,
    radiance I, stats)."""
    g = np.ascontiguousarray(dl_dhdr, dtype=np.float32)
    if g.shape != (be.height, be.width, 3):
        raise ValueError(f"dl_dhdr shape {g.shape} != {(be.height, be.width, 3)}")
    pairing = C.c_double(0.0)
    grad = np.zeros(5, np.float64)
    hdr = np.empty((be.height, be.width, 3), np.float32)
    st = _diff_call(be, "sing_diff_vjp", base, params, cam, opts, g.ctypes.data, C.byref(pairing), grad.ctypes.data,
                    hdr.ctypes.data)
    return grad, pairing.value, hdr, st


def diff_loss_grad_batch(be: "HipBackend", base: Uniforms, params, cams, target: np.ndarray,
                         weights: np.ndarray | None = None, opts: DiffOpts | None = None, want_hdr: bool = False):
    """sing_diff_loss_grad_batch: n parameter points (a sequence of DiffParams or an (n, 5)
    array) in one set of launches; cams is one DiffCamera or one per point; target is one
    (H, W, 3) image for all points or one per point, (n, H, W, 3) (weights likewise (H, W)
    or (n, H, W)).  Returns (losses (n,), gradients (n, 5), radiance (n, H, W, 3) or None,
    stats)."""
    plist = [p if isinstance(p, DiffParams) else DiffParams.from_array(p) for p in params]
    n = len(plist)
    clist = [cams] if isinstance(cams, DiffCamera) else list(cams)
    if n < 1 or n > DIFF_MAX_BATCH or len(clist) not in (1, n):
        raise ValueError(f"{n} parameter points (1..{DIFF_MAX_BATCH}) and {len(clist)} cameras (1 or n)")
    t = np.ascontiguousarray(target, dtype=np.float32)
    per_eval = t.ndim == 4
    img = (be.height, be.width)
    if t.shape != ((n,) + img + (3,) if per_eval else img + (3,)):
        raise ValueError(f"target shape {t.shape} != {img + (3,)} or {(n,) + img + (3,)}")
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
    if w is not None and w.shape != ((n,) + img if per_eval else img):
        raise ValueError(f"weights shape {w.shape} != {(n,) + img if per_eval else img}")
    parr = (DiffParams * n)(*plist)
    carr = (DiffCamera * len(clist))(*clist)
    loss = np.zeros(n, np.float64)
    grad = np.zeros((n, 5), np.float64)
    hdr = np.empty((n, be.height, be.width, 3), np.float32) if want_hdr else None
    o = DiffOpts.from_buffer_copy(bytes(opts)) if opts is not None else DiffOpts()
    o.target_per_eval = 1 if per_eval else 0
    st = DiffStats()
    be._check(lib().sing_diff_loss_grad_batch(be._ctx, C.byref(base), n, C.addressof(parr), C.addressof(carr),
                                              len(clist), C.byref(o), t.ctypes.data,
                                              None if w is None else w.ctypes.data, loss.ctypes.data,
                                              grad.ctypes.data, None if hdr is None else hdr.ctypes.data,
                                              C.byref(st)))
    return loss, grad, hdr, st


def adam_fit(be: "HipBackend", base: Uniforms, start, target: np.ndarray, cam: DiffCamera, steps: int = 200,
             lr=None, free=None, opts: DiffOpts | None = None, beta1: float = 0.9, beta2: float = 0.999,
             clip: float = 2.0, lr_floor: float = 0.003, callback=None):
    """Adam over the 5 scene parameters (entries with free == False stay fixed).

    The learning rate follows a cosine decay from lr to lr_floor * lr.  After 5
    steps each gradient entry is clipped to clip x its running RMS: pixels whose
    rays pass near the photon orbit have steep but finite derivatives, and one
    such spike would otherwise inflate Adam's second moment for hundreds of
    steps.  Returns (the parameters after the last update, history of (loss,
    params, grad) of every evaluated iterate).

    Several views: target (n, H, W, 3) with cam a sequence of n cameras fits one
    parameter point to all of them (loss and gradient summed over the views,
    evaluated in one sing_diff_loss_grad_batch call per step).
    """
    target = np.asarray(target)
    views = target.ndim == 4
    if views and (isinstance(cam, DiffCamera) or len(cam) != target.shape[0]):
        raise ValueError("several target views need one camera per view")
    x = np.array(start, np.float64)
    lr = np.array([0.03, 0.015, 0.075, 0.3, 0.015]) if lr is None else np.asarray(lr, np.float64)
    tied = opts is not None and opts.r_in_isco
    if free is None:
        free = np.array([True, True, not tied, True, True])
    free = np.asarray(free, bool)
    m = np.zeros(5)
    v = np.zeros(5)
    history = []
    for k in range(1, steps + 1):
        if views:
            losses, grads, _, _ = diff_loss_grad_batch(be, base, [x] * target.shape[0], list(cam), target, None, opts)
            loss, g = float(losses.sum()), grads.sum(axis=0)
        else:
            loss, g, _, _ = diff_loss_grad(be, base, DiffParams.from_array(x), cam, target, None, opts)
        if not (np.isfinite(loss) and np.all(np.isfinite(g))):
            raise FloatingPointError(f"non-finite loss/gradient at step {k}: {loss} {g} (params {x})")
        g = np.where(free, g, 0.0)
        history.append((loss, x.copy(), g.copy()))
        if callback is not None:
            callback(k, loss, x, g)
        if k > 5:
            lim = clip * np.sqrt(v / (1 - beta2 ** (k - 1)))
            g = np.clip(g, -lim, lim)
        m = beta1 * m + (1 - beta1) * g
        v = beta2 * v + (1 - beta2) * g * g
        sched = lr_floor + (1 - lr_floor) * 0.5 * (1 + np.cos(np.pi * k / steps))
        step = sched * lr * (m / (1 - beta1**k)) / (np.sqrt(v / (1 - beta2**k)) + 1e-300)
        x = x - np.where(free, step, 0.0)
        x[0] = min(max(x[0], 0.0), 0.99)
    return x, history


# --- batched render (include/sing_abi.h: sing_render_batch) -------------------
class BatchItem(C.Structure):
    _fields_ = [("uniforms", C.POINTER(Uniforms)), ("rgba", C.c_void_p), ("hdr", C.c_void_p), ("cls", C.c_void_p)]


class BatchOpts(C.Structure):
    _fields_ = [("precision", C.c_uint32), ("use_graph", C.c_uint32), ("autotune", C.c_uint32),
                ("step_batch", C.c_uint32), ("waves_per_cu", C.c_uint32), ("adapt_tol", C.c_float),
                ("turb_seed", C.c_uint32), ("record_budget_mib", C.c_uint32), ("rays_per_refill", C.c_uint32),
                ("queue_chunk", C.c_uint32), ("reserved", C.c_uint32 * 2)]


class BatchStats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("copy_ms", C.c_double), ("rays", C.c_uint64), ("steps", C.c_uint64),
                ("fp64_rays", C.c_uint64), ("rejected", C.c_uint64), ("step_batch", C.c_uint32),
                ("waves_per_cu", C.c_uint32), ("graph_replayed", C.c_uint32), ("groups", C.c_uint32),
                ("rays_per_refill", C.c_uint32), ("queue_chunk", C.c_uint32)]


_SIGS["sing_render_batch"] = (C.c_int, [C.c_void_p, C.POINTER(BatchItem), C.c_uint32, C.POINTER(BatchOpts),
                                        C.POINTER(BatchStats)])


def render_batch(be: "HipBackend", scenes: list, opts: BatchOpts | None = None, hdr: bool = False,
                 cls: bool = False) -> tuple[list[Frame], BatchStats]:
    """Render a list of Uniforms in one batched launch; returns one Frame per scene."""
    n = len(scenes)
    items = (BatchItem * n)()
    frames, keep = [], []
    for i, u in enumerate(scenes):
        W, H = u.width, u.height
        f = Frame(rgba=np.empty((H, W, 4), np.uint8), hdr=np.empty((H, W, 3), np.float32) if hdr else None,
                  cls=np.empty((H, W), np.uint8) if cls else None)
        uc = u.copy()
        keep.append(uc)
        items[i].uniforms = C.pointer(uc)
        items[i].rgba = f.rgba.ctypes.data
        items[i].hdr = f.hdr.ctypes.data if hdr else None
        items[i].cls = f.cls.ctypes.data if cls else None
        frames.append(f)
    st = BatchStats()
    o = opts if opts is not None else BatchOpts()
    be._check(lib().sing_render_batch(be._ctx, items, n, C.byref(o), C.byref(st)))
    for f in frames:
        f.stats = st
    return frames, st


# --- the CPU driver's diagnostics (include/sing_abi.h: sing_kerr_geometry, ...) --
class KerrGeometry(C.Structure):
    _fields_ = [(k, C.c_float) for k in (
        "mass", "spin", "spin_over_mass", "outer_horizon", "inner_horizon", "ergosphere_polar",
        "ergosphere_equatorial", "photon_sphere_prograde", "photon_sphere_retrograde", "isco_prograde",
        "isco_retrograde")]

    def as_dict(self) -> dict:
        return {k: float(getattr(self, k)) for k, _ in self._fields_}


class OrbitRow(C.Structure):
    _fields_ = [("step", C.c_int32), ("t", C.c_float), ("r", C.c_float), ("theta", C.c_float), ("phi", C.c_float)]


class FanOpts(C.Structure):
    _fields_ = [("metric", C.c_uint32), ("spin", C.c_float), ("n_rays", C.c_uint32), ("max_steps", C.c_uint32),
                ("b_min", C.c_float), ("b_max", C.c_float), ("reserved", C.c_uint32 * 2)]


class FanStats(C.Structure):
    _fields_ = [("total_steps", C.c_uint64), ("kernel_ms", C.c_double), ("setup_ms", C.c_double),
                ("reserved", C.c_uint64)]


_SIGS["sing_kerr_geometry"] = (C.c_int, [C.c_float, C.c_float, C.POINTER(KerrGeometry)])
_SIGS["sing_photon_orbit"] = (C.c_int, [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                        C.POINTER(OrbitRow), C.c_uint32, C.POINTER(C.c_uint32),
                                        C.POINTER(C.c_double)])
_SIGS["sing_benchmark_fan"] = (C.c_int, [C.c_void_p, C.POINTER(FanOpts), C.c_void_p, C.POINTER(FanStats)])


def kerr_geometry(mass: float, spin: float) -> dict:
    """cli/main.cpp:1415-1476 kerr-geometry scalars (float, host closed forms)."""
    g = KerrGeometry()
    rc = lib().sing_kerr_geometry(mass, spin, C.byref(g))
    if rc != SING_OK:
        raise SingError(rc, lib().sing_last_error(None).decode())
    return g.as_dict()


def photon_orbit(be: "HipBackend", mass: float, spin: float, r_init: float, orbits: float = 1.0,
                 h_step: float = 0.01) -> tuple[np.ndarray, np.ndarray, float]:
    """cli/main.cpp:1279-1410 photon-orbit on the device: (steps int32, rows (n, 4) float32 t r theta phi,
    kernel ms)."""
    bound = C.c_uint32(0)
    rc = lib().sing_photon_orbit(be._ctx, mass, spin, r_init, orbits, h_step, None, 0, C.byref(bound), None)
    if bound.value == 0:
        raise SingError(rc, lib().sing_last_error(be._ctx).decode())
    rows = (OrbitRow * bound.value)()
    n, ms = C.c_uint32(0), C.c_double(0.0)
    be._check(lib().sing_photon_orbit(be._ctx, mass, spin, r_init, orbits, h_step, rows, bound.value, C.byref(n),
                                      C.byref(ms)))
    a = np.frombuffer(rows, dtype=np.dtype([("step", "<i4"), ("v", "<f4", (4,))]), count=n.value)
    return a["step"].copy(), a["v"].copy(), ms.value


def benchmark_fan(be: "HipBackend", metric: int, spin: float = 0.0, n_rays: int = 100, max_steps: int = 5000,
                  b_min: float = -10.0, b_max: float = 10.0) -> tuple[int, np.ndarray, FanStats]:
    """cli/main.cpp:1492-1590 benchmark ray fan on the device: (total steps, per-ray steps, stats)."""
    o = FanOpts(metric=metric, spin=spin, n_rays=n_rays, max_steps=max_steps, b_min=b_min, b_max=b_max)
    per = np.zeros(n_rays, np.uint32)
    st = FanStats()
    be._check(lib().sing_benchmark_fan(be._ctx, C.byref(o), per.ctypes.data, C.byref(st)))
    return int(st.total_steps), per, st
