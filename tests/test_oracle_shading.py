"""Pins of the oracle's shading arithmetic (render/cuda/kernels/geodesic_kernel.cu:77-171,
266-286, 368-393; render/vulkan/shaders/bloom.hlsl:8-77, blit.hlsl:43-58).

* Blackbody and the disc's frequency-shift factors against the reference's own
  host twins (core/include/physics/disc.hpp:122-155, redshift.hpp:44-112),
  through the vectors in tests/golden/ref_shading.npz (written by
  tests/golden/make_golden.py from oracle/_ref) and, where the reference tree
  is present, the live oracle/_ref library.
* The Kerr disc's angular velocity against Bardeen, Press & Teukolsky's closed
  form Omega = sqrt(M) / (r^1.5 + a sqrt(M)).
* Band mask, flux clamp and star colour: known answers derived by hand from
  the .cu text at special points.
* Bloom: closed forms of the Karis average, the soft knee, the 7-tap blur and
  the composite.
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

from helpers import GOLDEN

V = np.load(GOLDEN / "ref_shading.npz")
F32 = np.float32


def ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def bb(orc, t):
    out = np.zeros(3, F32)
    orc.strict().orc_blackbody_f32(float(t), ptr(out))
    return out


# ---------------------------------------------------------------- blackbody
def test_blackbody_equals_reference_host_twin(orc):
    """Kernel variant (geodesic_kernel.cu:77-97) against blackbody_srgb_tanner_helland
    (disc.hpp:122-155) on 4,393 temperatures in [1000, 40000] K.  The kernel folds
    the fit's 1/255 into its constants, so values differ by rounding: at most
    8 x 2^-24 absolute.  One documented difference: at T = 6600 K exactly (t = 66)
    the kernel takes the t <= 66 branch (g from the log fit, r = 1) and the host
    twin the power-law branch."""
    T, ref = V["bb_T"], V["bb_rgb"]
    got = np.array([bb(orc, t) for t in T])
    at66 = T == F32(6600.0)
    assert at66.sum() == 1
    d = np.abs(got - ref).max(-1)
    assert d[~at66].max() <= 8 * 2.0**-24, (T[np.argmax(np.where(at66, 0, d))], d[~at66].max())
    # the branch point: kernel r = 1, g = 0.39008...*ln 66 - 0.63184... clamped to 1, b = 1
    np.testing.assert_array_equal(got[at66][0], [1.0, 1.0, 1.0])
    assert ref[at66][0][1] < 0.99  # host twin's power-law green at t = 66


def test_blackbody_live_reference(orc):
    if not orc.ref_available():
        pytest.skip("oracle/_ref not built (no reference tree)")
    R = orc.ref()
    R.ref_blackbody_th.argtypes = [C.c_float, C.c_void_p]
    rng = np.random.default_rng(7)
    out = np.zeros(3, F32)
    for t in rng.uniform(1000.0, 40000.0, 3000).astype(F32):
        if t == F32(6600.0):
            continue
        R.ref_blackbody_th(float(t), ptr(out))
        assert np.abs(bb(orc, t) - out).max() <= 8 * 2.0**-24, t


def test_blackbody_kernel_clamp():
    """The kernel clamps T to [1000, 40000] K before the fit (geodesic_kernel.cu:78);
    the host twin does not."""
    import oracle

    orc = oracle
    np.testing.assert_array_equal(bb(orc, 500.0), bb(orc, 1000.0))
    np.testing.assert_array_equal(bb(orc, 1.0e6), bb(orc, 40000.0))
    # t <= 19 (1900 K): no blue; t >= 66: blue saturated
    assert bb(orc, 1900.0)[2] == 0.0 and bb(orc, 7000.0)[2] == 1.0


# ---------------------------------------------------------------- redshift
def schw_shift(orc, r, uphi, f_cam, flags=3):
    out = np.zeros(3, F32)
    orc.strict().orc_schw_shift_f32(2.0, 1.0, float(r), float(uphi), float(f_cam), flags, ptr(out))
    return out


def kerr_shift(orc, a, r, cam_r, Lz, flags=3, f64=False):
    out = np.zeros(3, np.float64 if f64 else F32)
    fn = orc.strict().orc_kerr_shift_f64 if f64 else orc.strict().orc_kerr_shift_f32
    fn(1.0, float(a), float(r), float(cam_r), float(Lz), flags, ptr(out))
    return out


def test_schwarzschild_shift_factors_equal_reference(orc):
    """g_grav = schwarzschild_grav_redshift(rs, r, r_obs), omega =
    schwarzschild_keplerian_omega(M, r) and, where the host formula is defined
    (orbital speed v < 1, 1 - v cos psi > 1e-4), g_dop =
    schwarzschild_doppler_factor(rs, r, omega, cos psi): bit for bit (the
    kernel's extra clamps are inactive there).  Inside r = 3 M (v >= 1) the
    kernel's g_dop is 0, the host twin's NaN."""
    r, r_obs, cpsi = V["rs_r"], V["rs_r_obs"], V["rs_cos_psi"]
    n_checked = 0
    for i, x in enumerate(r):
        for j, o in enumerate(r_obs):
            f_cam = F32(1.0) - F32(2.0) / o
            s = schw_shift(orc, x, 0.0, f_cam)
            assert s[0] == V["rs_grav"][i, j], (x, o)
        for k, c in enumerate(cpsi):
            # kernel: cos psi = clamp(-u^phi r, -1, 1)
            uphi = F32(-c) / x
            s = schw_shift(orc, x, uphi, F32(1.0) - F32(2.0) / F32(30.0))
            assert s[2] == V["rs_omega"][i]
            ref = V["rs_doppler"][i, k]
            c_k = np.clip(-uphi * x, F32(-1.0), F32(1.0))
            if c_k != c:
                continue  # u^phi * r does not round back to cos psi: not the same input
            if np.isfinite(ref):
                assert s[1] == ref, (x, c)
                n_checked += 1
            else:
                assert x < 3.0 and s[1] == 0.0
    assert n_checked > 1000
    # flags off: both factors are 1
    np.testing.assert_array_equal(schw_shift(orc, 10.0, 0.01, 0.9, flags=0)[:2], [1.0, 1.0])


def test_intensity_scaling_is_g4_of_the_reference(orc):
    """disc_emit scales by g^4 = intensity_scaling(g) (redshift.hpp:106-109): with
    bands at their turbulence-0 value and a blackbody colour fixed by g T, the
    emission is bb(g T) * flux * 1.09 * g^4."""
    import singularity_amd as sa
    from helpers import scene_uniforms

    u = scene_uniforms(metric=sa.METRIC_SCHWARZSCHILD, spin=0.0, r_in=6.0, r_out=20.0, doppler=False,
                       redshift=False)
    u.disc_turbulence = 0.0
    rgb = np.zeros(3, F32)
    orc.strict().orc_disc_emit_f32(ptr(np.frombuffer(bytes(u), np.uint8).copy()), 6.0, 0.3, 0.9, 0.0, 5000.0,
                                   ptr(rgb))
    band = F32(0.65) + F32(0.55) * F32(0.8)
    np.testing.assert_allclose(rgb, bb(orc, 5000.0) * band, rtol=2e-7)  # g = 1, flux 1 at r_in
    for g in V["rs_g"][::8]:
        assert (F32(g) * F32(g)) * (F32(g) * F32(g)) == V["rs_intensity"][list(V["rs_g"]).index(g)]


def test_kerr_shift_factors(orc):
    """Kerr shading (geodesic_kernel.cu:368-386): omega_K is Bardeen-Press-Teukolsky's
    prograde Omega = sqrt(M)/(r^1.5 + a sqrt(M)) (double to 1e-14, float to 4e-7);
    g_grav is the Schwarzschild lapse ratio (the reference's own choice, equal to
    schwarzschild_grav_redshift(2M, r, r_cam)); g_dop is the Schwarzschild Doppler
    factor at omega_K with cos psi = L_z / r (schwarzschild_doppler_factor)."""
    R = orc.ref() if orc.ref_available() else None
    if R is not None:
        R.ref_grav_redshift.restype = C.c_float
        R.ref_grav_redshift.argtypes = [C.c_float] * 3
        R.ref_doppler_factor.restype = C.c_float
        R.ref_doppler_factor.argtypes = [C.c_float] * 4
    rng = np.random.default_rng(11)
    for _ in range(400):
        a = float(rng.choice([0.0, 0.5, 0.94, 0.998]))
        r = F32(rng.uniform(3.2, 40.0))
        Lz = F32(rng.uniform(-6.0, 6.0))
        d = kerr_shift(orc, a, r, 30.0, Lz, f64=True)
        bpt = 1.0 / (float(r) ** 1.5 + a)
        assert abs(d[2] - bpt) <= 1e-14 * bpt
        f = kerr_shift(orc, a, r, 30.0, Lz)
        assert abs(f[2] - bpt) <= 4e-7 * bpt
        if R is not None:
            assert f[0] == R.ref_grav_redshift(2.0, float(r), 30.0)
            c = float(np.clip(Lz / r, F32(-1.0), F32(1.0)))
            ref = R.ref_doppler_factor(2.0, float(r), float(f[2]), c)
            if np.isfinite(ref) and 1.0 - abs(c) > 1e-3:
                assert f[1] == ref, (a, r, Lz)


def test_disc_flux_clamp(orc):
    """clamp((r_in/r)^3, 0.05, 1.5) (geodesic_kernel.cu:281)."""
    fl = orc.strict().orc_disc_flux_f32
    assert fl(6.0, 6.0) == 1.0
    assert fl(6.0, 12.0) == 0.125
    assert fl(6.0, 5.0) == 1.5  # 1.728 clamped
    assert fl(6.0, 18.0) == F32(0.05)  # 1/27 clamped
    assert abs(fl(6.0, 9.0) - 8.0 / 27.0) < 1e-7


# ---------------------------------------------------------------- band mask
def band(orc, r, phi, rin=6.0, turb=1.0, t=0.0, ph1=0.0, ph2=0.0):
    return orc.strict().orc_band_mask_f32(r, phi, rin, turb, t, ph1, ph2)


def test_band_mask_known_answers(orc):
    """geodesic_kernel.cu:109-121 at hand-derived points:
    * turbulence 0: mix(0.8, ., 0) = 0.8, so 0.65 + 0.55*0.8 = 1.09 for any r, phi, t;
    * r = r_in, phi = 0, t = 0: rho = 0, shear = 4, band2 = 1, band1 = (1 + cos 4)/2,
      so 0.65 + 0.55 * ((1 + cos 4)/2)^2 = 0.666495;
    * the pattern turns at 0.8 omega: (r_in, phi = 0.8, t = 1) equals (r_in, 0, 0);
    * phase ph1 = pi - 4 makes band1 = 0 (value 0.65); ph1 = -4 gives band1 = 1, and
      ph2 = pi then band2 = 0 (0.65), ph2 = 0 band2 = 1 (1.2);
    * below r = 0.1 r_in, rho and omega stop at ln 0.1 and 0.1^-1.5; shear stops at 16
      below (r/r_in)^1.5 = 0.25;
    * range [0.65, 1.2] (the 0 and 1.3 clamps never act)."""
    for r, phi, t in ((6.0, 0.0, 0.0), (13.0, 2.0, 3.0), (40.0, -1.0, 0.5)):
        assert abs(band(orc, r, phi, turb=0.0, t=t) - 1.09) < 2e-7
    b1 = (1.0 + math.cos(4.0)) / 2.0
    assert abs(band(orc, 6.0, 0.0) - (0.65 + 0.55 * b1 * b1)) < 1e-6
    assert band(orc, 6.0, 0.8, t=1.0) == band(orc, 6.0, 0.0)
    assert abs(band(orc, 6.0, 0.0, ph1=math.pi - 4.0) - 0.65) < 1e-6
    assert abs(band(orc, 6.0, 0.0, ph1=-4.0, ph2=math.pi) - 0.65) < 1e-6
    assert abs(band(orc, 6.0, 0.0, ph1=-4.0) - 1.2) < 1e-6
    # r/r_in = 0.05 and 0.1 clamp to the same rho and omega; shear is 16 at both
    assert band(orc, 0.3, 1.0, t=0.7) == band(orc, 0.6, 1.0, t=0.7)
    rho, shear = math.log(0.1), 16.0
    want = 0.65 + 0.55 * ((0.5 + 0.5 * math.cos(10 * rho + 3 * 1.0 + shear)) ** 2
                          * (0.5 + 0.5 * math.cos(22 * rho - 1.7 * 1.0)))
    assert abs(band(orc, 0.6, 1.0) - want) < 2e-6
    rng = np.random.default_rng(3)
    vals = [band(orc, float(r), float(p), turb=float(tb), t=float(tt))
            for r, p, tb, tt in zip(rng.uniform(0.1, 60, 2000), rng.uniform(-7, 7, 2000), rng.uniform(0, 1, 2000),
                                    rng.uniform(0, 10, 2000))]
    assert min(vals) >= 0.65 - 1e-6 and max(vals) <= 1.2 + 1e-6


# ---------------------------------------------------------------- starfield
def star(orc, theta, phi):
    d = np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)], F32)
    out = np.zeros(3, F32)
    orc.strict().orc_starfield_f32(float(d[0]), float(d[1]), float(d[2]), ptr(out))
    return out


def test_star_colour_known_answer(orc):
    """Cell (ix, iy) = (384, 100) of the 768 x 384 equirectangular grid
    (geodesic_kernel.cu:135-171): seed = xorshift32(384 * 0x9E3779B9 + 100) =
    0xB0C0F101.  Low byte 0x01 < 4: a star.  Brightness byte 0xF1 = 241 ->
    (241/255)^4 = 0.797825; hue byte 0xC0 = 192 -> 0.752941; star colour
    (1 - hue/2, 0.8 + 0.8 hue (1 - hue), 0.4 + 0.6 hue) = (0.623529, 0.948836,
    0.851765); pixel = 0.6 * (base + brightness * star), base (0.0015, 0.002,
    0.005) = (0.299380, 0.455394, 0.410735).  The cell centre is at
    theta = (100.5/384) pi, phi = (384.5/768) 2 pi - pi."""
    assert orc.strict().orc_mix32(384, 100) == 0xB0C0F101
    got = star(orc, 100.5 / 384 * math.pi, 384.5 / 768 * 2 * math.pi - math.pi)
    np.testing.assert_allclose(got, [0.299380, 0.455394, 0.410735], atol=2e-6)
    # a cell without a star: 0.6 * base
    assert orc.strict().orc_mix32(1, 2) & 0xFF == 91
    np.testing.assert_allclose(star(orc, 2.5 / 384 * math.pi, 1.5 / 768 * 2 * math.pi - math.pi),
                               [0.0009, 0.0012, 0.003], rtol=1e-6)
    # cell (0, 0) hashes to 0: a "star" of brightness 0, so the sky colour again
    assert orc.strict().orc_mix32(0, 0) == 0
    np.testing.assert_allclose(star(orc, 0.5 / 384 * math.pi, 0.5 / 768 * 2 * math.pi - math.pi),
                               [0.0009, 0.0012, 0.003], rtol=1e-6)


# ---------------------------------------------------------------- bloom
W_BLOOM = np.array([0.071303, 0.131514, 0.189879, 0.214607, 0.189879, 0.131514, 0.071303], F32)
LUM = np.array([0.2126, 0.7152, 0.0722])


def test_bloom_closed_forms(orc):
    """bloom.hlsl:8-77 and blit.hlsl:43-58, stored in float (no half rounding):
    * a uniform 2x2 block's Karis average is its value; the soft knee scales it by
      lum / (lum + threshold): at lum = threshold, half;
    * the 7-tap blur of a constant image is the constant times the weight sum
      0.999999; of a unit impulse, the seven weights in order (clamp-to-edge
      folds the taps that fall off the image onto the edge texel);
    * the composite with zero bloom strength is ACES(hdr * exposure), with a
      constant bloom image it is ACES((hdr + b * strength) * exposure)."""
    thr = 0.8
    c = np.array([0.3, 0.9, 0.5])
    lum = float(LUM @ c)
    H, W = 8, 12
    img = np.broadcast_to(c, (H, W, 3)).astype(F32)
    r = orc.bloom_pipeline(img, thr, 0.0, 1.0, half_store=False)
    np.testing.assert_allclose(r["extract"], np.broadcast_to(c * lum / (lum + thr), (H // 2, W // 2, 3)), rtol=2e-6)
    # at lum = threshold the knee halves the colour
    c2 = c * (thr / lum)
    r2 = orc.bloom_pipeline(np.broadcast_to(c2, (H, W, 3)).astype(F32), thr, 0.0, 1.0, half_store=False)
    np.testing.assert_allclose(r2["extract"][0, 0], c2 / 2.0, rtol=2e-6)
    # Karis weights 1/(1 + lum): three texels at c and one black texel
    blk = np.zeros((2, 2, 3), F32)
    blk[0, 0] = blk[0, 1] = blk[1, 0] = c
    w = 1.0 / (1.0 + lum)
    avg = 3 * w * c / (3 * w + 1.0)
    la = float(LUM @ avg)
    r3 = orc.bloom_pipeline(blk, thr, 0.0, 1.0, half_store=False)
    np.testing.assert_allclose(r3["extract"][0, 0], avg * la / (la + thr), rtol=2e-6)
    # blur of a constant: times the weight sum
    assert abs(float(W_BLOOM.astype(np.float64).sum()) - 0.999999) < 1e-8  # float32 weights
    np.testing.assert_allclose(r["blur_h"], r["extract"] * 0.999999, rtol=2e-6)
    np.testing.assert_allclose(r["blur_v"], r["extract"] * 0.999999 * 0.999999, rtol=3e-6)
    # unit impulse through blur_h / blur_v directly
    L = orc.strict()
    Wb, Hb = 15, 9
    imp = np.zeros((Hb, Wb, 3), F32)
    imp[4, 7, 1] = 1.0
    out = np.zeros_like(imp)
    L.orc_bloom_blur(ptr(imp), Wb, Hb, 1, ptr(out), 0)
    np.testing.assert_array_equal(out[4, 4:11, 1], W_BLOOM[::-1])
    assert out[4, :4].max() == 0.0 and out[4, 11:].max() == 0.0 and out[:4].max() == 0.0
    L.orc_bloom_blur(ptr(imp), Wb, Hb, 0, ptr(out), 0)
    np.testing.assert_array_equal(out[1:8, 7, 1], W_BLOOM[::-1])
    # impulse on the edge texel: the three taps beyond the edge clamp onto it
    edge = np.zeros((Hb, Wb, 3), F32)
    edge[4, 0, 0] = 1.0
    L.orc_bloom_blur(ptr(edge), Wb, Hb, 1, ptr(out), 0)
    assert abs(out[4, 0, 0] - W_BLOOM[:4].astype(np.float64).sum()) < 1e-7
    # composite
    hdr = np.random.default_rng(5).uniform(0.0, 2.0, (H, W, 3)).astype(F32)
    r4 = orc.bloom_pipeline(hdr, 10.0, 0.0, 1.8, half_store=False)

    def aces_u8(x):
        x = x.astype(np.float64)
        y = np.clip((x * (2.51 * x + 0.03)) / (x * (2.43 * x + 0.59) + 0.14), 0.0, 1.0)
        s = np.where(y <= 0.0031308, 12.92 * y, 1.055 * y ** (1 / 2.4) - 0.055)
        return np.clip(s * 255.0 + 0.5, 0, 255).astype(np.uint8)

    assert np.abs(r4["rgba"][..., :3].astype(int) - aces_u8(hdr * F32(1.8)).astype(int)).max() <= 1
    bl = np.zeros((H // 2, W // 2, 3), F32) + F32(0.25)
    rgba = np.zeros((H, W, 4), np.uint8)
    zero = np.zeros((H, W, 3), F32)
    L.orc_bloom_composite(ptr(zero), ptr(bl), W, H, 0.55, 1.8, ptr(rgba))
    assert np.abs(rgba[..., :3].astype(int) - int(aces_u8(np.array([0.25 * 0.55 * 1.8]))[0])).max() <= 1
    assert (rgba[..., 3] == 255).all()
