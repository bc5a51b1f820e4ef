"""Closed-form pins of the oracle: Sobel, descriptor byte map, prior table,
ratio test, grid arithmetic (reference: src/Kernels/KerDescriptor.cu,
KerSupportMatches.cu:179, KerGrid.cu:33-36, KerComputeDisparity.cu:249-271)."""
import numpy as np


def _sobel_int(img):
    """128 + clamp(S >> 2, -128, 127): the integer form of the reference's
    float (uint8_t)(min(max(-128, S/4), 127) + 128) (KerDescriptor.cu:40-41)."""
    I = img.astype(np.int64)
    c = lambda du, dv: I[1 + dv:I.shape[0] - 1 + dv, 1 + du:I.shape[1] - 1 + du]
    su = c(-1, -1) + 2 * c(-1, 0) + c(-1, 1) - c(1, -1) - 2 * c(1, 0) - c(1, 1)
    sv = c(-1, -1) + 2 * c(0, -1) + c(1, -1) - c(-1, 1) - 2 * c(0, 1) - c(1, 1)
    du = np.zeros(img.shape, np.uint8)
    dv = np.zeros(img.shape, np.uint8)
    du[1:-1, 1:-1] = 128 + np.clip(su >> 2, -128, 127)
    dv[1:-1, 1:-1] = 128 + np.clip(sv >> 2, -128, 127)
    return du, dv


def test_sobel_float_form_equals_integer_form_exhaustively():
    s = np.arange(-1020, 1021)
    f = np.minimum(np.maximum(np.float32(-128), s.astype(np.float32) / np.float32(4)), np.float32(127))
    ref = (f + np.float32(128)).astype(np.uint8)  # truncation of a non-negative float
    assert np.array_equal(ref, (128 + np.clip(s >> 2, -128, 127)).astype(np.uint8))


def test_sobel_matches_closed_form(oracle):
    rng = np.random.default_rng(0)
    for H, W in [(7, 9), (40, 57), (3, 3)]:
        img = rng.integers(0, 256, (H, W), dtype=np.uint8)
        du, dv = oracle.sobel(img)
        edu, edv = _sobel_int(img)
        assert np.array_equal(du, edu) and np.array_equal(dv, edv)


def test_sobel_signs_on_ramps(oracle):
    """du is left minus right, dv top minus bottom (KerDescriptor.cu:25-37)."""
    u = np.arange(32)
    ramp_u = np.tile(2 * u, (16, 1)).astype(np.uint8)      # I = 2u: S_u = 4*(2(u-1) - 2(u+1)) = -16
    du, dv = oracle.sobel(ramp_u)
    assert (du[1:-1, 1:-1] == 124).all() and (dv[1:-1, 1:-1] == 128).all()
    ramp_v = ramp_u.T.copy()                                # I = 2v: S_v = -16
    du, dv = oracle.sobel(ramp_v)
    assert (dv[1:-1, 1:-1] == 124).all() and (du[1:-1, 1:-1] == 128).all()


# KerDescriptor.cu:80-95: (plane, du, dv) of the 16 descriptor bytes
BYTE_MAP = [("u", 0, -2), ("u", -2, -1), ("u", 0, -1), ("u", 2, -1), ("u", -1, 0), ("u", 0, 0),
            ("u", 0, 0), ("u", 1, 0), ("u", -2, 1), ("u", 0, 1), ("u", 2, 1), ("u", 0, 2),
            ("v", 0, -1), ("v", -1, 0), ("v", 1, 0), ("v", 0, 1)]


def test_descriptor_byte_map(oracle):
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, (23, 31), dtype=np.uint8)
    desc = oracle.descriptor(img)
    du, dv = _sobel_int(img)
    H, W = img.shape
    exp = np.zeros((H, W, 16), np.uint8)
    for k, (pl, x, y) in enumerate(BYTE_MAP):
        src = du if pl == "u" else dv
        exp[3:H - 3, 3:W - 3, k] = src[3 + y:H - 3 + y, 3 + x:W - 3 + x]
    assert np.array_equal(desc, exp)


def test_constant_image_descriptor_is_flat(oracle):
    desc = oracle.descriptor(np.full((20, 20), 99, np.uint8))
    assert (desc[3:-3, 3:-3] == 128).all() and (desc[:3] == 0).all() and (desc[:, -3:] == 0).all()


def test_prior_table(oracle):
    """P(delta) = (int)((-log(gamma + exp(-delta^2/2sigma^2)) + log gamma)/beta) at the
    effective parameters gamma 3, sigma 1, beta 0.02: -14.38, -9.21, -2.21 truncate to
    -14, -9, -2 (each > 0.2 from an integer, so any libm agrees)."""
    prm = oracle.params()
    assert oracle.plane_radius(prm) == 2
    assert [oracle.prior(prm, k) for k in range(3)] == [-14, -9, -2]
    exact = [(-np.log(3 + np.exp(-k * k / 2.0)) + np.log(3)) / 0.02 for k in range(3)]
    assert [int(x) for x in exact] == [-14, -9, -2]
    assert min(abs(x - round(x)) for x in exact) > 0.2


def test_ratio_test_float_equals_reference_double_exhaustively():
    """(float)m1 < 0.85 * (float)m2 is evaluated in double in the reference
    (KerSupportMatches.cu:179, SUPPORT_THRESHOLD 0.85); the oracle and kernels
    evaluate m1 < 0.85f * (float)m2 in float.  For integer m1 the accepted set
    is {m1 < t}, so the two agree for every (m1, m2) in [0, 16320]^2 iff
    ceil(t) agrees for every m2."""
    m2 = np.arange(0, 16321)
    t_double = 0.85 * m2.astype(np.float64)
    t_float = np.float32(0.85) * m2.astype(np.float32)
    assert np.array_equal(np.ceil(t_double), np.ceil(t_float.astype(np.float64)))
    # the tempting alternative (double)0.85f differs somewhere
    t_bad = np.float64(np.float32(0.85)) * m2
    assert not np.array_equal(np.ceil(t_double), np.ceil(t_bad))


def test_float_floor_division_equals_integer_division():
    """floor((float)x / (float)g) == floor(x / g) (KerGrid.cu:33-36,
    KerComputeDisparity.cu:238-239) for |x| < 2^20 and the grid sizes allowed."""
    x = np.arange(-(1 << 20), 1 << 20, dtype=np.int64)
    for g in (2, 5, 16, 20, 37, 64):
        f = np.floor(x.astype(np.float32) / np.float32(g)).astype(np.int64)
        assert np.array_equal(f, np.floor_divide(x, g)), g


def test_grid_dims(oracle):
    prm = oracle.params()
    for W, H in [(450, 375), (1242, 375), (2816, 1880), (40, 20), (41, 21)]:
        assert oracle.grid_dims(W, H, prm) == (-(-W // 20), -(-H // 20))
