"""Pins of the C oracle's single functions by what the mathematics fixes.

Each check is chosen so that a plausible slip -- a transposed operand, a wrong
sign or index, a dropped term, a mis-clamped coordinate -- fails it:
  * inverse transforms against their matrix forms (8-338..8-354, 8-355..8-374),
    at QPs where every intermediate shift is exact;
  * DC-only blocks give flat residuals;
  * luma / chroma interpolation against whole-plane numpy restatements of the
    6-tap / bilinear filters with edge replication;
  * the deblocking line filter: no-op conditions, bS-4 mirror symmetry, the
    opposite moves of p0 / q0 below bS 4, and
    hand-derived known answers of 8-460..8-485.
The end-to-end pins (every picture against the reference decoder) are in
test_oracle_golden.py.
"""
import numpy as np
import pytest

from oracle import refcpu

# Table 8-13 zig-zag (scan index -> raster position) and 8x8 zig-zag, used to place
# raster test matrices into scan order.
ZZ4 = [0, 1, 4, 8, 5, 2, 3, 6, 9, 12, 13, 10, 7, 11, 14, 15]
ZZ8 = [0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6, 7, 14,
       21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60,
       61, 54, 47, 55, 62, 63]
# normAdjust4x4 / normAdjust8x8 columns (8-315 / 8-318); flat weight scale 16.
V4 = [[10, 16, 13], [11, 18, 14], [13, 20, 16], [14, 23, 18], [16, 25, 20], [18, 29, 23]]
V8 = [[20, 18, 32, 19, 25, 24], [22, 19, 35, 21, 28, 26], [26, 23, 42, 24, 33, 31], [28, 25, 45, 26, 35, 33],
      [32, 28, 51, 30, 40, 38], [36, 32, 58, 34, 46, 43]]


def scale4(m):
    s = np.zeros((4, 4), np.int64)
    for i in range(4):
        for j in range(4):
            s[i, j] = 16 * V4[m][0 if (i % 2 == 0 and j % 2 == 0) else (1 if (i % 2 and j % 2) else 2)]
    return s


def scale8(m):
    s = np.zeros((8, 8), np.int64)
    for i in range(8):
        for j in range(8):
            if i % 4 == 0 and j % 4 == 0:
                c = 0
            elif i % 2 and j % 2:
                c = 1
            elif i % 4 == 2 and j % 4 == 2:
                c = 2
            elif (i % 4 == 0 and j % 2) or (i % 2 and j % 4 == 0):
                c = 3
            elif (i % 4 == 0 and j % 4 == 2) or (i % 4 == 2 and j % 4 == 0):
                c = 4
            else:
                c = 5
            s[i, j] = 16 * V8[m][c]
    return s


# Inverse transform matrices: output sample k = sum_n T[k, n] * coefficient n.
T4 = np.array([[1, 1, 1, 0.5], [1, 0.5, -1, -1], [1, -0.5, -1, 1], [1, -1, 1, -0.5]])
T8 = np.array([
    [1, 12 / 8, 1, 10 / 8, 1, 6 / 8, 0.5, 3 / 8],
    [1, 10 / 8, 0.5, -3 / 8, -1, -12 / 8, -1, -6 / 8],
    [1, 6 / 8, -0.5, -12 / 8, -1, 3 / 8, 1, 10 / 8],
    [1, 3 / 8, -1, -6 / 8, 1, 10 / 8, -0.5, -12 / 8],
    [1, -3 / 8, -1, 6 / 8, 1, -10 / 8, -0.5, 12 / 8],
    [1, -6 / 8, -0.5, 12 / 8, -1, -3 / 8, 1, -10 / 8],
    [1, -10 / 8, 0.5, 3 / 8, -1, 12 / 8, -1, 6 / 8],
    [1, -12 / 8, 1, -10 / 8, 1, -6 / 8, 0.5, -3 / 8]])


def to_scan(mat, zz):
    flat = np.asarray(mat).reshape(-1)
    return np.array([flat[zz[k]] for k in range(len(zz))], np.int32)


@pytest.mark.parametrize("qp", [24, 29, 30, 35, 41, 51])
def test_residual4x4_matrix_form(qp):
    rng = np.random.default_rng(qp)
    for _ in range(50):
        c = rng.integers(-40, 41, (4, 4))
        d = (c * scale4(qp % 6)) << (qp // 6 - 4)  # exact: every d is a multiple of 16
        want = np.floor((T4 @ d @ T4.T + 32) / 64).astype(np.int64)
        got = refcpu.residual4x4(to_scan(c, ZZ4), qp)
        np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("qp", [48, 49, 50, 51])
def test_residual8x8_matrix_form(qp):
    rng = np.random.default_rng(qp)
    for _ in range(30):
        c = rng.integers(-20, 21, (8, 8))
        d = (c * scale8(qp % 6)) << (qp // 6 - 6)  # exact: every d is a multiple of 64
        want = np.floor((T8 @ d @ T8.T + 32) / 64).astype(np.int64)
        got = refcpu.residual8x8(to_scan(c, ZZ8), qp)
        np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("qp", range(0, 52, 3))
def test_dc_only_blocks_are_flat(qp):
    for lv in (-7, -1, 1, 3, 25):
        s4 = np.zeros(16, np.int32)
        s4[0] = lv
        r4 = refcpu.residual4x4(s4, qp)
        assert (r4 == r4[0, 0]).all()
        s8 = np.zeros(64, np.int32)
        s8[0] = lv
        r8 = refcpu.residual8x8(s8, qp)
        assert (r8 == r8[0, 0]).all()
        # DC passed through unscaled (Intra16x16 / chroma): r = (dc + 32) >> 6
        r4p = refcpu.residual4x4(np.array([lv * 64] + [0] * 15, np.int32), qp, dc_pass=1)
        assert (r4p == (lv * 64 + 32) >> 6).all()


def test_residual_field_scan_differs_from_zigzag():
    """A sanity check only: the field scans themselves are pinned end to end by the MBAFF
    golden cases (field macroblocks; mbaff_t8_pcm has 8x8 transforms) in
    test_oracle_golden.py."""
    s = np.zeros(16, np.int32)
    s[2] = 10  # zig-zag puts scan 2 at (1,0), field scan at (0,1) -> transposed residuals
    a = refcpu.residual4x4(s, 30, field=0)
    b = refcpu.residual4x4(s, 30, field=1)
    assert not np.array_equal(a, b)


HAD4 = np.array([[1, 1, 1, 1], [1, 1, -1, -1], [1, -1, -1, 1], [1, -1, 1, -1]])


@pytest.mark.parametrize("qp", [36, 40, 45, 51])
def test_luma_dc_matrix_form(qp):
    rng = np.random.default_rng(qp)
    for _ in range(50):
        c = rng.integers(-30, 31, (4, 4))
        f = HAD4 @ c @ HAD4
        want = (f * 16 * V4[qp % 6][0]) << (qp // 6 - 6)
        np.testing.assert_array_equal(refcpu.luma_dc(to_scan(c, ZZ4), qp), want)


@pytest.mark.parametrize("qpc", [0, 10, 23, 29, 39])
def test_chroma_dc_matrix_form(qpc):
    rng = np.random.default_rng(qpc)
    H2 = np.array([[1, 1], [1, -1]])
    for _ in range(50):
        c = rng.integers(-30, 31, (2, 2))
        f = (H2 @ c @ H2).reshape(-1)
        want = ((f * 16 * V4[qpc % 6][0]) << (qpc // 6)) >> 5
        np.testing.assert_array_equal(refcpu.chroma_dc(c.reshape(-1), qpc), want)


def test_chroma_qp_mapping_invariants():
    prev = None
    for qpi in range(52):
        v = refcpu.chroma_qp(qpi, 0)
        if qpi < 30:
            assert v == qpi
        assert v <= qpi
        if prev is not None:
            assert 0 <= v - prev <= 1
        prev = v
    assert refcpu.chroma_qp(51, 0) == 39
    assert refcpu.chroma_qp(51, 12) == 39 and refcpu.chroma_qp(0, -12) == 0  # qPI clipped to [0, 51]
    assert refcpu.chroma_qp(20, 5) == 25


# ------------------------------------------------------------- interpolation
def _pad(ref, n):
    return np.pad(ref.astype(np.int64), n, mode="edge")


def luma_plane_numpy(ref, xf, yf):
    """All prediction samples of a plane for one fractional offset (8.4.2.2.1 with
    replicated borders), computed over the whole plane at once."""
    P = _pad(ref, 8)
    h_, w_ = ref.shape
    tap = np.array([1, -5, 20, 20, -5, 1])

    def at(dx, dy):  # full samples shifted by (dx, dy)
        return P[8 + dy:8 + dy + h_, 8 + dx:8 + dx + w_]

    def hfir(dy, base=0):  # horizontal 6-tap between x+base and x+base+1, row shift dy
        return sum(tap[k] * at(base + k - 2, dy) for k in range(6))

    def vfir(dx):
        return sum(tap[k] * at(dx, k - 2) for k in range(6))

    clip = lambda v: np.clip(v, 0, 255)
    G, H, M = at(0, 0), at(1, 0), at(0, 1)
    b = clip((hfir(0) + 16) >> 5)
    h = clip((vfir(0) + 16) >> 5)
    s = clip((hfir(1) + 16) >> 5)
    m = clip((vfir(1) + 16) >> 5)
    j1 = sum(tap[k] * hfir(k - 2) for k in range(6))
    j = clip((j1 + 512) >> 10)
    avg = lambda a, c: (a + c + 1) >> 1
    table = {(0, 0): G, (0, 1): avg(G, h), (0, 2): h, (0, 3): avg(M, h), (1, 0): avg(G, b), (2, 0): b,
             (3, 0): avg(H, b), (1, 1): avg(b, h), (3, 1): avg(b, m), (1, 3): avg(h, s), (3, 3): avg(s, m),
             (2, 1): avg(b, j), (2, 2): j, (2, 3): avg(j, s), (1, 2): avg(h, j), (3, 2): avg(j, m)}
    return table[(xf, yf)]


def test_luma_interpolation_matches_numpy_restatement():
    rng = np.random.default_rng(7)
    ref = rng.integers(0, 256, (13, 17)).astype(np.uint8)
    for xf in range(4):
        for yf in range(4):
            want = luma_plane_numpy(ref, xf, yf)
            for y in range(-3, 16, 2):
                for x in range(-3, 20, 2):
                    yy, xx = min(max(y, 0), 12), min(max(x, 0), 16)
                    # coordinates outside the plane behave as the clamped edge sample rows/cols
                    got = refcpu.luma_sample(ref, x, y, xf, yf)
                    if 0 <= x < 17 and 0 <= y < 13:
                        assert got == want[y, x], (x, y, xf, yf)
                    elif xf == 0 and yf == 0:
                        assert got == ref[yy, xx]


def test_luma_interpolation_far_outside_equals_edge():
    rng = np.random.default_rng(8)
    ref = rng.integers(0, 256, (8, 8)).astype(np.uint8)
    for xf in range(4):
        for yf in range(4):
            assert refcpu.luma_sample(ref, -100, -100, xf, yf) == ref[0, 0]
            assert refcpu.luma_sample(ref, 100, 100, xf, yf) == ref[7, 7]


def test_chroma_interpolation_matches_numpy_restatement():
    rng = np.random.default_rng(9)
    ref = rng.integers(0, 256, (9, 11)).astype(np.uint8)
    for xf in range(8):
        for yf in range(8):
            for y in range(-2, 11):
                for x in range(-2, 13):
                    yy, xx = np.clip([y, y + 1], 0, 8), np.clip([x, x + 1], 0, 10)
                    A, B = int(ref[yy[0], xx[0]]), int(ref[yy[0], xx[1]])
                    C, D = int(ref[yy[1], xx[0]]), int(ref[yy[1], xx[1]])
                    want = ((8 - xf) * (8 - yf) * A + xf * (8 - yf) * B + (8 - xf) * yf * C + xf * yf * D + 32) >> 6
                    assert refcpu.chroma_sample(ref, x, y, xf, yf) == want


# ------------------------------------------------------------- deblocking
def test_filter_noop_conditions():
    rng = np.random.default_rng(10)
    for _ in range(200):
        p = rng.integers(0, 256, 4)
        q = rng.integers(0, 256, 4)
        for bs in (1, 2, 3, 4):
            pp, qq = refcpu.filter_line(p, q, bs, 0, 10)  # indexA < 16: alpha = 0
            assert (pp == p).all() and (qq == q).all()
        pp, qq = refcpu.filter_line(p, q, 0, 0, 40)
        assert (pp == p).all() and (qq == q).all()


def test_filter_bs4_mirror_symmetry():
    """For bS 4 every formula is its own mirror image: swapping the p and q sides
    swaps the outputs.  (bS < 4 rounds delta towards -inf, so it is not exact there.)"""
    rng = np.random.default_rng(11)
    for _ in range(2000):
        base = int(rng.integers(20, 230))
        p = np.clip(base + rng.integers(-6, 7, 4), 0, 255)
        q = np.clip(base + rng.integers(-6, 7, 4) + int(rng.integers(-15, 16)), 0, 255)
        chroma = int(rng.integers(0, 2))
        qa = int(rng.integers(16, 52))
        pp, qq = refcpu.filter_line(p, q, 4, chroma, qa)
        qq2, pp2 = refcpu.filter_line(q, p, 4, chroma, qa)
        assert (pp == pp2).all() and (qq == qq2).all()


def test_filter_bs_below_4_moves_p0_and_q0_oppositely():
    """bS < 4 adds delta to p0 and subtracts it from q0: away from the clipping range
    p0 + q0 is preserved, and |delta| never exceeds tC <= tC0 + 2."""
    rng = np.random.default_rng(12)
    changed = 0
    for _ in range(3000):
        base = int(rng.integers(40, 216))
        p = np.clip(base + rng.integers(-6, 7, 4), 0, 255)
        q = np.clip(base + rng.integers(-6, 7, 4) + int(rng.integers(-15, 16)), 0, 255)
        bs = int(rng.integers(1, 4))
        chroma = int(rng.integers(0, 2))
        pp, qq = refcpu.filter_line(p, q, bs, chroma, int(rng.integers(16, 52)))
        assert int(pp[0]) + int(qq[0]) == int(p[0]) + int(q[0])
        assert abs(int(pp[0]) - int(p[0])) <= 27
        assert (pp[2:] == p[2:]).all() and (qq[2:] == q[2:]).all()
        if chroma:
            assert pp[1] == p[1] and qq[1] == q[1]
        changed += int(pp[0] != p[0])
    assert changed > 300


def test_filter_known_answers():
    p, q = [100] * 4, [110] * 4
    # bS 4, qPav 30 (alpha 25, beta 8): |p0-q0| = 10 >= (alpha >> 2) + 2, weak 3-tap form
    pp, qq = refcpu.filter_line(p, q, 4, 0, 30)
    assert list(pp) == [103, 100, 100, 100] and list(qq) == [108, 110, 110, 110]
    # bS 4, qPav 40 (alpha 80, beta 13): strong filter on both sides
    pp, qq = refcpu.filter_line(p, q, 4, 0, 40)
    assert list(pp) == [104, 103, 101, 100] and list(qq) == [106, 108, 109, 110]
    # bS 4 chroma: only p0 / q0 change, with the 3-tap form
    pp, qq = refcpu.filter_line(p, q, 4, 1, 40)
    assert list(pp) == [103, 100, 100, 100] and list(qq) == [108, 110, 110, 110]
    # bS 2, qPav 30: tC0 = 1, tC = 3, delta clipped to 3; p1 / q1 move by tC0
    pp, qq = refcpu.filter_line(p, q, 2, 0, 30)
    assert list(pp) == [103, 101, 100, 100] and list(qq) == [107, 109, 110, 110]
    # bS 2 chroma: tC = tC0 + 1 = 2, p1 / q1 untouched
    pp, qq = refcpu.filter_line(p, q, 2, 1, 30)
    assert list(pp) == [102, 100, 100, 100] and list(qq) == [108, 110, 110, 110]
    # FilterOffsetA shifts indexA: qPav 24 + 6 behaves as qPav 30 for alpha / tC0
    pp, qq = refcpu.filter_line(p, q, 2, 1, 24, 6, 6)
    assert list(pp) == [102, 100, 100, 100] and list(qq) == [108, 110, 110, 110]


def test_bgr24_closed_form_and_truncation():
    """Gray stays gray; the division truncates toward zero (C), not down."""
    y = np.full((2, 4), 16, np.uint8)
    cb = np.full((1, 2), 128, np.uint8)
    cr = np.full((1, 2), 128, np.uint8)
    out = refcpu.yuv420p_to_bgr24(y, cb, cr)
    assert (out[:, :12] == 0).all() and out.shape == (2, 12)
    y[:] = 235
    assert (refcpu.yuv420p_to_bgr24(y, cb, cr)[:, :12] == 254).all()  # 1164 * 219 / 1000 = 254.9 -> 254
    # b = (1164 * (17 - 16) + 2018 * (127 - 128)) / 1000 = -854 / 1000 -> 0 (trunc), clipped 0
    y[:] = 17
    cb[:] = 127
    px = refcpu.yuv420p_to_bgr24(y, cb, cr)[0, :3]
    assert list(px) == [0, (1164 + 391) // 1000, 1164 // 1000]
    # flip writes rows bottom-up
    y = np.arange(8, dtype=np.uint8).reshape(2, 4) * 30 + 16
    a = refcpu.yuv420p_to_bgr24(y, cb, cr, flip=0)
    b = refcpu.yuv420p_to_bgr24(y, cb, cr, flip=1)
    assert (a[0] == b[1]).all() and (a[1] == b[0]).all()
