"""Pins of the FP64 twin of the contraction cube filter (the checker of the
MFMA path): constants are reproduced exactly, rotating the environment by a
cube symmetry rotates the output, a full-sphere sum at a coarse level agrees
with the sum over every level-0 texel for rough lobes on a smooth
environment, and the near/far split agrees with the full-sphere sum."""
import numpy as np
import pytest

from oracle import hvx_oracle as O


def face_dirs(S):
    """World directions of every texel centre of a 6 x S x S cube (face-major)."""
    c = (2.0 * np.arange(S) + 1.0) / S - 1.0
    u, v = np.meshgrid(c, c)  # v rows, u columns
    one = np.ones_like(u)
    faces = [(one, -v, -u), (-one, -v, u), (u, one, v), (u, -one, -v), (u, -v, one), (-u, -v, -one)]
    d = np.stack([np.stack(f, axis=-1) for f in faces]).reshape(-1, 3)
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def env_from(fn, S):
    d = face_dirs(S)
    rgb = fn(d)
    return np.concatenate([rgb, np.ones((d.shape[0], 1))], axis=1).astype(np.float32).ravel()


def smooth(d):
    return np.stack([1.5 + 0.8 * d[:, 0] + 0.3 * d[:, 1] * d[:, 2], 1.0 + 0.5 * d[:, 1] ** 2,
                     0.7 + 0.4 * d[:, 2] - 0.2 * d[:, 0] * d[:, 1]], axis=1)


def test_level_table():
    """Per-mip kinds and roughness of a 64 -> 32 x 6 chain."""
    lv = O.contract_levels(1, 64, 32, 6)
    assert [l["kind"] for l in lv] == [O.CONTRACT_DELTA] + [O.CONTRACT_FULL] * 5
    assert lv[3]["alpha2"] == pytest.approx((3 / 5) ** 4)


@pytest.mark.parametrize("S,So,mips", [(64, 32, 6), (128, 64, 7)])
def test_constant_environment(S, So, mips):
    lvl = np.full(6 * S * S * 4, 2.25, np.float32)
    for m in range(mips):
        r = O.contract(1, lvl, S, So, mips, m, np.arange(min(6 * max(So >> m, 1) ** 2, 200), dtype=np.uint32))
        assert np.abs(r - 2.25).max() < 1e-12, m
    r = O.contract(0, lvl, S, 16, 1, 0)
    assert np.abs(r - 2.25).max() < 1e-12


def test_cube_rotation_symmetry():
    """Rotating the environment by a quarter turn about +Y (a cube symmetry)
    rotates the filtered result: out_rot(R d) = out(d)."""
    S, So, mips = 64, 32, 6
    rng = np.random.default_rng(5)
    centres = rng.normal(size=(6, 3))
    centres /= np.linalg.norm(centres, axis=1, keepdims=True)

    def env(d):
        acc = np.full((d.shape[0], 3), 0.2)
        for k, c in enumerate(centres):
            acc[:, k % 3] += 3.0 * np.exp(8.0 * (d @ c - 1.0))
        return acc

    R = np.array([[0.0, 0.0, 1.0], [0.0, 1.0, 0.0], [-1.0, 0.0, 0.0]])  # quarter turn about y
    a = env_from(env, S)
    b = env_from(lambda d: env(d @ R), S)  # b(d) = a(R^T d)
    dirs = face_dirs(So >> 2)
    for m in (1, 3):
        Sm = So >> m
        d = face_dirs(Sm)
        ids = np.arange(0, 6 * Sm * Sm, 7, dtype=np.uint32)
        ra = O.contract(1, a, S, So, mips, m, ids)
        # texel of R d (a texel centre again: R permutes the cube's texel grid)
        rd = d[ids] @ R.T
        D = face_dirs(Sm)
        jd = np.argmax(rd @ D.T, axis=1).astype(np.uint32)
        rb = O.contract(1, b, S, So, mips, m, jd)
        assert np.abs(rb - ra).max() <= 1e-10 * np.abs(ra).max(), m
    del dirs


def test_full_sum_matches_level0_integral_for_rough_lobes():
    S, So, mips = 128, 64, 7
    lvl = env_from(smooth, S)
    lv = O.contract_levels(1, S, So, mips)
    for m in (4, 5, 6):
        Sm = max(So >> m, 1)
        ids = np.arange(0, 6 * Sm * Sm, max(1, 6 * Sm * Sm // 24), dtype=np.uint32)
        r = O.contract(1, lvl, S, So, mips, m, ids)
        ideal = O.ideal(1, lvl, S, lv[m]["alpha2"], Sm, ids)
        assert np.abs(r - ideal).max() / np.abs(ideal).max() < 2e-3, m
    r = O.contract(0, lvl, S, 16, 1, 0)
    ideal = O.ideal(0, lvl, S, 0.0, 16, np.arange(6 * 256, dtype=np.uint32))
    assert np.abs(r - ideal).max() / np.abs(ideal).max() < 2e-3


def test_split_agrees_with_ideal_on_smooth_environment():
    """The near/far split (mip 1 of 256 -> 128 x 8) against the level-0 sum."""
    S, So, mips = 256, 128, 8
    lv = O.contract_levels(1, S, So, mips)
    assert lv[1]["kind"] == O.CONTRACT_SPLIT
    lvl = env_from(smooth, S)
    Sm = So >> 1
    ids = np.random.default_rng(2).integers(0, 6 * Sm * Sm, 40).astype(np.uint32)
    r = O.contract(1, lvl, S, So, mips, 1, ids)
    ideal = O.ideal(1, lvl, S, lv[1]["alpha2"], Sm, ids)
    assert np.abs(r - ideal).max() / np.abs(ideal).max() < 2e-3


def test_ideal_weight_is_the_reference_estimand():
    """The contraction's weight w(c) = D_GGX(h(c)) c (c = n.l, V = R = N) is
    the integrand the reference's estimator converges to: importance sampling
    H from D(h)(n.h) and weighting by n.l (CubeFilterSpecular.frag:46-72)
    estimates the integral of L(l) D(h) n.l over the integral of D(h) n.l.
    On a smooth environment the FP64 estimator with 16,384 Hammersley samples
    and the FP64 sum over every level-0 texel (hvxo_ideal_f64, which uses
    the product's weight) agree within 2e-4 at roughness 0.4, 0.6 and 1.0.
    Sums with a plausibly wrong weight (D alone, without n.l; or D at
    n.h^2 = c instead of (1 + c) / 2) miss the estimator by more than 10x
    that on the rough lobes, where the weight's shape matters."""
    S, So, mips, n = 128, 64, 6, 16384
    lvl = env_from(smooth, S)
    dirs = face_dirs(S)
    # solid angle of each texel: (2/S)^2 / (1 + u^2 + v^2)^(3/2)
    c = (2.0 * np.arange(S) + 1.0) / S - 1.0
    u, v = np.meshgrid(c, c)
    dw = np.tile(((2.0 / S) ** 2 / (1 + u * u + v * v) ** 1.5).ravel(), 6)
    rgb = lvl.reshape(-1, 4)[:, :3].astype(np.float64)
    rng = np.random.default_rng(11)
    miss = {"no n.l": 0.0, "D at c": 0.0}
    for m in (2, 3, 5):
        rough = m / (mips - 1)
        a2 = (rough * rough) ** 2
        Sm = So >> m
        ids = np.unique(rng.integers(0, 6 * Sm * Sm, 12)).astype(np.uint32)
        est = O.specular(lvl, S, So, mips, m, n, ids, dtype=np.float64)[:, :3]
        ref = O.ideal(1, lvl, S, a2, Sm, ids)
        scale = np.abs(ref).max()
        err = np.abs(est - ref).max() / scale
        assert err < 2e-4, (m, err)
        # the same sum in numpy: the weight the product uses, then wrong ones
        N = face_dirs(Sm)[ids]
        cl = N @ dirs.T  # n . l per (output, texel)

        def weighted(w):
            w = np.where(cl > 0, w, 0.0) * dw
            return (w @ rgb) / w.sum(axis=1, keepdims=True)

        def D_of(nh2):
            return a2 / (np.pi * (nh2 * (a2 - 1) + 1) ** 2)

        right = weighted(D_of(0.5 * (1 + cl)) * cl)
        assert np.abs(right - ref).max() <= 1e-9 * scale
        for name, wrong in (("no n.l", weighted(D_of(0.5 * (1 + cl)))), ("D at c", weighted(D_of(cl) * cl))):
            miss[name] = max(miss[name], np.abs(est - wrong).max() / scale)
    for name, e in miss.items():
        assert e > 2e-3, (name, e)


# ----------------------------------------------------------- bright-texel split
def _w(alpha2, c, diffuse=False):
    c = np.asarray(c, np.float64)
    if diffuse:
        return np.where(c > 0, c, 0.0)
    d = 0.5 * (1.0 + c) * (alpha2 - 1.0) + 1.0
    return np.where(c > 0, alpha2 / (np.pi * d * d) * c, 0.0)


def _dw(S):
    c = (2.0 * np.arange(S) + 1.0) / S - 1.0
    u, v = np.meshgrid(c, c)
    r2 = 1.0 + u * u + v * v
    return np.tile(((2.0 / S) ** 2 / (r2 * np.sqrt(r2))).ravel(), 6)


@pytest.mark.parametrize("alpha2", [1.5e-4, 2.4e-3, 1.23e-2, 0.1, 0.6, 0.9, 0.99, 0.9999, 1.0])
def test_weight_integral_matches_quadrature(alpha2):
    """The closed form (series near a^2 = 1) equals 2 pi int_0^1 w(c) dc by
    adaptive quadrature; the cosine weight integrates to pi."""
    from scipy.integrate import quad
    pts = [1.0 - 4.0 * alpha2] if alpha2 < 0.2 else None
    num = 2.0 * np.pi * quad(lambda c: float(_w(alpha2, c)), 0.0, 1.0, epsabs=0, epsrel=1e-13, limit=500,
                             points=pts)[0]
    assert O.weight_integral(1, alpha2) == pytest.approx(num, rel=1e-10)
    assert O.weight_integral(0, alpha2) == pytest.approx(np.pi, rel=1e-15)


def test_weight_integral_matches_level0_sum():
    """Summed over every texel centre of a 6 x 256^2 level, the weight of a
    lobe a few texels wide is within 1e-5 of the closed-form integral."""
    S = 256
    d, dw = face_dirs(S), _dw(S)
    rng = np.random.default_rng(3)
    for alpha2 in (2.4e-3, 1.23e-2, 0.1):
        for _ in range(3):
            n = rng.normal(size=3)
            n /= np.linalg.norm(n)
            z = float((_w(alpha2, d @ n) * dw).sum())
            assert z == pytest.approx(O.weight_integral(1, alpha2), rel=1e-5)


def _largest(level0):
    return np.maximum(np.maximum(level0[0::4], level0[1::4]), level0[2::4])


@pytest.mark.parametrize("K", [0, 1, 7, 64, 100000])
def test_hot_threshold_rule(K):
    """T is the lower edge of the lowest 2^(1/8)-wide bin (float pattern >> 20)
    such that at most K texels have a largest channel at or above it: at most
    K texels reach T, and more than K reach the next lower bin edge; 0 when
    every positive texel fits (K = 0: the edge above the brightest bin)."""
    S = 16
    rng = np.random.default_rng(K + 1)
    lvl = rng.uniform(0.0, 2.0, 6 * S * S * 4).astype(np.float32)
    lvl[4 * rng.integers(0, 6 * S * S, 40)] = rng.uniform(50.0, 5000.0, 40).astype(np.float32)
    lvl[4 * np.arange(5)] = 0.0
    lvl[4 * np.arange(5) + 1] = 0.0
    lvl[4 * np.arange(5) + 2] = 0.0
    m = _largest(lvl)
    T = O.hot_threshold(lvl, S, K)
    hot = (m > 0) & (m >= T)
    assert hot.sum() <= K
    if T == 0.0:
        assert (m > 0).sum() <= K
    elif np.isinf(T):
        top = (m.view(np.uint32) >> 20) == (m.view(np.uint32) >> 20).max()
        assert top.sum() > K
    else:
        tb = np.float32(T).view(np.uint32)
        assert tb & ((1 << 20) - 1) == 0
        lower = np.uint32(tb - (1 << 20)).view(np.float32)
        assert ((m > 0) & (m >= lower)).sum() > K


def test_hot_split_with_every_texel_hot_is_the_level0_sum():
    """When every positive texel is split off (T = 0), the clamped part is zero
    and the result is the exact level-0 sum over the closed-form integral: the
    level-0 sum times (discrete normaliser / integral)."""
    S, So, mips = 16, 8, 4
    rng = np.random.default_rng(11)
    lvl = rng.uniform(0.1, 3.0, 6 * S * S * 4).astype(np.float32)
    lvl[4 * 77] = 800.0
    assert O.hot_threshold(lvl, S, 4096) == 0.0
    d, dw = face_dirs(S), _dw(S)
    lv = O.contract_levels(1, S, So, mips)
    for m in range(1, mips):
        Sm = So >> m
        ids = np.arange(6 * Sm * Sm, dtype=np.uint32)
        got = O.contract(1, lvl, S, So, mips, m, ids, max_hot=4096)
        ref = O.ideal(1, lvl, S, lv[m]["alpha2"], Sm, ids)
        zc = O.weight_integral(1, lv[m]["alpha2"])
        n_o = face_dirs(Sm)
        z = (_w(lv[m]["alpha2"], n_o @ d.T) * dw[None, :]).sum(axis=1)
        np.testing.assert_allclose(got * (zc / z)[:, None], ref, rtol=1e-11)
    got = O.contract(0, lvl, S, 8, 1, 0, max_hot=4096)
    ref = O.ideal(0, lvl, S, 0.0, 8, np.arange(6 * 64, dtype=np.uint32))
    z = (_w(0, face_dirs(8) @ d.T, diffuse=True) * dw[None, :]).sum(axis=1)
    np.testing.assert_allclose(got * (np.pi / z)[:, None], ref, rtol=1e-11)


def test_hot_split_removes_a_point_source_error():
    """A single texel of radiance 1e5 on a smooth environment: the split
    contraction agrees with the level-0 sum at least 5x better than the
    plain contraction around it (on mips that sum coarser levels than level
    0), and the delta mip is untouched."""
    S, So, mips = 256, 128, 8
    lvl = env_from(smooth, S)
    src = 6 * S * S // 3 + S * 40 + 70
    lvl[4 * src:4 * src + 3] = 1e5
    n_src = face_dirs(S)[src]
    lv = O.contract_levels(1, S, So, mips)
    for m in (1, 2, 3):
        Sm = So >> m
        nd = face_dirs(Sm)
        ids = np.argsort(-(nd @ n_src))[:24].astype(np.uint32)
        ref = O.ideal(1, lvl, S, lv[m]["alpha2"], Sm, ids)
        e_plain = np.abs(O.contract(1, lvl, S, So, mips, m, ids) / ref - 1).max()
        e_hot = np.abs(O.contract(1, lvl, S, So, mips, m, ids, max_hot=16) / ref - 1).max()
        assert e_hot * 5 < e_plain, (m, e_hot, e_plain)
    ids = np.arange(6 * So * So, dtype=np.uint32)
    np.testing.assert_array_equal(O.contract(1, lvl, S, So, mips, 0, ids, max_hot=16),
                                  O.contract(1, lvl, S, So, mips, 0, ids))
