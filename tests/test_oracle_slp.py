"""CPU checks that pin the oracle's mixed-layer level (zdf_mxl) and
iso-neutral slopes (ldf_slp with ldf_slp_mxl) to closed forms:

* zdf_mxl: an unstratified column is mixed to its bottom (nmln = mbkt + 1);
  a density step at level L puts the base at L; negative N^2 counts as
  zero; a uniform N^2 crosses the criterion where its depth integral does;
  hmlp and hmlpt are the w- and t-depths of the base;
* ldf_slp below the mixed layer: for a density with a uniform horizontal
  gradient and uniform N^2 in a flat-bottomed box, every slope equals
  -g drho/dx / (N^2 (1 + rho)) (the Shapiro filter keeps a nearly linear
  field); with a steep gradient the limiter gives min(rn_slpmax, e3/7000)
  exactly; a horizontally uniform density gives zero slopes;
* in the mixed layer the u/v slopes taper linearly with depth from the
  slope at its base, (gdept - e3u(1)/2) / hmlpt, and the w slopes as
  gdepw / hmlp (the ldf_slp_mxl construction);
* the lateral boundary: the ring of a cyclic domain repeats the interior;
* the coastal and bottom tapers: with every unfiltered slope at the limiter,
  the filtered slope is that constant times the wet share of the filter
  weights times the taper factors;
* u-points between columns of different mixed-layer depth take the deeper
  one (level, taper depth and mask).
"""
import numpy as np
import pytest

from nemo_hip import configs
from oracle.oracle import Oracle

GRAV = 9.80665


E1, E2 = 10000.0, 15000.0


def box(jperio=0, jpi=30, jpj=24, jpk=12):
    """A flat-bottomed box, every interior column wet to level jpk-1, with
    e1 = 10 km and e2 = 15 km (so an e1/e2 mix-up changes the slopes)."""
    d = configs.basin(jpi, jpj, jpk, seed=3, land_frac=0.0, depth_min=5000.0, depth_max=5000.0, jperio=jperio,
                      e1=E1)
    A = d.arrays
    for n in ("e2t", "e2u", "e2v", "e2f"):
        A[n] = np.asfortranarray(np.full_like(A[n], E2))
    A["r1_e2u"], A["r1_e2v"] = 1.0 / A["e2u"], 1.0 / A["e2v"]
    for p in ("t", "u", "v", "f"):
        A["e1e2" + p] = A["e1" + p] * A["e2" + p]
        if "r1_e1e2" + p in A:
            A["r1_e1e2" + p] = 1.0 / A["e1e2" + p]
    A["e2_e1u"], A["e1_e2v"] = A["e2u"] / A["e1u"], A["e1v"] / A["e2v"]
    for n in A:
        A[n] = np.asfortranarray(A[n])
    return d


def fz(shape):
    return np.zeros(shape, order="F")


def nlb10_of(d):
    return configs.nlb10(d)


# ----------------------------------------------------------------------------
def test_zdf_mxl_closed_forms():
    d = configs.basin(26, 20, 14, seed=9)
    o = Oracle(d)
    t, *_ = configs.masks(d)
    gdept, gdepw = configs.depths(d)
    s2, s3 = (d.jpi, d.jpj), (d.jpi, d.jpj, d.jpk)
    nlb10 = nlb10_of(d)
    mbkt = np.maximum(d.arrays["k_bot"], 1)
    ss = t[:, :, 0]
    # unstratified (and unstable) columns are mixed to the bottom
    for val in (0.0, -3e-4):
        nm, hp, hpt = o.zdf_mxl(1, np.full(s3, val, order="F"), gdept, gdepw, nlb10)
        want = np.minimum(d.jpk - 1, mbkt) + 1
        assert np.array_equal(nm, np.maximum(want, nlb10))
        ii, jj = np.indices(s2)
        assert np.array_equal(hp, gdepw[ii, jj, nm - 1] * ss)
        assert np.array_equal(hpt, gdept[ii, jj, nm - 2] * ss)
    # unstable water above a strong step: negative N^2 counts as zero, so the
    # base is the step, not deeper
    rn2 = np.full(s3, -1.0, order="F")
    rn2[:, :, 5:] = 1.0
    nm, _, _ = o.zdf_mxl(1, rn2, gdept, gdepw, nlb10)
    assert np.array_equal(nm, np.maximum(np.minimum(5, mbkt) + 1, nlb10))
    # a density step at level L
    L = 6
    rn2 = fz(s3)
    rn2[:, :, L - 1:] = 1.0
    nm, _, _ = o.zdf_mxl(1, rn2, gdept, gdepw, nlb10)
    assert np.array_equal(nm, np.maximum(np.minimum(L - 1, mbkt) + 1, nlb10))
    # uniform N^2: the base where the integral from nlb10 first reaches the criterion
    c = GRAV * 0.01 / d.opts["rau0"]
    e3w = d.e3w_1d
    n0 = 0.37 * c / e3w[nlb10 - 1:nlb10 + 3].sum()
    nm, _, _ = o.zdf_mxl(1, np.full(s3, n0, order="F"), gdept, gdepw, nlb10)
    integ = np.cumsum(n0 * e3w[nlb10 - 1:d.jpk - 1])
    below = np.nonzero(integ < c)[0]
    klast = nlb10 + below[-1]
    assert abs(integ[below[-1]] - c) > 1e-6 * c and 3 <= len(below) <= 5
    assert np.array_equal(nm, np.maximum(np.minimum(klast, mbkt) + 1, nlb10))
    with pytest.raises(RuntimeError):
        o.zdf_mxl(1, rn2, gdept, gdepw, 1)


def uniform_gradient(d, axis, alpha, n0=1e-5, r0=2e-3, beta=0.0):
    """prd = r0 + alpha * (index along axis) + beta * level, N^2 = n0 on
    levels 2..jpk-1 (zero at 1 and jpk, as bn2 leaves them)."""
    s3 = (d.jpi, d.jpj, d.jpk)
    ix = np.indices(s3)
    prd = np.asfortranarray(r0 + alpha * (ix[axis] + 1.0) + beta * (ix[2] + 1.0))
    pn2 = fz(s3)
    pn2[:, :, 1:-1] = n0
    return prd, pn2


def slope_u(alpha, E, n0, r0, beta, c, k):
    """The u/v-point slope at 1-based coordinate c, level k: density at the
    face mean of the two t-points, zdzr = -(prd + 1) n0 / g."""
    return -alpha * GRAV / (E * n0 * (1 + r0 + alpha * (c + 0.5) + beta * k))


def slope_w(alpha, E, n0, r0, beta, c, k):
    """The w-point slope: zbw from the mean density of levels k and k-1."""
    return -alpha * GRAV / (E * n0 * (1 + r0 + alpha * c + beta * (k - 0.5)))


def slopes(o, d, prd, pn2, nmln, rn_slpmax=0.01):
    gdept, gdepw = configs.depths(d)
    ii, jj = np.indices((d.jpi, d.jpj))
    nm = np.full((d.jpi, d.jpj), nmln, dtype=np.int32, order="F")
    hmlp = np.asfortranarray(gdepw[ii, jj, nm - 1])
    hmlpt = np.asfortranarray(gdept[ii, jj, nm - 2])
    s3 = (d.jpi, d.jpj, d.jpk)
    out = [np.full(s3, 9.0, order="F") for _ in range(4)]
    for a in out:
        a[:, :, 0] = a[:, :, -1] = 0.0
    o.ldf_slp(1, prd, pn2, gdept, gdepw, nm, hmlp, hmlpt, rn_slpmax, *out)
    return out, gdept, gdepw


@pytest.mark.parametrize("axis", [0, 1])
def test_ldf_slp_uniform_gradient_below_mixed_layer(axis):
    d = box()
    o = Oracle(d)
    E, n0, r0, beta = (E1, E2)[axis], 1e-5, 2e-3, 1e-4
    alpha = 1e-4 * E * n0 / GRAV                      # slope ~ -1e-4
    prd, pn2 = uniform_gradient(d, axis, alpha, n0, r0, beta)
    (us, vs, wi, wj), _, _ = slopes(o, d, prd, pn2, nmln=2)
    along, across = (us, vs) if axis == 0 else (vs, us)
    walong, wacross = (wi, wj) if axis == 0 else (wj, wi)
    c = np.arange(d.jpi if axis == 0 else d.jpj) + 1.0
    sl = (slice(2, d.jpi - 3), slice(2, d.jpj - 3))   # 3x3 neighbourhood wet, away from the taper
    shape = [1, 1]
    shape[axis] = -1
    idx = c[sl[axis]]
    for k in range(2, d.jpk):                          # 1-based levels 2..jpk-1
        # at the bottom level (k = jpk-1) N^2 below is zero and zdzr divides
        # by 1 instead of 2; the vertical taper then halves the u/v slope
        taper = 0.5 if k == d.jpk - 1 else 1.0
        su = taper * slope_u(alpha, E, n0, r0, beta, idx, k)
        got_u = along[sl[0], sl[1], k - 1]
        assert np.allclose(got_u, su.reshape(shape), rtol=1e-9, atol=0), k
        if k >= 3:                                     # level 2 takes the mixed-layer base value
            sw = slope_w(alpha, E, n0, r0, beta, idx, k)
            got_w = walong[sl[0], sl[1], k - 1]
            assert np.allclose(got_w, sw.reshape(shape), rtol=1e-9, atol=0), k
        assert np.all(across[:, :, k - 1] == 0.0) and np.all(wacross[:, :, k - 1] == 0.0)
    sw3 = slope_w(alpha, E, n0, r0, beta, idx, 3)
    assert np.allclose(walong[sl[0], sl[1], 1], sw3.reshape(shape), rtol=1e-9, atol=0)
    # a horizontally uniform density has no slope
    prd0, _ = uniform_gradient(d, axis, 0.0, n0, r0)
    out, _, _ = slopes(o, d, prd0, pn2, nmln=2)
    for a in out:
        assert np.all(a == 0.0)


@pytest.mark.parametrize("axis", [0, 1])
def test_ldf_slp_limiter(axis):
    """A gradient 100x steeper than rn_slpmax: |slope| = min(rn_slpmax, e3/7000)."""
    d = box()
    o = Oracle(d)
    E, n0 = (E1, E2)[axis], 1e-5
    alpha = -1.0 * E * n0 / GRAV                     # raw slope ~ +1
    prd, pn2 = uniform_gradient(d, axis, alpha, n0)
    (us, vs, wi, wj), _, _ = slopes(o, d, prd, pn2, nmln=2, rn_slpmax=0.01)
    along, walong = (us, wi) if axis == 0 else (vs, wj)
    sl = (slice(2, d.jpi - 3), slice(2, d.jpj - 3))
    for k in range(2, d.jpk - 2):
        wu = min(0.01, d.e3t_1d[k - 1] / 7.0e3)
        # at k = 2 (= nmln) the w-slope is the mixed-layer base value, taken at level 3
        ww = min(0.01, d.e3w_1d[max(k, 3) - 1] / 7.0e3)
        assert np.allclose(along[sl[0], sl[1], k - 1], wu, rtol=1e-12, atol=0), k
        assert np.allclose(walong[sl[0], sl[1], k - 1], ww, rtol=1e-12, atol=0), k


def test_ldf_slp_mixed_layer_taper():
    d = box()
    o = Oracle(d)
    E, n0, r0, beta, M = E1, 1e-5, 2e-3, 1e-4, 6
    alpha = 1e-4 * E * n0 / GRAV
    prd, pn2 = uniform_gradient(d, 0, alpha, n0, r0, beta)
    (us, vs, wi, wj), gdept, gdepw = slopes(o, d, prd, pn2, nmln=M)
    c = np.arange(2, d.jpi - 3) + 1.0
    su_base = slope_u(alpha, E, n0, r0, beta, c, M)         # iku = nmln
    sw_base = slope_w(alpha, E, n0, r0, beta, c, M + 1)     # ik = nmln + 1
    e3u1 = d.arrays["e3u_n"][0, 0, 0]
    for k in range(2, M):              # inside the mixed layer: u-slope tapers with gdept
        f = (gdept[0, 0, k - 1] - 0.5 * e3u1) / gdept[0, 0, M - 2]
        assert np.allclose(us[2:-3, 2:-3, k - 1], (f * su_base)[:, None], rtol=1e-9, atol=0), k
    for k in range(2, M + 1):          # w-slope tapers with gdepw down to the base
        f = gdepw[0, 0, k - 1] / gdepw[0, 0, M - 1]
        assert np.allclose(wi[2:-3, 2:-3, k - 1], (f * sw_base)[:, None], rtol=1e-9, atol=0), k
    for k in range(M, d.jpk - 2):      # below: the interior slope
        assert np.allclose(us[2:-3, 2:-3, k - 1], slope_u(alpha, E, n0, r0, beta, c, k)[:, None], rtol=1e-9,
                           atol=0), k
    for k in range(M + 1, d.jpk - 1):
        assert np.allclose(wi[2:-3, 2:-3, k - 1], slope_w(alpha, E, n0, r0, beta, c, k)[:, None], rtol=1e-9,
                           atol=0), k


def test_ldf_slp_cyclic_ring_and_untouched_levels():
    d = configs.basin(28, 22, 10, seed=4, jperio=1)
    o = Oracle(d)
    rng = np.random.default_rng(5)
    s3 = (d.jpi, d.jpj, d.jpk)
    prd = np.asfortranarray(rng.uniform(-2e-3, 4e-3, s3))
    pn2 = np.asfortranarray(rng.uniform(1e-6, 1e-4, s3))
    (us, vs, wi, wj), _, _ = slopes(o, d, prd, pn2, nmln=3)
    for a in (us, vs, wi, wj):
        assert np.array_equal(a[0], a[-2]) and np.array_equal(a[-1], a[1])
        assert np.all(a[:, 0] == 0.0) and np.all(a[:, -1] == 0.0)      # closed north/south
        assert np.all(a[:, :, 0] == 0.0) and np.all(a[:, :, -1] == 0.0)
        assert np.isfinite(a).all() and np.abs(a).max() <= 0.01 * (1 + 1e-12)


def conv121(a):
    """The 1-2-1 x 1-2-1 sum of a 2D field at interior points (ring left 0)."""
    out = np.zeros_like(a)
    w = np.array([1.0, 2.0, 1.0])
    for di in range(3):
        for dj in range(3):
            out[1:-1, 1:-1] += w[di] * w[dj] * a[di:di + a.shape[0] - 2, dj:dj + a.shape[1] - 2]
    return out


def test_ldf_slp_coastal_and_topographic_tapers():
    """In a basin with coasts and a varying bottom, a density gradient 100x
    steeper than rn_slpmax makes every unfiltered slope the limited constant
    s0(k) = min(rn_slpmax, e3(k)/7000) on wet points and 0 on land.  The
    filtered slope is then s0 times the share of the 1-2-1 x 1-2-1 weights that
    fall on wet points, times the tapers: (umask(j+1) + umask(j-1))/2 for u
    across the coast and (umask(k) + umask(k+1))/2 at the bottom; for the w
    slopes wmask times (umask(i) + umask(i-1)) (vmask(j) + vmask(j-1)) / 4.
    A w-point none of whose four neighbouring u-points is wet has no gradient
    to average and stays 0."""
    d = configs.basin(30, 24, 12, seed=3, land_frac=0.3, e1=E1)
    o = Oracle(d)
    n0 = 1e-5
    alpha = -1.0 * E1 * n0 / GRAV
    prd, pn2 = uniform_gradient(d, 0, alpha, n0)
    (us, vs, wi, wj), _, _ = slopes(o, d, prd, pn2, nmln=2, rn_slpmax=0.01)
    _, um, vm, wm, _ = configs.masks(d)
    e3u, e3w = d.arrays["e3u_n"][0, 0], d.arrays["e3w_n"][0, 0]
    coast = 0
    for k in range(3, d.jpk - 1):                         # 1-based levels below the mixed layer
        q = k - 1
        s0u = min(0.01, e3u[q] / 7.0e3)
        s0w = min(0.01, e3w[q] / 7.0e3)
        u2 = um[:, :, q]
        tj = np.zeros_like(u2)
        tj[:, 1:-1] = 0.5 * (u2[:, 2:] + u2[:, :-2])
        tk = 0.5 * (u2 + um[:, :, q + 1])
        want_u = s0u * conv121(u2) / 16.0 * tj * tk
        want_u[0], want_u[-1], want_u[:, 0], want_u[:, -1] = 0, 0, 0, 0
        assert np.allclose(us[:, :, q], want_u, rtol=1e-12, atol=1e-18), k
        # unfiltered w slope: s0w where the w-point is wet and one of its four u-points is
        anyu = np.zeros_like(u2)
        anyu[1:, :] = (um[1:, :, q] + um[:-1, :, q] + um[1:, :, q - 1] + um[:-1, :, q - 1]) > 0
        zw = s0w * wm[:, :, q] * anyu
        zw[0], zw[-1], zw[:, 0], zw[:, -1] = 0, 0, 0, 0
        zck = np.zeros_like(u2)
        zck[1:, 1:] = (u2[1:, 1:] + u2[:-1, 1:]) * (vm[1:, 1:, q] + vm[1:, :-1, q]) * 0.25
        want_w = conv121(zw) / 16.0 * wm[:, :, q] * zck
        want_w[0], want_w[-1], want_w[:, 0], want_w[:, -1] = 0, 0, 0, 0
        assert np.allclose(wi[:, :, q], want_w, rtol=1e-12, atol=1e-18), k
        assert np.all(vs[:, :, q] == 0.0) and np.all(wj[:, :, q] == 0.0)
        coast += int(((want_u > 0) & (want_u < 0.999 * s0u)).sum())
    assert coast > 50                                     # many points feel a coast or the bottom


def test_ldf_slp_mixed_layer_between_columns_of_different_depth():
    """A band of columns (i = I0) whose mixed layer is deeper (nmln = M2) than
    elsewhere (M1): the u-points on both sides of the band take the deeper
    mixed layer (ldf_slp_mxl: iku = max(nmln(i), nmln(i+1)), the taper depth
    max(hmlpt(i), hmlpt(i+1)), the mixed-layer mask max of the two columns),
    and the 1-2-1 filter along i then mixes them with their shallow
    neighbours: (S + 2 S + D)/4, (S + 2 D + D)/4, (D + 2 D + S)/4, (D + 2 S + S)/4."""
    d = box()
    o = Oracle(d)
    E, n0, r0, beta, M1, M2, I0 = E1, 1e-5, 2e-3, 1e-4, 3, 7, 14
    alpha = 1e-4 * E * n0 / GRAV
    prd, pn2 = uniform_gradient(d, 0, alpha, n0, r0, beta)
    gdept, gdepw = configs.depths(d)
    nm = np.full((d.jpi, d.jpj), M1, dtype=np.int32, order="F")
    nm[I0 - 1, :] = M2                                    # 1-based column I0
    ii, jj = np.indices((d.jpi, d.jpj))
    hmlp = np.asfortranarray(gdepw[ii, jj, nm - 1])
    hmlpt = np.asfortranarray(gdept[ii, jj, nm - 2])
    s3 = (d.jpi, d.jpj, d.jpk)
    us, vs, wi, wj = [np.full(s3, 9.0, order="F") for _ in range(4)]
    for a in (us, vs, wi, wj):
        a[:, :, 0] = a[:, :, -1] = 0.0
    o.ldf_slp(1, prd, pn2, gdept, gdepw, nm, hmlp, hmlpt, 0.01, us, vs, wi, wj)
    e3u1 = d.arrays["e3u_n"][0, 0, 0]
    gt = gdept[0, 0]

    def zwz(c, k, M):
        """unfiltered u slope at the u-point of 1-based coordinate c whose
        mixed layer has its base at level M"""
        if k <= M - 1:
            return (gt[k - 1] - 0.5 * e3u1) / gt[M - 2] * slope_u(alpha, E, n0, r0, beta, c, M)
        return slope_u(alpha, E, n0, r0, beta, c, k)

    for k in range(2, d.jpk - 2):
        for c in range(I0 - 3, I0 + 3):                   # u-points around the band
            Ms = [M2 if cc in (I0 - 1, I0) else M1 for cc in (c - 1, c, c + 1)]
            want = (zwz(c - 1, k, Ms[0]) + 2 * zwz(c, k, Ms[1]) + zwz(c + 1, k, Ms[2])) / 4.0
            got = us[c - 1, 4:-4, k - 1]
            assert np.allclose(got, want, rtol=1e-9, atol=0), (k, c)
    # the deep band's own taper differs from the shallow one inside its mixed layer
    assert abs(us[I0 - 1, 6, 2] - us[I0 + 3, 6, 2]) > 0.1 * abs(us[I0 + 3, 6, 2])
