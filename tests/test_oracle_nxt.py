"""CPU checks that pin the oracle's time-step closure (div_hor, ssh_nxt, wzv,
ssh_swp, tra_nxt, dyn_nxt) to what the mathematics fixes: the discrete Gauss
theorem, exactly non-divergent flows, volume conservation, closed forms of
the vertical integral, the consistency of the vertical velocity at the surface
with the sea-level tendency, the exact behaviour of the Asselin filter on
linear and quadratic sequences, and the depth-mean constraints of the
time-splitting corrections.  A wrong index, sign, time step or dropped term
fails at least one of them.
"""
import numpy as np
import pytest

from nemo_hip import configs
from oracle.oracle import Oracle


def fz(shape):
    return np.zeros(shape, order="F")


def rough_domain(jperio=0, seed=11, opts=None, jpi=24, jpj=18, jpk=9):
    """A seeded basin whose face lengths and cell thicknesses vary in every
    direction, so telescoping sums only cancel with the right indices."""
    d = configs.basin(jpi, jpj, jpk, seed=seed, jperio=jperio, opts=opts)
    rng = np.random.default_rng(seed + 1000)
    A = d.arrays
    for n in ("e2u", "e1v"):
        A[n] = np.asfortranarray(A[n] * rng.uniform(0.8, 1.25, A[n].shape))
    A["r1_e1e2t"] = np.asfortranarray(A["r1_e1e2t"] * rng.uniform(0.8, 1.25, A["r1_e1e2t"].shape))
    for n in ("e3t_n", "e3u_n", "e3v_n"):
        A[n] = np.asfortranarray(A[n] * rng.uniform(0.9, 1.1, A[n].shape))
    return d


def masks(d):
    t, u, v, w, f = configs.masks(d)
    return t, u, v


def rnd(rng, shape, mask=None):
    a = rng.standard_normal(shape)
    if mask is not None:
        a = a * mask
    return np.asfortranarray(a)


def lbc_src(d, i, j):
    """0-based source of ring point (i,j) after the serial lbc_lnk, or None."""
    iper, jper = d.jperio in (1, 4, 6, 7), d.jperio in (2, 7)
    si, sj = i, j
    if j in (0, d.jpj - 1):
        if not jper:
            return None
        sj = d.jpj - 2 if j == 0 else 1
    if i in (0, d.jpi - 1):
        if not iper:
            return None
        si = d.jpi - 2 if i == 0 else 1
    return si, sj


def assert_ring_is_lbc(d, a, what):
    """Every ring point holds its lbc source value (cyclic) or zero (closed)."""
    for i in range(d.jpi):
        for j in range(d.jpj):
            if 0 < i < d.jpi - 1 and 0 < j < d.jpj - 1:
                continue
            s = lbc_src(d, i, j)
            want = np.zeros(a.shape[2:]) if s is None else a[s[0], s[1]]
            assert np.array_equal(a[i, j], want), f"{what}: ring point ({i + 1},{j + 1})"


# ----------------------------------------------------------------------------
# div_hor
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("jperio", [0, 7])
def test_div_hor_gauss_theorem(jperio):
    d = rough_domain(jperio)
    o = Oracle(d)
    A = d.arrays
    rng = np.random.default_rng(1)
    s3 = (d.jpi, d.jpj, d.jpk)
    un, vn = rnd(rng, s3), rnd(rng, s3)
    h = fz(s3)
    o.div_hor(1, un, vn, h)
    fu = A["e2u"][:, :, None] * A["e3u_n"] * un
    fv = A["e1v"][:, :, None] * A["e3v_n"] * vn
    vol = A["e3t_n"] / A["r1_e1e2t"][:, :, None]
    lhs = (h * vol)[1:-1, 1:-1, :-1].sum(axis=(0, 1))
    # net outflow through the faces of the interior block, per level
    rhs = ((fu[-2, 1:-1] - fu[0, 1:-1]).sum(axis=0) + (fv[1:-1, -2] - fv[1:-1, 0]).sum(axis=0))[:-1]
    scale = np.abs(fu).sum(axis=(0, 1))[:-1] + np.abs(fv).sum(axis=(0, 1))[:-1]
    assert np.all(np.abs(lhs - rhs) <= 1e-12 * scale)
    assert_ring_is_lbc(d, h, "hdivn")


def test_div_hor_of_a_streamfunction_flow_is_zero():
    d = rough_domain()
    o = Oracle(d)
    A = d.arrays
    rng = np.random.default_rng(2)
    s3 = (d.jpi, d.jpj, d.jpk)
    psi = rng.standard_normal(s3)
    un, vn = fz(s3), fz(s3)
    # transports e2u*e3u*un = -(psi(i,j) - psi(i,j-1)), e1v*e3v*vn = psi(i,j) - psi(i-1,j)
    un[:, 1:] = -(psi[:, 1:] - psi[:, :-1]) / (A["e2u"][:, 1:, None] * A["e3u_n"][:, 1:])
    vn[1:, :] = (psi[1:, :] - psi[:-1, :]) / (A["e1v"][1:, :, None] * A["e3v_n"][1:, :])
    h = fz(s3)
    o.div_hor(3, un, vn, h)
    scale = (np.abs(psi) * A["r1_e1e2t"][:, :, None] / A["e3t_n"]).max()
    assert np.abs(h[1:-1, 1:-1, :-1]).max() <= 1e-13 * scale
    # a purely divergent flow is not zero: the check above has teeth
    un2 = rnd(rng, s3)
    o.div_hor(3, un2, vn, h)
    assert np.abs(h[1:-1, 1:-1, :-1]).max() > 1e-6 * scale


def test_div_hor_leaves_level_jpk_alone():
    d = rough_domain(7)
    o = Oracle(d)
    rng = np.random.default_rng(3)
    s3 = (d.jpi, d.jpj, d.jpk)
    h = rnd(rng, s3)
    keep = h[1:-1, 1:-1, -1].copy()
    o.div_hor(1, rnd(rng, s3), rnd(rng, s3), h)
    assert np.array_equal(h[1:-1, 1:-1, -1], keep)
    assert_ring_is_lbc(d, h, "hdivn")


# ----------------------------------------------------------------------------
# ssh_nxt, wzv
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("kt,factor", [(1, 1.0), (2, 2.0), (9, 2.0)])
def test_ssh_nxt_fresh_water_only(kt, factor):
    """No flow: ssha = sshb - z2dt * emp / rau0, z2dt = rdt at the Euler step."""
    d = rough_domain()
    o = Oracle(d)
    s2, s3 = (d.jpi, d.jpj), (d.jpi, d.jpj, d.jpk)
    t, _, _ = masks(d)
    ssm = t.max(axis=2)
    rng = np.random.default_rng(4)
    sshb = rnd(rng, s2, ssm)
    emp = np.asfortranarray(np.full(s2, 1e-4))
    ssha = fz(s2)
    o.ssh_nxt(kt, fz(s3), fz(s3), fz(s3), sshb, emp, emp, ssha)
    z2dt = factor * d.opts["rdt"]
    want = (sshb - z2dt * 1e-4 / d.opts["rau0"]) * ssm
    assert np.allclose(ssha, want, rtol=1e-14, atol=1e-15 * z2dt * 1e-4 / d.opts["rau0"])


def test_ssh_nxt_conserves_volume_and_matches_surface_w():
    d = rough_domain()
    o = Oracle(d)
    A = d.arrays
    t, u, v = masks(d)
    ssm = t.max(axis=2)
    rng = np.random.default_rng(5)
    s2, s3 = (d.jpi, d.jpj), (d.jpi, d.jpj, d.jpk)
    un, vn = rnd(rng, s3, u), rnd(rng, s3, v)
    sshb = rnd(rng, s2, ssm)
    h, ssha, wn = fz(s3), fz(s2), fz(s3)
    kt = 4
    o.ssh_nxt(kt, un, vn, h, sshb, fz(s2), fz(s2), ssha)
    area = 1.0 / A["r1_e1e2t"]
    z2dt = 2.0 * d.opts["rdt"]
    dv = ((ssha - sshb) * area).sum()
    scale = z2dt * (np.abs(h) * A["e3t_n"] * area[:, :, None]).sum()
    assert abs(dv) <= 1e-12 * scale
    # wzv: w at the surface is the sea-level tendency
    o.wzv(kt, h, wn)
    wet = ssm > 0
    tend = (ssha - sshb) / z2dt
    assert np.allclose(wn[:, :, 0][wet], tend[wet], rtol=0, atol=1e-12 * np.abs(tend).max())


def test_wzv_closed_form():
    d = rough_domain()
    o = Oracle(d)
    A = d.arrays
    t, _, _ = masks(d)
    s3 = (d.jpi, d.jpj, d.jpk)
    c = 3.0e-7
    h = np.asfortranarray(np.full(s3, c))
    wn = fz(s3)
    o.wzv(1, h, wn)
    # wn(k) = -c * sum_{k' >= k, k' <= jpkm1} e3t tmask
    cum = np.cumsum((A["e3t_n"] * t)[:, :, :-1][:, :, ::-1], axis=2)[:, :, ::-1]
    assert np.allclose(wn[:, :, :-1], -c * cum, rtol=1e-13, atol=0)
    assert np.all(wn[:, :, -1] == 0.0)
    # after nit000 the bottom value is kept, not reset
    wn[:, :, -1] = 0.5
    o.wzv(5, h, wn)
    assert np.all(wn[:, :, -1] == 0.5)
    assert np.allclose(wn[:, :, :-1], 0.5 - c * cum, rtol=1e-13)


# ----------------------------------------------------------------------------
# Asselin filters: ssh_swp, tra_nxt
# ----------------------------------------------------------------------------
def dyadic(rng, shape, lo=-64, hi=64):
    """Values k/8: sums and doublings of them are exact in FP64."""
    return np.asfortranarray(rng.integers(lo, hi, shape) / 8.0)


def test_ssh_swp_filter():
    d = rough_domain()
    o = Oracle(d)
    atfp = d.opts["rn_atfp"]
    rng = np.random.default_rng(6)
    s2 = (d.jpi, d.jpj)
    n, dd = dyadic(rng, s2), dyadic(rng, s2)
    # linear sequence: the filter returns the middle value exactly
    b, nn, a = np.asfortranarray(n - dd), n.copy(order="F"), np.asfortranarray(n + dd)
    o.ssh_swp(3, b, nn, a)
    assert np.array_equal(b, n) and np.array_equal(nn, a)
    # curvature q: the filtered value is n + 2 atfp q
    q = dyadic(rng, s2)
    b, nn, a = np.asfortranarray(n + q), n.copy(order="F"), np.asfortranarray(n + q)
    o.ssh_swp(3, b, nn, a)
    assert np.allclose(b, n + 2 * atfp * q, rtol=1e-15, atol=1e-15)
    # Euler step: before is kept
    b0 = dyadic(rng, s2)
    b, nn = b0.copy(order="F"), n.copy(order="F")
    o.ssh_swp(1, b, nn, a)
    assert np.array_equal(b, b0) and np.array_equal(nn, a)


@pytest.mark.parametrize("jperio", [0, 7])
def test_tra_nxt_filter_and_swap(jperio):
    d = rough_domain(jperio)
    o = Oracle(d)
    atfp = d.opts["rn_atfp"]
    rng = np.random.default_rng(7)
    s4 = (d.jpi, d.jpj, d.jpk, 2)
    n, q = dyadic(rng, s4), dyadic(rng, s4)
    b, nn, a = np.asfortranarray(n - q), n.copy(order="F"), np.asfortranarray(n + q)
    b_jpk, n_jpk = b[:, :, -1].copy(), nn[:, :, -1].copy()
    o.tra_nxt(3, b, nn, a)
    I = (slice(1, -1), slice(1, -1), slice(0, -1))
    assert np.array_equal(b[I], n[I])            # linear: filtered = now
    assert np.array_equal(nn[I], (n + q)[I])     # now <- after
    assert np.array_equal(b[1:-1, 1:-1, -1], b_jpk[1:-1, 1:-1])   # level jpk untouched
    assert np.array_equal(nn[1:-1, 1:-1, -1], n_jpk[1:-1, 1:-1])
    for arr, nm in ((b, "tsb"), (nn, "tsn"), (a, "tsa")):
        for m in range(2):
            assert_ring_is_lbc(d, arr[..., m], nm)
    # curvature
    b, nn, a = np.asfortranarray(n + q), n.copy(order="F"), np.asfortranarray(n + q)
    o.tra_nxt(3, b, nn, a)
    assert np.allclose(b[I], (n + 2 * atfp * q)[I], rtol=1e-15, atol=1e-15)


def test_tra_nxt_euler_step():
    d = rough_domain(7)
    o = Oracle(d)
    rng = np.random.default_rng(8)
    s4 = (d.jpi, d.jpj, d.jpk, 2)
    b0, n0, a0 = dyadic(rng, s4), dyadic(rng, s4), dyadic(rng, s4)
    b, nn, a = b0.copy(order="F"), n0.copy(order="F"), a0.copy(order="F")
    o.tra_nxt(1, b, nn, a)
    assert np.array_equal(b, b0)
    for m in range(2):
        assert_ring_is_lbc(d, a[..., m], "tsa")
    assert np.array_equal(nn[:, :, :-1], a[:, :, :-1])
    assert np.array_equal(nn[:, :, -1], n0[:, :, -1])


# ----------------------------------------------------------------------------
# dyn_nxt
# ----------------------------------------------------------------------------
class Dyn:
    def __init__(self, jperio=0, seed=9, **opts):
        self.d = rough_domain(jperio, opts=opts)
        self.o = Oracle(self.d)
        t, u, v = masks(self.d)
        self.u, self.v = u, v
        A = self.d.arrays
        hu = (A["e3u_n"] * u)[:, :, :-1].sum(axis=2)
        hv = (A["e3v_n"] * v)[:, :, :-1].sum(axis=2)
        su, sv = (u.max(axis=2) > 0) * 1.0, (v.max(axis=2) > 0) * 1.0
        self.r1_hu = np.asfortranarray(su / (hu + 1.0 - su))
        self.r1_hv = np.asfortranarray(sv / (hv + 1.0 - sv))
        self.wu, self.wv = su > 0, sv > 0
        self.rng = np.random.default_rng(seed)
        self.s2, self.s3 = (self.d.jpi, self.d.jpj), (self.d.jpi, self.d.jpj, self.d.jpk)

    def mean_u(self, x):
        A = self.d.arrays
        return (A["e3u_n"] * x * self.u)[:, :, :-1].sum(axis=2) * self.r1_hu

    def mean_v(self, x):
        A = self.d.arrays
        return (A["e3v_n"] * x * self.v)[:, :, :-1].sum(axis=2) * self.r1_hv

    def fields(self):
        r = self.rng
        f = {k: rnd(r, self.s3, self.u if k[0] == "u" else self.v) for k in ("ub", "vb", "un", "vn", "ua", "va")}
        for k in ("ua_b", "va_b", "un_adv", "vn_adv", "ub_b", "vb_b", "un_b", "vn_b"):
            f[k] = rnd(r, self.s2)
        return f

    def run(self, kt, f):
        self.o.dyn_nxt(kt, f["ub"], f["vb"], f["un"], f["vn"], f["ua"], f["va"], f["ua_b"], f["va_b"], f["un_adv"],
                       f["vn_adv"], f["ub_b"], f["vb_b"], f["un_b"], f["vn_b"], self.r1_hu, self.r1_hv)


def test_dyn_nxt_filter_without_time_splitting():
    e = Dyn(ln_dynspg_ts=0)
    atfp = e.d.opts["rn_atfp"]
    n, q = dyadic(e.rng, e.s3) * e.u, dyadic(e.rng, e.s3) * e.u
    f = e.fields()
    f["ub"], f["un"], f["ua"] = np.asfortranarray(n - q), n.copy(order="F"), np.asfortranarray(n + q)
    f["vb"], f["vn"], f["va"] = f["ub"] * 0, f["un"] * 0, f["ua"] * 0
    e.run(4, f)
    K = slice(0, -1)
    assert np.array_equal(f["ub"][:, :, K], n[:, :, K])
    assert np.array_equal(f["un"][:, :, K], (n + q)[:, :, K])
    # barotropic now/before velocities are the depth means (a constant column -> itself)
    c = 0.75
    f = e.fields()
    f["ua"] = np.asfortranarray(np.full(e.s3, c) * e.u)
    f["ub"] = np.asfortranarray(np.full(e.s3, c) * e.u)
    f["un"] = f["ua"].copy(order="F")
    e.run(4, f)
    assert np.allclose(f["un_b"][e.wu], c, rtol=1e-14) and np.allclose(f["ub_b"][e.wu], c, rtol=1e-14)
    # curvature
    f = e.fields()
    f["ub"], f["un"], f["ua"] = np.asfortranarray(n + q), n.copy(order="F"), np.asfortranarray(n + q)
    e.run(4, f)
    assert np.allclose(f["ub"][1:-1, 1:-1, K], (n + 2 * atfp * q)[1:-1, 1:-1, K], rtol=1e-15, atol=1e-15)


@pytest.mark.parametrize("jperio", [0, 7])
def test_dyn_nxt_forward_time_splitting_depth_means(jperio):
    """ln_bt_fw: the new now velocity has the depth mean ua_b, the filtered
    before velocity the depth mean of the previous un_b."""
    e = Dyn(jperio, ln_dynspg_ts=1, ln_bt_fw=1)
    f = e.fields()
    un_b0, ua_b = f["un_b"].copy(), f["ua_b"].copy()
    e.run(5, f)
    w = e.wu.copy()
    w[0, :] = w[-1, :] = False
    w[:, 0] = w[:, -1] = False
    assert np.allclose(f["un_b"][w], ua_b[w], rtol=0, atol=1e-13 * np.abs(ua_b).max())
    assert np.allclose(e.mean_u(f["un"])[w], ua_b[w], rtol=0, atol=1e-13 * np.abs(ua_b).max())
    assert np.allclose(f["ub_b"][w], un_b0[w], rtol=0, atol=1e-13 * np.abs(un_b0).max())
    for nm in ("ua", "va", "un", "vn"):
        assert_ring_is_lbc(e.d, f[nm][:, :, :-1], nm)


def test_dyn_nxt_centred_time_splitting():
    """ln_bt_fw = F: before the filter, un gets back the depth mean it had
    before dyn_spg_ts replaced it by un_adv (mean - un_adv/hu + un_b), so the
    depth mean of the filtered velocity is the filter of the depth means."""
    e = Dyn(ln_dynspg_ts=1, ln_bt_fw=0)
    atfp = e.d.opts["rn_atfp"]
    f = e.fields()
    g = {k: v.copy(order="F") for k, v in f.items()}
    e.run(5, f)
    w = e.wu.copy()
    w[0, :] = w[-1, :] = False
    w[:, 0] = w[:, -1] = False
    mun = e.mean_u(g["un"]) - g["un_adv"] * e.r1_hu + g["un_b"]
    want = (1 - 2 * atfp) * mun + atfp * (e.mean_u(g["ub"]) + g["ua_b"])
    assert np.allclose(e.mean_u(f["ub"])[w], want[w], rtol=0, atol=1e-12 * np.abs(want).max())


def test_dyn_nxt_euler_step():
    e = Dyn(7, ln_dynspg_ts=1, ln_bt_fw=1)
    f = e.fields()
    ub0 = f["ub"].copy(order="F")
    ua_b = f["ua_b"].copy()
    e.run(1, f)
    assert np.array_equal(f["ub"], ub0)
    assert np.array_equal(f["un"][:, :, :-1], f["ua"][:, :, :-1])
    w = e.wu.copy()
    w[0, :] = w[-1, :] = False
    w[:, 0] = w[:, -1] = False
    assert np.allclose(f["un_b"][w], ua_b[w], rtol=0, atol=1e-13 * np.abs(ua_b).max())
    assert np.allclose(f["ub_b"], e.mean_u(ub0), rtol=1e-13, atol=1e-16)
