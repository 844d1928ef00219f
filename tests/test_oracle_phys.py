"""CPU checks that pin the oracle's density, pressure-gradient and lateral
viscosity routines to what the mathematics fixes:

* S-EOS: zero anomaly at the reference state (T = 10, S = 35); the linear
  case in closed form; the in situ minus potential density is the
  thermobaric term; the expansion ratios of rab_3d are the exact T and S
  derivatives of eos_insitu (central differences are exact on the quadratic
  polynomial);
* bn2: N^2 of a linear stratification with uniform ratios in closed form,
  and the w-point interpolation exact on ratios linear in depth;
* hpg_zco: no gradient for a horizontally uniform density; the closed form
  g c z of a density linear in x;
* dyn_ldf_lap: the discrete energy identity (the kinetic-energy tendency
  equals minus the dissipation of the divergence and curl parts), no curl
  part for a potential flow and no divergence part for a streamfunction
  flow.
"""
import numpy as np
import pytest

from nemo_hip import abi, configs
from oracle.oracle import Oracle


def fz(shape):
    return np.zeros(shape, order="F")


def domain(jperio=0, seed=11):
    """A seeded basin whose u- and v-point lengths vary in both directions
    (their inverses kept consistent), so the identities below only hold with
    the right indices."""
    d = configs.basin(24, 18, 9, seed=seed, jperio=jperio)
    rng = np.random.default_rng(seed + 1000)
    A = d.arrays
    for n in ("e2u", "e1v", "e1u", "e2v"):
        A[n] = np.asfortranarray(A[n] * rng.uniform(0.8, 1.25, A[n].shape))
    for n in ("e1u", "e2u", "e1v", "e2v"):
        A["r1_" + n] = np.asfortranarray(1.0 / A[n])
    return d


def with_eos(d, **kw):
    o = Oracle(d)
    o.eos_register(abi.seos(**kw))
    return o


# ----------------------------------------------------------------------------
def test_seos_reference_state_and_linear_case():
    d = domain()
    t, u, v, w, f = configs.masks(d)
    s3 = (d.jpi, d.jpj, d.jpk)
    dep = np.asfortranarray(np.random.default_rng(1).uniform(0, 5000, s3))
    ts = np.zeros(s3 + (2,), order="F")
    ts[..., 0], ts[..., 1] = 10.0, 35.0
    o = with_eos(d)
    prd, prhop = fz(s3), fz(s3)
    o.eos_insitu(ts, prd, dep)
    assert np.all(prd[:, :, :-1] == 0.0)
    o.eos_insitu_pot(ts, prd, prhop, dep)
    assert np.all(prd[:, :, :-1] == 0.0)
    assert np.array_equal(prhop[:, :, :-1], d.opts["rau0"] * t[:, :, :-1])
    # linear EOS: only rn_a0, rn_b0
    rng = np.random.default_rng(2)
    ts[..., 0] = rng.uniform(-2, 30, s3)
    ts[..., 1] = rng.uniform(30, 40, s3)
    o = with_eos(d, rn_lambda1=0.0, rn_lambda2=0.0, rn_mu1=0.0, rn_mu2=0.0, rn_nu=0.0)
    o.eos_insitu(ts, prd, dep)
    want = (-0.1655 * (ts[..., 0] - 10) + 0.76554 * (ts[..., 1] - 35)) / d.opts["rau0"] * t
    assert np.allclose(prd[:, :, :-1], want[:, :, :-1], rtol=1e-13, atol=1e-16)


def test_seos_thermobaric_term_and_expansion_ratios():
    d = domain()
    t, u, v, w, f = configs.masks(d)
    s3 = (d.jpi, d.jpj, d.jpk)
    rng = np.random.default_rng(3)
    dep = np.asfortranarray(rng.uniform(0, 5000, s3))
    ts = np.zeros(s3 + (2,), order="F")
    ts[..., 0] = rng.uniform(-2, 30, s3)
    ts[..., 1] = rng.uniform(30, 40, s3)
    o = with_eos(d)
    e = abi.SEOS_DEFAULTS
    rau0 = d.opts["rau0"]
    prd, prd2, prhop = fz(s3), fz(s3), fz(s3)
    o.eos_insitu(ts, prd, dep)
    o.eos_insitu_pot(ts, prd2, prhop, dep)
    K = slice(0, -1)
    assert np.allclose(prd[:, :, K], prd2[:, :, K], rtol=1e-13, atol=1e-17)
    zt, zs = ts[..., 0] - 10, ts[..., 1] - 35
    thermo = -(e["rn_a0"] * e["rn_mu1"] * zt + e["rn_b0"] * e["rn_mu2"] * zs) * dep / rau0 * t
    assert np.allclose((prd - (prhop - rau0 * t) / rau0)[:, :, K], thermo[:, :, K], rtol=1e-9, atol=1e-15)
    # rab_3d = (-d prd / dT, d prd / dS) at the t-point depth, exactly for the quadratic S-EOS
    pab = fz(s3 + (2,))
    o.rab_3d(ts, pab, dep)
    h = 0.5
    for n, sign in ((0, -1.0), (1, 1.0)):
        tp, tm = ts.copy(order="F"), ts.copy(order="F")
        tp[..., n] += h
        tm[..., n] -= h
        rp, rm = fz(s3), fz(s3)
        o.eos_insitu(tp, rp, dep)
        o.eos_insitu(tm, rm, dep)
        fd = sign * (rp - rm) / (2 * h)
        assert np.allclose(pab[:, :, K, n], fd[:, :, K], rtol=1e-9, atol=1e-12 * np.abs(fd).max())


def test_bn2_closed_forms():
    d = domain()
    t, u, v, w, f = configs.masks(d)
    s3 = (d.jpi, d.jpj, d.jpk)
    gdept, gdepw = configs.depths(d)
    o = with_eos(d)
    g = d.opts["grav"]
    alpha, beta = 2.0e-4, 7.6e-4
    pab = np.zeros(s3 + (2,), order="F")
    pab[..., 0], pab[..., 1] = alpha, beta
    ts = np.zeros(s3 + (2,), order="F")
    ts[..., 0] = 20.0 - 0.004 * gdept       # dT/dz = -0.004
    ts[..., 1] = 34.0 + 0.0005 * gdept      # dS/dz = +0.0005
    pn2 = fz(s3)
    o.bn2(ts, pab, pn2, gdept, gdepw)
    e3w = d.arrays["e3w_n"]
    # T(k-1) - T(k) = 0.004 * e3w(k), S(k-1) - S(k) = -0.0005 * e3w(k)
    want = g * (alpha * 0.004 + beta * 0.0005) * w
    K = slice(1, -1)
    assert np.allclose(pn2[:, :, K], want[:, :, K], rtol=1e-10, atol=1e-18)
    # ratios linear in depth: the w-point weight interpolates them exactly
    pab[..., 0] = 1.0e-4 + 3.0e-8 * gdept
    pab[..., 1] = 7.0e-4
    o.bn2(ts, pab, pn2, gdept, gdepw)
    want = g * ((1.0e-4 + 3.0e-8 * gdepw) * 0.004 + 7.0e-4 * 0.0005) * w
    assert np.allclose(pn2[:, :, K], want[:, :, K], rtol=1e-10, atol=1e-18)
    assert np.all(pn2[:, :, 0] == 0) and np.all(pn2[:, :, -1] == 0)


def test_hpg_zco_closed_forms():
    d = domain()
    o = Oracle(d)
    A = d.arrays
    s3 = (d.jpi, d.jpj, d.jpk)
    rng = np.random.default_rng(4)
    rhd = np.asfortranarray(np.broadcast_to(rng.standard_normal(d.jpk)[None, None, :], s3).copy())
    ua, va = fz(s3), fz(s3)
    o.hpg_zco(1, rhd, ua, va)
    assert np.all(ua == 0) and np.all(va == 0)
    # density linear in x: rhd(i+1) - rhd(i) = c / r1_e1u(i) (the metric the routine divides by)
    c = 1.0e-8
    x = np.concatenate([np.zeros((1, d.jpj)), np.cumsum(1.0 / A["r1_e1u"][:-1, :], axis=0)], axis=0)
    rhd = np.asfortranarray(np.broadcast_to((c * x)[:, :, None], s3).copy())
    ua, va = fz(s3), fz(s3)
    o.hpg_zco(1, rhd, ua, va)
    g = d.opts["grav"]
    e3w = A["e3w_n"]
    depth = 0.5 * e3w[:, :, :1] + np.concatenate([np.zeros(e3w.shape[:2] + (1,)), np.cumsum(e3w[:, :, 1:], axis=2)],
                                                  axis=2)
    want = -g * c * depth
    I = (slice(1, -1), slice(1, -1), slice(0, -1))
    assert np.allclose(ua[I], want[I], rtol=1e-10, atol=0)
    # this field also varies in j (rows differ): the j-gradient in the same closed form
    dj = np.zeros((d.jpi, d.jpj))
    dj[:, :-1] = c * (x[:, 1:] - x[:, :-1]) * A["r1_e2v"][:, :-1]
    want_v = -g * depth * dj[:, :, None]
    assert np.allclose(va[I], want_v[I], rtol=1e-8, atol=1e-9 * np.abs(want_v[I]).max())


class Ldf:
    def __init__(self):
        self.d = domain()
        self.o = Oracle(self.d)
        self.t, self.u, self.v, self.w, self.f = configs.masks(self.d)
        self.s3 = (self.d.jpi, self.d.jpj, self.d.jpk)
        rng = np.random.default_rng(5)
        self.ahmt = np.asfortranarray(rng.uniform(500, 2000, self.s3) * self.t)
        self.ahmf = np.asfortranarray(rng.uniform(500, 2000, self.s3) * self.f)
        self.rng = rng

    def trend(self, ub, vb, kpass=1):
        ua, va = fz(self.s3), fz(self.s3)
        self.o.dyn_ldf_lap(3, ub, vb, ua, va, kpass, self.ahmt, self.ahmf)
        return ua, va


def test_dyn_ldf_lap_energy_identity():
    e = Ldf()
    A = e.d.arrays
    ub = np.asfortranarray(e.rng.standard_normal(e.s3) * e.u)
    vb = np.asfortranarray(e.rng.standard_normal(e.s3) * e.v)
    ua, va = e.trend(ub, vb)
    e3u, e3v, e3t, e3f = A["e3u_n"], A["e3v_n"], A["e3t_n"], A["e3f_n"]
    K = slice(0, -1)
    # kinetic-energy tendency, weighted by the u / v cell volumes
    dke = ((A["e1u"] * A["e2u"])[:, :, None] * e3u * ub * ua)[:, :, K].sum() + \
          ((A["e1v"] * A["e2v"])[:, :, None] * e3v * vb * va)[:, :, K].sum()
    # flux divergence D at t-points and circulation C at f-points
    fu = A["e2u"][:, :, None] * e3u * ub
    fv = A["e1v"][:, :, None] * e3v * vb
    D = fz(e.s3)
    D[1:, 1:] = fu[1:, 1:] - fu[:-1, 1:] + fv[1:, 1:] - fv[1:, :-1]
    cu = A["e1u"][:, :, None] * ub
    cv = A["e2v"][:, :, None] * vb
    Cf = fz(e.s3)
    Cf[:-1, :-1] = cv[1:, :-1] - cv[:-1, :-1] - cu[:-1, 1:] + cu[:-1, :-1]
    diss = -((e.ahmt * A["r1_e1e2t"][:, :, None] / e3t * D * D)[:, :, K].sum() +
             (e.ahmf * e3f * A["r1_e1e2f"][:, :, None] * Cf * Cf)[:, :, K].sum())
    assert diss < 0
    assert abs(dke - diss) <= 1e-10 * abs(diss)
    # kpass 2 flips the sign
    ua2, va2 = e.trend(ub, vb, kpass=2)
    assert np.array_equal(ua2, -ua) and np.array_equal(va2, -va)


def test_dyn_ldf_lap_potential_and_streamfunction_flows():
    e = Ldf()
    A = e.d.arrays
    s3 = e.s3
    # potential flow: e1u u = phi(i+1) - phi(i), e2v v = phi(j+1) - phi(j) -> no curl
    phi = e.rng.standard_normal(s3)
    ub, vb = fz(s3), fz(s3)
    ub[:-1] = (phi[1:] - phi[:-1]) / A["e1u"][:-1, :, None]
    vb[:, :-1] = (phi[:, 1:] - phi[:, :-1]) / A["e2v"][:, :-1, None]
    ua, va = e.trend(ub, vb)
    ahmf0 = e.ahmf.copy(order="F")
    e.ahmf[:] = 0.0                      # the same trend with the curl part switched off
    ua0, va0 = e.trend(ub, vb)
    assert np.allclose(ua, ua0, rtol=0, atol=1e-12 * np.abs(ua0).max())
    assert np.allclose(va, va0, rtol=0, atol=1e-12 * np.abs(va0).max())
    e.ahmf[:] = ahmf0
    # streamfunction flow: e2u e3u u = -(psi(j) - psi(j-1)), e1v e3v v = psi(i) - psi(i-1) -> no divergence
    psi = e.rng.standard_normal(s3)
    ub, vb = fz(s3), fz(s3)
    ub[:, 1:] = -(psi[:, 1:] - psi[:, :-1]) / (A["e2u"][:, 1:, None] * A["e3u_n"][:, 1:])
    vb[1:, :] = (psi[1:, :] - psi[:-1, :]) / (A["e1v"][1:, :, None] * A["e3v_n"][1:, :])
    ua, va = e.trend(ub, vb)
    e.ahmt[:] = 0.0
    ua0, va0 = e.trend(ub, vb)
    assert np.allclose(ua, ua0, rtol=0, atol=1e-12 * np.abs(ua0).max())
    assert np.allclose(va, va0, rtol=0, atol=1e-12 * np.abs(va0).max())
    assert np.abs(ua0).max() > 0


# ----------------------------------------------------------------------------
def test_stp_ctl_maxima_tests_and_positions():
    d = domain()
    o = Oracle(d)
    t, u, v, w, f = configs.masks(d)
    s2, s3 = (d.jpi, d.jpj), (d.jpi, d.jpj, d.jpk)
    rng = np.random.default_rng(9)
    ssh = np.asfortranarray(rng.uniform(-1, 1, s2))
    un = np.asfortranarray(rng.uniform(-1, 1, s3))
    ts = np.zeros(s3 + (2,), order="F")
    ts[..., 0] = rng.uniform(0, 20, s3) * t
    ts[..., 1] = rng.uniform(33, 36, s3) * t
    st, zmax, loc, kind = o.stp_ctl(5, ssh, un, ts)
    assert st == 0 and kind == 0
    wet = t == 1
    assert zmax[0] == np.abs(ssh).max() and zmax[1] == np.abs(un).max()
    assert zmax[2] == -ts[..., 1][wet].min() and zmax[3] == ts[..., 1][wet].max()
    assert zmax[4] == -ts[..., 0][wet].min() and zmax[5] == ts[..., 0][wet].max()
    # a velocity of 12 m/s at (4,5,2) and a fresh cell at the first wet point in Fortran order
    un[3, 4, 1] = -12.0
    iw = np.argwhere(wet.transpose(2, 1, 0))[0][::-1]      # (i,j,k) 0-based, first in Fortran order
    ts[iw[0], iw[1], iw[2], 1] = -0.5
    st, zmax, loc, kind = o.stp_ctl(5, ssh, un, ts)
    assert st == 5 and kind == -3
    assert zmax[1] == 12.0 and list(loc[2:5]) == [4, 5, 2]
    assert zmax[2] == 0.5 and list(loc[5:8]) == [iw[0] + 1, iw[1] + 1, iw[2] + 1]
    # a NaN in the sea level is caught and located
    un[3, 4, 1] = 0.0
    ts[iw[0], iw[1], iw[2], 1] = 35.0
    ssh[7, 3] = np.nan
    st, zmax, loc, kind = o.stp_ctl(5, ssh, un, ts)
    assert st == 5 and np.isnan(zmax[0]) and list(loc[0:2]) == [8, 4]


# ----------------------------------------------------------------------------
def test_tra_sbc_closed_forms():
    d = domain()
    o = Oracle(d)
    s2, s3 = (d.jpi, d.jpj), (d.jpi, d.jpj, d.jpk)
    rau0, rcp = d.opts["rau0"], 3991.86795711963
    e3 = d.arrays["e3t_n"][:, :, 0]
    R = (slice(1, -1), slice(1, None))          # ji = 2..jpim1, jj = 2..jpj
    rng = np.random.default_rng(10)
    tsn = np.zeros(s3 + (2,), order="F")
    tsn[..., 0] = rng.uniform(0, 20, s3)
    tsn[..., 1] = rng.uniform(33, 36, s3)

    def run(kt, qns, qsr, sfx, emp, sbc, sbc_b):
        tsa = np.zeros(s3 + (2,), order="F")
        o.tra_sbc(kt, qns, qsr, sfx, emp, tsn, tsa, sbc, sbc_b)
        return tsa

    Q, Qs, F = 150.0, 50.0, 1.0e-3
    qns, qsr = np.full(s2, Q, order="F"), np.full(s2, Qs, order="F")
    sfx, emp = np.full(s2, F, order="F"), np.zeros(s2, order="F")
    sbc, sbc_b = np.zeros(s2 + (2,), order="F"), np.zeros(s2 + (2,), order="F")
    tsa = run(1, qns, qsr, sfx, emp, sbc, sbc_b)            # Euler start: zfact = 1, no before trend
    assert np.all(qns == Q + Qs) and np.all(qsr == 0)
    assert np.allclose(tsa[R + (0, 0)], (Q + Qs) / (rau0 * rcp) / e3[R], rtol=1e-14)
    assert np.allclose(tsa[R + (0, 1)], F / rau0 / e3[R], rtol=1e-14)
    out = np.ones(s2, bool)
    out[R] = False
    assert np.all(tsa[:, :, 0][out] == 0) and np.all(tsa[:, :, 1:] == 0)
    # next step, same forcing: the trapezoid of equal trends is the trend itself
    qns[:], qsr[:] = Q, Qs
    tsa2 = run(2, qns, qsr, sfx, emp, sbc, sbc_b)
    assert np.allclose(tsa2[R + (0,)], tsa[R + (0,)], rtol=1e-14)
    # constant-volume concentration/dilution by the fresh-water flux
    qns[:], qsr[:], sfx[:] = 0.0, 0.0, 0.0
    emp[:] = 2.0e-5
    sbc[:], sbc_b[:] = 0.0, 0.0
    tsa3 = run(1, qns, qsr, sfx, emp, sbc, sbc_b)
    for n in range(2):
        assert np.allclose(tsa3[R + (0, n)], 2.0e-5 / rau0 * tsn[R + (0, n)] / e3[R], rtol=1e-14)


# ----------------------------------------------------------------------------
# polynomial TEOS-10 / EOS-80 (TRA/eosbn2.F90:236-274, 399-437, 588-643)
# ----------------------------------------------------------------------------
POLY = [("teos10", abi.teos10, 0, 1028.21993233072), ("eos80", abi.eos80, 1, 1028.35011066567)]
DEG_RHO, DEG_AB = (6, 4, 2, 1), (5, 3, 1, 0)


@pytest.mark.parametrize("name,mk,scheme,check", POLY, ids=[p[0] for p in POLY])
def test_poly_eos_check_value(name, mk, scheme, check):
    """eos_insitu at z = 3000, T = 3, S = 35.5 gives the check value the
    reference documents (TRA/eosbn2.F90:205 and 208, quoted to 1e-11 kg/m3);
    every monomial is nonzero there, so a misplaced or mistyped coefficient
    above that size fails."""
    d = domain()
    t, *_ = configs.masks(d)
    s3 = (d.jpi, d.jpj, d.jpk)
    ts = np.zeros(s3 + (2,), order="F")
    ts[..., 0], ts[..., 1] = 3.0, 35.5
    o = Oracle(d)
    o.eos_register(mk())
    prd = np.full(s3, 7.0, order="F")
    o.eos_insitu(ts, prd, np.full(s3, 3000.0, order="F"))
    wet = t[:, :, :-1] > 0
    rho = (prd[:, :, :-1][wet] + 1.0) * d.opts["rau0"]
    assert np.abs(rho - check).max() < 1e-10
    assert np.all(prd[:, :, :-1][~wet] == 0.0)
    assert np.all(prd[:, :, -1] == 7.0)          # level jpk untouched


@pytest.mark.parametrize("name,mk,scheme,check", POLY, ids=[p[0] for p in POLY])
def test_poly_eos_potential_density(name, mk, scheme, check):
    """eos_insitu_pot: its prd is eos_insitu's; prhop does not depend on
    depth and is the in situ density at the surface."""
    d = domain()
    t, *_ = configs.masks(d)
    s3 = (d.jpi, d.jpj, d.jpk)
    rng = np.random.default_rng(21)
    ts = np.zeros(s3 + (2,), order="F")
    ts[..., 0] = rng.uniform(-2, 30, s3)
    ts[..., 1] = rng.uniform(30, 40, s3)
    dep = np.asfortranarray(rng.uniform(0, 5000, s3))
    o = Oracle(d)
    o.eos_register(mk())
    a, b, p = fz(s3), fz(s3), fz(s3)
    o.eos_insitu(ts, a, dep)
    o.eos_insitu_pot(ts, b, p, dep)
    assert np.array_equal(a, b)
    a0, p0 = fz(s3), fz(s3)
    o.eos_insitu_pot(ts, a0, p0, fz(s3))
    assert np.array_equal(p0, p)
    r1 = 1.0 / d.opts["rau0"]
    assert np.array_equal(a0[:, :, :-1], (p0[:, :, :-1] * r1 - 1.0) * t[:, :, :-1])
    assert np.all((p[:, :, :-1][t[:, :, :-1] > 0] > 990) & (p[:, :, :-1][t[:, :, :-1] > 0] < 1035))


@pytest.mark.parametrize("name,mk,scheme,check", POLY, ids=[p[0] for p in POLY])
def test_poly_rab_is_density_derivative(name, mk, scheme, check):
    """rab_3d's alpha = -1/rau0 drho/dT and beta = 1/rau0 drho/dS of
    eos_insitu (central differences, h = 1e-3; the alp/bet tables are the
    derivative tables rounded to 11 digits)."""
    d = domain()
    t, *_ = configs.masks(d)
    s3 = (d.jpi, d.jpj, d.jpk)
    rng = np.random.default_rng(22)
    ts = np.zeros(s3 + (2,), order="F")
    ts[..., 0] = rng.uniform(-2, 30, s3)
    ts[..., 1] = rng.uniform(30, 40, s3)
    dep = np.asfortranarray(rng.uniform(0, 5000, s3))
    o = Oracle(d)
    o.eos_register(mk())
    pab = np.zeros(s3 + (2,), order="F")
    o.rab_3d(ts, pab, dep)
    h = 1e-3

    def rd(dt, dsal):
        x = ts.copy(order="F")
        x[..., 0] += dt
        x[..., 1] += dsal
        r = fz(s3)
        o.eos_insitu(x, r, dep)
        return r

    alpha = -(rd(h, 0) - rd(-h, 0)) / (2 * h)
    beta = (rd(0, h) - rd(0, -h)) / (2 * h)
    sl = (slice(None), slice(None), slice(0, -1))
    assert np.allclose(pab[..., 0][sl], alpha[sl], rtol=1e-7, atol=1e-12)
    assert np.allclose(pab[..., 1][sl], beta[sl], rtol=1e-7, atol=1e-12)
    assert np.all(pab[..., 1][sl][t[sl] > 0] > 5e-4)      # haline contraction ~7.5e-4


@pytest.mark.parametrize("scheme", [0, 1])
def test_poly_tables_are_derivative_tables(scheme):
    """The alp and bet tables are -d/dT and 2 zs d/dS of the rho table in the
    scaled variables (alp_ijk = -(j+1) rho_i(j+1)k r1_T0, bet_ijk = (i+1)
    rho_(i+1)jk r1_S0 / 2), to the 11 digits they are tabulated with, and
    every table is zero beyond its degree."""
    d = domain()
    T = Oracle(d).eos_poly_table(scheme)
    rho, alp, bet = T["rho"], T["alp"], T["bet"]
    for k in range(4):
        for j in range(7):
            for i in range(7):
                if i + j > DEG_RHO[k]:
                    assert rho[k, j, i] == 0.0
                if i + j > DEG_AB[k]:
                    assert alp[k, j, i] == 0.0 and bet[k, j, i] == 0.0
                    continue
                da = -(j + 1) * rho[k, j + 1, i] * T["r1_T0"]
                db = (i + 1) * rho[k, j, i + 1] * T["r1_S0"] / 2
                assert abs(alp[k, j, i] - da) <= 2e-10 * abs(da) + 1e-14, (k, j, i)
                assert abs(bet[k, j, i] - db) <= 2e-10 * abs(db) + 1e-14, (k, j, i)
    assert T["r1_T0"] == 1 / 40 and T["r1_Z0"] == 1e-4


REF_EOS = "/root/reference/src/OCE/TRA/eosbn2.F90"


@pytest.mark.skipif(not __import__("os").path.exists(REF_EOS), reason="reference sources not present")
@pytest.mark.parametrize("scheme", [0, 1])
def test_poly_tables_match_reference_text(scheme):
    """Every coefficient eos_init assigns for the scheme equals the table entry
    (read as text from the reference's source, where it is present)."""
    import re
    lines = open(REF_EOS).read().splitlines()
    start = "CASE( np_teos10 )" if scheme == 0 else "CASE( np_eos80 )"
    stop = "CASE( np_eos80 )" if scheme == 0 else "CASE( np_seos )"
    got, on = {}, False
    for ln in lines[1200:]:
        if start in ln:
            on = True
            continue
        if on and stop in ln:
            break
        m = on and re.match(r"\s*(EOS|ALP|BET|rdeltaS|r1_S0)(\d?)(\d?)(\d?)\s*=\s*([-+0-9.eE]+)_wp(\s*/\s*([-+0-9.eE]+)_wp)?", ln)
        if m:
            v = float(m.group(5)) / (float(m.group(7)) if m.group(7) else 1.0)
            got[(m.group(1),) + tuple(int(c) for c in m.group(2, 3, 4) if c)] = v
    T = Oracle(domain()).eos_poly_table(scheme)
    names = {"EOS": "rho", "ALP": "alp", "BET": "bet"}
    n = 0
    for key, v in got.items():
        if key[0] in names:
            i, j, k = key[1:]
            assert T[names[key[0]]][k, j, i] == v, key
            n += 1
    assert n == 122
    assert T["rdeltaS"] == got[("rdeltaS",)] and T["r1_S0"] == got[("r1_S0",)]
    nz = sum(int(np.count_nonzero(T[x])) for x in names.values())
    assert nz == 122
