"""CPU checks that pin the oracle's two-band penetrative solar flux
(tra_qsr, TRA/traqsr.F90:134-276) and its extinction level
(trc_oce_ext_lev, trc_oce.F90) to closed forms:

* the heat contents of levels 1..nksr telescope: in a column deeper than
  the light reaches they sum to Qsr/(rau0 rcp) (1 - I(gdepw(nksr+1))/Qsr),
  with I(z)/Qsr = rn_abs e^(-z/rn_si0) + (1 - rn_abs) e^(-z/rn_si1), the
  documented two-band profile (traqsr.F90:87); in a column whose bottom is
  above that depth the bottom level absorbs the rest, so they sum to
  Qsr/(rau0 rcp);
* fraqsr_1lev = 1 - I(gdepw(2))/Qsr, and 1 where Qsr = 0;
* an Euler first step adds qsr_hc/e3t to the temperature trend of levels
  1..nksr only; a second step with the same forcing adds the same (the mean
  of before and now, qsr_hc_b = the first step's qsr_hc); a restart start
  (kqsr_b = 1) adds the mean of the given before field and the new one;
* nksr is the shallowest level whose w-depth below reaches
  rn_si1 (15 ln 10 + ln 100), found here by a search of the 1D profile.
"""
import numpy as np

from nemo_hip import configs
from oracle.oracle import Oracle

RCP = 3991.86795711963
ABS, SI0, SI1 = 0.58, 0.35, 23.0


def profile(z):
    return ABS * np.exp(-z / SI0) + (1.0 - ABS) * np.exp(-z / SI1)


def setup(depth, jpk):
    d = configs.basin(16, 14, jpk, seed=3, land_frac=0.0, depth_min=depth, depth_max=depth)
    o = Oracle(d)
    gdepw = configs.depths(d)[1]
    rng = np.random.default_rng(9)
    qsr = np.asfortranarray(rng.uniform(50.0, 300.0, (d.jpi, d.jpj)))
    qsr[4, 5] = 0.0
    return d, o, gdepw, qsr


def run(d, o, gdepw, qsr, nksr, kt=1, kqsr_b=0, hc=None, hc_b=None, tsa=None):
    s3 = (d.jpi, d.jpj, d.jpk)
    hc = np.zeros(s3, order="F") if hc is None else hc
    hc_b = np.zeros(s3, order="F") if hc_b is None else hc_b
    tsa = np.zeros(s3 + (2,), order="F") if tsa is None else tsa
    fr = np.zeros((d.jpi, d.jpj), order="F")
    o.tra_qsr(kt, nksr, qsr, gdepw, hc, hc_b, tsa, fr, kqsr_b=kqsr_b)
    return hc, hc_b, tsa, fr


def test_nksr_is_the_extinction_level():
    d, o, gdepw, _ = setup(5000.0, 31)
    zhext = SI1 * (np.log(10.0) * 15.0 + np.log(100.0))
    w = gdepw[5, 5, :]
    assert w[-1] > zhext
    assert o.qsr_nksr(gdepw) == int(np.argmax(w >= zhext))
    # a basin shallower than the extinction depth: every wet level
    d2, o2, g2, _ = setup(300.0, 12)
    assert o2.qsr_nksr(g2) == d2.jpk - 1


def test_heat_content_telescopes_and_level_one_fraction():
    for depth, jpk in ((5000.0, 31), (300.0, 12)):
        d, o, gdepw, qsr = setup(depth, jpk)
        r1 = 1.0 / (d.opts["rau0"] * RCP)
        nksr = o.qsr_nksr(gdepw)
        hc, _, _, fr = run(d, o, gdepw, qsr, nksr)
        wet = configs.masks(d)[0][1:-1, 1:-1, 0] > 0
        tot = hc[1:-1, 1:-1, :nksr].sum(axis=2)
        q = qsr[1:-1, 1:-1]
        if depth > 1000.0:
            want = q * r1 * (1.0 - profile(gdepw[1:-1, 1:-1, nksr]))
        else:
            want = q * r1
        assert wet.sum() > 100
        assert np.allclose(tot[wet], want[wet], rtol=1e-12, atol=0.0)
        assert not tot[~wet].any() and not hc[:, :, nksr:].any()
        f1 = 1.0 - profile(gdepw[1:-1, 1:-1, 1])
        ok = (q != 0.0) & wet
        assert np.allclose(fr[1:-1, 1:-1][ok], f1[ok], rtol=1e-13)
        assert fr[4, 5] == 1.0


def test_time_stepping():
    d, o, gdepw, qsr = setup(5000.0, 31)
    nksr = o.qsr_nksr(gdepw)
    e3t = d.arrays["e3t_n"]
    hc1, hb1, ta1, _ = run(d, o, gdepw, qsr, nksr, kt=1)
    assert not hb1.any()
    inc = np.zeros_like(e3t)
    inc[1:-1, 1:-1, :nksr] = hc1[1:-1, 1:-1, :nksr] / e3t[1:-1, 1:-1, :nksr]
    assert np.allclose(ta1[..., 0], inc, rtol=1e-14, atol=0.0) and not ta1[..., 1].any()
    hc2, hb2, ta2, _ = run(d, o, gdepw, qsr, nksr, kt=2, hc=hc1.copy(order="F"))
    assert np.array_equal(hb2, hc1) and np.array_equal(hc2, hc1)
    assert np.allclose(ta2[..., 0], inc, rtol=1e-14, atol=0.0)
    B = np.asfortranarray(np.full_like(hc1, 1e-5))
    hc3, hb3, ta3, _ = run(d, o, gdepw, qsr, nksr, kt=1, kqsr_b=1, hc_b=B.copy(order="F"))
    want = np.zeros_like(e3t)
    want[1:-1, 1:-1, :nksr] = 0.5 * (1e-5 + hc1[1:-1, 1:-1, :nksr]) / e3t[1:-1, 1:-1, :nksr]
    assert np.array_equal(hb3, B) and np.allclose(ta3[..., 0], want, rtol=1e-14, atol=0.0)


def test_sbc_then_qsr_with_ln_traqsr():
    """ln_traqsr = 1 (trasbc.F90:97-100): tra_sbc leaves qns and qsr as
    they are and its level-1 heat trend is that of qns alone; tra_sbc then
    tra_qsr put the column's heat content to qns + Qsr (1 - I(bottom)/Qsr),
    in units of rau0 rcp, where the merged run (ln_traqsr = 0) puts all of
    qns + qsr into level 1."""
    d0 = configs.basin(16, 14, 31, seed=3, land_frac=0.0, depth_min=5000.0, depth_max=5000.0)
    d1 = configs.basin(16, 14, 31, seed=3, land_frac=0.0, depth_min=5000.0, depth_max=5000.0,
                       opts=dict(ln_traqsr=1))
    gdepw = configs.depths(d1)[1]
    rng = np.random.default_rng(12)
    s2, s3 = (d1.jpi, d1.jpj), (d1.jpi, d1.jpj, d1.jpk)
    qns = np.asfortranarray(rng.uniform(-100.0, 100.0, s2))
    qsr = np.asfortranarray(rng.uniform(50.0, 300.0, s2))
    zero2 = np.zeros(s2, order="F")
    ts = np.zeros(s3 + (2,), order="F")
    ts[..., 0], ts[..., 1] = 10.0, 35.0

    def sbc(d, q, s):
        o = Oracle(d)                            # the oracle holds one registered domain
        a, b = q.copy(order="F"), s.copy(order="F")
        tsa = np.zeros(s3 + (2,), order="F")
        o.tra_sbc(1, a, b, zero2.copy(order="F"), zero2.copy(order="F"), ts, tsa,
                  np.zeros(s2 + (2,), order="F"), np.zeros(s2 + (2,), order="F"))
        return a, b, tsa

    a1, b1, t1 = sbc(d1, qns, qsr)
    assert np.array_equal(a1, qns) and np.array_equal(b1, qsr)
    _, _, t_qns = sbc(d0, qns, zero2)           # merged run with no solar part
    assert np.array_equal(t1, t_qns)
    a0, b0, t0 = sbc(d0, qns, qsr)               # merged: qns + qsr at level 1
    assert np.allclose(a0, qns + qsr, rtol=0, atol=0) and not b0.any()
    o1 = Oracle(d1)
    nksr = o1.qsr_nksr(gdepw)
    fr = np.zeros(s2, order="F")
    o1.tra_qsr(1, nksr, b1, gdepw, np.zeros(s3, order="F"), np.zeros(s3, order="F"), t1, fr)
    e3t = d1.arrays["e3t_n"]
    heat = (t1[..., 0] * e3t).sum(axis=2) * d1.opts["rau0"] * RCP
    want = qns + qsr * (1.0 - profile(gdepw[:, :, nksr]))
    want0 = (t0[..., 0] * e3t).sum(axis=2) * d0.opts["rau0"] * RCP
    wet = d1.arrays["k_bot"] > 0                 # wet columns (a land column takes no solar heat)
    wet[0, :] = wet[-1, :] = False
    wet[:, 0] = wet[:, -1] = False
    assert wet.sum() > 100
    assert np.allclose(heat[wet], want[wet], rtol=1e-12, atol=1e-9)
    assert np.allclose(want0[wet], (qns + qsr)[wet], rtol=1e-12, atol=1e-9)
