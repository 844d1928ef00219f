"""CPU checks that pin the oracle's compensated global sums (glob_sum,
glob_sum_full with DDPDD, lib_fortran.F90:300-332) and dia_hsb
(DIA/diahsb.F90:60-239, linear free surface):

* the double-double sum of terms spanning twenty decades with mixed signs
  and heavy cancellation equals math.fsum (the exactly rounded sum) of the
  same masked terms, where plain summation does not; glob_sum masks with
  tmask_i (wet interior columns), glob_sum_full with tmask_h (interior);
* dia_hsb: at the initial state every content variation vanishes to the
  rounding of its terms and the volume is the wet volume; adding c to sshn
  gives a volume variation c * surf_tot and a heat variation c * sum(T(1)
  surf); adding dT to T everywhere gives a heat variation dT * volume;
  moving heat between two cells at constant e3t T changes nothing; a
  uniform fresh-water flux sums to -emp surf_tot and a uniform surface
  heat trend q to q surf_tot; the surface advective flux of the linear
  free surface to w(1) sum(T(1) surf).
"""
import math

import numpy as np

from nemo_hip import configs
from oracle.oracle import Oracle


def masks(d):
    k_bot = d.arrays["k_bot"]
    h = np.zeros((d.jpi, d.jpj))
    h[1:-1, 1:-1] = 1.0
    return h * (k_bot > 0), h


def test_glob_sum_equals_fsum():
    d = configs.basin(40, 30, 6, seed=5, land_frac=0.3)
    o = Oracle(d)
    rng = np.random.default_rng(3)
    mi, mh = masks(d)
    for trial in range(4):
        mag = 10.0 ** rng.uniform(-10, 10, (d.jpi, d.jpj, 6))
        a = np.asfortranarray(mag * rng.choice([-1.0, 1.0], mag.shape))
        a[3, 4, 2] = 1e16                     # cancelling giants
        a[5, 6, 3] = -1e16
        for nk in (1, 6):
            for kind, m in ((0, mi), (1, mh)):
                terms = (a[:, :, :nk] * m[:, :, None]).ravel(order="F")
                want = math.fsum(terms)
                assert o.glob_sum(a, nk, kind) == want
        naive = float(np.sum((a * mi[:, :, None]).ravel(order="F")))
        if trial == 0:
            assert naive != math.fsum((a * mi[:, :, None]).ravel(order="F"))
    # the masks: a field of ones counts the wet interior columns, or the interior
    one = np.asfortranarray(np.ones((d.jpi, d.jpj, 1)))
    assert o.glob_sum(one, 1, 0) == mi.sum() and o.glob_sum(one, 1, 1) == (d.jpi - 2) * (d.jpj - 2)


def state(d, rng):
    s3 = (d.jpi, d.jpj, d.jpk)
    tm = d.tmask()
    ts = np.zeros(s3 + (2,), order="F")
    ts[..., 0] = rng.uniform(2, 20, s3) * tm
    ts[..., 1] = rng.uniform(34, 36, s3) * tm
    ssh = np.asfortranarray(rng.normal(0, 0.1, (d.jpi, d.jpj)) * tm[:, :, 0])
    return ts, ssh


def hsb(o, d, ts, ssh, ini, wn=None, emp=None, sbc=None, tsb=None):
    z2 = np.zeros((d.jpi, d.jpj), order="F")
    wn = np.zeros((d.jpi, d.jpj, d.jpk), order="F") if wn is None else wn
    emp = z2 if emp is None else emp
    sbc = np.zeros((d.jpi, d.jpj, 2), order="F") if sbc is None else sbc
    tsb = ts if tsb is None else tsb
    return o.dia_hsb_sums(ts, tsb, wn, ssh, emp, z2, None, sbc, ini)


def test_dia_hsb_closed_forms():
    d = configs.basin(30, 24, 12, seed=8, land_frac=0.25)
    o = Oracle(d)
    rng = np.random.default_rng(4)
    ts, ssh = state(d, rng)
    surf_tot, ini = o.dia_hsb_init(ts, ssh)
    mi, mh = masks(d)
    surf = d.arrays["e1e2t"] * mi
    assert surf_tot == math.fsum(surf.ravel(order="F"))
    e3t, tm = d.arrays["e3t_n"], d.tmask()
    vol = math.fsum((surf[:, :, None] * e3t * tm)[:, :, :-1].ravel(order="F"))
    heat = math.fsum((surf[:, :, None] * e3t * tm * ts[..., 0])[:, :, :-1].ravel(order="F"))
    s0 = hsb(o, d, ts, ssh, ini)
    assert s0[12] == vol
    # the initial state: every variation is the rounding of its terms
    assert s0[6] == 0.0 and s0[9] == 0.0
    assert abs(s0[10]) <= 1e-14 * heat and abs(s0[7]) <= 1e-14 * heat
    # sea level raised by c: volume c surf_tot, heat in the sea level c sum(T1 surf)
    c = 0.25
    s1 = hsb(o, d, ts, np.asfortranarray(ssh + c * mi), ini)
    assert math.isclose(s1[6], c * surf_tot, rel_tol=1e-13)
    assert math.isclose(s1[7], c * math.fsum((surf * ts[:, :, 0, 0]).ravel()), rel_tol=1e-12)
    # every temperature raised by dT: heat variation dT * volume, salt unchanged
    dT = 0.5
    ts2 = ts.copy(order="F")
    ts2[..., 0] += dT * tm
    s2 = hsb(o, d, ts2, ssh, ini)
    assert math.isclose(s2[10], dT * vol, rel_tol=1e-12)
    assert abs(s2[11]) <= 1e-14 * math.fsum(np.abs(surf[:, :, None] * e3t * ts[..., 1] * tm).ravel())
    # heat moved between two wet cells at constant e3t T: no change
    wet = np.argwhere((tm[:, :, :-1] > 0) & (mi[:, :, None] > 0))
    (i1, j1, k1), (i2, j2, k2) = wet[3], wet[-5]
    ts3 = ts.copy(order="F")
    q = 1.0
    ts3[i1, j1, k1, 0] += q / (surf[i1, j1] * e3t[i1, j1, k1])
    ts3[i2, j2, k2, 0] -= q / (surf[i2, j2] * e3t[i2, j2, k2])
    s3 = hsb(o, d, ts3, ssh, ini)
    assert abs(s3[10] - s0[10]) <= 1e-9
    # forcing: uniform fresh water and surface heat trend over the wet interior
    emp = np.asfortranarray(np.full((d.jpi, d.jpj), 2e-5))
    sbc = np.zeros((d.jpi, d.jpj, 2), order="F")
    sbc[..., 0] = 1e-4
    s4 = hsb(o, d, ts, ssh, ini, emp=emp, sbc=sbc)
    assert math.isclose(s4[0], -2e-5 * surf_tot, rel_tol=1e-13)
    assert math.isclose(s4[1], 1e-4 * surf_tot, rel_tol=1e-13) and s4[2] == 0.0
    # the surface advective flux w(1) T(1) of the linear free surface
    wn = np.zeros((d.jpi, d.jpj, d.jpk), order="F")
    wn[:, :, 0] = 1e-3
    s5 = hsb(o, d, ts, ssh, ini, wn=wn)
    assert math.isclose(s5[4], 1e-3 * math.fsum((surf * ts[:, :, 0, 0]).ravel()), rel_tol=1e-12)
