"""CPU checks that pin the oracle's tangent linear and adjoint of the tracer
step (one tracer: FCT with the nonosc limiter factors frozen, orders 2
and 4 horizontally and vertically; the composed step uses 2/2 and 4/4,
tra_ldf_lap, tra_zdf_imp, leapfrog tra_nxt):

* with the limiter factors of the state itself, the frozen-limiter FCT is
  tra_adv_fct bit for bit, and so is the whole frozen step the nonlinear
  step; the factors lie in [0, 1] and the limiter is active somewhere;
* every adjoint is the transpose of its tangent linear: <L x, y> equals
  <x, L^T y> to rounding for random x and y (dot-product tests), closed and
  cyclic, routine by routine and for the composed step;
* the gradient of J = 1/2 sum (SST - target)^2 after several frozen steps
  matches central finite differences of the same (linear) model.
"""
import numpy as np
import pytest

from nemo_hip import configs
from oracle.oracle import Oracle

KT = 5                      # a leapfrog step (kt > nit000)


def setup(jperio, seed=3):
    d = configs.basin(26, 20, 9, seed=seed, jperio=jperio)
    S = configs.make_state(d, seed=seed + 10)
    o = Oracle(d)
    s3 = (d.jpi, d.jpj, d.jpk)
    z = [np.zeros(s3, order="F") for _ in range(3)]
    o.tra_adv_trp(S["un"], S["vn"], S["wn"], *z)
    F = {"zun": z[0], "zvn": z[1], "zwn": z[2], "ahtu": S["ahtu"], "ahtv": S["ahtv"], "avt": S["avt"],
         "p2dt": 2.0 * d.opts["rdt"]}
    return d, S, o, F


def t1(S, name):
    return np.asfortranarray(S[name][..., 0])


def coef(o, d, F, tb, tn, kn=(2, 2)):
    s3 = (d.jpi, d.jpj, d.jpk)
    c = [np.ones(s3, order="F") for _ in range(3)]
    o.fct_coef(F["p2dt"], F["zun"], F["zvn"], F["zwn"], tb, tn, *c, kn=kn)
    return c


def tl_step(o, d, F, c, tb, tn):
    """one frozen-limiter step of the tracer: (tb, tn) -> (tb', tn')"""
    s3 = (d.jpi, d.jpj, d.jpk)
    ta = np.zeros(s3, order="F")
    o.fct_tl(F["p2dt"], F["zun"], F["zvn"], F["zwn"], tb, tn, *c, ta)
    o.tra_ldf_lap(KT, 1, "TRA", F["ahtu"], F["ahtv"], tb, ta, 1, 1)
    o.tra_zdf_imp(KT, 1, "TRA", F["p2dt"], tb, ta, 1, F["avt"], F["avt"])
    B, N, A = (np.zeros(s3 + (2,), order="F") for _ in range(3))
    B[..., 0], N[..., 0], A[..., 0] = tb, tn, ta
    o.tra_nxt(KT, B, N, A)
    return np.asfortranarray(B[..., 0]), np.asfortranarray(N[..., 0])


def ad_step(o, d, F, c, lb, ln):
    """adjoint of tl_step: dJ/d(tb', tn') -> dJ/d(tb, tn)"""
    s3 = (d.jpi, d.jpj, d.jpk)
    lb, ln = lb.copy(order="F"), ln.copy(order="F")
    la = np.zeros(s3, order="F")
    o.tra_nxt_ad(lb, ln, la)
    o.zdf_imp_ad(F["p2dt"], F["avt"], la, lb)
    o.ldf_lap_ad(F["ahtu"], F["ahtv"], la, lb)
    o.fct_ad(F["zun"], F["zvn"], F["zwn"], *c, la, lb, ln)
    return lb, ln


def rnd(rng, d):
    return np.asfortranarray(rng.standard_normal((d.jpi, d.jpj, d.jpk)))


def interior(d, a):
    """a with the ring and level jpk cleared (the adjoint inputs of the
    trend: the forward writes interior cells only)"""
    b = a.copy(order="F")
    b[0], b[-1], b[:, 0], b[:, -1], b[:, :, -1] = 0, 0, 0, 0, 0
    return b


def close(lhs, rhs, tol=1e-12):
    return abs(lhs - rhs) <= tol * max(abs(lhs), abs(rhs), 1e-300)


KN = [(2, 2), (4, 4), (2, 4), (4, 2)]


@pytest.mark.parametrize("jperio", [0, 7])
@pytest.mark.parametrize("kn", KN, ids=[f"{a}{b}" for a, b in KN])
def test_frozen_fct_with_own_factors_is_tra_adv_fct(jperio, kn):
    d, S, o, F = setup(jperio)
    tb, tn = t1(S, "tsb"), t1(S, "tsn")
    c = coef(o, d, F, tb, tn, kn)
    t, u, v, w, f = configs.masks(d)
    inner = (slice(1, -1), slice(1, -1), slice(0, -1))
    wz = w.copy()
    wz[:, :, 0] = 0.0                      # cz(k) is the factor of the face above level k, k >= 2
    for a, m in zip(c, (u, v, wz)):       # wet faces (land faces carry no flux; their factors are not bounded)
        val = a[inner][m[inner] > 0]
        assert val.min() >= 0.0 and val.max() <= 1.0
    assert (c[0][inner][u[inner] > 0] < 1.0).any(), "the limiter should act somewhere"
    want = np.asfortranarray(S["tsa"][..., 0])
    o.tra_adv_fct(KT, 1, "TRA", F["p2dt"], F["zun"], F["zvn"], F["zwn"], tb, tn, want, 1, *kn)
    got = np.asfortranarray(S["tsa"][..., 0])
    o.fct_tl(F["p2dt"], F["zun"], F["zvn"], F["zwn"], tb, tn, *c, got, kn=kn)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("jperio", [0, 7])
@pytest.mark.parametrize("kn", KN, ids=[f"{a}{b}" for a, b in KN])
def test_fct_dot_product_all_orders(jperio, kn):
    d, S, o, F = setup(jperio)
    rng = np.random.default_rng(50 + jperio)
    c = coef(o, d, F, t1(S, "tsb"), t1(S, "tsn"), kn)
    s3 = (d.jpi, d.jpj, d.jpk)
    xb, xn, y = rnd(rng, d), rnd(rng, d), interior(d, rnd(rng, d))
    ta = np.zeros(s3, order="F")
    o.fct_tl(F["p2dt"], F["zun"], F["zvn"], F["zwn"], xb, xn, *c, ta, kn=kn)
    ab, an = np.zeros(s3, order="F"), np.zeros(s3, order="F")
    o.fct_ad(F["zun"], F["zvn"], F["zwn"], *c, y, ab, an, kn=kn)
    assert close(np.vdot(ta, y), np.vdot(xb, ab) + np.vdot(xn, an)), kn


@pytest.mark.parametrize("jperio", [0, 7])
def test_dot_products(jperio):
    d, S, o, F = setup(jperio)
    rng = np.random.default_rng(40 + jperio)
    c = coef(o, d, F, t1(S, "tsb"), t1(S, "tsn"))
    s3 = (d.jpi, d.jpj, d.jpk)
    # FCT: (tb, tn) -> trend
    xb, xn, y = rnd(rng, d), rnd(rng, d), interior(d, rnd(rng, d))
    ta = np.zeros(s3, order="F")
    o.fct_tl(F["p2dt"], F["zun"], F["zvn"], F["zwn"], xb, xn, *c, ta)
    ab, an = np.zeros(s3, order="F"), np.zeros(s3, order="F")
    o.fct_ad(F["zun"], F["zvn"], F["zwn"], *c, y, ab, an)
    assert close(np.vdot(ta, y), np.vdot(xb, ab) + np.vdot(xn, an)), "fct"
    # laplacian: tb -> trend
    ta = np.zeros(s3, order="F")
    o.tra_ldf_lap(KT, 1, "TRA", F["ahtu"], F["ahtv"], xb, ta, 1, 1)
    ab = np.zeros(s3, order="F")
    o.ldf_lap_ad(F["ahtu"], F["ahtv"], y, ab)
    assert close(np.vdot(ta, y), np.vdot(xb, ab)), "ldf_lap"
    # implicit vertical diffusion: (tb, trend) -> after field
    xa = rnd(rng, d)
    ta = xa.copy(order="F")
    o.tra_zdf_imp(KT, 1, "TRA", F["p2dt"], xb, ta, 1, F["avt"], F["avt"])
    y2 = rnd(rng, d)
    la, ab = y2.copy(order="F"), np.zeros(s3, order="F")
    o.zdf_imp_ad(F["p2dt"], F["avt"], la, ab)
    assert close(np.vdot(ta, y2), np.vdot(xb, ab) + np.vdot(xa, la)), "zdf_imp"
    # leapfrog filter: (tb, tn, ta) -> (tb', tn')
    B, N, A = (np.zeros(s3 + (2,), order="F") for _ in range(3))
    B[..., 0], N[..., 0], A[..., 0] = xb, xn, xa
    o.tra_nxt(KT, B, N, A)
    yb, yn = rnd(rng, d), rnd(rng, d)
    lb, ln, la = yb.copy(order="F"), yn.copy(order="F"), np.zeros(s3, order="F")
    o.tra_nxt_ad(lb, ln, la)
    assert close(np.vdot(B[..., 0], yb) + np.vdot(N[..., 0], yn),
                 np.vdot(xb, lb) + np.vdot(xn, ln) + np.vdot(xa, la)), "tra_nxt"
    # the composed step
    b1, n1 = tl_step(o, d, F, c, xb, xn)
    lb, ln = ad_step(o, d, F, c, yb, yn)
    assert close(np.vdot(b1, yb) + np.vdot(n1, yn), np.vdot(xb, lb) + np.vdot(xn, ln)), "step"


@pytest.mark.parametrize("jperio", [0, 7])
def test_frozen_step_with_own_factors_is_the_step(jperio):
    d, S, o, F = setup(jperio)
    tb, tn = t1(S, "tsb"), t1(S, "tsn")
    c = coef(o, d, F, tb, tn)
    b1, n1 = tl_step(o, d, F, c, tb, tn)
    s3 = (d.jpi, d.jpj, d.jpk)
    ta = np.zeros(s3, order="F")
    o.tra_adv_fct(KT, 1, "TRA", F["p2dt"], F["zun"], F["zvn"], F["zwn"], tb, tn, ta, 1, 2, 2)
    o.tra_ldf_lap(KT, 1, "TRA", F["ahtu"], F["ahtv"], tb, ta, 1, 1)
    o.tra_zdf_imp(KT, 1, "TRA", F["p2dt"], tb, ta, 1, F["avt"], F["avt"])
    B, N, A = (np.zeros(s3 + (2,), order="F") for _ in range(3))
    B[..., 0], N[..., 0], A[..., 0] = tb, tn, ta
    o.tra_nxt(KT, B, N, A)
    assert np.array_equal(b1, B[..., 0]) and np.array_equal(n1, N[..., 0])


def test_gradient_matches_finite_differences():
    """J = 1/2 sum over interior wet surface points of (T_n - target)^2 after
    four frozen steps from tb = tn = T0 (ring filled by lbc); its gradient
    with respect to the interior of T0 against central differences."""
    d, S, o, F = setup(7)
    rng = np.random.default_rng(9)
    t, *_ = configs.masks(d)
    T0 = t1(S, "tsn")
    o.lbc_lnk(T0, "T")                 # a state's ring is the lbc fill of its interior
    s3 = (d.jpi, d.jpj, d.jpk)
    cs = []
    b, n = T0.copy(order="F"), T0.copy(order="F")
    for _ in range(4):
        c = coef(o, d, F, b, n)
        cs.append(c)
        b, n = tl_step(o, d, F, c, b, n)
    target = n[:, :, 0] + 0.1 * rng.standard_normal(n.shape[:2])
    w = np.zeros(s3)
    w[1:-1, 1:-1, 0] = t[1:-1, 1:-1, 0]

    def J(T):
        b, n = T.copy(order="F"), T.copy(order="F")
        for c in cs:
            b, n = tl_step(o, d, F, c, b, n)
        return 0.5 * float(np.sum(w[:, :, 0] * (n[:, :, 0] - target) ** 2)), n

    J0, nN = J(T0)
    lb = np.zeros(s3, order="F")
    ln = np.asfortranarray(w * np.where(np.arange(d.jpk)[None, None, :] == 0, nN - target[:, :, None], 0.0))
    for c in reversed(cs):
        lb, ln = ad_step(o, d, F, c, lb, ln)
    g = np.asfortranarray(lb + ln)
    o.lbc_adj(g, d.jpk, "T")

    def with_ring(x):
        y = x.copy(order="F")
        o.lbc_lnk(y, "T")
        return y

    for q in range(4):
        dirn = interior(d, rnd(rng, d))
        eps = 1e-3
        jp, _ = J(with_ring(T0 + eps * dirn))
        jm, _ = J(with_ring(T0 - eps * dirn))
        fd = (jp - jm) / (2 * eps)
        ad = float(np.vdot(g, dirn))
        assert abs(fd - ad) <= 1e-9 * max(abs(fd), 1e-12), (q, fd, ad)


def test_nonlinear_step_is_not_differentiable_at_finite_difference_scales():
    """Why the gradient is checked against the frozen-limiter model: central
    differences of the nonlinear FCT step (the limiter factors recomputed
    from each perturbed state) do not settle as the step shrinks from 1e-2
    to 1e-5 -- nonosc's min/max branches switch within every such step, so
    no linearisation, frozen or not, can match them to 1e-6 -- while those
    of the frozen model agree to rounding at every step."""
    d, S, o, F = setup(0)
    t, *_ = configs.masks(d)
    kn = (4, 4)
    T0 = t1(S, "tsn")
    o.lbc_lnk(T0, "T")
    s3 = (d.jpi, d.jpj, d.jpk)
    cs = []
    b, n = T0.copy(order="F"), T0.copy(order="F")
    for _ in range(4):
        cs.append(coef(o, d, F, b, n, kn=kn))
        b, n = tl_step(o, d, F, cs[-1], b, n)
    rng = np.random.default_rng(3)
    target = n + 0.1 * rng.standard_normal(s3)
    w = np.zeros(s3)
    w[1:-1, 1:-1, 0] = t[1:-1, 1:-1, 0]

    def J(T, frozen):
        b, n = T.copy(order="F"), T.copy(order="F")
        for c in cs:
            b, n = tl_step(o, d, F, c if frozen else coef(o, d, F, b, n, kn=kn), b, n)
        return 0.5 * float(np.sum(w * (n - target) ** 2))

    spread = {True: [], False: []}
    for q in range(4):
        dirn = interior(d, np.random.default_rng(100 + q).standard_normal(s3))
        o.lbc_lnk(dirn, "T")
        for frozen in (True, False):
            fd = [(J(T0 + e * dirn, frozen) - J(T0 - e * dirn, frozen)) / (2 * e) for e in (1e-2, 1e-3, 1e-4, 1e-5)]
            spread[frozen].append((max(fd) - min(fd)) / abs(fd[-1]))
    assert max(spread[True]) <= 1e-8, spread[True]
    assert min(spread[False]) >= 1e-2, spread[False]
