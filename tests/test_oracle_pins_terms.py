"""CPU pins of oracle terms that conservation laws alone leave open: the
rotated (iso-neutral) cross fluxes of tra_ldf_iso, the fourth-order
horizontal FCT flux, the enstrophy-conserving and mixed vorticity schemes,
the forcing, drag and time-filter terms of dyn_spg_ts, and the Zalesak
limiter next to land.  Each check is a closed form of the continuous or
discrete mathematics, chosen so that a flipped sign, a wrong coefficient or a
swapped metric fails it.
"""
import numpy as np
import pytest

from nemo_hip import abi, configs
from oracle.oracle import Oracle

from test_oracle_pins import fz, spg_inputs, uniform_domain


def box_domain(jpi, jpj, jpk, e1, e2, depth, jperio=0, opts=None, f0=0.0):
    """Flat box on uniform levels with e1 != e2 (land ring unless cyclic)."""
    e3 = np.full(jpk, depth / (jpk - 1))
    kt = configs._lbc_int(np.ones((jpi, jpj), np.int32), jperio)
    kb = configs._lbc_int(np.full((jpi, jpj), jpk - 1, np.int32), jperio)
    kb[kt == 0] = 0
    kt[kb == 0] = 0
    ff = np.full((jpi, jpj), f0)
    return configs._finish_domain(jpi, jpj, jpk, jperio, e1, e2, ff, ff, kt, kb, e3, e3, opts)


def periodic(a):
    """Doubly periodic ring copies (jperio = 7) of a 2D or 3D field."""
    a[0], a[-1] = a[-2], a[1]
    a[:, 0], a[:, -1] = a[:, -2], a[:, 1]
    return a


# ----------------------------------------------------------------------------
# tra_ldf_iso: rotated fluxes with non-zero slopes (traldf_iso.F90:247-366)
# ----------------------------------------------------------------------------
ISO_E1, ISO_E2, ISO_H, ISO_A = 2000.0, 1500.0, 450.0, 700.0


def iso_case(si, sj, tracer):
    jpi, jpj, jpk = 20, 18, 10
    d = box_domain(jpi, jpj, jpk, ISO_E1, ISO_E2, ISO_H)
    o = Oracle(d)
    _, u, v, w, _ = configs.masks(d)
    ahu, ahv = np.asfortranarray(ISO_A * u), np.asfortranarray(ISO_A * v)
    tb = np.asfortranarray(np.stack([tracer(d) * d.tmask()] * 2, axis=3))
    ta = fz(tb.shape)
    ah, akz = fz(u.shape), fz(u.shape)
    o.tra_ldf_iso(1, 1, "TRA", ahu, ahv, tb, tb, ta, 2, 1, np.asfortranarray(si * u), np.asfortranarray(sj * v),
                  np.asfortranarray(si * w), np.asfortranarray(sj * w), ah, akz)
    return d, tb[..., 0], ta[..., 0], ah, akz


def test_ldf_iso_cross_terms_closed_form():
    """T = x g(z) with g quadratic in z and uniform slopes (si, sj): the
    rotated operator of Redi (1982) in NEMO's small-slope form has the cross
    terms d/dx(-A si dT/dz) + d/dz(-A si dT/dx), i.e. -2 A si d2T/dxdz, which is
    -2 A si g'(z) here; the x-y term vanishes and dT/dy = 0 gives nothing in
    j.  With z positive up (level k below level k-1) the centred difference of
    a quadratic g is exact, so the interior trend equals the continuum value.
    At the top level (no flux through the surface) the horizontal part uses
    the one-sided g(1) - g(2) and the vertical part sees only the flux through
    w-level 2."""
    si, sj = 4e-3, -3e-3
    c0, c1, c2 = 0.7, -2e-3, 3e-6                         # g(z) = c0 + c1 z + c2 z^2, z < 0

    def gz(d):
        return c0 + c1 * d.zt + c2 * d.zt ** 2

    def tracer(d):
        d.zt = -(np.arange(d.jpk) + 0.5) * (ISO_H / (d.jpk - 1))
        x = (np.arange(d.jpi) * ISO_E1)[:, None, None]
        return 1e-5 * x * gz(d)[None, None, :] + 0.0 * np.zeros((d.jpi, d.jpj, d.jpk))

    d, tb, ta, _, _ = iso_case(si, sj, tracer)
    e3 = ISO_H / (d.jpk - 1)
    gp = c1 + 2.0 * c2 * d.zt                           # dg/dz at t-levels
    dTdx = 1e-5 * gz(d)                                 # dT/dx at t-levels
    inner = (slice(3, -3), slice(3, -3))
    # interior levels 2 .. jpk-2 (level jpk-1 is the bottom wet level)
    want = -2.0 * ISO_A * si * 1e-5 * gp[1:d.jpk - 2]
    got = ta[inner][:, :, 1:d.jpk - 2]
    assert np.allclose(got, want[None, None, :], rtol=1e-9, atol=0)
    # the top level: horizontal one-sided part plus the flux through w-level 2
    g = gz(d)
    top = -ISO_A * si * 1e-5 * (g[0] - g[1]) / e3 + ISO_A * si * 1e-5 * (g[0] + g[1]) / (2.0 * e3)
    assert np.allclose(ta[inner][:, :, 0], top, rtol=1e-9, atol=0)
    # same field rotated into j with e2 != e1: only si -> sj, x -> y change
    def tracer_y(d):
        d.zt = -(np.arange(d.jpk) + 0.5) * (ISO_H / (d.jpk - 1))
        y = (np.arange(d.jpj) * ISO_E2)[None, :, None]
        return 1e-5 * y * gz(d)[None, None, :] + np.zeros((d.jpi, d.jpj, d.jpk))
    d, tb, ta, _, _ = iso_case(si, sj, tracer_y)
    want = -2.0 * ISO_A * sj * 1e-5 * gp[1:d.jpk - 2]
    assert np.allclose(ta[inner][:, :, 1:d.jpk - 2], want[None, None, :], rtol=1e-9, atol=0)


def test_ldf_iso_annihilates_a_tracer_aligned_with_the_slopes():
    """A tracer constant along the iso-neutral surfaces, T = Gz (z + si x + sj y)
    (NEMO's slopes are minus dz/dx of the surfaces), has no rotated flux.  The
    operator is split: tra_ldf_iso does the explicit part and returns akz,
    which tra_zdf_imp adds to the vertical diffusivity.  Applied explicitly,
    the flux akz dT/dz through each w-level must cancel tra_ldf_iso's trend
    at every wet level of the interior columns."""
    si, sj, Gz = 5e-3, 3e-3, 2e-3

    def tracer(d):
        z = -(np.arange(d.jpk) + 0.5) * (ISO_H / (d.jpk - 1))
        x = (np.arange(d.jpi) * ISO_E1)[:, None, None]
        y = (np.arange(d.jpj) * ISO_E2)[None, :, None]
        return Gz * (z[None, None, :] + si * x + sj * y)

    d, tb, ta, ah, akz = iso_case(si, sj, tracer)
    assert np.allclose(ah[4:-4, 4:-4, 1:-1], ISO_A * (si * si + sj * sj), rtol=1e-12)
    e3 = ISO_H / (d.jpk - 1)
    w = configs.masks(d)[3]
    # explicit vertical flux akz (T(k-1) - T(k)) / e3w at w-levels 2..jpk-1
    fw = fz(tb.shape)
    fw[:, :, 1:] = akz[:, :, 1:] * (tb[:, :, :-1] - tb[:, :, 1:]) / e3 * w[:, :, 1:]
    zdf = fz(tb.shape)
    zdf[:, :, :-1] = (fw[:, :, :-1] - fw[:, :, 1:]) / e3
    inner = (slice(4, -4), slice(4, -4), slice(0, d.jpk - 1))
    total = (ta + zdf)[inner]
    scale = np.abs(ta[inner]).max()
    assert scale > 1e-9                                  # each part is far from zero at the surface
    assert np.abs(total).max() < 1e-11 * scale


# ----------------------------------------------------------------------------
# FCT order 4 horizontal (traadv_fct.F90:192-221)
# ----------------------------------------------------------------------------
def cell_mean_cubic(xc, dx, c):
    """Mean over [xc - dx/2, xc + dx/2] of p(x) = c0 + c1 x + c2 x^2 + c3 x^3."""
    P = lambda x: c[0] * x + c[1] * x ** 2 / 2 + c[2] * x ** 3 / 3 + c[3] * x ** 4 / 4
    return (P(xc + dx / 2) - P(xc - dx / 2)) / dx


def test_fct4_horizontal_flux_is_exact_for_cubic_cell_means():
    """The order-4 horizontal flux of tra_adv_fct uses the face value
    (T_i + T_i+1)/2 + (zltu_i - zltu_i+1)/2 with zltu the neighbour sum of
    differences times 1/6: the face value (7 (T_i + T_i+1) - (T_i-1 + T_i+2))/12,
    exact for the cell means of a cubic (the classical fourth-order
    finite-volume reconstruction).  With the limiter switched off (unit
    factors) the FCT trend is minus the divergence of that flux, so it equals
    -u (p(x_i+1/2) - p(x_i-1/2)) / dx exactly, in i and in j."""
    jpi, jpj, jpk = 24, 22, 3
    e1, e2, H = 1000.0, 800.0, 20.0
    d = box_domain(jpi, jpj, jpk, e1, e2, H)
    o = Oracle(d)
    t = d.tmask()
    _, um, vm, _, _ = configs.masks(d)
    e3 = H / (jpk - 1)
    cx = (0.3, 0.8, -0.5, 0.2)
    cy = (-0.1, 0.4, 0.3, -0.6)
    xs, ys = np.arange(jpi) / 10.0, np.arange(jpj) / 10.0
    px = cell_mean_cubic(xs, 0.1, cx)
    py = cell_mean_cubic(ys, 0.1, cy)
    ones = np.asfortranarray(np.ones((jpi, jpj, jpk)))
    for axis in (0, 1):
        tn = np.asfortranarray(np.broadcast_to((px[:, None, None] if axis == 0 else py[None, :, None]),
                                               (jpi, jpj, jpk)) * t)
        tb = np.asfortranarray(0.5 * tn)
        u0 = 0.3
        pun = np.asfortranarray(u0 * e2 * e3 * um) if axis == 0 else fz(t.shape)
        pvn = np.asfortranarray(u0 * e1 * e3 * vm) if axis == 1 else fz(t.shape)
        ta = fz(t.shape)
        o.fct_tl(600.0, pun, pvn, fz(t.shape), tb, tn, ones, ones, ones, ta, kn=(4, 2))
        P = lambda x, c: c[0] + c[1] * x + c[2] * x ** 2 + c[3] * x ** 3
        if axis == 0:
            want = -u0 * (P(xs + 0.05, cx) - P(xs - 0.05, cx)) / e1
            got = ta[4:-4, 3:-3, 0]
            assert np.allclose(got, want[4:-4, None], rtol=0, atol=1e-12 * np.abs(want).max())
        else:
            want = -u0 * (P(ys + 0.05, cy) - P(ys - 0.05, cy)) / e2
            got = ta[3:-3, 4:-4, 0]
            assert np.allclose(got, want[None, 4:-4], rtol=0, atol=1e-12 * np.abs(want).max())
        # the order-2 flux is not exact for a cubic: the test can tell them apart
        ta2 = fz(t.shape)
        o.fct_tl(600.0, pun, pvn, fz(t.shape), tb, tn, ones, ones, ones, ta2, kn=(2, 2))
        g2 = ta2[4:-4, 3:-3, 0] if axis == 0 else ta2[3:-3, 4:-4, 0]
        w2 = want[4:-4, None] if axis == 0 else want[None, 4:-4]
        assert np.abs(g2 - w2).max() > 1e-6 * np.abs(want).max()


# ----------------------------------------------------------------------------
# dyn_vor ENS and MIX (dynvor.F90:432-536, 169-171)
# ----------------------------------------------------------------------------
def vor_case(scheme, dyn, seed, flow):
    jpi, jpj, jpk = 26, 22, 3
    e1 = 1000.0
    d = uniform_domain(jpi, jpj, jpk, e1=e1, jperio=7, opts={"nvor_scheme": scheme, "n_dynadv": dyn})
    rng = np.random.default_rng(seed)
    d.arrays["ff_f"][...] = periodic(1e-4 * (1.0 + rng.standard_normal((jpi, jpj))))
    o = Oracle(d)
    _, u, v, _, _ = configs.masks(d)
    pot = periodic(rng.standard_normal((jpi, jpj, jpk)))
    un, vn = fz(pot.shape), fz(pot.shape)
    if flow == "nondivergent":      # streamfunction at f-points
        un[:, 1:] = -(pot[:, 1:] - pot[:, :-1]) / e1
        vn[1:, :] = (pot[1:, :] - pot[:-1, :]) / e1
    else:                           # velocity potential at t-points: no curl
        un[:-1, :] = (pot[1:, :] - pot[:-1, :]) / e1
        vn[:, :-1] = (pot[:, 1:] - pot[:, :-1]) / e1
    un, vn = periodic(un) * u, periodic(vn) * v
    ua, va = fz(un.shape), fz(un.shape)
    o.dyn_vor(1, un, vn, ua, va)
    return d, un, vn, periodic(ua), periodic(va)


def enstrophy_budget(d, un, vn, ua, va, q):
    """sum over f-points of q times the curl of the trend (summation over the
    unique points of the doubly periodic domain) and its absolute scale."""
    e1 = d.arrays["e1u"][0, 0]
    out = []
    for k in range(d.jpk - 1):
        dz = (va[1:, :-1, k] - va[:-1, :-1, k] - ua[:-1, 1:, k] + ua[:-1, :-1, k]) / e1
        qk = q[:, :, k] if q.ndim == 3 else q
        prod = (qk[:-1, :-1] * dz)[:-1, :-1]
        out.append((prod.sum(), np.abs(prod).sum()))
    return out


def curl_f(d, un, vn):
    e1 = d.arrays["e1u"][0, 0]
    c = np.zeros(un.shape)
    c[:-1, :-1] = (vn[1:, :-1] - vn[:-1, :-1] - un[:-1, 1:] + un[:-1, :-1]) / e1
    return c


@pytest.mark.parametrize("dyn", [abi.NP_LIN_DYN, abi.NP_VEC_C2])
def test_vor_ens_conserves_enstrophy(dyn):
    """Sadourny (1975): with a non-divergent flow the ENS vorticity term changes
    the relative vorticity without changing the sum of q^2 for ANY q at
    f-points (summation by parts turns the sum into q^2 times four t-point
    divergences).  Here q = f (linear dynamics, random f) or f + zeta (vector
    form).  ENE does not have this property, which the test also shows."""
    d, un, vn, ua, va = vor_case(abi.NP_ENS, dyn, 3, "nondivergent")
    f = d.arrays["ff_f"][..., None]
    q = np.broadcast_to(f + (curl_f(d, un, vn) if dyn == abi.NP_VEC_C2 else 0.0), un.shape)
    for s, sc in enstrophy_budget(d, un, vn, ua, va, q):
        assert sc > 0 and abs(s) < 1e-12 * sc
    d, un, vn, ua, va = vor_case(abi.NP_ENE, dyn, 3, "nondivergent")
    budgets = enstrophy_budget(d, un, vn, ua, va, q)
    assert max(abs(s) / sc for s, sc in budgets) > 1e-4


def test_vor_mix_is_ens_for_relative_and_ene_for_planetary_vorticity():
    """MIX (vector form) applies ENS to the relative vorticity and ENE to f.
    With f = 0 it must conserve enstrophy (the ENS identity); with a curl-free
    flow only the planetary part acts and it must do no work (the ENE
    identity); the swapped assignment fails both."""
    d, un, vn, ua, va = vor_case(abi.NP_MIX, abi.NP_VEC_C2, 5, "nondivergent")
    d.arrays["ff_f"][...] = 0.0
    o = Oracle(d)
    ua, va = fz(un.shape), fz(un.shape)
    o.dyn_vor(1, un, vn, ua, va)
    ua, va = periodic(ua), periodic(va)
    q = curl_f(d, un, vn)
    for s, sc in enstrophy_budget(d, un, vn, ua, va, q):
        assert sc > 0 and abs(s) < 1e-12 * sc
    d, un, vn, ua, va = vor_case(abi.NP_MIX, abi.NP_VEC_C2, 6, "irrotational")
    assert np.abs(curl_f(d, un, vn)[1:-2, 1:-2]).max() < 1e-15 * np.abs(un).max() / 1000.0
    sl = (slice(1, -1), slice(1, -1), slice(0, d.jpk - 1))
    work = (un * ua + vn * va)[sl].sum()
    scale = np.abs(un * ua + vn * va)[sl].sum()
    assert scale > 0 and abs(work) < 1e-12 * scale


# ----------------------------------------------------------------------------
# dyn_spg_ts: forcing, drag and time-filter terms (dynspg_ts.F90:358-639, 712-1269)
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("kt", [1, 3])
def test_spg_ts_uniform_channel_matches_scalar_recurrence(kt):
    """A cyclic E-W channel with walls north and south, f = 0, a flat bottom
    and fields uniform along the channel: every sub-step keeps the barotropic
    velocity U uniform (no divergence, no pressure gradient in x), so the
    generalised forward-backward loop reduces to a scalar recurrence:
        U <- U + dt (C U / H + F),   eta <- eta - dt emp / rho0,
    C = rCdU_bot (the sub-step's explicit bottom drag) and
    F = A0 + C (u_kb - U0) / H + tau / (rho0 H): the depth mean A0 of the 3D
    trend, the drag on the bottom baroclinic velocity and the wind.  The
    transport un_adv is the ts_wgt(2) mean of H times the AB3 extrapolation,
    and after the first step it is time-filtered with un_bf, ub2_b (atfp);
    ua becomes (ua_b - ub_b) / r2dt over the removed depth mean."""
    jpi, jpj, jpk = 14, 6, 6
    H, rdt, nn_baro = 50.0, 600.0, 30
    d = uniform_domain(jpi, jpj, jpk, e1=5000.0, depth=H, jperio=1, f0=0.0,
                       opts={"nvor_scheme": abi.NP_ENE, "nn_baro": nn_baro, "rdt": rdt})
    S = configs.make_state(d, seed=31)
    o = Oracle(d)
    F = spg_inputs(d, S, zero=True)
    _, u, v, _, _ = configs.masks(d)
    rau0, atfp = d.opts["rau0"], d.opts["rn_atfp"]
    U0, A0, dU, C, tau, emp, eta0 = 0.12, 3e-6, 0.02, -5e-3, 0.08, 2e-4, 0.05
    ub_b0, ub2, ubf = 0.09, 5.5, -0.7
    wet2 = u[:, :, 0]
    F["un_b"][...] = U0 * wet2
    F["ub_b"][...] = ub_b0 * wet2
    F["un"][...] = U0 * u
    F["un"][:, :, jpk - 2] += dU * wet2                  # bottom level mbku = jpk - 1
    F["ua"][...] = A0 * u
    F["rCdU_bot"][...] = C
    F["utau"][...] = tau
    F["emp"][...] = emp
    F["sshn"][...] = eta0 * d.tmask()[:, :, 0]
    F["ub2_b"][...] = ub2 * wet2
    F["un_bf"][...] = ubf * wet2
    for k in ("hu_n",):
        F[k][...] = H * wet2
    F["r1_hu_n"][...] = np.where(wet2 > 0, 1.0 / H, 0.0)
    o.dyn_spg_ts(kt, F)

    icycle, w1, w2 = o.ts_wgt(1, 1)
    dt = rdt / nn_baro
    Fc = A0 + C * dU / H + tau / (rau0 * H)
    Un, Ub, Ubb = U0, 0.0, 0.0
    eta = eta0
    ua_b = ssh = adv = 0.0
    for jn in range(1, icycle + 1):
        a1, a2, a3 = (1.0, 0.0, 0.0) if jn < 3 else (1.781105, -1.06221, 0.281105)
        ue = a1 * Un + a2 * Ub + a3 * Ubb
        adv += w2[jn - 1] * ue * H
        eta = eta - dt * emp / rau0
        Ua = Un + dt * (C * Un / H + Fc)
        ua_b += w1[jn - 1] * Ua
        ssh += w1[jn - 1] * eta
        Ubb, Ub, Un = Ub, Un, Ua
    if kt == 1:                                          # Euler start: no time filter, un_bf reset
        un_adv, un_bf, r2dt = adv, 0.0, rdt
    else:
        un_adv = 0.5 * (ub2 + adv - atfp * ubf)
        un_bf, r2dt = atfp * ubf + (adv - ub2), 2.0 * rdt
    inner = (slice(1, -1), slice(1, -1))
    assert abs(ua_b - U0) > 1e-4                          # the forcing did act
    assert np.allclose(F["ua_b"][inner], ua_b, rtol=1e-12, atol=0)
    assert np.allclose(F["ssha"][inner], ssh, rtol=1e-13, atol=0)
    assert np.allclose(F["un_adv"][inner], un_adv, rtol=1e-12, atol=0)
    assert np.allclose(F["un_bf"][inner], un_bf, rtol=1e-11, atol=1e-15)
    assert np.allclose(F["ub2_b"][inner], adv, rtol=1e-12, atol=0)
    assert np.abs(F["va_b"]).max() == 0.0 and np.abs(F["vn_adv"]).max() == 0.0
    wet = u[1:-1, 1:-1, :jpk - 1] > 0
    want_ua = (ua_b - ub_b0) / r2dt
    assert np.allclose(F["ua"][1:-1, 1:-1, :jpk - 1][wet], want_ua, rtol=1e-10, atol=1e-18)
    un_in = np.full(jpk - 1, U0)
    un_in[jpk - 2] += dU
    want_un = un_in + un_adv / H - U0
    assert np.allclose(F["un"][2:-2, 2:-2, :jpk - 1], want_un[None, None, :], rtol=1e-12, atol=0)


# ----------------------------------------------------------------------------
# nonosc next to land (traadv_fct.F90:354-426)
# ----------------------------------------------------------------------------
def zalesak_1d_walls(tb, tn, wet, u, dx, dt):
    """Textbook Zalesak (1979) FCT on a 1-D row of cells with land: fluxes only
    through faces between two wet cells, bounds from the wet cells of the
    3-point neighbourhood, limiter in the P/Q/R form."""
    n = len(tb)
    face = np.zeros(n)                                   # face i+1/2
    face[:-1] = wet[:-1] * wet[1:]
    uf = u * face
    f_low = np.where(uf > 0, uf * tb, uf * np.roll(tb, -1))
    f_low[-1] = 0.0
    im = lambda a: np.concatenate(([0.0], a[:-1]))
    t_td = (tb - dt / dx * (f_low - im(f_low))) * wet
    f_high = uf * 0.5 * (tn + np.roll(tn, -1))
    f_high[-1] = 0.0
    a = f_high - f_low
    big = 1e300
    hi = np.where(wet > 0, np.maximum(tb, t_td), -big)
    lo = np.where(wet > 0, np.minimum(tb, t_td), big)
    tmax = hi.copy()
    tmin = lo.copy()
    for sh in (1, -1):
        tmax[1:-1] = np.maximum(tmax[1:-1], np.roll(hi, sh)[1:-1])
        tmin[1:-1] = np.minimum(tmin[1:-1], np.roll(lo, sh)[1:-1])
    p_in = np.maximum(0, im(a)) - np.minimum(0, a)
    p_out = np.maximum(0, a) - np.minimum(0, im(a))
    with np.errstate(over="ignore", invalid="ignore"):
        r_in = np.where(wet > 0, np.minimum(1.0, (tmax - t_td) * dx / dt / (p_in + 1e-15)), 0.0)
        r_out = np.where(wet > 0, np.minimum(1.0, (t_td - tmin) * dx / dt / (p_out + 1e-15)), 0.0)
    ip = lambda x: np.concatenate((x[1:], [0.0]))
    c = np.where(a >= 0, np.minimum(ip(r_in), r_out), np.minimum(r_in, ip(r_out)))
    ac = c * a
    return -(f_low - im(f_low) + ac - im(ac)) / dx * wet


@pytest.mark.parametrize("u0", [0.45, -0.35])
def test_nonosc_next_to_land_matches_textbook_zalesak(u0):
    """A closed 1-D channel (one wet row, one wet level) cut by two land cells:
    the limiter's +-1e40 sentinels must keep land out of the bounds, and the
    faces that touch land carry no flux; the result is textbook Zalesak with
    walls, cell for cell."""
    n = 36
    jpi = n + 2
    kb = np.zeros((jpi, 3), np.int32)
    kb[1:-1, 1] = 1
    land = [12, 25]
    for i in land:
        kb[i, 1] = 0
    d = uniform_domain(jpi, 3, 2, e1=1.0, depth=1.0, k_bot=kb)
    o = Oracle(d)
    t = d.tmask()
    assert t[:, 1, 0].sum() == n - len(land)
    rng = np.random.default_rng(7)
    x = np.arange(jpi)
    prof = (np.where((x > 4) & (x < 10), 1.0, 0.0) + np.where((x > 15) & (x < 22), -0.6, 0.0)
            + 0.3 * np.sin(2 * np.pi * x / jpi) + 0.05 * rng.standard_normal(jpi))
    wet = t[:, 1, 0]
    tb3 = fz((jpi, 3, 2))
    tb3[:, 1, 0] = prof * wet
    tn3 = fz((jpi, 3, 2))
    tn3[:, 1, 0] = (0.7 * prof + 0.3 * np.roll(prof, 1)) * wet
    _, um, _, _, _ = configs.masks(d)
    pun = np.asfortranarray(u0 * um)
    z = fz(tb3.shape)
    tb = np.asfortranarray(np.stack([tb3, tb3], axis=3))
    tn = np.asfortranarray(np.stack([tn3, tn3], axis=3))
    ta = fz(tb.shape)
    dt = 1.0
    o.tra_adv_fct(1, 1, "TRA", dt, pun, z, z, tb, tn, ta, 2, 2, 2)
    ref = zalesak_1d_walls(tb3[:, 1, 0], tn3[:, 1, 0], wet, u0, 1.0, dt)
    assert np.abs(ref).max() > 0.1
    assert np.allclose(ta[1:-1, 1, 0, 0], ref[1:-1], rtol=0, atol=1e-13)


# ----------------------------------------------------------------------------
# vor_een next to land (dynvor.F90:580-603, 663-688)
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("e3f_opt", [0, 1])
def test_een_uniform_flow_along_a_coast(e3f_opt):
    """f-plane, flat box, uniform flow U across the basin (u = U on every wet
    u-point, 0 on the faces that touch the walls): at interior v-points the
    EEN Coriolis trend is -f U.  At a v-point next to an east or west wall two
    of its four u-faces are closed, and the two triads left see f/e3f at the
    wall f-point: with nn_een_e3f = 1 that is f/e3 (mean over the two wet
    t-points), giving -f U/2; with nn_een_e3f = 0 it is 2 f/e3 (sum of two
    wet thicknesses divided by 4), giving -2 f U/3."""
    f0, U = 1.2e-4, 0.25
    d = uniform_domain(18, 16, 4, f0=f0, opts={"nvor_scheme": abi.NP_EEN, "n_dynadv": abi.NP_LIN_DYN,
                                                "nn_een_e3f": e3f_opt})
    o = Oracle(d)
    _, u, v, _, _ = configs.masks(d)
    un = np.asfortranarray(U * u)
    vn = fz(u.shape)
    ua, va = fz(u.shape), fz(u.shape)
    o.dyn_vor(1, un, vn, ua, va)
    rows = slice(4, -4)
    assert np.allclose(va[4:-4, rows, :-1], -f0 * U, rtol=1e-13)
    coast = -f0 * U / 2.0 if e3f_opt == 1 else -2.0 * f0 * U / 3.0
    assert np.allclose(va[1, rows, :-1], coast, rtol=1e-13)              # west wall: v-points i = 2
    assert np.allclose(va[-2, rows, :-1], coast, rtol=1e-13)             # east wall: v-points i = jpi-1
    assert np.abs(ua[2:-2, rows, :-1]).max() < 1e-22
