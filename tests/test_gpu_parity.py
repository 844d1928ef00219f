"""Parity of the gfx950 kernels with the CPU oracle on identical inputs.

FP64 form: the bar is <= 1e-12 max relative error; every comparison also
records whether the result is bit-identical (it is expected to be: same
operation order, no FMA contraction on either side).  Masks and k-indices are
compared exactly.  FP32-storage form: relative error against the FP64 oracle
within the tolerance written in each test.

Cases: the 32x22x31 GYRE-like closed basin and small seeded basins with land
(closed, cyclic E-W, cyclic E-W + N-S) so every lbc_lnk branch is exercised,
and a 75-level cyclic basin (44x36x75), the depth of the full-size
configurations, for the column arrays sized by the level count.
"""
import numpy as np
import pytest

from nemo_hip import abi, configs
from oracle.oracle import Oracle

torch = pytest.importorskip("torch")

CASES = {
    "gyre": lambda opts=None: configs.gyre_like(opts=opts),
    "basin": lambda opts=None: configs.basin(40, 30, 12, seed=3, opts=opts),
    "cyc7": lambda opts=None: configs.basin(40, 30, 12, seed=4, jperio=7, opts=opts),
    "cyc1": lambda opts=None: configs.basin(36, 28, 10, seed=5, jperio=1, opts=opts),
    "deep75": lambda opts=None: configs.basin(44, 36, 75, seed=6, jperio=7, opts=opts),
}
CASE_IDS = list(CASES)


def dev(a, dt=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).ravel(order="F"))).to("cuda", dt)


def host(t, shape):
    return t.detach().cpu().numpy().astype(np.float64).reshape(shape, order="F")


def maxrel(got, want):
    scale = max(float(np.abs(want).max()), 1e-300)
    return float(np.abs(got - want).max()) / scale


def check(got, want, tol=1e-12, what=""):
    assert got.shape == want.shape, what
    assert np.all(np.isfinite(got)), f"{what}: non-finite values"
    if np.array_equal(got, want):
        return 0.0
    err = maxrel(got, want)
    assert err <= tol, f"{what}: max relative error {err:.3e} > {tol:.1e}"
    return err


class Env:
    def __init__(self, hip, case, opts=None, seed=7, **state_kw):
        self.d = CASES[case](opts)
        self.S = configs.make_state(self.d, seed=seed, **state_kw)
        self.o = Oracle(self.d)
        hip.register(self.d)
        self.h = hip
        self.s3 = (self.d.jpi, self.d.jpj, self.d.jpk)
        self.s4 = self.s3 + (2,)
        self.s2 = (self.d.jpi, self.d.jpj)

    def c(self, name):
        return self.S[name].copy(order="F")


def sync(h):
    torch.cuda.synchronize()
    h.sync()


# ----------------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("case", CASE_IDS)
def test_dom_msk_exact(hip, case):
    e = Env(hip, case)
    want = e.o.dom_msk()
    got = {}
    for k, v in want.items():
        got[k] = torch.zeros(v.size, dtype=torch.int32 if v.dtype == np.int32 else torch.float64, device="cuda")
    hip.dom_msk(**got)
    sync(hip)
    for k, v in want.items():
        g = got[k].cpu().numpy().reshape(v.shape, order="F")
        assert np.array_equal(g, v), f"{case}: {k} differs at {np.argwhere(g != v)[:5]}"


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["basin", "cyc7", "cyc1"])
@pytest.mark.parametrize("nat", list("TUVWF"))
def test_lbc_lnk_exact(hip, case, nat):
    e = Env(hip, case)
    a = np.asfortranarray(np.random.default_rng(1).standard_normal(e.s3))
    t = dev(a)
    hip.lbc_lnk(t, e.d.jpk, nat, 1.0, 0.0)
    sync(hip)
    e.o.lbc_lnk(a, nat)
    assert np.array_equal(host(t, e.s3), a)


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASE_IDS)
def test_tra_adv_trp(hip, case):
    e = Env(hip, case)
    z = [np.zeros(e.s3, order="F") for _ in range(3)]
    e.o.tra_adv_trp(e.S["un"], e.S["vn"], e.S["wn"], *z)
    g = [torch.zeros(int(np.prod(e.s3)), dtype=torch.float64, device="cuda") for _ in range(3)]
    hip.tra_adv_trp(dev(e.S["un"]), dev(e.S["vn"]), dev(e.S["wn"]), *g)
    sync(hip)
    for a, b, n in zip(g, z, "uvw"):
        check(host(a, e.s3), b, what=f"z{n}n")


def transports(e):
    z = [np.zeros(e.s3, order="F") for _ in range(3)]
    e.o.tra_adv_trp(e.S["un"], e.S["vn"], e.S["wn"], *z)
    return z


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASE_IDS)
@pytest.mark.parametrize("orders", [(2, 2), (4, 4), (2, 4), (4, 2)])
def test_tra_adv_fct(hip, case, orders):
    e = Env(hip, case)
    zun, zvn, zwn = transports(e)
    p2dt = 2.0 * e.d.opts["rdt"]
    want = e.c("tsa")
    e.o.tra_adv_fct(3, 1, "TRA", p2dt, zun, zvn, zwn, e.S["tsb"], e.S["tsn"], want, 2, *orders)
    ta = dev(e.S["tsa"])
    hip.tra_adv_fct(3, 1, "TRA", p2dt, dev(zun), dev(zvn), dev(zwn), dev(e.S["tsb"]), dev(e.S["tsn"]), ta, 2, *orders)
    sync(hip)
    check(host(ta, e.s4), want, what=f"tra_adv_fct {case} {orders}")


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["gyre", "basin"])
def test_tra_adv_fct_host_pointer_mode(hip, case):
    """The drop-in (host array) path: staging through HBM gives the same bits."""
    import nemo_hip
    e = Env(hip, case)
    zun, zvn, zwn = transports(e)
    want = e.c("tsa")
    e.o.tra_adv_fct(3, 1, "TRA", 3600.0, zun, zvn, zwn, e.S["tsb"], e.S["tsn"], want, 2, 4, 4)
    got = e.c("tsa")
    hip.set_pointer_mode(nemo_hip.NEMO_HIP_PTR_HOST)
    try:
        hip.tra_adv_fct(3, 1, "TRA", 3600.0, zun, zvn, zwn, e.S["tsb"], e.S["tsn"], got, 2, 4, 4)
    finally:
        hip.set_pointer_mode(nemo_hip.NEMO_HIP_PTR_DEVICE)
    check(got, want, what="host-pointer tra_adv_fct")


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASE_IDS)
def test_interp_4th_cpt(hip, case):
    e = Env(hip, case)
    tin = np.asfortranarray(e.S["tsn"][:, :, :, 0])
    want = np.zeros(e.s3, order="F")
    e.o.interp_4th_cpt(tin, want)
    out = torch.zeros(int(np.prod(e.s3)), dtype=torch.float64, device="cuda")
    hip.interp_4th_cpt(dev(tin), out)
    sync(hip)
    got = host(out, e.s3)
    sl = (slice(1, -1), slice(1, -1), slice(1, -1))      # interior columns, levels 2..jpkm1
    check(got[sl], want[sl], what="interp_4th_cpt")


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASE_IDS)
@pytest.mark.parametrize("kpass", [1, 2])
def test_tra_ldf_lap(hip, case, kpass):
    e = Env(hip, case)
    want = e.c("tsa")
    e.o.tra_ldf_lap(1, 1, "TRA", e.S["ahtu"], e.S["ahtv"], e.S["tsb"], want, 2, kpass)
    ta = dev(e.S["tsa"])
    hip.tra_ldf_lap(1, 1, "TRA", dev(e.S["ahtu"]), dev(e.S["ahtv"]), dev(e.S["tsb"]), ta, 2, kpass)
    sync(hip)
    check(host(ta, e.s4), want, what="tra_ldf_lap")


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASE_IDS)
@pytest.mark.parametrize("msc", [0, 1])
@pytest.mark.parametrize("kt", [1, 5])
def test_tra_ldf_iso(hip, case, msc, kt):
    e = Env(hip, case, opts={"ln_traldf_msc": msc})
    S = e.S
    ah_o, akz_o = np.zeros(e.s3, order="F"), np.zeros(e.s3, order="F")
    want = e.c("tsa")
    e.o.tra_ldf_iso(kt, 1, "TRA", S["ahtu"], S["ahtv"], S["tsb"], S["tsb"], want, 2, 1, S["uslp"], S["vslp"],
                    S["wslpi"], S["wslpj"], ah_o, akz_o)
    ta = dev(S["tsa"])
    ah, akz = dev(np.zeros(e.s3)), dev(np.zeros(e.s3))
    hip.tra_ldf_iso(kt, 1, "TRA", dev(S["ahtu"]), dev(S["ahtv"]), dev(S["tsb"]), dev(S["tsb"]), ta, 2, 1,
                    dev(S["uslp"]), dev(S["vslp"]), dev(S["wslpi"]), dev(S["wslpj"]), ah, akz)
    sync(hip)
    check(host(ta, e.s4), want, what="tra_ldf_iso pta")
    check(host(ah, e.s3), ah_o, what="ah_wslp2")
    check(host(akz, e.s3), akz_o, what="akz")


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASE_IDS)
@pytest.mark.parametrize("variant", ["tra", "tra_slp", "trc", "ddm"])
def test_tra_zdf_imp(hip, case, variant):
    opts = {"ln_zdfddm": 1} if variant == "ddm" else None
    e = Env(hip, case, opts=opts)
    S = e.S
    cd = "TRC" if variant == "trc" else "TRA"
    slp = 1 if variant == "tra_slp" else 0
    ah = np.asfortranarray(1e-3 * np.abs(S["wslpi"]) * 1e3)
    avs = np.asfortranarray(S["avt"] * 0.5)
    want = e.c("tsa")
    e.o.tra_zdf_imp(4, 1, cd, 3600.0, S["tsb"], want, 2, S["avt"], avs, slp, ah, ah)
    ta = dev(S["tsa"])
    hip.tra_zdf_imp(4, 1, cd, 3600.0, dev(S["tsb"]), ta, 2, dev(S["avt"]), dev(avs), slp, dev(ah), dev(ah))
    sync(hip)
    check(host(ta, e.s4), want, what=f"tra_zdf_imp {variant}")


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASE_IDS)
def test_tra_zdf_with_salinity_reset(hip, case):
    e = Env(hip, case)
    S = e.S
    tsa = e.c("tsa")
    tsa[:, :, :, 1] -= 1.0           # trend that drives S negative somewhere, incl. the halo
    want = tsa.copy(order="F")
    e.o.tra_zdf(1, S["tsb"], want, S["avt"], S["avs"])
    ta = dev(tsa)
    hip.tra_zdf(1, dev(S["tsb"]), ta, dev(S["avt"]), dev(S["avs"]))
    sync(hip)
    got = host(ta, e.s4)
    check(got, want, what="tra_zdf")
    assert (want[:, :, :, 1] == 0.1).any()


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASE_IDS)
@pytest.mark.parametrize("form", ["keg0", "keg1", "zad", "vec0", "vec1", "cen2"])
def test_dyn_adv(hip, case, form):
    e = Env(hip, case)
    S = e.S
    ua_o, va_o = e.c("ua"), e.c("va")
    ua, va = dev(S["ua"]), dev(S["va"])
    un, vn, wn = dev(S["un"]), dev(S["vn"]), dev(S["wn"])
    if form.startswith("keg"):
        k = int(form[-1])
        e.o.dyn_keg(1, k, S["un"], S["vn"], ua_o, va_o)
        hip.dyn_keg(1, k, un, vn, ua, va)
    elif form == "zad":
        e.o.dyn_zad(1, S["un"], S["vn"], S["wn"], ua_o, va_o)
        hip.dyn_zad(1, un, vn, wn, ua, va)
    elif form.startswith("vec"):
        k = int(form[-1])
        e.o.dyn_keg(1, k, S["un"], S["vn"], ua_o, va_o)
        e.o.dyn_zad(1, S["un"], S["vn"], S["wn"], ua_o, va_o)
        e.d.opts["nn_dynkeg"] = k
        hip.register(e.d)
        hip.dyn_adv(1, un, vn, wn, ua, va)
    else:
        e.o.dyn_adv_cen2(1, S["un"], S["vn"], S["wn"], ua_o, va_o)
        hip.dyn_adv_cen2(1, un, vn, wn, ua, va)
    sync(hip)
    check(host(ua, e.s3), ua_o, what=f"{form} ua")
    check(host(va, e.s3), va_o, what=f"{form} va")


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASE_IDS)
@pytest.mark.parametrize("scheme", ["ene", "ens", "mix", "een", "eet", "ent"])
@pytest.mark.parametrize("adv", [abi.NP_VEC_C2, abi.NP_FLX_C2, abi.NP_LIN_DYN])
def test_dyn_vor(hip, case, scheme, adv):
    e = Env(hip, case, opts={"nvor_scheme": abi.VOR_NAMES[scheme], "n_dynadv": adv, "nn_een_e3f": 1 if case != "basin" else 0})
    S = e.S
    ua_o, va_o = e.c("ua"), e.c("va")
    e.o.dyn_vor(1, S["un"], S["vn"], ua_o, va_o)
    ua, va = dev(S["ua"]), dev(S["va"])
    hip.dyn_vor(1, dev(S["un"]), dev(S["vn"]), ua, va)
    sync(hip)
    check(host(ua, e.s3), ua_o, what=f"vor {scheme} ua")
    check(host(va, e.s3), va_o, what=f"vor {scheme} va")


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASE_IDS)
@pytest.mark.parametrize("kt", [1, 2, 7])
def test_dyn_zdf(hip, case, kt):
    e = Env(hip, case)
    S = e.S
    ua_o, va_o = e.c("ua"), e.c("va")
    args = [S[n] for n in ("avm", "rCdU_bot", "ua_b", "va_b", "utau", "utau_b", "vtau", "vtau_b")]
    e.o.dyn_zdf(kt, S["ub"], S["vb"], ua_o, va_o, *args)
    ua, va = dev(S["ua"]), dev(S["va"])
    hip.dyn_zdf(kt, dev(S["ub"]), dev(S["vb"]), ua, va, *[dev(a) for a in args])
    sync(hip)
    check(host(ua, e.s3), ua_o, what="dyn_zdf ua")
    check(host(va, e.s3), va_o, what="dyn_zdf va")


@pytest.mark.gpu
@pytest.mark.parametrize("nk", [3, 4, 5, 6])
def test_column_solves_shallow(hip, nk):
    """The column solves on shallow basins: the back substitution runs by
    4-level segments from checkpoint pivots, so 1 to 4 solved levels leave
    no checkpoint and partial segments; FP64 bit-exact against the oracle."""
    d = configs.basin(30, 24, nk, seed=9)
    S = configs.make_state(d, seed=11)
    o = Oracle(d)
    hip.register(d)
    s3, s4 = (d.jpi, d.jpj, d.jpk), (d.jpi, d.jpj, d.jpk, 2)
    c = lambda n: np.array(S[n], order="F", copy=True)
    ua_o, va_o = c("ua"), c("va")
    args = [S[n] for n in ("avm", "rCdU_bot", "ua_b", "va_b", "utau", "utau_b", "vtau", "vtau_b")]
    o.dyn_zdf(2, S["ub"], S["vb"], ua_o, va_o, *args)
    ua, va = dev(S["ua"]), dev(S["va"])
    hip.dyn_zdf(2, dev(S["ub"]), dev(S["vb"]), ua, va, *[dev(a) for a in args])
    sync(hip)
    check(host(ua, s3), ua_o, what=f"dyn_zdf ua, {nk} levels")
    check(host(va, s3), va_o, what=f"dyn_zdf va, {nk} levels")
    want = c("tsa")
    avs = np.asfortranarray(S["avt"] * 0.5)
    ah = np.asfortranarray(np.abs(S["wslpi"]))
    o.tra_zdf_imp(4, 1, "TRA", 3600.0, S["tsb"], want, 2, S["avt"], avs, 1, ah, ah)
    ta = dev(S["tsa"])
    hip.tra_zdf_imp(4, 1, "TRA", 3600.0, dev(S["tsb"]), ta, 2, dev(S["avt"]), dev(avs), 1, dev(ah), dev(ah))
    sync(hip)
    check(host(ta, s4), want, what=f"tra_zdf_imp, {nk} levels")


SPG_2D = ["un_b", "vn_b", "ub_b", "vb_b", "sshn", "hu_n", "hv_n", "r1_hu_n", "r1_hv_n", "ht_n", "rCdU_bot", "utau",
          "vtau", "emp", "rnf", "ub2_b", "vb2_b", "un_bf", "vn_bf"]
SPG_OUT = ["ua_b", "va_b", "ssha", "un_adv", "vn_adv", "ub2_b", "vb2_b", "un_bf", "vn_bf"]


def spg_fields(e):
    F = {k: e.c(k) for k in SPG_2D}
    for k in ("ua", "va", "un", "vn"):
        F[k] = e.c(k)
    for k in ("ua_b", "va_b", "ssha", "un_adv", "vn_adv"):
        F[k] = np.zeros(e.s2, order="F")
    return F


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASE_IDS)
@pytest.mark.parametrize("scheme", ["ene", "een", "ens", "eet", "ent"])
def test_dyn_spg_ts(hip, case, scheme):
    e = Env(hip, case, opts={"nvor_scheme": abi.VOR_NAMES[scheme]})
    Fo = spg_fields(e)
    Fd = {k: dev(v) for k, v in Fo.items()}
    shapes = {k: (e.s3 if k in ("ua", "va", "un", "vn") else e.s2) for k in Fo}
    for kt in (1, 2, 3):                 # Euler start, then leapfrog with the time-filter state
        e.o.dyn_spg_ts(kt, Fo)
        hip.dyn_spg_ts(kt, Fd)
        sync(hip)
        for k in ["ua", "va", "un", "vn"] + SPG_OUT:
            check(host(Fd[k], shapes[k]), Fo[k], what=f"spg_ts {scheme} kt={kt} {k}")


@pytest.mark.gpu
def test_dyn_spg_ts_no_time_averaging(hip):
    """ln_bt_av = F: the barotropic *_e levels carry over between calls."""
    e = Env(hip, "basin", opts={"ln_bt_av": 0})
    Fo = spg_fields(e)
    Fd = {k: dev(v) for k, v in Fo.items()}
    shapes = {k: (e.s3 if k in ("ua", "va", "un", "vn") else e.s2) for k in Fo}
    for kt in (1, 2, 3):
        e.o.dyn_spg_ts(kt, Fo)
        hip.dyn_spg_ts(kt, Fd)
        sync(hip)
        for k in ["ua", "un"] + SPG_OUT:
            check(host(Fd[k], shapes[k]), Fo[k], what=f"spg_ts no-av kt={kt} {k}")


@pytest.mark.gpu
def test_ts_wgt(hip):
    e = Env(hip, "gyre")
    for av in (0, 1):
        for fw in (0, 1):
            a = hip.ts_wgt(av, fw)
            b = e.o.ts_wgt(av, fw)
            assert a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
