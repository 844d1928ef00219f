"""Fused steps (nemo_hip_step_dyn / nemo_hip_step_tra) against the sequence of
separate routine calls they replace.

FP64: against the oracle run as that sequence, <= 1e-12 max relative error
(bit-identical expected: the fused kernels round every stored intermediate
where the separate routines store it).  FP32 storage: against the separate
FP32 entry points, bit-identical, for the same reason.  The fused outputs start
from NaN-filled arrays, so every element they must write is checked to be
written.
"""
import numpy as np
import pytest

from nemo_hip import abi
from test_gpu_parity import CASE_IDS, SPG_OUT, Env, check, dev, host, spg_fields, sync

torch = pytest.importorskip("torch")


def nan_like(t):
    return torch.full_like(t, float("nan"))


def r2dt(d, kt):
    o = d.opts
    return o["rdt"] if (o["neuler"] == 0 and kt == o["nit000"]) else 2.0 * o["rdt"]


# ----------------------------------------------------------------------------
# momentum
# ----------------------------------------------------------------------------
def oracle_dyn_step(e, kt, F):
    S, o = e.S, e.o
    ua, va = np.zeros(e.s3, order="F"), np.zeros(e.s3, order="F")
    if e.d.opts["n_dynadv"] != abi.NP_LIN_DYN:
        o.dyn_adv(kt, F["un"], F["vn"], S["wn"], ua, va)
    o.dyn_vor(kt, F["un"], F["vn"], ua, va)
    F["ua"], F["va"] = ua, va
    o.dyn_spg_ts(kt, F)
    o.dyn_zdf(kt, S["ub"], S["vb"], ua, va, S["avm"], S["rCdU_bot"], F["ua_b"], F["va_b"], S["utau"], S["utau_b"],
              S["vtau"], S["vtau_b"])


DYN_OPTS = {
    "ene_vec": {},
    "een_vec": {"nvor_scheme": abi.NP_EEN},
    "ens_flx": {"nvor_scheme": abi.NP_ENS, "n_dynadv": abi.NP_FLX_C2},
    "ent_vec1": {"nvor_scheme": abi.NP_ENT, "nn_dynkeg": 1},
    "mix_lin": {"nvor_scheme": abi.NP_MIX, "n_dynadv": abi.NP_LIN_DYN},
}


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASE_IDS)
@pytest.mark.parametrize("opt", list(DYN_OPTS))
def test_step_dyn_fp64(hip, case, opt):
    e = Env(hip, case, opts=DYN_OPTS[opt])
    Fo = spg_fields(e)
    Fd = {k: dev(v) for k, v in Fo.items()}
    Fd["ua"], Fd["va"] = nan_like(Fd["ua"]), nan_like(Fd["va"])
    S = e.S
    args = [dev(S[n]) for n in ("ub", "vb", "wn", "avm")] + [dev(S["utau_b"]), dev(S["vtau_b"])]
    Fd["rCdU_bot"], Fd["utau"], Fd["vtau"] = dev(S["rCdU_bot"]), dev(S["utau"]), dev(S["vtau"])
    shapes = {k: (e.s3 if k in ("ua", "va", "un", "vn") else e.s2) for k in Fo}
    for kt in (1, 2, 3):
        oracle_dyn_step(e, kt, Fo)
        hip.step_dyn(kt, *args, Fd)
        sync(hip)
        for k in ["ua", "va", "un", "vn"] + SPG_OUT:
            check(host(Fd[k], shapes[k]), Fo[k], what=f"step_dyn {opt} kt={kt} {k}")


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["gyre", "cyc7"])
@pytest.mark.parametrize("opt", ["ene_vec", "een_vec", "ens_flx"])
def test_step_dyn_fp32_equals_separate_calls(hip, case, opt):
    e = Env(hip, case, opts=DYN_OPTS[opt])
    S = e.S
    f32 = lambda n: dev(S[n], torch.float32)
    base = spg_fields(e)
    Fa = {k: dev(v) for k, v in base.items()}
    for k in ("un", "vn"):
        Fa[k] = dev(base[k], torch.float32)
    Fb = {k: v.clone() for k, v in Fa.items()}
    Fb["ua"], Fb["va"] = torch.full_like(Fa["un"], float("nan")), torch.full_like(Fa["un"], float("nan"))
    ub, vb, wn, avm = f32("ub"), f32("vb"), f32("wn"), f32("avm")
    two = {n: dev(S[n]) for n in ("rCdU_bot", "utau", "utau_b", "vtau", "vtau_b")}
    Fa.update(rCdU_bot=two["rCdU_bot"], utau=two["utau"], vtau=two["vtau"])
    Fb.update(rCdU_bot=two["rCdU_bot"], utau=two["utau"], vtau=two["vtau"])
    for kt in (1, 2, 3):
        # separate calls
        ua, va = torch.zeros_like(Fa["un"]), torch.zeros_like(Fa["un"])
        if e.d.opts["n_dynadv"] != abi.NP_LIN_DYN:
            hip.dyn_adv(kt, Fa["un"], Fa["vn"], wn, ua, va, sp=True)
        hip.dyn_vor(kt, Fa["un"], Fa["vn"], ua, va, sp=True)
        Fa["ua"], Fa["va"] = ua, va
        hip.dyn_spg_ts(kt, Fa, sp=True)
        hip.dyn_zdf(kt, ub, vb, ua, va, avm, two["rCdU_bot"], Fa["ua_b"], Fa["va_b"], two["utau"], two["utau_b"],
                    two["vtau"], two["vtau_b"], sp=True)
        # fused
        hip.step_dyn(kt, ub, vb, wn, avm, two["utau_b"], two["vtau_b"], Fb, sp=True)
        sync(hip)
        for k in ["ua", "va", "un", "vn"] + SPG_OUT:
            a, b = Fa[k].cpu().numpy(), Fb[k].cpu().numpy()
            assert np.array_equal(a, b), f"step_dyn_sp {opt} kt={kt} {k}: max |diff| {np.nanmax(np.abs(a - b))}"


# ----------------------------------------------------------------------------
# tracers
# ----------------------------------------------------------------------------
def oracle_tra_step(e, kt, orders, ah, akz):
    S, o = e.S, e.o
    z = [np.zeros(e.s3, order="F") for _ in range(3)]
    tsa = np.zeros(e.s4, order="F")
    o.tra_adv_trp(S["un"], S["vn"], S["wn"], *z)
    o.tra_adv_fct(kt, 1, "TRA", r2dt(e.d, kt), *z, S["tsb"], S["tsn"], tsa, 2, *orders)
    o.tra_ldf_iso(kt, 1, "TRA", S["ahtu"], S["ahtv"], S["tsb"], S["tsb"], tsa, 2, 1, S["uslp"], S["vslp"],
                  S["wslpi"], S["wslpj"], ah, akz)
    o.tra_zdf(kt, S["tsb"], tsa, S["avt"], S["avt"], 1, ah, akz)
    return tsa


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASE_IDS)
@pytest.mark.parametrize("orders", [(4, 4), (2, 2), (4, 2), (2, 4)])
@pytest.mark.parametrize("msc", [0, 1])
def test_step_tra_fp64(hip, case, orders, msc):
    e = Env(hip, case, opts={"ln_traldf_msc": msc})
    S = e.S
    ah_o, akz_o = np.zeros(e.s3, order="F"), np.zeros(e.s3, order="F")
    ah_d, akz_d = torch.zeros(ah_o.size, dtype=torch.float64, device="cuda"), \
        torch.zeros(akz_o.size, dtype=torch.float64, device="cuda")
    f = {n: dev(S[n]) for n in ("un", "vn", "wn", "tsb", "tsn", "ahtu", "ahtv", "uslp", "vslp", "wslpi", "wslpj",
                                "avt")}
    inner = (slice(1, -1), slice(1, -1), slice(1, -1))
    for kt in (1, 2):
        want = oracle_tra_step(e, kt, orders, ah_o, akz_o)
        tsa = torch.full((want.size,), float("nan"), dtype=torch.float64, device="cuda")
        hip.step_tra(kt, f["un"], f["vn"], f["wn"], f["tsb"], f["tsn"], tsa, *orders, f["ahtu"], f["ahtv"],
                     f["uslp"], f["vslp"], f["wslpi"], f["wslpj"], ah_d, akz_d, f["avt"])
        sync(hip)
        check(host(tsa, e.s4), want, what=f"step_tra {orders} msc={msc} kt={kt} tsa")
        check(host(ah_d, e.s3)[inner], ah_o[inner], what=f"step_tra kt={kt} ah_wslp2")
        if msc:
            check(host(akz_d, e.s3)[inner], akz_o[inner], what=f"step_tra kt={kt} akz")


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["gyre", "basin", "cyc7"])
@pytest.mark.parametrize("orders", [(4, 4), (2, 2)])
def test_step_tra_fp32_equals_separate_calls(hip, case, orders):
    e = Env(hip, case)
    S = e.S
    f = {n: dev(S[n], torch.float32) for n in ("un", "vn", "wn", "tsb", "tsn", "ahtu", "ahtv", "uslp", "vslp",
                                               "wslpi", "wslpj", "avt")}
    n3 = e.d.jpi * e.d.jpj * e.d.jpk
    z = lambda n: torch.zeros(n, dtype=torch.float32, device="cuda")
    for kt in (1, 2):
        zun, zvn, zwn = z(n3), z(n3), z(n3)
        ah, akz = z(n3), z(n3)
        tsa = z(2 * n3)
        hip.tra_adv_trp(f["un"], f["vn"], f["wn"], zun, zvn, zwn, sp=True)
        hip.tra_adv_fct(kt, 1, "TRA", r2dt(e.d, kt), zun, zvn, zwn, f["tsb"], f["tsn"], tsa, 2, *orders, sp=True)
        hip.tra_ldf_iso(kt, 1, "TRA", f["ahtu"], f["ahtv"], f["tsb"], f["tsb"], tsa, 2, 1, f["uslp"], f["vslp"],
                        f["wslpi"], f["wslpj"], ah, akz, sp=True)
        hip.tra_zdf(kt, f["tsb"], tsa, f["avt"], f["avt"], 1, ah, akz, sp=True)
        tsa2 = torch.full_like(tsa, float("nan"))
        ah2, akz2 = z(n3), z(n3)
        hip.step_tra(kt, f["un"], f["vn"], f["wn"], f["tsb"], f["tsn"], tsa2, *orders, f["ahtu"], f["ahtv"],
                     f["uslp"], f["vslp"], f["wslpi"], f["wslpj"], ah2, akz2, f["avt"], sp=True)
        sync(hip)
        a, b = tsa.cpu().numpy(), tsa2.cpu().numpy()
        assert np.array_equal(a, b), f"step_tra_sp {orders} kt={kt}: max |diff| {np.nanmax(np.abs(a - b))}"
