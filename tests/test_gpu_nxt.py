"""Parity of the time-step closure kernels (div_hor, ssh_nxt, wzv, ssh_swp,
tra_nxt, dyn_nxt) with the oracle.

FP64: <= 1e-12 max relative error, bit-identical expected.  Inputs are
unmasked random fields, so every ring point and land point of the output is
compared, and the outputs start from random values, so what the routines must
leave alone (level jpk, the before fields of an Euler step) is checked too.
FP32 storage form: against the oracle on the float-rounded inputs, within the
tolerance of each test (the kernels compute in FP64 and round on store).
"""
import numpy as np
import pytest

from nemo_hip import NEMO_HIP_PTR_DEVICE, NEMO_HIP_PTR_HOST
from test_gpu_parity import CASE_IDS, Env, check, dev, host, sync

torch = pytest.importorskip("torch")

SP_TOL = 1e-6


class Fields:
    """Seeded random inputs; .o(name) the oracle's copy (FP64, or the FP64
    value of the float-rounded field when sp), .g(name) a device copy."""

    def __init__(self, e, seed, sp=False, **shapes):
        rng = np.random.default_rng(seed)
        self.sp = sp
        self.a = {}
        for n, s in shapes.items():
            x = np.asfortranarray(rng.standard_normal(s))
            if sp and len(s) >= 3:
                x = np.asfortranarray(x.astype(np.float32).astype(np.float64))
            self.a[n] = x
        self.s = shapes

    def o(self, n):
        return self.a[n].copy(order="F")

    def g(self, n):
        return dev(self.a[n], torch.float32 if (self.sp and len(self.s[n]) >= 3) else torch.float64)


def compare(got, want, sp, what):
    check(got, want, tol=SP_TOL if sp else 1e-12, what=what)


# ----------------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("case", CASE_IDS)
@pytest.mark.parametrize("sp", [False, True])
def test_div_hor(hip, case, sp):
    e = Env(hip, case)
    f = Fields(e, 11, sp, un=e.s3, vn=e.s3, hdivn=e.s3)
    want = f.o("hdivn")
    e.o.div_hor(3, f.o("un"), f.o("vn"), want)
    got = f.g("hdivn")
    hip.div_hor(3, f.g("un"), f.g("vn"), got, sp=sp)
    sync(hip)
    compare(host(got, e.s3), want, sp, f"div_hor {case}")


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASE_IDS)
@pytest.mark.parametrize("kt", [1, 2, 6])
@pytest.mark.parametrize("sp", [False, True])
def test_ssh_nxt(hip, case, kt, sp):
    e = Env(hip, case)
    f = Fields(e, 12 + kt, sp, un=e.s3, vn=e.s3, hdivn=e.s3, sshb=e.s2, emp_b=e.s2, emp=e.s2, ssha=e.s2)
    h, a = f.o("hdivn"), f.o("ssha")
    e.o.ssh_nxt(kt, f.o("un"), f.o("vn"), h, f.o("sshb"), f.o("emp_b"), f.o("emp"), a)
    gh, ga = f.g("hdivn"), f.g("ssha")
    hip.ssh_nxt(kt, f.g("un"), f.g("vn"), gh, f.g("sshb"), f.g("emp_b"), f.g("emp"), ga, sp=sp)
    sync(hip)
    compare(host(gh, e.s3), h, sp, f"ssh_nxt {case} kt={kt} hdivn")
    compare(host(ga, e.s2), a, sp, f"ssh_nxt {case} kt={kt} ssha")


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASE_IDS)
@pytest.mark.parametrize("kt", [1, 2, 6])
@pytest.mark.parametrize("sp", [False, True])
def test_wzv(hip, case, kt, sp):
    e = Env(hip, case)
    f = Fields(e, 20 + kt, sp, hdivn=e.s3, wn=e.s3)
    w = f.o("wn")
    e.o.wzv(kt, f.o("hdivn"), w)
    gw = f.g("wn")
    hip.wzv(kt, f.g("hdivn"), gw, sp=sp)
    sync(hip)
    compare(host(gw, e.s3), w, sp, f"wzv {case} kt={kt}")


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["gyre", "cyc7"])
@pytest.mark.parametrize("kt", [1, 2, 6])
def test_ssh_swp(hip, case, kt):
    e = Env(hip, case)
    f = Fields(e, 30 + kt, False, sshb=e.s2, sshn=e.s2, ssha=e.s2)
    b, n = f.o("sshb"), f.o("sshn")
    e.o.ssh_swp(kt, b, n, f.o("ssha"))
    gb, gn = f.g("sshb"), f.g("sshn")
    hip.ssh_swp(kt, gb, gn, f.g("ssha"))
    sync(hip)
    check(host(gb, e.s2), b, what=f"ssh_swp kt={kt} sshb")
    check(host(gn, e.s2), n, what=f"ssh_swp kt={kt} sshn")


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASE_IDS)
@pytest.mark.parametrize("kt", [1, 2, 6])
@pytest.mark.parametrize("sp", [False, True])
def test_tra_nxt(hip, case, kt, sp):
    e = Env(hip, case)
    f = Fields(e, 40 + kt, sp, tsb=e.s4, tsn=e.s4, tsa=e.s4)
    o = {n: f.o(n) for n in ("tsb", "tsn", "tsa")}
    e.o.tra_nxt(kt, o["tsb"], o["tsn"], o["tsa"])
    g = {n: f.g(n) for n in ("tsb", "tsn", "tsa")}
    hip.tra_nxt(kt, g["tsb"], g["tsn"], g["tsa"], sp=sp)
    sync(hip)
    for n in g:
        compare(host(g[n], e.s4), o[n], sp, f"tra_nxt {case} kt={kt} {n}")


DYN_NXT_OPTS = {
    "ts_fw": {"ln_dynspg_ts": 1, "ln_bt_fw": 1},
    "ts_centred": {"ln_dynspg_ts": 1, "ln_bt_fw": 0},
    "no_ts": {"ln_dynspg_ts": 0},
}
DYN3 = ("ub", "vb", "un", "vn", "ua", "va")
DYN2 = ("ua_b", "va_b", "un_adv", "vn_adv", "ub_b", "vb_b", "un_b", "vn_b")


def dyn_nxt_pair(hip, e, f, kt, sp):
    o = {n: f.o(n) for n in DYN3 + DYN2}
    S = e.S
    e.o.dyn_nxt(kt, *[o[n] for n in DYN3 + DYN2], S["r1_hu_n"], S["r1_hv_n"])
    g = {n: f.g(n) for n in DYN3 + DYN2}
    hip.dyn_nxt(kt, *[g[n] for n in DYN3 + DYN2], dev(S["r1_hu_n"]), dev(S["r1_hv_n"]), sp=sp)
    sync(hip)
    return o, g


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASE_IDS)
@pytest.mark.parametrize("opt", list(DYN_NXT_OPTS))
@pytest.mark.parametrize("kt", [1, 2, 6])
@pytest.mark.parametrize("sp", [False, True])
def test_dyn_nxt(hip, case, opt, kt, sp):
    e = Env(hip, case, opts=DYN_NXT_OPTS[opt])
    shapes = {n: e.s3 for n in DYN3}
    shapes.update({n: e.s2 for n in DYN2})
    f = Fields(e, 50 + kt, sp, **shapes)
    o, g = dyn_nxt_pair(hip, e, f, kt, sp)
    for n in DYN3:
        compare(host(g[n], e.s3), o[n], sp, f"dyn_nxt {case} {opt} kt={kt} {n}")
    for n in ("ub_b", "vb_b", "un_b", "vn_b"):
        compare(host(g[n], e.s2), o[n], sp, f"dyn_nxt {case} {opt} kt={kt} {n}")


@pytest.mark.gpu
def test_closure_host_pointer_mode(hip):
    """The same routines on host arrays (staged through the library)."""
    e = Env(hip, "cyc7")
    try:
        hip.set_pointer_mode(NEMO_HIP_PTR_HOST)
        f = Fields(e, 60, False, un=e.s3, vn=e.s3, hdivn=e.s3, wn=e.s3, sshb=e.s2, emp_b=e.s2, emp=e.s2,
                   ssha=e.s2, tsb=e.s4, tsn=e.s4, tsa=e.s4)
        h1, a1, w1 = f.o("hdivn"), f.o("ssha"), f.o("wn")
        h2, a2, w2 = f.o("hdivn"), f.o("ssha"), f.o("wn")
        e.o.ssh_nxt(4, f.o("un"), f.o("vn"), h1, f.o("sshb"), f.o("emp_b"), f.o("emp"), a1)
        e.o.wzv(4, h1, w1)
        hip.ssh_nxt(4, f.o("un"), f.o("vn"), h2, f.o("sshb"), f.o("emp_b"), f.o("emp"), a2)
        hip.wzv(4, h2, w2)
        check(h2, h1, what="host ssh_nxt hdivn")
        check(a2, a1, what="host ssh_nxt ssha")
        check(w2, w1, what="host wzv")
        t1 = {n: f.o(n) for n in ("tsb", "tsn", "tsa")}
        t2 = {n: f.o(n) for n in ("tsb", "tsn", "tsa")}
        e.o.tra_nxt(4, t1["tsb"], t1["tsn"], t1["tsa"])
        hip.tra_nxt(4, t2["tsb"], t2["tsn"], t2["tsa"])
        for n in t1:
            check(t2[n], t1[n], what=f"host tra_nxt {n}")
        shapes = {n: e.s3 for n in DYN3}
        shapes.update({n: e.s2 for n in DYN2})
        fd = Fields(e, 61, False, **shapes)
        o1 = {n: fd.o(n) for n in DYN3 + DYN2}
        o2 = {n: fd.o(n) for n in DYN3 + DYN2}
        e.o.dyn_nxt(4, *[o1[n] for n in DYN3 + DYN2], e.S["r1_hu_n"], e.S["r1_hv_n"])
        hip.dyn_nxt(4, *[o2[n] for n in DYN3 + DYN2], e.c("r1_hu_n"), e.c("r1_hv_n"))
        for n in DYN3 + ("ub_b", "vb_b", "un_b", "vn_b"):
            check(o2[n], o1[n], what=f"host dyn_nxt {n}")
    finally:
        hip.set_pointer_mode(NEMO_HIP_PTR_DEVICE)


@pytest.mark.gpu
def test_closure_argument_errors(hip):
    import nemo_hip
    e = Env(hip, "basin")
    x3 = dev(np.zeros(e.s3))
    x2 = dev(np.zeros(e.s2))
    with pytest.raises(nemo_hip.NemoHipError, match="NULL"):
        hip.div_hor(1, x3, None, x3)
    with pytest.raises(nemo_hip.NemoHipError, match="NULL"):
        hip.ssh_nxt(1, x3, x3, x3, x2, None, x2, x2)
    with pytest.raises(nemo_hip.NemoHipError, match="ua_b"):
        hip.dyn_nxt(3, x3, x3, x3, x3, x3, x3, None, None, None, None, x2, x2, x2, x2, x2, x2)
