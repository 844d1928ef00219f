"""The column solves with the forward-substituted values on chip (tuning
switch zdf_reg = 1; FP32 storage with jpk <= 76) against the same solves
through the trend array (zdf_reg = 0, the default), bit for bit: tra_zdf
(with and without the iso-neutral coefficient, ln_zdfddm's one tracer per
matrix, the salinity reset acting), dyn_zdf (implicit bottom drag, the
only form the library takes) at the Euler and a leapfrog step, and the fused
steps whose last kernels solve the columns; on every parity basin (12 to 75
levels) and on a 77-level basin, where the switch selects the same kernels
either way.  Against the FP64 oracle the FP32 solves stay within 2e-6."""
import numpy as np
import pytest

from nemo_hip import configs
from test_gpu_parity import CASES, Env, dev, host, spg_fields, sync

torch = pytest.importorskip("torch")

f32 = lambda a: dev(a, torch.float32)
CASE_LIST = list(CASES) + ["deep77"]


def env(hip, case, opts=None):
    if case == "deep77":
        CASES["deep77"] = lambda o=None: configs.basin(30, 26, 77, seed=9, opts=o)
        try:
            return Env(hip, "deep77", opts=opts)
        finally:
            del CASES["deep77"]
    return Env(hip, case, opts=opts)


def both(hip, fn):
    default = hip.tune_get("zdf_reg")
    try:
        hip.tune_set("zdf_reg", 1)
        a = fn()
        hip.tune_set("zdf_reg", 0)
        b = fn()
    finally:
        hip.tune_set("zdf_reg", default)
    return a, b


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASE_LIST)
@pytest.mark.parametrize("variant", ["tra", "tra_slp", "ddm", "reset"])
def test_tra_zdf_reg(hip, case, variant):
    e = env(hip, case, {"ln_zdfddm": 1} if variant == "ddm" else None)
    S = e.S
    slp = 1 if variant == "tra_slp" else 0
    ah = np.asfortranarray(np.abs(S["wslpi"]))
    avs = np.asfortranarray(S["avt"] * 0.5)
    tsa = e.c("tsa")
    if variant == "reset":        # a trend that drives S negative somewhere
        tsa[:, :, :, 1] -= 1.0

    def run():
        ta = f32(tsa)
        hip.tra_zdf(4, f32(S["tsb"]), ta, f32(S["avt"]), f32(avs), slp, f32(ah), f32(ah), sp=True)
        sync(hip)
        return host(ta, e.s4)

    a, b = both(hip, run)
    assert np.array_equal(a, b), f"{case} {variant}: max |diff| {np.abs(a - b).max():.3e}"
    want = tsa.copy(order="F")
    e.o.tra_zdf(4, S["tsb"], want, S["avt"], avs, slp, ah, ah)
    err = np.abs(a - want).max() / np.abs(want).max()
    assert err <= 2e-6, f"{case} {variant}: {err:.2e} against the FP64 oracle"


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASE_LIST)
@pytest.mark.parametrize("kt", [1, 3])
def test_dyn_zdf_reg(hip, case, kt):
    e = env(hip, case)
    S = e.S
    args = [S[n] for n in ("avm", "rCdU_bot", "ua_b", "va_b", "utau", "utau_b", "vtau", "vtau_b")]

    def run():
        ua, va = f32(S["ua"]), f32(S["va"])
        hip.dyn_zdf(kt, f32(S["ub"]), f32(S["vb"]), ua, va, f32(args[0]), *[dev(a) for a in args[1:]], sp=True)
        sync(hip)
        return host(ua, e.s3), host(va, e.s3)

    (ua1, va1), (ua0, va0) = both(hip, run)
    assert np.array_equal(ua1, ua0) and np.array_equal(va1, va0), case
    uo, vo = e.c("ua"), e.c("va")
    e.o.dyn_zdf(kt, S["ub"], S["vb"], uo, vo, *args)
    for g, w, n in ((ua1, uo, "ua"), (va1, vo, "va")):
        err = np.abs(g - w).max() / np.abs(w).max()
        assert err <= 2e-6, f"{case} kt={kt} {n}: {err:.2e} against the FP64 oracle"


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["gyre", "cyc7", "deep75", "deep77"])
def test_fused_steps_reg(hip, case):
    e = env(hip, case)
    S = e.S
    n3 = int(np.prod(e.s3))

    def run_dyn():
        Fd = {k: (f32(v) if k in ("ua", "va", "un", "vn") else dev(v)) for k, v in spg_fields(e).items()}
        hip.step_dyn(3, f32(S["ub"]), f32(S["vb"]), f32(S["wn"]), f32(S["avm"]), dev(S["utau_b"]),
                     dev(S["vtau_b"]), Fd, sp=True)
        sync(hip)
        return {k: v.cpu().numpy().copy() for k, v in Fd.items()}

    default = hip.tune_get("zdf_reg")
    try:
        hip.tune_set("zdf_reg", 1)
        hip.register(e.d)
        a = run_dyn()
        hip.tune_set("zdf_reg", 0)
        hip.register(e.d)     # the barotropic levels start afresh
        b = run_dyn()
    finally:
        hip.tune_set("zdf_reg", default)
    for k in a:
        assert np.array_equal(a[k], b[k]), f"step_dyn {case} {k}"

    f = {n: f32(S[n]) for n in ("un", "vn", "wn", "ahtu", "ahtv", "uslp", "vslp", "wslpi", "wslpj", "avt")}
    gb, gn = f32(S["tsb"]), f32(S["tsn"])

    def run_tra():
        tsa = torch.full((2 * n3,), float("nan"), dtype=torch.float32, device="cuda")
        ah, akz = torch.zeros(n3, dtype=torch.float32, device="cuda"), torch.zeros(n3, dtype=torch.float32,
                                                                                  device="cuda")
        hip.step_tra(3, f["un"], f["vn"], f["wn"], gb, gn, tsa, 4, 4, f["ahtu"], f["ahtv"], f["uslp"], f["vslp"],
                     f["wslpi"], f["wslpj"], ah, akz, f["avt"], sp=True)
        sync(hip)
        return tsa.cpu().numpy()

    a, b = both(hip, run_tra)
    assert np.all(np.isfinite(a)) and np.array_equal(a, b), f"step_tra {case}"
