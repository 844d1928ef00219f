"""Passive tracers through the tracer routines (TOP/TRP/trcadv.F90:127,
trcldf.F90:95-97, trczdf.F90:56 call them with cdtype = 'TRC' and kjpt =
jptra): tra_adv_fct, tra_ldf_lap, tra_ldf_iso and tra_zdf_imp with one,
three and five tracers (the kernels work on tracer pairs, so an odd count
takes the single-tracer remainder path) against the oracle.

FP64: <= 1e-12 max relative error (bit-identical expected).  FP32 storage
form of the stencil routines: the trends from zero within 2e-5 of their
maximum against the FP64 oracle on the same float-rounded inputs (the
stencils run in FP32).
"""
import numpy as np
import pytest

from test_gpu_parity import CASE_IDS, Env, check, dev, host, sync, transports

torch = pytest.importorskip("torch")


def passive(e, kjpt, seed):
    """kjpt tracers: T and S of the state first, then positive passive
    concentrations with structure in all three directions."""
    rng = np.random.default_rng(seed)
    t, *_ = __import__("nemo_hip").configs.masks(e.d)
    s4 = e.s3 + (kjpt,)
    out = []
    for name in ("tsb", "tsn", "tsa"):
        a = np.zeros(s4, order="F")
        for n in range(kjpt):
            if n < 2:
                a[..., n] = e.S[name][..., n]
            else:
                a[..., n] = (rng.uniform(0.0, 1.0, e.s3) + np.indices(e.s3)[n % 3] / 7.0) * t
        out.append(np.asfortranarray(a))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASE_IDS)
@pytest.mark.parametrize("kjpt", [1, 3, 5])
def test_trc_routines_fp64(hip, case, kjpt):
    e = Env(hip, case)
    S = e.S
    tb, tn, ta = passive(e, kjpt, 91 + kjpt)
    s4 = e.s3 + (kjpt,)
    p2dt = 2.0 * e.d.opts["rdt"]
    zun, zvn, zwn = transports(e)
    # advection
    want = ta.copy(order="F")
    e.o.tra_adv_fct(3, 1, "TRC", p2dt, zun, zvn, zwn, tb, tn, want, kjpt, 4, 4)
    g = dev(ta)
    hip.tra_adv_fct(3, 1, "TRC", p2dt, dev(zun), dev(zvn), dev(zwn), dev(tb), dev(tn), g, kjpt, 4, 4)
    sync(hip)
    check(host(g, s4), want, what=f"TRC tra_adv_fct kjpt={kjpt}")
    # laplacian diffusion
    want = ta.copy(order="F")
    e.o.tra_ldf_lap(3, 1, "TRC", S["ahtu"], S["ahtv"], tb, want, kjpt, 1)
    g = dev(ta)
    hip.tra_ldf_lap(3, 1, "TRC", dev(S["ahtu"]), dev(S["ahtv"]), dev(tb), g, kjpt, 1)
    sync(hip)
    check(host(g, s4), want, what=f"TRC tra_ldf_lap kjpt={kjpt}")
    # iso-neutral diffusion
    ah_o, akz_o = np.zeros(e.s3, order="F"), np.zeros(e.s3, order="F")
    want = ta.copy(order="F")
    e.o.tra_ldf_iso(3, 1, "TRC", S["ahtu"], S["ahtv"], tb, tb, want, kjpt, 1, S["uslp"], S["vslp"], S["wslpi"],
                    S["wslpj"], ah_o, akz_o)
    g = dev(ta)
    ah, akz = dev(np.zeros(e.s3)), dev(np.zeros(e.s3))
    hip.tra_ldf_iso(3, 1, "TRC", dev(S["ahtu"]), dev(S["ahtv"]), dev(tb), dev(tb), g, kjpt, 1, dev(S["uslp"]),
                    dev(S["vslp"]), dev(S["wslpi"]), dev(S["wslpj"]), ah, akz)
    sync(hip)
    check(host(g, s4), want, what=f"TRC tra_ldf_iso kjpt={kjpt}")
    check(host(akz, e.s3), akz_o, what="TRC akz")
    # implicit vertical diffusion (passive tracers use avt for every tracer)
    want = ta.copy(order="F")
    e.o.tra_zdf_imp(3, 1, "TRC", p2dt, tb, want, kjpt, S["avt"], S["avs"])
    g = dev(ta)
    hip.tra_zdf_imp(3, 1, "TRC", p2dt, dev(tb), g, kjpt, dev(S["avt"]), dev(S["avs"]))
    sync(hip)
    check(host(g, s4), want, what=f"TRC tra_zdf_imp kjpt={kjpt}")


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["gyre", "cyc7"])
@pytest.mark.parametrize("kjpt", [3])
def test_trc_routines_fp32(hip, case, kjpt):
    """The trends from zero (tsa = 0, as the step starts them), inputs
    rounded to float for both sides."""
    e = Env(hip, case)
    r = lambda a: np.asfortranarray(np.asarray(a).astype(np.float32).astype(np.float64))
    S = {k: r(e.S[k]) for k in ("ahtu", "ahtv", "uslp", "vslp", "wslpi", "wslpj")}
    tb, tn, _ = (r(a) for a in passive(e, kjpt, 97))
    s4 = e.s3 + (kjpt,)
    ta = np.zeros(s4, order="F")
    p2dt = 2.0 * e.d.opts["rdt"]
    zun, zvn, zwn = (r(a) for a in transports(e))
    f = lambda a: dev(a, torch.float32)

    def cmp(got, want, what):
        g = host(got, s4)
        err = float(np.abs(g - want).max() / max(np.abs(want).max(), 1e-300))
        assert np.isfinite(g).all() and err <= 2e-5, f"{what}: {err:.3e}"

    want = ta.copy(order="F")
    e.o.tra_adv_fct(3, 1, "TRC", p2dt, zun, zvn, zwn, tb, tn, want, kjpt, 4, 4)
    g = f(ta)
    hip.tra_adv_fct(3, 1, "TRC", p2dt, f(zun), f(zvn), f(zwn), f(tb), f(tn), g, kjpt, 4, 4, sp=True)
    sync(hip)
    cmp(g, want, "TRC tra_adv_fct fp32")
    want = ta.copy(order="F")
    ah_o, akz_o = np.zeros(e.s3, order="F"), np.zeros(e.s3, order="F")
    e.o.tra_ldf_iso(3, 1, "TRC", S["ahtu"], S["ahtv"], tb, tb, want, kjpt, 1, S["uslp"], S["vslp"], S["wslpi"],
                    S["wslpj"], ah_o, akz_o)
    g = f(ta)
    ah, akz = f(np.zeros(e.s3)), f(np.zeros(e.s3))
    hip.tra_ldf_iso(3, 1, "TRC", f(S["ahtu"]), f(S["ahtv"]), f(tb), f(tb), g, kjpt, 1, f(S["uslp"]), f(S["vslp"]),
                    f(S["wslpi"]), f(S["wslpj"]), ah, akz, sp=True)
    sync(hip)
    cmp(g, want, "TRC tra_ldf_iso fp32")
