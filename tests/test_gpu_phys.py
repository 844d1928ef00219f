"""Parity of the density (simplified EOS), pressure-gradient (hpg_zco) and
lateral-viscosity (dyn_ldf_lap) kernels with the oracle.

FP64: <= 1e-12 max relative error, bit-identical expected.  FP32 storage
form: the kernels compute in FP64 from the float inputs and round once, so
against the oracle on the same float-rounded inputs the bound is 1e-6.
"""
import numpy as np
import pytest

import nemo_hip
from nemo_hip import NEMO_HIP_PTR_DEVICE, NEMO_HIP_PTR_HOST, abi, configs
from test_gpu_parity import CASE_IDS, Env, check, dev, host, sync

torch = pytest.importorskip("torch")

SP_TOL = 1e-6


def rnd(rng, shape, sp, lo=None, hi=None):
    x = rng.standard_normal(shape) if lo is None else rng.uniform(lo, hi, shape)
    if sp:
        x = x.astype(np.float32).astype(np.float64)
    return np.asfortranarray(x)


def tracers(rng, e, sp):
    ts = np.zeros(e.s4, order="F")
    ts[..., 0] = rng.uniform(-2, 30, e.s3)
    ts[..., 1] = rng.uniform(30, 40, e.s3)
    if sp:
        ts = np.asfortranarray(ts.astype(np.float32).astype(np.float64))
    return ts


def g(a, sp):
    return dev(a, torch.float32 if sp else torch.float64)


def cmp(got, want, sp, what):
    check(got, want, tol=SP_TOL if sp else 1e-12, what=what)


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASE_IDS)
@pytest.mark.parametrize("sp", [False, True])
def test_eos(hip, case, sp):
    e = Env(hip, case)
    hip.eos_register(abi.seos())
    e.o.eos_register(abi.seos())
    rng = np.random.default_rng(71)
    ts = tracers(rng, e, sp)
    gdept, gdepw = configs.depths(e.d)
    if sp:
        gdept = np.asfortranarray(gdept.astype(np.float32).astype(np.float64))
        gdepw = np.asfortranarray(gdepw.astype(np.float32).astype(np.float64))
    # in situ density
    prd = rnd(rng, e.s3, sp)
    want = prd.copy(order="F")
    e.o.eos_insitu(ts, want, gdept)
    gp = g(prd, sp)
    hip.eos_insitu(g(ts, sp), gp, g(gdept, sp), sp=sp)
    # in situ and potential density
    want2, wantp = prd.copy(order="F"), prd.copy(order="F")
    e.o.eos_insitu_pot(ts, want2, wantp, gdept)
    gp2, gpp = g(prd, sp), g(prd, sp)
    hip.eos_insitu_pot(g(ts, sp), gp2, gpp, g(gdept, sp), sp=sp)
    # expansion ratios and N^2
    pab = rnd(rng, e.s4, sp)
    wab = pab.copy(order="F")
    e.o.rab_3d(ts, wab, gdept)
    gab = g(pab, sp)
    hip.rab_3d(g(ts, sp), gab, g(gdept, sp), sp=sp)
    pn2 = rnd(rng, e.s3, sp)
    wn2 = pn2.copy(order="F")
    e.o.bn2(ts, wab, wn2, gdept, gdepw)
    gn2 = g(pn2, sp)
    hip.bn2(g(ts, sp), gab, gn2, g(gdept, sp), g(gdepw, sp), sp=sp)
    sync(hip)
    cmp(host(gp, e.s3), want, sp, f"eos_insitu {case}")
    cmp(host(gp2, e.s3), want2, sp, f"eos_insitu_pot prd {case}")
    cmp(host(gpp, e.s3), wantp, sp, f"eos_insitu_pot prhop {case}")
    cmp(host(gab, e.s4), wab, sp, f"rab_3d {case}")
    if not sp:   # N^2 of FP32 ratios differs from the oracle's FP64 ratios by their rounding
        cmp(host(gn2, e.s3), wn2, sp, f"bn2 {case}")


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["gyre", "basin", "cyc7"])
@pytest.mark.parametrize("scheme", ["teos10", "eos80"])
@pytest.mark.parametrize("sp", [False, True])
def test_eos_poly(hip, case, scheme, sp):
    """Polynomial TEOS-10 / EOS-80: eos_insitu, eos_insitu_pot and rab_3d
    against the oracle, then N^2 from the kernel's own ratios."""
    e = Env(hip, case)
    mk = abi.teos10 if scheme == "teos10" else abi.eos80
    hip.eos_register(mk())
    e.o.eos_register(mk())
    rng = np.random.default_rng(72)
    ts = tracers(rng, e, sp)
    gdept, gdepw = configs.depths(e.d)
    if sp:
        gdept = np.asfortranarray(gdept.astype(np.float32).astype(np.float64))
        gdepw = np.asfortranarray(gdepw.astype(np.float32).astype(np.float64))
    prd = rnd(rng, e.s3, sp)
    want = prd.copy(order="F")
    e.o.eos_insitu(ts, want, gdept)
    gp = g(prd, sp)
    hip.eos_insitu(g(ts, sp), gp, g(gdept, sp), sp=sp)
    want2, wantp = prd.copy(order="F"), prd.copy(order="F")
    e.o.eos_insitu_pot(ts, want2, wantp, gdept)
    gp2, gpp = g(prd, sp), g(prd, sp)
    hip.eos_insitu_pot(g(ts, sp), gp2, gpp, g(gdept, sp), sp=sp)
    pab = rnd(rng, e.s4, sp)
    wab = pab.copy(order="F")
    e.o.rab_3d(ts, wab, gdept)
    gab = g(pab, sp)
    hip.rab_3d(g(ts, sp), gab, g(gdept, sp), sp=sp)
    sync(hip)
    cmp(host(gp, e.s3), want, sp, f"{scheme} eos_insitu {case}")
    cmp(host(gp2, e.s3), want2, sp, f"{scheme} eos_insitu_pot prd {case}")
    cmp(host(gpp, e.s3), wantp, sp, f"{scheme} eos_insitu_pot prhop {case}")
    cmp(host(gab, e.s4), wab, sp, f"{scheme} rab_3d {case}")
    if not sp:
        pn2 = rnd(rng, e.s3, sp)
        wn2 = pn2.copy(order="F")
        e.o.bn2(ts, wab, wn2, gdept, gdepw)
        gn2 = g(pn2, sp)
        hip.bn2(g(ts, sp), gab, gn2, g(gdept, sp), g(gdepw, sp), sp=sp)
        sync(hip)
        cmp(host(gn2, e.s3), wn2, sp, f"{scheme} bn2 {case}")
    hip.eos_register(abi.seos())


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["basin", "cyc7"])
def test_bn2_fp32_against_same_ratios(hip, case):
    """bn2 in FP32 storage on ratios both sides read as the same floats."""
    e = Env(hip, case)
    rng = np.random.default_rng(72)
    ts = tracers(rng, e, True)
    pab = rnd(rng, e.s4, True, 1e-5, 1e-3)
    gdept, gdepw = configs.depths(e.d)
    gdept = np.asfortranarray(gdept.astype(np.float32).astype(np.float64))
    gdepw = np.asfortranarray(gdepw.astype(np.float32).astype(np.float64))
    want = np.zeros(e.s3, order="F")
    e.o.bn2(ts, pab, want, gdept, gdepw)
    gn2 = g(np.zeros(e.s3, order="F"), True)
    hip.bn2(g(ts, True), g(pab, True), gn2, g(gdept, True), g(gdepw, True), sp=True)
    sync(hip)
    cmp(host(gn2, e.s3), want, True, f"bn2 fp32 {case}")


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASE_IDS)
@pytest.mark.parametrize("sp", [False, True])
def test_hpg_zco(hip, case, sp):
    e = Env(hip, case)
    rng = np.random.default_rng(73)
    rhd = rnd(rng, e.s3, sp)
    ua, va = rnd(rng, e.s3, sp), rnd(rng, e.s3, sp)
    wu, wv = ua.copy(order="F"), va.copy(order="F")
    e.o.hpg_zco(3, rhd, wu, wv)
    gu, gv = g(ua, sp), g(va, sp)
    hip.hpg_zco(3, g(rhd, sp), gu, gv, sp=sp)
    sync(hip)
    cmp(host(gu, e.s3), wu, sp, f"hpg_zco ua {case}")
    cmp(host(gv, e.s3), wv, sp, f"hpg_zco va {case}")


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASE_IDS)
@pytest.mark.parametrize("kpass", [1, 2])
@pytest.mark.parametrize("sp", [False, True])
def test_dyn_ldf_lap(hip, case, kpass, sp):
    e = Env(hip, case)
    rng = np.random.default_rng(74 + kpass)
    t, u, v, w, f = configs.masks(e.d)
    ub, vb = rnd(rng, e.s3, sp), rnd(rng, e.s3, sp)
    ahmt = np.asfortranarray(rnd(rng, e.s3, sp, 500, 2000) * t)
    ahmf = np.asfortranarray(rnd(rng, e.s3, sp, 500, 2000) * f)
    ua, va = rnd(rng, e.s3, sp), rnd(rng, e.s3, sp)
    wu, wv = ua.copy(order="F"), va.copy(order="F")
    e.o.dyn_ldf_lap(3, ub, vb, wu, wv, kpass, ahmt, ahmf)
    gu, gv = g(ua, sp), g(va, sp)
    hip.dyn_ldf_lap(3, g(ub, sp), g(vb, sp), gu, gv, kpass, g(ahmt, sp), g(ahmf, sp), sp=sp)
    sync(hip)
    cmp(host(gu, e.s3), wu, sp, f"dyn_ldf_lap ua {case}")
    cmp(host(gv, e.s3), wv, sp, f"dyn_ldf_lap va {case}")


@pytest.mark.gpu
def test_phys_host_pointer_mode_and_errors(hip):
    e = Env(hip, "cyc1")
    hip.eos_register(abi.seos())
    e.o.eos_register(abi.seos())
    rng = np.random.default_rng(75)
    ts = tracers(rng, e, False)
    gdept, _ = configs.depths(e.d)
    try:
        hip.set_pointer_mode(NEMO_HIP_PTR_HOST)
        a, b = np.zeros(e.s3, order="F"), np.zeros(e.s3, order="F")
        e.o.eos_insitu(ts, a, gdept)
        hip.eos_insitu(ts, b, gdept)
        check(b, a, what="host eos_insitu")
        ua, va = rnd(rng, e.s3, False), rnd(rng, e.s3, False)
        wu, wv = ua.copy(order="F"), va.copy(order="F")
        e.o.hpg_zco(2, a, wu, wv)
        hip.hpg_zco(2, b, ua, va)
        check(ua, wu, what="host hpg ua")
        check(va, wv, what="host hpg va")
    finally:
        hip.set_pointer_mode(NEMO_HIP_PTR_DEVICE)
    with pytest.raises(nemo_hip.NemoHipError, match="neos"):
        bad = abi.seos()
        bad.neos = 5
        hip.eos_register(bad)
    x3 = dev(np.zeros(e.s3))
    with pytest.raises(nemo_hip.NemoHipError, match="kpass"):
        hip.dyn_ldf_lap(1, x3, x3, x3, x3, 3, x3, x3)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["gyre", "cyc7"])
@pytest.mark.parametrize("sp", [False, True])
def test_stp_ctl(hip, case, sp):
    """Maxima, tests, status and positions equal the oracle's (a maximum is
    exact in any order), for a healthy state, a blow-up and a NaN."""
    e = Env(hip, case)
    t, u, v, w, f = configs.masks(e.d)
    rng = np.random.default_rng(76)
    ssh = np.asfortranarray(rng.uniform(-1, 1, e.s2))
    un = rnd(rng, e.s3, sp)
    ts = tracers(rng, e, sp)
    ts[..., 0] *= t
    ts[..., 1] *= t
    wet = t == 1
    iw = np.argwhere(wet.transpose(2, 1, 0))[0][::-1]

    def both():
        want = e.o.stp_ctl(5, ssh, un, ts)
        got = hip.stp_ctl(5, dev(ssh), g(un, sp), g(ts, sp), sp=sp)
        return want, got

    for scenario in ("healthy", "blow-up", "nan"):
        if scenario == "blow-up":
            un[3, 4, 1] = 12.0
            ts[iw[0], iw[1], iw[2], 1] = -0.5
        if scenario == "nan":
            un[3, 4, 1] = 0.0
            ts[iw[0], iw[1], iw[2], 1] = 35.0
            ssh[5, 3] = np.nan
        (st0, z0, l0, k0), (st1, z1, l1, k1, msg) = both()
        assert st1 == st0, (scenario, st0, st1, msg)
        assert np.array_equal(z1, z0, equal_nan=True), (scenario, z0, z1)
        assert k1 == k0
        if st0:
            assert np.array_equal(l1, l0), (scenario, l0, l1)
            assert "stp_ctl" in msg


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["basin", "cyc7"])
@pytest.mark.parametrize("sp", [False, True])
def test_tra_sbc(hip, case, sp):
    e = Env(hip, case)
    rng = np.random.default_rng(77)
    ts = tracers(rng, e, sp)
    fields = {n: np.asfortranarray(rng.standard_normal(e.s2) * sc) for n, sc in
              (("qns", 100.0), ("qsr", 50.0), ("sfx", 1e-3), ("emp", 1e-5))}
    sbc, sbc_b = np.zeros(e.s2 + (2,), order="F"), np.zeros(e.s2 + (2,), order="F")
    o = {n: v.copy(order="F") for n, v in fields.items()}
    o.update(sbc=sbc.copy(order="F"), sbc_b=sbc_b.copy(order="F"))
    gd = {n: dev(v) for n, v in o.items()}
    tsa = rnd(rng, e.s4, sp)
    wta, gta = tsa.copy(order="F"), g(tsa, sp)
    for kt in (1, 2, 3):
        for n in ("qns", "qsr"):            # the forcing module refreshes the fluxes every step
            o[n][:] = fields[n]
            gd[n].copy_(dev(fields[n]))
        torch.cuda.synchronize()            # the refresh is on torch's stream, the library runs on its own
        e.o.tra_sbc(kt, o["qns"], o["qsr"], o["sfx"], o["emp"], ts, wta, o["sbc"], o["sbc_b"])
        hip.tra_sbc(kt, gd["qns"], gd["qsr"], gd["sfx"], gd["emp"], g(ts, sp), gta, gd["sbc"], gd["sbc_b"], sp=sp)
        sync(hip)
        cmp(host(gta, e.s4), wta, sp, f"tra_sbc {case} kt={kt} tsa")
        for n in ("qns", "qsr"):
            check(host(gd[n], e.s2), o[n], what=f"tra_sbc {n}")
        for n in ("sbc", "sbc_b"):
            check(host(gd[n], e.s2 + (2,)), o[n], tol=SP_TOL if sp else 1e-12, what=f"tra_sbc {n}")


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["basin", "cyc7"])
@pytest.mark.parametrize("sp", [False, True])
def test_tra_qsr(hip, case, sp):
    """tra_qsr (two bands) against the oracle over an Euler first step, two
    leapfrog steps and a restart start; the device's exp differs from the
    host's libm by a few units in the last place, so FP64 is held to 1e-12"""
    e = Env(hip, case)
    rng = np.random.default_rng(78)
    gdepw = configs.depths(e.d)[1]
    if sp:
        gdepw = np.asfortranarray(gdepw.astype(np.float32).astype(np.float64))
    nksr = e.o.qsr_nksr(gdepw)
    qsr = np.asfortranarray(rng.uniform(0.0, 300.0, e.s2))
    qsr[3, 4] = 0.0
    hc, hb = np.zeros(e.s3, order="F"), np.zeros(e.s3, order="F")
    fr = np.zeros(e.s2, order="F")
    ghc, ghb, gfr, gq, gw = dev(hc), dev(hb), dev(fr), dev(qsr), g(gdepw, sp)
    tsa = rnd(rng, e.s4, sp)
    wta, gta = tsa.copy(order="F"), g(tsa, sp)
    for kt, kb in ((1, 0), (2, 0), (3, 0), (1, 1)):
        e.o.tra_qsr(kt, nksr, qsr, gdepw, hc, hb, wta, fr, kqsr_b=kb)
        hip.tra_qsr(kt, nksr, gq, gw, ghc, ghb, gta, gfr, kqsr_b=kb, sp=sp)
        sync(hip)
        cmp(host(gta, e.s4), wta, sp, f"tra_qsr {case} kt={kt} tsa")
        for n, a, b in (("qsr_hc", ghc, hc), ("qsr_hc_b", ghb, hb)):
            check(host(a, e.s3), b, tol=SP_TOL if sp else 1e-12, what=f"tra_qsr {case} kt={kt} {n}")
        check(host(gfr, e.s2), fr, tol=SP_TOL if sp else 1e-12, what=f"tra_qsr {case} kt={kt} fraqsr_1lev")
    with pytest.raises(nemo_hip.NemoHipError, match="nksr"):
        hip.tra_qsr(2, 0, gq, gw, ghc, ghb, gta, gfr, sp=sp)


@pytest.mark.gpu
@pytest.mark.parametrize("sp", [False, True])
@pytest.mark.parametrize("mode", ["device", "host"])
def test_tra_sbc_then_qsr(hip, sp, mode):
    """step.F90's order, tra_sbc then tra_qsr, with ln_traqsr = 1 registered:
    tra_sbc leaves qsr to tra_qsr (trasbc.F90:97-100), so the solar heat is
    spread over the levels down to nksr instead of entering level 1 through
    qns.  Checked against the oracle run in the same order, in device- and
    host-pointer mode and both storage forms."""
    e = Env(hip, "basin", opts=dict(ln_traqsr=1))
    rng = np.random.default_rng(79)
    ts = tracers(rng, e, sp)
    gdepw = configs.depths(e.d)[1]
    if sp:
        gdepw = np.asfortranarray(gdepw.astype(np.float32).astype(np.float64))
    nksr = e.o.qsr_nksr(gdepw)
    assert nksr > 1
    fields = {n: np.asfortranarray(rng.standard_normal(e.s2) * sc) for n, sc in
              (("qns", 100.0), ("sfx", 1e-3), ("emp", 1e-5))}
    fields["qsr"] = np.asfortranarray(rng.uniform(0.0, 300.0, e.s2))
    o = {n: v.copy(order="F") for n, v in fields.items()}
    o.update(sbc=np.zeros(e.s2 + (2,), order="F"), sbc_b=np.zeros(e.s2 + (2,), order="F"),
             hc=np.zeros(e.s3, order="F"), hb=np.zeros(e.s3, order="F"), fr=np.zeros(e.s2, order="F"))
    tsa = rnd(rng, e.s4, sp)
    wta = tsa.copy(order="F")
    fdt = np.float32 if sp else np.float64
    if mode == "device":
        gd = {n: dev(v) for n, v in o.items()}
        gta, gts, gw = g(tsa, sp), g(ts, sp), g(gdepw, sp)
    else:
        gd = {n: v.copy(order="F") for n, v in o.items()}
        gta = np.asfortranarray(tsa.astype(fdt))
        gts, gw = np.asfortranarray(ts.astype(fdt)), np.asfortranarray(gdepw.astype(fdt))
    if mode == "host":
        hip.set_pointer_mode(NEMO_HIP_PTR_HOST)
    try:
        _sbc_qsr_steps(hip, e, mode, sp, o, fields, ts, gdepw, nksr, wta, gd, gta, gts, gw)
    finally:
        hip.set_pointer_mode(NEMO_HIP_PTR_DEVICE)
    assert np.abs(o["hc"][:, :, 1:nksr]).max() > 0.0     # the heat reached below level 1


def _sbc_qsr_steps(hip, e, mode, sp, o, fields, ts, gdepw, nksr, wta, gd, gta, gts, gw):
    for kt in (1, 2, 3):
        e.o.tra_sbc(kt, o["qns"], o["qsr"], o["sfx"], o["emp"], ts, wta, o["sbc"], o["sbc_b"])
        e.o.tra_qsr(kt, nksr, o["qsr"], gdepw, o["hc"], o["hb"], wta, o["fr"])
        hip.tra_sbc(kt, gd["qns"], gd["qsr"], gd["sfx"], gd["emp"], gts, gta, gd["sbc"], gd["sbc_b"], sp=sp)
        hip.tra_qsr(kt, nksr, gd["qsr"], gw, gd["hc"], gd["hb"], gta, gd["fr"], sp=sp)
        if mode == "device":
            sync(hip)
            got = lambda a, s: host(a, s)
        else:
            got = lambda a, s: np.asarray(a, dtype=np.float64).reshape(s, order="F")
        cmp(got(gta, e.s4), wta, sp, f"tra_sbc+tra_qsr kt={kt} tsa")
        # qsr survives tra_sbc, and qns is not given the solar part
        check(got(gd["qsr"], e.s2), fields["qsr"], what="qsr kept")
        check(got(gd["qns"], e.s2), fields["qns"], what="qns without qsr")
        check(got(gd["hc"], e.s3), o["hc"], tol=SP_TOL if sp else 1e-12, what=f"qsr_hc kt={kt}")
        check(got(gd["fr"], e.s2), o["fr"], tol=SP_TOL if sp else 1e-12, what=f"fraqsr_1lev kt={kt}")
