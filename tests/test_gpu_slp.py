"""Parity of the mixed-layer level (zdf_mxl) and the iso-neutral slopes
(ldf_slp) with the oracle, on stratified states whose density comes from
the oracle's equation of state and N^2 from its bn2.

FP64: <= 1e-12 max relative error (bit-identical expected); nmln is
compared exactly.  FP32 storage form: the kernels compute in FP64 from the
float inputs and round once, so against the oracle on the same
float-rounded inputs the bound is 1e-6 (nmln, hmlp and hmlpt exact).
"""
import numpy as np
import pytest

import nemo_hip
from nemo_hip import NEMO_HIP_PTR_DEVICE, NEMO_HIP_PTR_HOST, abi, configs
from test_gpu_parity import CASE_IDS, Env, check, dev, host, sync
from test_oracle_slp import nlb10_of

torch = pytest.importorskip("torch")

SP_TOL = 1e-6


def f32(a):
    return np.asfortranarray(np.asarray(a).astype(np.float32).astype(np.float64))


def g(a, sp):
    return dev(a, torch.float32 if sp else torch.float64)


def stratified(e, eos, sp, seed=81):
    """prd, pn2, gdept, gdepw of a warm-over-cold state with small horizontal
    noise, a well-mixed upper layer over the western half and strong noise
    over the southern half (steep and unstable columns)."""
    rng = np.random.default_rng(seed)
    gdept, gdepw = configs.depths(e.d)
    ts = np.zeros(e.s4, order="F")
    ts[..., 0] = 2.0 + 18.0 * np.exp(-gdept / 700.0) + 0.05 * rng.standard_normal(e.s3)
    ts[..., 1] = 34.6 + 0.4 * gdept / 5000.0 + 0.01 * rng.standard_normal(e.s3)
    half = e.d.jpi // 2
    for n in range(2):
        ts[:half, :, :4, n] = ts[:half, :, 3:4, n]
    ts[:, : e.d.jpj // 2, :, 0] += 0.6 * rng.standard_normal((e.d.jpi, e.d.jpj // 2, e.d.jpk))
    if sp:
        ts, gdept, gdepw = f32(ts), f32(gdept), f32(gdepw)
    e.o.eos_register(eos)
    prd = np.zeros(e.s3, order="F")
    e.o.eos_insitu(ts, prd, gdept)
    pab = np.zeros(e.s4, order="F")
    e.o.rab_3d(ts, pab, gdept)
    pn2 = np.zeros(e.s3, order="F")
    e.o.bn2(ts, pab, pn2, gdept, gdepw)
    if sp:
        prd, pn2 = f32(prd), f32(pn2)
    return prd, pn2, gdept, gdepw


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASE_IDS)
@pytest.mark.parametrize("sp", [False, True])
def test_zdf_mxl(hip, case, sp):
    e = Env(hip, case)
    prd, pn2, gdept, gdepw = stratified(e, abi.seos(), sp)
    nlb10 = nlb10_of(e.d)
    wn, wh, wt = e.o.zdf_mxl(1, pn2, gdept, gdepw, nlb10)
    nm = torch.full(e.s2, -5, dtype=torch.int32, device="cuda").t().contiguous()
    hp = torch.full((e.d.jpi * e.d.jpj,), 7.0, dtype=torch.float64, device="cuda")
    ht = torch.full((e.d.jpi * e.d.jpj,), 7.0, dtype=torch.float64, device="cuda")
    hip.zdf_mxl(1, g(pn2, sp), g(gdept, sp), g(gdepw, sp), nlb10, nm, hp, ht, sp=sp)
    sync(hip)
    got_n = nm.cpu().numpy().reshape(-1).reshape(e.s2, order="F")
    assert np.array_equal(got_n, wn), f"nmln {case}"
    assert len(np.unique(wn)) >= 3, "the state should give several mixed-layer levels"
    check(host(hp, e.s2), wh, tol=0.0, what=f"hmlp {case}")
    check(host(ht, e.s2), wt, tol=0.0, what=f"hmlpt {case}")


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASE_IDS)
@pytest.mark.parametrize("eos", ["seos", "teos10"])
@pytest.mark.parametrize("sp", [False, True])
def test_ldf_slp(hip, case, eos, sp):
    e = Env(hip, case)
    prd, pn2, gdept, gdepw = stratified(e, abi.seos() if eos == "seos" else abi.teos10(), sp)
    nlb10 = nlb10_of(e.d)
    nmln, hmlp, hmlpt = e.o.zdf_mxl(1, pn2, gdept, gdepw, nlb10)
    rng = np.random.default_rng(82)
    init = [np.asfortranarray(rng.uniform(-1e-3, 1e-3, e.s3)) for _ in range(4)]
    if sp:
        init = [f32(a) for a in init]
    want = [a.copy(order="F") for a in init]
    e.o.ldf_slp(1, prd, pn2, gdept, gdepw, nmln, hmlp, hmlpt, 0.01, *want)
    got = [g(a, sp) for a in init]
    nm = dev(nmln.astype(np.float64)).to(torch.int32)
    hip.ldf_slp(1, g(prd, sp), g(pn2, sp), g(gdept, sp), g(gdepw, sp), nm, dev(hmlp), dev(hmlpt), 0.01, *got,
                sp=sp)
    sync(hip)
    for name, a, w in zip(("uslp", "vslp", "wslpi", "wslpj"), got, want):
        check(host(a, e.s3), w, tol=SP_TOL if sp else 1e-12, what=f"{name} {case} {eos}")
    # steep (limited before the filter) and gentle slopes are both present
    wi = np.abs(want[2][:, :, 1:-1])
    assert wi.max() > 5e-3 and ((wi > 0) & (wi < 1e-3)).any()


@pytest.mark.gpu
def test_slp_host_pointer_mode_and_errors(hip):
    e = Env(hip, "cyc1")
    prd, pn2, gdept, gdepw = stratified(e, abi.seos(), False)
    nlb10 = nlb10_of(e.d)
    wn, wh, wt = e.o.zdf_mxl(1, pn2, gdept, gdepw, nlb10)
    w = [np.zeros(e.s3, order="F") for _ in range(4)]
    e.o.ldf_slp(1, prd, pn2, gdept, gdepw, wn, wh, wt, 0.02, *w)
    try:
        hip.set_pointer_mode(NEMO_HIP_PTR_HOST)
        nm = np.zeros(e.s2, dtype=np.int32, order="F")
        hp, ht = np.zeros(e.s2, order="F"), np.zeros(e.s2, order="F")
        hip.zdf_mxl(1, pn2, gdept, gdepw, nlb10, nm, hp, ht)
        assert np.array_equal(nm, wn)
        check(hp, wh, what="host hmlp")
        out = [np.zeros(e.s3, order="F") for _ in range(4)]
        hip.ldf_slp(1, prd, pn2, gdept, gdepw, nm, hp, ht, 0.02, *out)
        for a, b in zip(out, w):
            check(a, b, what="host ldf_slp")
    finally:
        hip.set_pointer_mode(NEMO_HIP_PTR_DEVICE)
    x3 = dev(np.zeros(e.s3))
    x2 = dev(np.zeros(e.s2))
    n2 = torch.zeros(e.s2[0] * e.s2[1], dtype=torch.int32, device="cuda")
    with pytest.raises(nemo_hip.NemoHipError, match="nlb10"):
        hip.zdf_mxl(1, x3, x3, x3, 1, n2, x2, x2)
    with pytest.raises(nemo_hip.NemoHipError, match="rn_slpmax"):
        hip.ldf_slp(1, x3, x3, x3, x3, n2, x2, x2, 0.0, x3, x3, x3, x3)
