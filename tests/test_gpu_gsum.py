"""GPU parity of the compensated global sums and dia_hsb.

* nemo_hip_glob_sum (kind 0 glob_sum, kind 1 glob_sum_full) equals
  math.fsum of the same masked terms, bit for bit, on terms spanning twenty
  decades with cancelling giants, for 2D and 3D arrays, FP64 and FP32
  storage (FP32 values are summed as FP64 terms), in device and host pointer
  modes, and on a 512x512x75 array;
* nemo_hip_dia_hsb_init writes the oracle's initial-state arrays and
  surf_tot bit for bit; nemo_hip_dia_hsb's thirteen sums equal the oracle's
  (sequential DDPDD) over three calls with forcing, in both storage forms,
  and the accumulated forcing and diagnostics follow diahsb.F90:93-233
  from those sums.
"""
import math

import numpy as np
import pytest

import nemo_hip
from nemo_hip import NEMO_HIP_PTR_DEVICE, NEMO_HIP_PTR_HOST, abi, configs
from oracle.oracle import Oracle
from test_gpu_parity import sync

torch = pytest.importorskip("torch")

RCP = 3991.86795711963


def masks(d):
    h = np.zeros((d.jpi, d.jpj))
    h[1:-1, 1:-1] = 1.0
    return h * (d.arrays["k_bot"] > 0), h


def wide(rng, shape):
    mag = 10.0 ** rng.uniform(-10, 10, shape)
    a = mag * rng.choice([-1.0, 1.0], shape)
    a.flat[7 * shape[0] + 3] = 1e16
    a.flat[11 * shape[0] + 5] = -1e16
    return np.asfortranarray(a)


@pytest.mark.gpu
@pytest.mark.parametrize("sp", [False, True])
def test_glob_sum_equals_fsum(hip, sp):
    d = configs.basin(70, 50, 9, seed=5, land_frac=0.3, jperio=0)
    hip.register(d)
    rng = np.random.default_rng(31)
    mi, mh = masks(d)
    fdt, tdt = (np.float32, torch.float32) if sp else (np.float64, torch.float64)
    for trial in range(3):
        a = wide(rng, (d.jpi, d.jpj, d.jpk)).astype(fdt)
        ga = torch.from_numpy(a.ravel(order="F").copy()).cuda()
        for nk in (1, d.jpk):
            for kind, m in ((0, mi), (1, mh)):
                out = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
                hip.glob_sum(ga, nk, kind, out=out, sp=sp)
                sync(hip)
                want = math.fsum((a[:, :, :nk].astype(np.float64) * m[:, :, None]).ravel(order="F"))
                got = float(out.cpu()[0])
                assert got == want, (trial, nk, kind, got, want)
    # host pointer mode
    hip.set_pointer_mode(NEMO_HIP_PTR_HOST)
    try:
        a = wide(rng, (d.jpi, d.jpj, d.jpk)).astype(fdt)
        got = hip.glob_sum(a, d.jpk, 1, sp=sp)
    finally:
        hip.set_pointer_mode(NEMO_HIP_PTR_DEVICE)
    assert got == math.fsum((a.astype(np.float64) * mh[:, :, None]).ravel(order="F"))
    with pytest.raises(nemo_hip.NemoHipError, match="nk"):
        hip.glob_sum(ga, d.jpk + 1, 0, out=out, sp=sp)


@pytest.mark.gpu
def test_glob_sum_large(hip):
    h = hip
    d = configs.basin(512, 512, 75, seed=9, land_frac=0.3)
    h.register(d)
    rng = np.random.default_rng(8)
    a = np.asfortranarray(rng.standard_normal((d.jpi, d.jpj, d.jpk)) * 10.0 ** rng.integers(-6, 7, (1, 1, d.jpk)))
    ga = torch.from_numpy(a.ravel(order="F").copy()).cuda()
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    h.glob_sum(ga, d.jpk, 0, out=out)
    sync(h)
    mi, _ = masks(d)
    assert float(out.cpu()[0]) == math.fsum((a * mi[:, :, None]).ravel(order="F"))


def hsb_inputs(d, rng, sp):
    s2, s3 = (d.jpi, d.jpj), (d.jpi, d.jpj, d.jpk)
    tm = d.tmask()
    f = (lambda x: np.asfortranarray(x.astype(np.float32).astype(np.float64))) if sp else np.asfortranarray
    ts = np.zeros(s3 + (2,), order="F")
    ts[..., 0] = rng.uniform(2, 20, s3)
    ts[..., 1] = rng.uniform(34, 36, s3)
    ts = f(ts)
    ssh = np.asfortranarray(rng.normal(0, 0.1, s2))
    return ts, ssh, tm


@pytest.mark.gpu
@pytest.mark.parametrize("sp", [False, True])
@pytest.mark.parametrize("traqsr", [0, 1])
def test_dia_hsb(hip, sp, traqsr):
    d = configs.basin(60, 44, 17, seed=12, land_frac=0.3, opts=dict(ln_traqsr=traqsr, rdt=900.0))
    o = Oracle(d)
    hip.register(d)
    rng = np.random.default_rng(13)
    ts, ssh, tm = hsb_inputs(d, rng, sp)
    s2, s3 = (d.jpi, d.jpj), (d.jpi, d.jpj, d.jpk)
    tdt = torch.float32 if sp else torch.float64
    dev = lambda a, t=torch.float64: torch.from_numpy(np.ascontiguousarray(a.ravel(order="F"))).to("cuda", t)
    host = lambda t, shape: t.double().cpu().numpy().reshape(shape, order="F")
    surf_tot, ini = o.dia_hsb_init(ts, ssh)
    G = {n: torch.full((int(np.prod(v.shape)),), float("nan"), dtype=torch.float64, device="cuda")
         for n, v in ini.items()}
    st = hip.dia_hsb_init(sp=sp, tsn=dev(ts, tdt), sshn=dev(ssh), **G)
    sync(hip)
    assert st.surf_tot == surf_tot
    for n, v in ini.items():
        assert np.array_equal(host(G[n], v.shape), v), n
    frc = dict(v=0.0, t=0.0, s=0.0, wn_t=0.0, wn_s=0.0)
    rau0, rdt = d.opts["rau0"], d.opts["rdt"]
    for kt in (1, 2, 3):
        ts = np.asfortranarray(ts + rng.normal(0, 0.01, ts.shape))
        if sp:
            ts = np.asfortranarray(ts.astype(np.float32).astype(np.float64))
        tsb = np.asfortranarray(ts + rng.normal(0, 0.01, ts.shape))
        if sp:
            tsb = np.asfortranarray(tsb.astype(np.float32).astype(np.float64))
        wn = np.asfortranarray(rng.normal(0, 1e-4, s3))
        if sp:
            wn = np.asfortranarray(wn.astype(np.float32).astype(np.float64))
        ssh = np.asfortranarray(ssh + rng.normal(0, 0.01, s2))
        emp = np.asfortranarray(rng.normal(0, 1e-5, s2))
        rnf = np.asfortranarray(np.abs(rng.normal(0, 1e-6, s2)))
        qsr = np.asfortranarray(rng.uniform(0, 300, s2))
        sbc = np.asfortranarray(rng.normal(0, 1e-4, s2 + (2,)))
        want = o.dia_hsb_sums(ts, tsb, wn, ssh, emp, rnf, qsr if traqsr else None, sbc, ini)
        hip.dia_hsb(kt, st, sp=sp, tsn=dev(ts, tdt), tsb=dev(tsb, tdt), wn=dev(wn, tdt), sshn=dev(ssh),
                    emp=dev(emp), rnf=dev(rnf), qsr=dev(qsr), sbc_tsc=dev(sbc), **G)
        got = np.array(st.sums[:])
        assert np.array_equal(got, want), (kt, got - want)
        # diahsb.F90:93-233 from the sums
        z = want
        ztrd_t = z[1] + (1.0 / (rau0 * RCP)) * z[3] if traqsr else z[1]
        frc["v"] += (1.0 / rau0) * z[0] * rdt
        frc["t"] += ztrd_t * rdt
        frc["s"] += z[2] * rdt
        frc["wn_t"] += -z[4] * rdt
        frc["wn_s"] += -z[5] * rdt
        assert (st.frc_v, st.frc_t, st.frc_s, st.frc_wn_t, st.frc_wn_s) == (
            frc["v"], frc["t"], frc["s"], frc["wn_t"], frc["wn_s"])
        hc1 = (z[10] - frc["t"]) + z[7]
        assert st.bgtemper == hc1 / z[12]
        assert st.bgheatfx == hc1 * rau0 * RCP / (surf_tot * kt * rdt)
        assert st.bgvolssh == (z[6] - frc["v"]) * 1.e-9
        assert st.bgmissal == (z[8] - frc["wn_s"]) / z[12]
