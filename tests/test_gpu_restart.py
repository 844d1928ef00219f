"""Restart files (rst_write / rst_read with ts_rst, IOM/restart.F90:133-309,
DYN/dynspg_ts.F90:1391-1463):

* a run of 8 steps equals, bit for bit, 4 steps, a restart written at
  step 4, and 4 more steps in a context registered with nit000 = 5 and a
  leapfrog start, the restarted state read into NaN-filled fields and the
  depth-mean velocities recomputed as DOM/istate.F90 does; the restarted
  state at step 8 also matches the oracle's 8-step trajectory (1e-12);
* the files round-trip every variable exactly, in device and host pointer
  modes and in the FP32 storage form (and a file of one form reads into the
  other, widened or rounded);
* a changed time step forces an Euler restart (before = now, the ts_rst
  fields zero); a missing or mismatched file and ln_bt_av = F are errors.
"""
import json
import os
from dataclasses import replace

import numpy as np
import pytest

import nemo_hip
from nemo_hip import NEMO_HIP_PTR_DEVICE, NEMO_HIP_PTR_HOST, configs
from test_gpu_parity import check
from test_gpu_trajectory import CASES
from trajectory import OUT_2D, OUT_3D, OUT_4D, gpu_run, initial, oracle_run

torch = pytest.importorskip("torch")

RST_3D = ["ub", "vb", "un", "vn", "rhop"]
RST_4D = ["tsb", "tsn"]
RST_2D = ["sshb", "sshn", "ub2_b", "vb2_b", "un_bf", "vn_bf"]


def restarted(d, **opts):
    d2 = replace(d, opts=dict(d.opts, **opts))
    d2.e3t_1d, d2.e3w_1d = d.e3t_1d, d.e3w_1d
    return d2


def nan_like(t):
    return torch.full_like(t, float("nan"))


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["gyre", "cyc7", "basin_een"])
def test_restart_continues_bit_for_bit(hip, case, tmp_path):
    d = CASES[case]()
    S = configs.make_state(d, seed=17)
    hip.register(d)
    full = gpu_run(hip, d, S, (8,))
    hip.register(d)
    X = {}
    gpu_run(hip, d, S, (4,), keep=X)
    rhop = torch.zeros_like(X["un"])
    path = tmp_path / "rst_0004"
    hip.rst_write(path, 4, rhop=rhop, **{k: X[k] for k in RST_3D + RST_4D + RST_2D if k != "rhop"})
    man = json.load(open(path / "manifest.json"))
    assert man["kt"] == 4 and man["rdt"] == d.opts["rdt"] and [man["jpi"], man["jpj"], man["jpk"]] == [d.jpi, d.jpj,
                                                                                                        d.jpk]
    assert sorted(v["name"] for v in man["variables"]) == sorted(
        ["ub", "vb", "tb", "sb", "sshb", "un", "vn", "tn", "sn", "sshn", "rhop", "ub2_b", "vb2_b", "un_bf", "vn_bf"])
    # a fresh run restarted from the file at step 5
    d2 = restarted(d, nit000=5, neuler=1)
    hip.register(d2)
    dev = lambda a, t: torch.from_numpy(np.ascontiguousarray(np.asarray(a).ravel(order="F"))).to("cuda", t)
    Y = {k: dev(v, torch.float64) for k, v in initial(S).items()}
    for k in RST_3D[:-1] + RST_4D + RST_2D + ["un_b", "vn_b", "ub_b", "vb_b"]:
        Y[k] = nan_like(Y[k])
    rhop2 = nan_like(rhop)
    kt, ne = hip.rst_read(path, rhop=rhop2, r1_hu_n=Y["r1_hu_n"], r1_hv_n=Y["r1_hv_n"],
                          **{k: Y[k] for k in RST_3D[:-1] + RST_4D + RST_2D + ["un_b", "vn_b", "ub_b", "vb_b"]})
    assert (kt, ne) == (4, 1)
    rest = gpu_run(hip, d2, S, (8,), kt0=5, X0=Y)
    ident = True
    for k in OUT_3D + OUT_4D + OUT_2D:
        check(rest[8][k], full[8][k], what=f"{case} restarted step 8 {k}")
        ident &= bool(np.array_equal(rest[8][k], full[8][k]))
    assert ident, "the restarted run should equal the continuous run bit for bit"
    want = oracle_run(d, S, (8,))[8]
    for k in OUT_3D + OUT_4D + OUT_2D:
        check(rest[8][k], want[k], what=f"{case} restarted step 8 {k} against the oracle's trajectory")


@pytest.mark.gpu
@pytest.mark.parametrize("sp", [False, True])
def test_restart_round_trip_and_euler_restart(hip, sp, tmp_path):
    d = CASES["cyc7"]()
    hip.register(d)
    rng = np.random.default_rng(5)
    n2, n3 = d.jpi * d.jpj, d.jpi * d.jpj * d.jpk
    t3 = torch.float32 if sp else torch.float64
    F = {k: torch.from_numpy(rng.standard_normal(n3)).to("cuda", t3) for k in RST_3D}
    F.update({k: torch.from_numpy(rng.standard_normal(2 * n3)).to("cuda", t3) for k in RST_4D})
    F.update({k: torch.from_numpy(rng.standard_normal(n2)).to("cuda", torch.float64) for k in RST_2D})
    path = tmp_path / "rt"
    hip.rst_write(path, 12, sp=sp, **F)
    G = {k: nan_like(v) for k, v in F.items()}
    assert hip.rst_read(path, sp=sp, **G) == (12, 1)
    for k in F:
        assert torch.equal(G[k], F[k]), k
    # the other storage form reads the same file: widened (f4 -> f8) or rounded (f8 -> f4)
    t3o = torch.float64 if sp else torch.float32
    H = {k: torch.full(v.shape, float("nan"), dtype=t3o if k in RST_3D + RST_4D else torch.float64, device="cuda")
         for k, v in F.items()}
    assert hip.rst_read(path, sp=not sp, **H) == (12, 1)
    for k in RST_3D + RST_4D:
        assert torch.equal(H[k], F[k].to(t3o)), k
    # host pointer mode reads the same bits
    hp = {k: np.full(v.numel(), np.nan, dtype=np.float32 if (sp and k in RST_3D + RST_4D) else np.float64)
          for k, v in F.items()}
    hip.set_pointer_mode(NEMO_HIP_PTR_HOST)
    try:
        assert hip.rst_read(path, sp=sp, **hp) == (12, 1)
    finally:
        hip.set_pointer_mode(NEMO_HIP_PTR_DEVICE)
    for k in F:
        assert np.array_equal(hp[k], F[k].cpu().numpy()), k
    # a changed time step: Euler restart, before = now, ts_rst fields zero
    hip.register(restarted(d, rdt=d.opts["rdt"] / 2))
    E = {k: nan_like(v) for k, v in F.items()}
    assert hip.rst_read(path, sp=sp, **E) == (12, 0)
    assert torch.equal(E["ub"], F["un"]) and torch.equal(E["vb"], F["vn"]) and torch.equal(E["tsb"], F["tsn"])
    assert torch.equal(E["sshb"], F["sshn"])
    for k in ("ub2_b", "vb2_b", "un_bf", "vn_bf"):
        assert torch.count_nonzero(E[k]) == 0, k


@pytest.mark.gpu
def test_restart_errors(hip, tmp_path):
    d = CASES["gyre"]()
    hip.register(d)
    n2, n3 = d.jpi * d.jpj, d.jpi * d.jpj * d.jpk
    z = lambda n: torch.zeros(n, dtype=torch.float64, device="cuda")
    F = {k: z(n3) for k in RST_3D}
    F.update({k: z(2 * n3) for k in RST_4D})
    F.update({k: z(n2) for k in RST_2D})
    with pytest.raises(nemo_hip.NemoHipError, match="manifest"):
        hip.rst_read(tmp_path / "none", **F)
    hip.rst_write(tmp_path / "a", 3, **F)
    os.remove(tmp_path / "a" / "sn.bin")
    with pytest.raises(nemo_hip.NemoHipError, match="sn.bin"):
        hip.rst_read(tmp_path / "a", **F)
    with pytest.raises(nemo_hip.NemoHipError, match="NULL"):
        hip.rst_write(tmp_path / "b", 3, **{k: v for k, v in F.items() if k != "rhop"})
    hip.rst_write(tmp_path / "c", 3, **F)
    hip.register(CASES["cyc7"]())      # the shape check comes before any field is touched
    with pytest.raises(nemo_hip.NemoHipError, match="registered"):
        hip.rst_read(tmp_path / "c", **F)
    hip.register(restarted(d, ln_bt_av=0))
    with pytest.raises(nemo_hip.NemoHipError, match="ln_bt_av"):
        hip.rst_write(tmp_path / "d", 3, **F)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["device", "host"])
def test_restart_solar_flux_variables(hip, mode, tmp_path):
    """tra_qsr's restart variables (traqsr.F90:287-292, 136-139, 424-425):
    rst_write stores qsr_hc under the name qsr_hc_b and fraqsr_1lev;
    rst_read reads them into qsr_hc_b and fraqsr_1lev and reports
    kqsr_b = 1; a file without them leaves both untouched with kqsr_b = 0"""
    d = CASES["cyc7"]()
    hip.register(d)
    rng = np.random.default_rng(6)
    n2, n3 = d.jpi * d.jpj, d.jpi * d.jpj * d.jpk
    F = {k: rng.standard_normal(n3) for k in RST_3D}
    F.update({k: rng.standard_normal(2 * n3) for k in RST_4D})
    F.update({k: rng.standard_normal(n2) for k in RST_2D})
    hc, fr = rng.standard_normal(n3), rng.uniform(0, 1, n2)
    if mode == "device":
        to = lambda a: torch.from_numpy(a).to("cuda")
        back = lambda t: t.cpu().numpy()
    else:
        to = lambda a: a.copy()
        back = lambda a: a
        hip.set_pointer_mode(NEMO_HIP_PTR_HOST)
    try:
        G = {k: to(v) for k, v in F.items()}
        hip.rst_write(tmp_path / "q", 7, qsr_hc=to(hc), fraqsr_1lev=to(fr), **G)
        man = json.load(open(tmp_path / "q" / "manifest.json"))
        assert {"qsr_hc_b", "fraqsr_1lev"} <= {v["name"] for v in man["variables"]}
        hb, fr2 = to(np.full(n3, np.nan)), to(np.full(n2, np.nan))
        H = {k: to(np.full_like(v, np.nan)) for k, v in F.items()}
        assert hip.rst_read(tmp_path / "q", qsr_hc_b=hb, fraqsr_1lev=fr2, **H) == (7, 1)
        assert hip.kqsr_b == 1
        assert np.array_equal(back(hb), hc) and np.array_equal(back(fr2), fr)
        # a restart without them: an Euler start of tra_qsr
        hip.rst_write(tmp_path / "p", 7, **G)
        hb2, fr3 = to(np.full(n3, 3.0)), to(np.full(n2, 4.0))
        assert hip.rst_read(tmp_path / "p", qsr_hc_b=hb2, fraqsr_1lev=fr3, **H) == (7, 1)
        assert hip.kqsr_b == 0
        assert np.all(back(hb2) == 3.0) and np.all(back(fr3) == 4.0)
    finally:
        hip.set_pointer_mode(NEMO_HIP_PTR_DEVICE)
