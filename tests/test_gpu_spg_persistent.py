"""dyn_spg_ts's barotropic sub-steps in one cooperative launch (tuning
switch spg_persistent = 1: one 1024-thread workgroup per CU; 2: 256-thread
workgroups at full occupancy) against the two launches per sub-step
(spg_persistent = 0, the default): every output of three calls (Euler
start, leapfrog with the time-filter state) bit for bit, on closed and
cyclic basins, for every vorticity scheme's barotropic Coriolis term, in
both storage forms of the 3D fields; and against the oracle (to the
tolerance of the two-kernel loop's own parity test, 1e-12) on a 1030x600
basin, where the grid is smaller than the tile count and workgroups loop
over tiles."""
import numpy as np
import pytest

from nemo_hip import abi, configs
from test_gpu_parity import CASE_IDS, Env, check, dev, host, spg_fields, SPG_OUT, sync

torch = pytest.importorskip("torch")


def loop_launches(hip):
    """launches of the cooperative kernel since the last timing_reset"""
    return hip.timing_get("dyn_spg_ts/k_spg_loop")[1]


def run(hip, e, sp, persistent):
    hip.tune_set("spg_persistent", persistent)
    Fo = spg_fields(e)
    t = torch.float32 if sp else torch.float64
    Fd = {k: dev(v, t if k in ("ua", "va", "un", "vn") else torch.float64) for k, v in Fo.items()}
    out = []
    hip.timing_reset()
    hip.timing_enable(True)
    try:
        for kt in (1, 2, 3):
            hip.dyn_spg_ts(kt, Fd, sp=sp)
            sync(hip)
            out.append({k: v.cpu().numpy().copy() for k, v in Fd.items()})
    finally:
        hip.timing_enable(False)
    # the switch selects the path: three cooperative launches, or none
    assert loop_launches(hip) == (3 if persistent else 0)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASE_IDS)
@pytest.mark.parametrize("scheme", ["ene", "een", "ent"])
@pytest.mark.parametrize("sp", [False, True])
@pytest.mark.parametrize("mode", [1, 2])
def test_persistent_equals_two_kernel_loop(hip, case, scheme, sp, mode):
    e = Env(hip, case, opts={"nvor_scheme": abi.VOR_NAMES[scheme]})
    default = hip.tune_get("spg_persistent")
    try:
        a = run(hip, e, sp, mode)
        hip.register(e.d)       # the levels start afresh for the second run
        b = run(hip, e, sp, 0)
    finally:
        hip.tune_set("spg_persistent", default)
    for kt, (x, y) in enumerate(zip(a, b), 1):
        for k in x:
            assert np.array_equal(x[k], y[k]), f"{case} {scheme} sp={sp} kt={kt} {k}"


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [1, 2])
def test_persistent_large_grid_against_oracle(hip, mode):
    e = Env(hip, "basin", opts={"nvor_scheme": abi.NP_ENE})
    d = configs.basin(1030, 600, 4, seed=41, opts={"nvor_scheme": abi.NP_ENE})
    from oracle.oracle import Oracle
    S = configs.make_state(d, seed=42)
    o = Oracle(d)
    hip.register(d)
    e.d, e.S, e.o = d, S, o
    e.s3, e.s2 = (d.jpi, d.jpj, d.jpk), (d.jpi, d.jpj)
    Fo = spg_fields(e)
    Fd = {k: dev(v) for k, v in Fo.items()}
    shapes = {k: (e.s3 if k in ("ua", "va", "un", "vn") else e.s2) for k in Fo}
    default = hip.tune_get("spg_persistent")
    hip.tune_set("spg_persistent", mode)
    hip.timing_reset()
    hip.timing_enable(True)
    try:
        for kt in (1, 2):
            o.dyn_spg_ts(kt, Fo)
            hip.dyn_spg_ts(kt, Fd)
            sync(hip)
            for k in ["ua", "va", "un", "vn"] + SPG_OUT:
                check(host(Fd[k], shapes[k]), Fo[k], what=f"spg_ts 1030x600 mode {mode} kt={kt} {k}")
    finally:
        hip.timing_enable(False)
        hip.tune_set("spg_persistent", default)
    assert loop_launches(hip) == 2
