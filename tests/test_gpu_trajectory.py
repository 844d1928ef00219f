"""Whole time steps on the GPU against the oracle (tests/trajectory.py): the
hot path closed by ssh_nxt, wzv, div_hor, tra_nxt, dyn_nxt and ssh_swp, so the
state advances on the device from step to step.

FP64: every prognostic field after steps 1 (Euler), 2 (first leapfrog) and 8
within 1e-12 max relative error of the oracle's trajectory (bit-identical
expected).  FP32 storage form: the same trajectory within the tolerances
below after 8 steps.
"""
import json

import numpy as np
import pytest

from nemo_hip import abi, configs
from test_gpu_parity import check
from trajectory import OUT_2D, OUT_3D, OUT_4D, gpu_run, oracle_run

torch = pytest.importorskip("torch")

CASES = {
    "gyre": lambda: configs.gyre_like(),
    "cyc7": lambda: configs.basin(40, 30, 12, seed=4, jperio=7),
    "basin_een": lambda: configs.basin(40, 30, 12, seed=3, opts={"nvor_scheme": 4}),
}
MARKS = (1, 2, 8)


@pytest.mark.gpu
@pytest.mark.parametrize("case", list(CASES))
def test_trajectory_fp64(hip, case):
    d = CASES[case]()
    S = configs.make_state(d, seed=17)
    want = oracle_run(d, S, MARKS)
    hip.register(d)
    got = gpu_run(hip, d, S, MARKS)
    ident = True
    for kt in MARKS:
        for k in OUT_3D + OUT_4D + OUT_2D:
            check(got[kt][k], want[kt][k], what=f"{case} step {kt} {k}")
            ident &= bool(np.array_equal(got[kt][k], want[kt][k]))
    print(json.dumps({"case": case, "bit_identical": ident}))


@pytest.mark.gpu
@pytest.mark.parametrize("case,eos", [("gyre", "seos"), ("cyc7", "seos"), ("cyc7", "teos10")])
def test_trajectory_with_density_and_viscosity_fp64(hip, case, eos):
    """The step with the iso-neutral slopes of the before state (rab_3d, bn2,
    zdf_mxl, eos_insitu, ldf_slp), the density and its pressure gradient,
    dyn_ldf_lap and tra_sbc added."""
    d = CASES[case]()
    S = configs.make_state(d, seed=19)
    mk = abi.seos if eos == "seos" else abi.teos10
    want = oracle_run(d, S, MARKS, physics=True, eos=mk())
    hip.register(d)
    got = gpu_run(hip, d, S, MARKS, physics=True, eos=mk())
    ident = True
    for kt in MARKS:
        for k in OUT_3D + OUT_4D + OUT_2D:
            check(got[kt][k], want[kt][k], what=f"{case} {eos} physics step {kt} {k}")
            ident &= bool(np.array_equal(got[kt][k], want[kt][k]))
    print(json.dumps({"case": case, "eos": eos, "physics": True, "bit_identical": ident}))


# max relative error after 8 steps, FP32 storage against the FP64 oracle
# (measured on MI355X: tracers 4.4e-7, velocities 2.2e-7, w 4.8e-7, sea level
# 2.6e-7, barotropic velocities 1.3e-7, un_bf 1.1e-6)
SP_TOL = {"tsn": 5e-6, "tsb": 5e-6, "un": 3e-6, "vn": 3e-6, "ub": 3e-6, "vb": 3e-6, "wn": 5e-6,
          "sshn": 3e-6, "sshb": 3e-6, "un_b": 2e-6, "vn_b": 2e-6, "ub_b": 2e-6, "vb_b": 2e-6,
          "ub2_b": 2e-6, "un_bf": 1e-5}


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["gyre", "cyc7"])
def test_trajectory_fp32(hip, case):
    d = CASES[case]()
    S = configs.make_state(d, seed=17)
    want = oracle_run(d, S, (8,))
    hip.register(d)
    got = gpu_run(hip, d, S, (8,), sp=True)
    errs = {}
    for k, tol in SP_TOL.items():
        a, b = got[8][k], want[8][k]
        assert np.all(np.isfinite(a)), k
        errs[k] = float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
    print(json.dumps({"case": case, "maxrel": errs}))
    for k, tol in SP_TOL.items():
        assert errs[k] <= tol, (k, errs[k])
