"""Long and deep trajectories on the GPU against the oracle.

* The 32x22x31 GYRE-like basin (rdt = 1800 s, nn_baro = 30) for 100 whole
  FP64 time steps: every prognostic field within 1e-9 of the oracle's
  trajectory (bit-identical expected; the bar is the plain form's 100-step
  one), at steps 1, 10 and 100.
* 75-level basins with land, FCT 4/4, iso-neutral diffusion and the whole
  step (closed and doubly cyclic), 4 steps: bit-identical to the oracle.
"""
import json

import numpy as np
import pytest

from nemo_hip import configs
from test_gpu_parity import check
from trajectory import OUT_2D, OUT_3D, OUT_4D, gpu_run, oracle_run

torch = pytest.importorskip("torch")


@pytest.mark.gpu
def test_gyre_100_steps_fp64(hip):
    d = configs.gyre_like()
    S = configs.make_state(d, seed=23)
    marks = (1, 10, 100)
    want = oracle_run(d, S, marks)
    hip.register(d)
    got = gpu_run(hip, d, S, marks)
    worst, ident = 0.0, True
    for kt in marks:
        for k in OUT_3D + OUT_4D + OUT_2D:
            worst = max(worst, check(got[kt][k], want[kt][k], tol=1e-9, what=f"gyre step {kt} {k}"))
            ident &= bool(np.array_equal(got[kt][k], want[kt][k]))
    print(json.dumps({"steps": 100, "max_relerr": worst, "bit_identical": ident}))


@pytest.mark.gpu
@pytest.mark.parametrize("jperio", [0, 7])
def test_75_levels(hip, jperio):
    d = configs.basin(64, 48, 75, seed=31, jperio=jperio)
    S = configs.make_state(d, seed=32)
    marks = (1, 2, 4)
    want = oracle_run(d, S, marks)
    hip.register(d)
    got = gpu_run(hip, d, S, marks)
    for kt in marks:
        for k in OUT_3D + OUT_4D + OUT_2D:
            g, w = got[kt][k], want[kt][k]
            assert np.array_equal(g, w), f"jperio={jperio} step {kt} {k}: max |diff| {np.abs(g - w).max():.3e}"
