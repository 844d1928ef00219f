"""Error of the FP32-storage form against the FP64 oracle along a trajectory.

Workload: the 32x22x31 GYRE-like closed basin (f = 1e-4 + 2e-11 y, flat
4500 m, tanh levels, T = 4 + 21 exp(-z/800) and S = 35 plus seeded N(0, 0.01)
noise), rdt = 1800 s with an Euler first step, dyn_spg_ts with nn_baro = 30,
FCT 4/4, iso-neutral laplacian 1000 m2/s on the generator's slopes, Kz =
1e-5 m2/s, the generator's constant wind stress and fresh-water flux.

Each step is the whole time step of tests/trajectory.py: ssh_nxt, wzv, the
hot path (ua = va = 0, dyn_adv, dyn_vor, dyn_spg_ts, div_hor + wzv, dyn_zdf;
tsa = 0, tra_adv_fct, tra_ldf_iso, tra_zdf), then tra_nxt, dyn_nxt and
ssh_swp, so velocities, w, the sea level and the tracers all evolve.

The FP64 trajectory is the oracle; the FP32 trajectory is the library's
storage form on the device.  Reported per field: max |x32 - x64| / max |x64|
over the domain after the listed steps (T and S of the now level, the now
velocities u and v, the now sea level sshn).

Run directly (needs a GPU):  python tests/precision_study.py [steps...]
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (HERE, ROOT, os.path.join(ROOT, "nemo-fmi-devel_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

UMAX, DIVERGENT = 0.2, 0.05


def setup(seed=20261202):
    from nemo_hip import configs
    d = configs.gyre_like(opts={"rdt": 1800.0, "nn_baro": 30})
    S = configs.make_state(d, seed=seed, kz_const=1e-5, aht=1000.0, umax=UMAX, divergent=DIVERGENT)
    return d, S


def fields(out):
    return {kt: {"T": o["tsn"][..., 0], "S": o["tsn"][..., 1], "u": o["un"], "v": o["vn"], "ssh": o["sshn"]}
            for kt, o in out.items()}


def errors(marks=(1, 10, 100)):
    import nemo_hip
    from trajectory import gpu_run, oracle_run
    d, S = setup()
    a = fields(oracle_run(d, S, marks))
    h = nemo_hip.get(0, nemo_hip.NEMO_HIP_PTR_DEVICE)
    h.register(d)
    b = fields(gpu_run(h, d, S, marks, sp=True))
    res = {}
    for kt in marks:
        res[kt] = {}
        for f in ("T", "S", "u", "v", "ssh"):
            ref = a[kt][f]
            res[kt][f] = float(np.abs(b[kt][f] - ref).max() / max(np.abs(ref).max(), 1e-300))
    return res


if __name__ == "__main__":
    marks = tuple(int(x) for x in sys.argv[1:]) or (1, 10, 100)
    print(json.dumps(errors(marks), indent=1))
