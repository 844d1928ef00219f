"""nemo_hip_glob_sum_noalloc, the glob_sum that allocates nothing (its
partial sums in a static device array), against nemo_hip_glob_sum on the
same inputs: two calls in a row, the second on a smaller input (a smaller
registered domain, fewer levels, the other mask), each bit-identical to the
allocating entry point and to math.fsum; host-pointer mode is refused."""
import math

import numpy as np
import pytest

import nemo_hip
from nemo_hip import NEMO_HIP_PTR_DEVICE, NEMO_HIP_PTR_HOST, configs

torch = pytest.importorskip("torch")


def masks(d):
    h = np.zeros((d.jpi, d.jpj))
    h[1:-1, 1:-1] = 1.0
    return h * (d.arrays["k_bot"] > 0), h


@pytest.mark.gpu
def test_glob_sum_noalloc_matches_glob_sum(hip):
    rng = np.random.default_rng(41)
    for (jpi, jpj, jpk), nk, kind in (((300, 260, 40), 40, 0), ((90, 70, 12), 5, 1)):
        d = configs.basin(jpi, jpj, jpk, seed=jpi, land_frac=0.3)
        hip.register(d)
        a = np.asfortranarray(rng.standard_normal((jpi, jpj, jpk)) * 10.0 ** rng.uniform(-8, 8, (jpi, jpj, jpk)))
        ga = torch.from_numpy(a.ravel(order="F").copy()).cuda()
        o1 = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
        o2 = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
        hip.glob_sum_noalloc(ga, nk, kind, o1)
        hip.glob_sum(ga, nk, kind, out=o2)
        torch.cuda.synchronize()
        hip.sync()
        got, ref = float(o1.cpu()[0]), float(o2.cpu()[0])
        m = masks(d)[kind]
        assert got == ref == math.fsum((a[:, :, :nk] * m[:, :, None]).ravel(order="F")), (jpi, got, ref)
    hip.set_pointer_mode(NEMO_HIP_PTR_HOST)
    try:
        with pytest.raises(nemo_hip.NemoHipError, match="device-pointer mode"):
            hip.glob_sum_noalloc(np.zeros(d.jpi * d.jpj * d.jpk), 1, 0, np.zeros(1))
    finally:
        hip.set_pointer_mode(NEMO_HIP_PTR_DEVICE)
