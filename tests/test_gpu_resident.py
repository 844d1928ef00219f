"""Resident fields (nemo_hip_field_alloc / push / pull / free): the device
mirrors a resident-mode caller keeps in HBM between calls.  A routine run on
pushed fields must give the oracle's result, bit for bit, after a pull, and
the calls must reject addresses and sizes they do not own."""
import numpy as np
import pytest

import nemo_hip
from test_gpu_parity import Env, check

torch = pytest.importorskip("torch")


@pytest.mark.gpu
def test_resident_fields_run_a_routine(hip):
    e = Env(hip, "basin")
    S = e.S
    want = e.c("tsa")
    e.o.tra_ldf_lap(1, 1, "TRA", S["ahtu"], S["ahtv"], S["tsb"], want, 2, 1)
    f = {k: hip.field_alloc(S[k].nbytes) for k in ("ahtu", "ahtv", "tsb", "tsa")}
    try:
        for k, v in f.items():
            hip.field_push(v, S[k])
        hip.tra_ldf_lap(1, 1, "TRA", f["ahtu"], f["ahtv"], f["tsb"], f["tsa"], 2, 1)
        got = np.zeros(e.s4, order="F")
        hip.field_pull(got, f["tsa"])
        check(got, want, what="tra_ldf_lap on resident fields")
        back = np.zeros(e.s4, order="F")
        hip.field_pull(back, f["tsb"])
        assert np.array_equal(back, S["tsb"])
    finally:
        for v in f.values():
            hip.field_free(v)


@pytest.mark.gpu
def test_resident_fields_reject_foreign_and_oversized(hip):
    f = hip.field_alloc(4096)
    a = np.arange(512, dtype=np.float64)
    try:
        hip.field_push(f, a)                       # exactly the field's size
        with pytest.raises(nemo_hip.NemoHipError, match="exceed"):
            hip.field_push(f, np.zeros(513))       # one element too many
        with pytest.raises(nemo_hip.NemoHipError, match="exceed"):
            hip.field_pull(np.zeros(513), f)
        inner = nemo_hip.DeviceField(f.ptr + 8, 4088)
        out = np.zeros(511)
        hip.field_pull(out, inner)                 # a sub-range of a field is accepted
        assert np.array_equal(out, a[1:])
        t = torch.zeros(8, dtype=torch.float64, device="cuda")
        with pytest.raises(nemo_hip.NemoHipError, match="not a field"):
            hip.field_push(nemo_hip.DeviceField(t.data_ptr(), 64), np.zeros(8))
        with pytest.raises(nemo_hip.NemoHipError, match="not returned"):
            hip.field_free(t.data_ptr())
    finally:
        hip.field_free(f)
    with pytest.raises(nemo_hip.NemoHipError, match="not returned"):
        hip.field_free(f)                          # twice
