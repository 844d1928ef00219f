"""Precision form: FP32 storage against the FP64 oracle along 100 whole time
steps (hot path and closure) of the 32x22x31 GYRE-like basin
(tests/precision_study.py).

Tolerance: max relative error after 100 steps of 5e-6 for the tracers and
3e-6 for u, v and the sea level (measured on MI355X: T 5.1e-7, S 1.8e-6,
u 4.8e-7, v 4.9e-7, ssh 2.6e-7; the barotropic sub-stepping is FP64 in both
forms).
"""
import json

import numpy as np
import pytest

from precision_study import errors


@pytest.mark.gpu
def test_precision_form_error_growth(hip):
    res = errors((1, 10, 100))
    print(json.dumps(res))
    for kt, e in res.items():
        assert all(np.isfinite(v) for v in e.values()), f"step {kt}: {e}"
    assert res[100]["T"] <= 5e-6 and res[100]["S"] <= 5e-6, res[100]
    assert max(res[100]["u"], res[100]["v"], res[100]["ssh"]) <= 3e-6, res[100]
