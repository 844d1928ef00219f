"""Precision form: FP32 storage against the FP64 oracle along a 100-step
trajectory of the 32x22x31 GYRE-like basin (tests/precision_study.py).

Tolerance: max relative error after 100 steps of 1e-5 for the tracers and
5e-6 for u, v and the sea level (measured on MI355X: T 2.3e-6, S 4.9e-6,
u 4.5e-7, v 2.9e-7, ssh 9.6e-8; the barotropic sub-stepping is FP64 in both
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
    assert res[100]["T"] <= 1e-5 and res[100]["S"] <= 1e-5, res[100]
    assert max(res[100]["u"], res[100]["v"], res[100]["ssh"]) <= 5e-6, res[100]
