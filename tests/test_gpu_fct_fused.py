"""The fused FCT kernel (tra_adv_fct on closed domains in one k-march per
tile and tracer) against the oracle and against the flux / limiter chain.

* FP64 drop-in tra_adv_fct with either workgroup size: bit-identical to
  the oracle for the four order pairs on basins whose sizes leave partial tiles in i and j, on the
  GYRE-like basin, on a 6-level, a 75-level and a 90-level basin, with 1, 2
  and 3 tracers (a pair and a single-tracer remainder);
* both storage forms: the fused kernel, with 256 or 512 threads per tile
  (tuning switch fct_fused = 1, 2), equals the chain (fct_fused = 0) bit
  for bit, for the drop-in routine (trend read) and for the fused step
  (trend from zero, transports formed where used); so do the chain's
  forms, its vertical half as the kernels k_fct_v, k_fct_lim (fct_vlim =
  0) or in one bottom-up march with 256 or 512 threads per tile
  (fct_vlim = 1, 2).
"""
import numpy as np
import pytest

from nemo_hip import configs
from oracle.oracle import Oracle
from test_gpu_parity import check, dev, host, sync

torch = pytest.importorskip("torch")

BASINS = {
    "gyre": lambda: configs.gyre_like(),
    "odd": lambda: configs.basin(131, 37, 14, seed=21),
    "deep": lambda: configs.basin(70, 45, 75, seed=22),
    "tiny": lambda: configs.basin(9, 7, 6, seed=23),
    # deeper than the FP32 register-held column form (76 levels): the FP32
    # chain takes its per-cell / per-column kernels
    "deeper": lambda: configs.basin(40, 33, 90, seed=24),
}
ORDERS = [(2, 2), (4, 4), (2, 4), (4, 2)]


def setup(hip, name, kjpt=2, seed=5):
    d = BASINS[name]()
    S = configs.make_state(d, seed=seed)
    hip.register(d)
    o = Oracle(d)
    s3 = (d.jpi, d.jpj, d.jpk)
    z = [np.zeros(s3, order="F") for _ in range(3)]
    o.tra_adv_trp(S["un"], S["vn"], S["wn"], *z)
    rng = np.random.default_rng(seed)
    if kjpt == 2:
        tb, tn, ta = S["tsb"], S["tsn"], S["tsa"]
    else:
        tb = np.asfortranarray(rng.uniform(0, 10, s3 + (kjpt,)) * d.tmask()[..., None])
        tn = np.asfortranarray(tb + rng.normal(0, 0.1, tb.shape) * d.tmask()[..., None])
        ta = np.asfortranarray(rng.normal(0, 1e-6, tb.shape))
    return d, S, o, z, tb, tn, ta


def with_switches(hip, sw, fn):
    defaults = {k: hip.tune_get(k) for k in sw}
    try:
        for k, v in sw.items():
            hip.tune_set(k, v)
        return fn()
    finally:
        for k, v in defaults.items():
            hip.tune_set(k, v)


def with_mode(hip, mode, fn):
    return with_switches(hip, {"fct_fused": mode}, fn)


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(BASINS))
@pytest.mark.parametrize("orders", ORDERS)
@pytest.mark.parametrize("mode", [1, 2])
def test_fused_fct_fp64_oracle(hip, name, orders, mode):
    d, S, o, z, tb, tn, ta = setup(hip, name)
    p2dt = 2.0 * d.opts["rdt"]
    want = ta.copy(order="F")
    o.tra_adv_fct(3, 1, "TRA", p2dt, *z, tb, tn, want, 2, *orders)

    def run():
        got = dev(ta)
        hip.tra_adv_fct(3, 1, "TRA", p2dt, *[dev(a) for a in z], dev(tb), dev(tn), got, 2, *orders)
        sync(hip)
        return host(got, want.shape)

    g = with_mode(hip, mode, run)
    assert np.array_equal(g, want), f"{name} {orders} mode {mode}: max |diff| {np.abs(g - want).max():.3e}"


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(BASINS))
@pytest.mark.parametrize("orders", ORDERS)
@pytest.mark.parametrize("vlim", [1, 2])
def test_chain_march_fp64_oracle(hip, name, orders, vlim):
    """the chain with its vertical half in one bottom-up march (fct_vlim 1,
    2), FP64 drop-in tra_adv_fct, against the oracle bit for bit"""
    d, S, o, z, tb, tn, ta = setup(hip, name)
    p2dt = 2.0 * d.opts["rdt"]
    want = ta.copy(order="F")
    o.tra_adv_fct(3, 1, "TRA", p2dt, *z, tb, tn, want, 2, *orders)

    def run():
        got = dev(ta)
        hip.tra_adv_fct(3, 1, "TRA", p2dt, *[dev(a) for a in z], dev(tb), dev(tn), got, 2, *orders)
        sync(hip)
        return host(got, want.shape)

    g = with_switches(hip, {"fct_fused": 0, "fct_vlim": vlim}, run)
    assert np.array_equal(g, want), f"{name} {orders} vlim {vlim}: max |diff| {np.abs(g - want).max():.3e}"


@pytest.mark.gpu
@pytest.mark.parametrize("kjpt", [1, 3])
def test_fused_fct_tracer_counts(hip, kjpt):
    d, S, o, z, tb, tn, ta = setup(hip, "odd", kjpt=kjpt)
    want = ta.copy(order="F")
    o.tra_adv_fct(3, 1, "TRC", 1200.0, *z, tb, tn, want, kjpt, 4, 4)
    got = dev(ta)
    hip.tra_adv_fct(3, 1, "TRC", 1200.0, *[dev(a) for a in z], dev(tb), dev(tn), got, kjpt, 4, 4)
    sync(hip)
    g = host(got, want.shape)
    assert np.array_equal(g, want), f"kjpt={kjpt}: max |diff| {np.abs(g - want).max():.3e}"


def both_paths(hip, fn):
    """fn() with the fused kernel (both workgroup sizes, which must agree),
    then with the chain in its two forms (which must agree); returns the
    fused and the chain results"""
    a = with_mode(hip, 1, fn)
    a2 = with_mode(hip, 2, fn)
    assert np.array_equal(a, a2, equal_nan=True), "the fused kernel's two workgroup sizes differ"
    b = with_switches(hip, {"fct_fused": 0, "fct_vlim": 0}, fn)
    for v in (1, 2):
        bv = with_switches(hip, {"fct_fused": 0, "fct_vlim": v}, fn)
        assert np.array_equal(b, bv, equal_nan=True), \
            f"the chain's march form {v} differs: max |diff| {np.nanmax(np.abs(b - bv)):.3e}"
    return a, b


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["gyre", "odd", "deep", "tiny", "deeper"])
@pytest.mark.parametrize("sp", [False, True])
@pytest.mark.parametrize("orders", [(4, 4), (2, 2)])
def test_fused_equals_chain(hip, name, sp, orders):
    d, S, o, z, tb, tn, ta = setup(hip, name)
    t = torch.float32 if sp else torch.float64
    zz = [dev(a, t) for a in z]
    gb, gn = dev(tb, t), dev(tn, t)

    def dropin():
        out = dev(ta, t)
        hip.tra_adv_fct(3, 1, "TRA", 1200.0, *zz, gb, gn, out, 2, *orders, sp=sp)
        sync(hip)
        return out.cpu().numpy()

    a, b = both_paths(hip, dropin)
    assert np.array_equal(a, b), f"drop-in {name} sp={sp} {orders}: max |diff| {np.abs(a - b).max():.3e}"
    f = {n: dev(S[n], t) for n in ("un", "vn", "wn", "ahtu", "ahtv", "uslp", "vslp", "wslpi", "wslpj", "avt")}
    n3 = d.jpi * d.jpj * d.jpk

    def step():
        tsa = torch.full((2 * n3,), float("nan"), dtype=t, device="cuda")
        ah, akz = torch.zeros(n3, dtype=t, device="cuda"), torch.zeros(n3, dtype=t, device="cuda")
        hip.step_tra(3, f["un"], f["vn"], f["wn"], gb, gn, tsa, *orders, f["ahtu"], f["ahtv"], f["uslp"],
                     f["vslp"], f["wslpi"], f["wslpj"], ah, akz, f["avt"], sp=sp)
        sync(hip)
        return tsa.cpu().numpy()

    a, b = both_paths(hip, step)
    assert np.all(np.isfinite(a))
    assert np.array_equal(a, b), f"step {name} sp={sp} {orders}: max |diff| {np.nanmax(np.abs(a - b)):.3e}"
