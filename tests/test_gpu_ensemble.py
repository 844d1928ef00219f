"""An ensemble of domains of different sizes (one library context each),
stepped by the fused step entry points, recorded once into a HIP graph and
replayed:

* every replay gives, bit for bit, what the same calls give when issued one
  by one (FP32 storage); the replays are interleaved with eager work and
  synchronisations, and the recording refuses the calls it cannot replay
  faithfully;
* a plan (nemo_hip_plan_*) of the same step over closed members and a
  cyclic one: every position of the closed members' launch sequence is one
  batched launch over all their workgroups, and plan launches, eager or
  recorded into a graph, equal the calls issued one by one, bit for bit;
* 24 members of uneven sizes (56 levels, closed and cyclic, each its own
  seeded land mask, bathymetry and initial state) advanced by whole time
  steps (the hot path and its closure), the step recorded once for all
  members into a graph, of the calls or of a plan of them: after 8 steps
  each member equals the oracle's trajectory of the same member, bit for
  bit in FP64 and within the FP32 storage form's tolerances.
"""
import json

import numpy as np
import pytest

import nemo_hip
from nemo_hip import abi, configs

torch = pytest.importorskip("torch")

# (jpi, jpj, jpk, jperio, seed): closed and cyclic members of different shapes
MEMBERS = [(40, 30, 12, 0, 3), (36, 28, 10, 7, 5), (52, 20, 14, 0, 6)]
OPTS = dict(nvor_scheme=abi.NP_ENE, n_dynadv=1, nn_dynkeg=0, ln_traldf_msc=0, nn_baro=30, rdt=600.0, neuler=0,
            nit000=1)


def _stepper(h, d, seed):
    from bench import Stepper
    S = configs.make_state_device(d, seed=seed)
    return Stepper(h, d, S, True, True)


def _state(st):
    out = {"tsa": st.tsa, "ua": st.ua, "va": st.va, "un": st.un_dyn, "vn": st.vn_dyn}
    for k in ("ua_b", "va_b", "ssha", "un_adv", "vn_adv", "ub2_b", "vb2_b", "un_bf", "vn_bf"):
        out[k] = st.f2[k]
    return {k: v.detach().cpu().numpy().copy() for k, v in out.items()}


@pytest.mark.gpu
def test_ensemble_graph_replays_match_eager_calls(hip):
    ctxs, eager, graph = [], [], []
    g = None
    try:
        for (jpi, jpj, jpk, jperio, seed) in MEMBERS:
            c = hip.context_create()
            ctxs.append(c)
            hip.context_select(c)
            d = configs.basin(jpi, jpj, jpk, seed=seed, jperio=jperio, opts=dict(OPTS))
            hip.register(d)
            eager.append(_stepper(hip, d, seed + 100))
            graph.append(_stepper(hip, d, seed + 100))
        # steps 1 and 2 (Euler start, switch to leapfrog) are issued eagerly
        for kt in (1, 2):
            for c, a, b in zip(ctxs, eager, graph):
                hip.context_select(c)
                a.step(kt)
                b.step(kt)
        hip.sync()
        # what recording refuses
        hip.graph_begin()
        hip.context_select(ctxs[0])
        with pytest.raises(nemo_hip.NemoHipError, match="first two steps"):
            graph[0].step(2)
        with pytest.raises(nemo_hip.NemoHipError, match="recording"):
            hip.sync()
        hip.graph_free(hip.graph_end())
        # one ensemble step recorded, then replayed for kt = 3, 4, 5
        hip.graph_begin()
        for c, b in zip(ctxs, graph):
            hip.context_select(c)
            b.step(3)
        g = hip.graph_end()
        for kt in (3, 4, 5):
            for c, a in zip(ctxs, eager):
                hip.context_select(c)
                a.step(kt)
            hip.graph_launch(g)
            hip.sync()
            torch.cuda.synchronize()
            for m, (a, b) in enumerate(zip(eager, graph)):
                sa, sb = _state(a), _state(b)
                for k in sa:
                    assert np.all(np.isfinite(sa[k])), f"member {m} {k} kt {kt}: non-finite"
                    assert np.array_equal(sa[k], sb[k]), f"member {m} field {k} differs after replay at kt {kt}"
    finally:
        if g is not None:
            hip.graph_free(g)
        hip.context_select(0)
        del eager, graph
        for c in ctxs:
            hip.context_destroy(c)


@pytest.mark.gpu
@pytest.mark.parametrize("vlim", [0, 1])
def test_ensemble_plan_replays_match_eager_calls(hip, vlim):
    """a plan of the fused FP32 step over closed and cyclic members: each
    position of the closed members' launch sequence is one batched launch,
    the cyclic member forms a group of its own; plan launches, eager or
    recorded into a graph, equal the calls issued one by one (with the FCT
    chain's vertical half as kernels or as one march, switch fct_vlim)"""
    members = MEMBERS + [(44, 26, 9, 0, 8)]
    ctxs, eager, batched, plan, g = [], [], [], None, None
    vlim0 = hip.tune_get("fct_vlim")
    hip.tune_set("fct_vlim", vlim)
    try:
        for (jpi, jpj, jpk, jperio, seed) in members:
            c = hip.context_create()
            ctxs.append(c)
            hip.context_select(c)
            d = configs.basin(jpi, jpj, jpk, seed=seed, jperio=jperio, opts=dict(OPTS))
            hip.register(d)
            eager.append(_stepper(hip, d, seed + 100))
            batched.append(_stepper(hip, d, seed + 100))
        for kt in (1, 2):
            for c, a, b in zip(ctxs, eager, batched):
                hip.context_select(c)
                a.step(kt)
                b.step(kt)
        hip.sync()
        # what tracing refuses, and an empty plan
        hip.plan_begin()
        hip.context_select(ctxs[0])
        with pytest.raises(nemo_hip.NemoHipError, match="first two steps"):
            batched[0].step(2)
        with pytest.raises(nemo_hip.NemoHipError, match="no launch"):
            hip.plan_end()
        hip.plan_begin()
        for c, b in zip(ctxs, batched):
            hip.context_select(c)
            b.step(3)
        plan = hip.plan_end()
        info = hip.plan_info(plan)
        assert info["members"] == len(members) and info["groups"] == 2, info
        assert info["batched"] > 0 and info["launches"] < info["ops"] * len(members), info
        for kt in (3, 4, 5, 6):
            for c, a in zip(ctxs, eager):
                hip.context_select(c)
                a.step(kt)
            if kt == 5:       # from here on through a graph of the plan
                hip.graph_begin()
                hip.plan_launch(plan)
                g = hip.graph_end()
            if kt >= 5:
                hip.graph_launch(g)
            else:
                hip.plan_launch(plan)
            hip.sync()
            torch.cuda.synchronize()
            for m, (a, b) in enumerate(zip(eager, batched)):
                sa, sb = _state(a), _state(b)
                for k in sa:
                    assert np.array_equal(sa[k], sb[k]), f"member {m} field {k} differs after the plan at kt {kt}"
        print(json.dumps({"plan": info}))
    finally:
        hip.tune_set("fct_vlim", vlim0)
        if g is not None:
            hip.graph_free(g)
        if plan is not None:
            hip.plan_free(plan)
        hip.context_select(0)
        del eager, batched
        for c in ctxs:
            hip.context_destroy(c)


# 24 members: i x j from 24 x 20 to 70 x 58, 56 levels; every third one
# east-west cyclic
RNG = np.random.default_rng(20261203)
UNEVEN = [(int(RNG.integers(24, 71)), int(RNG.integers(20, 59)), 56, 7 if m % 3 == 2 else 0, 200 + m)
          for m in range(24)]
# FP32 storage against the FP64 oracle after 8 steps (the trajectory tests'
# bars, tracers at the 1e-5 relative bar of the precision form)
SP_TOL = {"tsn": 1e-5, "tsb": 1e-5, "un": 3e-6, "vn": 3e-6, "sshn": 3e-6, "un_b": 2e-6, "vn_b": 2e-6}


@pytest.mark.gpu
@pytest.mark.parametrize("sp", [False, True], ids=["fp64", "fp32"])
@pytest.mark.parametrize("mode", ["graph", "plan"])
def test_uneven_ensemble_graph_against_oracle(hip, sp, mode):
    from trajectory import OUT_2D, OUT_3D, OUT_4D, HotStep, oracle_run
    last = 8
    ctxs, members, doms, g, plan, info = [], [], [], None, None, None
    try:
        for (jpi, jpj, jpk, jperio, seed) in UNEVEN:
            c = hip.context_create()
            ctxs.append(c)
            hip.context_select(c)
            d = configs.basin(jpi, jpj, jpk, seed=seed, jperio=jperio, opts=dict(rdt=600.0))
            hip.register(d)
            S = configs.make_state(d, seed=seed + 1)
            doms.append((d, S))
            members.append(HotStep(hip, d, S, sp))
        for kt in (1, 2):                      # Euler start and first leapfrog: eager
            for c, m in zip(ctxs, members):
                hip.context_select(c)
                m.step(kt)
        if mode == "plan":        # one batched launch per position of the members' step
            hip.plan_begin()
            for c, m in zip(ctxs, members):
                hip.context_select(c)
                m.step(3)
            plan = hip.plan_end()
            info = hip.plan_info(plan)
            assert info["members"] == len(members) and info["batched"] > 0, info
            hip.graph_begin()
            hip.plan_launch(plan)
            g = hip.graph_end()
        else:
            hip.graph_begin()
            for c, m in zip(ctxs, members):
                hip.context_select(c)
                m.step(3)
            g = hip.graph_end()
        for kt in range(3, last + 1):
            hip.graph_launch(g)
        hip.sync()
        torch.cuda.synchronize()
        worst = {}
        for i, ((d, S), m) in enumerate(zip(doms, members)):
            got = m.fields()
            want = oracle_run(d, S, (last,))[last]
            for k in OUT_3D + OUT_4D + OUT_2D:
                a, b = got[k], want[k]
                assert np.all(np.isfinite(a)), f"member {i} {k}: non-finite"
                if not sp:
                    assert np.array_equal(a, b), f"member {i} {d.jpi}x{d.jpj} {k}: max |diff| {np.abs(a - b).max():.3e}"
                elif k in SP_TOL:
                    e = float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
                    worst[k] = max(worst.get(k, 0.0), e)
                    assert e <= SP_TOL[k], f"member {i} {d.jpi}x{d.jpj} {k}: relative error {e:.2e}"
        print(json.dumps({"members": len(members), "steps": last, "sp": sp, "mode": mode, "plan": info,
                          "max_relerr": worst}))
    finally:
        if g is not None:
            hip.graph_free(g)
        if plan is not None:
            hip.plan_free(plan)
        hip.context_select(0)
        del members
        for c in ctxs:
            hip.context_destroy(c)
