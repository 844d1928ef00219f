"""Autotuner of the library's kernel variants and workgroup shapes.

The library's tuning switches (nemo_hip_tune_set) choose among kernel
variants that give the same bits: the FCT as a chain of kernels or as one
k-marched kernel per tile with 256 or 512 threads per 64 x TY tile
(fct_fused 0 / 1 / 2), the barotropic sub-steps as two launches per
sub-step or one cooperative launch with 1024- or 256-thread workgroups
(spg_persistent 0 / 1 / 2), the column solves through the trend array or
on chip (zdf_reg 0 / 1), and the chain's vertical half as kernels or as one
bottom-up march with 256 or 512 threads per tile (fct_vlim 0 / 1 / 2).  Which is fastest depends on the domain (its size,
level count and storage form), so tune() times a caller's own step with each
candidate, by coordinate descent over the switches (each switch's values
with the others at their best so far, repeated until nothing changes), and
returns the fastest setting; save() / load() / apply() keep settings in a
JSON file keyed by workload, so a run applies the setting tuned for its
workload without timing again.
"""
import json
import os
import time

# every switch and its candidate values, in the order the descent visits them
SWITCHES = {
    "fct_fused": (0, 1, 2),
    "spg_persistent": (0, 1, 2),
    "zdf_reg": (0, 1),
    "fct_vlim": (0, 1, 2),
}


def _time(h, step, steps, warmup):
    for _ in range(warmup):
        step()
    h.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    h.sync()
    return (time.perf_counter() - t0) * 1e3 / steps


def tune(h, step, switches=None, steps=3, warmup=1, rounds=2, log=None):
    """Coordinate descent over the switches.  step(): issues one step of the
    workload on the library stream (the caller's own fields and time step
    index).  Returns (best setting, best ms per step, every measurement as a
    list of (setting, ms))."""
    switches = dict(SWITCHES if switches is None else switches)
    best = {k: h.tune_get(k) for k in switches}
    for k, v in best.items():
        h.tune_set(k, v)
    best_ms = _time(h, step, steps, warmup)
    trials = [(dict(best), best_ms)]
    if log:
        log(f"autotune: start {best} {best_ms:.3f} ms")
    for _ in range(rounds):
        changed = False
        for k, values in switches.items():
            for v in values:
                if v == best[k]:
                    continue
                cand = dict(best, **{k: v})
                for kk, vv in cand.items():
                    h.tune_set(kk, vv)
                ms = _time(h, step, steps, warmup)
                trials.append((cand, ms))
                if log:
                    log(f"autotune: {cand} {ms:.3f} ms")
                if ms < best_ms:
                    best, best_ms, changed = cand, ms, True
            for kk, vv in best.items():
                h.tune_set(kk, vv)
        if not changed:
            break
    return best, best_ms, trials


def apply(h, setting):
    """set every switch of `setting` (a dict)"""
    for k, v in setting.items():
        h.tune_set(k, int(v))


def load(path, workload):
    """the setting saved for `workload`, or None"""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get(workload, {}).get("setting")


def save(path, workload, setting, ms, trials=None, note=""):
    """add or replace the entry of `workload` in the JSON file at path"""
    data = {}
    if os.path.exists(path):
        with open(path) as f:
            data = json.load(f)
    data[workload] = {"setting": setting, "ms_per_step": ms, "note": note,
                      "trials": [{"setting": s, "ms_per_step": m} for s, m in (trials or [])]}
    with open(path, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
