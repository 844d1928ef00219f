"""The autotuner's coordinate descent and its settings file, on the CPU with
a stand-in for the library handle whose step time is a known function of
the switches."""
import json

from nemo_hip import autotune


class FakeLib:
    """tune_set / tune_get / sync, and a step whose cost depends on the switches"""

    def __init__(self, cost):
        self.sw = {k: 0 for k in autotune.SWITCHES}
        self.cost = cost
        self.clock = 0.0

    def tune_set(self, k, v):
        self.sw[k] = v

    def tune_get(self, k):
        return self.sw[k]

    def sync(self):
        pass

    def step(self):
        self.clock += self.cost(self.sw) * 1e-3     # cost in ms, clock in s


def run(cost, monkeypatch):
    h = FakeLib(cost)
    monkeypatch.setattr(autotune.time, "perf_counter", lambda: h.clock)
    return h, autotune.tune(h, h.step, steps=2, warmup=0)


def test_descent_finds_the_separable_optimum(monkeypatch):
    # fastest: fct_fused 2, spg_persistent 0, zdf_reg 1, fct_vlim 1
    cost = lambda s: (10 + [3, 2, 1][s["fct_fused"]] + [0, 5, 9][s["spg_persistent"]] + [1, 0][s["zdf_reg"]]
                      + [2, 0, 4][s["fct_vlim"]])
    h, (best, ms, trials) = run(cost, monkeypatch)
    assert best == {"fct_fused": 2, "spg_persistent": 0, "zdf_reg": 1, "fct_vlim": 1}
    assert abs(ms - 11.0) < 1e-9
    assert h.sw == best                      # the best setting is left applied
    assert len(trials) >= 1 + 2 + 2 + 1 + 2


def test_descent_keeps_the_start_when_it_is_best(monkeypatch):
    cost = lambda s: 1.0 + s["fct_fused"] + s["spg_persistent"] + s["zdf_reg"] + s["fct_vlim"]
    h, (best, ms, trials) = run(cost, monkeypatch)
    assert best == {k: 0 for k in autotune.SWITCHES} and abs(ms - 1.0) < 1e-9


def test_save_load_apply(tmp_path):
    p = str(tmp_path / "tuning.json")
    autotune.save(p, "fp32_fused 1024x1024x75", {"fct_fused": 1, "spg_persistent": 0, "zdf_reg": 0}, 13.8,
                  [({"fct_fused": 0}, 13.9)], note="n")
    autotune.save(p, "fp64 1024x1024x75", {"fct_fused": 0}, 25.8)
    assert autotune.load(p, "fp32_fused 1024x1024x75") == {"fct_fused": 1, "spg_persistent": 0, "zdf_reg": 0}
    assert autotune.load(p, "fp64 1024x1024x75") == {"fct_fused": 0}
    assert autotune.load(p, "other") is None and autotune.load(str(tmp_path / "none.json"), "x") is None
    data = json.load(open(p))
    assert data["fp32_fused 1024x1024x75"]["trials"][0]["ms_per_step"] == 13.9
    h = FakeLib(lambda s: 0)
    autotune.apply(h, {"fct_fused": 2, "zdf_reg": 1})
    assert h.sw["fct_fused"] == 2 and h.sw["zdf_reg"] == 1


def test_committed_settings_name_known_switches():
    import os
    p = os.path.join(os.path.dirname(autotune.__file__), "tuning.json")
    if not os.path.exists(p):
        return
    for wl, e in json.load(open(p)).items():
        for k, v in e["setting"].items():
            assert k in autotune.SWITCHES and v in autotune.SWITCHES[k], (wl, k, v)
