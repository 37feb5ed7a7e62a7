"""The schedule autotuner: coordinate descent over knobs that do not change a
bit of any result, the bit-identity guard, and the settings cache.

`dev` = "cuda:0" runs on the MI355X (marked gpu); `dev` = "cpu" runs the CPU
emulation build.  The last two tests are host logic only.
"""
import pytest
import torch

from mgpcg_amd import scenes
from test_solve import KNOB_DEFAULTS, _emu, set_knobs


def test_tune_small_space(dev, tmp_path):
    from mgpcg_amd import autotune
    from mgpcg_amd import systems

    emu = _emu(dev)
    lo, hi = (16, 24) if emu else (32, 64)
    reqs = [systems.upload(scenes.c5_request(3000 + i, lo=lo, hi=hi, L=3), dev)
            for i in range(3)]
    space = {"max_grid": [2048, 64], "fuse_p": [0, 1]}
    start = autotune.current(list(space))
    seen = []
    try:
        out = autotune.tune(lambda i: reqs[i % len(reqs)], 3, concurrency=2, requests=3,
                            warmup=1, space=space, max_passes=1, restore=True, device=dev,
                            log=seen.append)
    finally:
        set_knobs(**KNOB_DEFAULTS)
    assert autotune.current(list(space)) == start  # restore=True
    # the start, then every other value of each knob
    assert len(out["trials"]) == 1 + sum(len(v) - 1 for v in space.values()) == len(seen)
    assert out["trials"][0]["knobs"] == start
    for k, v in out["knobs"].items():
        assert v in space[k]
    assert out["solves_per_s"] > 0
    assert out["solves_per_s"] >= max(t["solves_per_s"] for t in out["trials"]) / 1.01 - 1e-9

    path = str(tmp_path / "tuned.json")
    key = autotune.cache_key([32, 32, 32], 3, "mixed", 2, device=dev)
    autotune.save(path, key, out)
    try:
        assert autotune.load(path, key) == out["knobs"]
        assert autotune.current(list(space)) == out["knobs"]
        assert autotune.load(path, "absent") is None
    finally:
        set_knobs(**KNOB_DEFAULTS)


def test_tune_rejects_a_setting_that_changes_results(monkeypatch):
    """Host logic: a candidate whose results differ from the first
    measurement's stops the tuner, and the starting settings come back."""
    from mgpcg_amd import autotune

    state = {"n": 0, "applied": []}
    monkeypatch.setattr(autotune, "current", lambda names=None: {"fuse_p": 0})
    monkeypatch.setattr(autotune, "apply", lambda k: state["applied"].append(dict(k)))

    def served(*a, **k):  # the untimed round, the start, then a candidate
        state["n"] += 1
        return 100.0, {0: torch.tensor([1.0 if state["n"] <= 2 else 2.0])}

    monkeypatch.setattr(autotune, "_served", served)
    with pytest.raises(RuntimeError, match="changed the result"):
        autotune.tune(lambda i: None, 3, space={"fuse_p": [0, 1]}, restore=True, device="cpu")
    assert state["applied"][-1] == {"fuse_p": 0}


def test_tune_keeps_only_clear_gains(monkeypatch):
    """Host logic: a value is kept only when it beats the best by min_gain;
    a pass that changes nothing ends the descent."""
    from mgpcg_amd import autotune

    rates = {(2048, 0): 100.0, (64, 0): 100.5, (4096, 0): 120.0, (4096, 1): 119.0,
             (64, 1): 99.0, (2048, 1): 98.0}
    cur = {"max_grid": 2048, "fuse_p": 0}
    monkeypatch.setattr(autotune, "current", lambda names=None: dict(cur))
    monkeypatch.setattr(autotune, "apply", lambda k: cur.update(k))
    monkeypatch.setattr(autotune, "_served",
                        lambda *a, **k: (rates[(cur["max_grid"], cur["fuse_p"])],
                                         {0: torch.zeros(1)}))
    out = autotune.tune(lambda i: None, 3, space={"max_grid": [2048, 64, 4096], "fuse_p": [0, 1]},
                        min_gain=0.01, device="cpu")
    assert out["knobs"] == {"max_grid": 4096, "fuse_p": 0}
    assert out["solves_per_s"] == 120.0 and out["start"]["solves_per_s"] == 100.0
    assert cur == out["knobs"]  # the winner stays applied
