"""The launch-shape autotuner (pdx.autotune, the counterpart of the reference's ENGINE_SETTINGS,
src/engine/StateEngine.jl:7-56) on the device: every candidate setting gives the same forces as
the defaults (bitwise for the spread and transform sizes and the overlap mode, to FP32 rounding
of the row sums for the force lanes), the tuner applies its fastest candidate, and a recorded
table is looked up by class."""
import numpy as np
import pytest

import pdx
from conftest import dev
from pdx import autotune, systems

pytestmark = pytest.mark.gpu


def test_settings_keep_forces_and_tuner_applies_best(cuda, tmp_path):
    import torch
    s = systems.rock_salt(20, 2.82, 0.1, seed=2)
    N, box = s["N"], s["box"]
    spec = pdx.PME(coulomb_prefactor=14.399645, real_cutoff=9.0, real_space_tolerance=1e-5, mesh=(60, 60, 60), order=5)
    lj = pdx.LennardJones(epsilon=0.01, sigma=2.8, cutoff=9.0)
    X, Y, Z, Q = (dev(s[k]) for k in ("x", "y", "z", "q"))

    def forces(setting):
        eng = pdx.ForceEngine(N, box, lj, electrostatics=spec, charges=s["q"], skin=1.0, settings=setting)
        f = [torch.empty(N, device="cuda") for _ in range(3)]
        for _ in range(2):
            eng.evaluate(X, Y, Z, Q, *f, None, None, energy=False)
        eng.status()
        return torch.stack(f, 1).double().cpu().numpy()

    F0 = forces({"force_lanes": 4, "overlap": "fft", "spread_threads": 512, "transform_threads": 512})
    for setting in ({"force_lanes": 4, "overlap": "spread", "spread_threads": 256, "transform_threads": 1024},
                    {"force_lanes": 4, "overlap": "fft", "spread_threads": 512, "transform_threads": 512,
                     "build_lanes": 8}):
        assert np.array_equal(forces(setting), F0), setting
    for lanes in (1, 2, 8):
        F = forces({"force_lanes": lanes})
        assert np.abs(F - F0).max() <= 2e-6 * np.abs(F0).max(), lanes
    eng = pdx.ForceEngine(N, box, lj, electrostatics=spec, charges=s["q"], skin=1.0)
    cands = [{"force_lanes": fl, "overlap": ov} for fl in (2, 4) for ov in ("fft", "spread")]
    res = autotune.tune(eng, X, Y, Z, Q, cands=cands, steps=3, warmup=1, rounds=1)
    assert len(res["table"]) == 4
    assert round(res["ms"], 4) == min(ms for _, ms in res["table"])
    assert autotune.current(eng)["force_lanes"] == res["best"]["force_lanes"]
    assert autotune.current(eng)["overlap"] == res["best"]["overlap"]
    path = str(tmp_path / "settings.json")
    table = autotune.record(eng, res, table={}, path=path)
    found = autotune.settings_for(N, box, 9.0, True, table=autotune.load(path))
    assert found == {k: v for k, v in res["best"].items() if k in autotune.KNOBS}
