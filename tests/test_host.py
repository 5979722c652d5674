"""Host logic of the pdx package (CPU only)."""
import math

import numpy as np
import pytest

import pdx
from pdx import systems


def test_fft_friendly_and_alpha():
    assert pdx.fft_friendly(113) == 120
    assert pdx.fft_friendly(57) == 60
    assert pdx.fft_friendly(11) == 12
    a = pdx.alpha_from_tolerance(9.0, 1e-5)
    assert math.erfc(a * 9.0) == pytest.approx(1e-5, rel=1e-9)
    assert a == pytest.approx(0.347046, abs=1e-6)
    with pytest.raises(pdx.ArgumentError):
        pdx.alpha_from_tolerance(9.0, 1.5)


def test_spec_validation():
    with pytest.raises(pdx.ArgumentError):
        pdx.PME(coulomb_prefactor=1.0, real_cutoff=4.0, alpha=0.9)  # no mesh
    with pytest.raises(pdx.ArgumentError):
        pdx.PME(coulomb_prefactor=1.0, real_cutoff=4.0, mesh=(8, 8, 8))  # no alpha
    s = pdx.PME(coulomb_prefactor=14.399645, real_cutoff=9.0, real_space_tolerance=1e-5, mesh_spacing=1.0, order=5)
    assert s.C == 14.399645
    c = s.c_spec()
    assert c.order == 5 and c.mesh[0] == 0 and c.mesh_spacing == 1.0


def test_systems():
    s = systems.fcc_lattice(4000, 0.8442, 0.05, seed=1)
    L = s["box"][0]
    assert L == pytest.approx((4000 / 0.8442) ** (1 / 3))
    for k in "xyz":
        assert s[k].dtype == np.float32 and s[k].min() >= -L / 2 and s[k].max() <= L / 2
    r = systems.rock_salt(40, 2.82, 0.1, seed=0)
    assert r["N"] == 64000 and r["q"].sum() == 0
    assert r["box"][0] == pytest.approx(112.8)
    v = systems.rock_salt(10, 2.82, 0.1, seed=0, vacancy_fraction=0.05)
    assert v["N"] == 1000 - 2 * 25 and v["q"].sum() == 0
    assert pdx.default_capacity(4000, s["box"], 2.9) == math.ceil(1.5 * 4 * math.pi / 3 * 2.9 ** 3 * 0.8442 + 64)


def test_esp_descriptor_rules_and_refusals():
    """The ESP descriptor's defaults (src/workflow/Electrostatics.jl:177-215,288-301): at 1e-4 the
    split shape of Table 2 (12.024), width 6, window shape c(5e-5), spacing pi rc / c; the mesh
    refuses, before any device work, a window wider than 2 c1/pi - 1/2 (Spec.jl:74-75), a width
    this build does not compile, and a missing shape."""
    c, P, c1 = pdx.esp_rule(1e-4)
    assert c == pytest.approx(12.024, rel=5e-4) and P == 6 and c1 == pytest.approx(12.762, rel=5e-4)
    assert pdx.esp_rule(1e-3)[1] == 5 and pdx.esp_rule(1e-5)[1] == 8 and pdx.esp_rule(1e-4, 7)[1] == 7
    s = pdx.ESP(coulomb_prefactor=14.4, real_cutoff=9.0, tolerance=1e-4)
    assert s.order == 6 and s.influence == "esp"
    assert s.mesh_spacing == pytest.approx(math.pi * 9.0 / c)
    assert s.resolve_mesh((112.8, 112.8, 112.8)) == (48, 48, 48)
    cs = s.c_spec()
    assert cs.influence == 4 and cs.split_shape == pytest.approx(c) and cs.window_shape == pytest.approx(c1)
    assert pdx.ESP(coulomb_prefactor=1.0, real_cutoff=4.0, influence="deconvolution").influence == "esp_deconvolution"
    with pytest.raises(pdx.ArgumentError, match="supports at most"):
        pdx.ElectrostaticsSolver(pdx.ESP(coulomb_prefactor=1.0, real_cutoff=4.0, split_shape=12.0, window_shape=10.0,
                                         order=8, mesh=(24, 24, 24)), (14.0, 15.0, 16.0), np.zeros(10))
    with pytest.raises(pdx.ArgumentError, match="not compiled"):
        pdx.ElectrostaticsSolver(pdx.ESP(coulomb_prefactor=1.0, real_cutoff=4.0, tolerance=1e-6, order=10,
                                         mesh=(24, 24, 24)), (14.0, 15.0, 16.0), np.zeros(10))
    with pytest.raises(pdx.ArgumentError):
        pdx.ESP(coulomb_prefactor=1.0, real_cutoff=4.0, tolerance=None)
    with pytest.raises(pdx.ArgumentError):
        pdx.ESP(coulomb_prefactor=1.0, real_cutoff=4.0, influence="ik")


def test_autotune_classes_and_table(tmp_path):
    """The reference's classes (engine_settings, src/engine/StateEngine.jl:47-56): size N <= 30 000,
    <= 300 000, larger; neighbours rho 4 pi/3 rc^3 with borders 9.8, 36, 72; a saved table is read
    back and looked up by class."""
    from pdx import autotune
    L = (4000 / 0.8442) ** (1 / 3)
    assert autotune.classify(4000, (L, L, L), 2.5) == (1, 3)       # LJ fluid: 55 neighbours
    assert autotune.classify(64000, (112.8,) * 3, 9.0) == (2, 4)   # rock salt: 135 within 9 A
    assert autotune.classify(486400, (225.6,) * 3, 10.0) == (3, 4)
    assert autotune.classify(1000, (100.0,) * 3, 2.5) == (1, 1)
    path = str(tmp_path / "t.json")
    autotune.save({"2,4,mesh": {"force_lanes": 4, "overlap": "fft"}}, path)
    t = autotune.load(path)
    assert autotune.settings_for(64000, (112.8,) * 3, 9.0, True, table=t)["force_lanes"] == 4
    assert autotune.settings_for(64000, (112.8,) * 3, 9.0, False, table=t) is None
    assert autotune.load(str(tmp_path / "missing.json")) == {}
