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
