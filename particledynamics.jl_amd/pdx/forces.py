"""Pair-force descriptors and the device pair model (host side of libpdx's pair kernels).

``LennardJones`` / ``WCA`` / ``SoftRepulsive`` mirror the workflow descriptors of
src/workflow/Forces.jl:37-60 and the device models of src/engine/PairModels.jl:13-54;
``PairForce`` owns the compiled model (``pdx_pair_create``) and evaluates forces,
per-particle half energies / half virials and FP64 totals over a neighbor list's rows.
A ``coulomb`` real-space term (Ewald split) turns the model into the reference's
``ChargedPairs`` (PairModels.jl:84-110).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

from . import _lib as L


@dataclass(frozen=True)
class LennardJones:
    epsilon: float = 1.0
    sigma: float = 1.0
    cutoff: float = 2.5
    shift: bool = False  # potential-shift modifier (energy only)
    model = L.PDX_PAIR_LJ


@dataclass(frozen=True)
class WCA:
    epsilon: float = 1.0
    sigma: float = 1.0
    model = L.PDX_PAIR_WCA

    @property
    def cutoff(self):
        return (1.2599211 ** 0.5) * self.sigma


@dataclass(frozen=True)
class SoftRepulsive:
    epsilon: float = 1.0
    sigma: float = 1.0
    model = L.PDX_PAIR_SOFT

    @property
    def cutoff(self):
        return self.sigma


@dataclass(frozen=True)
class RealSpaceCoulomb:
    """Short-range Ewald Coulomb C q_i q_j erfc(alpha r)/r inside ``cutoff``."""
    C: float
    cutoff: float
    alpha: float
    shift: bool = False


class PairForce:
    def __init__(self, model=None, coulomb: RealSpaceCoulomb | None = None, lanes: int = 0):
        spec = L.pdx_pair_spec()
        if model is None:
            spec.model = L.PDX_PAIR_NONE
        else:
            spec.model = model.model
            spec.epsilon = float(model.epsilon)
            spec.sigma = float(model.sigma)
            spec.cutoff = float(getattr(model, "cutoff", 0.0))
            spec.shift = 1 if getattr(model, "shift", False) else 0
        if coulomb is not None:
            spec.coulomb = 1
            spec.coulomb_shift = 1 if coulomb.shift else 0
            spec.coulomb_C = float(coulomb.C)
            spec.coulomb_cutoff = float(coulomb.cutoff)
            spec.alpha = float(coulomb.alpha)
        spec.lanes = int(lanes)
        h = C.c_void_p()
        L.check(L.lib.pdx_pair_create(C.byref(spec), C.byref(h)))
        self._h = h
        self.model = model
        self.coulomb = coulomb

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            L.lib.pdx_pair_destroy(h)
            self._h = None

    def info(self):
        out = (C.c_double * 4)()
        L.check(L.lib.pdx_pair_info(self._h, out))
        return dict(n_energy_terms=int(out[0]), n_force_terms=int(out[1]), energy_fit_error=out[2],
                    force_fit_error=out[3])

    def compute(self, nl, fx, fy, fz, Epot=None, V=None, totals=None, energy=True, add=False, stream=None):
        """Enqueue the pair kernel over ``nl``'s rows (positions from ``nl.update``)."""
        flags = (L.PDX_ENERGY if energy else 0) | (L.PDX_ADD if add else 0)
        L.check(L.lib.pdx_pair_forces(self._h, nl.handle, L.ptr(fx), L.ptr(fy), L.ptr(fz), L.ptr(Epot),
                                      L.ptr(V), L.ptr(totals), flags, L.stream_handle(stream)))
