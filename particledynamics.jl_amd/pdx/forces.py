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


class MixedSigmaLJ:
    """Per-particle diameters, sigma_ij = (sigma_i + sigma_j)/2, rc = rcut_factor sigma_ij, one
    epsilon (``MixedSigmaLJ``, src/engine/PairModels.jl:16,61-65); ``shift``: u - u(rc), the same
    constant for every pair."""
    model = L.PDX_PAIR_MIXED_LJ

    def __init__(self, sigmas, epsilon=1.0, rcut_factor=2.5, shift=False):
        import numpy as np
        self.sigmas = np.ascontiguousarray(sigmas, dtype=np.float32)
        self.epsilon, self.rcut_factor, self.shift = float(epsilon), float(rcut_factor), bool(shift)

    @property
    def cutoff(self):
        return self.rcut_factor * float(self.sigmas.max())


class MixedSigmaWCA(MixedSigmaLJ):
    """``MixedSigmaWCA`` (PairModels.jl:17,66-70): the mixed-sigma pairs with u + epsilon; the
    reference takes the factor as given (2^(1/6) for a WCA cut)."""
    model = L.PDX_PAIR_MIXED_WCA

    def __init__(self, sigmas, epsilon=1.0, rcut_factor=2 ** (1 / 6)):
        super().__init__(sigmas, epsilon, rcut_factor, False)


class PairMatrixLJ:
    """sigma, epsilon and rc by type pair (``PairMatrixLJ{SHIFT}``, PairModels.jl:18,71-77):
    ``types`` 1-based per particle, ``sigma`` / ``epsilon`` / ``rcut`` ntypes x ntypes."""
    model = L.PDX_PAIR_MATRIX_LJ

    def __init__(self, types, sigma, epsilon, rcut, shift=False):
        import numpy as np
        self.types = np.ascontiguousarray(types, dtype=np.int32)
        self.sigma = np.asarray(sigma, dtype=np.float64)
        self.epsilon_m = np.asarray(epsilon, dtype=np.float64)
        self.rcut = np.asarray(rcut, dtype=np.float64)
        self.shift = bool(shift)
        nt = self.sigma.shape[0]
        if self.sigma.shape != (nt, nt) or self.epsilon_m.shape != (nt, nt) or self.rcut.shape != (nt, nt):
            raise L.DimensionMismatch("sigma, epsilon and rcut must be ntypes x ntypes")

    @property
    def cutoff(self):
        return float(self.rcut.max())


class PairMatrixWCA(PairMatrixLJ):
    """``PairMatrixWCA`` (PairModels.jl:19,78-81): the matrix pairs with u + epsilon_ab."""
    model = L.PDX_PAIR_MATRIX_WCA

    def __init__(self, types, sigma, epsilon, rcut):
        super().__init__(types, sigma, epsilon, rcut, False)


@dataclass(frozen=True)
class RealSpaceCoulomb:
    """Short-range Coulomb C q_i q_j E_s(r)/r inside ``cutoff``: the Gaussian (Ewald) split
    E_s = erfc(alpha r), or the prolate split of ESP (``split="prolate"``, bandlimit ``shape``),
    whose E_s vanishes at the cutoff (src/electrostatics/Splits.jl:91-130)."""
    C: float
    cutoff: float
    alpha: float
    shift: bool = False
    split: str = "gaussian"
    shape: float = 0.0


ACCUMULATION = {"fp32": 0, "fixed64": 1, "fp64": 2}


def _sig(lib):
    vp, i32, f64 = C.c_void_p, C.c_int32, C.c_double
    P = C.POINTER
    for name, args in {
        "pdx_pair_set_accumulation": [vp, i32],
        "pdx_bonds_create": [i32, i32, P(C.c_int32), P(C.c_int32), f64, f64, P(vp)],
        "pdx_bonds_destroy": [vp],
        "pdx_pair_set_bonds": [vp, vp],
        "pdx_pair_set_bond_exclusion": [vp, i32],
        "pdx_bonds_status": [vp, P(C.c_int32), i32, vp],
        "pdx_pair_set_traversal": [vp, i32],
        "pdx_pair_set_exceptions": [vp, i32, i32, P(C.c_int32), P(C.c_int32), P(C.c_double)],
    }.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int


_sig(L.lib)


class DomainError(ValueError):
    """A FENE bond reached r >= R0 (the reference's ``DomainError``)."""


class FENEBonds:
    """FENE bonds u = -k R0^2/2 ln(1 - r^2/R0^2) (``FENEParams``, src/Definitions.jl:93-101;
    kernels in src/bonded/FENEBondCUDA.jl). ``pairs``: (nb, 2) 0-based particle ids."""

    def __init__(self, n, pairs, k=30.0, R0=1.5):
        import numpy as np
        b = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        bi, bj = np.ascontiguousarray(b[:, 0]), np.ascontiguousarray(b[:, 1])
        h = C.c_void_p()
        P = C.POINTER(C.c_int32)
        L.check(L.lib.pdx_bonds_create(int(n), len(b), bi.ctypes.data_as(P), bj.ctypes.data_as(P), float(k),
                                       float(R0), C.byref(h)))
        self._h = h
        self._release = L.adopt(self, h, L.lib.pdx_bonds_destroy)
        self.n, self.nbond, self.k, self.R0 = int(n), len(b), float(k), float(R0)
        self.pairs = b.copy()

    @property
    def handle(self):
        return self._h


    def status(self, reset=True, stream=None):
        """Raise ``DomainError`` naming the first particle whose bond left the FENE domain
        in the evaluations since the last reset; returns -1 otherwise."""
        v = C.c_int32()
        rc = L.lib.pdx_bonds_status(self._h, C.byref(v), 1 if reset else 0, L.stream_handle(stream))
        if rc == L.PDX_ERR_COORDINATE:
            raise DomainError(L.last_error())
        L.check(rc)
        return v.value


class PairForce:
    def __init__(self, model=None, coulomb: RealSpaceCoulomb | None = None, lanes: int = 0,
                 accumulation: str = "fp32", bonds: FENEBonds | None = None, traversal: str = "rows"):
        spec = L.pdx_pair_spec()
        if model is None:
            spec.model = L.PDX_PAIR_NONE
        elif isinstance(model, MixedSigmaLJ):
            spec.model = model.model
            spec.epsilon = model.epsilon
            spec.sigma = 1.0
            spec.cutoff = model.rcut_factor
            spec.shift = 1 if model.shift else 0
        elif isinstance(model, PairMatrixLJ):
            spec.model = model.model
            spec.shift = 1 if model.shift else 0
        else:
            spec.model = model.model
            spec.epsilon = float(model.epsilon)
            spec.sigma = float(model.sigma)
            spec.cutoff = float(getattr(model, "cutoff", 0.0))
            spec.shift = 1 if getattr(model, "shift", False) else 0
        if coulomb is not None:
            if coulomb.split not in ("gaussian", "prolate"):
                raise L.ArgumentError("the real-space split must be 'gaussian' or 'prolate'")
            spec.coulomb = 2 if coulomb.split == "prolate" else 1
            spec.coulomb_shift = 1 if coulomb.shift else 0
            spec.coulomb_C = float(coulomb.C)
            spec.coulomb_cutoff = float(coulomb.cutoff)
            spec.alpha = float(coulomb.alpha)
            spec.split_shape = float(coulomb.shape)
        spec.lanes = int(lanes)
        h = C.c_void_p()
        L.check(L.lib.pdx_pair_create(C.byref(spec), C.byref(h)))
        self._h = h
        self._release = L.adopt(self, h, L.lib.pdx_pair_destroy)
        self.model = model
        if isinstance(model, MixedSigmaLJ):
            L.check(L.lib.pdx_pair_set_sigmas(h, len(model.sigmas), model.sigmas.ctypes.data_as(C.POINTER(C.c_float))))
        elif isinstance(model, PairMatrixLJ):
            self._set_types(model)
        self.coulomb = coulomb
        self.bonds = None
        self.set_accumulation(accumulation)
        self.set_traversal(traversal)
        if bonds is not None:
            self.set_bonds(bonds)

    def set_lanes(self, lanes: int):
        """Lanes per particle of the row traversal (0 automatic; the force lanes of the engine
        settings), set before a graph capture."""
        L.check(L.lib.pdx_pair_set_lanes(self._h, int(lanes)))
        self.lanes = int(lanes)

    def set_traversal(self, mode: str):
        """"rows": the list's ELL rows; "cells": the list's cell grid directly, one wave per
        particle with ballot-compacted candidates (no rows are read)."""
        if mode not in ("rows", "cells"):
            raise L.ArgumentError("traversal must be 'rows' or 'cells'")
        L.check(L.lib.pdx_pair_set_traversal(self._h, 0 if mode == "rows" else 1))
        self.traversal = mode

    def set_accumulation(self, mode: str):
        """Precision of the per-particle sums: "fp32", "fixed64" (64-bit fixed point, exact)
        or "fp64"; the pair arithmetic itself stays FP32."""
        if mode not in ACCUMULATION:
            raise L.ArgumentError("accumulation must be 'fp32', 'fixed64' or 'fp64'")
        L.check(L.lib.pdx_pair_set_accumulation(self._h, ACCUMULATION[mode]))
        self.accumulation = mode

    def set_exceptions(self, n, exceptions):
        """Electrostatic exceptions as (i, j) or (i, j, scale) with 1-based particle ids, the
        reference's convention (normalize_exceptions, src/electrostatics/Spec.jl:200-224)."""
        import numpy as np
        ii, jj, ss = [], [], []
        for item in exceptions:
            if len(item) not in (2, 3):
                raise L.ArgumentError("Each electrostatic exception must be (i, j) or (i, j, scale).")
            ii.append(int(item[0]) - 1)
            jj.append(int(item[1]) - 1)
            ss.append(float(item[2]) if len(item) == 3 else 0.0)
        a = np.ascontiguousarray(ii, dtype=np.int32)
        b = np.ascontiguousarray(jj, dtype=np.int32)
        c = np.ascontiguousarray(ss, dtype=np.float64)
        L.check(L.lib.pdx_pair_set_exceptions(self._h, int(n), len(a), a.ctypes.data_as(C.POINTER(C.c_int32)),
                                              b.ctypes.data_as(C.POINTER(C.c_int32)),
                                              c.ctypes.data_as(C.POINTER(C.c_double))))

    def set_bond_exclusion(self, exclude: bool):
        """Bonded pairs feel no pair term (the reference's ``exclude_bonded``)."""
        L.check(L.lib.pdx_pair_set_bond_exclusion(self._h, 1 if exclude else 0))
        self.exclude_bonded = bool(exclude)

    def set_bonds(self, bonds: FENEBonds | None):
        L.check(L.lib.pdx_pair_set_bonds(self._h, bonds.handle if bonds is not None else None))
        self.bonds = bonds  # kept alive while attached


    def info(self):
        out = (C.c_double * 4)()
        L.check(L.lib.pdx_pair_info(self._h, out))
        return dict(n_energy_terms=int(out[0]), n_force_terms=int(out[1]), energy_fit_error=out[2],
                    force_fit_error=out[3])

    def _set_types(self, model):
        import numpy as np
        cm = [np.ascontiguousarray(np.asarray(m, dtype=np.float64).reshape(-1, order="F"))
              for m in (model.sigma, model.epsilon_m, model.rcut)]
        D = C.POINTER(C.c_double)
        L.check(L.lib.pdx_pair_set_types(self._h, len(model.types), model.sigma.shape[0],
                                         model.types.ctypes.data_as(C.POINTER(C.c_int32)), cm[0].ctypes.data_as(D),
                                         cm[1].ctypes.data_as(D), cm[2].ctypes.data_as(D)))

    def set_sigmas(self, sigmas):
        import numpy as np
        s = np.ascontiguousarray(sigmas, dtype=np.float32)
        L.check(L.lib.pdx_pair_set_sigmas(self._h, len(s), s.ctypes.data_as(C.POINTER(C.c_float))))

    def cutoff(self):
        """The largest pair cutoff of the model (``pdx_pair_cutoff``)."""
        r = C.c_double()
        L.check(L.lib.pdx_pair_cutoff(self._h, C.byref(r)))
        return r.value

    def compute_at(self, nl, x, y, z, q, fx, fy, fz, Epot=None, V=None, totals=None, energy=True, add=False,
                   stream=None):
        """Forces at the given coordinates (``pdx_pair_compute``): x, y, z, q must be the arrays
        the list's last ``update`` / ``check`` packed."""
        flags = (L.PDX_ENERGY if energy else 0) | (L.PDX_ADD if add else 0)
        L.check(L.lib.pdx_pair_compute(self._h, nl.handle, L.ptr(x), L.ptr(y), L.ptr(z), L.ptr(q), L.ptr(fx),
                                       L.ptr(fy), L.ptr(fz), L.ptr(Epot), L.ptr(V), L.ptr(totals), flags,
                                       L.stream_handle(stream)))

    def compute(self, nl, fx, fy, fz, Epot=None, V=None, totals=None, energy=True, add=False, stream=None):
        """Enqueue the pair kernel over ``nl``'s rows (positions from ``nl.update``)."""
        flags = (L.PDX_ENERGY if energy else 0) | (L.PDX_ADD if add else 0)
        L.check(L.lib.pdx_pair_forces(self._h, nl.handle, L.ptr(fx), L.ptr(fy), L.ptr(fz), L.ptr(Epot),
                                      L.ptr(V), L.ptr(totals), flags, L.stream_handle(stream)))
