"""Smooth PME on libpdx: specification, solver and the FP64 Ewald reference.

``ElectrostaticsSpec`` and ``PME`` mirror src/electrostatics/Spec.jl:31-102 and the workflow
descriptor src/workflow/Electrostatics.jl:45-63; ``ElectrostaticsSolver`` mirrors
src/electrostatics/Solver.jl:75-158 (``compute_mesh``, ``mesh_transform``, ``mesh_gather``,
``set_solver_charges``, ``refresh_geometry``, ``self_energy``, ``background_energy``).
The solver holds no references to particle arrays: they are passed on every call.
"""
from __future__ import annotations

import ctypes
import ctypes as C
import math
from dataclasses import dataclass, field

from . import _lib as L


def fft_friendly(n: int) -> int:
    """Smallest integer >= n whose prime factors are in {2, 3, 5, 7} (Spec.jl:122-134)."""
    m = max(int(n), 1)
    while True:
        r = m
        for p in (2, 3, 5, 7):
            while r % p == 0:
                r //= p
        if r == 1:
            return m
        m += 1


def alpha_from_tolerance(cutoff: float, tol: float) -> float:
    """alpha with erfc(alpha r_c) = tol, by bisection (Spec.jl:154-163)."""
    if not (0 < tol < 1):
        raise L.ArgumentError("real_space_tolerance must lie in (0, 1).")
    lo, hi = 0.0, 10.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if math.erfc(mid) - tol > 0:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi) / cutoff


@dataclass
class ElectrostaticsSpec:
    coulomb_prefactor: float
    real_cutoff: float
    alpha: float | None = None
    real_space_tolerance: float | None = None
    order: int = 4
    mesh: tuple | None = None
    mesh_spacing: float | None = None
    relative_permittivity: float = 1.0
    net_charge: str = "error"
    real_space_modifier: str = "none"
    net_charge_tolerance: float = 1e-8
    influence: str = "spme"
    alias_order: int = 2
    split_shape: float | None = None   # ESP: the prolate split's bandlimit c
    window_shape: float | None = None  # ESP: the prolate window's shape c1 (``order`` is its width)

    @property
    def is_esp(self):
        return self.influence in ("esp", "esp_deconvolution")

    def __post_init__(self):
        if self.mesh is None and self.mesh_spacing is None:
            raise L.ArgumentError("Provide either `mesh=(K1,K2,K3)` or `mesh_spacing`.")
        if self.influence in ("esp", "esp_deconvolution"):
            if self.split_shape is None or self.window_shape is None:
                raise L.ArgumentError("ESP requires a prolate split and a prolate window.")
        elif self.alpha is None and self.real_space_tolerance is None:
            raise L.ArgumentError("Provide `alpha` or `real_space_tolerance`.")
        if self.net_charge not in ("error", "background"):
            raise L.ArgumentError("net_charge must be :error or :background")
        if self.real_space_modifier not in ("none", "potential_shift"):
            raise L.ArgumentError("real_space_modifier must be :none or :potential_shift")
        if self.influence not in ("spme", "p3m_ad", "p3m_ik", "esp", "esp_deconvolution"):
            raise L.ArgumentError(f"Unknown influence kind {self.influence}.")
        if not 0 <= int(self.alias_order) <= 6:
            raise L.ArgumentError("alias_order must be between 0 and 6.")

    @property
    def C(self):
        return self.coulomb_prefactor / self.relative_permittivity

    def resolve_mesh(self, box):
        """(K1, K2, K3): the given mesh, or the smallest 2,3,5,7-smooth count covering each box
        length at the spacing, at least the order (resolve_mesh, Spec.jl:141-146)."""
        if self.mesh is not None:
            return tuple(int(k) for k in self.mesh)
        return tuple(max(fft_friendly(int(math.ceil(b / self.mesh_spacing - 1e-9))), int(self.order)) for b in box)

    def resolved_alpha(self):
        if self.is_esp:
            return None
        return float(self.alpha) if self.alpha is not None else alpha_from_tolerance(self.real_cutoff,
                                                                                     self.real_space_tolerance)

    def c_spec(self):
        s = L.pdx_mesh_spec()
        s.coulomb_prefactor = float(self.coulomb_prefactor)
        s.relative_permittivity = float(self.relative_permittivity)
        s.real_cutoff = float(self.real_cutoff)
        s.alpha = float(self.alpha) if self.alpha is not None else 0.0
        s.real_space_tolerance = float(self.real_space_tolerance or 0.0)
        s.order = int(self.order)
        if self.mesh is not None:
            for d in range(3):
                s.mesh[d] = int(self.mesh[d])
        else:
            s.mesh[0] = 0
        s.mesh_spacing = float(self.mesh_spacing or 0.0)
        s.influence = {"spme": L.PDX_INFLUENCE_SPME, "p3m_ad": L.PDX_INFLUENCE_P3M_AD,
                       "p3m_ik": L.PDX_INFLUENCE_P3M_IK, "esp": L.PDX_INFLUENCE_ESP,
                       "esp_deconvolution": L.PDX_INFLUENCE_ESP_DECONVOLUTION}[self.influence]
        s.split_shape = float(self.split_shape or 0.0)
        s.window_shape = float(self.window_shape or 0.0)
        s.alias_order = int(self.alias_order)
        s.net_charge = 1 if self.net_charge == "background" else 0
        s.net_charge_tolerance = float(self.net_charge_tolerance)
        s.real_space_modifier = 1 if self.real_space_modifier == "potential_shift" else 0
        return s


def PPPM(*, coulomb_prefactor, real_cutoff, alpha=None, real_space_tolerance=None, mesh=None, mesh_spacing=None,
        order=5, differentiation="ad", alias_order=2, relative_permittivity=1.0, net_charge="error",
        real_space_modifier="none"):
    """Workflow descriptor of P3M with the Hockney-Eastwood optimal influence function for
    analytic ("ad") or ik ("ik") differentiation (PPPM, src/workflow/Electrostatics.jl:64-100)."""
    if differentiation not in ("ad", "ik"):
        raise L.ArgumentError("differentiation must be :ad or :ik")
    return ElectrostaticsSpec(coulomb_prefactor=coulomb_prefactor, real_cutoff=real_cutoff, alpha=alpha,
                              real_space_tolerance=real_space_tolerance, order=order, mesh=mesh,
                              mesh_spacing=mesh_spacing, relative_permittivity=relative_permittivity,
                              net_charge=net_charge, real_space_modifier=real_space_modifier,
                              influence="p3m_" + differentiation, alias_order=alias_order)


def PME(*, coulomb_prefactor, real_cutoff, alpha=None, real_space_tolerance=None, mesh=None, mesh_spacing=None,
        order=4, relative_permittivity=1.0, net_charge="error", real_space_modifier="none"):
    """Workflow descriptor of smooth PME (src/workflow/Electrostatics.jl:45-63)."""
    return ElectrostaticsSpec(coulomb_prefactor=coulomb_prefactor, real_cutoff=real_cutoff, alpha=alpha,
                              real_space_tolerance=real_space_tolerance, order=order, mesh=mesh,
                              mesh_spacing=mesh_spacing, relative_permittivity=relative_permittivity,
                              net_charge=net_charge, real_space_modifier=real_space_modifier)


def esp_rule(tolerance, order=None):
    """(split shape c, window width P, window shape c1) of the ESP rules for a target relative force
    error in an FP32 state (``pdx_esp_rule``; src/workflow/Electrostatics.jl:177-215,288-301)."""
    out = (ctypes.c_double * 3)()
    L.check(L.lib.pdx_esp_rule(float(tolerance), int(order or 0), out))
    return float(out[0]), int(out[1]), float(out[2])


def ESP(*, coulomb_prefactor, real_cutoff, tolerance=1e-4, split_shape=None, window_shape=None, order=None,
        influence="optimal", alias_order=2, mesh=None, mesh_spacing=None, relative_permittivity=1.0,
        net_charge="error", real_space_modifier="none"):
    """Ewald summation with prolates (the workflow descriptor ESP, src/workflow/Electrostatics.jl:
    102-154, 288-301): a prolate split of 1/r whose real-space part vanishes at ``real_cutoff`` and
    a prolate assignment window of ``order`` grid points. Defaults from ``tolerance``: the split
    shape c with psi(1)/||psi|| = tolerance, the width of Table 2 of Liang et al., the window shape
    c(tolerance/2) (raised where an FP32 mesh would not resolve it) and the spacing pi rc / c.
    ``influence``: "optimal" (AD-optimal, alias sums over |n_d| <= alias_order) or "deconvolution"."""
    if split_shape is None or window_shape is None or order is None:
        if tolerance is None:
            raise L.ArgumentError("ESP needs `tolerance`, or explicit `split_shape`, `window_shape` and `order`.")
        c, P, c1 = esp_rule(tolerance, order)
        split_shape = c if split_shape is None else split_shape
        order = P if order is None else order
        window_shape = c1 if window_shape is None else window_shape
    if mesh is None and mesh_spacing is None:
        mesh_spacing = math.pi * float(real_cutoff) / float(split_shape)
    if influence not in ("optimal", "deconvolution"):
        raise L.ArgumentError(f"ESP influence must be :optimal or :deconvolution; got {influence}.")
    return ElectrostaticsSpec(coulomb_prefactor=coulomb_prefactor, real_cutoff=real_cutoff, order=int(order), mesh=mesh,
                              mesh_spacing=mesh_spacing, relative_permittivity=relative_permittivity,
                              net_charge=net_charge, real_space_modifier=real_space_modifier,
                              influence="esp" if influence == "optimal" else "esp_deconvolution",
                              alias_order=alias_order, split_shape=float(split_shape),
                              window_shape=float(window_shape))


class ElectrostaticsSolver:
    def __init__(self, spec: ElectrostaticsSpec, box, charges):
        """``charges``: host sequence in storage order (set_solver_charges!, Solver.jl:166-185)."""
        import numpy as np
        self.spec = spec
        q = np.ascontiguousarray(np.asarray(charges, dtype=np.float32))
        self.N = int(q.size)
        h = C.c_void_p()
        cs = spec.c_spec()
        L.check(L.lib.pdx_mesh_create(C.byref(cs), self.N, L.dbl3(box), C.byref(h)))
        self._h = h
        self._release = L.adopt(self, h, L.lib.pdx_mesh_destroy)
        self.box = tuple(float(b) for b in box)
        self.set_charges(q)
        info = (C.c_double * 9)()
        L.check(L.lib.pdx_mesh_info(self._h, info))
        self.mesh = (int(info[0]), int(info[1]), int(info[2]))
        self.order = int(info[3])
        self.alpha = info[4] if not spec.is_esp else None
        self.C = info[5]
        self.self_coefficient = info[6]
        self.plane_lines = int(info[8])
        self.spread_threads, self.transform_threads = 512, 1024


    @property
    def handle(self):
        return self._h

    def set_charges(self, charges):
        import numpy as np
        q = np.ascontiguousarray(np.asarray(charges, dtype=np.float32))
        L.check(L.lib.pdx_mesh_set_charges(self._h, q.ctypes.data_as(C.POINTER(C.c_float)), int(q.size)))
        self.charges = q
        self.qtotal = float(q.astype(np.float64).sum())
        self.qsq = float((q.astype(np.float64) ** 2).sum())

    def refresh_geometry(self, box):
        L.check(L.lib.pdx_mesh_set_box(self._h, L.dbl3(box)))
        self.box = tuple(float(b) for b in box)

    def real_space(self):
        """The matching short-range term for a ``PairForce``."""
        from .forces import RealSpaceCoulomb
        if self.spec.is_esp:
            return RealSpaceCoulomb(C=self.C, cutoff=self.spec.real_cutoff, alpha=0.0, split="prolate",
                                    shape=self.spec.split_shape)
        return RealSpaceCoulomb(C=self.C, cutoff=self.spec.real_cutoff, alpha=self.alpha,
                                shift=self.spec.real_space_modifier == "potential_shift")

    # ---- enqueue
    def spread(self, x, y, z, q, order=None, stream=None):
        L.check(L.lib.pdx_mesh_spread(self._h, L.ptr(x), L.ptr(y), L.ptr(z), L.ptr(q), L.ptr(order),
                                      L.stream_handle(stream)))

    def spread_cells(self, nl, stream=None):
        """Owner-computes spread from ``nl``'s cells and packed (x, y, z, q) (no atomics)."""
        L.check(L.lib.pdx_mesh_spread_cells(self._h, nl.handle, L.stream_handle(stream)))

    def set_spread_threads(self, threads):
        """Workgroup size of the owner-computes spreads that take every particle reaching a tile
        (sparse meshes; the dense meshes' own-particle form picks its own): 512, 1024 or 256."""
        L.check(L.lib.pdx_mesh_set_spread_threads(self._h, int(threads)))
        self.spread_threads = int(threads)

    def set_transform_threads(self, threads):
        """Workgroup size of the plane passes of the solve (1024 or 512): the same potential,
        the FP64 energy and virial to FP64 rounding."""
        L.check(L.lib.pdx_mesh_set_transform_threads(self._h, int(threads)))
        self.transform_threads = int(threads)

    def spread_snapshot(self, nl, stream=None):
        """Owner-computes spread from ``nl``'s cell snapshot (may run beside ``nl.build``)."""
        L.check(L.lib.pdx_mesh_spread_snapshot(self._h, nl.handle, L.stream_handle(stream)))

    def solve(self, energy=True, stream=None):
        L.check(L.lib.pdx_mesh_solve(self._h, L.PDX_ENERGY if energy else 0, L.stream_handle(stream)))

    def mesh_transform(self, x, y, z, q, energy=True, order=None, stream=None):
        self.spread(x, y, z, q, order, stream)
        self.solve(energy, stream)

    def mesh_gather(self, x, y, z, q, fx, fy, fz, Epot=None, Vglobal=None, energy=True, order=None, set=False,
                    stream=None):
        L.check(L.lib.pdx_mesh_gather(self._h, L.ptr(x), L.ptr(y), L.ptr(z), L.ptr(q), L.ptr(order), L.ptr(fx),
                                      L.ptr(fy), L.ptr(fz), L.ptr(Epot), L.ptr(Vglobal),
                                      L.PDX_ENERGY if energy else 0, 1 if set else 0, L.stream_handle(stream)))

    def gather_cells(self, nl, fx, fy, fz, Epot=None, Vglobal=None, energy=True, set=False, stream=None):
        """``mesh_gather`` at the positions of the last ``spread_cells`` / ``spread_snapshot`` from
        ``nl``: one workgroup per mesh tile gathers the particles that spread listed under it
        from the tile's potential staged in LDS (smooth PME and P3M-ad)."""
        L.check(L.lib.pdx_mesh_gather_cells(self._h, nl.handle, L.ptr(fx), L.ptr(fy), L.ptr(fz), L.ptr(Epot),
                                            L.ptr(Vglobal), L.PDX_ENERGY if energy else 0, 1 if set else 0,
                                            L.stream_handle(stream)))

    def compute_mesh(self, x, y, z, q, fx, fy, fz, Epot=None, Vglobal=None, energy=True, order=None, set=False,
                     stream=None):
        self.mesh_transform(x, y, z, q, energy, order, stream)
        self.mesh_gather(x, y, z, q, fx, fy, fz, Epot, Vglobal, energy, order, set, stream)

    # ---- boundary readbacks
    def totals(self, stream=None):
        out = (C.c_double * 9)()
        L.check(L.lib.pdx_mesh_totals(self._h, out, L.stream_handle(stream)))
        return dict(mesh_energy=out[0], virial=list(out[1:7]), self_energy=out[7], background_energy=out[8])

    def self_energy(self):
        """-(C/2) self_coefficient sum q^2 (self_energy, Solver.jl:421-423; the coefficient
        2 alpha/sqrt(pi) of the Gaussian split or chi(0)/rc of the prolate one)."""
        return -0.5 * self.C * self.self_coefficient * self.qsq

    def background_energy(self):
        """-C background_coefficient Q^2 / V with the neutralising background, else 0."""
        if self.spec.net_charge != "background":
            return 0.0
        info = (C.c_double * 9)()
        L.check(L.lib.pdx_mesh_info(self._h, info))
        return info[7] * self.qtotal

    def fixed_grid(self, stream=None):
        import torch
        K1, K2, K3 = self.mesh
        out = torch.empty(K3, K2, K1, dtype=torch.float64, device="cuda")
        L.check(L.lib.pdx_mesh_copy_grid(self._h, 0, L.ptr(out), L.stream_handle(stream)))
        return out

    def potential_grid(self, stream=None):
        import torch
        K1, K2, K3 = self.mesh
        out = torch.empty(K3, K2, K1, dtype=torch.float32, device="cuda")
        L.check(L.lib.pdx_mesh_copy_grid(self._h, 1, L.ptr(out), L.stream_handle(stream)))
        return out

    def fft_r2c(self, grid, stream=None):
        """R2C of a float grid shaped (K3, K2, K1) -> complex64 (K3, K2, K1//2+1)."""
        import torch
        K1, K2, K3 = self.mesh
        out = torch.empty(K3, K2, K1 // 2 + 1, dtype=torch.complex64, device="cuda")
        L.check(L.lib.pdx_mesh_fft(self._h, -1, L.ptr(grid.contiguous()), L.ptr(out), L.stream_handle(stream)))
        return out

    def fft_c2r(self, spec, stream=None):
        import torch
        K1, K2, K3 = self.mesh
        out = torch.empty(K3, K2, K1, dtype=torch.float32, device="cuda")
        L.check(L.lib.pdx_mesh_fft(self._h, 1, L.ptr(spec.contiguous()), L.ptr(out), L.stream_handle(stream)))
        return out


def ewald_fp64(x, y, z, q, box, C, alpha, digits=14.0, background=False, sample=None, want_energy=True,
               stream=None, F=None, no_alloc=False):
    """FP64 Ewald reference on the device (``pdx_ewald_fp64``). x..q: float64 device tensors;
    sample: int32 device tensor of 0-based ids; F: optional float64 device tensor [nsample, 3]
    for the forces. no_alloc: call ``pdx_ewald_fp64_noalloc``, which takes its workspace from a
    static device array instead of allocating (pass F to keep the whole call allocation-free).
    Returns (F[nsample,3] tensor or None, totals dict)."""
    import torch
    n = x.numel()
    ns = 0 if sample is None else sample.numel()
    if F is None:
        F = torch.zeros(max(ns, 1), 3, dtype=torch.float64, device="cuda")
    elif ns and (F.dtype != torch.float64 or not F.is_contiguous() or F.numel() < 3 * ns):
        raise L.DimensionMismatch("F must be a contiguous float64 tensor of at least nsample x 3")
    out = (ctypes.c_double * 7)()
    fn = L.lib.pdx_ewald_fp64_noalloc if no_alloc else L.lib.pdx_ewald_fp64
    L.check(fn(n, L.ptr(x), L.ptr(y), L.ptr(z), L.ptr(q), L.dbl3(box), float(C), float(alpha),
                                 float(digits), 1 if background else 0, ns, L.ptr(sample), L.ptr(F),
                                 1 if want_energy else 0, out, L.stream_handle(stream)))
    tot = dict(energy=out[0], virial=list(out[1:7])) if want_energy else None
    return (F[:ns] if ns else None), tot
