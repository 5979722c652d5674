"""ctypes wrapper of the FP64 CPU oracle (oracle.c) — TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg, as the
checker. See oracle.c's header for the reference file each function restates and for the
checks that pin it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER
    d, i = C.c_double, C.c_int
    lib.or_neighbor_rows.argtypes = [i, i, P(d), P(d), P(d), P(d), P(d), i, P(i), P(i), P(d)]
    lib.or_pair_direct.argtypes = [i, P(d), P(d), P(d), P(d), i, d, d, d, i, P(i), P(d)]
    lib.or_pair_typed.argtypes = [i, i, P(d), P(d), P(d), P(d), i, i, d, d, d, P(d), d, P(i), i, P(d), P(d), P(d),
                                  i, P(i), P(d)]
    lib.or_erfc.argtypes = [d]
    lib.or_erfc.restype = d
    lib.or_ewald.argtypes = [i, P(d), P(d), P(d), d, d, d, i, i, P(i), P(i), P(d), P(d), P(d), P(d)]
    lib.or_ewald_sampled.argtypes = [i, P(d), P(d), P(d), P(d), P(d), d, d, d, i, P(i), P(d)]
    lib.or_bspline.argtypes = [i, d, P(d), P(d)]
    lib.or_bspline.restype = i
    lib.or_cell_pairs.argtypes = [i, P(d), P(d), P(d), P(d), P(d), i, d, d, d, i, d, d, d, P(d), P(d), P(d)]
    lib.or_cell_pairs.restype = C.c_long
    lib.or_mesh_spread.argtypes = [i, P(d), P(d), P(d), P(d), P(d), P(i), i, P(d)]
    lib.or_mesh_gather.argtypes = [i, P(d), P(d), P(d), P(d), P(d), P(i), i, P(d), P(d), P(d)]
    lib.or_kg_forces.argtypes = [i, P(d), P(d), P(d), P(d), d, d, i, P(i), P(i), d, d, P(d), P(d), P(d), P(d), i]
    lib.or_kg_forces.restype = C.c_long
    lib.or_num_threads.restype = i
    lib.or_set_threads.argtypes = [i]
    lib.or_set_threads.restype = i
    return lib


_lib = _load()


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(C.POINTER(C.c_int))


def num_threads():
    return _lib.or_num_threads()


def set_threads(n):
    """OpenMP threads of later oracle calls; returns the previous count."""
    return _lib.or_set_threads(int(n))


def erfc(x):
    return _lib.or_erfc(float(x))


def neighbor_rows(coords, box, radii, cap=None):
    """Rows (lists of 0-based ids, ascending) by brute force (test/neighbor_reference.jl:6-16).
    coords: tuple of D arrays; radii: per particle. Returns (rows, counts, margin)."""
    D = len(coords)
    n = len(coords[0])
    cs = [_d(c) for c in coords] + [_d(np.zeros(1))] * (3 - D)
    bx = _d(np.asarray(list(box) + [1.0] * (3 - D)))
    rr = _d(radii)
    cap = n if cap is None else cap
    counts = np.zeros(n, dtype=np.int32)
    rows = np.zeros(n * max(cap, 1), dtype=np.int32)
    margin = C.c_double()
    _lib.or_neighbor_rows(n, D, cs[0][1], cs[1][1], cs[2][1], bx[1], rr[1], cap,
                          counts.ctypes.data_as(C.POINTER(C.c_int)), rows.ctypes.data_as(C.POINTER(C.c_int)),
                          C.byref(margin))
    out = [list(rows[i * cap:i * cap + min(counts[i], cap)]) for i in range(n)]
    return out, counts, margin.value


def pair_direct(x, y, z, box, kind="lj", sigma=1.0, epsilon=1.0, cutoff=2.5, ids=None):
    """FP64 direct pair sums at the sampled ids. Returns dict F (n,3), energy (n), virial (n,6), counts."""
    k = {"lj": 1, "wca": 2, "soft": 3}[kind]
    n = len(x)
    ids = np.arange(n) if ids is None else np.asarray(ids)
    xa, ya, za, bx = _d(x), _d(y), _d(z), _d(box)
    ia = _i(ids)
    out = np.zeros((len(ids), 12))
    _lib.or_pair_direct(n, xa[1], ya[1], za[1], bx[1], k, float(sigma), float(epsilon), float(cutoff), len(ids),
                        ia[1], out.ctypes.data_as(C.POINTER(C.c_double)))
    return dict(F=out[:, 0:3], energy=out[:, 3], virial=out[:, 4:10], pair_scale=out[:, 10], counts=out[:, 11])


def pair_typed(coords, box, kind="lj", style="uniform", sigma=1.0, epsilon=1.0, cutoff=2.5, sigma_particle=None,
               factor=1.0, typeid=None, sigma_pair=None, epsilon_pair=None, rcut_pair=None, ids=None):
    """FP64 direct sums of the typed pair models (test/pair_reference_gpu.jl:8-64). coords: (x, y)
    or (x, y, z); style "uniform", "mixed" (per-particle sigma, rc = factor sigma_ij) or "matrix"
    (1-based typeid, ntypes x ntypes sigma / epsilon / rcut). Returns dict F (n, D), energy (n),
    virial (n, 6)."""
    D = len(coords)
    n = len(coords[0])
    k = {"lj": 1, "wca": 2, "soft": 3}[kind]
    st = {"uniform": 1, "mixed": 2, "matrix": 3}[style]
    ids = np.arange(n) if ids is None else np.asarray(ids)
    cs = [_d(c) for c in coords] + [_d(np.zeros(n))] * (3 - D)
    bx = _d(list(box)[:D] + [1.0] * (3 - D))
    sp = _d(sigma_particle if sigma_particle is not None else np.ones(1))
    ti = _i(typeid if typeid is not None else np.ones(1, dtype=np.int32))
    nt = 1 if sigma_pair is None else np.asarray(sigma_pair).shape[0]
    m = [_d(np.asarray(a, dtype=np.float64).T.reshape(-1)) if a is not None else _d(np.ones(1))
         for a in (sigma_pair, epsilon_pair, rcut_pair)]  # column-major
    ia = _i(ids)
    out = np.zeros((len(ids), 10))
    rc = _lib.or_pair_typed(n, D, cs[0][1], cs[1][1], cs[2][1], bx[1], k, st, float(sigma), float(epsilon),
                            float(cutoff), sp[1], float(factor), ti[1], int(nt), m[0][1], m[1][1], m[2][1], len(ids),
                            ia[1], out.ctypes.data_as(C.POINTER(C.c_double)))
    if rc != 0:
        raise ValueError("type id out of range")
    return dict(F=out[:, :D], energy=out[:, 3], virial=out[:, 4:10])


def ewald(R, q, H, C_=1.0, alpha=None, digits=16.0, background=False, exceptions=()):
    """FP64 Ewald (test/ewald_reference.jl:50-126). R: (n,3); H: 3x3 cell (columns = box vectors).
    Returns (E, F (n,3), W 3x3)."""
    R = np.asarray(R, dtype=np.float64)
    n = R.shape[0]
    Ra = _d(R.reshape(-1))
    qa = _d(q)
    Ha = _d(np.asarray(H, dtype=np.float64).T.reshape(-1))  # column-major
    ne = len(exceptions)
    ei = _i([e[0] for e in exceptions] or [0])
    ej = _i([e[1] for e in exceptions] or [0])
    es = _d([(e[2] if len(e) == 3 else 0.0) for e in exceptions] or [0.0])
    E = C.c_double()
    F = np.zeros(3 * n)
    W = np.zeros(9)
    _lib.or_ewald(n, Ra[1], qa[1], Ha[1], float(C_), float(alpha) if alpha else 0.0, float(digits),
                  1 if background else 0, ne, ei[1], ej[1], es[1], C.byref(E),
                  F.ctypes.data_as(C.POINTER(C.c_double)), W.ctypes.data_as(C.POINTER(C.c_double)))
    return E.value, F.reshape(n, 3), W.reshape(3, 3).T


def ewald_sampled(x, y, z, q, box, C_=1.0, alpha=0.35, digits=14.0, ids=None):
    """FP64 Ewald forces on the sampled ids of an orthorhombic box. Returns F (ns, 3)."""
    n = len(x)
    ids = np.arange(n) if ids is None else np.asarray(ids)
    xa, ya, za, qa, bx = _d(x), _d(y), _d(z), _d(q), _d(box)
    ia = _i(ids)
    F = np.zeros(3 * len(ids))
    _lib.or_ewald_sampled(n, xa[1], ya[1], za[1], qa[1], bx[1], float(C_), float(alpha), float(digits), len(ids),
                          ia[1], F.ctypes.data_as(C.POINTER(C.c_double)))
    return F.reshape(-1, 3)


def bspline(P, u):
    w = np.zeros(P)
    dw = np.zeros(P)
    k0 = _lib.or_bspline(int(P), float(u), w.ctypes.data_as(C.POINTER(C.c_double)),
                         dw.ctypes.data_as(C.POINTER(C.c_double)))
    return k0, w, dw


def cell_pairs(x, y, z, q, box, kind="none", sigma=1.0, epsilon=1.0, cutoff=2.5, shift=False, C_=0.0, alpha=0.0,
               rc_coul=0.0):
    """FP64 cell-list pair evaluation: LJ / WCA (kind "lj" / "wca" / "none") plus the real-space
    Ewald term C q_i q_j erfc(alpha r)/r inside rc_coul when C_ != 0. Returns dict F (n,3),
    E (n, half energies), V (n,6 half virials), pairs (each counted once)."""
    k = {"none": 0, "lj": 1, "wca": 2}[kind]
    n = len(x)
    xa, ya, za, bx = _d(x), _d(y), _d(z), _d(box)
    qa = _d(q if q is not None else np.zeros(n))
    F = np.zeros(3 * n)
    E = np.zeros(n)
    V = np.zeros(6 * n)
    P = C.POINTER(C.c_double)
    pairs = _lib.or_cell_pairs(n, xa[1], ya[1], za[1], qa[1], bx[1], k, float(sigma), float(epsilon), float(cutoff),
                               1 if shift else 0, float(C_), float(alpha), float(rc_coul), F.ctypes.data_as(P),
                               E.ctypes.data_as(P), V.ctypes.data_as(P))
    return dict(F=F.reshape(n, 3), E=E, V=V.reshape(n, 6), pairs=int(pairs))


def vv_steps(x, y, z, vx, vy, vz, box, kind="lj", sigma=1.0, epsilon=1.0, cutoff=2.5, dt=0.005, mass=1.0,
             nsteps=1, shift=False):
    """FP64 velocity Verlet of one box with the direct pair forces: per step v += c F,
    x += dt v (wrapped to [-L/2, L/2)), F(x), v += c F, c = dt / (2 mass)
    (_nve_kick_drift3_kernel!, src/simulation/NVEFusion.jl:3-24; _nve_kick3_kernel!,
    src/SimulationCore.jl:581-591). Returns x, v (n, 3), F (n, 3), potential energy (LJ
    shifted to zero at the cutoff when ``shift``), kinetic energy, sum r.F over pairs.
    Pinned in tests/test_oracle.py by a dimer against a converged ODE solution, time
    reversibility and energy conservation."""
    L = np.asarray(box, dtype=np.float64)
    X = np.stack([np.asarray(a, dtype=np.float64) for a in (x, y, z)], 1)
    V = np.stack([np.asarray(a, dtype=np.float64) for a in (vx, vy, vz)], 1)
    c = 0.5 * dt / mass

    def forces(X):
        d = pair_direct(X[:, 0], X[:, 1], X[:, 2], L, kind=kind, sigma=sigma, epsilon=epsilon, cutoff=cutoff)
        return d

    d = forces(X)
    for _ in range(nsteps):
        V = V + c * d["F"]
        X = X + dt * V
        X = X - L * (X >= 0.5 * L) + L * (X < -0.5 * L)
        d = forces(X)
        V = V + c * d["F"]
    pe = d["energy"].sum()
    if kind == "lj" and shift:
        s6 = (sigma / cutoff) ** 6
        pe -= 0.5 * 4 * epsilon * (s6 * s6 - s6) * d["counts"].sum()
    ke = 0.5 * mass * (V ** 2).sum()
    w = d["virial"][:, :3].sum()
    return X, V, d["F"], pe, ke, w


def kg_forces(x, y, z, box, bonds, sigma=1.0, epsilon=1.0, k=30.0, R0=1.5, exclude_bonded=False):
    """FP64 Kremer-Grest forces: WCA over all pairs (bonded ones included) plus FENE bonds.
    bonds: (nb, 2) int array, 0-based. Returns dict F (n,3), E (n), energy, W (6), pairs;
    raises ValueError naming the first bond with r >= R0."""
    n = len(x)
    xa, ya, za, bx = _d(x), _d(y), _d(z), _d(box)
    b = np.ascontiguousarray(bonds, dtype=np.int32).reshape(-1, 2)
    bi, bj = _i(b[:, 0].copy()), _i(b[:, 1].copy())
    F = np.zeros(3 * n)
    E = np.zeros(n)
    W = np.zeros(6)
    e = C.c_double()
    P = C.POINTER(C.c_double)
    r = _lib.or_kg_forces(n, xa[1], ya[1], za[1], bx[1], float(sigma), float(epsilon), len(b), bi[1], bj[1],
                          float(k), float(R0), F.ctypes.data_as(P), E.ctypes.data_as(P), C.byref(e),
                          W.ctypes.data_as(P), 1 if exclude_bonded else 0)
    if r < 0:
        raise ValueError(f"FENE bond {-1 - r} has r >= R0")
    return dict(F=F.reshape(n, 3), E=E, energy=e.value, W=W, pairs=int(r))
