"""One force evaluation on the device: list check + predicated rebuild, pair forces (LJ/WCA
and real-space Coulomb), and the PME mesh — the work of ``evaluate_forces_into_f!``
(src/simulation/ForceEvaluation.jl:40-76) on the reference's ELL path (``Run.enqueue_force!``
+ ``Run.enqueue_ell_mesh!``, src/engine/Run.jl:179-238)."""
from __future__ import annotations

import math

from . import _lib as L
from .electrostatics import ElectrostaticsSolver
from .forces import PairForce
from .neighbors import DeviceNeighborWorkspace
from .clusters import ClusterPairEngine


def default_capacity(N, box, rlist):
    """ceil(1.5 (4 pi/3) rlist^3 rho + 64) (src/engine/StateEngine.jl:463); in 2-D (two box
    lengths) the disc, ceil(1.5 pi rlist^2 rho + 64)."""
    if len(box) == 2:
        return int(math.ceil(1.5 * math.pi * rlist ** 2 * N / (box[0] * box[1]) + 64))
    rho = N / (box[0] * box[1] * box[2])
    return int(math.ceil(1.5 * (4.0 * math.pi / 3.0) * rlist ** 3 * rho + 64))


class ForceEngine:
    def __init__(self, N, box, pair_model=None, electrostatics=None, charges=None, skin=0.4, cap=None,
                 interval=20, lanes=0, diagnostics=False, spread="snapshot", correction=None, accumulation="fp32",
                 bonds=None, exceptions=(), traversal="rows", backend="ell", prune_buffer=None, cluster_capacity=0,
                 overlap_with="fft", gather="tiles", exclude_bonded=False, settings=None):
        """backend "ell": the cell list and ELL rows (``pdx_nl_*`` + ``pdx_pair_forces``); "cluster":
        the cluster-pair list and tiles (``pdx_cluster_*``, the reference's ``ClusterBackend``,
        src/engine/ClusterBackend.jl:1-40); its mesh spread takes the cluster order."""
        import torch
        if overlap_with not in ("spread", "solve", "fft"):
            raise L.ArgumentError("overlap_with must be 'spread', 'solve' or 'fft'")
        # "spread": the pair forces run beside the whole mesh branch (spread and solve);
        # "solve": the spread runs first on the main stream, the pairs beside the solve only;
        # "fft": the spread runs on the mesh stream beside the rebuild, the pairs wait for it and
        # run beside the transforms
        self.overlap_with = overlap_with
        if backend not in ("ell", "cluster"):
            raise L.ArgumentError("backend must be 'ell' or 'cluster'")
        self.backend = backend
        self.N = int(N)
        self.box = tuple(float(b) for b in box)
        if len(self.box) == 2 and (electrostatics is not None or backend != "ell"):
            raise L.ArgumentError("2-D systems (two box lengths) run on the ELL rows without a mesh")
        self.solver = ElectrostaticsSolver(electrostatics, box, charges) if electrostatics is not None else None
        rc = 0.0
        if pair_model is not None:
            rc = max(rc, float(pair_model.cutoff))
        if self.solver is not None:
            rc = max(rc, float(electrostatics.real_cutoff))
        self.cutoff = rc
        self.skin = float(skin)
        self.pairs = PairForce(pair_model, coulomb=self.solver.real_space() if self.solver else None, lanes=lanes,
                               accumulation=accumulation, bonds=bonds, traversal=traversal)
        if exclude_bonded:
            if bonds is None:
                raise L.ArgumentError("exclude_bonded needs bonds")
            self.pairs.set_bond_exclusion(True)
        if backend == "cluster":
            self.nl = None
            self.cluster = ClusterPairEngine(N, box, self.pairs, skin, prune_buffer, capacity=cluster_capacity)
            spread = "atomic"  # the owner-computes spread walks the ELL list's cells
        else:
            cap = cap if cap is not None else default_capacity(N, box, rc + skin)
            self.nl = DeviceNeighborWorkspace(N, box, rc, skin, cap, interval=interval, diagnostics=diagnostics)
            self.cluster = None
        self._exceptions = list(exceptions)
        self._bond_pairs = None if bonds is None else bonds.pairs
        if exceptions:
            if self.solver is None:
                raise L.ArgumentError("electrostatic exceptions need a mesh")
            self.pairs.set_exceptions(self.N, exceptions)
        self.pair_totals = torch.zeros(7, dtype=torch.float64, device="cuda")
        self.Vglobal = torch.zeros(6, dtype=torch.float32, device="cuda")
        # the mesh branch runs on its own stream (default priority), forked after the list
        # check and joined before the gather adds the mesh forces to the pair forces
        self.mesh_stream = torch.cuda.Stream() if self.solver is not None else None
        if self.solver is not None:
            # a spread that shares the CUs with the pair kernel runs better in smaller workgroups
            self.solver.set_spread_threads(256 if overlap_with == "spread" else 512)
            # the pair kernel runs beside the transforms: they take 512-thread workgroups, which
            # share the CUs with it better than 1024-thread ones (slower alone, faster together)
            self.solver.set_transform_threads(512)
        # "snapshot": the owner-computes spread over the cell snapshot of the last completed
        # build, forked right after the list check so that it runs beside the rebuild; "cells":
        # the same spread over the current cells, after the whole list update; "atomic": the
        # global-atomic spread in the list's order
        if spread not in ("snapshot", "cells", "atomic"):
            raise L.ArgumentError("spread must be 'snapshot', 'cells' or 'atomic'")
        self.spread_mode = spread
        # "tiles": after an owner-computes spread, the gather runs tile by tile from the particle
        # lists that spread wrote (smooth PME / P3M-ad); "particles": the per-particle gather
        if gather not in ("tiles", "particles"):
            raise L.ArgumentError("gather must be 'tiles' or 'particles'")
        self._tile_gather = (gather == "tiles" and electrostatics is not None
                             and getattr(electrostatics, "influence", "spme") != "p3m_ik")
        # learned force correction of a coarse mesh (an MLP of the charge-grid features, formed
        # from the grid inside the network's kernel): it runs on the main stream once the spread
        # is done, beside the transforms, and adds its output to the forces; energies and
        # virials are not corrected
        self.correction = correction
        if correction is not None and self.solver is None:
            raise L.ArgumentError("a learned correction needs a mesh")
        self._graph = None
        self._snap_ready = False
        # launch shapes (pdx.autotune): a setting dict, or "auto" for the measured table's setting
        # of this system's class (none: the defaults above)
        if settings is not None:
            from . import autotune
            if settings == "auto":
                settings = autotune.settings_for(self.N, self.box, self.cutoff, self.solver is not None)
            if settings:
                autotune.apply(self, settings)

    def evaluate(self, x, y, z, q, fx, fy, fz, Epot=None, V=None, energy=True, stream=None, overlap=True,
                 kick_drift=None):
        """Enqueue one force evaluation (no host synchronisation). ``kick_drift``: dict(vx, vy, vz,
        dt, mass, inv_mass) moves x, y, z first by the velocity-Verlet kick-drift with the forces
        fx, fy, fz hold (fused into the list check), as ``step`` does."""
        import torch
        main = stream if stream is not None else torch.cuda.current_stream()
        early = (self.cluster is None and self.solver is not None and overlap and self.spread_mode == "snapshot"
                 and self.overlap_with in ("spread", "fft"))
        if kick_drift is not None and self.cluster is not None:
            raise L.ArgumentError("the fused velocity-Verlet step runs on the ELL backend")

        def check():
            if kick_drift is None:
                self.nl.check(x, y, z, q, stream=main)
            else:
                kd = kick_drift
                self.nl.check_kick_drift(x, y, z, q, kd["vx"], kd["vy"], kd["vz"], fx, fy, fz, kd["dt"],
                                         kd.get("mass", 1.0), kd.get("inv_mass"), stream=main)

        if self.cluster is not None:
            self.cluster.update(x, y, z, q, stream=main)
            order = self.cluster.order_ptr
        elif early and self._snap_ready:
            check()  # the rebuild kernels follow after the fork
            order = self.nl._arrays.particle_ids_sorted
        elif kick_drift is not None:
            check()
            self.nl.build(stream=main)
            if early:
                self._snap_ready = True
                early = False
            order = self.nl._arrays.particle_ids_sorted
        else:
            self.nl.update(x, y, z, q, stream=main)
            if early:
                # the next check snapshots this first build; this evaluation spreads from the
                # current cells after the whole update
                self._snap_ready = True
                early = False
            order = self.nl._arrays.particle_ids_sorted
        if self.solver is None:
            self._pairs(x, y, z, q, fx, fy, fz, Epot, V, energy, main)
            return
        side = self.mesh_stream if overlap else main
        spread_stream = main if self.overlap_with == "solve" else side
        if overlap and spread_stream is side:
            side.wait_stream(main)
        if early:
            self.solver.spread_snapshot(self.nl, stream=spread_stream)
        elif self.spread_mode in ("cells", "snapshot") and self.cluster is None:
            self.solver.spread_cells(self.nl, stream=spread_stream)
        else:
            self.solver.spread(x, y, z, q, order, stream=spread_stream)
        if early:
            self.nl.build(stream=main)
        if overlap and spread_stream is main:
            side.wait_stream(main)
        if overlap and self.overlap_with == "fft":
            spread_done = torch.cuda.Event()
            spread_done.record(side)
            main.wait_event(spread_done)
        self.solver.solve(energy, stream=side)
        self._pairs(x, y, z, q, fx, fy, fz, Epot, V, energy, main)
        # the main stream has the spread behind it when it ran the spread or waited for it
        spread_seen = spread_stream is main or (overlap and self.overlap_with == "fft")
        if self.correction is not None and spread_seen:
            self.correction.forward_mesh(self.solver, x, y, z, q, fx=fx, fy=fy, fz=fz, stream=main)
        if overlap:
            main.wait_stream(side)
        if self.correction is not None and overlap and not spread_seen:
            self.correction.forward_mesh(self.solver, x, y, z, q, fx=fx, fy=fy, fz=fz, stream=main)
        if self._tile_gather and self.cluster is None and self.spread_mode in ("cells", "snapshot"):
            # the owner-computes spread listed the particles per tile: gather them tile by tile
            self.solver.gather_cells(self.nl, fx, fy, fz, Epot, self.Vglobal if energy else None, energy,
                                     set=False, stream=main)
        else:
            self.solver.mesh_gather(x, y, z, q, fx, fy, fz, Epot, self.Vglobal if energy else None, energy,
                                    order=order, set=False, stream=main)

    def _pairs(self, x, y, z, q, fx, fy, fz, Epot, V, energy, stream):
        tot = self.pair_totals if energy else None
        if self.cluster is not None:
            self.cluster.forces(fx, fy, fz, Epot, V, tot, energy=energy, add=False, stream=stream)
        else:
            self.pairs.compute_at(self.nl, x, y, z, q, fx, fy, fz, Epot, V, tot, energy=energy, add=False,
                                  stream=stream)

    def step(self, x, y, z, q, vx, vy, vz, fx, fy, fz, dt, mass=1.0, inv_mass=None, Epot=None, V=None,
             energy=False, stream=None, overlap=True):
        """One velocity-Verlet step (the reference's NVE step, src/engine/Run.jl:247-294): kick-drift
        fused into the list check, the force evaluation of the new positions, the closing kick
        (which also advances the list's step counter). fx, fy, fz must hold the forces of the
        current positions (an ``evaluate`` first); they hold the new forces afterwards."""
        import torch
        main = stream if stream is not None else torch.cuda.current_stream()
        kd = dict(vx=vx, vy=vy, vz=vz, dt=dt, mass=mass, inv_mass=inv_mass)
        # (a learned correction is part of the forces the closing kick applies: evaluate adds it)
        self.evaluate(x, y, z, q, fx, fy, fz, Epot, V, energy, stream=main, overlap=overlap, kick_drift=kd)
        self.nl.kick(vx, vy, vz, fx, fy, fz, dt, mass, inv_mass, stream=main)

    def capture_step(self, x, y, z, q, vx, vy, vz, fx, fy, fz, dt, mass=1.0, inv_mass=None, Epot=None, V=None,
                     energy=False, overlap=True):
        """Capture one ``step`` into a hipGraph; ``replay()`` then advances one step per launch. The
        warm-up step before the capture advances the system by one step."""
        import torch
        self.step(x, y, z, q, vx, vy, vz, fx, fy, fz, dt, mass, inv_mass, Epot, V, energy, overlap=overlap)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.step(x, y, z, q, vx, vy, vz, fx, fy, fz, dt, mass, inv_mass, Epot, V, energy, overlap=overlap)
        torch.cuda.synchronize()
        self._graph = g
        return g

    def reorder(self, x, y, z, q=None, *extra, tags=None):
        """Spatial reordering (OrderingWorkspace / enqueue_ordering!, src/engine/Ordering.jl:19-79),
        at a host boundary: the particles are permuted into the spatial order of the last build
        (cells, then position inside a cell), so storage neighbours are space neighbours and the
        order no longer depends on how the particles were labelled. x, y, z, q and every tensor in
        ``extra`` (velocities, forces, per-particle arrays of N rows) are permuted in place, as are
        ``tags`` (slot -> original id, int tensor); the engine's own per-particle data (charges,
        typed pair data, bonds, exceptions) follows. Returns the permutation (slot -> old slot).
        The list rebuilds at the next evaluation."""
        import numpy as np
        import torch
        if self.cluster is not None:
            raise L.ArgumentError("spatial reordering runs on the ELL backend")
        torch.cuda.synchronize()
        self.nl.status(check=True)
        perm = self.nl.spatial_order()
        for t in (x, y, z, q) + tuple(extra) + ((tags,) if tags is not None else ()):
            if t is not None:
                t.copy_(t[perm])
        p = perm.cpu().numpy()
        inv = np.empty_like(p)
        inv[p] = np.arange(len(p))
        model = self.pairs.model
        from .forces import MixedSigmaLJ, PairMatrixLJ
        if isinstance(model, MixedSigmaLJ):
            model.sigmas = np.ascontiguousarray(model.sigmas[p])
            self.pairs.set_sigmas(model.sigmas)
        elif isinstance(model, PairMatrixLJ):
            model.types = np.ascontiguousarray(model.types[p])
            self.pairs._set_types(model)
        if self.pairs.bonds is not None:
            from .forces import FENEBonds
            b = self.pairs.bonds
            pairs = inv[self._bond_pairs]
            self._bond_pairs = pairs
            nb = FENEBonds(self.N, pairs, k=b.k, R0=b.R0)
            self.pairs.set_bonds(nb)
        if self._exceptions:
            self._exceptions = [(int(inv[i - 1]) + 1, int(inv[j - 1]) + 1) + tuple(r) for i, j, *r in self._exceptions]
            self.pairs.set_exceptions(self.N, self._exceptions)
        if self.solver is not None:
            self.solver.set_charges(self.solver.charges[p])
        self.nl.invalidate()
        self._snap_ready = False
        return perm

    def status(self, check=True, stream=None):
        """Boundary readback of the short-range list (raises on a latched capacity / coordinate error)."""
        if self.cluster is not None:
            return self.cluster.status(check=check, stream=stream)
        return self.nl.status(check=check, stream=stream)

    def capture(self, x, y, z, q, fx, fy, fz, Epot=None, V=None, energy=True, overlap=True):
        """Capture one evaluation into a hipGraph (torch.cuda.CUDAGraph) over these buffers;
        ``replay()`` then re-runs it with one launch. Every enqueue is capture-safe: no
        allocation, synchronisation or memset inside."""
        import torch
        self.evaluate(x, y, z, q, fx, fy, fz, Epot, V, energy, overlap=overlap)  # warm: workspaces sized
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.evaluate(x, y, z, q, fx, fy, fz, Epot, V, energy, overlap=overlap)
        torch.cuda.synchronize()
        self._graph = g
        return g

    def replay(self):
        self._graph.replay()

    def totals(self, stream=None):
        """FP64 energy and virial of the last energy evaluation (boundary readback)."""
        self.status(check=True, stream=stream)
        p = self.pair_totals.cpu().tolist()
        E = p[0]
        W = p[1:7]
        out = dict(pair_energy=p[0], pair_virial=p[1:7])
        if self.solver is not None:
            m = self.solver.totals(stream=stream)
            E += m["mesh_energy"] + m["background_energy"]
            W = [a + b for a, b in zip(W, m["virial"])]
            out.update(mesh_energy=m["mesh_energy"], mesh_virial=m["virial"], self_energy=m["self_energy"],
                       background_energy=m["background_energy"])
        out.update(energy=E, virial=W)
        return out
