"""Force-evaluation benchmark of the ionic rock-salt workload on one MI355X.

Workload (the configuration the headline metric is quoted on): 64 000 ions, 40^3 rock-salt
lattice, a = 2.82 A, charges +1/-1, Gaussian jitter 0.1 A; LJ eps = 0.01 eV, sigma = 2.8 A,
cutoff 9 A; smooth PME with C = 14.399645 eV A/e^2, real cutoff 9 A, erfc(alpha rc) = 1e-5,
mesh spacing 1.0 A (120^3), B-spline order 5; FP32 arithmetic, FP64 energy / virial sums.

One step = one force evaluation with energy and full virial: list coverage check (with the
predicated rebuild), LJ + real-space Coulomb pair kernel, spread, FFT + influence + inverse
FFT, gather. Inputs are resident in HBM before timing starts. Prints ONE JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "particledynamics.jl_amd"))

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
DT_FS = 1.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-error", action="store_true", help="skip the FP64 Ewald force-error measurement")
    p.add_argument("--lanes", type=int, default=0)
    p.add_argument("--no-clusters", action="store_true", help="skip the cluster-pair backend leg")
    p.add_argument("--skin", type=float, default=1.0)
    p.add_argument("--overlap-with", choices=("spread", "solve", "fft"), default="fft",
                   help="what the pair kernel overlaps: the whole mesh branch, the solve only, or the transforms only")
    p.add_argument("--spread", choices=("snapshot", "cells", "atomic"), default="snapshot")
    p.add_argument("--traversal", choices=("rows", "cells"), default="rows",
                   help="pair traversal: the list's ELL rows, or its cell grid without rows")
    p.add_argument("--no-learned", action="store_true", help="skip the learned-correction comparison")
    p.add_argument("--no-replicas", action="store_true", help="skip the batched-replica run")
    p.add_argument("--no-precision", action="store_true", help="skip the bead-spring accumulation study")
    p.add_argument("--no-published", action="store_true",
                   help="skip the Coulomb force calls at the reference's published methods and meshes")
    p.add_argument("--copy-calibration", action="store_true",
                   help="run only a plain 256 MiB device copy (five launches): the reference reading for the "
                        "FETCH_SIZE / WRITE_SIZE counters")
    p.add_argument("--mode", choices=("graph", "eager", "serial"), default="graph",
                   help="graph: hipGraph replay of the overlapped evaluation; eager: overlapped "
                        "streams without a graph; serial: one stream")
    return p.parse_args()


def pair_traffic():
    """HBM bytes per pair-kernel launch on this workload, as recorded by the counter summary
    in profiles/ (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by the reading of a
    plain copy of known size); None when that summary is absent."""
    path = os.path.join(ROOT, "profiles", "k_pair_rocksalt64k_hbm_traffic.txt")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch"):
                    return int(line.split()[-1])
    except OSError:
        pass
    return None


def pair_flops_per_visit(npe, npf, energy=True):
    """FP32 operations the pair kernel spends per visited row entry (each pair is visited
    from both ends): periodic differences and r^2 (8), LJ force (9), Coulomb force with the
    force polynomial (8 + 2 npf), force accumulation (6); with energy: LJ energy (4), Coulomb
    energy with its polynomial (4 + 2 npe), energy sum (1), six virial products (15)."""
    f = 8 + 9 + 8 + 2 * npf + 6
    if energy:
        f += 4 + 4 + 2 * npe + 1 + 15
    return f


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np
    import torch

    import pdx
    from pdx import systems

    torch.cuda.set_device(0)
    if args.copy_calibration:
        n = 64 << 20  # 64 Mi floats = 256 MiB read + 256 MiB written per launch
        src = torch.rand(n, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read_per_launch": 4 * n, "copy_bytes_written_per_launch": 4 * n,
                          "launches": 5}))
        return
    sysd = systems.rock_salt(40, 2.82, 0.1, seed=0)
    N, box = sysd["N"], sysd["box"]
    C = 14.399645
    spec = pdx.PME(coulomb_prefactor=C, real_cutoff=9.0, real_space_tolerance=1e-5, mesh_spacing=1.0, order=5)
    lj = pdx.LennardJones(epsilon=0.01, sigma=2.8, cutoff=9.0)
    eng = pdx.ForceEngine(N, box, lj, electrostatics=spec, charges=sysd["q"], skin=args.skin, lanes=args.lanes,
                          spread=args.spread, traversal=args.traversal, overlap_with=args.overlap_with)
    X, Y, Z, Q = (torch.from_numpy(sysd[k]).cuda() for k in ("x", "y", "z", "q"))
    fx, fy, fz, E = (torch.empty(N, device="cuda") for _ in range(4))
    V = torch.empty(6, N, device="cuda")
    stream = torch.cuda.current_stream()

    if args.mode == "graph":
        eng.capture(X, Y, Z, Q, fx, fy, fz, E, V, energy=True)

        def step():
            eng.replay()
    else:
        def step():
            eng.evaluate(X, Y, Z, Q, fx, fy, fz, E, V, energy=True, overlap=args.mode == "eager")

    for _ in range(max(args.warmup, 1)):
        step()
    torch.cuda.synchronize()
    st = eng.nl.status()

    # ---- timed region: exactly K steps
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms = 1e3 * dt / args.steps
    evals = args.steps / dt

    # ---- per-kernel durations with events on the launch stream (same work, split by phase)
    ev = {k: [] for k in ("update", "pairs", "spread", "solve", "gather")}
    nl, pf, sol = eng.nl, eng.pairs, eng.solver
    order = nl._arrays.particle_ids_sorted
    reps = max(10, min(args.steps, 50))
    for _ in range(reps):
        marks = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
        marks[0].record(stream)
        nl.update(X, Y, Z, Q)
        marks[1].record(stream)
        pf.compute(nl, fx, fy, fz, E, V, eng.pair_totals, energy=True)
        marks[2].record(stream)
        if eng.spread_mode in ("cells", "snapshot"):
            sol.spread_cells(nl)
        else:
            sol.spread(X, Y, Z, Q, order=order)
        marks[3].record(stream)
        sol.solve(energy=True)
        marks[4].record(stream)
        if eng._tile_gather and eng.spread_mode in ("cells", "snapshot"):
            sol.gather_cells(nl, fx, fy, fz, E, eng.Vglobal, energy=True)
        else:
            sol.mesh_gather(X, Y, Z, Q, fx, fy, fz, E, eng.Vglobal, energy=True, order=order)
        marks[5].record(stream)
        torch.cuda.synchronize()
        for k, (a, b) in zip(ev, zip(marks[:-1], marks[1:])):
            ev[k].append(a.elapsed_time(b))
    phase_ms = {k: float(np.mean(v)) for k, v in ev.items()}

    # ---- the pair kernel alone: `reps` back-to-back launches bracketed by HIP events on the
    # stream they are launched on (the last `reps` pair dispatches of the run). Energies and
    # virials are computed as in the evaluation; without a totals buffer the one-block FP64 sum
    # of the kernel's partials is not launched, so only k_pair is timed
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(reps):
        pf.compute(nl, fx, fy, fz, E, V, None, energy=True)
    b.record(stream)
    torch.cuda.synchronize()
    pair_ms = a.elapsed_time(b) / reps

    # forced rebuild cost (decide + full rebuild)
    rb = []
    for _ in range(10):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        nl.rebuild(X, Y, Z, Q)
        b.record(stream)
        torch.cuda.synchronize()
        rb.append(a.elapsed_time(b))
    rebuild_ms = float(np.median(rb))
    interval = st.interval
    md_ms = ms + rebuild_ms / max(interval, 1)
    ns_day = (1e3 / md_ms) * DT_FS * 86400 / 1e6

    # ---- roofline of the dominant kernel (the pair kernel: FP32 VALU)
    counts = nl.counts().to(torch.int64)
    visits = int(counts.sum().item())
    info = pf.info()
    fpv = pair_flops_per_visit(info["n_energy_terms"], info["n_force_terms"], True)
    pair_flop = visits * fpv
    achieved = pair_flop / (pair_ms * 1e-3) / 1e12
    roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": pair_traffic(),
                "kernel": "k_pair<LJ,coulomb,energy>", "duration_us": round(1e3 * pair_ms, 2), "launches_timed": reps,
                "note": "FP32 VALU peak; flops counted per row visit; duration = mean over back-to-back "
                        "launches of the kernel alone (HIP events on its stream); traffic = HBM bytes per launch "
                        "from rocprofv3 FETCH_SIZE / WRITE_SIZE scaled by a 256 MiB copy",
                "flop_per_launch": pair_flop, "row_visits": visits, "flop_per_visit": fpv,
                "algorithmic_bytes_per_launch": visits * 4 + N * 60}

    tot = eng.totals()
    out = {"metric": "force evals/s + ns/day at matched force err: learned-corrected vs plain PME",
           "value": round(evals, 2), "unit": "force evals/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "rock-salt 64000 ions, LJ + SPME 120^3 P=5, rc 9 A, tol 1e-5, plain PME",
                      "N": N, "mesh": list(sol.mesh), "order": sol.order, "alpha": sol.alpha, "skin": args.skin,
                      "cap": nl.cap, "lanes": args.lanes, "mode": args.mode, "spread": args.spread,
                      "traversal": args.traversal, "overlap_with": args.overlap_with,
                      "parallelism": "single GPU"},
           "ns_per_day": round(ns_day, 2), "ns_per_day_note": f"dt=1 fs, one rebuild every {interval} steps",
           "phase_ms": {k: round(v, 4) for k, v in phase_ms.items()}, "rebuild_ms": round(rebuild_ms, 4),
           "energy_eV": tot["energy"], "roofline": roofline}

    if not args.no_error:
        # Coulomb force error of plain PME against FP64 Ewald on 256 sampled ions
        e2 = pdx.ForceEngine(N, box, None, electrostatics=spec, charges=sysd["q"], skin=args.skin)
        gx, gy, gz = (torch.empty(N, device="cuda") for _ in range(3))
        e2.evaluate(X, Y, Z, Q, gx, gy, gz, None, None, energy=False)
        rng = np.random.default_rng(1)
        ids = np.sort(rng.choice(N, 256, replace=False)).astype(np.int32)
        xs = [t.double() for t in (X, Y, Z, Q)]
        Fref, _ = pdx.ewald_fp64(*xs, box, C, 0.25, digits=12.0, sample=torch.from_numpy(ids).cuda(),
                                 want_energy=False)
        Fg = torch.stack([gx, gy, gz], 1).double()[torch.from_numpy(ids).long().cuda()]
        err = math.sqrt(((Fg - Fref) ** 2).sum().item() / (Fref ** 2).sum().item())
        out["force_rel_rms_vs_ewald"] = err
        out["force_error_sample"] = "Coulomb forces of 256 ions vs FP64 Ewald (alpha 0.25, 12 digits)"

    if not args.no_published:
        out["published_coulomb"] = published_leg(args.steps, args.warmup, args.skin)
    if not args.no_learned:
        out["learned"] = learned_leg(args.steps, args.warmup, args.skin)
    if not args.no_replicas:
        out["replicas"] = replica_leg()
    if not args.no_precision:
        out["precision"] = precision_leg(args.steps, args.warmup)
    if not args.no_clusters:
        out["clusters"] = cluster_leg(args.steps, args.warmup, args.skin, nl, X, Y, Z, pair_ms)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(sysd, C)
    print(json.dumps(out))


def _time_events(fn, reps, stream):
    import torch
    for _ in range(3):
        fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(reps):
        fn()
    b.record(stream)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def pairs_within(nl, X, Y, Z, box, rc):
    """Unique pairs closer than rc, counted over the ELL rows of nl (GPU, torch)."""
    import torch
    N, cap = nl.N, nl.cap
    rows = nl.neighbors_flat().view(cap, N).long() - 1
    cnt = nl.counts().long()
    P = torch.stack([X, Y, Z], 1).double()
    Lb = torch.tensor(box, dtype=torch.float64, device="cuda")
    total = 0
    for t0 in range(0, cap, 32):
        r = rows[t0:t0 + 32]
        valid = (torch.arange(t0, min(t0 + 32, cap), device="cuda")[:, None] < cnt[None, :]) & (r >= 0)
        d = P[r.clamp(min=0)] - P[None, :, :]
        d -= Lb * torch.round(d / Lb)
        total += int(((d * d).sum(-1) < rc * rc)[valid].sum().item())
    return total // 2


def cluster_leg(steps, warmup, skin, nl_rows, X, Y, Z, rows_pair_ms):
    """The cluster-pair backend (8 x 8 Newton-3 tiles) on the same 64 000-ion PME evaluation, and
    its force pass against the ELL rows' on that system and on the 4000-atom LJ fluid."""
    import numpy as np
    import torch

    import pdx
    from pdx import systems
    sysd = systems.rock_salt(40, 2.82, 0.1, seed=0)
    N, box = sysd["N"], sysd["box"]
    C = 14.399645
    spec = pdx.PME(coulomb_prefactor=C, real_cutoff=9.0, real_space_tolerance=1e-5, mesh_spacing=1.0, order=5)
    lj = pdx.LennardJones(epsilon=0.01, sigma=2.8, cutoff=9.0)
    eng = pdx.ForceEngine(N, box, lj, electrostatics=spec, charges=sysd["q"], skin=skin, backend="cluster")
    Q = torch.from_numpy(sysd["q"]).cuda()
    fx, fy, fz, E = (torch.empty(N, device="cuda") for _ in range(4))
    V = torch.empty(6, N, device="cuda")
    eng.capture(X, Y, Z, Q, fx, fy, fz, E, V, energy=True)
    for _ in range(max(warmup, 1)):
        eng.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.replay()
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    eng.totals()
    stream = torch.cuda.current_stream()
    cl = eng.cluster
    tot = eng.pair_totals
    pass_ms = _time_events(lambda: cl.forces(fx, fy, fz, E, V, tot, energy=True), 50, stream)
    pass_f_ms = _time_events(lambda: cl.forces(fx, fy, fz, energy=False), 50, stream)
    upd_ms = _time_events(lambda: cl.update(X, Y, Z, Q), 50, stream)

    def rebuild():
        cl.request_build()
        cl.update(X, Y, Z, Q)
        cl.forces(fx, fy, fz, E, V, tot, energy=True)
    build_ms = _time_events(rebuild, 10, stream)
    tc = cl.tile_counts()
    inrc = pairs_within(nl_rows, X, Y, Z, box, 9.0)
    evals = 64 * tc["working_tiles"]
    fpp = 109
    c2 = {"ms_per_eval": round(ms, 4), "evals_per_s": round(1e3 / ms, 2), "pair_pass_ms": round(pass_ms, 4),
          "pair_pass_force_only_ms": round(pass_f_ms, 4), "update_ms": round(upd_ms, 4),
          "build_prune_ms": round(build_ms, 4), "rows_pair_pass_ms": round(rows_pair_ms, 4),
          "outer_entries": tc["outer"], "working_entries": tc["working"], "working_tiles": tc["working_tiles"],
          "tile_pair_evaluations": evals, "unique_pairs_within_cutoff": inrc,
          "tile_efficiency": round(inrc / max(evals, 1), 4),
          "tile_tflops_evaluated": round(evals * fpp / (pass_ms * 1e-3) / 1e12, 3),
          "tile_tflops_in_cutoff": round(inrc * fpp / (pass_ms * 1e-3) / 1e12, 3),
          "flop_per_pair": fpp, "skin": skin, "prune_buffer": cl.prune_buffer, "waves": cl.info["waves"]}
    # the 4000-atom LJ fluid (reduced units, rc 2.5, skin 0.4): rows against tiles, force pass with energies
    s1 = systems.fcc_lattice(4000, 0.8442, 0.05, seed=0)
    X1, Y1, Z1 = (torch.from_numpy(s1[k]).cuda() for k in ("x", "y", "z"))
    pf1 = pdx.PairForce(pdx.LennardJones())
    n1 = 4000
    g = [torch.empty(n1, device="cuda") for _ in range(4)]
    V1 = torch.empty(6, n1, device="cuda")
    t1 = torch.zeros(7, dtype=torch.float64, device="cuda")
    ws = pdx.DeviceNeighborWorkspace(n1, s1["box"], 2.5, 0.4, pdx.default_capacity(n1, s1["box"], 2.9))
    ws.update(X1, Y1, Z1)
    rows1 = _time_events(lambda: pf1.compute(ws, g[0], g[1], g[2], g[3], V1, t1, energy=True), 50, stream)
    cl1 = pdx.ClusterPairEngine(n1, s1["box"], pf1, 0.4)
    cl1.update(X1, Y1, Z1)
    cl1.forces(g[0], g[1], g[2], g[3], V1, t1, energy=True)
    tiles1 = _time_events(lambda: cl1.forces(g[0], g[1], g[2], g[3], V1, t1, energy=True), 50, stream)
    tc1 = cl1.tile_counts()
    c1 = {"rows_pair_pass_ms": round(rows1, 4), "cluster_pair_pass_ms": round(tiles1, 4),
          "working_tiles": tc1["working_tiles"], "tile_pair_evaluations": 64 * tc1["working_tiles"]}
    del eng
    return {"c2": c2, "c1_lj_fluid_4000": c1,
            "note": "cluster tiles evaluate every pair of an 8 x 8 tile; rows visit only list entries"}


def coulomb_error(X, Y, Z, Q, box, spec, C, skin, correction=None, nsample=256):
    """Relative RMS Coulomb-force error of one mesh setting (optionally with the learned
    correction) against FP64 Ewald (alpha 0.25, 12 digits) on sampled ions."""
    import numpy as np
    import torch

    import pdx
    N = X.numel()
    eng = pdx.ForceEngine(N, box, None, electrostatics=spec, charges=Q.cpu().numpy(), skin=skin,
                          correction=correction)
    g = [torch.empty(N, device="cuda") for _ in range(3)]
    eng.evaluate(X, Y, Z, Q, *g, None, None, energy=False)
    ids = torch.from_numpy(np.sort(np.random.default_rng(1).choice(N, nsample, replace=False)).astype(np.int32)).cuda()
    Fref, _ = pdx.ewald_fp64(*(t.double() for t in (X, Y, Z, Q)), box, C, 0.25, digits=12.0, sample=ids,
                             want_energy=False)
    Fg = torch.stack(g, 1).double()[ids.long()]
    return math.sqrt(((Fg - Fref) ** 2).sum().item() / (Fref ** 2).sum().item())


# The reference's published Coulomb force calls (docs/reference/manual/electrostatics.md:248-253):
# 64 000 ions, measured force error 1e-4, rc 9 A, FP32, RTX 3090, one force call each.
PUBLISHED_64K = (("SPME", "spme", 150, 1.14), ("P3M-AD", "p3m_ad", 150, 1.09), ("P3M-ik", "p3m_ik", 120, 1.15))


def published_leg(steps, warmup, skin):
    """The reference's published 64 000-ion Coulomb force calls at the same methods and meshes, on
    64 000 +-1 e ions (rock salt with 0.5 A random displacements: disordered like the reference's
    melt, whose relative force errors a near-perfect lattice would inflate), Coulomb only, rc 9 A:
    per method the order and real-space tolerance (alpha) with the smallest measured force error
    against FP64 Ewald are kept, and one force call (list check with predicated rebuild, real-space
    pairs, spread, solve, gather; forces only) is timed as a graph replay. "matched" says whether
    our error is within the reference's 1e-4."""
    import numpy as np
    import torch

    import pdx
    from pdx import systems
    C = 14.399645
    sysd = systems.rock_salt(40, 2.82, 0.5, seed=0)
    N, box = sysd["N"], sysd["box"]
    X, Y, Z, Q = (torch.from_numpy(sysd[k]).cuda() for k in ("x", "y", "z", "q"))

    def spec_of(kind, K, P, tol):
        if kind == "spme":
            return pdx.PME(coulomb_prefactor=C, real_cutoff=9.0, real_space_tolerance=tol, mesh=(K, K, K), order=P)
        return pdx.PPPM(coulomb_prefactor=C, real_cutoff=9.0, real_space_tolerance=tol, mesh=(K, K, K), order=P,
                        differentiation=kind[-2:])

    res = {"N": N, "system": "64 000 +-1 e ions (rock salt, 0.5 A random displacements), Coulomb only, rc 9 A",
           "reference": "RTX 3090, dense +-1 e melt, measured force error 1e-4 (electrostatics.md:248-253)"}
    for name, kind, K, ref_ms in PUBLISHED_64K:
        trials = []
        for P in ((4, 5) if kind == "spme" else (5,)):
            for tol in (3e-5, 1e-5, 3e-6, 1e-6):
                spec = spec_of(kind, K, P, tol)
                trials.append((coulomb_error(X, Y, Z, Q, box, spec, C, skin), tol, spec, P))
        err, tol, spec, _ = min(trials, key=lambda t: t[0])
        eng = pdx.ForceEngine(N, box, None, electrostatics=spec, charges=sysd["q"], skin=skin)
        fx, fy, fz = (torch.empty(N, device="cuda") for _ in range(3))
        eng.capture(X, Y, Z, Q, fx, fy, fz, None, None, energy=False)
        for _ in range(max(warmup, 1)):
            eng.replay()
        torch.cuda.synchronize()
        a = time.perf_counter()
        for _ in range(steps):
            eng.replay()
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - a) / steps
        res[name] = {"mesh": K, "order": eng.solver.order, "real_space_tolerance": tol,
                     "alpha": round(spec.resolved_alpha(), 5), "ms_per_force_call": round(ms, 4),
                     "force_rel_rms_vs_ewald": err, "matched": bool(err <= 1e-4),
                     "reference_ms": ref_ms, "reference_over_ours": round(ref_ms / ms, 2),
                     "errors_by_order_and_tolerance": {f"P{P} {t:g}": e for e, t, _, P in trials}}
        del eng
    return res


def learned_leg(steps, warmup, skin):
    """Coarse-mesh PME plus the learned correction against plain PME on the ionic workload
    with 5% of the sites vacated in neutral pairs (60 800 ions): the network (80-64-64-3,
    BF16 MFMA) is trained on the device on residuals the method generates (coarse 2.0 A / P=4
    against fine 0.5 A / P=6 PME, three training configurations), then each variant's full
    force evaluation (LJ + real space + mesh [+ correction], energy and virial) is timed as a
    graph replay and its Coulomb-force error measured against FP64 Ewald on a held-out
    configuration."""
    import numpy as np
    import torch

    import pdx
    from pdx import learned, systems
    C = 14.399645
    mk = lambda h, P: pdx.PME(coulomb_prefactor=C, real_cutoff=9.0, real_space_tolerance=1e-5, mesh_spacing=h,
                              order=P)
    plain, coarse, fine = mk(1.0, 5), mk(2.0, 4), mk(0.5, 6)
    lj = pdx.LennardJones(epsilon=0.01, sigma=2.8, cutoff=9.0)
    train_sys = [systems.rock_salt(40, 2.82, 0.1, seed=s, vacancy_fraction=0.05) for s in (11, 12, 13)]
    t0 = time.perf_counter()
    F, T = learned.residual_dataset(train_sys, coarse, fine, skin)
    mlp = learned.MLP(seed=0)
    mu, isd, sc = learned.fit_standardisation(F, T)
    mlp.set_norm(mu, isd, sc)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    nsteps, batch = 4000, 8192
    hist = learned.train(mlp, F, T, steps=nsteps, batch=batch, lr=2e-3)
    torch.cuda.synchronize()
    t2 = time.perf_counter()

    sysd = systems.rock_salt(40, 2.82, 0.1, seed=10, vacancy_fraction=0.05)
    N, box = sysd["N"], sysd["box"]
    X, Y, Z, Q = (torch.from_numpy(sysd[k]).cuda() for k in ("x", "y", "z", "q"))
    res = {"N": N, "train": {"configs": len(train_sys), "samples": int(F.shape[0]), "steps": nsteps,
                             "batch": batch, "data_s": round(t1 - t0, 2), "train_s": round(t2 - t1, 2),
                             "final_loss": hist[-1] if hist else None}}
    for name, spec, corr in (("plain", plain, None), ("coarse", coarse, None), ("corrected", coarse, mlp)):
        eng = pdx.ForceEngine(N, box, lj, electrostatics=spec, charges=sysd["q"], skin=skin, correction=corr)
        fx, fy, fz, E = (torch.empty(N, device="cuda") for _ in range(4))
        V = torch.empty(6, N, device="cuda")
        eng.capture(X, Y, Z, Q, fx, fy, fz, E, V, energy=True)
        for _ in range(max(warmup, 1)):
            eng.replay()
        torch.cuda.synchronize()
        a = time.perf_counter()
        for _ in range(steps):
            eng.replay()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - a) / steps
        err = coulomb_error(X, Y, Z, Q, box, spec, C, skin, correction=corr)
        res[name] = {"mesh": list(eng.solver.mesh), "order": eng.solver.order, "ms_per_eval": round(1e3 * dt, 4),
                     "evals_per_s": round(1.0 / dt, 2), "force_rel_rms_vs_ewald": err}
        del eng
    res["speedup_corrected_vs_plain"] = round(res["plain"]["ms_per_eval"] / res["corrected"]["ms_per_eval"], 3)
    res["error_ratio_corrected_vs_plain"] = round(res["corrected"]["force_rel_rms_vs_ewald"] /
                                                  res["plain"]["force_rel_rms_vs_ewald"], 3)
    res["matched"] = res["corrected"]["force_rel_rms_vs_ewald"] <= res["plain"]["force_rel_rms_vs_ewald"]
    return res


def replica_leg(nrep=512, nsteps=1000, seed=0):
    """512 independent LJ / WCA boxes (n log-uniform in [256, 8192], density uniform in
    [0.70, 0.90], model drawn per box), each advanced 1000 velocity-Verlet steps (dt 0.005)
    in one launch captured as a hipGraph, after the launch shape is autotuned on 10 steps."""
    import numpy as np
    import torch

    import pdx
    from pdx import replicas, systems
    b = systems.replica_batch(nrep, seed=seed)
    rb = pdx.ReplicaBatch(b["specs"])
    st = replicas.device_state(b)
    rb.forces(st, energy=False)
    best, times = replicas.autotune(rb, st, steps=10)
    work = replicas.clone_state(st)
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        rb.run(work, 1)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    work = replicas.clone_state(st)
    out = torch.zeros(nrep, 4, dtype=torch.float64, device="cuda")
    e0 = rb.forces(replicas.clone_state(st)).cpu().numpy()
    with torch.cuda.graph(g):
        rb.run(work, nsteps, dt=0.005, energy=True, out=out)
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    z = torch.cuda.Event(enable_timing=True)
    a.record()
    g.replay()
    z.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(z)
    e1 = out.cpu().numpy()
    ns = np.array([sp["n"] for sp in b["specs"]])
    drift = np.abs((e1[:, 0] + e1[:, 1]) - (e0[:, 0] + e0[:, 1])) / ns
    return {"replicas": nrep, "atoms": int(ns.sum()), "steps": nsteps, "ms": round(ms, 3),
            "atom_steps_per_s": round(float(ns.sum()) * nsteps / (ms * 1e-3), 1),
            "replica_steps_per_s": round(nrep * nsteps / (ms * 1e-3), 1),
            "models": {"lj": int(sum(sp["model"] == "lj" for sp in b["specs"])),
                       "wca": int(sum(sp["model"] == "wca" for sp in b["specs"]))},
            "launch": {"block": best[0], "cells_per_cutoff": best[1], "autotune_ms_10_steps": times},
            "energy_drift_per_atom_max": float(drift.max()), "energy_drift_per_atom_median": float(np.median(drift))}


def precision_leg(steps, warmup, nchains=10486, length=100):
    """Kremer-Grest melt of 1 048 600 beads (10 486 random-walk chains of 100, bond 0.97,
    density 0.85, pushed off with capped WCA + FENE descent): one force evaluation (list
    check + WCA over all pairs + FENE, energy and virial) per mode of the per-particle sums
    (FP32, 64-bit fixed point, FP64), each timed as a graph replay, and its error against
    the FP64 oracle on all beads."""
    import numpy as np
    import torch

    import pdx
    from pdx import melt, systems
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    sysd = systems.kg_chains(nchains, length=length, density=0.85, seed=0)
    t0 = time.perf_counter()
    X, Y, Z, bonds = melt.push_off(sysd, iterations=300)
    t_push = time.perf_counter() - t0
    N, box = sysd["N"], sysd["box"]
    t0 = time.perf_counter()
    ref = oracle.kg_forces(sysd["x"], sysd["y"], sysd["z"], box, sysd["bonds"])
    t_ref = time.perf_counter() - t0
    Fn = np.linalg.norm(ref["F"])
    res = {"N": N, "chains": nchains, "bonds": int(len(sysd["bonds"])), "box": box[0],
           "push_off_s": round(t_push, 2), "fp64_oracle_s": round(t_ref, 2), "wca_pairs": ref["pairs"]}
    for mode in ("fp32", "fixed64", "fp64"):
        eng = pdx.ForceEngine(N, box, pdx.WCA(epsilon=1.0, sigma=1.0), skin=0.4, accumulation=mode, bonds=bonds,
                              cap=2 * pdx.default_capacity(N, box, 2 ** (1 / 6) + 0.4))
        f = [torch.empty(N, device="cuda") for _ in range(3)]
        E = torch.empty(N, device="cuda")
        V = torch.empty(6, N, device="cuda")
        eng.capture(X, Y, Z, None, *f, E, V, energy=True)
        for _ in range(max(warmup, 1)):
            eng.replay()
        torch.cuda.synchronize()
        a = time.perf_counter()
        for _ in range(steps):
            eng.replay()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - a) / steps
        eng.evaluate(X, Y, Z, None, *f, E, V, energy=True)
        tot = eng.totals()
        F = torch.stack(f, 1).double().cpu().numpy()
        d = F - ref["F"]
        W = np.array(tot["pair_virial"])
        visits = int(eng.nl.counts().to(torch.int64).sum().item())
        res[mode] = {"ms_per_eval": round(1e3 * dt, 4), "evals_per_s": round(1.0 / dt, 2),
                     "force_rel_rms": float(np.linalg.norm(d) / Fn), "force_max_abs_err": float(np.abs(d).max()),
                     "energy_rel_err": float(abs(tot["pair_energy"] - ref["energy"]) / abs(ref["energy"])),
                     "virial_rel_err": float(np.linalg.norm(W - ref["W"]) / np.linalg.norm(ref["W"])),
                     "row_visits": visits}
        bonds.status()
        del eng
    return res


def cpu_baseline(sysd, C):
    """The FP64 CPU oracle (Ewald + LJ direct sums) timed on a bounded sample of ions and
    scaled to one full force evaluation: the structure factors are timed alone (they are
    computed once per evaluation), the per-ion real-space, reciprocal and LJ sums on S ions,
    then t_full = t_structure + N * t_per_ion."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import oracle
    x, y, z, q = (sysd[k].astype(np.float64) for k in ("x", "y", "z", "q"))
    box = sysd["box"]
    N = len(x)
    S = 256
    ids = np.sort(np.random.default_rng(2).choice(N, S, replace=False))
    t0 = time.perf_counter()
    oracle.ewald_sampled(x, y, z, q, box, C_=C, alpha=0.15, digits=12.0, ids=ids[:0])
    t_struct = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.ewald_sampled(x, y, z, q, box, C_=C, alpha=0.15, digits=12.0, ids=ids)
    t_ewald = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.pair_direct(x, y, z, box, kind="lj", sigma=2.8, epsilon=0.01, cutoff=9.0, ids=ids)
    t_lj = time.perf_counter() - t0
    per_ion = (max(t_ewald - t_struct, 0.0) + t_lj) / S
    t_full = t_struct + N * per_ion
    return {"value": round(1.0 / t_full, 6), "unit": "force evals/s", "cores": oracle.num_threads(),
            "kind": "port",
            "sample": f"FP64 Ewald (alpha 0.15, 12 digits) + LJ direct sums: structure factors {t_struct:.2f} s, "
                      f"forces on {S} of {N} ions {t_ewald - t_struct + t_lj:.2f} s, scaled to all ions: "
                      f"{t_full:.1f} s per evaluation"}


if __name__ == "__main__":
    main()
