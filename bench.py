"""Force-evaluation benchmark on one MI355X.

Headline workload (the learned-correction comparison the metric names): charged rock salt of
80^3 sites with 5 % of them vacated in neutral pairs (486 400 ions, a = 2.82 A, +-1 e, Gaussian
jitter 0.1 A), LJ eps = 0.01 eV, sigma = 2.8 A, cutoff 9 A, plus Ewald-split Coulomb
(C = 14.399645 eV A / e^2). Two ways to reach a 1e-4 relative RMS Coulomb force error against
FP64 Ewald: plain smooth PME tuned over real-space cutoff, order and mesh, and coarse PME
(2.0 A, order 4) plus the learned per-ion correction trained on the device on residuals the
method generates (fine 0.5 A, order 6). The headline value is the force evaluations/s of the
faster variant that reaches the error (one evaluation: list check with predicated rebuild, LJ +
real-space pairs, spread, FFT / influence / inverse FFT, gather [+ correction], energy and full
virial, replayed as one hipGraph; inputs resident in HBM). Prints ONE JSON line.
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


def _log(msg):
    """Progress of the long legs on stderr (stdout carries only the JSON line)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)

FP32_PEAK_TFLOPS = 157.3
DT_FS = 1.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-clusters", action="store_true", help="skip the cluster-pair backend leg")
    p.add_argument("--c5-sites", type=int, default=80, help="lattice sites per side of the headline system")
    p.add_argument("--no-small-c5", action="store_true", help="skip the 60 800-ion learned comparison")
    p.add_argument("--no-c2", action="store_true", help="skip the 64 000-ion plain-SPME leg")
    p.add_argument("--quick", action="store_true", help="headline search over rc 9 A, order 5 only")
    p.add_argument("--pair-only", action="store_true",
                   help="run only the headline system's pair kernel (`steps` launches) for profiling")
    p.add_argument("--pair-rc", type=float, default=9.0, help="real-space cutoff of --pair-only")
    p.add_argument("--only", choices=("c2", "published", "replicas", "precision"), default=None,
                   help="run only one extra leg (the profiling runs): the 64 000-ion plain SPME leg with its cluster "
                        "backend, the published Coulomb force calls, the 512 replicas or the bead-spring study")
    p.add_argument("--eval-only", default=None, metavar="RC,ORDER,MESH",
                   help="run only `steps` graph replays of the headline system's plain PME evaluation at this "
                        "real-space cutoff, order and cubic mesh (the profiling runs), e.g. 10,6,160")
    p.add_argument("--skin", type=float, default=1.0)
    p.add_argument("--no-reorder", action="store_true",
                   help="keep the headline systems in their generated (lattice) storage order")
    p.add_argument("--no-replicas", action="store_true", help="skip the batched-replica run")
    p.add_argument("--no-precision", action="store_true", help="skip the bead-spring accumulation study")
    p.add_argument("--no-published", action="store_true",
                   help="skip the Coulomb force calls at the reference's published methods and meshes")
    p.add_argument("--copy-calibration", action="store_true",
                   help="run only a plain 256 MiB device copy (five launches): the reference reading for the "
                        "FETCH_SIZE / WRITE_SIZE counters")
    return p.parse_args()


def pair_traffic(N):
    """HBM bytes per pair-kernel launch of the N-ion system, as recorded by the counter summary in
    profiles/ (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by the reading of a plain copy
    of known size); None when that summary is absent."""
    name = {486400: "k_pair_rocksalt486k_hbm_traffic.txt", 64000: "k_pair_rocksalt64k_hbm_traffic.txt"}.get(N)
    if name is None:
        return None
    try:
        with open(os.path.join(ROOT, "profiles", name)) as f:
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
    import torch

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
    if args.pair_only:
        pair_only(args)
        return
    if args.eval_only:
        eval_only(args)
        return
    if args.only:
        leg = {"c2": ("c2_rock_salt_64k", lambda: c2_leg(args)),
               "published": ("published_coulomb", lambda: published_leg(args.steps, args.warmup, args.skin)),
               "replicas": ("replicas", replica_leg),
               "precision": ("precision", lambda: precision_leg(args.steps, args.warmup))}[args.only]
        print(json.dumps({leg[0]: leg[1]()}))
        return
    head = headline_leg(args.c5_sites, args.steps, args.warmup, args.skin, quick=args.quick,
                        reorder=not args.no_reorder)
    out = head.pop("line")
    if not args.no_small_c5:
        _log("60 800-ion learned comparison")
        out["learned_60800"] = headline_leg(40, args.steps, args.warmup, args.skin, quick=args.quick,
                                            reorder=not args.no_reorder)["line"]
        for k in ("metric", "unit", "n_gpus", "steps", "warmup", "higher_is_better", "scaling", "vs_baseline",
                  "dtype", "data", "roofline"):
            out["learned_60800"].pop(k, None)
    if not args.no_c2:
        _log("64 000-ion plain SPME leg")
        out["c2_rock_salt_64k"] = c2_leg(args)
    if not args.no_published:
        _log("published Coulomb force calls")
        out["published_coulomb"] = published_leg(args.steps, args.warmup, args.skin)
    if not args.no_replicas:
        _log("512 replicas")
        out["replicas"] = replica_leg()
    if not args.no_precision:
        _log("bead-spring precision study")
        out["precision"] = precision_leg(args.steps, args.warmup)
    if not args.no_cpu_baseline:
        _log("CPU baseline")
        out["cpu_baseline"] = cpu_baseline(head["sysd"], head["plain_spec"])
    print(json.dumps(out))


C_ION = 14.399645
TARGET_ERR = 1e-4


def _ion_system(nsites, seed):
    from pdx import systems
    return systems.rock_salt(nsites, 2.82, 0.1, seed=seed, vacancy_fraction=0.05)


def _tensors(sysd):
    import torch
    return [torch.from_numpy(sysd[k]).cuda() for k in ("x", "y", "z", "q")]


def _ewald_reference(P, box, ids):
    """FP64 Ewald Coulomb forces on the sampled ions (device kernel). alpha = 17 / L keeps the real
    space inside one image (rcut = 5.3 / alpha < L / 2) and about 48 000 half-space wave vectors."""
    import pdx
    alpha = 17.0 / min(box)
    F, _ = pdx.ewald_fp64(*(t.double() for t in P), box, C_ION, alpha, digits=12.0, sample=ids, want_energy=False)
    return F


def _coulomb_error(P, box, spec, skin, ids, Fref, correction=None):
    """Relative RMS Coulomb-force error of one mesh setting (+ correction) on the sampled ions."""
    import torch

    import pdx
    N = P[0].numel()
    eng = pdx.ForceEngine(N, box, None, electrostatics=spec, charges=P[3].cpu().numpy(), skin=skin,
                          correction=correction)
    g = [torch.empty(N, device="cuda") for _ in range(3)]
    eng.evaluate(*P, *g, None, None, energy=False)
    Fg = torch.stack(g, 1).double()[ids.long()]
    err = math.sqrt(((Fg - Fref) ** 2).sum().item() / (Fref ** 2).sum().item())
    del eng
    return err


TUNED = {}  # (N, rc, order, mesh, corrected) -> autotune result of that evaluation


def _time_eval(P, box, spec, skin, steps, warmup, correction=None, keep=False, tune=False, settings=None):
    """Full evaluation (LJ + Coulomb, energy and virial) as a graph replay: ms per evaluation.
    tune: first measure the launch-shape candidates of pdx.autotune on this system and keep the
    fastest (set-up, outside the timing)."""
    import torch

    import pdx
    from pdx import autotune
    N = P[0].numel()
    lj = pdx.LennardJones(epsilon=0.01, sigma=2.8, cutoff=9.0)
    eng = pdx.ForceEngine(N, box, lj, electrostatics=spec, charges=P[3].cpu().numpy(), skin=skin,
                          correction=correction, settings=settings)
    if tune:
        key = (N, spec.real_cutoff, spec.order, eng.solver.mesh[0], correction is not None)
        if key not in TUNED:
            TUNED[key] = autotune.tune(eng, *P, steps=10, warmup=3, rounds=2)
            _log(f"tuned N {N} rc {spec.real_cutoff} order {spec.order} mesh {eng.solver.mesh[0]}"
                 f"{' + correction' if correction is not None else ''}: {TUNED[key]['best']} "
                 f"{TUNED[key]['ms']:.4f} ms")
        else:
            autotune.apply(eng, TUNED[key]["best"])
    bufs = [torch.empty(N, device="cuda") for _ in range(4)] + [torch.empty(6, N, device="cuda")]
    eng.capture(*P, *bufs, energy=True)
    for _ in range(max(warmup, 1)):
        eng.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.replay()
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    eng.status()
    if keep:
        return ms, eng, bufs
    del eng
    return ms


def _spec(rc, P, K):
    import pdx
    return pdx.PME(coulomb_prefactor=C_ION, real_cutoff=rc, real_space_tolerance=1e-5, mesh=(K, K, K), order=P)


def plain_search(P, box, skin, ids, Fref, quick=False):
    """Plain smooth PME at the target error: for each real-space cutoff and order, the smallest
    2,3,5,7-smooth cubic mesh whose measured error reaches 1e-4 (alpha from erfc(alpha rc) = 1e-5),
    then its full evaluation timed; the fastest one is the plain variant."""
    import pdx
    L = box[0]
    trials, feasible = [], []
    rcs = (9.0,) if quick else (9.0, 10.0)
    orders = (5,) if quick else (4, 5, 6)
    for rc in rcs:
        for order in orders:
            Ks = sorted({pdx.fft_friendly(int(math.ceil(L / h))) for h in
                         (1.6, 1.45, 1.3, 1.2, 1.1, 1.0, 0.92, 0.85, 0.78, 0.7, 0.62, 0.55)})
            for K in Ks:
                err = _coulomb_error(P, box, _spec(rc, order, K), skin, ids, Fref)
                _log(f"plain PME rc {rc} order {order} mesh {K}: error {err:.3e}")
                trials.append({"rc": rc, "order": order, "mesh": K, "err": err})
                if err <= TARGET_ERR:
                    ms = _time_eval(P, box, _spec(rc, order, K), skin, 20, 3, tune=True)
                    feasible.append({"rc": rc, "order": order, "mesh": K, "err": err, "ms_per_eval": round(ms, 4)})
                    trials[-1]["ms_per_eval"] = round(ms, 4)
                    break
    best = min(feasible, key=lambda r: r["ms_per_eval"]) if feasible else None
    return best, trials


def train_correction(nsites, skin, coarse, fine, configs=2, steps=4000, batch=8192):
    """The learned correction trained on the device on residuals F_fine - F_coarse of `configs`
    training configurations (seeds 11, 12, ...) of the workload's construction."""
    import torch

    from pdx import learned
    train = [_ion_system(nsites, 11 + k) for k in range(configs)]
    _log(f"training data: {configs} configurations of {train[0]['N']} ions")
    t0 = time.perf_counter()
    F, T = learned.residual_dataset(train, coarse, fine, skin)
    mlp = learned.MLP(seed=0)
    mu, isd, sc = learned.fit_standardisation(F, T)
    mlp.set_norm(mu, isd, sc)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    _log(f"training the correction on {int(F.shape[0])} samples")
    hist = learned.train(mlp, F, T, steps=steps, batch=batch, lr=2e-3)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    info = {"configs": configs, "samples": int(F.shape[0]), "steps": steps, "batch": batch,
            "data_s": round(t1 - t0, 2), "train_s": round(t2 - t1, 2), "final_loss": hist[-1] if hist else None}
    del F, T
    return mlp, info


def spatially_ordered(sysd, skin):
    """The system relabelled into the spatial order of the engine's cell list (ForceEngine.reorder,
    the reference's spatial reordering, src/engine/Ordering.jl:19-79, which its published timings
    use): storage neighbours are then space neighbours. Host-side set-up of the inputs."""
    import numpy as np
    import torch

    import pdx
    N, box = sysd["N"], sysd["box"]
    X, Y, Z, Q = _tensors(sysd)
    eng = pdx.ForceEngine(N, box, pdx.LennardJones(epsilon=0.01, sigma=2.8, cutoff=9.0),
                          electrostatics=_spec(10.0, 4, pdx.fft_friendly(int(math.ceil(box[0] / 2.0)))),
                          charges=sysd["q"], skin=skin)
    f = [torch.empty(N, device="cuda") for _ in range(3)]
    eng.evaluate(X, Y, Z, Q, *f, None, None, energy=False)
    perm = eng.reorder(X, Y, Z, Q).cpu().numpy()
    del eng
    out = dict(sysd)
    for k in ("x", "y", "z", "q", "mass"):
        out[k] = np.ascontiguousarray(np.asarray(sysd[k])[perm])
    return out


def headline_leg(nsites, steps, warmup, skin, quick=False, reorder=True):
    """Learned-corrected coarse PME against plain PME tuned to the same 1e-4 force error, on the
    held-out configuration (seed 10) of `nsites`^3 sites with 5 % vacated (spatially ordered
    unless reorder is False)."""
    import numpy as np
    import torch

    import pdx
    sysd = _ion_system(nsites, 10)
    if reorder:
        sysd = spatially_ordered(sysd, skin)
    N, box = sysd["N"], sysd["box"]
    P = _tensors(sysd)
    ids = torch.from_numpy(np.sort(np.random.default_rng(1).choice(N, 256, replace=False)).astype(np.int32)).cuda()
    _log(f"headline system {N} ions: FP64 Ewald reference of 256 ions")
    t0 = time.perf_counter()
    Fref = _ewald_reference(P, box, ids)
    torch.cuda.synchronize()
    t_ref = time.perf_counter() - t0
    best, trials = plain_search(P, box, skin, ids, Fref, quick=quick)
    mk = lambda h, order: pdx.PME(coulomb_prefactor=C_ION, real_cutoff=9.0, real_space_tolerance=1e-5,
                                  mesh_spacing=h, order=order)
    coarse, fine = mk(2.0, 4), mk(0.5, 6)
    mlp, train_info = train_correction(nsites, skin, coarse, fine, configs=1 if quick else 2)
    _log("coarse mesh and corrected errors")
    err_coarse = _coulomb_error(P, box, coarse, skin, ids, Fref)
    err_corr = _coulomb_error(P, box, coarse, skin, ids, Fref, correction=mlp)
    err_fine = _coulomb_error(P, box, fine, skin, ids, Fref)
    ms_coarse = _time_eval(P, box, coarse, skin, 20, 3, tune=True)
    ms_corr = _time_eval(P, box, coarse, skin, 20, 3, correction=mlp, tune=True)
    corrected = {"mesh": coarse.resolve_mesh(box)[0], "order": 4, "rc": 9.0, "err": err_corr,
                 "ms_per_eval": round(ms_corr, 4), "train": train_info}
    coarse_only = {"mesh": coarse.resolve_mesh(box)[0], "order": 4, "rc": 9.0, "err": err_coarse,
                   "ms_per_eval": round(ms_coarse, 4)}
    fine_ref = {"mesh": fine.resolve_mesh(box)[0], "order": 6, "err": err_fine}
    corrected_ok = err_corr <= TARGET_ERR
    plain_ok = best is not None
    if corrected_ok and (not plain_ok or ms_corr < best["ms_per_eval"]):
        variant, spec, corr = "learned-corrected coarse PME", coarse, mlp
        chosen = corrected
    elif plain_ok:
        variant, spec, corr = "plain PME", _spec(best["rc"], best["order"], best["mesh"]), None
        chosen = best
    else:  # nothing reaches the target: report the most accurate plain trial
        t = min(trials, key=lambda r: r["err"])
        variant, spec, corr = "plain PME (target not reached)", _spec(t["rc"], t["order"], t["mesh"]), None
        chosen = t
    _log(f"timing {steps} evaluations of {variant}")
    # the headline: exactly `steps` evaluations of the chosen variant, graph replays, at the launch
    # shapes the autotuner measured fastest for it
    from pdx import autotune
    lj = pdx.LennardJones(epsilon=0.01, sigma=2.8, cutoff=9.0)
    eng = pdx.ForceEngine(N, box, lj, electrostatics=spec, charges=sysd["q"], skin=skin, correction=corr)
    tkey = (N, spec.real_cutoff, spec.order, eng.solver.mesh[0], corr is not None)
    if tkey in TUNED:
        autotune.apply(eng, TUNED[tkey]["best"])
    bufs = [torch.empty(N, device="cuda") for _ in range(4)] + [torch.empty(6, N, device="cuda")]
    eng.capture(*P, *bufs, energy=True)
    for _ in range(max(warmup, 1)):
        eng.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.replay()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms = 1e3 * dt / steps
    st = eng.status()
    roofline = pair_roofline(eng, P, box, steps)
    md = md_leg(eng, sysd, P, bufs, steps, warmup)
    line = {"metric": "force evals/s + ns/day at matched force err: learned-corrected vs plain PME",
            "value": round(steps / dt, 2), "unit": "force evals/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"learned PME correction: rock salt {nsites}^3 sites, 5% vacated ({N} ions), "
                                   "LJ + Ewald Coulomb, 1e-4 relative RMS Coulomb force error vs FP64 Ewald",
                       "N": N, "variant": variant, "mesh": list(eng.solver.mesh), "order": eng.solver.order,
                       "real_cutoff": eng.solver.spec.real_cutoff, "alpha": eng.solver.alpha, "skin": skin,
                       "cap": eng.nl.cap, "rebuild_interval": st.interval, "mode": "graph",
                       "storage_order": "spatial (ForceEngine.reorder at set-up)" if reorder else "lattice",
                       "launch": autotune.current(eng), "parallelism": "single GPU"},
            "force_rel_rms_vs_ewald": chosen["err"],
            "force_error_sample": "Coulomb forces of 256 ions vs FP64 Ewald (device, alpha 17/L, 12 digits)",
            "ns_per_day": md["ns_per_day"], "md_step": md,
            "comparison": {"target": TARGET_ERR, "plain_pme": best, "corrected": corrected, "coarse_only": coarse_only,
                           "fine_training_target": fine_ref, "matched": bool(corrected_ok),
                           "speedup_corrected_vs_plain": (round(best["ms_per_eval"] / ms_corr, 3) if plain_ok
                                                          else None),
                           "error_ratio_corrected_vs_target": round(err_corr / TARGET_ERR, 2),
                           "plain_search": trials, "ewald_reference_s": round(t_ref, 2),
                           "autotune": {f"N{k[0]} rc{k[1]} P{k[2]} K{k[3]}{' corrected' if k[4] else ''}":
                                        {"best": v["best"], "ms": round(v["ms"], 4),
                                         "candidates": [[c, m] for c, m in v["table"]]}
                                        for k, v in TUNED.items() if k[0] == N}},
            "roofline": roofline}
    del eng
    plain_spec = _spec(chosen["rc"], chosen["order"], chosen["mesh"]) if "rc" in chosen else spec
    line["comparison"]["frontier"] = frontier(P, box, skin, trials, corrected, coarse_only, quick)
    line["storage_order"] = storage_order_leg(nsites, skin, plain_spec, line["ms_per_step"] if reorder else None)
    return {"line": line, "sysd": sysd, "plain_spec": (best or chosen)}


def frontier(P, box, skin, trials, corrected, coarse_only, quick=False):
    """The measured error / speed front of plain PME around the corrected variant: for each error
    level (the corrected variant's error, 1e-3, 3e-4), the smallest mesh of every searched cutoff
    and order that reaches it, timed with the measured launch table; then the fastest plain
    configuration at the corrected variant's error, against which the corrected variant is judged."""
    levels = sorted({corrected["err"], 1e-3, 3e-4}, reverse=True)
    picked = {}
    for lev in levels:
        for rc, order in sorted({(t["rc"], t["order"]) for t in trials}):
            ok = [t for t in trials if t["rc"] == rc and t["order"] == order and t["err"] <= lev]
            if ok:
                t = min(ok, key=lambda r: r["mesh"])
                picked[(rc, order, t["mesh"])] = t
        if quick:
            break
    points = []
    for (rc, order, K), t in sorted(picked.items()):
        ms = t.get("ms_per_eval")
        if ms is None:
            ms = round(_time_eval(P, box, _spec(rc, order, K), skin, 20, 3, settings="auto"), 4)
        points.append({"rc": rc, "order": order, "mesh": K, "err": t["err"], "ms_per_eval": ms})
        _log(f"frontier: rc {rc} order {order} mesh {K}: error {t['err']:.3e}, {ms:.4f} ms")
    points.sort(key=lambda r: r["err"])
    pareto, best_ms = [], math.inf
    for r in sorted(points, key=lambda r: r["err"]):  # non-dominated: faster than every more accurate point
        if r["ms_per_eval"] < best_ms:
            pareto.append(r)
            best_ms = r["ms_per_eval"]
    at = [r for r in points if r["err"] <= corrected["err"]]
    eq = min(at, key=lambda r: r["ms_per_eval"]) if at else None
    return {"plain": points, "plain_pareto": pareto, "corrected": {k: corrected[k] for k in ("mesh", "order", "rc",
                                                                                               "err", "ms_per_eval")},
            "coarse_only": coarse_only, "plain_at_corrected_error": eq,
            "speedup_corrected_vs_plain_at_its_error": (round(eq["ms_per_eval"] / corrected["ms_per_eval"], 3)
                                                         if eq else None),
            "note": "plain timings use the measured launch table (settings 'auto'); the search's target "
                    "configurations keep their own tuned times"}


def storage_order_leg(nsites, skin, spec, ms_spatial):
    """The chosen plain evaluation on the same configuration stored three ways: lattice order, a
    random permutation, and that permutation put back into spatial order by ForceEngine.reorder
    (label-independent, so it lands on the same order as the headline's)."""
    import numpy as np
    lat = _ion_system(nsites, 10)
    box = lat["box"]
    perm = np.random.default_rng(5).permutation(lat["N"])
    shuf = dict(lat)
    for k in ("x", "y", "z", "q", "mass"):
        shuf[k] = np.ascontiguousarray(np.asarray(lat[k])[perm])
    out = {"spatial (headline)": ms_spatial}
    for name, sysd in (("lattice", lat), ("shuffled", shuf), ("shuffled, then reordered", None)):
        if sysd is None:
            sysd = spatially_ordered(shuf, skin)
        ms = _time_eval(_tensors(sysd), box, spec, skin, 20, 3, settings="auto")
        out[name] = round(ms, 4)
        _log(f"storage order {name}: {ms:.4f} ms")
    out["note"] = ("ms per evaluation of the plain configuration, graph replays, measured launch table; "
                   "'shuffled' is a random relabelling of the same ions")
    return out


def md_leg(eng, sysd, P, bufs, steps, warmup):
    """ns/day of the complete velocity-Verlet step (kick-drift fused into the list check, the
    evaluation, the closing kick) at dt = 1 fs with the ions' masses, graph replays."""
    import numpy as np
    import torch
    N = sysd["N"]
    mass = sysd["mass"]
    kT = 300 * 8.617333e-5
    rng = np.random.default_rng(3)
    v = rng.normal(0, 1, (N, 3)) * np.sqrt(kT / mass)[:, None]
    v -= (v * mass[:, None]).sum(0) / mass.sum()
    X, Y, Z = (t.clone() for t in P[:3])
    Vv = [torch.from_numpy(v[:, d].astype(np.float32)).cuda() for d in range(3)]
    inv = torch.from_numpy((1.0 / mass).astype(np.float32)).cuda()
    f = [torch.empty(N, device="cuda") for _ in range(3)]
    dt_int = DT_FS / 10.1805  # 1 fs in the A / eV / amu time unit
    eng.evaluate(X, Y, Z, P[3], *f, None, None, energy=False)
    eng.capture_step(X, Y, Z, P[3], *Vv, *f, dt_int, inv_mass=inv)
    for _ in range(max(warmup, 1)):
        eng.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.replay()
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / steps
    st = eng.status()
    return {"ms_per_step": round(1e3 * sec, 4), "steps_per_s": round(1.0 / sec, 2),
            "ns_per_day": round((1.0 / sec) * DT_FS * 86400 / 1e6, 2), "dt_fs": DT_FS,
            "builds_during_steps": int(st.builds), "note": "forces only in the MD step (no energies)"}


def pair_roofline(eng, P, box, reps):
    """FP32 VALU roofline of the pair kernel (the dominant kernel of the evaluation): algorithmic
    FLOPs = unique pairs within the cutoffs x the per-pair count of the force, energy and virial
    (8 for the distance and periodic image, 17 LJ, 8 + 2 npf Coulomb force, 4 + 2 npe Coulomb
    energy, 6 force and 16 energy / virial accumulation), over the kernel's average duration
    (HIP events on its stream, back-to-back launches of the kernel alone)."""
    import torch
    nl, pf = eng.nl, eng.pairs
    X, Y, Z, Q = P
    N = X.numel()
    f = [torch.empty(N, device="cuda") for _ in range(4)]
    V = torch.empty(6, N, device="cuda")
    stream = torch.cuda.current_stream()
    nl.update(X, Y, Z, Q)
    reps = max(10, min(reps, 50))
    for _ in range(3):
        pf.compute_at(nl, X, Y, Z, Q, *f, V, None, energy=True)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(reps):
        pf.compute_at(nl, X, Y, Z, Q, *f, V, None, energy=True)
    b.record(stream)
    torch.cuda.synchronize()
    pair_ms = a.elapsed_time(b) / reps
    rc = max(9.0, eng.solver.spec.real_cutoff) if eng.solver else 9.0
    pairs = pairs_within(nl, X, Y, Z, box, rc)
    info = pf.info()
    fpp = 8 + 17 + 8 + 2 * info["n_force_terms"] + 4 + 2 * info["n_energy_terms"] + 6 + 16
    flop = pairs * fpp
    achieved = flop / (pair_ms * 1e-3) / 1e12
    visits = int(nl.counts().to(torch.int64).sum().item())
    vpv = pair_flops_per_visit(info["n_energy_terms"], info["n_force_terms"], True)
    return {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": pair_traffic(N),
            "kernel": "k_pair<LJ,coulomb,energy>", "duration_us": round(1e3 * pair_ms, 2), "launches_timed": reps,
            "unique_pairs_within_cutoff": pairs, "flop_per_pair": fpp, "flop_per_launch": flop,
            "algorithmic_bytes_per_launch": visits * 4 + N * 60,
            "visits": {"row_visits": visits, "flop_per_visit": vpv,
                       "tflops_per_visit_count": round(visits * vpv / (pair_ms * 1e-3) / 1e12, 3)},
            "note": "FP32 VALU peak (no MFMA in this kernel); each pair is visited from both ends and the rows "
                    "hold the list skin, so the visit-based rate is higher; traffic = HBM bytes per launch from "
                    "rocprofv3 FETCH_SIZE / WRITE_SIZE scaled by a 256 MiB copy"}


def pair_only(args):
    """Only the pair kernel of the headline's plain configuration (the profiling runs): list build,
    then `steps` pair launches with energies."""
    import numpy as np
    import torch

    import pdx
    sysd = _ion_system(args.c5_sites, 10)
    if not args.no_reorder:
        sysd = spatially_ordered(sysd, args.skin)
    N, box = sysd["N"], sysd["box"]
    P = _tensors(sysd)
    spec = _spec(args.pair_rc, 5, pdx.fft_friendly(int(math.ceil(box[0]))))
    lj = pdx.LennardJones(epsilon=0.01, sigma=2.8, cutoff=9.0)
    eng = pdx.ForceEngine(N, box, lj, electrostatics=spec, charges=sysd["q"], skin=args.skin, settings="auto")
    f = [torch.empty(N, device="cuda") for _ in range(4)]
    V = torch.empty(6, N, device="cuda")
    eng.nl.update(*P)
    for _ in range(args.steps):
        eng.pairs.compute_at(eng.nl, *P, *f, V, None, energy=True)
    torch.cuda.synchronize()
    eng.status()
    from pdx import autotune
    print(json.dumps({"pair_only": True, "N": N, "launches": args.steps, "launch": autotune.current(eng),
                      "row_visits": int(eng.nl.counts().to(torch.int64).sum().item())}))


def eval_only(args):
    """`steps` graph replays (after `warmup`) of the headline system's full plain PME evaluation
    at the given cutoff, order and mesh, at the launch shapes of the measured settings table
    (pdx/engine_settings.json, what the headline's tuning chose), for kernel traces and counters."""
    import torch
    from pdx import autotune
    rc, order, K = args.eval_only.split(",")
    sysd = _ion_system(args.c5_sites, 10)
    if not args.no_reorder:
        sysd = spatially_ordered(sysd, args.skin)
    P = _tensors(sysd)
    ms, eng, _ = _time_eval(P, sysd["box"], _spec(float(rc), int(order), int(K)), args.skin, args.steps, args.warmup,
                            keep=True, settings="auto")
    torch.cuda.synchronize()
    print(json.dumps({"eval_only": True, "N": sysd["N"], "rc": float(rc), "order": int(order), "mesh": int(K),
                      "replays": args.steps, "ms_per_eval": round(ms, 4), "launch": autotune.current(eng)}))
    del eng


def c2_leg(args):
    """The 64 000-ion ionic workload (plain SPME 120^3, P = 5, rc 9 A): evaluation time, per-phase
    event timings, the forced rebuild and its force error; and the cluster-pair backend on it."""
    import numpy as np
    import torch

    import pdx
    from pdx import systems
    sysd = systems.rock_salt(40, 2.82, 0.1, seed=0)
    N, box = sysd["N"], sysd["box"]
    spec = pdx.PME(coulomb_prefactor=C_ION, real_cutoff=9.0, real_space_tolerance=1e-5, mesh_spacing=1.0, order=5)
    P = _tensors(sysd)
    ms, eng, bufs = _time_eval(P, box, spec, args.skin, args.steps, args.warmup, keep=True)
    fx, fy, fz, E, V = bufs
    X, Y, Z, Q = P
    stream = torch.cuda.current_stream()
    ev = {k: [] for k in ("update", "pairs", "spread", "solve", "gather")}
    nl, pf, sol = eng.nl, eng.pairs, eng.solver
    for _ in range(20):
        marks = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
        marks[0].record(stream)
        nl.update(X, Y, Z, Q)
        marks[1].record(stream)
        pf.compute_at(nl, X, Y, Z, Q, fx, fy, fz, E, V, eng.pair_totals, energy=True)
        marks[2].record(stream)
        sol.spread_cells(nl)
        marks[3].record(stream)
        sol.solve(energy=True)
        marks[4].record(stream)
        sol.gather_cells(nl, fx, fy, fz, E, eng.Vglobal, energy=True)
        marks[5].record(stream)
        torch.cuda.synchronize()
        for k, (a, b) in zip(ev, zip(marks[:-1], marks[1:])):
            ev[k].append(a.elapsed_time(b))
    rb = []
    for _ in range(10):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        nl.rebuild(X, Y, Z, Q)
        b.record(stream)
        torch.cuda.synchronize()
        rb.append(a.elapsed_time(b))
    ids = torch.from_numpy(np.sort(np.random.default_rng(1).choice(N, 256, replace=False)).astype(np.int32)).cuda()
    Fref = _ewald_reference(P, box, ids)
    err = _coulomb_error(P, box, spec, args.skin, ids, Fref)
    pair_ms = _time_events(lambda: pf.compute_at(nl, X, Y, Z, Q, fx, fy, fz, E, V, None, energy=True), 50, stream)
    out = {"N": N, "mesh": list(sol.mesh), "order": sol.order, "alpha": sol.alpha, "ms_per_eval": round(ms, 4),
           "evals_per_s": round(1e3 / ms, 2), "phase_ms": {k: round(float(np.mean(v)), 4) for k, v in ev.items()},
           "rebuild_ms": round(float(np.median(rb)), 4), "pair_kernel_ms": round(pair_ms, 4),
           "force_rel_rms_vs_ewald": err}
    del eng
    if not args.no_clusters:
        ws = pdx.DeviceNeighborWorkspace(N, box, 9.0, args.skin, pdx.default_capacity(N, box, 9.0 + args.skin))
        ws.update(X, Y, Z, Q)
        out["clusters"] = cluster_leg(args.steps, args.warmup, args.skin, ws, X, Y, Z, pair_ms)
    return out


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


def coulomb_reference(X, Y, Z, Q, box, C, nsample=256):
    """FP64 Ewald Coulomb forces on `nsample` seeded ions (device kernel, alpha 17/L at 12 digits:
    the real space stays inside one image) and their ids."""
    import numpy as np
    import torch

    import pdx
    N = X.numel()
    ids = torch.from_numpy(np.sort(np.random.default_rng(1).choice(N, nsample, replace=False)).astype(np.int32)).cuda()
    Fref, _ = pdx.ewald_fp64(*(t.double() for t in (X, Y, Z, Q)), box, C, 17.0 / min(box), digits=12.0, sample=ids,
                             want_energy=False)
    return ids, Fref


def coulomb_error(X, Y, Z, Q, box, spec, C, skin, correction=None, ref=None):
    """Relative RMS Coulomb-force error of one mesh setting (optionally with the learned
    correction) against FP64 Ewald on sampled ions (``coulomb_reference``, or ``ref``)."""
    import torch

    import pdx
    N = X.numel()
    eng = pdx.ForceEngine(N, box, None, electrostatics=spec, charges=Q.cpu().numpy(), skin=skin,
                          correction=correction)
    g = [torch.empty(N, device="cuda") for _ in range(3)]
    eng.evaluate(X, Y, Z, Q, *g, None, None, energy=False)
    ids, Fref = ref if ref is not None else coulomb_reference(X, Y, Z, Q, box, C)
    Fg = torch.stack(g, 1).double()[ids.long()]
    del eng
    return math.sqrt(((Fg - Fref) ** 2).sum().item() / (Fref ** 2).sum().item())


# The reference's published Coulomb force calls (docs/reference/manual/electrostatics.md:248-256):
# ions in a dense +-1 e melt, rc 9 A, FP32, RTX 3090, one force call each at a measured force
# error of 1e-4 (or 1e-5): (name, method, sites per side, mesh, reference ms, target error).
PUBLISHED = (("SPME", "spme", 40, 150, 1.14, 1e-4), ("P3M-AD", "p3m_ad", 40, 150, 1.09, 1e-4),
             ("P3M-ik", "p3m_ik", 40, 120, 1.15, 1e-4), ("ESP", "esp", 40, 50, 0.74, 1e-4),
             ("ESP 1e-5", "esp", 40, 63, 1.6, 1e-5), ("ESP 512k", "esp", 80, 100, 7.6, 1e-4),
             ("SPME 512k", "spme", 80, 300, 13.5, 1e-4))


def published_leg(steps, warmup, skin):
    """The reference's published Coulomb force calls at the same methods, meshes and sizes, on
    64 000 (40^3 sites) and 512 000 (80^3) +-1 e ions (rock salt with 0.5 A random displacements:
    disordered like the reference's melt, whose relative force errors a near-perfect lattice would
    inflate), Coulomb only, rc 9 A: per method the setting with the smallest measured force error
    against FP64 Ewald is kept (PME / P3M: order and real-space tolerance; ESP: the loosest tolerance
    of its rules that meets the target, which sets the split and window shapes and the width), and
    one force call (list
    check with predicated rebuild, real-space pairs, spread, solve, gather; forces only) is timed
    as a graph replay. "matched" says whether our error is within the reference's target."""
    import torch

    import pdx
    from pdx import systems
    C = 14.399645
    res = {"systems": "+-1 e ions (rock salt, 0.5 A random displacements), Coulomb only, rc 9 A",
           "reference": "RTX 3090, dense +-1 e melt, measured force error 1e-4 / 1e-5 (electrostatics.md:248-256)"}
    cache = {}

    def spec_of(kind, K, P, tol):
        if kind == "spme":
            return pdx.PME(coulomb_prefactor=C, real_cutoff=9.0, real_space_tolerance=tol, mesh=(K, K, K), order=P)
        if kind == "esp":
            return pdx.ESP(coulomb_prefactor=C, real_cutoff=9.0, tolerance=tol, mesh=(K, K, K))
        return pdx.PPPM(coulomb_prefactor=C, real_cutoff=9.0, real_space_tolerance=tol, mesh=(K, K, K), order=P,
                        differentiation=kind[-2:])

    for name, kind, sites, K, ref_ms, target in PUBLISHED:
        if sites not in cache:
            sysd = systems.rock_salt(sites, 2.82, 0.5, seed=0)
            P4 = tuple(torch.from_numpy(sysd[k]).cuda() for k in ("x", "y", "z", "q"))
            cache[sites] = (sysd, P4, coulomb_reference(*P4, sysd["box"], C))
        sysd, (X, Y, Z, Q), ref = cache[sites]
        N, box = sysd["N"], sysd["box"]
        trials = []
        if kind == "esp":
            for tol in (target, target / 2, target / 4):
                if pdx.esp_rule(tol)[1] > 8:  # widths above 8 are not compiled
                    continue
                spec = spec_of(kind, K, None, tol)
                trials.append((coulomb_error(X, Y, Z, Q, box, spec, C, skin, ref=ref), tol, spec, spec.order))
        else:
            for P in ((4, 5) if kind == "spme" else (5,)):
                for tol in (3e-5, 1e-5, 3e-6, 1e-6):
                    spec = spec_of(kind, K, P, tol)
                    trials.append((coulomb_error(X, Y, Z, Q, box, spec, C, skin, ref=ref), tol, spec, P))
        if kind == "esp":
            # the loosest tolerance that meets the target: its rules give the narrowest window
            # (cheapest spread and gather) at this mesh; the most accurate one if none does
            ok = [t for t in trials if t[0] <= target]
            err, tol, spec, _ = max(ok, key=lambda t: t[1]) if ok else min(trials, key=lambda t: t[0])
        else:
            err, tol, spec, _ = min(trials, key=lambda t: t[0])
        _log(f"published {name}: error {err:.2e}")
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
        entry = {"N": N, "mesh": K, "order": eng.solver.order, "ms_per_force_call": round(ms, 4),
                 "force_rel_rms_vs_ewald": err, "target": target, "matched": bool(err <= target),
                 "reference_ms": ref_ms, "reference_over_ours": round(ref_ms / ms, 2)}
        if kind == "esp":
            entry.update(tolerance=tol, split_shape=round(spec.split_shape, 4), window_shape=round(spec.window_shape, 4),
                         errors_by_tolerance={f"{t:g}": e for e, t, _, _ in trials})
        else:
            entry.update(real_space_tolerance=tol, alpha=round(spec.resolved_alpha(), 5),
                         errors_by_order_and_tolerance={f"P{P} {t:g}": e for e, t, _, P in trials})
        res[name] = entry
        del eng
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


def cpu_baseline(sysd, plain):
    """The same algorithm on the host in FP64 (oracle/: cell-list LJ + real-space Ewald, B-spline
    spread and analytic-derivative gather, pocketfft R2C / C2R through scipy) at the headline's
    plain PME configuration, one full evaluation of the headline system timed on all host cores
    (its influence function is set-up work, outside the timing); plus a one-thread figure on the
    60 800-ion system of the same construction (mesh scaled with the box)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import mesh as cpu_mesh
    import oracle
    import pdx
    rc, order, K = plain["rc"], plain["order"], plain["mesh"]
    alpha = pdx.alpha_from_tolerance(rc, 1e-5)
    lj = dict(kind="lj", sigma=2.8, epsilon=0.01, cutoff=9.0)

    def run(s, Kk):
        x, y, z, q = (s[k].astype(np.float64) for k in ("x", "y", "z", "q"))
        m = cpu_mesh.Mesh((Kk, Kk, Kk), s["box"], C_ION, alpha, order)
        t0 = time.perf_counter()
        cpu_mesh.evaluate(x, y, z, q, s["box"], C_ION, alpha, rc, (Kk, Kk, Kk), order, lj=lj, mesh=m)
        return time.perf_counter() - t0

    cores = oracle.num_threads()
    t_all = run(sysd, K)
    small = _ion_system(40, 10)
    k_small = pdx.fft_friendly(int(math.ceil(K * small["box"][0] / sysd["box"][0])))
    prev = oracle.set_threads(1)
    t_one = run(small, k_small)
    oracle.set_threads(prev)
    return {"value": round(1.0 / t_all, 4), "unit": "force evals/s", "cores": cores, "kind": "port",
            "sample": f"one full FP64 evaluation of the {sysd['N']}-ion headline system (same algorithm: rc {rc} A, "
                      f"order {order}, {K}^3 mesh): {t_all:.2f} s on {cores} threads",
            "single_thread": {"N": small["N"], "mesh": k_small, "s_per_eval": round(t_one, 2),
                              "evals_per_s": round(1.0 / t_one, 4)}}


if __name__ == "__main__":
    main()
