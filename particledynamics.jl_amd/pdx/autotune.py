"""Launch-shape autotuner of the force evaluation: the MI355X counterpart of the reference's
ENGINE_SETTINGS table (src/engine/StateEngine.jl:7-56), which keeps, per size class and
neighbours-per-atom class, the measured best builder lanes, force lanes, skin and fusion.

Here a setting holds the knobs of one ``ForceEngine`` evaluation:
  * ``force_lanes``: lanes per particle of the pair kernel's row traversal;
  * ``build_lanes``: lanes per particle of the row build;
  * ``overlap``: how the pair kernel shares the device with the mesh branch ("fft": it waits for
    the spread and runs beside the transforms; "spread": it runs beside the whole mesh branch);
  * ``spread_threads``, ``transform_threads``: workgroup sizes of the owner-computes spread that
    takes every particle reaching a tile (sparse meshes) and of the plane passes.
``tune`` measures candidate settings on the caller's own system (graph replays of the captured
evaluation, as the bench times it) and applies the fastest; ``classify`` gives the reference's
classes (size: N <= 30 000, <= 300 000, larger; neighbours rho 4 pi/3 rc^3 below 9.8, 36, 72, or
above) so that measured settings can be stored in a table (JSON: ``save`` / ``load``) and looked
up for a new system of the same class (``settings_for``).
"""
from __future__ import annotations

import itertools
import json
import math
import os
import time

from . import _lib as L

TABLE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "engine_settings.json")
KNOBS = ("force_lanes", "build_lanes", "overlap", "spread_threads", "transform_threads")


def classify(N, box, cutoff):
    """(size class, neighbours class, mesh flag) of a system (engine_settings, StateEngine.jl:47-56)."""
    if len(box) == 2:
        nb = N / (box[0] * box[1]) * math.pi * cutoff ** 2
    else:
        nb = N / (box[0] * box[1] * box[2]) * 4.0 * math.pi / 3.0 * cutoff ** 3
    size = 1 if N <= 30_000 else (2 if N <= 300_000 else 3)
    dens = 1 if nb < 9.8 else (2 if nb < 36.0 else (3 if nb < 72.0 else 4))
    return size, dens


def _key(size, dens, mesh):
    return f"{size},{dens},{'mesh' if mesh else 'pairs'}"


def load(path=TABLE_PATH):
    """The measured table: {class key: setting}; empty when there is none."""
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save(table, path=TABLE_PATH):
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def settings_for(N, box, cutoff, mesh, table=None):
    """The table's setting for the class of this system, or None."""
    table = load() if table is None else table
    size, dens = classify(N, box, cutoff)
    s = table.get(_key(size, dens, mesh))
    return {k: v for k, v in s.items() if k in KNOBS} if s else None


def current(engine):
    """The setting an engine runs with now."""
    s = {"force_lanes": getattr(engine.pairs, "lanes", 0),
         "build_lanes": getattr(engine.nl, "row_lanes_set", 0) if engine.nl else 0, "overlap": engine.overlap_with}
    if engine.solver is not None:
        s.update(spread_threads=engine.solver.spread_threads, transform_threads=engine.solver.transform_threads)
    return s


def apply(engine, setting):
    """Apply a setting to an engine (before capturing its graph). Unknown or absent knobs are kept."""
    if engine.nl is None:
        raise L.ArgumentError("settings apply to the ELL backend")
    if "force_lanes" in setting:
        engine.pairs.set_lanes(int(setting["force_lanes"]))
    if setting.get("build_lanes"):
        engine.nl.set_row_lanes(int(setting["build_lanes"]))
    if "overlap" in setting:
        if setting["overlap"] not in ("spread", "solve", "fft"):
            raise L.ArgumentError("overlap must be 'spread', 'solve' or 'fft'")
        engine.overlap_with = setting["overlap"]
    if engine.solver is not None:
        if "spread_threads" in setting:
            engine.solver.set_spread_threads(int(setting["spread_threads"]))
        if "transform_threads" in setting:
            engine.solver.set_transform_threads(int(setting["transform_threads"]))
    engine._graph = None


def candidates(engine, force_lanes=(1, 2, 4, 8), overlaps=("fft", "spread"), spread_threads=(256, 512),
               transform_threads=(512, 1024)):
    """The grid of settings ``tune`` measures by default (the transform sizes only where the mesh
    takes the plane passes)."""
    out = []
    if engine.solver is None:
        for fl in force_lanes:
            out.append({"force_lanes": fl})
        return out
    tt = transform_threads if engine.solver.plane_lines else (512,)
    for fl, ov, st, t in itertools.product(force_lanes, overlaps, spread_threads, tt):
        out.append({"force_lanes": fl, "overlap": ov, "spread_threads": st, "transform_threads": t})
    return out


def time_setting(engine, x, y, z, q, bufs, setting, steps=10, warmup=3, energy=True):
    """ms per evaluation of one setting: apply, capture the evaluation, replay."""
    import torch
    apply(engine, setting)
    fx, fy, fz, E, V = bufs
    engine.capture(x, y, z, q, fx, fy, fz, E, V, energy=energy)
    for _ in range(warmup):
        engine.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        engine.replay()
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    engine.status(check=True)
    return ms


def tune(engine, x, y, z, q, cands=None, steps=10, warmup=3, energy=True, rounds=2):
    """Measure every candidate setting (``rounds`` interleaved passes, the best time of each kept,
    so a drift of the device clock does not favour one candidate) on this engine and inputs,
    apply the fastest, and return {"best": setting, "ms": its time, "table": [(setting, ms)]}."""
    import torch
    N = engine.N
    bufs = [torch.empty(N, device="cuda") for _ in range(4)] + [torch.empty(6, N, device="cuda")]
    cands = candidates(engine) if cands is None else list(cands)
    base = current(engine)
    best_ms = [math.inf] * len(cands)
    for _ in range(max(1, rounds)):
        for i, c in enumerate(cands):
            s = dict(base)
            s.update(c)
            best_ms[i] = min(best_ms[i], time_setting(engine, x, y, z, q, bufs, s, steps, warmup, energy))
    i = min(range(len(cands)), key=lambda k: best_ms[k])
    best = dict(base)
    best.update(cands[i])
    apply(engine, best)
    return {"best": best, "ms": best_ms[i],
            "table": [(dict(base, **c), round(ms, 4)) for c, ms in zip(cands, best_ms)]}


def record(engine, result, table=None, path=TABLE_PATH):
    """Store a tuning result under the engine's class in the table (and the file)."""
    table = load(path) if table is None else table
    size, dens = classify(engine.N, engine.box, engine.cutoff)
    entry = dict(result["best"])
    entry["ms_measured"] = round(result["ms"], 4)
    entry["N_measured"] = engine.N
    table[_key(size, dens, engine.solver is not None)] = entry
    save(table, path)
    return table
