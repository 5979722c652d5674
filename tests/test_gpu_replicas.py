"""Batched replicas: forces against the FP64 direct sums, short trajectories against the
FP64 velocity-Verlet oracle, every replica of a batch bitwise equal to the same box run
alone, launch shapes, graph replay, energy conservation."""
import numpy as np
import pytest

import oracle
import pdx
from pdx import replicas, systems

pytestmark = pytest.mark.gpu


def _host(state, k, off):
    a, b = off[k], off[k + 1]
    return {n: state[n][a:b].cpu().numpy().astype(np.float64) for n in state}


def _small_batch(seed=3, nrep=12, n_range=(256, 1500)):
    return systems.replica_batch(nrep, seed=seed, n_range=n_range)


def test_forces_match_direct_sums(cuda):
    b = _small_batch()
    for stencil in (1, 2):
        rb = pdx.ReplicaBatch(b["specs"], block=256, stencil=stencil)
        st = replicas.device_state(b)
        out = rb.forces(st).cpu().numpy()
        for k, s in enumerate(b["specs"]):
            h = _host(st, k, rb.offsets)
            d = oracle.pair_direct(h["x"], h["y"], h["z"], s["box"], kind=s["model"], cutoff=s.get("cutoff", 2.5))
            F = np.stack([h["fx"], h["fy"], h["fz"]], 1)
            err = np.linalg.norm(F - d["F"]) / np.linalg.norm(d["F"])
            assert err < 2e-5, (stencil, k, s["model"], s["n"], err)
            pe = d["energy"].sum()
            if s["model"] == "lj":
                s6 = (1.0 / 2.5) ** 6
                pe -= 0.5 * 4 * (s6 * s6 - s6) * d["counts"].sum()
            assert abs(out[k, 0] - pe) < 1e-5 * max(1.0, abs(pe)), (k, out[k, 0], pe)
            assert abs(out[k, 2] - d["virial"][:, :3].sum()) < 1e-5 * max(1.0, np.abs(d["virial"]).sum())
            assert out[k, 3] == s["n"]
    info = rb.info()
    print(f"[replicas] {info}", flush=True)


def test_all_pairs_boxes(cuda):
    """Boxes with fewer than three cells a side take the all-pairs traversal."""
    specs = [{"model": "lj", "n": 300, "box": (6.8, 6.8, 6.8), "cutoff": 2.5},
             {"model": "wca", "n": 256, "box": (6.6, 6.6, 6.6)}]
    rng = np.random.default_rng(4)
    xs = [rng.random((s["n"], 3)) * s["box"][0] - s["box"][0] / 2 for s in specs]
    X = np.concatenate(xs).astype(np.float32)
    bd = dict(x=X[:, 0], y=X[:, 1], z=X[:, 2], vx=np.zeros(len(X)), vy=np.zeros(len(X)), vz=np.zeros(len(X)))
    rb = pdx.ReplicaBatch(specs)
    assert rb.info()["allpairs_replicas"] == 1
    st = replicas.device_state(bd)
    rb.forces(st)
    for k, s in enumerate(specs):
        h = _host(st, k, rb.offsets)
        d = oracle.pair_direct(h["x"], h["y"], h["z"], s["box"], kind=s["model"], cutoff=s.get("cutoff", 2.5))
        F = np.stack([h["fx"], h["fy"], h["fz"]], 1)
        assert np.linalg.norm(F - d["F"]) / np.linalg.norm(d["F"]) < 2e-5


def test_trajectory_matches_fp64_verlet(cuda):
    b = _small_batch(seed=5, nrep=6, n_range=(256, 600))
    rb = pdx.ReplicaBatch(b["specs"])
    st = replicas.device_state(b)
    h0 = [_host(st, k, rb.offsets) for k in range(rb.nrep)]
    rb.forces(st, energy=False)
    rb.run(st, 20, dt=0.005)
    for k, s in enumerate(b["specs"]):
        h = _host(st, k, rb.offsets)
        X, V, F, *_ = oracle.vv_steps(h0[k]["x"], h0[k]["y"], h0[k]["z"], h0[k]["vx"], h0[k]["vy"], h0[k]["vz"],
                                      s["box"], kind=s["model"], cutoff=s.get("cutoff", 2.5), dt=0.005, nsteps=20)
        L = s["box"][0]
        d = np.stack([h["x"], h["y"], h["z"]], 1) - X
        d -= L * np.round(d / L)
        dv = np.stack([h["vx"], h["vy"], h["vz"]], 1) - V
        print(f"[replicas] {s['model']} n={s['n']} max|dx|={np.abs(d).max():.2e} max|dv|={np.abs(dv).max():.2e}",
              flush=True)
        assert np.abs(d).max() < 1e-4
        assert np.abs(dv).max() < 1e-3
        assert np.all(np.abs(np.stack([h["x"], h["y"], h["z"]], 1)) <= L / 2)


def test_batch_equals_each_box_alone(cuda):
    b = _small_batch(seed=7, nrep=24, n_range=(256, 3000))
    rb = pdx.ReplicaBatch(b["specs"])
    st = replicas.device_state(b)
    init = replicas.clone_state(st)
    rb.forces(st, energy=False)
    out = rb.run(st, 100, dt=0.005, energy=True).cpu().numpy()
    for k in (0, 5, 11, 23):
        s = b["specs"][k]
        a, e = rb.offsets[k], rb.offsets[k + 1]
        one = pdx.ReplicaBatch([s])
        st1 = {n: v[a:e].clone() for n, v in init.items()}
        one.forces(st1, energy=False)
        o1 = one.run(st1, 100, dt=0.005, energy=True).cpu().numpy()
        for n in st:
            assert np.array_equal(st[n][a:e].cpu().numpy(), st1[n].cpu().numpy()), (k, n)
        assert np.array_equal(out[k], o1[0]), k


def test_launch_shapes_and_graph_replay(cuda):
    import torch
    b = _small_batch(seed=9, nrep=10, n_range=(256, 2000))
    rb = pdx.ReplicaBatch(b["specs"], block=256, stencil=1)
    st0 = replicas.device_state(b)
    rb.forces(st0, energy=False)
    ref = replicas.clone_state(st0)
    rb.run(ref, 30)
    # the workgroup size does not change a replica's arithmetic
    for blk in (512, 1024):
        rb.set_launch(blk, 1)
        t = replicas.clone_state(st0)
        rb.run(t, 30)
        for n in t:
            assert torch.equal(t[n], ref[n]), (blk, n)
    # a graph of one 10-step launch, replayed three times, equals 30 eager steps
    rb.set_launch(256, 1)
    g_state = replicas.clone_state(st0)
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        rb.run(g_state, 10)  # warm on the capture stream
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g_state = replicas.clone_state(st0)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        rb.run(g_state, 10)
    for _ in range(3):
        graph.replay()
        torch.cuda.synchronize()
    for n in g_state:
        assert torch.equal(g_state[n], ref[n]), n
    best, times = replicas.autotune(rb, st0, steps=5, candidates=[(256, 1), (512, 2)])
    print(f"[replicas] autotune {best} {times}", flush=True)
    assert best in [(256, 1), (512, 2)]


def test_energy_conservation(cuda):
    """Over 1000 steps the total energy of each replica drifts as the FP64 integration of
    the same box does (the drift of these hot starts at dt = 0.005 is integration error,
    O(dt^2), not rounding)."""
    b = _small_batch(seed=11, nrep=4, n_range=(256, 700))
    rb = pdx.ReplicaBatch(b["specs"])
    st = replicas.device_state(b)
    h0 = [_host(st, k, rb.offsets) for k in range(rb.nrep)]
    e0 = rb.forces(st).cpu().numpy()
    e1 = rb.run(st, 1000, dt=0.005, energy=True).cpu().numpy()
    for k, s in enumerate(b["specs"]):
        E0, E1 = e0[k, 0] + e0[k, 1], e1[k, 0] + e1[k, 1]
        drift = (E1 - E0) / s["n"]
        h = h0[k]
        kw = dict(kind=s["model"], cutoff=s.get("cutoff", 2.5), shift=s["model"] == "lj", dt=0.005)
        *_, pa, ka, _ = oracle.vv_steps(h["x"], h["y"], h["z"], h["vx"], h["vy"], h["vz"], s["box"], nsteps=0, **kw)
        *_, pb, kb, _ = oracle.vv_steps(h["x"], h["y"], h["z"], h["vx"], h["vy"], h["vz"], s["box"], nsteps=1000,
                                        **kw)
        ref = ((pb + kb) - (pa + ka)) / s["n"]
        print(f"[replicas] {s['model']} n={s['n']} E/n {E0 / s['n']:.4f} drift {drift:.3e} (FP64 {ref:.3e})",
              flush=True)
        assert abs(E0 - (pa + ka)) < 1e-5 * abs(pa + ka) + 1e-3
        assert abs(drift - ref) < 1e-3 + 0.1 * abs(ref), (k, drift, ref)
