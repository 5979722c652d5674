"""Device cell list and ELL rows against the FP64 brute-force oracle
(test/neighbor_reference.jl:6-16; test/test_neighbor_integrity.jl:6-75)."""
import math

import numpy as np
import pytest

import oracle
import pdx
from conftest import dev
from pdx import systems

pytestmark = pytest.mark.gpu


def _ulp32(v):
    return float(np.spacing(np.float32(v)))


def _gpu_rows(x, y, z, box, cutoff, skin, cap, allpairs=False):
    X, Y, Z = dev(x), dev(y), dev(z)
    ws = pdx.build_neighbors_dense(X, Y, Z, box, cutoff, skin, cap, allpairs=allpairs)
    return ws, ws.rows_host()


@pytest.mark.parametrize("N", [1, 2, 33])
def test_rows_small_random(cuda, N):
    print(f"[rows] random N={N}", flush=True)
    rng = np.random.default_rng(0xE110 + N)
    L = 20.0
    c = [(20 * rng.random(N) - 10).astype(np.float32) for _ in range(3)]
    rl = float(np.float32(3) + np.float32(0.4))
    ws, rows = _gpu_rows(*c, (L, L, L), 3.0, 0.4, max(N - 1, 1))
    exp, counts, _ = oracle.neighbor_rows(tuple(a.astype(np.float64) for a in c), (L, L, L), np.full(N, rl))
    assert rows == [list(r) for r in exp]
    st = ws.state()
    assert st[8] == max(counts.max(), 0)  # required capacity


def test_rows_fcc_exact(cuda):
    """Bit-exact pair sets at the LJ-fluid size, tie band empty (no pair within 4 ulp of rl^2)."""
    cutoff, skin = 2.5, 0.4
    rl = float(np.float32(cutoff) + np.float32(skin))
    band = 4 * _ulp32(np.float32(rl) * np.float32(rl))
    for seed in range(20):
        s = systems.fcc_lattice(4000, 0.8442, 0.05, seed=seed)
        exp, counts, margin = oracle.neighbor_rows((s["x"].astype(np.float64), s["y"].astype(np.float64),
                                                    s["z"].astype(np.float64)), s["box"], np.full(4000, rl))
        if margin > band:
            break
    print(f"[rows] fcc seed={seed} margin={margin:.3e} band={band:.3e}", flush=True)
    assert margin > band
    cap = pdx.default_capacity(4000, s["box"], rl)
    ws, rows = _gpu_rows(s["x"], s["y"], s["z"], s["box"], cutoff, skin, cap)
    assert rows == [list(r) for r in exp]
    assert ws.status().required == counts.max()
    assert int(ws.counts().sum().item()) == int(counts.sum())


def test_sorted_ids_and_offsets(cuda):
    s = systems.fcc_lattice(4000, 0.8442, 0.05, seed=3)
    ws, _ = _gpu_rows(s["x"], s["y"], s["z"], s["box"], 2.5, 0.4, 200)
    ids = ws.particle_ids_sorted().cpu().numpy()
    off = ws.cell_offsets().cpu().numpy()
    cell = ws.cell_of_particle().cpu().numpy()
    assert off[0] == 1 and off[-1] == 4001
    assert sorted(ids.tolist()) == list(range(1, 4001))
    for c in range(ws.ncell):
        seg = ids[off[c] - 1:off[c + 1] - 1]
        assert np.all(np.diff(seg) > 0)  # stable: ascending id inside a cell
        assert np.all(cell[seg - 1] == c)


def test_allpairs_rows(cuda):
    rng = np.random.default_rng(7)
    N, L = 50, 8.0
    c = [(L * rng.random(N) - L / 2).astype(np.float32) for _ in range(3)]
    rl = float(np.float32(2.5) + np.float32(0.4))
    ws, rows = _gpu_rows(*c, (L, L, L), 2.5, 0.4, N - 1, allpairs=True)
    exp, _, _ = oracle.neighbor_rows(tuple(a.astype(np.float64) for a in c), (L, L, L), np.full(N, rl))
    assert rows == [list(r) for r in exp]


def test_capacity_overflow_and_recovery(cuda):
    """test/test_neighbor_integrity.jl:41-75: overflow rejects rows and reports the required cap."""
    import torch
    L = 20.0
    x = torch.zeros(4, device="cuda")
    y = torch.zeros(4, device="cuda")
    z = torch.zeros(4, device="cuda")
    with pytest.raises(pdx.NeighborCapacityError) as e:
        pdx.build_neighbors_dense(x, y, z, (L, L, L), 2.0, 0.3, 2)
    assert e.value.capacity == 2 and e.value.required == 3
    ok = pdx.build_neighbors_dense(x, y, z, (L, L, L), 2.0, 0.3, 3)
    assert ok.status().required == 3
    assert (ok.counts().cpu().numpy() == 3).all()
    # a valid list, then an overflowing configuration: error latched, refs untouched
    xs = torch.tensor([-6.0, -2.0, 2.0, 6.0], device="cuda")
    ws = pdx.build_neighbors_dense(xs, y, z, (L, L, L), 2.0, 0.3, 2)
    refs = ws._arrays
    before = pdx.neighbors.device_view(refs.rref_x, (4,), torch.float32).clone()
    ws.update(x, y, z)
    with pytest.raises(pdx.NeighborCapacityError) as e:
        ws.status()
    assert e.value.required == 3
    assert torch.equal(pdx.neighbors.device_view(refs.rref_x, (4,), torch.float32), before)
    # recover: valid configuration again, rearm, rebuild with the same capacity
    ws.rearm()
    ws.update(xs, y, z)
    st = ws.status()
    assert st.valid and st.required == 0


def test_coordinate_error_latches(cuda):
    import torch
    x = torch.tensor([0.0, 1.0, float("nan")], device="cuda")
    y = torch.zeros(3, device="cuda")
    with pytest.raises(pdx.ArgumentError, match="finite"):
        pdx.build_neighbors_dense(x, y, y.clone(), (20.0, 20.0, 20.0), 2.0, 0.3, 4)


def test_rebuild_policy(cuda):
    """max |dr|^2 > skin^2/4 requests a build; smaller moves keep the list; the age limit too."""
    import torch
    s = systems.fcc_lattice(4000, 0.8442, 0.05, seed=1)
    X, Y, Z = dev(s["x"]), dev(s["y"]), dev(s["z"])
    ws = pdx.DeviceNeighborWorkspace(4000, s["box"], 2.5, 0.4, 200, interval=20)
    ws.update(X, Y, Z)
    assert ws.status().builds == 1
    ws.advance()
    ws.update(X + 0.099, Y, Z)  # 0.099^2 < 0.04
    assert ws.status().builds == 1
    ws.advance()
    ws.update(X + 0.201, Y, Z)  # 0.201^2 > 0.04
    st = ws.status()
    assert st.builds == 2 and st.interval == 18  # round(0.9 * 20)
    for _ in range(18):
        ws.advance()
    ws.update(X + 0.201, Y, Z)
    assert ws.status().builds == 3  # aged out


def test_rows_and_forces_2d(cuda):
    """2-D lists (DenseNeighborCUDA.jl:77, D = 2): rows bit-exact against the 2-D brute force
    (tie band empty), and LJ forces / energies against the 2-D direct sums of
    test/pair_reference_gpu.jl (oracle.pair_typed, D = 2); z is absent (NULL)."""
    import torch
    cutoff, skin = 2.5, 0.3
    rl = float(np.float32(cutoff) + np.float32(skin))
    band = 4 * _ulp32(np.float32(rl) * np.float32(rl))
    n, L = 3000, 60.0  # area density 0.83
    for seed in range(20):
        rng = np.random.default_rng(100 + seed)
        m = int(math.ceil(math.sqrt(n)))
        g = np.stack(np.meshgrid(np.arange(m), np.arange(m), indexing="ij"), -1).reshape(-1, 2)[:n]
        P = (g + 0.5) * (L / m) - L / 2 + rng.normal(0, 0.05, (n, 2))
        P = (P - L * np.floor(P / L + 0.5)).astype(np.float32)
        exp, counts, margin = oracle.neighbor_rows((P[:, 0].astype(np.float64), P[:, 1].astype(np.float64)), (L, L),
                                                   np.full(n, rl))
        if margin > band:
            break
    assert margin > band
    X, Y = dev(P[:, 0]), dev(P[:, 1])
    cap = pdx.default_capacity(n, (L, L), rl)
    ws = pdx.build_neighbors_dense(X, Y, None, (L, L), cutoff, skin, cap)
    assert ws.rows_host() == [list(r) for r in exp]
    eng = pdx.ForceEngine(n, (L, L), pdx.LennardJones(1.0, 1.0, 2.5), skin=skin)
    f = [torch.empty(n, device="cuda") for _ in range(3)]
    E = torch.empty(n, device="cuda")
    eng.evaluate(X, Y, None, None, *f, E, None, energy=True)
    tot = eng.totals()
    ref = oracle.pair_typed((P[:, 0].astype(np.float64), P[:, 1].astype(np.float64)), (L, L), kind="lj", cutoff=2.5)
    F = torch.stack(f[:2], 1).double().cpu().numpy()
    err = np.linalg.norm(F - ref["F"]) / np.linalg.norm(ref["F"])
    print(f"[rows 2d] seed={seed} force {err:.2e}", flush=True)
    assert err < 2e-5
    assert torch.count_nonzero(f[2]).item() == 0
    assert abs(tot["energy"] - ref["energy"].sum()) < 1e-5 * abs(ref["energy"].sum())
    with pytest.raises(pdx.ArgumentError):
        pdx.ForceEngine(n, (L, L), None, electrostatics=pdx.PME(coulomb_prefactor=1.0, real_cutoff=2.5, alpha=1.0,
                                                               mesh=(32, 32, 32), order=4), charges=np.zeros(n))


@pytest.mark.parametrize("case", ["fcc4000", "rocksalt64k"])
def test_rows_staged_equal_direct_and_lanes(cuda, case, monkeypatch):
    """The row build staged in LDS writes the same rows, slot for slot, as the build that stores
    straight from the search (the default), for every builder-lane count (rows are identical across
    builder lanes, src/engine/StateEngine.jl:15-19), in storage and spatial order; so does the
    cell-staged build (PDX_ROW_CELLS=1)."""
    import torch
    if case == "fcc4000":
        s = systems.fcc_lattice(4000, 0.8442, 0.05, seed=3)
        cutoff, skin = 2.5, 0.4
    else:
        s = systems.rock_salt(40, 2.82, 0.1, seed=0)
        cutoff, skin = 9.0, 1.0
    N, box = len(s["x"]), s["box"]
    cap = pdx.default_capacity(N, box, cutoff + skin)
    X, Y, Z = dev(s["x"]), dev(s["y"]), dev(s["z"])

    def rows(stage, lanes=None, cells=False):
        monkeypatch.setenv("PDX_ROW_STAGE", "1" if stage else "0")
        monkeypatch.setenv("PDX_ROW_CELLS", "1" if cells else "0")
        ws = pdx.DeviceNeighborWorkspace(N, box, cutoff, skin, cap)
        if lanes:
            ws.set_row_lanes(lanes)
        ws.update(X, Y, Z, None)
        st = ws.status()
        cnt = ws.counts().clone()
        flat = ws.neighbors_flat().view(cap, N).clone()
        mask = torch.arange(cap, device="cuda")[:, None] < cnt[None, :]
        return cnt, torch.where(mask, flat, torch.zeros_like(flat)), st.builds

    c0, r0, b0 = rows(False)  # k_rows
    assert b0 == 1
    for stage, lanes in ((True, None), (True, 1), (True, 2), (True, 32), (False, 8)):
        c1, r1, _ = rows(stage, lanes)
        assert torch.equal(c0, c1) and torch.equal(r0, r1), (stage, lanes)
    # the cell-staged build (one workgroup per cell, candidates from LDS)
    for lanes in (None, 1, 4, 8, 32):
        c1, r1, _ = rows(False, lanes, cells=True)
        assert torch.equal(c0, c1) and torch.equal(r0, r1), ("cells", lanes)
