"""The allocation-free FP64 Ewald entry point (pdx_ewald_fp64_noalloc) through the C ABI against
pdx_ewald_fp64: the same kernels on a static device workspace, so forces, energy and virial
must be bit for bit equal, on two calls in a row (the second on a smaller system, which must
not see the first call's scratch contents), with and without the neutralising background; a
wave-vector box beyond the static workspace is refused with an argument error."""
import numpy as np
import pytest

import oracle
import pdx
from pdx import systems

pytestmark = pytest.mark.gpu


def _dev64(a):
    import torch
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def _call(R, q, box, C, alpha, digits, ids, background, no_alloc):
    import torch
    xs = [_dev64(R[:, d]) for d in range(3)]
    F = torch.full((ids.numel(), 3), float("nan"), dtype=torch.float64, device="cuda")
    F, tot = pdx.ewald_fp64(*xs, _dev64(q), box, C, alpha, digits=digits, background=background, sample=ids,
                            F=F, no_alloc=no_alloc)
    return F.cpu().numpy(), tot


@pytest.mark.parametrize("background", [False, True])
def test_noalloc_equals_allocating_entry_point(cuda, background):
    import torch
    cases = [  # (particles, box, C, alpha, digits, seed): the second is smaller in every respect
        (240, (16.0, 17.0, 18.0), 1.7, 0.55, 13.0, 5),
        (48, (9.0, 9.5, 10.0), 1.1, 0.7, 11.0, 8),
    ]
    for n, box, C, alpha, digits, seed in cases:
        R, q = systems.random_cloud(n, box, seed=seed)
        if background:
            q = q + 0.03
        ids = torch.arange(0, n, 3, dtype=torch.int32, device="cuda")
        Fa, ta = _call(R, q, box, C, alpha, digits, ids, background, no_alloc=False)
        Fb, tb = _call(R, q, box, C, alpha, digits, ids, background, no_alloc=True)
        print(f"[noalloc] n={n} box={box} background={background}: E={tb['energy']:.12e}", flush=True)
        assert np.array_equal(Fa, Fb), n
        assert ta["energy"] == tb["energy"], n
        assert ta["virial"] == tb["virial"], n
        # and both are the FP64 Ewald sum (the CPU oracle, test/ewald_reference.jl:50-126)
        Eref, Fref, _ = oracle.ewald(R, q, np.diag(box), C_=C, alpha=alpha, digits=14, background=background)
        assert np.abs(Fb - Fref[::3]).max() < 1e-9 * np.abs(Fref).max()
        assert tb["energy"] == pytest.approx(Eref, rel=1e-10)


def test_noalloc_refuses_a_box_beyond_its_workspace(cuda):
    import torch
    R, q = systems.random_cloud(16, (40.0, 40.0, 40.0), seed=2)
    ids = torch.arange(4, dtype=torch.int32, device="cuda")
    # kmax L / 2 pi = 2 * 3.0 * sqrt(14 ln 10) * 40 / 2 pi ~ 218 -> ~ 2e7 box indices
    with pytest.raises(pdx.ArgumentError, match="static scratch"):
        _call(R, q, (40.0, 40.0, 40.0), 1.0, 3.0, 14.0, ids, False, no_alloc=True)
