"""Edge cases of the HIP action and solve against the CPU oracle, through the C-ABI: a cell without atoms (no
dyadic term), atoms with numax 0 (one SHO coefficient each), and a truncation radius so small that every column
group holds its source block alone (every stencil neighbour absent).  Tolerances as elsewhere: FP64 action 1e-12,
FP32 action 2e-5 (max-norm, relative to the FP64 oracle), solves to the true residual 1e-9 (oracle action)."""
import numpy as np
import pytest

from angstromcube_amd import gfa as G
from angstromcube_amd import synthetic as S
from oracle import oracle as O
from tests.test_solve_gpu import true_residuals

pytestmark = pytest.mark.gpu

EDGE = [
    ("no_atoms", lambda: S.make_system(21, (4, 4, 4), ngauss=4, natoms=0, bc=(1, 1, 1)), dict(r_trunc=6.0),
     complex(-0.4, 0.1)),
    ("numax0", lambda: S.make_system(22, (4, 4, 4), ngauss=4, natoms=3, numax=0, bc=(1, 1, 1)), dict(r_trunc=6.0),
     complex(-0.4, 0.1)),
    ("one_block_per_column", lambda: S.make_system(23, (4, 4, 4), ngauss=4, natoms=2, numax=2, bc=(0, 0, 0)),
     dict(r_trunc=0.1, nsources=(2, 2, 1)), complex(-0.4, 0.1)),
]


@pytest.mark.parametrize("name,mk,kw,E", EDGE)
def test_edge_case_action_matches_oracle(device, name, mk, kw, E):
    import torch
    plan = G.GreenPlan(mk(), **kw)
    plan.set_inputs(E, (0.0, 0.0, 0.0))
    if name == "one_block_per_column":
        assert plan.nnzb == plan.nCols   # every column group is its source block alone
    ora = O.Oracle(plan, E, (0.0, 0.0, 0.0))
    x = np.random.default_rng(31).standard_normal((plan.nnzb, 2, 64, 64))
    y_ref = ora.action(x)
    print(f"case {name}: nnzb {plan.nnzb}, columns {plan.nCols}, atom images {plan.nAtomImages}", flush=True)
    for fp, tol in ((64, 1e-12), (32, 2e-5)):
        dt = torch.float64 if fp == 64 else torch.float32
        xd = torch.from_numpy(x).to(device=device, dtype=dt)
        yd = torch.empty_like(xd)
        plan.multiply(yd, xd, fp_bits=fp)
        torch.cuda.synchronize()
        err = np.abs(yd.double().cpu().numpy() - y_ref).max() / np.abs(y_ref).max()
        assert err < tol, f"{name}: FP{fp} relative error {err:.3e}"


@pytest.mark.parametrize("name,mk,kw,E", EDGE)
def test_edge_case_solves_meet_the_tolerance(device, name, mk, kw, E):
    import torch
    plan = G.GreenPlan(mk(), **kw)
    plan.set_inputs(E, (0.0, 0.0, 0.0))
    ora = O.Oracle(plan, E, (0.0, 0.0, 0.0))
    X = {}
    for fp, method in ((64, "tfqmr"), (32, "cocg")):
        Xd = torch.empty((plan.nnzb, 2, 64, 64), dtype=torch.float64, device=device)
        _, st = plan.solve(fp_bits=fp, tolerance=1e-9, max_iterations=2000, max_refinements=6, X=Xd, method=method,
                           inner_tolerance=1e-30 if method == "cocg" else 0.0, fp32_max_iterations=1200)
        torch.cuda.synchronize()
        X[fp] = Xd.cpu().numpy()
        r = true_residuals(plan, ora, X[fp])
        print(f"{name} fp{fp} {method}: {st['iterations']} applications, residual {r.max():.2e}", flush=True)
        assert st["converged_columns"] == st["total_columns"]
        assert r.max() <= 1e-9
    assert np.abs(X[32] - X[64]).max() <= 1e-6 * np.abs(X[64]).max()
