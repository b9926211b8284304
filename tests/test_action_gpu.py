"""HIP action vs the CPU oracle (which itself equals the reference's CPU action, see
test_oracle_vs_reference.py), through the C-ABI, on seeded inputs.

Tolerances: FP64 1e-12 relative (max-norm), FP32 2e-5 relative (max-norm) against the FP64 oracle."""
import numpy as np
import pytest

from angstromcube_amd import gfa as G
from angstromcube_amd import synthetic as S
from oracle import oracle as O
from tests.cases import cases

pytestmark = pytest.mark.gpu


def _rand(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape)


@pytest.mark.parametrize("name,mk,kw,E,k", cases())
def test_action_fp64_matches_oracle(device, name, mk, kw, E, k):
    import torch
    sysm = mk()
    plan = G.GreenPlan(sysm, **kw)
    plan.set_inputs(E, k)
    ora = O.Oracle(plan, E, k)
    x = _rand((plan.nnzb, 2, 64, 64), 11)
    y_ref = ora.action(x)
    xd = torch.from_numpy(x).to(device)
    yd = torch.empty_like(xd)
    print(f"case {name} nnzb {plan.nnzb} images {plan.nAtomImages}", flush=True)
    flops = plan.multiply(yd, xd, fp_bits=64)
    torch.cuda.synchronize()
    y = yd.cpu().numpy()
    err = np.abs(y - y_ref).max() / np.abs(y_ref).max()
    assert flops > 0
    assert err < 1e-12, f"{name}: FP64 relative error {err:.3e}"
    # each term on its own: potential only is y(x) with kinetic/dyadic switched off -> check via linearity
    y2 = torch.empty_like(xd)
    plan.multiply(y2, 2.0 * xd, fp_bits=64)
    torch.cuda.synchronize()
    assert np.abs(y2.cpu().numpy() - 2 * y).max() <= 1e-13 * np.abs(y).max()


@pytest.mark.parametrize("name,mk,kw,E,k", cases(small_only=True))
def test_action_fp32_close_to_oracle(device, name, mk, kw, E, k):
    import torch
    sysm = mk()
    plan = G.GreenPlan(sysm, **kw)
    plan.set_inputs(E, k)
    ora = O.Oracle(plan, E, k)
    x = _rand((plan.nnzb, 2, 64, 64), 12).astype(np.float32)
    y_ref = ora.action(x.astype(np.float64))
    xd = torch.from_numpy(x).to(device)
    yd = torch.empty_like(xd)
    plan.multiply(yd, xd, fp_bits=32)
    torch.cuda.synchronize()
    err = np.abs(yd.cpu().numpy() - y_ref).max() / np.abs(y_ref).max()
    assert err < 2e-5, f"{name}: FP32 relative error {err:.3e}"


def test_action_repeated_calls_are_identical(device):
    import torch
    name, mk, kw, E, k = cases(small_only=True)[0]
    plan = G.GreenPlan(mk(), **kw)
    plan.set_inputs(E, k)
    xd = torch.from_numpy(_rand((plan.nnzb, 2, 64, 64), 13)).to(device)
    y1, y2 = torch.empty_like(xd), torch.empty_like(xd)
    plan.multiply(y1, xd)
    plan.multiply(y2, xd)
    torch.cuda.synchronize()
    assert torch.equal(y1, y2), "the action must be deterministic"


def test_action_rejects_bad_arguments(device):
    import torch
    name, mk, kw, E, k = cases(small_only=True)[0]
    plan = G.GreenPlan(mk(), **kw)
    xd = torch.zeros((plan.nnzb, 2, 64, 64), dtype=torch.float64, device=device)
    yd = torch.empty_like(xd)
    with pytest.raises(G.GfaError, match="energy"):
        plan.multiply(yd, xd)  # no energy set yet
    plan.set_inputs(E, k)
    with pytest.raises(G.GfaError):
        plan.multiply(yd, xd, fp_bits=16)
    with pytest.raises(G.GfaError, match="in place"):
        G._check(G.lib().gfa_action(plan.h, 64, G.C.c_void_p(xd.data_ptr()), G.C.c_void_p(xd.data_ptr()), None, None))


@pytest.mark.parametrize("kind,columns,lines", [(0, 16, 1), (0, 16, 2), (0, 16, 4), (0, 32, 1), (0, 32, 2), (0, 32, 4),
                                               (1, 32, 0), (1, 64, 0), (2, 16, 3), (2, 32, 3), (2, 64, 3),
                                               (2, 16, 2), (2, 32, 2), (2, 64, 2)])
def test_every_tile_shape_matches_oracle(device, kind, columns, lines):
    """Each tile shape the autotuner may pick gives the oracle's action (FP64 1e-12, FP32 2e-5), with periodic
    images (8th order) and with the 16th-order stencil."""
    import torch
    geoms = [c for c in cases() if c[0] in ("periodic8_k", "isolated6_fd16", "mixed_bc", "wrap16_numax5")]
    geoms.append(("wrap8_fd8", lambda: S.make_system(9, (8, 8, 8), natoms=4, numax=3, bc=(1, 1, 1)),
                  dict(r_trunc=3.0, nsources=(2, 1, 1), fd_range=(4, 4, 4)), complex(-0.4, 0.05), (0, 0, 0)))
    for name, mk, kw, E, k in geoms:
        plan = G.GreenPlan(mk(), **kw)
        plan.set_inputs(E, k)
        x = _rand((plan.nnzb, 2, 64, 64), 21)
        y_ref = O.Oracle(plan, E, k).action(x)
        for fp, tol, dt in ((64, 1e-12, torch.float64), (32, 2e-5, torch.float32)):
            if kind == 2 and columns == 64 and fp == 64:
                continue   # no such FP64 shape
            plan.set_tile_shape(fp, kind, columns, lines)
            xd = torch.from_numpy(x).to(device=device, dtype=dt)
            yd = torch.empty_like(xd)
            print(f"tile {kind}:{columns}x{lines} {name} fp{fp}", flush=True)
            plan.multiply(yd, xd, fp_bits=fp)
            torch.cuda.synchronize()
            err = np.abs(yd.double().cpu().numpy() - y_ref).max() / np.abs(y_ref).max()
            assert err < tol, f"{name} fp{fp} tile {columns}x{lines}: {err:.3e}"


@pytest.mark.parametrize("kind,columns,lines", [(1, 32, 0), (2, 32, 2), (0, 16, 2)])
def test_matrix_core_dyadic_matches_oracle(device, kind, columns, lines):
    """The dense SHO projection and add-back on the matrix cores (16x16x4 MFMA in the action's precision, the
    add-back in its own kernel after the local one) give the oracle's action (FP64 1e-12, FP32 2e-5) with periodic
    images at a k-point, the 16th-order stencil, mixed boundary conditions and numax 5 atoms; and a mixed COCG
    solve with them (its (p, A p) then comes from a pass of its own) meets the residual 1e-9."""
    import torch
    geoms = [c for c in cases() if c[0] in ("periodic8_k", "isolated6_fd16", "mixed_bc", "wrap16_numax5")]
    for name, mk, kw, E, k in geoms:
        plan = G.GreenPlan(mk(), **kw)
        plan.set_inputs(E, k)
        x = _rand((plan.nnzb, 2, 64, 64), 24)
        y_ref = O.Oracle(plan, E, k).action(x)
        for fp, tol, dt in ((64, 1e-12, torch.float64), (32, 2e-5, torch.float32)):
            plan.set_tile_shape(fp, kind, columns, lines)
            plan.set_dyadic_engine(fp, 1)
            assert plan.dyadic_engine(fp) == 1
            xd = torch.from_numpy(x).to(device=device, dtype=dt)
            yd = torch.empty_like(xd)
            print(f"matrix-core dyadic term, tile {kind}:{columns}x{lines} {name} fp{fp}", flush=True)
            plan.multiply(yd, xd, fp_bits=fp)
            torch.cuda.synchronize()
            err = np.abs(yd.double().cpu().numpy() - y_ref).max() / np.abs(y_ref).max()
            assert err < tol, f"{name} fp{fp}: {err:.3e}"
        if name == "wrap16_numax5":
            _, st = plan.solve(fp_bits=32, tolerance=1e-9, max_iterations=3000, max_refinements=6,
                               inner_tolerance=1e-30, method="cocg")
            print(f"  mixed COCG with the matrix-core dyadic term: {st['converged_columns']}/{st['total_columns']}, "
                  f"max residual {st['max_residual']:.2e}", flush=True)
            assert st["converged_columns"] == st["total_columns"] and st["max_residual"] <= 1e-9
        plan.close()


def test_autotune_keeps_the_action_exact(device):
    import torch
    name, mk, kw, E, k = cases(small_only=True)[1]
    plan = G.GreenPlan(mk(), **kw)
    plan.set_inputs(E, k)
    x = _rand((plan.nnzb, 2, 64, 64), 22)
    xd = torch.from_numpy(x).to(device)
    yd = torch.empty_like(xd)
    assert plan.tile_shape(64)[3] == 0   # not tuned yet
    best = plan.autotune(yd, xd, fp_bits=64, reps=2)
    assert best > 0
    kind, columns, lines, tuned = plan.tile_shape(64)
    assert tuned == 1 and (kind, columns) in [(0, 16), (0, 32), (1, 32), (1, 64), (2, 16), (2, 32)]
    plan.multiply(yd, xd, fp_bits=64)
    torch.cuda.synchronize()
    y_ref = O.Oracle(plan, E, k).action(x)
    assert np.abs(yd.cpu().numpy() - y_ref).max() <= 1e-12 * np.abs(y_ref).max()
    with pytest.raises(G.GfaError, match="tile"):
        plan.set_tile_shape(64, 0, 24, 2)
    with pytest.raises(G.GfaError, match="tile"):
        plan.set_tile_shape(64, 1, 16)
    plan.set_tile_shape(32, 1, 64)
    assert plan.tile_shape(32) == (1, 64, 0, 1)


def test_action_parts_compose(device):
    """Projection (part 1) followed by the local kernel (part 2) is the whole action, bit for bit."""
    import torch
    name, mk, kw, E, k = cases(small_only=True)[1]
    plan = G.GreenPlan(mk(), **kw)
    plan.set_inputs(E, k)
    xd = torch.from_numpy(_rand((plan.nnzb, 2, 64, 64), 23)).to(device)
    y_all, y_parts = torch.empty_like(xd), torch.empty_like(xd)
    plan.multiply(y_all, xd, fp_bits=64)
    plan.multiply_part(y_parts, xd, 1, fp_bits=64)
    plan.multiply_part(y_parts, xd, 2, fp_bits=64)
    torch.cuda.synchronize()
    assert torch.equal(y_all, y_parts)
    with pytest.raises(G.GfaError, match="part"):
        plan.multiply_part(y_parts, xd, 4, fp_bits=64)
