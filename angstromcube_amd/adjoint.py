"""Adjoint gradients of a contour functional of the Green function, behind torch.autograd.Function.

    L(V, {h_a, s_a}) = Im sum_c w_c Tr[G_c],   (H - E_c S) G_c = B_c,

with one column group c per contour point (or instance) and the trace over its diagonal (source) block.  Since
dG = -A^-1 dA G, dL = -Im sum_c w_c <lambda_c, dA G_c> with lambda_c = A_c^-H B_c.  H and S are Hermitian, so
A_c^H = H - conj(E_c) S: the backward pass is one FP64 solve of the same plan at the conjugate energies and the
library's fused conj(lambda) G reduction (gfa_gradient); no iterates are stored.

The reference has no derivatives; the forward value follows its density export G_ii / pi
(include/green_action.hxx:226-238) times pi.
"""
from __future__ import annotations

import numpy as np
import torch

from . import gfa as G


def _split_hs(system, hs_flat):
    out, o = [], 0
    for ia in range(system.natoms):
        n = G.nsho_of(system.xyzZinso[ia, 5])
        out.append(np.ascontiguousarray(hs_flat[o:o + 2 * n * n]).reshape(2, n, n))
        o += 2 * n * n
    return out


def flatten_hs(hs):
    return np.concatenate([np.asarray(m, dtype=np.float64).reshape(-1) for m in hs])


def solve_fp64(plan, E, tol, maxit, method="tfqmr", fixed=0):
    """X of (H - E_c S) X = B for every column group (FP64 device tensor) and the solve statistics.

    fixed (an iteration count per column group, or one for all): run every group for exactly that many operator
    applications (no stopping test, no refinement) and require the true residual to be at most tol at the end.
    Solves of nearby operators then stop at the same iterates, so their difference carries no stopping-point
    noise (used for finite differences)."""
    plan.update_energies(E)
    X = torch.empty((plan.nnzb, 2, 64, 64), dtype=torch.float64, device="cuda")
    if np.any(np.asarray(fixed) > 0):
        caps = np.broadcast_to(np.asarray(fixed, dtype=np.int32), (plan.nCols,))
        rho, st = plan.solve(fp_bits=64, tolerance=1e-300, max_iterations=int(caps.max()) + 8, max_refinements=0,
                             X=X, method=method, group_max_iterations=caps)
        ok = st["max_residual"] <= tol
    else:
        rho, st = plan.solve(fp_bits=64, tolerance=tol, max_iterations=maxit, max_refinements=4, X=X, method=method)
        ok = st["converged_columns"] == st["total_columns"]
    if not ok:
        raise G.GfaError(f"solve did not converge: {st['converged_columns']}/{st['total_columns']} columns, "
                         f"max residual {st['max_residual']:.2e}")
    return X, rho


def trace_value(rho, w):
    """Im sum_c w_c Tr[G_c] from the exported densities rho = G_ii / pi."""
    return float(np.imag(np.sum(np.asarray(w) * np.pi * rho.sum(axis=1))))


class ContourTrace(torch.autograd.Function):
    """L = Im sum_c w_c Tr[G_c] as a function of the grid potential V (ng_z, ng_y, ng_x) and the flattened atom
    matrices (h, s per atom, as given to gfa_set_atom_matrices).  Arguments after those two are constants:
    the plan (GreenPlan, one column group per energy), energies E[nCols], weights w[nCols], tolerance, iterations
    and the Krylov method ("cocg" needs real Bloch phases: Gamma or a wrapped cell; the adjoint system then is
    complex symmetric too)."""

    @staticmethod
    def forward(ctx, V, hs_flat, plan, E, w, tol=1e-11, maxit=4000, method="tfqmr", fixed=0):
        sysm = plan.system
        plan.update_potential(V.detach().cpu().numpy(), _split_hs(sysm, hs_flat.detach().cpu().numpy()))
        X, rho = solve_fp64(plan, E, tol, maxit, method, fixed)
        ctx.plan, ctx.E, ctx.w, ctx.tol, ctx.maxit = plan, np.asarray(E, dtype=complex), np.asarray(w, dtype=complex), tol, maxit
        ctx.method, ctx.fixed = method, fixed
        ctx.X = X
        return torch.tensor(trace_value(rho, w), dtype=torch.float64)

    @staticmethod
    def backward(ctx, gL):
        plan, E, w = ctx.plan, ctx.E, ctx.w
        Lam, _ = solve_fp64(plan, np.conj(E), ctx.tol, ctx.maxit, ctx.method, ctx.fixed)   # (H - conj(E) S) Lambda = B
        plan.update_energies(E)
        gV, ghs = plan.gradient(ctx.X, Lam, E, w)
        g = float(gL)
        return (torch.from_numpy(gV * g), torch.from_numpy(flatten_hs(ghs) * g), None, None, None, None, None, None,
                None)


def contour_trace(V, hs_flat, plan, E, w, tol=1e-11, maxit=4000, method="tfqmr", fixed=0):
    return ContourTrace.apply(V, hs_flat, plan, E, w, tol, maxit, method, fixed)
