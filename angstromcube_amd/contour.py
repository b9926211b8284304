"""The energy contour around the Green-function solve, over the C-ABI (include/gfa.h).

Mirrors src/energy_contour.cxx: the quadratures (Gauss_Legendre_quadrature, Gauss_Fermi_Dirac_quadrature), the
energy mesh (get_energy_mesh, :139-253) and the density accumulation of energy_contour::Integrator::integrate
(:273-451) up to the 4x4x4 density, with the energy points of the contour solved as column groups of batched
plans on the device instead of one at a time.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import gfa as G


def gauss_legendre(n):
    """(x, w) of the Gauss-Legendre rule on [-1, 1] with n points, n rounded up as the reference's tables are."""
    x, w = np.zeros(128), np.zeros(128)
    m = G.lib().gfa_gauss_legendre(int(n), G._dptr(x), G._dptr(w))
    if m < 0:
        G._check(-m)
    return x[:m].copy(), w[:m].copy()


def gauss_fermi_dirac(n):
    """(x, w) of the Gauss rule for the weight 1 / (1 + exp(30 x)) on [-1, infinity), n <= 16 points."""
    x, w = np.zeros(16), np.zeros(16)
    m = G.lib().gfa_gauss_fermi_dirac(int(n), G._dptr(x), G._dptr(w))
    if m < 0:
        G._check(-m)
    return x[:m].copy(), w[:m].copy()


def energy_mesh(kBT=1e-2, eBot=-1.0, nBot=5, nPar=15, nFer=9, nPol=3):
    """Contour points relative to the Fermi level and their complex weights. The defaults are the reference's
    (energy_contour.temperature, .band.bottom, .bottom, .parallel, .fermidirac, .matsubara at scale 1,
    src/energy_contour.cxx:245-251)."""
    L = G.lib()
    n = C.c_int32()
    G._check(L.gfa_energy_mesh(kBT, eBot, nBot, nPar, nFer, nPol, None, None, 0, C.byref(n)))
    E, w = np.zeros((n.value, 2)), np.zeros((n.value, 2))
    G._check(L.gfa_energy_mesh(kBT, eBot, nBot, nPar, nFer, nPol, G._dptr(E), G._dptr(w), n.value, C.byref(n)))
    return E[:, 0] + 1j * E[:, 1], w[:, 0] + 1j * w[:, 1]


def contour_density(system, r_trunc, E, wE, kpoints=((0.0, 0.0, 0.0, 1.0),), E_Fermi=0.0, nsources=(1, 1, 1),
                    sources=None, fd_range=(4, 4, 4), r_proj=6.0, max_groups=64, dVol=None, fp_bits=64,
                    tolerance=1e-9, max_iterations=4000, max_refinements=4, method="tfqmr", inner_tolerance=0.0,
                    fp32_max_iterations=0, autotune=False, fp32_progress_check=0):
    """rho[nsources][64] = Re sum_E w_E sum_k w_k G_ii / pi and the reference's response density, for the source
    blocks of the description (a box of nsources at the cell centre, or an explicit list), the system's potential
    and raw atom matrices; returns (rho, rho_response, stats)."""
    L = G.lib()
    d, keep = G.plan_desc(system, r_trunc, fd_range, nsources, sources, r_proj)
    V = np.ascontiguousarray(system.V, dtype=np.float64)
    hs = np.concatenate([np.ascontiguousarray(m, dtype=np.float64).reshape(-1) for m in system.hs]) \
        if system.natoms else np.zeros(1)
    E = np.asarray(E, dtype=complex).reshape(-1)
    wE = np.broadcast_to(np.asarray(wE, dtype=complex), E.shape)
    e2 = np.ascontiguousarray(np.stack([E.real, E.imag], axis=1))
    w2 = np.ascontiguousarray(np.stack([wE.real, wE.imag], axis=1))
    kp = np.ascontiguousarray(np.asarray(kpoints, dtype=np.float64).reshape(-1, 4))
    o = G.ContourOpts()
    o.E, o.wE, o.nE = G._dptr(e2), G._dptr(w2), E.size
    o.k, o.nk = G._dptr(kp), kp.shape[0]
    o.E_Fermi = E_Fermi
    o.dVol = system.dvol if dVol is None else dVol
    o.max_groups = max_groups
    s = o.solve
    s.fp_bits, s.max_iterations, s.tolerance, s.max_refinements = fp_bits, max_iterations, tolerance, max_refinements
    s.check_every, s.inner_tolerance, s.autotune = 8, inner_tolerance, 1 if autotune else 0
    s.fp32_max_iterations = fp32_max_iterations
    s.fp32_progress_check = fp32_progress_check
    s.method = {"tfqmr": 0, "cocg": 1}[method]
    nsrc = len(np.asarray(sources).reshape(-1, 3)) if sources is not None else None
    if nsrc is None:   # the box of source blocks, as the plan builder places it
        p = G.GreenPlan(system, r_trunc, fd_range, nsources, None, r_proj)
        nsrc = p.nCols
        p.close()
    rho, res = np.zeros((nsrc, 64)), np.zeros((nsrc, 64))
    st = G.ContourStats()
    G._check(L.gfa_contour_density(C.byref(d), G._dptr(V), G._dptr(hs), C.byref(o), G._dptr(rho), G._dptr(res),
                                   C.byref(st)))
    del keep
    return rho, res, {k: getattr(st, k) for k, _ in st._fields_}
