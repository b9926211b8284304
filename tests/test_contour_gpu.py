"""The contour integration gfa_contour_density on the GPU.

Pins: (1) the batched driver (energies as column groups, several energy batches per k-point) equals the
reference's loop structure (src/energy_contour.cxx:364-393: per energy, per k-point, one solve of all sources,
rho_E = sum_k w_k rho_Ek, rho = Re sum_E w_E rho_E, response = Re(-sum rho_E / -sum E) over all but the last two
points), run here as separate single-energy FP64 plans, to 1e-8 relative; (2) one instance of that loop equals the
oracle's TFQMR (the CPU restatement) to 1e-8; (3) E_Fermi only shifts the energies; (4) the mixed-precision
driver agrees with FP64 within the conditioning (1e-6)."""
import numpy as np
import pytest

from angstromcube_amd import contour as K
from angstromcube_amd import gfa as G
from angstromcube_amd import synthetic as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SOURCES = [(1, 1, 1), (2, 1, 1)]
KPTS = [(0.0, 0.0, 0.0, 0.5), (0.25, 0.0, 0.0, 0.5)]


def system():
    return S.make_system(0xA43C0000, (4, 4, 4), ngauss=4, natoms=2, numax=2)


def loop_density(sysm, E, wE, kpts, E_F, tol):
    """The reference's loop: one plan and solve per (energy, k-point), all sources as columns."""
    nsrc = len(SOURCES)
    rho_c = np.zeros((nsrc, 64), complex)
    res_c = np.zeros((nsrc, 64), complex)
    res_point = 0
    per_instance = {}
    for e, (En, w) in enumerate(zip(E, wE)):
        rho_E = np.zeros((nsrc, 64), complex)
        for k in kpts:
            p = G.GreenPlan(sysm, r_trunc=6.0, sources=SOURCES)
            p.set_inputs(En + E_F, k[:3])
            rho, st = p.solve(fp_bits=64, tolerance=tol, max_iterations=4000, max_refinements=4)
            assert st["converged_columns"] == st["total_columns"]
            per_instance[(e, k[:3])] = rho
            rho_E += k[3] * rho
            p.close()
        rho_c += w * rho_E
        if e < len(E) - 2:
            res_c -= rho_E
            res_point -= En + E_F
    res_point = 1 / res_point if res_point != 0 else 1
    return rho_c.real, (res_c * res_point).real, per_instance


def test_contour_density_equals_the_reference_loop(device):
    sysm = system()
    E, wE = K.energy_mesh(kBT=0.02, eBot=-1.0, nBot=2, nPar=3, nFer=2, nPol=1)
    assert len(E) == 8
    E_F = -0.1
    print(f"contour: {len(E)} energy points x {len(KPTS)} k-points x {len(SOURCES)} sources", flush=True)
    rho, res, st = K.contour_density(sysm, 6.0, E, wE, KPTS, E_Fermi=E_F, sources=SOURCES, max_groups=4,
                                     tolerance=1e-11, max_refinements=4)
    assert st["solves"] == 4 * len(KPTS) and st["sources"] == len(SOURCES)
    assert st["converged_columns"] == st["total_columns"] == len(E) * len(KPTS) * 64 * len(SOURCES)
    rho_l, res_l, inst = loop_density(sysm, E, wE, KPTS, E_F, 1e-11)
    e1 = np.abs(rho - rho_l).max() / np.abs(rho_l).max()
    e2 = np.abs(res - res_l).max() / np.abs(res_l).max()
    print(f"batched vs loop: density {e1:.2e}, response {e2:.2e}")
    assert e1 < 1e-8 and e2 < 1e-8
    # one instance of the loop against the oracle's TFQMR on the CPU
    e, k = 3, KPTS[1]
    p = G.GreenPlan(sysm, r_trunc=6.0, sources=SOURCES)
    ora = O.Oracle(p, E[e] + E_F, k[:3])
    Xo, _, _ = ora.tfqmr(tol=1e-12, maxit=4000, seed=5)
    sub = p.export("SUBSET")
    rho_o = np.stack([np.diag(Xo[b, 0]) + 1j * np.diag(Xo[b, 1]) for b in sub]) / np.pi
    e3 = np.abs(inst[(e, k[:3])] - rho_o).max() / np.abs(rho_o).max()
    print(f"instance vs oracle TFQMR: {e3:.2e}")
    assert e3 < 1e-8


def test_fermi_level_shifts_the_energies(device):
    sysm = system()
    E, wE = K.energy_mesh(kBT=0.02, eBot=-0.8, nBot=1, nPar=2, nFer=1, nPol=1)
    a = K.contour_density(sysm, 6.0, E, wE, KPTS[:1], E_Fermi=0.15, sources=SOURCES, tolerance=1e-11)
    b = K.contour_density(sysm, 6.0, E + 0.15, wE, KPTS[:1], E_Fermi=0.0, sources=SOURCES, tolerance=1e-11)
    for x, y in zip(a[:2], b[:2]):
        assert np.abs(x - y).max() <= 1e-9 * np.abs(y).max()


def test_mixed_precision_contour(device):
    sysm = system()
    E, wE = K.energy_mesh(kBT=0.02, eBot=-1.0, nBot=2, nPar=3, nFer=2, nPol=1)
    r64, q64, _ = K.contour_density(sysm, 6.0, E, wE, KPTS, sources=SOURCES, tolerance=1e-10)
    r32, q32, st = K.contour_density(sysm, 6.0, E, wE, KPTS, sources=SOURCES, tolerance=1e-10, fp_bits=32,
                                     max_refinements=6, inner_tolerance=1e-4)
    assert st["converged_columns"] == st["total_columns"]
    assert np.abs(r32 - r64).max() <= 1e-6 * np.abs(r64).max()
    assert np.abs(q32 - q64).max() <= 1e-6 * np.abs(q64).max()
