"""Grid Hamiltonian files (angstromcube_amd/hamiltonian.py): the reference's Hmt.xml / Hmt.json input.

Pins: (1) the reference's own test input test/Hmt_vacuum_8x8x8.xml (a copy under tests/golden/hamiltonian)
loads as the reference's loader reads it (grid, spacing, boundary, 512 zero potential values, no atoms);
(2) writing a system with atoms in the reference's XML and JSON formats and loading it back returns the values to
the precision the format prints (positions %.12f, matrices %.15e, potential %.15f); (3) plans built from a
loaded file are array-for-array the reference's plans for the same inputs, also with anisotropic spacings, and
the oracle's action on them (pinned to the reference's action elsewhere) equals the reference's own action;
(4) on the GPU, the action of a plan from a loaded file equals the oracle's."""
import os

import numpy as np
import pytest

from angstromcube_amd import gfa as G
from angstromcube_amd import hamiltonian as H
from angstromcube_amd import synthetic as S
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
VACUUM = os.path.join(HERE, "golden", "hamiltonian", "Hmt_vacuum_8x8x8.xml")
REF_SO = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libref_green.so")


def test_reference_vacuum_input_loads():
    h = H.load_hamiltonian(VACUUM)
    assert h.ng == (8, 8, 8) and h.hg == (1.0, 1.0, 1.0) and h.bc == (0, 0, 0)
    assert h.natoms == 0 and h.V.shape == (8, 8, 8) and not h.V.any()
    p = G.GreenPlan(h, r_trunc=2.0)
    assert p.nCols == 1 and p.nnzb > 0


@pytest.mark.parametrize("fmt", ["xml", "json"])
def test_round_trip(tmp_path, fmt):
    sysm = S.make_system(0xA43C0002, (4, 3, 5), ngauss=6, natoms=3, numax=2, bc=(1, 0, 2))
    ham = H.from_system(sysm)
    f = tmp_path / f"Hmt.{fmt}"
    H.write_hamiltonian(ham, f, fmt)
    back = H.load_hamiltonian(f)
    assert back.ng == ham.ng and back.bc == ham.bc and np.allclose(back.hg, ham.hg, rtol=0, atol=1e-16)
    assert np.abs(back.xyzZinso - ham.xyzZinso).max() <= 1e-12
    assert np.abs(back.V - ham.V).max() <= 1e-15
    for a, b in zip(back.hs, ham.hs):
        assert a.shape == b.shape and np.abs(a - b).max() <= 1e-15 * max(1.0, np.abs(b).max())


def test_xml_written_like_the_reference(tmp_path):
    """The writer's output for the vacuum Hamiltonian parses to the same data as the reference's file."""
    f = tmp_path / "Hmt.xml"
    H.write_hamiltonian(H.load_hamiltonian(VACUUM), f)
    a, b = H.load_hamiltonian(f), H.load_hamiltonian(VACUUM)
    assert (a.ng, a.hg, a.bc, a.natoms) == (b.ng, b.hg, b.bc, b.natoms) and np.array_equal(a.V, b.V)
    text = f.read_text()
    for tag in ("<grid_Hamiltonian", "<sho_atoms number=\"0\">", "<spacing x=", "<boundary x=", "<potential nx=\"8\""):
        assert tag in text


def test_short_potential_is_zero_filled(tmp_path):
    f = tmp_path / "Hmt.xml"
    with open(VACUUM) as g:
        f.write_text(g.read().replace("0 0 0 0 0 0 0 0  0 0 0 0 0 0 0 0", "0", 1))
    with pytest.warns(UserWarning, match="set to zero"):
        h = H.load_hamiltonian(f)
    assert h.V.shape == (8, 8, 8) and not h.V.any()


@pytest.mark.skipif(not os.path.exists(REF_SO), reason="reference build oracle/_ref absent")
@pytest.mark.parametrize("hg", [(0.25, 0.25, 0.25), (0.25, 0.3, 0.2)])
def test_plan_and_action_from_a_file_equal_the_reference(tmp_path, hg):
    from oracle import refgreen as R
    sysm = S.make_system(0xA43C0002, (5, 4, 6), ngauss=6, natoms=3, numax=2, bc=(1, 0, 0))
    ham = H.from_system(sysm)
    ham.hg = hg
    f = tmp_path / "Hmt.xml"
    H.write_hamiltonian(ham, f)
    h = H.load_hamiltonian(f)
    mine = G.GreenPlan(h, r_trunc=2.5, nsources=(2, 1, 1))
    ref = R.RefPlan(h, r_trunc=2.5, nsrc=(2, 1, 1))
    assert (mine.nnzb, mine.nRows, mine.nCols, mine.nAtoms, mine.nAtomImages) == \
        (ref.nnzb, ref.nRows, ref.nCols, ref.nAtoms, ref.nAtomImages)
    b = ref.bsr()
    assert np.array_equal(b["colindx"], mine.export("COLINDX"))
    for dd, D in enumerate("XYZ"):
        kr = ref.kinetic(dd)
        assert np.array_equal(kr["lists"], mine.export("KIN_LIST_" + D))
        assert kr["prefactor"] == mine.export("KIN_PREFACTOR")[dd] == -0.5 / hg[dd] ** 2
    E = complex(-0.4, 0.1)
    mine.set_inputs(E)
    ref.set_inputs(E)
    x = np.random.default_rng(4).standard_normal((mine.nnzb, 2, 64, 64))
    y_ora = O.Oracle(mine, E).action(x)
    y_ref = ref.action(x)
    assert np.abs(y_ora - y_ref).max() <= 1e-13 * np.abs(y_ref).max()


@pytest.mark.gpu
def test_gpu_action_of_a_loaded_hamiltonian(tmp_path, device):
    import torch
    sysm = S.make_system(0xA43C0002, (5, 4, 6), ngauss=6, natoms=3, numax=3, bc=(1, 0, 0))
    ham = H.from_system(sysm)
    ham.hg = (0.25, 0.3, 0.2)
    f = tmp_path / "Hmt.json"
    H.write_hamiltonian(ham, f, "json")
    h = H.load_hamiltonian(f)
    plan = G.GreenPlan(h, r_trunc=2.5, nsources=(2, 1, 1))
    E = complex(-0.4, 0.1)
    plan.set_inputs(E)
    x = np.random.default_rng(5).standard_normal((plan.nnzb, 2, 64, 64))
    xd = torch.from_numpy(x).to(device)
    yd = torch.empty_like(xd)
    plan.multiply(yd, xd)
    torch.cuda.synchronize()
    y_ref = O.Oracle(plan, E).action(x)
    assert np.abs(yd.cpu().numpy() - y_ref).max() <= 1e-12 * np.abs(y_ref).max()
