"""Golden fixtures written by tests/golden/make_golden.py from the reference's own CPU action: the product's
plan and the oracle (CPU) and the HIP action (GPU) are checked against reference outputs without needing
the reference build at run time."""
import ast
import glob
import os

import numpy as np
import pytest

from angstromcube_amd import gfa as G
from angstromcube_amd import synthetic as S
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))
NAMES = [os.path.basename(f)[:-4] for f in FIXTURES]


def load(name):
    f = np.load(os.path.join(HERE, "golden", name + ".npz"), allow_pickle=False)
    skw = ast.literal_eval(str(f["system"]))
    pkw = ast.literal_eval(str(f["plan"]))
    seed, nb = skw.pop("seed"), skw.pop("nb")
    sysm = S.make_system(seed, nb, **skw)
    nsrc = pkw.pop("nsrc", (1, 1, 1))
    plan = G.GreenPlan(sysm, nsources=nsrc, **pkw)
    E = complex(*f["E"])
    k = tuple(f["k"])
    return f, sysm, plan, E, k


def x_of(f, plan):
    n = plan.nnzb * 2 * 64 * 64
    return S.hash_uniform(int(f["x_seed"]), n).reshape(plan.nnzb, 2, 64, 64)


def check_y(f, y, rtol):
    scale = np.abs(f["sample_y"]).max()
    got = y.reshape(-1)[f["sample_index"]]
    assert np.abs(got - f["sample_y"]).max() <= rtol * scale
    assert abs(y.sum() - float(f["y_sum"])) <= rtol * np.abs(y).sum()
    assert abs((y * y).sum() - float(f["y_sq"])) <= rtol * float(f["y_sq"]) * 4
    w = S.hash_uniform(int(f["x_seed"]) + 1, y.size)
    assert abs(np.dot(w, y.reshape(-1)) - float(f["y_weighted"])) <= rtol * np.abs(y).sum()


def test_fixtures_present():
    assert len(FIXTURES) >= 5


@pytest.mark.parametrize("name", NAMES)
def test_plan_matches_golden(name):
    f, sysm, plan, E, k = load(name)
    sizes = f["sizes"]
    assert (plan.nRows, plan.nCols, plan.nnzb, plan.nAtoms, plan.nAtomImages) == tuple(sizes[:5])
    if "bsr_colindx" in f.files:
        for key, exp in (("bsr_RowStart", "ROW_START"), ("bsr_colindx", "COLINDX"), ("bsr_veff_index", "VEFF_INDEX"),
                         ("bsr_subset", "SUBSET"), ("bsr_rowCubePos", "ROW_POS"), ("kin_list_x", "KIN_LIST_X"),
                         ("kin_list_y", "KIN_LIST_Y"), ("kin_list_z", "KIN_LIST_Z"),
                         ("dyadic_add_colIndex", "ADD_INDEX"), ("dyadic_prj_colIndex", "PRJ_INDEX"),
                         ("dyadic_AtomImagePos", "IMAGE_POS")):
            assert np.array_equal(f[key], plan.export(exp)), key


@pytest.mark.parametrize("name", [n for n in NAMES if not n.startswith("wrap32")])
def test_oracle_matches_golden(name):
    f, sysm, plan, E, k = load(name)
    y = O.Oracle(plan, E, k).action(x_of(f, plan))
    check_y(f, y, 1e-13)


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_hip_action_matches_golden(device, name):
    import torch
    f, sysm, plan, E, k = load(name)
    plan.set_inputs(E, k)
    x = torch.from_numpy(x_of(f, plan)).to(device)
    y = torch.empty_like(x)
    print(f"golden {name}: nnzb {plan.nnzb}", flush=True)
    plan.multiply(y, x, fp_bits=64)
    torch.cuda.synchronize()
    check_y(f, y.cpu().numpy(), 1e-12)
