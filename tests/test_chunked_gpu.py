"""Chunked solves of many source blocks (gfa_solve_chunked): the source columns split over plans that fit a
memory budget, as the reference distributes them over ranks (src/green_function.cxx:255-354).  Column groups are
independent, so two chunks give the single plan's densities: bit for bit with COCG (no random vector), and to the
solve tolerance with tfQMR (its shadow residual is seeded per block of the plan, which differs between the plans)."""
import numpy as np
import pytest

from angstromcube_amd import gfa as G
from angstromcube_amd import synthetic as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", ["cocg", "tfqmr"])
def test_two_chunks_equal_one_plan(device, method):
    sysm = S.make_system(21, (12, 12, 12), ngauss=8, natoms=6, numax=3)   # periodic, wrapped by r = 4 Bohr
    E = complex(-0.4, 0.08)
    kw = dict(fp_bits=64, tolerance=1e-9, max_iterations=3000, max_refinements=2, method=method)
    plan = G.GreenPlan(sysm, r_trunc=4.0, nsources=(2, 2, 2))
    plan.set_inputs(E)
    sb, pb = plan.solve_memory(64, method)
    rho1, s1 = plan.solve(**kw)
    plan.close()
    rho2, s2 = G.solve_chunked(sysm, 4.0, E, max_bytes=(sb + pb) // 2, nsources=(2, 2, 2), **kw)
    print(f"{method}: one plan of 8 sources ({(sb + pb) / 1e6:.0f} MB), chunks {s2['chunks']} x "
          f"{s2['sources_per_chunk']} sources, largest chunk {s2['largest_chunk_bytes'] / 1e6:.0f} MB")
    assert s1["converged_columns"] == s2["converged_columns"] == s2["total_columns"] == 512
    assert s2["chunks"] == 2 and s2["sources_per_chunk"] == 4
    assert s2["largest_chunk_bytes"] <= (sb + pb) // 2 * 1.25   # the budget holds up to the estimate's accuracy
    d = np.abs(rho1 - rho2).max() / np.abs(rho1).max()
    print(f"  max |rho_chunked - rho_single| / max |rho| = {d:.2e}")
    if method == "cocg":
        assert d == 0.0
    else:
        assert d <= 1e-7
