"""gfa_action_noalloc: the action with its workspace in the library's fixed per-device scratch array.

Called through the C-ABI on two plans in a row, the second with a smaller input, in both precisions: its output
equals gfa_action's bit for bit (the same kernels on the same inputs, deterministic reductions), the free device
memory is the same before and after the calls, a second call after the scratch held another plan's data gives
the same result, and a plan whose workspace exceeds the array's 64 MiB is refused with GFA_ERR_ARGUMENT."""
import ctypes as C

import numpy as np
import pytest

from angstromcube_amd import gfa as G
from angstromcube_amd import synthetic as S

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20


def noalloc(plan, y, x, fp_bits):
    import torch
    fl = C.c_double()
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    return G.lib().gfa_action_noalloc(plan.h, fp_bits, C.c_void_p(y.data_ptr()), C.c_void_p(x.data_ptr()), st,
                                      C.byref(fl))


def plans():
    # the 8^3-block periodic cell (512 blocks, 4 atoms numax 3 with periodic images), then a smaller one
    big = G.GreenPlan(S.small_periodic(), r_trunc=6.0)
    big.set_inputs(complex(-0.5, 0.05))
    small = G.GreenPlan(S.make_system(0xA43C0000, (4, 4, 4), ngauss=4, natoms=2, numax=2), r_trunc=6.0)
    small.set_inputs(complex(-0.4, 0.2), (0.25, 0.0, -0.125))
    return big, small


def test_noalloc_action_equals_action(device):
    import torch
    big, small = plans()
    for fp in (64, 32):
        dt = torch.float64 if fp == 64 else torch.float32
        for plan in (big, small):
            plan.prepare(fp)
        results = []
        for plan, seed in ((big, 1), (small, 2)):
            print(f"fp{fp} plan nnzb {plan.nnzb}", flush=True)
            g = torch.Generator(device="cpu").manual_seed(seed)
            x = torch.randn((plan.nnzb, 2, 64, 64), dtype=torch.float64, generator=g).to(device=device, dtype=dt)
            y_ref = torch.empty_like(x)
            plan.multiply(y_ref, x, fp_bits=fp)
            y = torch.full_like(x, float("nan"))
            torch.cuda.synchronize()
            free0 = torch.cuda.mem_get_info()[0]
            assert noalloc(plan, y, x, fp) == G.GFA_OK, G.lib().gfa_last_error()
            torch.cuda.synchronize()
            free1 = torch.cuda.mem_get_info()[0]
            assert free1 == free0, (free0, free1)
            assert torch.equal(y, y_ref), float((y - y_ref).abs().max())
            results.append((plan, x, y_ref))
        # the scratch now holds the small plan's data: the big plan again gives the same result
        plan, x, y_ref = results[0]
        y = torch.full_like(x, float("nan"))
        assert noalloc(plan, y, x, fp) == G.GFA_OK, G.lib().gfa_last_error()
        torch.cuda.synchronize()
        assert torch.equal(y, y_ref)
    big.close()
    small.close()


def test_noalloc_refuses_a_plan_beyond_the_scratch(device):
    import torch
    # 18 source blocks of a cell with 64 numax-5 atoms: the addition coefficients alone exceed 64 MiB
    sysm = S.make_system(5, (8, 8, 8), ngauss=4, natoms=64, numax=5)
    plan = G.GreenPlan(sysm, r_trunc=6.0, nsources=(3, 3, 2))
    plan.set_inputs(complex(-0.5, 0.1))
    need = plan.sizes.natomcoeffs * plan.nCols * 128 * 8
    print(f"plan nnzb {plan.nnzb}, addition coefficients {need / 2**20:.0f} MiB", flush=True)
    assert need > SCRATCH_BYTES
    plan.prepare(64)
    x = torch.zeros((plan.nnzb, 2, 64, 64), dtype=torch.float64, device=device)
    y = torch.empty_like(x)
    assert noalloc(plan, y, x, 64) == G.GFA_ERR_ARGUMENT
    assert b"allocation-free" in G.lib().gfa_last_error()
    plan.close()


def test_noalloc_needs_the_plan_on_the_device(device):
    import torch
    plan = G.GreenPlan(S.make_system(3, (4, 4, 4), ngauss=2, natoms=1, numax=1), r_trunc=6.0)
    plan.set_inputs(complex(-0.5, 0.1))
    x = torch.zeros((plan.nnzb, 2, 64, 64), dtype=torch.float64, device=device)
    y = torch.empty_like(x)
    assert noalloc(plan, y, x, 64) == G.GFA_ERR_STATE
    plan.close()
