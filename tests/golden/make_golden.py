"""Writes the golden fixtures of tests/golden/ from the reference's own CPU action, compiled from the
reference sources into oracle/_ref (make -C oracle ref).  Run in a checkout where the reference exists:
    python tests/golden/make_golden.py
Each fixture holds the inputs' recipe (system seed, sizes, E, k, the seed of the input vector x) and
reference outputs: the plan's index arrays or their sizes, the action y = (H - E S) x at 4096 sampled
entries, and whole-array checksums of y."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from angstromcube_amd import synthetic as S  # noqa: E402
from oracle import refgreen as R  # noqa: E402

FIXTURES = {
    # name: (system kwargs, plan kwargs, E, k, store full plan arrays)
    "periodic8_gamma": (dict(seed=0xA43C0000, nb=(8, 8, 8), ngauss=8, natoms=4, numax=3),
                        dict(r_trunc=6.0), complex(-0.5, 0.05), (0.0, 0.0, 0.0), True),
    "periodic8_k": (dict(seed=0xA43C0000, nb=(8, 8, 8), ngauss=8, natoms=4, numax=3),
                    dict(r_trunc=6.0), complex(-0.5, 0.05), (0.25, -1 / 3.0, 0.1), False),
    "isolated6_fd16": (dict(seed=7, nb=(6, 6, 6), natoms=5, numax=2, bc=(0, 0, 0)),
                       dict(r_trunc=3.0, nsrc=(2, 1, 2), fd_range=(8, 8, 8)), complex(-0.3, 0.02), (0, 0, 0), True),
    "mixed_bc": (dict(seed=8, nb=(6, 5, 7), natoms=5, numax=3, bc=(1, 2, 3)),
                 dict(r_trunc=3.0, nsrc=(2, 2, 1)), complex(-0.3, 0.1), (0.1, 0, 0), True),
    "wrap32_r10_numax5": (dict(seed=0xA43C0001, nb=(32, 32, 32), ngauss=64, natoms=64, numax=5),
                          dict(r_trunc=10.0), complex(-0.3, 0.02), (0, 0, 0), False),
}

X_SEED = 0x601D


def system(kw):
    kw = dict(kw)
    return S.make_system(kw.pop("seed"), kw.pop("nb"), **kw)


def main():
    assert R.available(), "build the reference harness first: make -C oracle ref"
    for name, (skw, pkw, E, k, full) in FIXTURES.items():
        sysm = system(skw)
        rp = R.RefPlan(sysm, **pkw)
        rp.set_inputs(E, k)
        n = rp.nnzb * 2 * 64 * 64
        x = S.hash_uniform(X_SEED, n).reshape(rp.nnzb, 2, 64, 64)
        y = rp.action(x)
        rng = np.random.default_rng(1)
        idx = np.sort(rng.choice(n, size=4096, replace=False)).astype(np.int64)
        w = S.hash_uniform(X_SEED + 1, n)
        out = dict(system=np.array(repr(skw)), plan=np.array(repr(pkw)), E=np.array([E.real, E.imag]),
                   k=np.array(k, dtype=float), x_seed=np.array(X_SEED), sizes=np.array(rp.sizes, dtype=np.int64),
                   sample_index=idx, sample_y=y.reshape(-1)[idx], y_sum=np.array(y.sum()),
                   y_sq=np.array((y * y).sum()), y_weighted=np.array(np.dot(w, y.reshape(-1))))
        if full:
            b = rp.bsr()
            out.update({"bsr_" + key: v for key, v in b.items()})
            for dd, D in enumerate("xyz"):
                kp = rp.kinetic(dd)
                out["kin_start_" + D] = kp["rowStart"]
                out["kin_list_" + D] = kp["lists"]
            dy = rp.dyadic()
            out.update({"dyadic_" + key: v for key, v in dy.items()})
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
        print(name, rp.sizes[:7], os.path.getsize(os.path.join(HERE, name + ".npz")))


if __name__ == "__main__":
    main()
