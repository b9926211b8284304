"""Writes tests/golden/contour/quadrature.npz from the reference's own energy-contour functions
(Gauss_Legendre_quadrature, Gauss_Fermi_Dirac_quadrature, get_energy_mesh of src/energy_contour.cxx), compiled
from the reference sources into oracle/_ref (make -C oracle ref).  Run in a checkout where the reference exists:
    python tests/golden/make_contour_golden.py
The fixture holds, per request, the number of points the reference returns and its nodes and weights."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))

# energy-mesh parameter sets: (kBT, eBot, nBot, nPar, nFer, nPol)
MESHES = [(1e-2, -1.0, 5, 15, 9, 3), (2e-2, -0.8, 3, 7, 4, 1), (1e-2, -1.0, 0, 12, 0, 0), (5e-3, -1.2, 4, 10, 6, -2),
          (1e-2, -1.0, 33, 20, 16, 2)]
GL_N = list(range(1, 81))
FD_N = list(range(1, 17))


def main():
    R = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_green.so"))
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    out = {}
    for n in GL_N:
        x, w = np.zeros(128), np.zeros(128)
        m = R.ref_gauss_legendre(n, p(x), p(w))
        out[f"gl_{n}"] = np.stack([x[:m], w[:m]])
    for n in FD_N:
        x, w = np.zeros(128), np.zeros(128)
        m = R.ref_gauss_fermi_dirac(n, p(x), p(w))
        out[f"fd_{n}"] = np.stack([x[:m], w[:m]])
    R.ref_energy_mesh.argtypes = [C.c_double, C.c_double, C.c_uint, C.c_uint, C.c_uint, C.c_int, C.c_void_p,
                                  C.c_void_p, C.c_int]
    for i, (kBT, eBot, nBot, nPar, nFer, nPol) in enumerate(MESHES):
        E, w = np.zeros((256, 2)), np.zeros((256, 2))
        m = R.ref_energy_mesh(kBT, eBot, nBot, nPar, nFer, nPol, p(E), p(w), 256)
        out[f"mesh_{i}"] = np.concatenate([E[:m], w[:m]], axis=1)
    out["meshes"] = np.array(MESHES)
    np.savez_compressed(os.path.join(HERE, "contour", "quadrature.npz"), **out)
    print("wrote", os.path.join(HERE, "contour", "quadrature.npz"))


if __name__ == "__main__":
    main()
