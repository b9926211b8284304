"""ctypes wrapper of oracle/_ref/libref_green.so, the reference's CPU action compiled from its
own sources (oracle/Makefile, target `ref`).  Test infrastructure only."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_SO = os.path.join(HERE, "_ref", "libref_green.so")

_lib = None


def available() -> bool:
    return os.path.exists(REF_SO)


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(REF_SO)
        vp, i32p = C.c_void_p, C.POINTER(C.c_int32)
        L.ref_plan_create.restype = vp
        L.ref_plan_create.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_int8), C.POINTER(C.c_double),
                                      C.POINTER(C.c_double), C.c_int, C.c_double, i32p, i32p,
                                      C.c_double, C.c_int]
        L.ref_plan_destroy.argtypes = [vp]
        L.ref_plan_sizes.argtypes = [vp, C.POINTER(C.c_int64)]
        L.ref_plan_bsr.argtypes = [vp] + [C.c_void_p] * 7
        L.ref_plan_kinetic.argtypes = [vp, C.c_int, C.c_void_p, C.c_void_p, C.POINTER(C.c_double)]
        L.ref_plan_dyadic.argtypes = [vp] + [C.c_void_p] * 14
        L.ref_set_potential.argtypes = [vp, C.c_void_p]
        L.ref_set_atom_matrices.argtypes = [vp, C.POINTER(C.c_void_p)]
        L.ref_set_energy.argtypes = [vp, C.c_double, C.c_double, C.c_double]
        L.ref_set_kpoint.argtypes = [vp, C.POINTER(C.c_double)]
        L.ref_action64.restype = C.c_double
        L.ref_action64.argtypes = [vp, C.c_void_p, C.c_void_p]
        L.ref_action32.restype = C.c_double
        L.ref_action32.argtypes = [vp, C.c_void_p, C.c_void_p]
        L.ref_action_part64.argtypes = [vp, C.c_int, C.c_void_p, C.c_void_p]
        L.ref_selftest.argtypes = [C.c_int, C.c_int]
        L.ref_density_matrices.argtypes = [vp, C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class RefPlan:
    """A reference plan built by green_function::construct_Green_function."""

    def __init__(self, system, r_trunc, fd_range=(4, 4, 4), nsrc=(1, 1, 1), r_proj=6.0, echo=0):
        L = lib()
        ng = (C.c_uint32 * 3)(*system.ng)
        bc = (C.c_int8 * 3)(*system.bc)
        hg = (C.c_double * 3)(*(getattr(system, "hg", None) or (system.h,) * 3))
        self._xyz = np.ascontiguousarray(system.xyzZinso, dtype=np.float64)
        fd = (C.c_int32 * 3)(*fd_range)
        ns = (C.c_int32 * 3)(*nsrc)
        self.h = L.ref_plan_create(ng, bc, hg, self._xyz.ctypes.data_as(C.POINTER(C.c_double)),
                                   system.natoms, float(r_trunc), fd, ns, float(r_proj), echo)
        if not self.h:
            raise RuntimeError("reference plan construction failed")
        s = (C.c_int64 * 24)()
        L.ref_plan_sizes(self.h, s)
        self.sizes = list(s)
        self.nRows, self.nCols, self.nnzb = s[0], s[1], s[2]
        self.nAtoms, self.nAtomImages = s[3], s[4]
        self.system = system

    def __del__(self):
        if getattr(self, "h", None):
            lib().ref_plan_destroy(self.h)
            self.h = None

    def bsr(self):
        d = dict(RowStart=np.zeros(self.nRows + 1, np.uint32), colindx=np.zeros(self.nnzb, np.uint16),
                 subset=np.zeros(self.nCols, np.uint32), veff_index=np.zeros(self.nnzb, np.int32),
                 rowCubePos=np.zeros((self.nRows, 4), np.float32), colCubePos=np.zeros((self.nCols, 4), np.float32),
                 global_source_indices=np.zeros(self.nCols, np.int64))
        lib().ref_plan_bsr(self.h, *[_p(d[k]) for k in ("RowStart", "colindx", "subset", "veff_index",
                                                        "rowCubePos", "colCubePos", "global_source_indices")])
        return d

    def kinetic(self, dd):
        n, nz = self.sizes[7 + dd], self.sizes[10 + dd]
        rs, li, pf = np.zeros(n + 1, np.uint32), np.zeros(nz, np.int32), C.c_double()
        lib().ref_plan_kinetic(self.h, dd, _p(rs), _p(li), C.byref(pf))
        return dict(rowStart=rs, lists=li, prefactor=pf.value, fd_range=self.sizes[13 + dd])

    def dyadic(self):
        nai, na, nnzb, nrhs = self.nAtomImages, self.nAtoms, self.nnzb, self.nCols
        d = dict(AtomImagePos=np.zeros((nai, 4)), AtomImageLmax=np.zeros(nai, np.int8),
                 AtomImageStarts=np.zeros(nai + 1, np.uint32), AtomImageIndex=np.zeros(nai, np.uint32),
                 AtomImageShift=np.zeros((nai, 4), np.int8), AtomStarts=np.zeros(na + 1, np.uint32),
                 AtomLmax=np.zeros(na, np.int8), global_atom_ids=np.zeros(na, np.int64),
                 add_rowStart=np.zeros(nnzb + 1, np.uint32), add_colIndex=np.zeros(self.sizes[16], np.uint32),
                 prj_rowStart=np.zeros((nrhs, nai + 1), np.uint32), prj_colIndex=np.zeros(self.sizes[17], np.uint32),
                 sum_rowStart=np.zeros(na + 1, np.uint32), sum_colIndex=np.zeros(self.sizes[18], np.uint32))
        keys = ("AtomImagePos", "AtomImageLmax", "AtomImageStarts", "AtomImageIndex", "AtomImageShift",
                "AtomStarts", "AtomLmax", "global_atom_ids", "add_rowStart", "add_colIndex",
                "prj_rowStart", "prj_colIndex", "sum_rowStart", "sum_colIndex")
        lib().ref_plan_dyadic(self.h, *[_p(d[k]) for k in keys])
        return d

    def set_inputs(self, E, k=(0.0, 0.0, 0.0)):
        sysm = self.system
        L = lib()
        V = np.ascontiguousarray(sysm.V, dtype=np.float64)
        L.ref_set_potential(self.h, _p(V))
        self._hs = [np.ascontiguousarray(m, dtype=np.float64) for m in sysm.hs]
        arr = (C.c_void_p * len(self._hs))(*[m.ctypes.data for m in self._hs])
        L.ref_set_atom_matrices(self.h, arr)
        L.ref_set_energy(self.h, float(E.real), float(E.imag), sysm.dvol)
        kk = (C.c_double * 3)(*k)
        L.ref_set_kpoint(self.h, kk)

    def action(self, x, fp=64):
        """y = (H - E S) x for x of shape (nnzb, 2, 64, 64)."""
        dt = np.float64 if fp == 64 else np.float32
        x = np.ascontiguousarray(x, dtype=dt)
        assert x.shape == (self.nnzb, 2, 64, 64)
        y = np.zeros_like(x)
        fn = lib().ref_action64 if fp == 64 else lib().ref_action32
        fn(self.h, _p(y), _p(x))
        return y

    def density_matrices(self, G):
        """The reference's atomic density matrices of G (nnzb, 2, 64, 64), concatenated per atom."""
        G = np.ascontiguousarray(G, dtype=np.float64)
        assert G.shape == (self.nnzb, 2, 64, 64)
        out = np.zeros(64 * 64 * 128)
        n = lib().ref_density_matrices(self.h, _p(G), _p(out))
        return out[:n].copy()

    def action_part(self, which, x, y=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x) if y is None else np.ascontiguousarray(y, dtype=np.float64).copy()
        lib().ref_action_part64(self.h, which, _p(y), _p(x))
        return y
