"""ctypes wrapper of oracle/_build/liboracle.so, the CPU restatement of the action and solve.

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's CPU baseline use it as
the checker; the product never does.  The oracle takes the plan's index structures from a
GreenPlan (whose equality with the reference's plan builder is checked separately) and computes
all value-dependent data (potential blocks, atom matrices, phases) itself.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liboracle.so")


class OraclePlan(C.Structure):
    _fields_ = [("nb", C.c_uint32 * 3), ("hg", C.c_double * 3), ("r_proj", C.c_double),
                ("nnzb", C.c_uint64), ("nRows", C.c_uint32), ("nCols", C.c_uint32),
                ("colindx", C.c_void_p), ("rowindx", C.c_void_p), ("veff_index", C.c_void_p), ("row_pos", C.c_void_p),
                ("nlists", C.c_uint32 * 3), ("kin_start", C.c_void_p * 3), ("kin_list", C.c_void_p * 3),
                ("fd_range", C.c_int32 * 3), ("scale_T", C.c_double),
                ("nAtoms", C.c_uint32), ("nAtomImages", C.c_uint32),
                ("image_pos", C.c_void_p), ("image_lmax", C.c_void_p), ("image_starts", C.c_void_p),
                ("image_shift", C.c_void_p), ("add_start", C.c_void_p), ("add_index", C.c_void_p),
                ("prj_start", C.c_void_p), ("prj_index", C.c_void_p), ("sum_start", C.c_void_p),
                ("sum_index", C.c_void_p), ("atom_starts", C.c_void_p), ("atom_lmax", C.c_void_p),
                ("atom_orig", C.c_void_p), ("V_grid", C.c_void_p), ("hs", C.POINTER(C.c_void_p)),
                ("xyzZinso", C.c_void_p), ("E_re", C.c_double), ("E_im", C.c_double), ("dVol", C.c_double),
                ("k", C.c_double * 3)]


_lib = None


def available():
    return os.path.exists(ORACLE_SO)


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(ORACLE_SO)
        L.oracle_create.restype = C.c_void_p
        L.oracle_create.argtypes = [C.POINTER(OraclePlan)]
        L.oracle_destroy.argtypes = [C.c_void_p]
        L.oracle_action.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        L.oracle_shadow.restype = C.c_double
        L.oracle_shadow.argtypes = [C.c_uint64, C.c_uint64]
        L.oracle_set_sum_order.argtypes = [C.c_int]
        L.oracle_set_sum_order.restype = None
        L.oracle_tfqmr.restype = C.c_int
        L.oracle_tfqmr.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_uint64, C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def _p(a):
    return C.c_void_p(a.ctypes.data)


class Oracle:
    """CPU action/solve for the plan structure of `plan` (a GreenPlan) with energy E and k-point k."""

    def __init__(self, plan, E, k=(0.0, 0.0, 0.0), atom_orig=None):
        sysm = plan.system
        self.plan = plan
        keep = {}
        for name in ("COLINDX", "ROWINDX", "VEFF_INDEX", "ROW_POS", "IMAGE_POS", "IMAGE_LMAX", "IMAGE_STARTS",
                     "IMAGE_SHIFT", "ADD_START", "ADD_INDEX", "PRJ_START", "PRJ_INDEX", "SUM_START", "SUM_INDEX",
                     "ATOM_STARTS", "ATOM_LMAX", "ATOM_IDS", "SUBSET",
                     "KIN_START_X", "KIN_START_Y", "KIN_START_Z", "KIN_LIST_X", "KIN_LIST_Y", "KIN_LIST_Z"):
            keep[name] = np.ascontiguousarray(plan.export(name))
        # original atom of each contributing atom copy: the atom whose id is recorded (ids are row indices)
        ids = keep["ATOM_IDS"]
        keep["ATOM_ORIG"] = np.ascontiguousarray(ids.astype(np.int32) if atom_orig is None else atom_orig, dtype=np.int32)
        keep["V"] = np.ascontiguousarray(sysm.V, dtype=np.float64)
        keep["XYZ"] = np.ascontiguousarray(sysm.xyzZinso, dtype=np.float64)
        self._hs = [np.ascontiguousarray(m, dtype=np.float64) for m in sysm.hs]
        self._hsp = (C.c_void_p * max(1, len(self._hs)))(*[m.ctypes.data for m in self._hs])
        s = plan.sizes
        op = OraclePlan()
        for d in range(3):
            op.nb[d] = sysm.ng[d] // 4
            op.hg[d] = (getattr(sysm, "hg", None) or (sysm.h,) * 3)[d]
            op.nlists[d] = s.nlists[d]
            op.kin_start[d] = keep["KIN_START_" + "XYZ"[d]].ctypes.data
            op.kin_list[d] = keep["KIN_LIST_" + "XYZ"[d]].ctypes.data
            op.fd_range[d] = s.fd_range[d]
            op.k[d] = k[d]
        op.r_proj = 6.0 if not hasattr(plan, "r_proj") else plan.r_proj
        op.nnzb, op.nRows, op.nCols = s.nnzb, s.nRows, s.nCols
        op.colindx = keep["COLINDX"].ctypes.data
        op.rowindx = keep["ROWINDX"].ctypes.data
        op.veff_index = keep["VEFF_INDEX"].ctypes.data
        op.row_pos = keep["ROW_POS"].ctypes.data
        op.scale_T = 1.0
        op.nAtoms, op.nAtomImages = s.nAtoms, s.nAtomImages
        for f, n in (("image_pos", "IMAGE_POS"), ("image_lmax", "IMAGE_LMAX"), ("image_starts", "IMAGE_STARTS"),
                     ("image_shift", "IMAGE_SHIFT"), ("add_start", "ADD_START"), ("add_index", "ADD_INDEX"),
                     ("prj_start", "PRJ_START"), ("prj_index", "PRJ_INDEX"), ("sum_start", "SUM_START"),
                     ("sum_index", "SUM_INDEX"), ("atom_starts", "ATOM_STARTS"), ("atom_lmax", "ATOM_LMAX"),
                     ("atom_orig", "ATOM_ORIG")):
            setattr(op, f, keep[n].ctypes.data)
        op.V_grid = keep["V"].ctypes.data
        op.hs = self._hsp
        op.xyzZinso = keep["XYZ"].ctypes.data
        op.E_re, op.E_im, op.dVol = float(E.real), float(E.imag), sysm.dvol
        self._keep = keep
        self._op = op
        self.h = lib().oracle_create(C.byref(op))
        if not self.h:
            raise RuntimeError("oracle_create failed")
        self.nnzb = int(s.nnzb)
        self.nCols = int(s.nCols)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_destroy(self.h)
            self.h = None

    def action(self, x, which=-1, y=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert x.shape == (self.nnzb, 2, 64, 64)
        y = np.zeros_like(x) if y is None else np.ascontiguousarray(y, dtype=np.float64).copy()
        lib().oracle_action(self.h, which, _p(y), _p(x))
        return y

    def tfqmr(self, tol=1e-9, maxit=400, seed=0x5EEDA43C, reverse_sums=False):
        """TFQMR from X = 0 for the unit columns; reverse_sums sums the inner products over blocks in descending
        order (the same algorithm with other rounding)."""
        X = np.zeros((self.nnzb, 2, 64, 64))
        res = np.zeros(64 * self.nCols)
        sub = self._keep["SUBSET"]
        lib().oracle_set_sum_order(1 if reverse_sums else 0)
        try:
            m = lib().oracle_tfqmr(self.h, _p(sub), tol, maxit, seed, _p(X), _p(res))
        finally:
            lib().oracle_set_sum_order(0)
        return X, res, m
