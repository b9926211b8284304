"""ctypes binding of libgfa.so (include/gfa.h) for tests and the benchmark.

The product path is the HIP library; this module only moves pointers.  Device buffers are torch
tensors on the current HIP device.  Every call raises GfaError with the library's message on a
non-zero status, and loading fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# GFA_LIBRARY selects another build of the library (e.g. a tuning variant)
LIB_PATH = os.environ.get("GFA_LIBRARY") or os.path.join(HERE, "libgfa.so")

# export ids, in the order of the enum in include/gfa.h
EXPORTS = ["ROW_START", "COLINDX", "SUBSET", "VEFF_INDEX", "ROW_POS", "COL_POS", "SOURCE_IDS",
           "KIN_START_X", "KIN_START_Y", "KIN_START_Z", "KIN_LIST_X", "KIN_LIST_Y", "KIN_LIST_Z",
           "IMAGE_POS", "IMAGE_LMAX", "IMAGE_STARTS", "IMAGE_ATOM", "IMAGE_SHIFT", "ATOM_STARTS",
           "ATOM_LMAX", "ATOM_IDS", "ADD_START", "ADD_INDEX", "PRJ_START", "PRJ_INDEX", "SUM_START",
           "SUM_INDEX", "ROWINDX", "V_CELL", "ATOM_MATRICES", "IMAGE_PHASE", "PHASE", "KIN_PREFACTOR"]
EXPORT_ID = {n: i for i, n in enumerate(EXPORTS)}

C_SYMBOLS = ["gfa_plan_create", "gfa_plan_destroy", "gfa_plan_query", "gfa_plan_export", "gfa_set_potential",
             "gfa_set_atom_matrices", "gfa_set_energy", "gfa_set_energies", "gfa_set_kpoint", "gfa_action", "gfa_action_multiply", "gfa_action_benchmark", "gfa_action_noalloc", "gfa_autotune", "gfa_set_tile_shape", "gfa_get_tile_shape", "gfa_set_dyadic_engine", "gfa_get_dyadic_engine", "gfa_action_part", "gfa_gradient",
             "gfa_solve", "gfa_solver_release", "gfa_solve_memory", "gfa_plan_prepare", "gfa_atom_density_matrices", "gfa_gauss_legendre", "gfa_gauss_fermi_dirac", "gfa_energy_mesh", "gfa_contour_density", "gfa_solve_chunked",
             "gfa_last_error"]


GFA_OK, GFA_ERR_ARGUMENT, GFA_ERR_DEVICE, GFA_ERR_STATE, GFA_ERR_MEMORY = 0, 1, 2, 3, 4


class GfaError(RuntimeError):
    pass


class PlanDesc(C.Structure):
    _fields_ = [("ng", C.c_uint32 * 3), ("bc", C.c_int8 * 3), ("reserved0", C.c_int8), ("hg", C.c_double * 3),
                ("xyzZinso", C.POINTER(C.c_double)), ("natoms", C.c_int32), ("echo", C.c_int32),
                ("r_trunc", C.c_double), ("r_proj", C.c_double), ("fd_range", C.c_int32 * 3),
                ("nsources", C.c_int32 * 3), ("sources", C.POINTER(C.c_int32)), ("nsources_explicit", C.c_int32),
                ("reserved1", C.c_int32), ("scale_T", C.c_double), ("source_radii", C.POINTER(C.c_double))]


class Sizes(C.Structure):
    _fields_ = [("nnzb", C.c_uint64), ("nRows", C.c_uint32), ("nCols", C.c_uint32), ("nAtoms", C.c_uint32),
                ("nAtomImages", C.c_uint32), ("natomcoeffs", C.c_uint64), ("ncoeffs", C.c_uint64),
                ("nlists", C.c_uint32 * 3), ("fd_range", C.c_int32 * 3), ("nlist_entries", C.c_uint64 * 3),
                ("sho_add_nnz", C.c_uint64), ("sho_prj_nnz", C.c_uint64), ("sho_sum_nnz", C.c_uint64),
                ("bc", C.c_int8 * 3), ("reserved", C.c_int8), ("nchunks", C.c_uint32), ("device_bytes", C.c_uint64),
                ("solver_bytes", C.c_uint64)]


class SolveOpts(C.Structure):
    _fields_ = [("fp_bits", C.c_int32), ("max_iterations", C.c_int32), ("tolerance", C.c_double),
                ("max_refinements", C.c_int32), ("check_every", C.c_int32), ("shadow_seed", C.c_uint64),
                ("X_device", C.c_void_p), ("stream", C.c_void_p), ("inner_tolerance", C.c_double),
                ("echo", C.c_int32), ("eager", C.c_int32), ("autotune", C.c_int32),
                ("fp32_max_iterations", C.c_int32), ("residuals", C.POINTER(C.c_double)), ("halfsteps", C.POINTER(C.c_int32)),
                ("method", C.c_int32), ("fp32_progress_check", C.c_int32), ("group_max_iterations", C.POINTER(C.c_int32))]


class ContourOpts(C.Structure):
    _fields_ = [("E", C.POINTER(C.c_double)), ("wE", C.POINTER(C.c_double)), ("nE", C.c_int32), ("nk", C.c_int32),
                ("k", C.POINTER(C.c_double)), ("E_Fermi", C.c_double), ("dVol", C.c_double),
                ("max_groups", C.c_int32), ("reserved", C.c_int32), ("solve", SolveOpts)]


class ContourStats(C.Structure):
    _fields_ = [("solves", C.c_int32), ("sources", C.c_int32), ("converged_columns", C.c_int32),
                ("total_columns", C.c_int32), ("max_residual", C.c_double), ("seconds", C.c_double)]


class ChunkedStats(C.Structure):
    _fields_ = [("sources", C.c_int32), ("sources_per_chunk", C.c_int32), ("chunks", C.c_int32),
                ("converged_columns", C.c_int32), ("total_columns", C.c_int32), ("iterations", C.c_int32),
                ("max_residual", C.c_double), ("seconds", C.c_double), ("largest_chunk_bytes", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("refinements", C.c_int32), ("converged_columns", C.c_int32),
                ("total_columns", C.c_int32), ("max_residual", C.c_double), ("seconds", C.c_double),
                ("flops", C.c_double), ("actions", C.c_int32), ("host_syncs", C.c_int32), ("graph_captures", C.c_int32),
                ("fp64_groups", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """Load libgfa.so; raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GfaError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "or make -C angstromcube_amd/csrc")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.gfa_plan_create.argtypes = [C.POINTER(PlanDesc), C.POINTER(vp)]
        L.gfa_plan_destroy.argtypes = [vp]
        L.gfa_plan_query.argtypes = [vp, C.POINTER(Sizes)]
        L.gfa_plan_export.argtypes = [vp, C.c_int, vp, C.c_uint64]
        L.gfa_set_potential.argtypes = [vp, C.POINTER(C.c_double)]
        L.gfa_set_atom_matrices.argtypes = [vp, C.c_int32, C.POINTER(C.c_double), C.c_int32]
        L.gfa_set_energy.argtypes = [vp, C.c_double, C.c_double, C.c_double]
        L.gfa_set_energies.argtypes = [vp, C.POINTER(C.c_double), C.c_double]
        L.gfa_set_kpoint.argtypes = [vp, C.POINTER(C.c_double)]
        L.gfa_action.argtypes = [vp, C.c_int, vp, vp, vp, C.POINTER(C.c_double)]
        L.gfa_action_multiply.argtypes = [vp, C.c_int, vp, vp, vp, C.c_uint32, C.c_uint32, C.c_uint32, vp, C.c_int,
                                          C.POINTER(C.c_double)]
        L.gfa_action_benchmark.argtypes = [vp, C.c_int, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.gfa_action_noalloc.argtypes = [vp, C.c_int, vp, vp, vp, C.POINTER(C.c_double)]
        L.gfa_autotune.argtypes = [vp, C.c_int, vp, vp, vp, C.c_int, C.POINTER(C.c_double)]
        L.gfa_set_tile_shape.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
        ip = C.POINTER(C.c_int)
        L.gfa_set_dyadic_engine.argtypes = [vp, C.c_int, C.c_int]
        L.gfa_get_dyadic_engine.argtypes = [vp, C.c_int, ip]
        L.gfa_get_tile_shape.argtypes = [vp, C.c_int, ip, ip, ip, ip]
        L.gfa_action_part.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp]
        L.gfa_gradient.argtypes = [vp, vp, vp, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                   C.POINTER(C.c_double), vp]
        L.gfa_solve.argtypes = [vp, C.POINTER(SolveOpts), C.POINTER(C.c_double), C.POINTER(SolveStats)]
        L.gfa_solver_release.argtypes = [vp]
        L.gfa_solve_memory.argtypes = [vp, C.POINTER(SolveOpts), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.gfa_plan_prepare.argtypes = [vp, C.c_int]
        dp = C.POINTER(C.c_double)
        L.gfa_atom_density_matrices.argtypes = [vp, vp, C.c_double, dp, vp]
        L.gfa_gauss_legendre.argtypes = [C.c_int32, dp, dp]
        L.gfa_gauss_fermi_dirac.argtypes = [C.c_int32, dp, dp]
        L.gfa_energy_mesh.argtypes = [C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_int32, C.c_int32, dp, dp,
                                      C.c_int32, C.POINTER(C.c_int32)]
        L.gfa_contour_density.argtypes = [vp, dp, dp, C.POINTER(ContourOpts), dp, dp, C.POINTER(ContourStats)]
        L.gfa_solve_chunked.argtypes = [C.POINTER(PlanDesc), dp, dp, C.c_double, C.c_double, dp, C.POINTER(SolveOpts),
                                        C.c_uint64, dp, C.POINTER(ChunkedStats)]
        L.gfa_last_error.restype = C.c_char_p
        for name in C_SYMBOLS:
            getattr(L, name).restype = C.c_int if name != "gfa_last_error" else C.c_char_p
        _lib = L
    return _lib


def _check(status):
    if status != 0:
        raise GfaError(f"gfa status {status}: {lib().gfa_last_error().decode()}")


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def plan_desc(system, r_trunc, fd_range=(4, 4, 4), nsources=(1, 1, 1), sources=None, r_proj=6.0, echo=0,
              source_radii=None):
    """The gfa_plan_desc of a system; returns (desc, arrays the desc points into)."""
    d = PlanDesc()
    keep = []
    hg = getattr(system, "hg", None) or (system.h,) * 3   # per-direction spacings where the system has them
    for k in range(3):
        d.ng[k] = system.ng[k]
        d.bc[k] = system.bc[k]
        d.hg[k] = hg[k]
        d.fd_range[k] = fd_range[k]
        d.nsources[k] = nsources[k]
    xyz = np.ascontiguousarray(system.xyzZinso, dtype=np.float64)
    keep.append(xyz)
    d.xyzZinso = _dptr(xyz)
    d.natoms = system.natoms
    d.echo = echo
    d.r_trunc = r_trunc
    d.r_proj = r_proj
    d.scale_T = 1.0
    src = None
    if sources is not None:
        src = np.ascontiguousarray(np.asarray(sources, dtype=np.int32).reshape(-1, 3))
        keep.append(src)
        d.sources = src.ctypes.data_as(C.POINTER(C.c_int32))
        d.nsources_explicit = src.shape[0]
    if source_radii is not None:
        radii = np.ascontiguousarray(source_radii, dtype=np.float64)
        if src is None or radii.shape != (src.shape[0],):
            raise GfaError("one truncation radius per explicit source block is needed")
        keep.append(radii)
        d.source_radii = _dptr(radii)
    return d, keep


def nsho_of(numax):
    n = int(numax)
    return (n + 1) * (n + 2) * (n + 3) // 6


class GreenPlan:
    """A Green-function plan: BSR structure, kinetic lists, atom images and the per-(E,k) data.

    Mirrors the reference's call surface: construct_Green_function (constructor), update_potential,
    update_energy_parameter, update_phases, multiply (action_t::multiply) and solve (green_solver_t::solve).
    """

    def __init__(self, system, r_trunc, fd_range=(4, 4, 4), nsources=(1, 1, 1), sources=None, r_proj=6.0,
                 echo=0, source_radii=None):
        L = lib()
        d, self._keep = plan_desc(system, r_trunc, fd_range, nsources, sources, r_proj, echo, source_radii)
        h = C.c_void_p()
        _check(L.gfa_plan_create(C.byref(d), C.byref(h)))
        self.h = h
        self.system = system
        self.refresh_sizes()

    def refresh_sizes(self):
        s = Sizes()
        _check(lib().gfa_plan_query(self.h, C.byref(s)))
        self.sizes = s
        self.nnzb, self.nRows, self.nCols = s.nnzb, s.nRows, s.nCols
        self.nAtoms, self.nAtomImages = s.nAtoms, s.nAtomImages
        return s

    def close(self):
        if getattr(self, "h", None):
            lib().gfa_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown: module globals may already be gone
            pass

    def export(self, name):
        s = self.sizes
        nai, na, nnzb = s.nAtomImages, s.nAtoms, s.nnzb
        ncell = int(np.prod([n // 4 for n in self.system.ng]))
        spec = {
            "ROW_START": (np.uint32, (s.nRows + 1,)), "COLINDX": (np.uint16, (nnzb,)),
            "SUBSET": (np.uint32, (s.nCols,)), "VEFF_INDEX": (np.int32, (nnzb,)),
            "ROW_POS": (np.float32, (s.nRows, 4)), "COL_POS": (np.float32, (s.nCols, 4)),
            "SOURCE_IDS": (np.int64, (s.nCols,)),
            "IMAGE_POS": (np.float64, (nai, 4)), "IMAGE_LMAX": (np.int8, (nai,)),
            "IMAGE_STARTS": (np.uint32, (nai + 1,)), "IMAGE_ATOM": (np.uint32, (nai,)),
            "IMAGE_SHIFT": (np.int8, (nai, 4)), "ATOM_STARTS": (np.uint32, (na + 1,)),
            "ATOM_LMAX": (np.int8, (na,)), "ATOM_IDS": (np.int64, (na,)),
            "ADD_START": (np.uint32, (nnzb + 1,)), "ADD_INDEX": (np.uint32, (s.sho_add_nnz,)),
            "PRJ_START": (np.uint32, (s.nCols, nai + 1)), "PRJ_INDEX": (np.uint32, (s.sho_prj_nnz,)),
            "SUM_START": (np.uint32, (na + 1,)), "SUM_INDEX": (np.uint32, (s.sho_sum_nnz,)),
            "ROWINDX": (np.uint32, (nnzb,)), "V_CELL": (np.float64, (ncell, 64)),
            "IMAGE_PHASE": (np.float64, (nai, 2)), "PHASE": (np.float64, (3, 2, 2)),
            "KIN_PREFACTOR": (np.float64, (3,)),
        }
        for k, d in enumerate("XYZ"):
            spec["KIN_START_" + d] = (np.uint32, (s.nlists[k] + 1,))
            spec["KIN_LIST_" + d] = (np.int32, (s.nlist_entries[k],))
        if name == "ATOM_MATRICES":
            starts = self.export("ATOM_STARTS")
            lm = self.export("ATOM_LMAX")
            n = int(sum(2 * ((l + 1) * (l + 2) * (l + 3) // 6) ** 2 for l in lm.astype(int)))
            spec[name] = (np.float64, (n,))
            del starts
        dt, shape = spec[name]
        a = np.zeros(shape, dtype=dt)
        _check(lib().gfa_plan_export(self.h, EXPORT_ID[name], a.ctypes.data_as(C.c_void_p), a.nbytes))
        return a

    # ---- per-system and per-(E,k) updates ----
    def update_potential(self, V=None, hs=None):
        """Potential on the full grid and raw (normalised-basis) atom matrices [2][nc][nc] per atom."""
        sysm = self.system
        V = sysm.V if V is None else V
        hs = sysm.hs if hs is None else hs
        Vc = np.ascontiguousarray(V, dtype=np.float64)
        _check(lib().gfa_set_potential(self.h, _dptr(Vc)))
        for ia, m in enumerate(hs):
            m = np.ascontiguousarray(m, dtype=np.float64)
            _check(lib().gfa_set_atom_matrices(self.h, ia, _dptr(m), m.shape[-1]))

    def update_energy_parameter(self, E, dVol=None):
        dVol = self.system.dvol if dVol is None else dVol
        _check(lib().gfa_set_energy(self.h, float(E.real), float(E.imag), float(dVol)))

    def update_energies(self, E, dVol=None):
        """One energy per column group (batched contour instances)."""
        dVol = self.system.dvol if dVol is None else dVol
        e = np.zeros((self.nCols, 2))
        E = np.broadcast_to(np.asarray(E, dtype=complex), (self.nCols,))
        e[:, 0], e[:, 1] = E.real, E.imag
        _check(lib().gfa_set_energies(self.h, _dptr(e), float(dVol)))

    def update_phases(self, k):
        kk = (C.c_double * 3)(*k)
        _check(lib().gfa_set_kpoint(self.h, kk))

    def set_inputs(self, E, k=(0.0, 0.0, 0.0)):
        """E: one energy, or one per column group."""
        self.update_potential()
        if np.ndim(E) == 0:
            self.update_energy_parameter(E)
        else:
            self.update_energies(E)
        self.update_phases(k)

    # ---- device work ----
    def multiply(self, y, x, fp_bits=64, stream=None):
        """y = (H - E S) x for torch device tensors of shape (nnzb, 2, 64, 64); returns the flop count."""
        import torch
        dt = torch.float64 if fp_bits == 64 else torch.float32
        for t in (x, y):
            if t.dtype != dt or not t.is_cuda or not t.is_contiguous() or tuple(t.shape) != (self.nnzb, 2, 64, 64):
                raise GfaError("x and y must be contiguous device tensors of shape (nnzb, 2, 64, 64)")
        fl = C.c_double()
        st = C.c_void_p(stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream)
        _check(lib().gfa_action(self.h, fp_bits, C.c_void_p(y.data_ptr()), C.c_void_p(x.data_ptr()), st, C.byref(fl)))
        return fl.value

    def autotune(self, y, x, fp_bits=64, reps=3, stream=None):
        """Time the action's tile shapes on device tensors x, y (y is overwritten); returns the best ms."""
        import torch
        best = C.c_double()
        st = C.c_void_p(stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream)
        _check(lib().gfa_autotune(self.h, fp_bits, C.c_void_p(y.data_ptr()), C.c_void_p(x.data_ptr()), st, reps,
                                  C.byref(best)))
        return best.value

    def multiply_part(self, y, x, part, fp_bits=64, stream=None):
        """part 1: projection + atom matrices, 2: local kernel (uses the last part-1 coefficients), 3: both."""
        import torch
        st = C.c_void_p(stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream)
        _check(lib().gfa_action_part(self.h, fp_bits, part, C.c_void_p(y.data_ptr()), C.c_void_p(x.data_ptr()), st))

    def atom_density_matrices(self, X, factor=1.0, stream=None):
        """Per contributing atom (ATOM_IDS order) the matrix factor * Im(phase <p_i|X|p_j>) of an FP64 device
        tensor X in the plan layout, the right projection over every source block of the plan."""
        import torch
        lm = self.export("ATOM_LMAX").astype(int)
        ncs = [nsho_of(l) for l in lm]
        D = np.zeros(sum(n * n for n in ncs))
        st = C.c_void_p(stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream)
        _check(lib().gfa_atom_density_matrices(self.h, C.c_void_p(X.data_ptr()), float(factor), _dptr(D), st))
        out, o = [], 0
        for n in ncs:
            out.append(D[o:o + n * n].reshape(n, n))
            o += n * n
        return out

    def gradient(self, X, Lam, E, w, stream=None):
        """dL/dV (grid array) and dL/d(h, s) per atom (list of (2, nc, nc)) of L = Im sum_c w_c Tr[G_c], from the
        FP64 solutions X of the forward and Lam of the adjoint system (device tensors in the plan layout)."""
        import torch
        sysm = self.system
        E = np.broadcast_to(np.asarray(E, dtype=complex), (self.nCols,))
        w = np.broadcast_to(np.asarray(w, dtype=complex), (self.nCols,))
        e = np.ascontiguousarray(np.stack([E.real, E.imag], axis=1))
        ww = np.ascontiguousarray(np.stack([w.real, w.imag], axis=1))
        gV = np.zeros(tuple(reversed(sysm.ng)))
        ncs = [nsho_of(sysm.xyzZinso[ia, 5]) for ia in range(sysm.natoms)]
        ghs = np.zeros(sum(2 * n * n for n in ncs))
        st = C.c_void_p(stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream)
        for t in (X, Lam):
            if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous() or tuple(t.shape) != (self.nnzb, 2, 64, 64):
                raise GfaError("X and Lambda must be contiguous FP64 device tensors of shape (nnzb, 2, 64, 64)")
        _check(lib().gfa_gradient(self.h, C.c_void_p(X.data_ptr()), C.c_void_p(Lam.data_ptr()), _dptr(e), _dptr(ww),
                                  _dptr(gV), _dptr(ghs), st))
        out, o = [], 0
        for n in ncs:
            out.append(ghs[o:o + 2 * n * n].reshape(2, n, n))
            o += 2 * n * n
        return gV, out

    def set_tile_shape(self, fp_bits, kind, columns, lines=0):
        _check(lib().gfa_set_tile_shape(self.h, fp_bits, kind, columns, lines))

    def tile_shape(self, fp_bits):
        """(kind, columns, lines, tuned) of the local kernel that the action of fp_bits launches."""
        v = [C.c_int() for _ in range(4)]
        _check(lib().gfa_get_tile_shape(self.h, fp_bits, *[C.byref(t) for t in v]))
        return tuple(t.value for t in v)

    def set_dyadic_engine(self, fp_bits, engine):
        """0: the SHO projection and add-back in their sum-factorised forms on the vector ALUs; 1: in their dense forms
        on the matrix cores (the add-back in a kernel after the local one)."""
        _check(lib().gfa_set_dyadic_engine(self.h, fp_bits, engine))

    def dyadic_engine(self, fp_bits):
        e = C.c_int()
        _check(lib().gfa_get_dyadic_engine(self.h, fp_bits, C.byref(e)))
        return e.value

    def prepare(self, fp_bits):
        """Upload the plan and allocate the action's workspace of fp_bits (later actions allocate nothing)."""
        _check(lib().gfa_plan_prepare(self.h, fp_bits))
        self.refresh_sizes()

    def release_solver(self):
        """Free the solver state (solution, work vectors, captured graph); the plan stays usable."""
        _check(lib().gfa_solver_release(self.h))
        self.refresh_sizes()

    def solve_memory(self, fp_bits=64, method="tfqmr"):
        """(solver bytes at the solve's largest, plan device bytes) of a solve with these options."""
        o = SolveOpts()
        o.fp_bits = fp_bits
        o.method = {"tfqmr": 0, "cocg": 1}[method]
        sb, pb = C.c_uint64(), C.c_uint64()
        _check(lib().gfa_solve_memory(self.h, C.byref(o), C.byref(sb), C.byref(pb)))
        return sb.value, pb.value

    def solve(self, fp_bits=64, tolerance=1e-9, max_iterations=400, max_refinements=4, check_every=8,
              shadow_seed=0, X=None, stream=None, inner_tolerance=0.0, echo=0, eager=False, autotune=False, fp32_max_iterations=0, method="tfqmr",
              group_max_iterations=None, fp32_progress_check=0):
        """Solve (H - E S) X = B; returns (rho[nCols, 64] complex, stats dict)."""
        import torch
        o = SolveOpts()
        o.fp_bits = fp_bits
        o.max_iterations = max_iterations
        o.tolerance = tolerance
        o.max_refinements = max_refinements
        o.check_every = check_every
        o.shadow_seed = shadow_seed
        o.inner_tolerance = inner_tolerance
        o.echo = echo
        o.eager = 1 if eager else 0
        o.autotune = 1 if autotune else 0
        o.fp32_max_iterations = fp32_max_iterations
        o.fp32_progress_check = fp32_progress_check
        o.method = {"tfqmr": 0, "cocg": 1}[method]
        o.X_device = C.c_void_p(X.data_ptr()) if X is not None else None
        caps = None
        if group_max_iterations is not None:
            caps = np.ascontiguousarray(np.broadcast_to(np.asarray(group_max_iterations, dtype=np.int32), (self.nCols,)))
            o.group_max_iterations = caps.ctypes.data_as(C.POINTER(C.c_int32))
        o.stream = C.c_void_p(stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream)
        rho = np.zeros((self.nCols, 64, 2))
        res = np.zeros(64 * self.nCols)
        hst = np.zeros(64 * self.nCols, dtype=np.int32)
        o.residuals = _dptr(res)
        o.halfsteps = hst.ctypes.data_as(C.POINTER(C.c_int32))
        st = SolveStats()
        _check(lib().gfa_solve(self.h, C.byref(o), _dptr(rho), C.byref(st)))
        self.refresh_sizes()
        d = st.as_dict()
        d["column_residuals"] = res
        d["column_halfsteps"] = hst
        return rho[..., 0] + 1j * rho[..., 1], d


def solve_chunked(system, r_trunc, E, k=(0.0, 0.0, 0.0), max_bytes=0, fd_range=(4, 4, 4), nsources=(1, 1, 1),
                  sources=None, r_proj=6.0, fp_bits=64, tolerance=1e-9, max_iterations=2000, max_refinements=4,
                  method="tfqmr", inner_tolerance=0.0, fp32_max_iterations=0):
    """The Green function of every source block of a plan description at one energy and k-point, in chunks of
    source blocks that fit max_bytes of device memory (gfa_solve_chunked; 0: one chunk).  Returns
    (rho[nsrc, 64] complex, stats dict)."""
    d, keep = plan_desc(system, r_trunc, fd_range, nsources, sources, r_proj)
    o = SolveOpts()
    o.fp_bits, o.max_iterations, o.tolerance, o.max_refinements = fp_bits, max_iterations, tolerance, max_refinements
    o.method = {"tfqmr": 0, "cocg": 1}[method]
    o.inner_tolerance, o.fp32_max_iterations = inner_tolerance, fp32_max_iterations
    V = np.ascontiguousarray(system.V, dtype=np.float64)
    hs = np.concatenate([np.asarray(m, dtype=np.float64).reshape(-1) for m in system.hs]) if system.natoms else \
        np.zeros(1)
    kk = np.ascontiguousarray(k, dtype=np.float64)
    base = GreenPlan(system, r_trunc, fd_range, nsources, sources, r_proj)
    nsrc = base.nCols
    base.close()
    rho = np.zeros((nsrc, 64, 2))
    st = ChunkedStats()
    _check(lib().gfa_solve_chunked(C.byref(d), _dptr(V), _dptr(hs), float(E.real), float(E.imag), _dptr(kk),
                                   C.byref(o), int(max_bytes), _dptr(rho), C.byref(st)))
    del keep
    return rho[..., 0] + 1j * rho[..., 1], st.as_dict()

