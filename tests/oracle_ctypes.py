"""ctypes bindings for the CPU oracle (oracle/liboracle.so) and for the
reference shim (oracle/_ref/libref.so).  Test infrastructure only."""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref.so")

_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


_oracle = None
_ref = None


def oracle_lib():
    global _oracle
    if _oracle is None:
        lib = C.CDLL(ORACLE_SO)
        _sig(lib, "oc_last_error", C.c_char_p, [])
        _sig(lib, "oc_create", C.c_void_p, [C.c_char_p, C.c_double])
        _sig(lib, "oc_set_precision", None, [C.c_void_p, C.c_int])
        _sig(lib, "oc_table_digest", None, [C.c_void_p, _u64p])
        _sig(lib, "oc_destroy", None, [C.c_void_p])
        _sig(lib, "oc_info", None, [C.c_void_p, _i32p])
        _sig(lib, "oc_core", C.c_double, [C.c_void_p])
        _sig(lib, "oc_hf", None, [C.c_void_p, _u64p])
        _sig(lib, "oc_diagonal", C.c_double, [C.c_void_p, _u64p])
        _sig(lib, "oc_off_diagonal", C.c_double, [C.c_void_p, _u64p, _u64p])
        _sig(lib, "oc_column", C.c_long, [C.c_void_p, _u64p, _u64p, _f64p, C.c_long])
        _sig(lib, "oc_lowest_eig", None, [C.c_int, _f64p, C.POINTER(C.c_double), _f64p, _f64p])
        _sig(lib, "oc_eig_update", C.c_int, [C.c_int, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p])
        _sig(lib, "oc_solver_init", C.c_int,
             [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_void_p])
        _sig(lib, "oc_solver_step", C.c_int, [C.c_void_p, C.c_long])
        _sig(lib, "oc_solver_set_update", C.c_int, [C.c_void_p, C.c_int])
        _sig(lib, "oc_ls_update", None, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, _f64p])
        _sig(lib, "oc_solver_report", None, [C.c_void_p, _f64p])
        _sig(lib, "oc_solver_timers", None, [C.c_void_p, _f64p])
        _sig(lib, "oc_table_dump", C.c_long, [C.c_void_p, _u64p, _f64p, _f64p, C.c_void_p, C.c_long])
        _sig(lib, "oc_candidates", C.c_long, [C.c_void_p, _u64p, _f64p, _f64p, C.c_long])
        _sig(lib, "oc_to_fixed", None, [C.c_double, _i64p])
        _sig(lib, "oc_fixed_to_double", C.c_double, [C.c_int64, C.c_int64])
        _sig(lib, "oc_dd_ops", None, [C.c_double] * 4 + [_f64p])
        _sig(lib, "oc_wf_reset", None, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int])
        _sig(lib, "oc_wf_set_scalars", None, [C.c_void_p, _f64p, C.c_double])
        _sig(lib, "oc_wf_state", None, [C.c_void_p, _f64p])
        _sig(lib, "oc_wf_update", None, [C.c_void_p, C.c_int, _u64p, _f64p])
        _sig(lib, "oc_energy_gradient", C.c_int,
             [C.c_void_p, _u64p, _f64p, C.c_long, _f64p, _f64p, C.POINTER(C.c_double)])
        _sig(lib, "oc_pick", C.c_int, [C.c_void_p, C.c_long, _u64p, _f64p, _f64p, _f64p, C.c_int, _u64p])
        _oracle = lib
    return _oracle


def ref_available():
    return os.path.exists(REF_SO)


def ref_lib():
    global _ref
    if _ref is None:
        lib = C.CDLL(REF_SO)
        _sig(lib, "ref_last_error", C.c_char_p, [])
        _sig(lib, "ref_ham_create", C.c_void_p, [C.c_char_p, C.c_double])
        _sig(lib, "ref_ham_destroy", None, [C.c_void_p])
        _sig(lib, "ref_ham_info", None, [C.c_void_p, _i32p])
        _sig(lib, "ref_hf", None, [C.c_void_p, _u64p])
        _sig(lib, "ref_diagonal", C.c_double, [C.c_void_p, _u64p])
        _sig(lib, "ref_off_diagonal", C.c_double, [C.c_void_p, _u64p, _u64p])
        _sig(lib, "ref_column", C.c_long, [C.c_void_p, _u64p, _u64p, _f64p, C.c_long])
        _sig(lib, "ref_submatrix", None, [C.c_void_p, C.c_int, _u64p, _f64p])
        _sig(lib, "ref_ls_update", C.c_double, [C.c_double, C.c_double, C.c_double, C.c_double])
        _sig(lib, "ref_pick", None, [C.c_int, C.c_long, _u64p, _f64p, _f64p, C.c_double,
                                     C.c_double, C.c_int, _u64p, _f64p, _f64p])
        _sig(lib, "ref_wf_create", C.c_void_p, [C.c_int, C.c_long, C.c_int])
        _sig(lib, "ref_wf_destroy", None, [C.c_void_p])
        _sig(lib, "ref_wf_scalars", None, [C.c_void_p, C.c_int, _f64p])
        _sig(lib, "ref_wf_update_x", None, [C.c_void_p, C.c_int, _u64p, _f64p])
        _sig(lib, "ref_wf_update_z", C.c_long,
             [C.c_void_p, C.c_int, _i64p, _u64p, _f64p, _u64p, _f64p, C.c_double,
              _u64p, _f64p, _f64p, C.c_long])
        _sig(lib, "ref_wf_fetch", C.c_int, [C.c_void_p, _u64p, _f64p, _f64p])
        _ref = lib
    return _ref


class Oracle:
    """Thin object wrapper over one oracle context."""

    def __init__(self, fcidump: str, threshold: float = 0.0, precision: str = "fp64"):
        self.lib = oracle_lib()
        self.h = self.lib.oc_create(fcidump.encode(), threshold)
        if not self.h:
            raise RuntimeError(self.lib.oc_last_error().decode())
        info = np.zeros(4, np.int32)
        self.lib.oc_info(self.h, info)
        self.nw, self.nso, self.nelec, self.ms2 = (int(v) for v in info)
        self.set_precision(precision)

    def set_precision(self, precision: str):
        """'fp64' or 'mixed' (fp32 off-diagonal column elements, as the device)."""
        if precision not in ("fp64", "mixed"):
            raise ValueError(precision)
        self.lib.oc_set_precision(self.h, 1 if precision == "mixed" else 0)

    def close(self):
        if self.h:
            self.lib.oc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def det(self, d):
        a = np.zeros(2, np.uint64)
        a[: self.nw] = np.asarray(d, dtype=np.uint64).reshape(-1)[: self.nw]
        return a

    def hf(self):
        a = np.zeros(2, np.uint64)
        self.lib.oc_hf(self.h, a)
        return a[: self.nw].copy()

    def diagonal(self, d):
        return self.lib.oc_diagonal(self.h, self.det(d))

    def off_diagonal(self, d1, d2):
        return self.lib.oc_off_diagonal(self.h, self.det(d1), self.det(d2))

    def column(self, d):
        cap = 1 << 20
        dets = np.zeros(cap * self.nw, np.uint64)
        vals = np.zeros(cap, np.float64)
        n = self.lib.oc_column(self.h, self.det(d), dets, vals, cap)
        return dets[: n * self.nw].reshape(n, self.nw), vals[:n]

    def solver_init(self, k, M, z_thr, max_size, nthreads=1, ref=None):
        refp = None
        if ref is not None:
            self._ref = self.det(ref)
            refp = self._ref.ctypes.data
        rc = self.lib.oc_solver_init(self.h, k, M, z_thr, float(max_size), nthreads, refp)
        if rc:
            raise RuntimeError(self.lib.oc_last_error().decode())

    def step(self, n):
        return self.lib.oc_solver_step(self.h, n)

    def set_update(self, mode):
        """0: eigen update, 1: line search (one coordinate per iteration)."""
        if self.lib.oc_solver_set_update(self.h, mode):
            raise RuntimeError(self.lib.oc_last_error().decode())

    def ls_update(self, h, x, z, xx_hi, xx_lo=0.0):
        out = np.zeros(3)
        self.lib.oc_ls_update(h, x, z, xx_hi, xx_lo, out)
        return out[0], (out[1], out[2])

    def report(self):
        out = np.zeros(17)
        self.lib.oc_solver_report(self.h, out)
        keys = ["energy", "dx_max", "x_size", "z_size", "cols_size", "updates", "scale",
                "xx", "xx_lo", "scale_lo", "xz", "iterations", "max_residual",
                "max_abs_evec", "diff_inv_iter", "error", "eig_unconverged"]
        return dict(zip(keys, out.tolist()))

    def table(self):
        cap = 1 << 26
        n = self.lib.oc_table_dump(self.h, np.zeros(1, np.uint64), np.zeros(1), np.zeros(1), None, 0)
        dets = np.zeros(max(n, 1) * self.nw, np.uint64)
        x = np.zeros(max(n, 1))
        z = np.zeros(max(n, 1))
        zf = np.zeros(2 * max(n, 1), np.int64)
        self.lib.oc_table_dump(self.h, dets, x, z, zf.ctypes.data, n)
        return dets[: n * self.nw].reshape(n, self.nw), x[:n], z[:n], zf[: 2 * n].reshape(n, 2)

    def table_digest(self):
        """{entries, key-hash sum, entry-hash sum} (see mcdfci_table_digest)."""
        out = np.zeros(3, np.uint64)
        self.lib.oc_table_digest(self.h, out)
        return tuple(int(v) for v in out)

    def candidates(self):
        cap = 1 << 20
        dets = np.zeros(cap * self.nw, np.uint64)
        x = np.zeros(cap)
        z = np.zeros(cap)
        n = self.lib.oc_candidates(self.h, dets, x, z, cap)
        return dets[: n * self.nw].reshape(n, self.nw), x[:n], z[:n]

    def energy_gradient(self, dets, x):
        """(E, dE/dh, dE/d(pq|rs)) of the wavefunction (dets, x) at the
        canonical positions of the unique integrals."""
        m = self.nso // 2
        dets = np.ascontiguousarray(np.asarray(dets, dtype=np.uint64).reshape(-1, self.nw))
        x = np.ascontiguousarray(x, dtype=np.float64)
        gh = np.zeros(m * m)
        ge = np.zeros(m ** 4)
        e = C.c_double()
        rc = self.lib.oc_energy_gradient(self.h, dets.reshape(-1), x, len(x), gh, ge, C.byref(e))
        if rc:
            raise RuntimeError("zero wavefunction")
        return e.value, gh.reshape(m, m), ge.reshape(m, m, m, m)


class RefHam:
    """The reference's HamiltonianMolecule<N> built from its own sources."""

    def __init__(self, fcidump: str, threshold: float = 0.0):
        self.lib = ref_lib()
        self.h = self.lib.ref_ham_create(fcidump.encode(), threshold)
        if not self.h:
            raise RuntimeError(self.lib.ref_last_error().decode())
        info = np.zeros(4, np.int32)
        self.lib.ref_ham_info(self.h, info)
        self.nw, self.nso, self.nelec, self.ms2 = (int(v) for v in info)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.ref_ham_destroy(self.h)
            self.h = None

    def det(self, d):
        a = np.zeros(2, np.uint64)
        a[: self.nw] = np.asarray(d, dtype=np.uint64).reshape(-1)[: self.nw]
        return a

    def hf(self):
        a = np.zeros(2, np.uint64)
        self.lib.ref_hf(self.h, a)
        return a[: self.nw].copy()

    def diagonal(self, d):
        return self.lib.ref_diagonal(self.h, self.det(d))

    def off_diagonal(self, d1, d2):
        return self.lib.ref_off_diagonal(self.h, self.det(d1), self.det(d2))

    def column(self, d):
        cap = 1 << 20
        dets = np.zeros(cap * self.nw, np.uint64)
        vals = np.zeros(cap, np.float64)
        n = self.lib.ref_column(self.h, self.det(d), dets, vals, cap)
        return dets[: n * self.nw].reshape(n, self.nw), vals[:n]

    def submatrix(self, dets):
        dets = np.ascontiguousarray(np.asarray(dets, dtype=np.uint64).reshape(-1))
        k = len(dets) // self.nw
        H = np.zeros(k * k)
        self.lib.ref_submatrix(self.h, k, dets, H)
        return H.reshape(k, k)

