"""MI355X-native coordinate-descent FCI (CDFCI) hot path.

Python mirror of the C ABI in ``include/mcdfci.h`` (loaded with ctypes from
the in-tree ``libmcdfci.so``).  The class :class:`Cdfci` follows the
reference's solver surface: construction from an FCIDUMP with the
``input.json`` solver options (src/run_wrapper.h:64-122), ``init`` (the outer
``CDFCI::solve`` initialisation, src/solver.h:853-873), ``step`` (inner
iterations, src/solver.h:701-759) and ``report`` (the report row,
src/solver.h:761-798), plus the test hooks ``column``, ``submatrix``,
``eig_update``, ``table`` and ``candidates``.

There is no CPU fallback: if the HIP library is missing or cannot be
loaded, importing the bindings raises.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

__all__ = ["Cdfci", "McdfciError", "load_library", "LIB_PATH", "PREC_FP64", "PREC_MIXED"]

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmcdfci.so")
DRIVER_PATH = os.path.join(HERE, "bin", "cdfci_omp")

PREC_FP64 = 0
PREC_MIXED = 1
UPDATE_EIG = 0  # coordinate_update: subspace eigen-update
UPDATE_LS = 1   # coordinate_update: line search, one coordinate per iteration

_ERRORS = {
    1: "input", 2: "out of device memory", 3: "table full", 4: "HIP runtime",
    5: "numeric", 6: "candidates", 7: "unsupported",
}


class McdfciError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{_ERRORS.get(code, code)}] {msg}")
        self.code = code


class Opts(C.Structure):
    _fields_ = [("num_of_coordinate", C.c_int), ("estimate", C.c_int), ("z_threshold", C.c_double),
                ("max_memory", C.c_double), ("threshold", C.c_double), ("precision", C.c_int),
                ("device", C.c_int), ("coordinate_update", C.c_int)]


class Ham(C.Structure):
    _fields_ = [("norb", C.c_int), ("nelec", C.c_int), ("ms2", C.c_int), ("core", C.c_double),
                ("h", C.POINTER(C.c_double)), ("eri", C.POINTER(C.c_double)),
                ("orbital_energy", C.POINTER(C.c_double))]


class HamUhf(C.Structure):
    _fields_ = [("norb", C.c_int), ("nelec", C.c_int), ("ms2", C.c_int), ("core", C.c_double),
                ("h_alpha", C.POINTER(C.c_double)), ("h_beta", C.POINTER(C.c_double)),
                ("eri_aa", C.POINTER(C.c_double)), ("eri_bb", C.POINTER(C.c_double)),
                ("eri_ab", C.POINTER(C.c_double)), ("orbital_energy", C.POINTER(C.c_double))]


class Report(C.Structure):
    _fields_ = [("energy", C.c_double), ("dx_max", C.c_double), ("x_size", C.c_int64),
                ("z_size", C.c_int64), ("cols_size", C.c_int64), ("updates", C.c_int64),
                ("scale", C.c_double), ("xx", C.c_double), ("xz", C.c_double),
                ("iterations", C.c_int64), ("max_residual", C.c_double),
                ("curr_max_abs_evec", C.c_double), ("sum_iterations", C.c_int64),
                ("sum_inv_iter", C.c_int64), ("curr_inv_iter", C.c_int64),
                ("sum_diff_inv_iter", C.c_double), ("curr_diff_inv_iter", C.c_double),
                ("max_diff_inv_iter", C.c_double), ("table_slots", C.c_int64),
                ("max_size", C.c_double), ("candidates", C.c_int64), ("eig_unconverged", C.c_int64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_lib = None

_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")

# name -> (restype, argtypes); every symbol include/mcdfci.h declares.
SIGNATURES = {
    "mcdfci_last_error": (C.c_char_p, [C.c_void_p]),
    "mcdfci_create": (C.c_int, [C.POINTER(Ham), C.POINTER(Opts), C.POINTER(C.c_void_p)]),
    "mcdfci_create_uhf": (C.c_int, [C.POINTER(HamUhf), C.POINTER(Opts), C.POINTER(C.c_void_p)]),
    "mcdfci_create_from_fcidump": (C.c_int, [C.c_char_p, C.POINTER(Opts), C.POINTER(C.c_void_p)]),
    "mcdfci_destroy": (C.c_int, [C.c_void_p]),
    "mcdfci_info": (C.c_int, [C.c_void_p, _i64p]),
    "mcdfci_hartree_fock": (C.c_int, [C.c_void_p, _u64p]),
    "mcdfci_init": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mcdfci_set_z_threshold": (C.c_int, [C.c_void_p, C.c_double]),
    "mcdfci_step": (C.c_int, [C.c_void_p, C.c_int64]),
    "mcdfci_report_get": (C.c_int, [C.c_void_p, C.POINTER(Report)]),
    "mcdfci_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "mcdfci_timers": (C.c_int, [C.c_void_p, _f64p]),
    "mcdfci_kernel_times": (C.c_int, [C.c_void_p, _f64p, C.POINTER(C.c_int64)]),
    "mcdfci_set_graph": (C.c_int, [C.c_void_p, C.c_int]),
    "mcdfci_autotune": (C.c_int, [C.c_void_p, C.c_int, _i64p]),
    "mcdfci_set_launch": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "mcdfci_access_ceiling": (C.c_int, [C.c_void_p, C.c_int, C.c_int64, C.POINTER(C.c_double)]),
    "mcdfci_checkpoint_save": (C.c_int, [C.c_void_p, C.c_char_p]),
    "mcdfci_checkpoint_load": (C.c_int, [C.c_void_p, C.c_char_p]),
    "mcdfci_column": (C.c_int, [C.c_void_p, _u64p, _u64p, _f64p, C.c_int64, C.POINTER(C.c_int64)]),
    "mcdfci_submatrix": (C.c_int, [C.c_void_p, C.c_int, _u64p, _f64p]),
    "mcdfci_eig_update": (C.c_int, [C.c_void_p, C.c_int, _f64p, _f64p, _f64p, _f64p, _f64p]),
    "mcdfci_eig_update_noalloc": (C.c_int, [C.c_void_p, C.c_int, _f64p, _f64p, _f64p, _f64p, _f64p]),
    "mcdfci_table_dump": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                    C.POINTER(C.c_int64)]),
    "mcdfci_table_digest": (C.c_int, [C.c_void_p, _u64p]),
    "mcdfci_candidates": (C.c_int, [C.c_void_p, _u64p, _f64p, _f64p, C.c_int64, C.POINTER(C.c_int64)]),
    "mcdfci_energy_gradient": (C.c_int, [C.c_void_p, _f64p, _f64p, C.POINTER(C.c_double)]),
    "mcdfci_energy_fp64": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_double)]),
}


def load_library(path: str = LIB_PATH):
    """Load libmcdfci.so (raises OSError when it is missing: no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise OSError(f"{path} is not built; run __graft_entry__.build() (make -C mcdfci_amd/csrc)")
        lib = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


class Cdfci:
    """One device solver context (one HIP device, one stream)."""

    def __init__(self, fcidump=None, *, ham=None, ham_uhf=None, num_of_coordinate=1, estimate=100, z_threshold=0.0,
                 max_memory=1.0, threshold=1e-13, precision=PREC_FP64, device=0, coordinate_update=UPDATE_EIG):
        self.lib = load_library()
        o = Opts(int(num_of_coordinate), int(estimate), float(z_threshold), float(max_memory),
                 float(threshold), int(precision), int(device), int(coordinate_update))
        h = C.c_void_p()
        dp = C.POINTER(C.c_double)
        if fcidump is not None:
            rc = self.lib.mcdfci_create_from_fcidump(os.fsencode(fcidump), C.byref(o), C.byref(h))
        elif ham_uhf is not None:
            # (norb, nelec, ms2, core, h_alpha, h_beta, eri_aa, eri_bb, eri_ab)
            norb, nelec, ms2, core = ham_uhf[:4]
            self._uhf = [np.ascontiguousarray(a, dtype=np.float64) for a in ham_uhf[4:9]]
            hs = HamUhf(int(norb), int(nelec), int(ms2), float(core),
                        *[a.ctypes.data_as(dp) for a in self._uhf], None)
            rc = self.lib.mcdfci_create_uhf(C.byref(hs), C.byref(o), C.byref(h))
        else:
            norb, nelec, ms2, core, h1, eri = ham
            self._h1 = np.ascontiguousarray(h1, dtype=np.float64)
            self._eri = np.ascontiguousarray(eri, dtype=np.float64)
            hs = Ham(int(norb), int(nelec), int(ms2), float(core),
                     self._h1.ctypes.data_as(C.POINTER(C.c_double)),
                     self._eri.ctypes.data_as(C.POINTER(C.c_double)), None)
            rc = self.lib.mcdfci_create(C.byref(hs), C.byref(o), C.byref(h))
        if rc:
            raise McdfciError(rc, self.lib.mcdfci_last_error(None).decode())
        self.h = h
        info = np.zeros(8, np.int64)
        self._check(self.lib.mcdfci_info(self.h, info))
        self.nwords, self.nso, self.nelec, self.ms2, self.n_enum, self.k, self.has_orbital_energy, uhf = (
            int(v) for v in info[:8])
        self.uhf = bool(uhf)
        self.norb = self.nso // 2
        self.estimate = int(estimate)

    def _check(self, rc):
        if rc:
            raise McdfciError(rc, self.lib.mcdfci_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.mcdfci_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def hartree_fock(self):
        d = np.zeros(2, np.uint64)
        self._check(self.lib.mcdfci_hartree_fock(self.h, d))
        return d[: self.nwords].copy()

    def init(self, ref_det=None):
        if ref_det is None:
            self._check(self.lib.mcdfci_init(self.h, None))
        else:
            self._ref = np.zeros(2, np.uint64)
            self._ref[: self.nwords] = np.asarray(ref_det, dtype=np.uint64).reshape(-1)[: self.nwords]
            self._check(self.lib.mcdfci_init(self.h, self._ref.ctypes.data))

    def step(self, n=1):
        self._check(self.lib.mcdfci_step(self.h, int(n)))

    def report(self):
        r = Report()
        self._check(self.lib.mcdfci_report_get(self.h, C.byref(r)))
        return r.as_dict()

    def set_timing(self, on=True):
        self._check(self.lib.mcdfci_set_timing(self.h, 1 if on else 0))

    def timers(self):
        t = np.zeros(6)
        self._check(self.lib.mcdfci_timers(self.h, t))
        return t

    def kernel_times(self):
        """Seconds per kernel {pick, submatrix, eig, scatter, pass2, merge} and the
        number of timed iterations (timing mode only)."""
        t = np.zeros(6)
        n = C.c_int64()
        self._check(self.lib.mcdfci_kernel_times(self.h, t, C.byref(n)))
        return dict(zip(["pick", "submatrix", "eig", "scatter", "pass2", "merge"], t.tolist())), n.value

    def set_graph(self, iterations_per_graph):
        """Replay steps from a HIP graph of this many captured iterations (0: off)."""
        self._check(self.lib.mcdfci_set_graph(self.h, int(iterations_per_graph)))

    def autotune(self, rounds=3):
        """Time the scatter and pick launch forms on real iterations and keep the
        fastest; returns {"elements_per_thread", "tree_pick", "iterations"}."""
        ch = np.zeros(3, np.int64)
        self._check(self.lib.mcdfci_autotune(self.h, int(rounds), ch))
        return {"elements_per_thread": int(ch[0]), "tree_pick": bool(ch[1]), "iterations": int(ch[2])}

    def set_launch(self, elements_per_thread, tree_pick):
        """Pin the scatter's elements per thread and the pick form (result-neutral)."""
        self._check(self.lib.mcdfci_set_launch(self.h, int(elements_per_thread), 1 if tree_pick else 0))

    def save(self, path):
        """Write a checkpoint (table, scalars, counters, candidates) to path."""
        self._check(self.lib.mcdfci_checkpoint_save(self.h, os.fsencode(path)))

    def load(self, path):
        """Resume from a checkpoint written by save() (in place of init())."""
        self._check(self.lib.mcdfci_checkpoint_load(self.h, os.fsencode(path)))

    def access_ceiling(self, mode, n_ops=1 << 25):
        """Device seconds for n_ops random-slot operations on this context's table
        (mode 0 insert, 1 update, 2 read, 3 update by one FP64 atomic, 4 update by two
        int64 atomics).  Overwrites the table: call init() after."""
        sec = C.c_double()
        self._check(self.lib.mcdfci_access_ceiling(self.h, int(mode), int(n_ops), C.byref(sec)))
        return sec.value

    def column(self, det):
        d = np.zeros(2, np.uint64)
        d[: self.nwords] = np.asarray(det, dtype=np.uint64).reshape(-1)[: self.nwords]
        cap = max(self.n_enum, 1)
        dets = np.zeros(cap * self.nwords, np.uint64)
        vals = np.zeros(cap)
        n = C.c_int64()
        self._check(self.lib.mcdfci_column(self.h, d, dets, vals, cap, C.byref(n)))
        m = min(n.value, cap)
        return dets[: m * self.nwords].reshape(m, self.nwords), vals[:m]

    def submatrix(self, dets):
        dets = np.ascontiguousarray(np.asarray(dets, dtype=np.uint64).reshape(-1))
        k = len(dets) // self.nwords
        H = np.zeros(k * k)
        self._check(self.lib.mcdfci_submatrix(self.h, k, dets, H))
        return H.reshape(k, k)

    def eig_update(self, H, x, z, xx, scale, xz):
        """Subspace update on given inputs; returns (dx, xx_dd, scale_dd)."""
        H = np.ascontiguousarray(H, dtype=np.float64).reshape(-1)
        x = np.ascontiguousarray(x, dtype=np.float64)
        z = np.ascontiguousarray(z, dtype=np.float64)
        k = len(x)
        s = np.array([xx[0], xx[1], scale[0], scale[1], xz[0], xz[1]], dtype=np.float64)
        dx = np.zeros(k)
        self._check(self.lib.mcdfci_eig_update(self.h, k, H, x, z, s, dx))
        return dx, (s[0], s[1]), (s[2], s[3])

    def table(self):
        n = C.c_int64()
        self._check(self.lib.mcdfci_table_dump(self.h, None, None, None, 0, C.byref(n)))
        m = n.value
        dets = np.zeros(max(m, 1) * self.nwords, np.uint64)
        x = np.zeros(max(m, 1))
        z = np.zeros(max(m, 1))
        self._check(self.lib.mcdfci_table_dump(self.h, dets.ctypes.data, x.ctypes.data, z.ctypes.data, m,
                                               C.byref(n)))
        m = min(m, n.value)
        return dets[: m * self.nwords].reshape(m, self.nwords), x[:m], z[:m]

    def table_digest(self):
        """{entries, key-hash sum, entry-hash sum} of the whole table, order-independent
        and exact in z (mcdfci_table_digest)."""
        out = np.zeros(3, np.uint64)
        self._check(self.lib.mcdfci_table_digest(self.h, out))
        return tuple(int(v) for v in out)

    def candidates(self):
        cap = self.k * self.estimate
        dets = np.zeros(cap * self.nwords, np.uint64)
        x = np.zeros(cap)
        z = np.zeros(cap)
        n = C.c_int64()
        self._check(self.lib.mcdfci_candidates(self.h, dets, x, z, cap, C.byref(n)))
        m = min(n.value, cap)
        return dets[: m * self.nwords].reshape(m, self.nwords), x[:m], z[:m]

    def energy_gradient(self):
        """(E, dE/dh, dE/d(pq|rs)) of the current wavefunction: the gradient of
        the variational energy with respect to the unique integrals at their
        canonical positions (h: p >= q; eri: p >= q, r >= s, pair(pq) >=
        pair(rs)); other entries are 0."""
        n = self.norb
        gh = np.zeros(n * n)
        ge = np.zeros(n ** 4)
        e = C.c_double()
        self._check(self.lib.mcdfci_energy_gradient(self.h, gh, ge, C.byref(e)))
        return e.value, gh.reshape(n, n), ge.reshape(n, n, n, n)

    def energy_fp64(self, integrals=None):
        """FP64 Rayleigh quotient x.Hx / x.x of the current wavefunction (fp64
        integrals whatever the precision of the run).  integrals = (h, eri, core)
        evaluates the same x under those integrals instead of the context's."""
        e = C.c_double()
        if integrals is None:
            self._check(self.lib.mcdfci_energy_fp64(self.h, None, C.byref(e)))
        else:
            h1, eri, core = integrals
            h1 = np.ascontiguousarray(h1, dtype=np.float64)
            eri = np.ascontiguousarray(eri, dtype=np.float64)
            hs = Ham(self.norb, self.nelec, self.ms2, float(core), h1.ctypes.data_as(C.POINTER(C.c_double)),
                     eri.ctypes.data_as(C.POINTER(C.c_double)), None)
            self._check(self.lib.mcdfci_energy_fp64(self.h, C.byref(hs), C.byref(e)))
        return e.value
