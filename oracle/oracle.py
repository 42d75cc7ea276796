"""ctypes binding of the CPU oracle (test infrastructure only).

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
never by the product.  See pdmd_oracle.c for what it restates.
"""
import ctypes
import os

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(ORACLE_DIR, "_build", "liboracle.so")

EVENT_DTYPE = np.dtype([("coll", "<i8"), ("i", "<i4"), ("j", "<i4"), ("type", "<i4"), ("pad", "<i4"), ("t", "<f8")])

_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)


class OrcStats(ctypes.Structure):
    _fields_ = [
        ("coll", ctypes.c_int64), ("numghosts", ctypes.c_int64), ("nupdates", ctypes.c_int64),
        ("nforced", ctypes.c_int64), ("nevents", ctypes.c_int64 * 31),
        ("t", ctypes.c_double), ("tfalse", ctypes.c_double), ("interval", ctypes.c_double),
        ("interval_max", ctypes.c_double), ("t_fact", ctypes.c_double), ("n_forced", ctypes.c_double),
        ("sortsize", ctypes.c_double), ("tbin_off", ctypes.c_double), ("nbin", ctypes.c_int32),
        ("num_cell", ctypes.c_int32), ("sig_max_all", ctypes.c_double), ("hdelr", ctypes.c_double),
        ("avg_nbrs", ctypes.c_double), ("rng_ctr", ctypes.c_int64 * 2),
    ]


class OrcObs(ctypes.Structure):
    _fields_ = [
        ("ered", ctypes.c_double), ("tred", ctypes.c_double), ("sumvel", ctypes.c_double),
        ("hb_alpha", ctypes.c_int32), ("hb_ii", ctypes.c_int32), ("hb_ij", ctypes.c_int32), ("pad", ctypes.c_int32),
        ("ehh_ii", ctypes.c_double), ("ehh_ij", ctypes.c_double), ("rg_avg", ctypes.c_double),
        ("e2e_avg", ctypes.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: make -C oracle")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_char_p, ctypes.c_int, _i32p, _i32p, _i32p, _i32p, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_uint64, ctypes.c_int, _f64p, _i32p, ctypes.c_char_p,
                                 ctypes.c_int]
        L.orc_create_ex.restype = ctypes.c_void_p
        L.orc_create_ex.argtypes = [ctypes.c_char_p, ctypes.c_int, _i32p, _i32p, _i32p, _i32p, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, _f64p, _i32p,
                                    ctypes.c_char_p, ctypes.c_int]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_run.restype = ctypes.c_int64
        L.orc_run.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64]
        L.orc_rebuild.argtypes = [ctypes.c_void_p]
        L.orc_nbeads.argtypes = [ctypes.c_void_p]
        L.orc_get_state.argtypes = [ctypes.c_void_p, _f64p, _i32p, _i32p, _i32p, _f64p, _i32p, _i32p]
        L.orc_stats_get.argtypes = [ctypes.c_void_p, ctypes.POINTER(OrcStats)]
        L.orc_observables.argtypes = [ctypes.c_void_p, ctypes.POINTER(OrcObs)]
        L.orc_ev_code.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]
        L.orc_nb_up.argtypes = [ctypes.c_void_p, ctypes.c_int32, _i32p, ctypes.c_int32]
        L.orc_nb_dn.argtypes = [ctypes.c_void_p, ctypes.c_int32, _i32p, ctypes.c_int32]
        L.orc_cell_of.argtypes = [ctypes.c_void_p, ctypes.c_int32]
        L.orc_predict.restype = ctypes.c_double
        L.orc_predict.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, _i32p]
        L.orc_checkover.argtypes = [ctypes.c_void_p]
        L.orc_check.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_table.restype = ctypes.c_double
        L.orc_table.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_bond.restype = ctypes.c_double
        L.orc_bond.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.orc_rng_u01.restype = ctypes.c_double
        L.orc_rng_u01.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_uint64]
        L.orc_log.restype = ctypes.c_double
        L.orc_log.argtypes = [ctypes.c_double]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t)) if a is not None else None


class Oracle:
    """One system in the CPU oracle."""

    def __init__(self, system, seed, param_dir, canon=True, bptnr=None, species=None, no_hbs=False, no_h=False,
                 chaptype2=False):
        self.cl = np.ascontiguousarray(system["chain_len"], np.int32)
        self.rt = np.ascontiguousarray(system["restype"], np.int32)
        self.hp = np.ascontiguousarray(system["hp"], np.int32)
        self.sv = np.ascontiguousarray(system["sv"], np.float64)
        self.sp = None if species is None else np.ascontiguousarray(species, np.int32)
        N = 4 * int(self.cl.sum())
        self.bp = np.zeros(N, np.int32) if bptnr is None else np.ascontiguousarray(bptnr, np.int32)
        err = ctypes.create_string_buffer(256)
        self.h = lib().orc_create_ex(param_dir.encode(), len(self.cl), _p(self.cl, ctypes.c_int32),
                                     _p(self.sp, ctypes.c_int32), _p(self.rt, ctypes.c_int32),
                                     _p(self.hp, ctypes.c_int32), float(system["boxl"]), float(system["tstar"]),
                                     int(seed), 1 if canon else 0, (1 if no_hbs else 0) | (2 if no_h else 0) | (4 if chaptype2 else 0), _p(self.sv, ctypes.c_double),
                                     _p(self.bp, ctypes.c_int32), err, 256)
        if not self.h:
            raise RuntimeError("orc_create failed: " + err.value.decode())
        self.N = N

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def run(self, nev, log_cap=0):
        log = np.zeros(max(log_cap, 1), EVENT_DTYPE)
        r = lib().orc_run(self.h, int(nev), log.ctypes.data_as(ctypes.c_void_p), int(log_cap))
        if r < 0:
            raise RuntimeError(f"oracle self-check failed ({r})")
        return log[: min(log_cap, nev)] if log_cap else None

    def rebuild(self):
        """List update forced at the current time (orc_rebuild)."""
        lib().orc_rebuild(self.h)

    def state(self):
        N = self.N
        out = dict(sv=np.zeros(6 * N), identity=np.zeros(N, np.int32), bptnr=np.zeros(N, np.int32),
                   extra=np.zeros(4 * N, np.int32), tim=np.zeros(N + 3), nptnr=np.zeros(N + 3, np.int32),
                   coltype=np.zeros(N + 3, np.int32))
        lib().orc_get_state(self.h, _p(out["sv"], ctypes.c_double), _p(out["identity"], ctypes.c_int32),
                            _p(out["bptnr"], ctypes.c_int32), _p(out["extra"], ctypes.c_int32),
                            _p(out["tim"], ctypes.c_double), _p(out["nptnr"], ctypes.c_int32),
                            _p(out["coltype"], ctypes.c_int32))
        return out

    def stats(self):
        s = OrcStats()
        lib().orc_stats_get(self.h, ctypes.byref(s))
        return s

    def observables(self):
        o = OrcObs()
        lib().orc_observables(self.h, ctypes.byref(o))
        return o

    def ev_code(self, a, b):
        return lib().orc_ev_code(self.h, a, b)

    def predict(self, i, j):
        ty = ctypes.c_int32(0)
        t = lib().orc_predict(self.h, i, j, ctypes.byref(ty))
        return t, ty.value

    def nb_up(self, i):
        buf = np.zeros(1024, np.int32)
        n = lib().orc_nb_up(self.h, i, _p(buf, ctypes.c_int32), 1024)
        return buf[:n].copy()

    def nb_dn(self, i):
        buf = np.zeros(1024, np.int32)
        n = lib().orc_nb_dn(self.h, i, _p(buf, ctypes.c_int32), 1024)
        return buf[:n].copy()

    def checkover(self):
        return lib().orc_checkover(self.h)

    def check(self, sv=None):
        """checkover.f split by kind and check_nc_int.f: dict of counts."""
        out = np.zeros(9, np.int64)
        keep = None if sv is None else np.ascontiguousarray(sv, np.float64)
        lib().orc_check(self.h, None if keep is None else keep.ctypes.data, out.ctypes.data)
        k = ("overlaps", "bonds_long", "bonds_short", "first_i", "first_j", "boundbad", "unboundbad", "ss_counted",
             "ss_total")
        return {n: int(v) for n, v in zip(k, out)}

    def table(self, which, k, kk=0):
        return lib().orc_table(self.h, which, k, kk)

    def bond(self, which, q):
        return lib().orc_bond(self.h, which, q)


def rng_u01(seed, stream, ctr):
    return lib().orc_rng_u01(seed, stream, ctr)


def orc_log(x):
    return lib().orc_log(x)
