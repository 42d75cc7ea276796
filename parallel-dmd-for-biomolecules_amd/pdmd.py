"""Python host binding of the MI355X PRIME20 DMD engine (ctypes over include/pdmd.h).

Mirrors the reference driver's path (parallel-dmd-PRIME20/code/main.F90): a
system is set up from the reference's state arrays and parameter tables,
then events are committed on the GPU in the reference's serial order.  The
compute runs in libpdmd.so (gfx950 HIP); this module only moves arguments.
Loading fails loudly when the library has not been built.
"""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
PARAM_DIR = os.path.join(PKG_DIR, "parameters")
LIB_PATH = os.path.join(PKG_DIR, "libpdmd.so")

_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_f64p = ctypes.POINTER(ctypes.c_double)


NO_HBS = 1  # pdmd_system_desc.flags: build flag -Dno_hbs
NO_H = 2  # pdmd_system_desc.flags: build flag -Dno_h (every side chain polar)
CHAPTYPE2 = 4  # pdmd_system_desc.flags: build flag -Dchaptype=2 (second species unbonded)


class SystemDesc(ctypes.Structure):
    _fields_ = [
        ("nchains", ctypes.c_int32),
        ("chain_len", _i32p),
        ("species", _i32p),
        ("restype", _i32p),
        ("hp", _i32p),
        ("boxl", ctypes.c_double),
        ("tstar", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("canon", ctypes.c_int32),
        ("list_capacity", ctypes.c_int32),
        ("flags", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("sv", _f64p),
        ("bptnr", _i32p),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("coll", ctypes.c_int64),
        ("pair_events", ctypes.c_int64),
        ("numghosts", ctypes.c_int64),
        ("nupdates", ctypes.c_int64),
        ("nforced", ctypes.c_int64),
        ("nevents", ctypes.c_int64 * 31),
        ("t", ctypes.c_double),
        ("tfalse", ctypes.c_double),
        ("interval", ctypes.c_double),
        ("interval_max", ctypes.c_double),
        ("status", ctypes.c_int32),
        ("max_list", ctypes.c_int32),
        ("nbeads", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("avg_up", ctypes.c_double),
        ("avg_dn", ctypes.c_double),
        ("phase_cycles", ctypes.c_int64 * 40),
        ("device_ms", ctypes.c_double),
        ("rebuild_ms", ctypes.c_double),
        ("rebuilds", ctypes.c_int64),
        ("waves", ctypes.c_int32),
        ("cell_factor", ctypes.c_int32),
        ("cells_axis", ctypes.c_int32),
        ("pad2", ctypes.c_int32),
    ]


class RequestInfo(ctypes.Structure):
    _fields_ = [
        ("state", ctypes.c_int32),
        ("status", ctypes.c_int32),
        ("ncoll", ctypes.c_int64),
        ("done", ctypes.c_int64),
        ("nbeads", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("t_submit", ctypes.c_double),
        ("t_admit", ctypes.c_double),
        ("t_done", ctypes.c_double),
        ("stats", Stats),
    ]


class Event(ctypes.Structure):
    _fields_ = [
        ("coll", ctypes.c_int64),
        ("i", ctypes.c_int32),
        ("j", ctypes.c_int32),
        ("type", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("t", ctypes.c_double),
    ]


EVENT_DTYPE = np.dtype([("coll", "<i8"), ("i", "<i4"), ("j", "<i4"), ("type", "<i4"), ("pad", "<i4"), ("t", "<f8")])


class Obs(ctypes.Structure):
    _fields_ = [
        ("ered", ctypes.c_double),
        ("tred", ctypes.c_double),
        ("sumvel", ctypes.c_double),
        ("hb_alpha", ctypes.c_int32),
        ("hb_ii", ctypes.c_int32),
        ("hb_ij", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("ehh_ii", ctypes.c_double),
        ("ehh_ij", ctypes.c_double),
        ("rg_avg", ctypes.c_double),
        ("e2e_avg", ctypes.c_double),
    ]


class Check(ctypes.Structure):
    _fields_ = [
        ("overlaps", ctypes.c_int64),
        ("bonds_long", ctypes.c_int64),
        ("bonds_short", ctypes.c_int64),
        ("first_i", ctypes.c_int32),
        ("first_j", ctypes.c_int32),
        ("boundbad", ctypes.c_int64),
        ("unboundbad", ctypes.c_int64),
        ("ss_counted", ctypes.c_int64),
        ("ss_total", ctypes.c_int64),
    ]


EXPORTS = [
    "pdmd_version", "pdmd_last_error", "pdmd_ctx_create", "pdmd_ctx_destroy", "pdmd_system_add",
    "pdmd_system_remove", "pdmd_nbeads", "pdmd_system_bytes", "pdmd_system_bytes_ctx", "pdmd_advance", "pdmd_advance_async", "pdmd_wait",
    "pdmd_download", "pdmd_stats", "pdmd_log_enable", "pdmd_log_read", "pdmd_predict_pairs", "pdmd_observables",
    "pdmd_generate", "pdmd_generate_seq", "pdmd_last_launch_ms", "pdmd_pair_codes",
    "pdmd_server_create", "pdmd_server_destroy", "pdmd_server_submit", "pdmd_server_step", "pdmd_server_pending",
    "pdmd_server_result", "pdmd_server_release", "pdmd_advance_budget_async", "pdmd_advance_to_output",
    "pdmd_set_waves", "pdmd_autotune_waves", "pdmd_check", "pdmd_set_graph_launch", "pdmd_rebuild",
    "pdmd_observables_noalloc", "pdmd_list_capacity", "pdmd_lists", "pdmd_set_system_tuning",
    "pdmd_autotune_systems",
]

_lib = None


def lib():
    """The loaded libpdmd.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libpdmd.so not built: run make -C {PKG_DIR} (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        L.pdmd_version.restype = ctypes.c_char_p
        L.pdmd_last_error.restype = ctypes.c_char_p
        L.pdmd_system_bytes.restype = ctypes.c_int64
        L.pdmd_ctx_create.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]
        L.pdmd_ctx_destroy.argtypes = [ctypes.c_void_p]
        L.pdmd_system_add.argtypes = [ctypes.c_void_p, ctypes.POINTER(SystemDesc), _i32p]
        L.pdmd_system_remove.argtypes = [ctypes.c_void_p, ctypes.c_int32]
        L.pdmd_nbeads.argtypes = [ctypes.c_void_p, ctypes.c_int32]
        L.pdmd_system_bytes.argtypes = [ctypes.POINTER(SystemDesc)]
        L.pdmd_system_bytes_ctx.restype = ctypes.c_int64
        L.pdmd_system_bytes_ctx.argtypes = [ctypes.c_void_p, ctypes.POINTER(SystemDesc)]
        L.pdmd_advance.argtypes = [ctypes.c_void_p, ctypes.c_int32, _i32p, _i64p]
        L.pdmd_advance_async.argtypes = [ctypes.c_void_p, ctypes.c_int32, _i32p, _i64p]
        L.pdmd_wait.argtypes = [ctypes.c_void_p]
        L.pdmd_last_launch_ms.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]
        L.pdmd_set_waves.argtypes = [ctypes.c_void_p, ctypes.c_int32]
        L.pdmd_set_graph_launch.argtypes = [ctypes.c_void_p, ctypes.c_int32]
        L.pdmd_rebuild.argtypes = [ctypes.c_void_p, ctypes.c_int32, _i32p]
        L.pdmd_check.argtypes = [ctypes.c_void_p, ctypes.c_int32, _f64p, ctypes.POINTER(Check)]
        L.pdmd_autotune_waves.argtypes = [ctypes.c_void_p, ctypes.c_int32, _i32p, ctypes.c_int64, _i32p, _f64p]
        L.pdmd_download.argtypes = [ctypes.c_void_p, ctypes.c_int32, _f64p, _i32p, _i32p, _i32p, _f64p, _i32p, _i32p]
        L.pdmd_stats.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(Stats)]
        L.pdmd_log_enable.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]
        L.pdmd_log_read.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, _i32p]
        L.pdmd_predict_pairs.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, _i32p, _i32p, _f64p, _i32p]
        L.pdmd_observables.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(Obs)]
        L.pdmd_observables_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(Obs)]
        L.pdmd_list_capacity.argtypes = [ctypes.c_void_p, ctypes.c_int32]
        L.pdmd_set_system_tuning.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
        L.pdmd_autotune_systems.argtypes = [ctypes.c_void_p, ctypes.c_int32, _i32p, ctypes.c_int64, _i32p, _i32p,
                                            _f64p]
        L.pdmd_lists.argtypes = [ctypes.c_void_p, ctypes.c_int32, _i32p, _i32p, _i32p, _i32p]
        L.pdmd_generate.argtypes = [ctypes.c_char_p, ctypes.c_int32, _i32p, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_uint64, _i32p, _i32p, _f64p, _f64p]
        L.pdmd_server_create.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_double,
                                         ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]
        L.pdmd_advance_budget_async.argtypes = [ctypes.c_void_p, ctypes.c_int32, _i32p, _i64p, ctypes.c_double]
        L.pdmd_advance_to_output.argtypes = [ctypes.c_void_p, ctypes.c_int32, _i32p, _i64p]
        L.pdmd_server_destroy.argtypes = [ctypes.c_void_p]
        L.pdmd_server_submit.argtypes = [ctypes.c_void_p, ctypes.POINTER(SystemDesc), ctypes.c_int64, _i64p]
        L.pdmd_server_step.argtypes = [ctypes.c_void_p, _i32p]
        L.pdmd_server_pending.argtypes = [ctypes.c_void_p, _i32p, _i32p, _i64p]
        L.pdmd_server_result.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(RequestInfo), _f64p]
        L.pdmd_server_release.argtypes = [ctypes.c_void_p, ctypes.c_int64]
        L.pdmd_pair_codes.argtypes = [ctypes.POINTER(SystemDesc), ctypes.c_int32, _i32p, _i32p, _i32p]
        L.pdmd_generate_seq.argtypes = [ctypes.c_char_p, ctypes.c_int32, _i32p, _i32p, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_uint64, _i32p, _i32p, _f64p, _f64p]
        _lib = L
    return _lib


class PdmdError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise PdmdError(f"pdmd error {rc}: {lib().pdmd_last_error().decode()}")


def _p(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype)) if a is not None else None


def generate(chain_len, density=0.01, tstar=0.20, seed=1, param_dir=PARAM_DIR, sequence=None):
    """Seeded synthetic PRIME20 system (sequences, coils, MB velocities).

    sequence: optional side-chain identities (9..28) per residue, all chains
    concatenated; None draws them from the seed."""
    cl = np.ascontiguousarray(chain_len, dtype=np.int32)
    nres = int(cl.sum())
    N = 4 * nres
    restype = np.zeros(nres, np.int32)
    hp = np.zeros(nres, np.int32)
    sv = np.zeros(6 * N, np.float64)
    boxl = ctypes.c_double(0.0)
    seq = None
    if sequence is not None:
        seq = np.ascontiguousarray(sequence, dtype=np.int32)
        if seq.size != nres:
            raise ValueError("sequence length differs from sum(chain_len)")
    _check(lib().pdmd_generate_seq(param_dir.encode(), len(cl), _p(cl, ctypes.c_int32),
                                   None if seq is None else _p(seq, ctypes.c_int32), density, tstar, seed,
                                   _p(restype, ctypes.c_int32), _p(hp, ctypes.c_int32), _p(sv, ctypes.c_double),
                                   ctypes.byref(boxl)))
    return {"chain_len": cl, "restype": restype, "hp": hp, "sv": sv, "boxl": boxl.value, "tstar": tstar,
            "density": density}


def pair_codes(system, a, b, species=None, no_hbs=False, no_h=False, chaptype2=False):
    """Static pair codes ev_code(a, b) (1-based beads) of a system; host only."""
    d, keep = Engine.make_desc(system, 0, species=species, no_hbs=no_hbs, no_h=no_h, chaptype2=chaptype2)
    a = np.ascontiguousarray(a, np.int32)
    b = np.ascontiguousarray(b, np.int32)
    out = np.zeros(a.size, np.int32)
    _check(lib().pdmd_pair_codes(ctypes.byref(d), a.size, _p(a, ctypes.c_int32), _p(b, ctypes.c_int32),
                                 _p(out, ctypes.c_int32)))
    return out


class Engine:
    """A device context holding resident systems (pdmd_ctx)."""

    def __init__(self, param_dir=PARAM_DIR, waves=1):
        self._ctx = ctypes.c_void_p()
        _check(lib().pdmd_ctx_create(param_dir.encode(), ctypes.byref(self._ctx)))
        self._keep = {}
        self.nbeads = {}
        if waves != 1:
            self.set_waves(waves)

    def set_waves(self, waves):
        """Wavefronts per system for later advances (pdmd_set_waves)."""
        _check(lib().pdmd_set_waves(self._ctx, int(waves)))

    def rebuild(self, ids):
        """Neighbour lists and calendar rebuilt now, at each system's
        current time (pdmd_rebuild)."""
        a = np.ascontiguousarray(ids, np.int32)
        _check(lib().pdmd_rebuild(self._ctx, len(a), _p(a, ctypes.c_int32)))

    def set_graph_launch(self, on=True):
        """Replay a captured HIP graph of the event-loop launch while the
        launch shape stays the same (pdmd_set_graph_launch)."""
        _check(lib().pdmd_set_graph_launch(self._ctx, 1 if on else 0))

    def check(self, sid, sv=None):
        """checkover.f / check_nc_int.f on the device (pdmd_check): dict of
        counts; sv replaces the resident positions/velocities when given."""
        out = Check()
        keep = None if sv is None else np.ascontiguousarray(sv, np.float64)
        _check(lib().pdmd_check(self._ctx, sid, None if keep is None else _p(keep, ctypes.c_double), ctypes.byref(out)))
        return {n: int(getattr(out, n)) for n, _ in Check._fields_}

    def set_system_tuning(self, sid, waves=0, cell_factor=0):
        _check(lib().pdmd_set_system_tuning(self._ctx, sid, waves, cell_factor))

    def autotune_systems(self, ids, probe_events=3000):
        """pdmd_autotune_systems: per-system wavefronts and rebuild cell
        factor; returns (waves, cell_factor, rates[nsys, 4])."""
        ids = np.ascontiguousarray(ids, np.int32)
        n = len(ids)
        w, cf, r = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(4 * n)
        _check(lib().pdmd_autotune_systems(self._ctx, n, _p(ids, ctypes.c_int32), int(probe_events),
                                           _p(w, ctypes.c_int32), _p(cf, ctypes.c_int32), _p(r, ctypes.c_double)))
        return w, cf, r.reshape(n, 4)

    def autotune_waves(self, ids, probe_events=2000):
        """Pick the wavefronts per system by timing short advances of the
        listed systems (pdmd_autotune_waves); returns (waves, {waves: rate})."""
        a = np.ascontiguousarray(ids, np.int32)
        w = ctypes.c_int32(0)
        rates = np.zeros(4, np.float64)
        _check(lib().pdmd_autotune_waves(self._ctx, len(a), _p(a, ctypes.c_int32), int(probe_events), ctypes.byref(w),
                                         _p(rates, ctypes.c_double)))
        return w.value, {k: float(r) for k, r in zip((1, 2, 4, 8), rates) if r > 0}

    def close(self):
        if self._ctx:
            lib().pdmd_ctx_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def make_desc(system, seed, canon=True, bptnr=None, species=None, list_capacity=0, no_hbs=False, no_h=False,
                  chaptype2=False):
        keep = dict(cl=np.ascontiguousarray(system["chain_len"], np.int32),
                    rt=np.ascontiguousarray(system["restype"], np.int32),
                    hp=np.ascontiguousarray(system["hp"], np.int32),
                    sv=np.ascontiguousarray(system["sv"], np.float64),
                    bp=None if bptnr is None else np.ascontiguousarray(bptnr, np.int32),
                    sp=None if species is None else np.ascontiguousarray(species, np.int32))
        d = SystemDesc(len(keep["cl"]), _p(keep["cl"], ctypes.c_int32), _p(keep["sp"], ctypes.c_int32),
                       _p(keep["rt"], ctypes.c_int32), _p(keep["hp"], ctypes.c_int32), float(system["boxl"]),
                       float(system["tstar"]), int(seed), 1 if canon else 0, int(list_capacity),
                       (NO_HBS if no_hbs else 0) | (NO_H if no_h else 0) | (CHAPTYPE2 if chaptype2 else 0), 0,
                       _p(keep["sv"], ctypes.c_double), _p(keep["bp"], ctypes.c_int32))
        return d, keep

    def add_system(self, system, seed, canon=True, bptnr=None, species=None, list_capacity=0, no_hbs=False,
                   no_h=False, chaptype2=False):
        d, keep = self.make_desc(system, seed, canon, bptnr, species, list_capacity, no_hbs, no_h, chaptype2)
        sid = ctypes.c_int32(0)
        _check(lib().pdmd_system_add(self._ctx, ctypes.byref(d), ctypes.byref(sid)))
        self._keep[sid.value] = keep
        self.nbeads[sid.value] = lib().pdmd_nbeads(self._ctx, sid.value)
        return sid.value

    def remove_system(self, sid):
        _check(lib().pdmd_system_remove(self._ctx, sid))
        self._keep.pop(sid, None)
        self.nbeads.pop(sid, None)

    def advance(self, ids, nevents, wait=True):
        ids = np.ascontiguousarray(np.atleast_1d(ids), np.int32)
        nev = np.ascontiguousarray(np.broadcast_to(np.asarray(nevents, np.int64), ids.shape), np.int64)
        fn = lib().pdmd_advance if wait else lib().pdmd_advance_async
        _check(fn(self._ctx, len(ids), _p(ids, ctypes.c_int32), _p(nev, ctypes.c_int64)))

    def wait(self):
        _check(lib().pdmd_wait(self._ctx))

    def last_launch_ms(self):
        ms = ctypes.c_float(0.0)
        _check(lib().pdmd_last_launch_ms(self._ctx, ctypes.byref(ms)))
        return ms.value

    def download(self, sid):
        N = self.nbeads[sid]
        out = dict(sv=np.zeros(6 * N), identity=np.zeros(N, np.int32), bptnr=np.zeros(N, np.int32),
                   extra=np.zeros(4 * N, np.int32), tim=np.zeros(N + 3), nptnr=np.zeros(N + 3, np.int32),
                   coltype=np.zeros(N + 3, np.int32))
        _check(lib().pdmd_download(self._ctx, sid, _p(out["sv"], ctypes.c_double),
                                   _p(out["identity"], ctypes.c_int32), _p(out["bptnr"], ctypes.c_int32),
                                   _p(out["extra"], ctypes.c_int32), _p(out["tim"], ctypes.c_double),
                                   _p(out["nptnr"], ctypes.c_int32), _p(out["coltype"], ctypes.c_int32)))
        return out

    def stats(self, sid):
        s = Stats()
        _check(lib().pdmd_stats(self._ctx, sid, ctypes.byref(s)))
        return s

    def lists(self, sid):
        """Neighbour rows as the last rebuild left them: (up, dn), lists of
        1-based partner arrays per bead (pdmd_lists)."""
        N = self.nbeads[sid]
        cap = lib().pdmd_list_capacity(self._ctx, sid)
        nu, nd = np.zeros(N, np.int32), np.zeros(N, np.int32)
        up, dn = np.zeros(N * cap, np.int32), np.zeros(N * cap, np.int32)
        _check(lib().pdmd_lists(self._ctx, sid, _p(nu, ctypes.c_int32), _p(up, ctypes.c_int32),
                                _p(nd, ctypes.c_int32), _p(dn, ctypes.c_int32)))
        up, dn = up.reshape(N, cap), dn.reshape(N, cap)
        return [up[k, :nu[k]].copy() for k in range(N)], [dn[k, :nd[k]].copy() for k in range(N)]

    def log_enable(self, sid, capacity):
        _check(lib().pdmd_log_enable(self._ctx, sid, capacity))

    def log_read(self, sid, maxn):
        out = np.zeros(maxn, EVENT_DTYPE)
        n = ctypes.c_int32(0)
        _check(lib().pdmd_log_read(self._ctx, sid, maxn, out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(n)))
        return out[: n.value]

    def predict_pairs(self, sid, i, j):
        i = np.ascontiguousarray(i, np.int32)
        j = np.ascontiguousarray(j, np.int32)
        tij = np.zeros(len(i))
        ty = np.zeros(len(i), np.int32)
        _check(lib().pdmd_predict_pairs(self._ctx, sid, len(i), _p(i, ctypes.c_int32), _p(j, ctypes.c_int32),
                                        _p(tij, ctypes.c_double), _p(ty, ctypes.c_int32)))
        return tij, ty

    def observables(self, sid):
        o = Obs()
        _check(lib().pdmd_observables(self._ctx, sid, ctypes.byref(o)))
        return o

    def observables_noalloc(self, sid):
        """pdmd_observables_noalloc: the same values, no device allocation."""
        o = Obs()
        _check(lib().pdmd_observables_noalloc(self._ctx, sid, ctypes.byref(o)))
        return o


class Server:
    """Continuous-batching request loop on an Engine's context (pdmd_server_*)."""

    def __init__(self, engine, mem_budget_bytes, quantum, slice_ms=0.0, max_resident=0):
        self.engine = engine
        self._h = ctypes.c_void_p()
        _check(lib().pdmd_server_create(engine._ctx, int(mem_budget_bytes), int(quantum), float(slice_ms),
                                        int(max_resident), ctypes.byref(self._h)))

    def close(self):
        if self._h:
            lib().pdmd_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, system, ncoll, seed, canon=True, species=None, bptnr=None, no_hbs=False, no_h=False,
               chaptype2=False):
        d, keep = Engine.make_desc(system, seed, canon, bptnr, species, no_hbs=no_hbs, no_h=no_h, chaptype2=chaptype2)
        rid = ctypes.c_int64(0)
        _check(lib().pdmd_server_submit(self._h, ctypes.byref(d), int(ncoll), ctypes.byref(rid)))
        return rid.value

    def step(self):
        n = ctypes.c_int32(0)
        _check(lib().pdmd_server_step(self._h, ctypes.byref(n)))
        return n.value

    def pending(self):
        q, r, c = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_int64(0)
        _check(lib().pdmd_server_pending(self._h, ctypes.byref(q), ctypes.byref(r), ctypes.byref(c)))
        return q.value, r.value, c.value

    def result(self, rid, with_state=False):
        info = RequestInfo()
        _check(lib().pdmd_server_result(self._h, int(rid), ctypes.byref(info), None))
        sv = None
        if with_state and info.state >= 2:
            sv = np.zeros(6 * info.nbeads)
            _check(lib().pdmd_server_result(self._h, int(rid), ctypes.byref(info), _p(sv, ctypes.c_double)))
        return info, sv

    def release(self, rid):
        return lib().pdmd_server_release(self._h, int(rid))
