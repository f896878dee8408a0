"""ctypes binding of ``libpphpc.so`` (C ABI of ``include/pphpc.h``).

The library is loaded from ``pphpc_amd/lib/libpphpc.so`` inside this package;
there is no fallback: if the HIP library is missing, importing works but any
call raises :class:`RuntimeError` telling how to build it.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

PP_SUCCESS = 0
PP_UNKNOWN_ARGS = -1
PP_INVALID_ARGS = -2
PP_LIBRARY_ERROR = -3
PP_UNABLE_TO_OPEN_PARAMS_FILE = -4
PP_INVALID_PARAMS_FILE = -5
PP_UNABLE_SAVE_STATS = -6
PP_OUT_OF_RESOURCES = -8
PP_ENERGY_OVERFLOW = -9   # no counterpart in the reference: energy limit

# Event sums of pp_sim_counters / pp_batch_counters: agents at the start of
# each iteration, alive after the move, after actions, occupied cells, grass
# cells eaten, births.
COUNTER_FIELDS = ("agent_steps", "live", "out", "occupied", "eaten", "births")

# Field order of PPParameters (opencl/pp/pp_common.in.h:129-157) and the keys
# of the parameter file that fill them (pp_common.c:55-167).
PARAM_FIELDS = (
    "init_sheep", "sheep_gain_from_food", "sheep_reproduce_threshold",
    "sheep_reproduce_prob", "init_wolves", "wolves_gain_from_food",
    "wolves_reproduce_threshold", "wolves_reproduce_prob", "grass_restart",
    "grid_x", "grid_y", "grid_xy", "iters",
)
PARAM_KEYS = (
    "INIT_SHEEP", "SHEEP_GAIN_FROM_FOOD", "SHEEP_REPRODUCE_THRESHOLD",
    "SHEEP_REPRODUCE_PROB", "INIT_WOLVES", "WOLVES_GAIN_FROM_FOOD",
    "WOLVES_REPRODUCE_THRESHOLD", "WOLVES_REPRODUCE_PROB", "GRASS_RESTART",
    "GRID_X", "GRID_Y", "ITERS",
)
# Field order of PPStatistics (pp_common.in.h:109-124).
STAT_FIELDS = ("sheep", "wolves", "grass", "sheep_en", "wolves_en",
               "grass_en", "errors")


class Params(ctypes.Structure):
    _fields_ = [(f, ctypes.c_uint32) for f in PARAM_FIELDS]

    def as_dict(self):
        return {f: getattr(self, f) for f in PARAM_FIELDS}

    @classmethod
    def from_values(cls, **kw):
        p = cls()
        for f in PARAM_FIELDS:
            if f == "grid_xy":
                continue
            setattr(p, f, int(kw[f]))
        p.grid_xy = p.grid_x * p.grid_y
        return p

    def copy(self, **changes):
        d = self.as_dict()
        d.update(changes)
        return Params.from_values(**d)


class Stats(ctypes.Structure):
    _fields_ = [(f, ctypes.c_uint64) for f in STAT_FIELDS]

    def as_tuple(self):
        return tuple(getattr(self, f) for f in STAT_FIELDS)


class PPError(RuntimeError):
    def __init__(self, code: int, what: str = ""):
        self.code = code
        super().__init__(f"{what}: {strerror(code)} ({code})" if what else
                         f"{strerror(code)} ({code})")


_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def library_path() -> str:
    return os.path.join(_HERE, "lib", "libpphpc.so")


def _lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    path = library_path()
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is not built: run `make -C pphpc_amd/csrc` (or "
            "__graft_entry__.build()). There is no fallback path.")
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER
    u32, u64, i32 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
    vp = ctypes.c_void_p
    sig = {
        "pp_load_params": (i32, [ctypes.c_char_p, P(Params)]),
        "pp_stats_save": (i32, [ctypes.c_char_p, P(Stats), u32, P(Params)]),
        "pp_stats_save_f32": (i32, [ctypes.c_char_p, P(Stats), u32,
                                    P(Params)]),
        "pp_strerror": (ctypes.c_char_p, [i32]),
        "pp_sim_create": (i32, [P(Params), u64, u64, i32, P(vp)]),
        "pp_sim_init": (i32, [vp, P(Stats)]),
        "pp_sim_step": (i32, [vp, P(Stats)]),
        "pp_sim_run": (i32, [vp, u32, P(Stats)]),
        "pp_sim_simulate": (i32, [vp, P(Stats)]),
        "pp_sim_step_count": (u32, [vp]),
        "pp_sim_num_agents": (u64, [vp]),
        "pp_sim_error_step": (u32, [vp]),
        "pp_sim_agent_steps": (u64, [vp]),
        "pp_sim_capacity": (u64, [vp]),
        "pp_sim_counters": (i32, [vp, P(u64)]),
        "pp_sim_get_state": (i32, [vp, vp, vp, vp, vp]),
        "pp_sim_set_profiling": (i32, [vp, i32]),
        "pp_sim_prof_count": (i32, [vp]),
        "pp_sim_prof_get": (i32, [vp, i32, P(ctypes.c_char_p), P(u64),
                                  P(ctypes.c_double)]),
        "pp_sim_prof_reset": (i32, [vp]),
        "pp_sim_prof_export": (i32, [vp, ctypes.c_char_p]),
        "pp_sim_info": (i32, [vp, P(u32), P(u32), P(u32), P(u32)]),
        "pp_sim_destroy": (None, [vp]),
        "pp_sim_sync": (i32, [vp]),
        "pp_sim_set_sort": (i32, [vp, u32]),
        "pp_sim_simulate_noalloc": (i32, [P(Params), u64, u64, i32,
                                          P(Stats)]),
        "pp_sim_dump": (i32, [vp, ctypes.c_char_p, ctypes.c_char_p, u32,
                              i32]),
        "pp_dump_state": (i32, [ctypes.c_char_p, ctypes.c_char_p, i32, u32,
                                u32, P(Params), u64, vp, vp, vp, vp]),
        "pp_batch_footprint": (u64, [u64, u64, u32, u32]),
        "pp_batch_create": (i32, [u64, u64, u32, u32, i32, P(vp)]),
        "pp_batch_add": (i32, [vp, P(Params), u64, u64, P(u32)]),
        "pp_batch_run": (i32, [vp, u32]),
        "pp_batch_replica": (i32, [vp, u32, P(i32), P(u32), P(i32), P(u64)]),
        "pp_batch_rows": (i32, [vp, u32, u32, u32, P(Stats)]),
        "pp_batch_get_state": (i32, [vp, u32, P(u32), vp, vp, vp, vp]),
        "pp_batch_release": (i32, [vp, u32]),
        "pp_batch_resources": (i32, [vp, P(u64), P(u32), P(u32), P(u32)]),
        "pp_batch_tick": (u32, [vp]),
        "pp_batch_counters": (i32, [vp, P(u64)]),
        "pp_batch_info": (i32, [vp, P(u32), P(u32), P(u32), P(u32)]),
        "pp_batch_set_shape": (i32, [vp, u32, u32]),
        "pp_batch_tune": (i32, [vp, u32, P(u32), P(u32), P(u32),
                                P(ctypes.c_double)]),
        "pp_batch_set_sort": (i32, [vp, u32]),
        "pp_batch_set_launch": (i32, [vp, u32, u32]),
        "pp_batch_get_launch": (i32, [vp, P(u32), P(u32)]),
        "pp_batch_set_profiling": (i32, [vp, i32]),
        "pp_batch_prof_get": (i32, [vp, i32, P(ctypes.c_char_p), P(u64),
                                    P(ctypes.c_double)]),
        "pp_batch_prof_reset": (i32, [vp]),
        "pp_batch_prof_count": (i32, [vp]),
        "pp_batch_sync": (i32, [vp]),
        "pp_batch_destroy": (None, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def strerror(code: int) -> str:
    try:
        return _lib().pp_strerror(int(code)).decode()
    except RuntimeError:
        return f"error {code}"


def load_params(path: Optional[str] = None) -> Params:
    """Parse a parameter file (pp_load_params, pp_common.c:39-195)."""
    p = Params()
    rc = _lib().pp_load_params(path.encode() if path else None,
                               ctypes.byref(p))
    if rc:
        raise PPError(rc, f"load_params({path!r})")
    return p


def stats_save(path: Optional[str], rows, params: Params,
               f32: bool = False) -> None:
    """Write the statistics file (pp_stats_save, pp_common.c:205-240); f32:
    as the reference prints it (32-bit counts, single-precision means)."""
    arr = (Stats * len(rows))(*rows) if not isinstance(rows, ctypes.Array) \
        else rows
    fn = _lib().pp_stats_save_f32 if f32 else _lib().pp_stats_save
    rc = fn(path.encode() if path else None, arr, len(arr),
            ctypes.byref(params))
    if rc:
        raise PPError(rc, f"stats_save({path!r})")


def simulate_noalloc(params: Params, seed: int = 0, capacity: int = 0,
                     device: int = -1) -> np.ndarray:
    """pp_sim_simulate_noalloc: rows 0..iters of one simulation run in the
    library's fixed device scratch (no device allocation), as an (iters+1, 7)
    uint64 array."""
    out = (Stats * (params.iters + 1))()
    rc = _lib().pp_sim_simulate_noalloc(ctypes.byref(params), seed, capacity,
                                        device, out)
    if rc:
        raise PPError(rc, "simulate_noalloc")
    return stats_rows_to_array(out)


def stats_rows_to_array(rows) -> np.ndarray:
    """(n, 7) uint64 array of statistics rows."""
    return np.array([r.as_tuple() for r in rows], dtype=np.uint64).reshape(
        -1, len(STAT_FIELDS))


class Simulation:
    """One simulation resident on a GPU (a ``pp_sim_t`` handle)."""

    def __init__(self, params: Params, seed: int = 0, capacity: int = 0,
                 device: int = -1):
        self.params = params
        self.seed = seed
        h = ctypes.c_void_p()
        rc = _lib().pp_sim_create(ctypes.byref(params), seed, capacity,
                                  device, ctypes.byref(h))
        if rc:
            raise PPError(rc, "pp_sim_create")
        self._h = h

    # -- lifecycle ---------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            _lib().pp_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc, what):
        if rc:
            raise PPError(rc, what)

    # -- simulation --------------------------------------------------------
    def init(self) -> Stats:
        s = Stats()
        self._check(_lib().pp_sim_init(self._h, ctypes.byref(s)), "init")
        return s

    def step(self) -> Stats:
        s = Stats()
        self._check(_lib().pp_sim_step(self._h, ctypes.byref(s)), "step")
        return s

    def run(self, iters: int, collect: bool = True):
        """Run ``iters`` iterations; returns their stats rows (1..iters)."""
        rows = (Stats * iters)() if collect and iters else None
        rc = _lib().pp_sim_run(self._h, iters, rows)
        if rc:
            raise PPError(rc, f"run({iters}) stopped at iteration "
                              f"{self.error_step}")
        return list(rows) if rows is not None else []

    def run_array(self, iters: int) -> np.ndarray:
        return stats_rows_to_array(self.run(iters))

    def simulate(self):
        """init + ITERS iterations: rows 0..ITERS (pp_gpu.c:1857-1866)."""
        n = self.params.iters + 1
        rows = (Stats * n)()
        rc = _lib().pp_sim_simulate(self._h, rows)
        if rc:
            raise PPError(rc, f"simulate stopped at iteration "
                              f"{self.error_step}")
        return list(rows)

    # -- queries -----------------------------------------------------------
    @property
    def step_count(self) -> int:
        return _lib().pp_sim_step_count(self._h)

    @property
    def num_agents(self) -> int:
        return _lib().pp_sim_num_agents(self._h)

    @property
    def error_step(self) -> int:
        return _lib().pp_sim_error_step(self._h)

    @property
    def agent_steps(self) -> int:
        return _lib().pp_sim_agent_steps(self._h)

    @property
    def capacity(self) -> int:
        return _lib().pp_sim_capacity(self._h)

    def counters(self) -> dict:
        """Exact sums over completed iterations since init."""
        out = (ctypes.c_uint64 * 6)()
        self._check(_lib().pp_sim_counters(self._h, out), "counters")
        return dict(zip(COUNTER_FIELDS, tuple(out)))

    def info(self) -> dict:
        a, b, c, d = (ctypes.c_uint32() for _ in range(4))
        self._check(_lib().pp_sim_info(self._h, ctypes.byref(a),
                                       ctypes.byref(b), ctypes.byref(c),
                                       ctypes.byref(d)), "info")
        return {"sort_passes": a.value, "sort_tile": b.value,
                "action_tile": c.value, "grid_move": d.value}

    def state(self):
        """(cell u32[n], energy i32[n], type u8[n], grass u32[C])."""
        n = self.num_agents
        # refresh n after any pending work
        self.sync()
        n = self.num_agents
        cell = np.empty(n, np.uint32)
        energy = np.empty(n, np.int32)
        typ = np.empty(n, np.uint8)
        grass = np.empty(self.params.grid_xy, np.uint32)
        self._check(_lib().pp_sim_get_state(
            self._h, cell.ctypes.data_as(ctypes.c_void_p),
            energy.ctypes.data_as(ctypes.c_void_p),
            typ.ctypes.data_as(ctypes.c_void_p),
            grass.ctypes.data_as(ctypes.c_void_p)), "get_state")
        return cell, energy, typ, grass

    def sync(self):
        self._check(_lib().pp_sim_sync(self._h), "sync")

    def set_sort(self, mode: int):
        """0: radix sort every iteration; 8..11: row buckets of 2**mode
        cells; 1: automatic (the default). Results do not depend on it."""
        self._check(_lib().pp_sim_set_sort(self._h, mode), "set_sort")

    # -- profiling ---------------------------------------------------------
    def set_profiling(self, enable: bool):
        self._check(_lib().pp_sim_set_profiling(self._h, int(enable)),
                    "set_profiling")

    def profile(self) -> dict:
        """{kernel: (launches, total_ns)}"""
        out = {}
        for i in range(_lib().pp_sim_prof_count(self._h)):
            name = ctypes.c_char_p()
            launches = ctypes.c_uint64()
            ns = ctypes.c_double()
            self._check(_lib().pp_sim_prof_get(
                self._h, i, ctypes.byref(name), ctypes.byref(launches),
                ctypes.byref(ns)), "prof_get")
            out[name.value.decode()] = (launches.value, ns.value)
        return out

    def profile_reset(self):
        self._check(_lib().pp_sim_prof_reset(self._h), "prof_reset")

    def profile_export(self, path: str):
        self._check(_lib().pp_sim_prof_export(self._h, path.encode()),
                    "prof_export")

    def dump(self, agents_path=None, cells_path=None, dump_type=0x01,
             append=False):
        """Appends the text dump of the current state (pp_sim_dump)."""
        enc = (lambda x: x.encode() if x is not None else None)
        self._check(_lib().pp_sim_dump(self._h, enc(agents_path),
                                       enc(cells_path), dump_type,
                                       int(append)), "dump")


# Replica states (PP_REPLICA_* of pphpc.h).
RUNNING, DONE, FAILED = 1, 2, 3


class Batch:
    """Replicas of different grids, parameters, seeds and iteration counts
    advancing together on one GPU (a ``pp_batch_t`` handle). A replica's
    statistics rows are those of the same simulation run alone."""

    def __init__(self, agent_capacity: int, cell_capacity: int,
                 max_replicas: int, max_grass_restart: int, device: int = -1):
        h = ctypes.c_void_p()
        rc = _lib().pp_batch_create(agent_capacity, cell_capacity,
                                    max_replicas, max_grass_restart, device,
                                    ctypes.byref(h))
        if rc:
            raise PPError(rc, "pp_batch_create")
        self._h = h
        self.params = {}

    @staticmethod
    def footprint(agent_capacity: int, cell_capacity: int, max_replicas: int,
                  max_grass_restart: int) -> int:
        """Device bytes a batch with these limits allocates."""
        return _lib().pp_batch_footprint(agent_capacity, cell_capacity,
                                         max_replicas, max_grass_restart)

    def close(self):
        if getattr(self, "_h", None):
            _lib().pp_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc, what):
        if rc:
            raise PPError(rc, what)

    def add(self, params: Params, seed: int, reserve: int = 0):
        """Admit a replica; returns its id, or None when the batch has no
        room for it now (PP_OUT_OF_RESOURCES)."""
        rid = ctypes.c_uint32()
        rc = _lib().pp_batch_add(self._h, ctypes.byref(params), seed, reserve,
                                 ctypes.byref(rid))
        if rc == PP_OUT_OF_RESOURCES:
            return None
        self._check(rc, "add")
        self.params[rid.value] = params
        return rid.value

    def run(self, ticks: int):
        self._check(_lib().pp_batch_run(self._h, ticks), f"run({ticks})")

    def replica(self, rid: int) -> dict:
        st, steps, err, ast = (ctypes.c_int(), ctypes.c_uint32(),
                               ctypes.c_int(), ctypes.c_uint64())
        self._check(_lib().pp_batch_replica(
            self._h, rid, ctypes.byref(st), ctypes.byref(steps),
            ctypes.byref(err), ctypes.byref(ast)), "replica")
        return {"state": st.value, "steps": steps.value, "error": err.value,
                "agent_steps": ast.value}

    def rows(self, rid: int, first: int = 0, n: Optional[int] = None):
        if n is None:
            n = self.replica(rid)["steps"] + 1 - first
        out = (Stats * n)()
        self._check(_lib().pp_batch_rows(self._h, rid, first, n, out), "rows")
        return list(out)

    def rows_array(self, rid: int) -> np.ndarray:
        return stats_rows_to_array(self.rows(rid))

    def state(self, rid: int):
        """(cell, energy, type, grass) of a running replica, cells local to
        its grid."""
        p = self.params[rid]
        n = ctypes.c_uint32()
        self._check(_lib().pp_batch_get_state(self._h, rid, ctypes.byref(n),
                                              None, None, None, None), "state")
        cell = np.empty(n.value, np.uint32)
        energy = np.empty(n.value, np.int32)
        typ = np.empty(n.value, np.uint8)
        grass = np.empty(p.grid_x * p.grid_y, np.uint32)
        self._check(_lib().pp_batch_get_state(
            self._h, rid, ctypes.byref(n),
            cell.ctypes.data_as(ctypes.c_void_p),
            energy.ctypes.data_as(ctypes.c_void_p),
            typ.ctypes.data_as(ctypes.c_void_p),
            grass.ctypes.data_as(ctypes.c_void_p)), "state")
        return cell, energy, typ, grass

    def release(self, rid: int):
        self._check(_lib().pp_batch_release(self._h, rid), "release")
        self.params.pop(rid, None)

    def resources(self) -> dict:
        a, c, s, n = (ctypes.c_uint64(), ctypes.c_uint32(), ctypes.c_uint32(),
                      ctypes.c_uint32())
        self._check(_lib().pp_batch_resources(
            self._h, ctypes.byref(a), ctypes.byref(c), ctypes.byref(s),
            ctypes.byref(n)), "resources")
        return {"agents_free": a.value, "cells_free_run": c.value,
                "slots_free": s.value, "active": n.value}

    @property
    def tick(self) -> int:
        return _lib().pp_batch_tick(self._h)

    def counters(self) -> dict:
        out = (ctypes.c_uint64 * 6)()
        self._check(_lib().pp_batch_counters(self._h, out), "counters")
        return dict(zip(COUNTER_FIELDS, tuple(out)))

    def info(self) -> dict:
        a, b, c, d = (ctypes.c_uint32() for _ in range(4))
        self._check(_lib().pp_batch_info(self._h, ctypes.byref(a),
                                         ctypes.byref(b), ctypes.byref(c),
                                         ctypes.byref(d)), "info")
        return {"sort_passes": a.value, "sort_tile": b.value,
                "action_tile": c.value, "grid_move": d.value}

    def set_shape(self, keys_per_thread: int, max_digit_bits: int):
        self._check(_lib().pp_batch_set_shape(self._h, keys_per_thread,
                                              max_digit_bits), "set_shape")

    def set_sort(self, mode: int):
        """0: radix sort every iteration; 8..11: row buckets of 2**mode
        cells; 1: automatic (the default)."""
        self._check(_lib().pp_batch_set_sort(self._h, mode), "set_sort")

    def tune(self, trial_ticks: int = 32) -> dict:
        s, k, b, t = (ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32(),
                      ctypes.c_double())
        self._check(_lib().pp_batch_tune(self._h, trial_ticks, ctypes.byref(s),
                                         ctypes.byref(k), ctypes.byref(b),
                                         ctypes.byref(t)), "tune")
        out = {"bucket_shift": s.value, "keys_per_thread": k.value,
               "max_digit_bits": b.value, "sec_per_tick": t.value}
        out.update(self.launch())
        return out

    def set_launch(self, tile_rounds: int, bucket_eighths: int):
        self._check(_lib().pp_batch_set_launch(self._h, tile_rounds,
                                               bucket_eighths), "set_launch")

    def launch(self) -> dict:
        r, e = ctypes.c_uint32(), ctypes.c_uint32()
        self._check(_lib().pp_batch_get_launch(self._h, ctypes.byref(r),
                                               ctypes.byref(e)), "launch")
        return {"tile_rounds": r.value, "bucket_eighths": e.value}

    def set_profiling(self, enable: bool):
        self._check(_lib().pp_batch_set_profiling(self._h, int(enable)),
                    "set_profiling")

    def profile(self) -> dict:
        out = {}
        for i in range(_lib().pp_batch_prof_count(self._h)):
            name = ctypes.c_char_p()
            launches = ctypes.c_uint64()
            ns = ctypes.c_double()
            self._check(_lib().pp_batch_prof_get(
                self._h, i, ctypes.byref(name), ctypes.byref(launches),
                ctypes.byref(ns)), "prof_get")
            out[name.value.decode()] = (launches.value, ns.value)
        return out

    def sync(self):
        self._check(_lib().pp_batch_sync(self._h), "sync")
