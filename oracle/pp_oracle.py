"""ctypes wrapper of the CPU oracle (``oracle/build/libpp_oracle.so``).

TEST INFRASTRUCTURE ONLY: imported by the tests, the smoke check and the
CPU-baseline leg of the benchmark, always as the checker, never as the thing
measured for the GPU path. See ``pp_oracle.c`` for the reference file:line each
function restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(_HERE, "build")
_LIB = None

PARAM_FIELDS = (
    "init_sheep", "sheep_gain_from_food", "sheep_reproduce_threshold",
    "sheep_reproduce_prob", "init_wolves", "wolves_gain_from_food",
    "wolves_reproduce_threshold", "wolves_reproduce_prob", "grass_restart",
    "grid_x", "grid_y", "grid_xy", "iters",
)
STAT_FIELDS = ("sheep", "wolves", "grass", "sheep_en", "wolves_en",
               "grass_en", "errors")


class OParams(ctypes.Structure):
    _fields_ = [(f, ctypes.c_uint32) for f in PARAM_FIELDS]


class OStats(ctypes.Structure):
    _fields_ = [(f, ctypes.c_uint64) for f in STAT_FIELDS]


def build():
    """Build the oracle with its Makefile (gcc; hipcc host-only for the
    rocRAND cross-check program)."""
    subprocess.run(["make", "-C", _HERE], check=True,
                   stdout=subprocess.DEVNULL)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(BUILD, "libpp_oracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P = ctypes.POINTER
        u32, u64, i32, vp = (ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int,
                             ctypes.c_void_p)
        sig = {
            "ppo_philox4x32_10": (None, [P(u32), P(u32), P(u32)]),
            "ppo_load_params": (i32, [ctypes.c_char_p, P(OParams)]),
            "ppo_stats_save": (i32, [ctypes.c_char_p, P(OStats), u32, u32]),
            "ppo_create": (vp, [P(OParams), u64, u64]),
            "ppo_destroy": (None, [vp]),
            "ppo_stats_get": (None, [vp, P(OStats)]),
            "ppo_step": (i32, [vp]),
            "ppo_run": (i32, [vp, u32, P(OStats)]),
            "ppo_step_count": (u32, [vp]),
            "ppo_num_agents": (u64, [vp]),
            "ppo_state_get": (None, [vp, vp, vp, vp, vp]),
            "ppo_state_set": (i32, [vp, u64, vp, vp, vp, vp]),
            "ppo_last_events": (None, [vp, P(u64)]),
            "ppo_agent_steps": (u64, [vp]),
            "ppo_set_fault": (None, [vp, i32]),
            "ppo_dump": (i32, [vp, ctypes.c_char_p, ctypes.c_char_p, u32,
                               i32]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().ppo_philox4x32_10(c, k, o)
    return tuple(o)


def params_from(p) -> OParams:
    """OParams from any object with the PPParameters field names."""
    o = OParams()
    for f in PARAM_FIELDS:
        if f == "grid_xy":
            continue
        setattr(o, f, int(getattr(p, f) if not isinstance(p, dict) else p[f]))
    o.grid_xy = o.grid_x * o.grid_y
    return o


def load_params(path: str) -> tuple[int, OParams]:
    p = OParams()
    rc = lib().ppo_load_params(path.encode(), ctypes.byref(p))
    return rc, p


def stats_tuple(s: OStats):
    return tuple(getattr(s, f) for f in STAT_FIELDS)


class Oracle:
    def __init__(self, params, seed=0, capacity=0):
        self.params = params_from(params)
        n0 = self.params.init_sheep + self.params.init_wolves
        if capacity == 0:
            capacity = max(4 * n0, 16777216)
        self.capacity = capacity
        self._h = lib().ppo_create(ctypes.byref(self.params), seed, capacity)
        if not self._h:
            raise ValueError("oracle rejected the parameters")

    def close(self):
        if getattr(self, "_h", None):
            lib().ppo_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def stats(self):
        s = OStats()
        lib().ppo_stats_get(self._h, ctypes.byref(s))
        return stats_tuple(s)

    def step(self) -> int:
        return lib().ppo_step(self._h)

    def run(self, iters):
        """(rc, uint64 array of rows 1..done)."""
        rows = (OStats * max(iters, 1))()
        rc = lib().ppo_run(self._h, iters, rows)
        done = self.step_count
        return rc, np.array([stats_tuple(r) for r in rows[:iters]],
                            dtype=np.uint64).reshape(-1, 7)

    @property
    def step_count(self):
        return lib().ppo_step_count(self._h)

    @property
    def num_agents(self):
        return lib().ppo_num_agents(self._h)

    @property
    def agent_steps(self):
        return lib().ppo_agent_steps(self._h)

    def state(self):
        n = self.num_agents
        cell = np.empty(n, np.uint32)
        energy = np.empty(n, np.int32)
        typ = np.empty(n, np.uint8)
        grass = np.empty(self.params.grid_xy, np.uint32)
        vp = ctypes.c_void_p
        lib().ppo_state_get(self._h, cell.ctypes.data_as(vp),
                            energy.ctypes.data_as(vp), typ.ctypes.data_as(vp),
                            grass.ctypes.data_as(vp))
        return cell, energy, typ, grass

    def set_state(self, cell, energy, typ, grass):
        cell = np.ascontiguousarray(cell, np.uint32)
        energy = np.ascontiguousarray(energy, np.int32)
        typ = np.ascontiguousarray(typ, np.uint8)
        grass = np.ascontiguousarray(grass, np.uint32)
        vp = ctypes.c_void_p
        rc = lib().ppo_state_set(self._h, len(cell), cell.ctypes.data_as(vp),
                                 energy.ctypes.data_as(vp),
                                 typ.ctypes.data_as(vp),
                                 grass.ctypes.data_as(vp))
        if rc:
            raise ValueError("state above capacity")

    def last_events(self):
        out = (ctypes.c_uint64 * 5)()
        lib().ppo_last_events(self._h, out)
        return dict(zip(("starved", "grass_eaten", "sheep_eaten",
                         "sheep_births", "wolf_births"), tuple(out)))

    def set_fault(self, mode: int):
        lib().ppo_set_fault(self._h, mode)

    def dump(self, agents_path, cells_path, dump_type=0x01, append=False):
        enc = (lambda x: x.encode() if x is not None else None)
        return lib().ppo_dump(self._h, enc(agents_path), enc(cells_path),
                              dump_type, int(append))


def simulate(params, seed=0, capacity=0, iters=None):
    """rows 0..iters as an (iters+1, 7) uint64 array (oracle)."""
    o = Oracle(params, seed, capacity)
    it = o.params.iters if iters is None else iters
    row0 = np.array([o.stats()], dtype=np.uint64)
    rc, rows = o.run(it)
    o.close()
    if rc:
        raise RuntimeError(f"oracle stopped with {rc}")
    return np.vstack([row0, rows])


# ------------------------------------------------------------------------
# Reference side of the statistical tests (oracle/pp_statref.c)

_SREF = None


def statref_lib():
    global _SREF
    if _SREF is None:
        path = os.path.join(BUILD, "libpp_statref.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.ppr_run.restype = ctypes.c_int
        L.ppr_run.argtypes = [ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32,
                              ctypes.c_uint32, ctypes.c_void_p]
        L.ppr_run_cells.restype = ctypes.c_int
        L.ppr_run_cells.argtypes = [ctypes.POINTER(ctypes.c_uint32),
                                    ctypes.c_uint32, ctypes.c_uint32,
                                    ctypes.c_void_p]
        L.ppr_mt_outputs.restype = None
        L.ppr_mt_outputs.argtypes = [ctypes.c_uint32, ctypes.c_uint32,
                                     ctypes.c_void_p]
        _SREF = L
    return _SREF


def statref_run(params, seed, iters=None, rules="cells"):
    """(iters+1, 7) uint64 rows of one replica of the independent simulator
    (MT19937): rules "cells" = the Java single-thread semantics (per-cell
    shuffled acting order, a wolf takes the first live sheep of the list),
    "netlogo" = the NetLogo rules (global random order, a wolf takes a
    uniformly random sheep)."""
    p = params_from(params)
    it = p.iters if iters is None else iters
    arr = (ctypes.c_uint32 * 11)(p.init_sheep, p.sheep_gain_from_food,
                                 p.sheep_reproduce_threshold,
                                 p.sheep_reproduce_prob, p.init_wolves,
                                 p.wolves_gain_from_food,
                                 p.wolves_reproduce_threshold,
                                 p.wolves_reproduce_prob, p.grass_restart,
                                 p.grid_x, p.grid_y)
    rows = np.zeros((it + 1, 7), np.uint64)
    fn = {"cells": statref_lib().ppr_run_cells,
          "netlogo": statref_lib().ppr_run}[rules]
    rc = fn(arr, seed, it, rows.ctypes.data_as(ctypes.c_void_p))
    if rc:
        raise ValueError("statref rejected the parameters")
    return rows


def mt_outputs(seed, n):
    out = np.zeros(n, np.uint32)
    statref_lib().ppr_mt_outputs(seed, n, out.ctypes.data_as(ctypes.c_void_p))
    return out
