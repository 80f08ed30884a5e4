"""Autotuner for launch shapes (no counterpart in the reference, which has one launch shape).

It times candidate launches of the library on the device, on the caller's own batches (outputs are
accumulated by every trial, so the caller resets them afterwards; `DeviceBatch.reset_output`), and keeps
the fastest choice in a cache keyed by (GPU architecture, element type, d, n, batch-size bucket). The cache
is process-wide, guarded by a lock, and optionally persisted as JSON (path argument or the environment
variable KRONMULT_TUNE_CACHE).

Choices:
  * one batch: the kernel family (register-resident sweep where a plan exists, matrix-core sweep for
    d=6 n=4 FP64, packed fiber sweep for small vectors, fiber sweep) and, for the register, matrix-core and
    packed families, workgroups per CU;
  * several batches (a grouped call): the persistent launch -- workgroups per CU and per-group tile length
    (entries per tile, from a quarter of the entries the launch processes in one pass, which also shrinks
    the pass and its LDS, to four passes) -- against one launch per group with each group's own best
    shape.
"""
from __future__ import annotations

import json
import math
import os
import threading
from typing import Dict, List, Optional

from . import api

_LOCK = threading.Lock()
_CACHE: Dict[str, dict] = {}
_LOADED: set = set()

LDS_BUDGET = 64 * 1024      # the persistent launch's LDS budget per workgroup (kron_group.hpp host side)
GROUP_THREADS = 256


def _torch():
    import torch
    return torch


def bucket(nb: int) -> int:
    """Batch sizes are tuned per power-of-two bucket."""
    return 1 << max(0, math.ceil(math.log2(max(1, nb))))


def arch(device) -> str:
    torch = _torch()
    props = torch.cuda.get_device_properties(device)
    return getattr(props, "gcnArchName", props.name).split(":")[0]


def key_single(arch_name: str, dtype: str, d: int, n: int, nb: int) -> str:
    return f"{arch_name}|{dtype}|d{d}n{n}|nb{bucket(nb)}"


def entries_per_pass(d: int, n: int, itemsize: int) -> int:
    """Entries the persistent launch processes at once for (d, n) (same rule as the library)."""
    nfib = n ** (d - 1)
    size = n ** d
    padded = size + size // n if n in (4, 8) else size   # kron_group.hpp: padded rows, 16-byte factor pad
    per_entry = (padded + d * n * n + 2) * itemsize + (d + 1) * 8   # + the pass's pointer slots
    e = max(1, GROUP_THREADS // nfib)
    while e > 1 and e * per_entry > LDS_BUDGET:
        e -= 1
    return e


class Autotuner:
    def __init__(self, cache_path: Optional[str] = None, reps: int = 5):
        self.path = cache_path or os.environ.get("KRONMULT_TUNE_CACHE")
        self.reps = reps
        if self.path:
            self._load(self.path)

    # ---- cache -----------------------------------------------------------------------------------------
    @staticmethod
    def _load(path: str) -> None:
        with _LOCK:
            if path in _LOADED or not os.path.exists(path):
                _LOADED.add(path)
                return
            with open(path) as f:
                _CACHE.update(json.load(f))
            _LOADED.add(path)

    def _store(self, key: str, value: dict) -> None:
        with _LOCK:
            _CACHE[key] = value
            if self.path:
                tmp = self.path + f".tmp{os.getpid()}"
                os.makedirs(os.path.dirname(os.path.abspath(self.path)), exist_ok=True)
                with open(tmp, "w") as f:
                    json.dump(_CACHE, f, indent=1, sort_keys=True)
                os.replace(tmp, self.path)

    @staticmethod
    def lookup(key: str) -> Optional[dict]:
        with _LOCK:
            v = _CACHE.get(key)
            return None if v is None else dict(v)

    # ---- timing ----------------------------------------------------------------------------------------
    def _time_us(self, fn) -> float:
        """Device time of one call of fn(stream): `reps` calls captured in one graph on a side stream, the
        best of three replays, so the host's launch cost (a Python call per launch) is not timed."""
        torch = _torch()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            fn(s)
        s.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            for _ in range(self.reps):
                fn(s)
        best = float("inf")
        with torch.cuda.stream(s):
            for _ in range(3):
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record(s)
                g.replay()
                b.record(s)
                b.synchronize()
                best = min(best, a.elapsed_time(b) * 1e3 / self.reps)
        return best

    # ---- one batch -------------------------------------------------------------------------------------
    def tune_single(self, db, force: bool = False) -> dict:
        """Best kronmult_tuning_t (as a dict) for one DeviceBatch; cached."""
        key = key_single(arch(db.device), db.dtype_name, db.d, db.n, db.nb)
        hit = None if force else self.lookup(key)
        if hit is not None:
            return hit
        cands = [{"family": api.FAMILY_FIBER}]
        try:
            api.launch_info(db.d, db.n, db.itemsize, db.nb, {"family": api.FAMILY_REGISTER})
            cands += [{"family": api.FAMILY_REGISTER, "workgroups_per_cu": w} for w in (1, 2, 4, 8, 16)]
        except api.KronmultError:
            pass
        try:
            api.launch_info(db.d, db.n, db.itemsize, db.nb, {"family": api.FAMILY_PACKED})
            cands += [{"family": api.FAMILY_PACKED, "workgroups_per_cu": w} for w in (4, 8, 16)]
        except api.KronmultError:
            pass
        try:
            api.launch_info(db.d, db.n, db.itemsize, db.nb, {"family": api.FAMILY_WAVE})
            cands += [{"family": api.FAMILY_WAVE}] + [{"family": api.FAMILY_WAVE, "entries_per_workgroup": e}
                                                      for e in (1, 2, 4)]
        except api.KronmultError:
            pass
        try:
            api.launch_info(db.d, db.n, db.itemsize, db.nb, {"family": api.FAMILY_MATRIX})
            cands += [{"family": api.FAMILY_MATRIX, "workgroups_per_cu": w} for w in (2, 3, 4)]
        except api.KronmultError:
            pass
        best, best_t = None, float("inf")
        for c in cands:
            t = self._time_us(lambda st: db.run(sync=False, tuning=c, stream=st))
            if t < best_t:
                best, best_t = c, t
        choice = dict(best, us=round(best_t, 2))
        self._store(key, choice)
        return choice

    # ---- grouped call ----------------------------------------------------------------------------------
    def tune_grouped(self, dbs: List, force: bool = False) -> dict:
        """Best strategy for a grouped call over DeviceBatches `dbs`: {"strategy", "workgroups_per_cu",
        "entries_per_tile": [...], "tunings": [...], "us"}; cached."""
        a = arch(dbs[0].device)
        key = "grouped|" + "+".join(key_single(a, db.dtype_name, db.d, db.n, db.nb) for db in dbs)
        hit = None if force else self.lookup(key)
        if hit is not None:
            return hit
        dt = dbs[0].dtype_name
        # one launch per group, each with its own best shape
        tunings = [self.tune_single(db) for db in dbs]
        tun_clean = [{k: v for k, v in t.items() if k != "us"} for t in tunings]
        groups = [db.group(tuning=t) for db, t in zip(dbs, tun_clean)]
        structs = api.grouped_structs(groups, dt)
        best = {"strategy": "separate", "workgroups_per_cu": 0, "entries_per_tile": [0] * len(dbs),
                "tunings": tun_clean}
        best_t = self._time_us(lambda st: api.kronmult_batched_grouped(groups, dtype=dt, strategy="separate",
                                                                       stream=st, structs=structs))
        # persistent launch: grid and tile length, the same multiple of each group's pass for all groups ...
        mults = (0.25, 0.5, 1, 2, 4)
        tile = lambda db, m: max(1, int(entries_per_pass(db.d, db.n, db.itemsize) * m))

        def time_persistent(wpc, epw):
            groups = [db.group(tuning=t, entries_per_tile=e) for db, t, e in zip(dbs, tun_clean, epw)]
            structs = api.grouped_structs(groups, dt)
            return self._time_us(lambda st: api.kronmult_batched_grouped(
                groups, dtype=dt, strategy="persistent", workgroups_per_cu=wpc, stream=st, structs=structs))

        best_p, best_pt = None, float("inf")
        for wpc in (4, 8, 16, 32):
            for mult in mults:
                epw = [tile(db, mult) for db in dbs]
                t = time_persistent(wpc, epw)
                if t < best_pt:
                    best_pt, best_p = t, (wpc, list(epw))
        # ... then each group's tile on its own, the others held (coordinate descent, at most two passes
        # over the groups): the groups share the grid but not their tiles, and a short tile suits a group
        # whose pass holds many entries
        wpc, epw = best_p
        for _ in range(2):
            moved = False
            for g, db in enumerate(dbs):
                for mult in (0.125,) + mults:
                    trial = list(epw)
                    trial[g] = tile(db, mult)
                    if trial[g] == epw[g]:
                        continue
                    t = time_persistent(wpc, trial)
                    if t < best_pt:
                        best_pt, epw, moved = t, trial, True
            if not moved:
                break
        if best_pt < best_t:
            best_t = best_pt
            best = {"strategy": "persistent", "workgroups_per_cu": wpc, "entries_per_tile": epw,
                    "tunings": tun_clean}
        choice = dict(best, us=round(best_t, 2))
        self._store(key, choice)
        return choice


def grouped_call_from_choice(dbs, choice: dict):
    """(groups, structs, kwargs) for api.kronmult_batched_grouped from a tune_grouped() choice."""
    groups = [db.group(tuning=t, entries_per_tile=e)
              for db, t, e in zip(dbs, choice["tunings"], choice["entries_per_tile"])]
    structs = api.grouped_structs(groups, dbs[0].dtype_name)
    return groups, structs, {"strategy": choice["strategy"], "workgroups_per_cu": choice["workgroups_per_cu"]}
