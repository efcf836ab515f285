"""Launch-schedule autotuner for the HIP kernels.

The library exposes a small table of schedule knobs (srf_tune_set / srf_tune_get in
include/srf_hip.h): the point chunk of the weight-gradient GEMMs (NeRF MLP and TensoRF colour
network), the grid cap of the factor gathers and the run length of the gather backwards.  None
changes a result beyond the order of FP32 atomic additions.  The best value depends on the
workload's shape (point counts, number of layer jobs, how many consecutive samples share texel
corners), so it is measured, not modelled: :meth:`Autotuner.tune` times the caller's own step
function under each candidate with HIP events on the current stream and keeps the fastest.

Results are kept per workload key in memory and, when the tuner is given a path, in a JSON cache,
so a later process applies them without measuring again (:meth:`Autotuner.apply`).
"""
from __future__ import annotations

import json
from pathlib import Path

from . import _lib

KNOBS = {'dw_chunk': 0, 'color_dw_chunk': 1, 'gather_blocks': 2, 'gather_run': 3}

# candidate values per knob (0 = the library's built-in choice)
CANDIDATES = {
    'dw_chunk': [0, 2048, 4096, 8192, 16384, 32768],
    'color_dw_chunk': [0, 2304, 3072, 4608, 6144, 8192, 16384],
    'gather_blocks': [0, 2048, 4096, 16384, 32768],
    'gather_run': [0, 8, 16, 64, 128],
}


def set_knob(name: str, value: int) -> None:
    """Sets a schedule knob (0 restores the built-in choice); raises SrfError out of range."""
    _lib.call('srf_tune_set', KNOBS[name], int(value))


def get_knob(name: str) -> int:
    return int(_lib.lib().srf_tune_get(KNOBS[name]))


def reset() -> None:
    for name in KNOBS:
        set_knob(name, 0)


class Autotuner:
    """Measures and applies schedule knobs per workload key (e.g. 'nerf_train_4096rays')."""

    def __init__(self, cache_path: str | Path | None = None):
        self.path = Path(cache_path) if cache_path else None
        self.results: dict[str, dict] = {}
        if self.path and self.path.exists():
            self.results = json.loads(self.path.read_text())

    def _save(self) -> None:
        if self.path:
            self.path.parent.mkdir(parents=True, exist_ok=True)
            self.path.write_text(json.dumps(self.results, indent=1, sort_keys=True) + '\n')

    def apply(self, workload: str) -> bool:
        """Sets every knob recorded for `workload`; False if nothing is recorded."""
        rec = self.results.get(workload)
        if not rec:
            return False
        for name, r in rec.items():
            set_knob(name, r['chosen'])
        return True

    def tune_param(self, workload: str, name: str, candidates, run, reps: int = 2, force: bool = False) -> dict:
        """Host-level schedule parameter (e.g. the ray chunk of a full-image render): times
        `run(value)` for each candidate (one untimed call, then `reps` calls between HIP events)
        and records the fastest; returns {'chosen': v, 'ms': {v: ms per call}}."""
        import torch
        rec = self.results.setdefault(workload, {})
        if name in rec and not force:
            return rec[name]
        times = {}
        for v in candidates:
            run(v)
            start = torch.cuda.Event(enable_timing=True)
            end = torch.cuda.Event(enable_timing=True)
            start.record()
            for _ in range(reps):
                run(v)
            end.record()
            end.synchronize()
            times[v] = start.elapsed_time(end) / reps
        chosen = min(times, key=times.get)
        rec[name] = {'chosen': chosen, 'ms': {str(k): round(t, 3) for k, t in times.items()}}
        self._save()
        return rec[name]

    def tune(self, workload: str, knob: str, step, candidates=None, reps: int = 3, force: bool = False) -> dict:
        """Times `step()` (one full unit of the workload, launched on the current stream) under each
        candidate value of `knob`: one untimed call, then `reps` calls between two HIP events.  The
        fastest value is set and recorded; returns {'chosen': v, 'ms': {v: ms per call}}."""
        import torch
        rec = self.results.setdefault(workload, {})
        if knob in rec and not force:
            set_knob(knob, rec[knob]['chosen'])
            return rec[knob]
        cands = list(candidates if candidates is not None else CANDIDATES[knob])
        previous = get_knob(knob)
        times = {}
        try:
            for v in cands:
                set_knob(knob, v)
                step()
                start = torch.cuda.Event(enable_timing=True)
                end = torch.cuda.Event(enable_timing=True)
                start.record()
                for _ in range(reps):
                    step()
                end.record()
                end.synchronize()
                times[v] = start.elapsed_time(end) / reps
        except Exception:
            set_knob(knob, previous)
            raise
        chosen = min(times, key=times.get)
        set_knob(knob, chosen)
        rec[knob] = {'chosen': chosen, 'ms': {str(k): round(t, 3) for k, t in times.items()}}
        self._save()
        return rec[knob]
