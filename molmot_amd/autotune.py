"""Launch-shape autotuner for the gfx950 kernels.

Some kernels have more than one valid launch shape and the fastest one
depends on the problem size and on the device:

* the Lindblad VALU propagator: lanes per instance (64, 128 or 256; one,
  two or four wavefronts share one density matrix and each lane holds fewer
  of its N^2 elements, trading registers for barriers);
* the rate-equation map / point kernels: threads per block (64, 128, 256;
  the per-thread LDS slab limits the largest block).

``Tuner.pick`` times every candidate on the caller's own workload (HIP events
on the current stream, after one warm-up launch), keeps the fastest, and
caches the choice in a JSON file keyed by device, kernel family and problem
shape, so later processes reuse it without timing.  The cache path is
``$MOLMOT_AMD_TUNE_CACHE`` or ``~/.cache/molmot_amd/tune.json``; ``None``
disables the file.  The reference has no counterpart: molmot_cuda fixes its
launch shapes (``block_size`` arguments, molmot_cuda/rate_equations_cuda.py:448-544).
"""
from __future__ import annotations

import json
import os
import threading
from typing import Callable, Dict, Iterable, Optional

import torch

_DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "molmot_amd", "tune.json")


def _device_key() -> str:
    p = torch.cuda.get_device_properties(torch.cuda.current_device())
    arch = getattr(p, "gcnArchName", "") or ""
    return f"{p.name}|{arch}|{p.multi_processor_count}"


class Tuner:
    """Times candidate launch shapes and remembers the winner per key."""

    def __init__(self, cache_path: Optional[str] = "default"):
        if cache_path == "default":
            cache_path = os.environ.get("MOLMOT_AMD_TUNE_CACHE", _DEFAULT_CACHE)
        self.cache_path = cache_path
        self._lock = threading.Lock()
        self._table: Dict[str, dict] = {}
        if cache_path and os.path.exists(cache_path):
            try:
                with open(cache_path) as f:
                    self._table = json.load(f)
            except (OSError, ValueError):
                self._table = {}

    def _save(self) -> None:
        if not self.cache_path:
            return
        os.makedirs(os.path.dirname(self.cache_path) or ".", exist_ok=True)
        tmp = self.cache_path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self._table, f, indent=1, sort_keys=True)
        os.replace(tmp, self.cache_path)

    def lookup(self, key: str):
        e = self._table.get(f"{_device_key()}|{key}")
        return None if e is None else e["best"]

    def pick(self, key: str, candidates: Iterable, apply: Callable[[object], None], run: Callable[[], None],
             repeats: int = 3, retune: bool = False, tie: float = 0.03):
        """Apply each candidate, time ``run`` (median of ``repeats`` launches
        after a warm-up) and leave the fastest applied; a candidate listed
        earlier wins over a later one that is faster by less than ``tie``
        (relative), so list the default first.  Returns (best, {candidate: ms})."""
        full = f"{_device_key()}|{key}"
        cands = list(candidates)
        if not retune and full in self._table and self._table[full]["best"] in cands:
            best = self._table[full]["best"]
            apply(best)
            return best, {k: v for k, v in self._table[full]["ms"].items()}
        stream = torch.cuda.current_stream()
        times = {}
        for c in cands:
            apply(c)
            run()
            ts = []
            for _ in range(max(1, repeats)):
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record(stream)
                run()
                b.record(stream)
                b.synchronize()
                ts.append(a.elapsed_time(b))
            ts.sort()
            times[c] = ts[len(ts) // 2]
        fastest = min(times.values())
        best = next(c for c in cands if times[c] <= fastest * (1.0 + tie))
        apply(best)
        with self._lock:
            self._table[full] = {"best": best, "ms": {str(k): v for k, v in times.items()}}
            self._save()
        return best, times


_TUNER: Optional[Tuner] = None


def default_tuner() -> Tuner:
    global _TUNER
    if _TUNER is None:
        _TUNER = Tuner()
    return _TUNER


def tune_lindblad_lanes(cls, precision: str, kv: torch.Tensor, phi0: torch.Tensor, dt: float = 0.05,
                        n_steps: int = 20, tuner: Optional[Tuner] = None, retune: bool = False):
    """Pick lanes per instance for the VALU Lindblad propagator of ``cls`` at
    ``precision`` ("fp64" or "fp32") on a sample of the caller's instances
    (all of kv/phi0; a short run of n_steps).  Leaves the choice set on the
    class; returns (lanes, {lanes: ms})."""
    from .obe.lindblad import lindblad_rk4_device
    if precision not in ("fp64", "fp32"):
        raise ValueError("the VALU propagator takes precision 'fp64' or 'fp32'")
    t = tuner or default_tuner()
    key = f"lindblad_valu|{cls.n_ground}+{cls.n_excited}|{precision}|n={kv.numel()}"
    cands = (64, 128, 256) if cls.N * cls.N <= 256 else (256, 128, 64)   # the default first
    return t.pick(key, cands, lambda c: cls.set_lanes(precision, c),
                  lambda: lindblad_rk4_device(cls, kv, phi0, dt, n_steps, n_steps // 2, precision), retune=retune)


def tune_rate_block(plan, v: torch.Tensor, z: torch.Tensor, B_gradient: float, tuner: Optional[Tuner] = None,
                    retune: bool = False):
    """Pick threads per block for the rate-equation map kernel of ``plan``
    (a HIPMolData) on the caller's (v, z) grid.  Returns (block, {block: ms})."""
    from .hip.rate_equations import rate_eq_map_device
    t = tuner or default_tuner()
    cands = [b for b in (256, 128, 64) if plan.block_ok(b)]
    key = f"rate_map|{plan.n_ground}+{plan.n_excited}|n={v.numel()}x{z.numel()}"
    return t.pick(key, cands, plan.set_block, lambda: rate_eq_map_device(plan, v, z, B_gradient), retune=retune)
