"""Launch-shape autotuner.

For a shape bucket (a hashable key describing the work: kernel, pattern size,
tap radius, spots-per-pattern scale, ...) it times each candidate
configuration on the current device with HIP events and keeps the fastest.
Results are cached per device in memory and, when ``DIFFSIMS_AMD_TUNE_CACHE``
names a file, in that JSON file so later processes skip the search.
"""

from __future__ import annotations

import json
import os
import threading
from typing import Callable, Hashable, Optional, Sequence

import torch

__all__ = ["Autotuner"]


class Autotuner:
    def __init__(self, reps: int = 3, cache_file: Optional[str] = None):
        self.reps = int(reps)
        self.cache_file = cache_file if cache_file is not None else os.environ.get(
            "DIFFSIMS_AMD_TUNE_CACHE")
        self._lock = threading.Lock()
        self._cache = {}
        self.log = []          # (key, candidate, ms) of every timing, for reports
        if self.cache_file and os.path.exists(self.cache_file):
            try:
                with open(self.cache_file) as f:
                    self._cache = {k: tuple(v) if isinstance(v, list) else v
                                   for k, v in json.load(f).items()}
            except (OSError, ValueError):
                self._cache = {}

    @staticmethod
    def _key(key: Hashable) -> str:
        dev = torch.cuda.get_device_name(torch.cuda.current_device())
        return f"{dev}|{key!r}"

    def best(self, key: Hashable, candidates: Sequence, run: Callable,
             feasible: Optional[Callable] = None):
        """Fastest candidate for ``key``; ``run(c)`` launches candidate c on the
        current stream, ``feasible(c)`` (optional) runs it once and says
        whether its result is usable."""
        k = self._key(key)
        with self._lock:
            # a winner cached from another candidate list is timed again
            if k in self._cache and self._cache[k] in list(candidates):
                return self._cache[k]
        stream = torch.cuda.current_stream()
        best_c, best_ms = None, float("inf")
        for c in candidates:
            if feasible is not None and not feasible(c):
                continue
            run(c)                                   # warm-up
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(self.reps):
                run(c)
            e1.record(stream)
            e1.synchronize()
            ms = e0.elapsed_time(e1) / self.reps
            self.log.append((key, c, ms))
            if ms < best_ms:
                best_c, best_ms = c, ms
        if best_c is None:
            raise RuntimeError(f"no feasible launch configuration for {key!r}")
        with self._lock:
            self._cache[k] = best_c
            if self.cache_file:
                try:
                    with open(self.cache_file, "w") as f:
                        json.dump({kk: list(v) if isinstance(v, tuple) else v
                                   for kk, v in self._cache.items()}, f, indent=1)
                except OSError:
                    pass
        return best_c
