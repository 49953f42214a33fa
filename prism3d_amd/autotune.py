"""Online autotuner of the chemistry kernel configuration.

The fixed-point driver calls the chemistry solve once per iteration on the
same grid.  After a warm-up call (the cold first iteration is not
representative), the tuner hands out each candidate configuration for one
call, scores it by device time per accepted BDF step (calls differ in how
many steps their cells need, so time alone does not compare them), and then
keeps the best.  Choices are cached per (device, network, cell-count bucket)
in a small JSON file so later runs start with the tuned configuration.

A configuration is (variant, waves_per_cu, rescue_frac) of p3d_solve_opts,
so the tuner picks both the cells per workgroup and the LU shape: variants
2 and 3 are one cell per 64-lane workgroup with the LU in registers (lane =
row) or in LDS with unrolled elimination steps; variant 4 is the persistent
one-cell-per-lane kernel (64 cells per wavefront, statically ordered sparse
LU) with `waves_per_cu` resident wavefronts per CU that hands its long-tail
cells over once `rescue_frac` of its lanes are idle; variant 6 is the
persistent one-cell-per-wavefront kernel (lane = species, the same sparse LU
with its rows in registers), `waves_per_cu` capping its resident wavefronts
(0: as many as fit).
"""

import json
import math
import os

DEFAULT_CANDIDATES = ((6, 0, 0.0), (6, 2, 0.0), (4, 8, 0.95), (4, 8, 0.99), (4, 12, 0.95),
                      (2, 0, 0.0))


def _cache_path():
    base = os.environ.get("P3D_TUNE_CACHE") or os.path.join(
        os.path.expanduser("~"), ".cache", "prism3d_amd", "chem_tune.json")
    return base


def _bucket(n_cells):
    """Cell counts within a factor of two share a tuning entry."""
    return int(round(math.log2(max(n_cells, 1))))


class ChemTuner:
    def __init__(self, n_cells, device_name, network_tag="bundled31",
                 candidates=DEFAULT_CANDIDATES, warmup_calls=1, cache=True):
        self.key = f"{device_name}|{network_tag}|{_bucket(n_cells)}"
        self.candidates = [tuple(c) for c in candidates]
        self.warmup_calls = int(warmup_calls)
        self.cache = cache
        self.calls = 0
        self.scores = {}      # config -> seconds per accepted step
        self.best = None
        if cache:
            cached = self._load().get(self.key)
            if cached is not None:
                self.best = tuple(cached["config"])

    # -- cache --------------------------------------------------------------
    def _load(self):
        try:
            with open(_cache_path()) as f:
                return json.load(f)
        except (OSError, ValueError):
            return {}

    def _store(self):
        if not self.cache:
            return
        path = _cache_path()
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            data = self._load()
            data[self.key] = {"config": list(self.best),
                              "scores": {",".join(map(str, k)): v
                                         for k, v in self.scores.items()}}
            with open(path, "w") as f:
                json.dump(data, f, indent=1)
        except OSError:
            pass

    # -- protocol -----------------------------------------------------------
    def next_config(self):
        """Configuration for the next chemistry call."""
        if self.best is not None:
            return self.best
        if self.calls < self.warmup_calls:
            return self.candidates[0]
        untried = [c for c in self.candidates if c not in self.scores]
        return untried[0] if untried else self.candidates[0]

    def record(self, config, seconds, accepted_steps):
        """Time of the call that used `config` and its total accepted steps."""
        self.calls += 1
        if self.best is not None or self.calls <= self.warmup_calls:
            return
        self.scores[tuple(config)] = seconds / max(float(accepted_steps), 1.0)
        if all(c in self.scores for c in self.candidates):
            self.best = min(self.scores, key=self.scores.get)
            self._store()

    @property
    def tuning(self):
        return self.best is None
