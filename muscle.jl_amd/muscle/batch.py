"""Batched contractions: many pairwise binary_einsums in one persistent launch.

    batch = ContractionBatch([(c1, a1, b1), (c2, a2, b2), ...])   # device Tensors
    batch.autotune()          # optional: pick tile shapes per size class
    batch.run()               # one hipGraph-replayed persistent launch
    batch.run("loop")         # the same pairs as plain per-call launches

Each pair has the semantics of binary_einsum!(c, a, b) (batch, contracted,
open, summed and diagonal modes).  The reference has no batched form: its
callers loop over binary_einsum (src/Operations/binary_einsum.jl:33-70).
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .errors import ArgumentError
from .ops import _descs, _device_of, _handle_for
from .tensor import Tensor

MODES = {"persistent": 0, "loop": 1, "persistent_nograph": 2}
TILES = {"auto": -1, "64x64": 0, "32x128": 1, "128x32": 2}


class ContractionBatch:
    def __init__(self, pairs, tiles=None, complex_mult: int = 3):
        pairs = list(pairs)
        if not pairs:
            raise ArgumentError("empty batch")
        self.device = _device_of(*[t for p in pairs for t in p])
        self._pairs = pairs  # keep the tensors alive
        n = len(pairs)
        cs = (_lib.TensorDesc * n)()
        as_ = (_lib.TensorDesc * n)()
        bs = (_lib.TensorDesc * n)()
        self._keep = []
        for i, (c, a, b) in enumerate(pairs):
            dc, da, db = _descs(c, a, b)
            self._keep += [dc, da, db]
            cs[i], as_[i], bs[i] = dc.desc, da.desc, db.desc
        tl = None
        if tiles is not None:
            tl = (ctypes.c_int32 * n)(*[TILES[t] if isinstance(t, str) else int(t) for t in tiles])
        self._h = _handle_for(self.device)
        self._complex_mult = int(complex_mult)
        _lib.set_option(self.device, self._h, 0, self._complex_mult)
        self._b = ctypes.c_void_p()
        _lib.check(_lib.lib.muscle_hip_batch_create(self._h, n, cs, as_, bs, tl, ctypes.byref(self._b)))

    def run(self, mode: str = "persistent") -> None:
        # the library runs the batch in the complex product form it was built with
        h = _handle_for(self.device)
        _lib.check(_lib.lib.muscle_hip_batch_run(h, self._b, MODES[mode]))

    def info(self) -> dict:
        inf = _lib.BatchInfo()
        _lib.check(_lib.lib.muscle_hip_batch_get_info(self._b, ctypes.byref(inf)))
        return inf.as_dict()

    def pair_info(self, i: int) -> dict:
        """Route (empty, tiles, split_k, shallow, solo or skinny), tile, size class, K
        slices and fused shape of pair i."""
        inf = _lib.BatchPairInfo()
        _lib.check(_lib.lib.muscle_hip_batch_get_pair_info(self._b, int(i), ctypes.byref(inf)))
        out = inf.as_dict()
        out["route"] = _lib.PAIR_ROUTES[out["route"]]
        return out

    def autotune(self, samples: int = 16, reps: int = 3) -> None:
        h = _handle_for(self.device)
        _lib.check(_lib.lib.muscle_hip_batch_autotune(h, self._b, int(samples), int(reps)))

    def close(self) -> None:
        if self._b:
            _lib.check(_lib.lib.muscle_hip_batch_destroy(self._b))
            self._b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tuning_save(path: str, device: int = 0) -> None:
    _lib.check(_lib.lib.muscle_hip_tuning_save(_lib.handle(device), path.encode()))


def tuning_load(path: str, device: int = 0) -> None:
    _lib.check(_lib.lib.muscle_hip_tuning_load(_lib.handle(device), path.encode()))


def tuning_class(m: int, n: int) -> int:
    return int(_lib.lib.muscle_hip_tuning_class(int(m), int(n)))
