"""Batched form: many particle sets of uneven size through one HIP graph.

``SnapshotBatch(sizes, des_ngb, K)`` owns the per-snapshot indices and
neighbour lists of libmeshoid_hip.so's batch runtime (mhip_batch_*).  Each
``run(pos_list, field_list)`` computes, per snapshot, the smoothing length,
density and the kernel-weighted least-squares gradient of every field, i.e.
for each snapshot what ``Meshoid(pos, des_ngb=k).D(f)`` computes
(src/meshoid/Meshoid.py:147-170,237-268), with all snapshots in one graph
launch and the field contraction of all snapshots in one kernel.
"""
from __future__ import annotations

import ctypes

import torch

from . import engine
from ._lib import lib, raise_for


def _ptrs(ts):
    arr = (ctypes.c_void_p * len(ts))()
    for i, t in enumerate(ts):
        arr[i] = t.data_ptr()
    return arr


class SnapshotBatch:
    """sizes: particles per snapshot; des_ngb: int or one per snapshot;
    K: field components per particle; nbranch: most concurrent streams (the
    first run times 1, 2 and 4 of them, up to nbranch)."""

    def __init__(self, sizes, des_ngb, K, nbranch=4, device=None):
        self.sizes = [int(n) for n in sizes]
        ns = len(self.sizes)
        if isinstance(des_ngb, int):
            des_ngb = [des_ngb] * ns
        self.des_ngb = [int(k) for k in des_ngb]
        if len(self.des_ngb) != ns:
            raise ValueError("one des_ngb per snapshot")
        self.K = int(K)
        self.device = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        # graphs are captured on a stream of the batch's own (the legacy
        # default stream cannot be captured)
        self._stream = torch.cuda.Stream(self.device)
        self._ctx = engine.context(self.device)
        n_arr = (ctypes.c_int64 * ns)(*self.sizes)
        k_arr = (ctypes.c_int * ns)(*self.des_ngb)
        h = ctypes.c_void_p()
        raise_for(lib.mhip_batch_create(self._ctx, ns, n_arr, k_arr, self.K, int(nbranch), ctypes.byref(h)),
                  self._ctx)
        self.h = h
        self.out = None

    def _alloc(self):
        dev = self.device
        self.out = {
            "h": [torch.empty(n, dtype=torch.float32, device=dev) for n in self.sizes],
            "rho": [torch.empty(n, dtype=torch.float32, device=dev) for n in self.sizes],
            "grad": [torch.empty((n, self.K, 3), dtype=torch.float32, device=dev) for n in self.sizes],
            "status": [torch.empty(n, dtype=torch.int32, device=dev) for n in self.sizes],
        }

    def run(self, pos, fields, graph=True, out=None):
        """pos: list of (n_s, 3) float32 device tensors; fields: list of
        (n_s, K) float32 device tensors.  Returns dict of per-snapshot lists
        h, rho, grad (n_s, K, 3), status."""
        if len(pos) != len(self.sizes) or len(fields) != len(self.sizes):
            raise ValueError("one position and one field array per snapshot")
        for p, f, n in zip(pos, fields, self.sizes):
            engine._need(p, (torch.float32,), "pos")
            engine._need(f, (torch.float32,), "fields")
            if tuple(p.shape) != (n, 3) or tuple(f.shape) != (n, self.K):
                raise ValueError("snapshot shapes differ from the batch layout")
        if out is not None:
            self.out = out
        elif self.out is None:
            self._alloc()
        o = self.out
        cur = torch.cuda.current_stream(self.device)
        self._stream.wait_stream(cur)
        with torch.cuda.stream(self._stream):
            ctx = engine.context(self.device)
            rc = lib.mhip_batch_run(ctx, self.h, _ptrs(pos), _ptrs(fields), _ptrs(o["h"]), _ptrs(o["rho"]),
                                    _ptrs(o["grad"]), _ptrs(o["status"]), int(bool(graph)))
            raise_for(rc, ctx)
        cur.wait_stream(self._stream)
        return o

    def force(self, candidate=-1, nbranch=0):
        """Test hook, before the first run: use candidate index ``candidate``
        of the tuning list (-1: timed) and ``nbranch`` streams (0: timed)."""
        raise_for(lib.mhip_batch_force(self.h, int(candidate), int(nbranch)))

    def info(self):
        """The first run's timings and choices: every contraction candidate
        ("tile": fields per thread, "wide": threads per query on the VALU,
        "mfma": the same on the matrix cores) with its ms, the candidate in
        use, and the branch streams in use with the timings of 1, 2 and 4."""
        nc, choice, nb = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        kind = (ctypes.c_int * 4)()
        param = (ctypes.c_int * 4)()
        ms = (ctypes.c_float * 4)()
        bms = (ctypes.c_float * 3)()
        raise_for(lib.mhip_batch_tuning(self.h, ctypes.byref(nc), kind, param, ms, ctypes.byref(choice),
                                        ctypes.byref(nb), bms))
        names = {0: "tile", 1: "wide", 2: "mfma"}
        cands = [{"kernel": names[kind[i]], "param": int(param[i]), "ms": round(float(ms[i]), 4)}
                 for i in range(nc.value)]
        ch = cands[choice.value] if 0 <= choice.value < nc.value else None
        return {"candidates": cands, "choice": ch, "branches": nb.value,
                "branch_ms": {n: round(float(bms[i]), 3) for i, n in enumerate((1, 2, 4)) if bms[i] > 0}}

    def __del__(self):
        h = getattr(self, "h", None)
        if h and lib is not None:
            lib.mhip_batch_destroy(h)
            self.h = None
