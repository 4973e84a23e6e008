"""Serving contraction-tree jobs with continuous batching (libmuscle_hip's
server: admission by free device memory, device-resident intermediates,
one batched persistent launch per scheduler tick across all admitted jobs).

    srv = Server(pool_gib=200)
    jid = srv.submit(leaves, nodes)       # leaves: device Tensors; nodes: [(lhs, rhs, out_inds), ...]
    res = srv.wait(jid)                   # timestamps, status
    out = srv.result(jid, inds)           # root tensor as a muscle.Tensor (copied out)
    srv.release(jid)

A job is what a tensor-network caller evaluates with successive
binary_einsum calls (src/Operations/binary_einsum.jl:33-51): operands are
numbered leaves first, then nodes; every node's operands precede it; the last
node is the root.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .errors import ArgumentError
from .tensor import Index, Tensor, empty_col_major

MODES = {"batched": 0, "loop": 1}


class Server:
    def __init__(self, pool_gib: float = 64.0, max_active: int = 64, mode: str = "batched", device: int = 0):
        self.device = device
        with torch.cuda.device(device):
            self._s = ctypes.c_void_p()
            _lib.check(_lib.lib.muscle_hip_server_create(int(pool_gib * 2 ** 30), int(max_active), MODES[mode],
                                                         ctypes.byref(self._s)))
        self._jobs = {}  # job id -> (keepalive, root labels, root shape, dtype)

    def submit(self, leaves, nodes) -> int:
        """leaves: list of muscle.Tensor (device, one element type); nodes: list of
        (lhs, rhs, out_inds).  Index extents come from the leaves."""
        return self.submit_prepared(self.prepare(leaves, nodes))

    def prepare(self, leaves, nodes) -> "PreparedJob":
        """Builds the C job description once (submit_prepared then only queues it)."""
        if not leaves or not nodes:
            raise ArgumentError("a job needs leaves and nodes")
        labels, extents = {}, []
        for t in leaves:
            for i, n in zip(t.inds(), t.shape):
                if i not in labels:
                    labels[i] = len(labels)
                    extents.append(int(n))
        node_inds = []
        for lhs, rhs, out in nodes:
            out = [x if isinstance(x, Index) else Index(x) for x in out]
            for i in out:
                if i not in labels:
                    raise ArgumentError(f"node output index {i} is not an index of the leaves")
            node_inds.append(out)
        nl, nn = len(leaves), len(nodes)
        dtype = leaves[0].dtype
        ldesc = (_lib.TensorDesc * nl)()
        keep = []
        for k, t in enumerate(leaves):
            d = t.parent()
            if d.dtype != dtype or d.device.type != "cuda":
                raise ArgumentError("leaves must be device tensors of one element type")
            dd = _lib.Desc(_lib.DTYPE_CODE[d.dtype], d.data_ptr(), [labels[i] for i in t.inds()], list(d.shape),
                           list(d.stride()))
            keep.append(dd)
            ldesc[k] = dd.desc
        lhs = (ctypes.c_int32 * nn)(*[int(n[0]) for n in nodes])
        rhs = (ctypes.c_int32 * nn)(*[int(n[1]) for n in nodes])
        offs, modes = [0], []
        for out in node_inds:
            modes += [labels[i] for i in out]
            offs.append(len(modes))
        off = (ctypes.c_int32 * (nn + 1))(*offs)
        md = (ctypes.c_int32 * max(len(modes), 1))(*modes)
        ext = (ctypes.c_int64 * max(len(extents), 1))(*extents)
        job = _lib.JobDesc(nl, nn, ldesc, lhs, rhs, off, md, len(extents), ext)
        root = node_inds[-1]
        return PreparedJob(job, (keep, ldesc, lhs, rhs, off, md, ext), leaves, root,
                           [extents[labels[i]] for i in root], dtype)

    def submit_prepared(self, pj: "PreparedJob") -> int:
        jid = ctypes.c_int64()
        _lib.check(_lib.lib.muscle_hip_server_submit(self._s, ctypes.byref(pj.job), ctypes.byref(jid)))
        self._jobs[jid.value] = (pj.leaves, pj.root, pj.shape, pj.dtype)
        return jid.value

    def wait(self, job_id: int, timeout: float = -1.0) -> dict:
        r = _lib.JobResult()
        st = _lib.lib.muscle_hip_server_wait(self._s, int(job_id), float(timeout), ctypes.byref(r))
        if st not in (_lib.OK, _lib.NOT_READY):
            _lib.check(st)
        out = r.as_dict()
        out["latency_s"] = r.done_s - r.submit_s if st == _lib.OK else None
        return out

    def result(self, job_id: int) -> Tensor:
        leaves, root, shape, dtype = self._jobs[job_id]
        with torch.cuda.device(self.device):
            out = empty_col_major(shape, dtype, torch.device("cuda", self.device))
            _lib.check(_lib.lib.muscle_hip_server_copy_result(self._s, int(job_id), ctypes.c_void_p(out.data_ptr())))
        return Tensor(out, root)

    def release(self, job_id: int) -> None:
        _lib.check(_lib.lib.muscle_hip_server_release(self._s, int(job_id)))
        self._jobs.pop(job_id, None)

    def stats(self) -> dict:
        s = _lib.ServerStats()
        _lib.check(_lib.lib.muscle_hip_server_get_stats(self._s, ctypes.byref(s)))
        return s.as_dict()

    def clock(self) -> float:
        return float(_lib.lib.muscle_hip_server_clock(self._s))

    def close(self) -> None:
        if self._s:
            _lib.check(_lib.lib.muscle_hip_server_destroy(self._s))
            self._s = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PreparedJob:
    __slots__ = ("job", "_keep", "leaves", "root", "shape", "dtype")

    def __init__(self, job, keep, leaves, root, shape, dtype):
        self.job, self._keep, self.leaves, self.root, self.shape, self.dtype = job, keep, leaves, root, shape, dtype


DTYPES = {"c32": torch.complex64, "c64": torch.complex128}


class LeafBank:
    """Device-resident random leaf data for synthetic tree jobs.

    Leaves are column-major views (every index of dimension 2) into one
    region per (element type, scale exponent e): iid complex normal entries
    with E|z|^2 = 4^-e, so that the scale a tree needs at each leaf
    (workloads.leaf_scale_exponents) keeps its intermediates near unit size.
    A region is as large as the largest leaf placed in it; offsets are seeded."""

    def __init__(self, device: int = 0, seed: int = 0):
        self.device = torch.device("cuda", device)
        self.gen = torch.Generator(device=self.device)
        self.gen.manual_seed(seed)
        self.regions = {}

    def reserve(self, specs) -> None:
        """Allocates every region the jobs need (call before timing)."""
        from .workloads import leaf_scale_exponents
        need = {}
        for spec in specs:
            for inds, e in zip(spec.leaves, leaf_scale_exponents(spec)):
                k = (spec.dtype, e)
                need[k] = max(need.get(k, 0), len(inds))
        for (dt, e), r in sorted(need.items()):
            cur = self.regions.get((dt, e))
            if cur is not None and cur.numel() >= (1 << r):
                continue
            self.regions[(dt, e)] = None
            x = torch.randn(1 << r, dtype=DTYPES[dt], device=self.device, generator=self.gen)
            self.regions[(dt, e)] = x.mul_(2.0 ** -e)

    def bytes(self) -> int:
        return sum(x.numel() * x.element_size() for x in self.regions.values())

    def leaves(self, spec, rng) -> list:
        from .workloads import leaf_scale_exponents
        self.reserve([spec])
        out = []
        for inds, e in zip(spec.leaves, leaf_scale_exponents(spec)):
            pool = self.regions[(spec.dtype, e)]
            r = len(inds)
            off = int(rng.integers(0, pool.numel() - (1 << r) + 1))
            v = pool.as_strided((2,) * r, tuple(1 << i for i in range(r)), off)
            out.append(Tensor(v, [Index(i) for i in inds]))
        return out


def tree_nodes(spec):
    return [(a, b, [Index(i) for i in out]) for a, b, out in spec.nodes]
