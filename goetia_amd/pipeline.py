"""Batched insert engine and tile-shape tuning (gt_pipeline_* / gt_tune_* /
gt_autotune of libgoetia_amd).

``Pipeline`` streams batches of host reads into a dBG: each batch is copied
to pinned memory, moved to the device on a copy stream while the previous
batch runs, and processed by a hipGraph captured once per buffer (batch sizes
live in device memory, so the same graph serves every batch).  Per-read n_new
equals ``dBG.insert_sequences`` over the same reads in the same order.
"""
import ctypes as C

import numpy as np

from . import _native as N
from .reads import ReadBatch


class Pipeline:
    def __init__(self, graph, max_reads, max_bases):
        self.graph = graph
        self._h = C.c_void_p()
        self._pending = []
        N.call('gt_pipeline_new', graph.S.handle, graph.K, int(graph.canonical), graph.mode,
               int(max_reads), int(max_bases), C.byref(self._h))
        self.max_reads = int(max_reads)
        self.max_bases = int(max_bases)

    def submit(self, batch, n_new=None):
        """Queues a ReadBatch; returns the array that receives its per-read
        n_new (complete after ``flush``)."""
        if not isinstance(batch, ReadBatch):
            batch = ReadBatch.from_strings(list(batch))
        out = n_new if n_new is not None else np.zeros(max(batch.n_reads, 1), np.uint64)
        N.call('gt_pipeline_submit', self._h, N.ptr(batch.seq), N.ptr(batch.offsets),
               batch.n_reads, N.ptr(out))
        self._pending.append(out)
        return out

    def flush(self):
        N.call('gt_pipeline_flush', self._h)
        self._pending.clear()

    def close(self):
        if self._h and self._h.value:
            N.load().gt_pipeline_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.flush()
        self.close()


def tune_set(kind, per, blocks_per_cu):
    N.call('gt_tune_set', int(kind), int(per), int(blocks_per_cu))


def tune_get(kind):
    a, b = C.c_int(), C.c_int()
    N.call('gt_tune_get', int(kind), C.byref(a), C.byref(b))
    return a.value, b.value


def autotune(graph, reads, stream=None):
    """Times every tile shape on `reads` (DeviceReads) for the graph's
    storage kind, keeps the fastest and returns (per, blocks_per_cu)."""
    from .dbg import _stream_ptr
    a, b = C.c_int(), C.c_int()
    N.call('gt_autotune', graph.S.handle, graph.K, int(graph.canonical), graph.mode,
           N.ptr(reads.seq), N.ptr(reads.offsets), reads.n_reads, reads.n_bases,
           C.byref(a), C.byref(b), _stream_ptr(stream))
    return a.value, b.value
