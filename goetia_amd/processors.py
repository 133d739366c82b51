"""goetia's chunk driver for the dBG path: IntervalCounter (metrics.hh:101-145)
and InserterProcessor<dBG> (processors.hh:248-344) over FileProcessor
(processors.hh:45-236), on the native inserter (gt_inserter_*).

advance() inserts reads until the k-mers of the reads since the last
interval reach `interval`, returning right after that read with the running
(n_sequences, time, remaining) as goetia's does; reads are parsed and
inserted in batches on the GPU instead of one insert_sequence call each.
"""
import ctypes as C

from . import _native as N
from .parsing import FastxParser


class IntervalCounter:
    """goetia::IntervalCounter: poll(incr) adds incr and returns True, resetting
    the counter, once it reaches the interval."""
    DEFAULT_INTERVAL = 500000

    def __init__(self, interval=DEFAULT_INTERVAL):
        self.interval = interval
        self._counter = 0
        self._total = 0

    def counter(self):
        return self._counter

    def total(self):
        return self._total

    def poll(self, incr=1):
        self._counter += incr
        self._total += incr
        if self._counter >= self.interval:
            self._counter = 0
            return True
        return False


class InserterProcessor:
    """goetia::InserterProcessor<dBG<S, Sh>> for one graph."""
    DEFAULT_INTERVAL = IntervalCounter.DEFAULT_INTERVAL

    def __init__(self, graph, interval=DEFAULT_INTERVAL, verbose=False, batch_bases=64 << 20):
        self.inserter = graph
        self.verbose = verbose
        self._interval = interval
        self._h = N.vp()
        N.call('gt_inserter_new', graph.S.handle, graph.K, int(graph.canonical), graph.mode, int(interval),
               int(batch_bases), C.byref(self._h))

    @classmethod
    def build(cls, graph, interval=DEFAULT_INTERVAL, verbose=False):
        return cls(graph, interval, verbose)

    def __del__(self):
        try:
            if self._h.value:
                N.load().gt_inserter_free(self._h)
                self._h = N.vp()
        except Exception:
            pass

    def advance(self, parser):
        """FileProcessor::advance: (n_sequences, time, remaining)."""
        n, t, rem = C.c_uint64(), C.c_uint64(), C.c_int()
        N.call('gt_inserter_advance', self._h, parser._h, C.byref(n), C.byref(t), C.byref(rem))
        return n.value, t.value, bool(rem.value)

    def process(self, filename, strict=False, min_length=0):
        """FileProcessor::process (processors.hh:100-126): the whole file;
        (n_sequences, time)."""
        parser = filename if isinstance(filename, FastxParser) else FastxParser(filename, strict, min_length)
        remaining = True
        n = t = 0
        while remaining:
            n, t, remaining = self.advance(parser)
        return n, t

    def _stats(self):
        n, t, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
        N.call('gt_inserter_stats', self._h, C.byref(n), C.byref(t), C.byref(c))
        return n.value, t.value, c.value

    def n_sequences(self):
        return self._stats()[0]

    def time_elapsed(self):
        return self._stats()[1]

    def interval(self):
        return self._interval
