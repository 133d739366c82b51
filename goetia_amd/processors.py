"""goetia's chunk driver for the dBG path: IntervalCounter (metrics.hh:101-145),
InserterProcessor<dBG> (processors.hh:248-344) on the native inserter
(gt_inserter_*) and FilterProcessor<Filter> (processors.hh:355-440), both over
FileProcessor (processors.hh:45-236).

advance() processes reads until the k-mers of the reads since the last
interval reach `interval`, returning right after that read with the running
(n_sequences, time, remaining) as goetia's does; reads are parsed and
processed in batches on the GPU instead of one call each.
"""
import ctypes as C

import numpy as np

from . import _native as N
from .parsing import FastxParser, SplitPairedReader, fastx_text
from .reads import ReadBatch


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

    def __init__(self, graph, interval=DEFAULT_INTERVAL, verbose=False, batch_bases=0):
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
        """FileProcessor::advance over a FastxParser or a SplitPairedReader
        (processors.hh:180-231): (n_sequences, time, remaining)."""
        n, t, rem = C.c_uint64(), C.c_uint64(), C.c_int()
        entry = 'gt_inserter_advance_pairs' if isinstance(parser, SplitPairedReader) else 'gt_inserter_advance'
        N.call(entry, self._h, parser._h, C.byref(n), C.byref(t), C.byref(rem))
        return n.value, t.value, bool(rem.value)

    def process(self, filename, strict=False, min_length=0, right_filename=None, force_name_match=False):
        """FileProcessor::process (processors.hh:69-126): a whole file, or a
        split pair of files; (n_sequences, time)."""
        if isinstance(filename, (FastxParser, SplitPairedReader)):
            parser = filename
        elif right_filename is not None:
            parser = SplitPairedReader(filename, right_filename, strict, min_length, force_name_match)
        else:
            parser = FastxParser(filename, strict, min_length)
        remaining = True
        n = t = 0
        while remaining:
            n, t, remaining = self.advance(parser)
        return n, t

    def chunked_process(self, file, right_file=None):
        """pythonize_processors.py:12-33: advance through a file (or a
        parser), yielding (n_sequences, time, parser.n_skipped()) after every
        advance that consumed reads."""
        if isinstance(file, (str, bytes)):
            parser = FastxParser(file) if right_file is None else SplitPairedReader(file, right_file)
        else:
            parser = file
        advancing = True
        prev = 0
        while advancing:
            n, t, advancing = self.advance(parser)
            if n > prev:
                yield n, t, parser.n_skipped()
            prev = n

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


class FilterProcessor:
    """goetia::FilterProcessor<Filter> (processors.hh:355-440): every read goes
    through the filter (StreamingSolidFilter or DiginormFilter) and the reads
    that pass are written to output_filename with Record::write_fastx
    (parsing.hh:38-49).  A read's time is its k-mer count (0 below K, which
    the reference's filter rejects with SequenceLengthException); advance()
    returns at the read that brings the IntervalCounter to the interval, and
    the filter runs on the GPU over the reads up to there, so the graph at a
    return holds exactly the reads before it."""
    DEFAULT_INTERVAL = IntervalCounter.DEFAULT_INTERVAL

    def __init__(self, filter, output_filename, interval=DEFAULT_INTERVAL, verbose=False, batch_reads=1 << 16,
                 batch_bases=16 << 20):
        self.filter = filter
        self.K = filter.K
        self.verbose = verbose
        self._interval = int(interval)
        self._timer = IntervalCounter(interval)
        self._out = open(output_filename, 'w')
        self._n_sequences = 0
        self._n_passed = 0
        self._pending = None        # [batch, names, qualities, times, next read]
        self._batch_reads, self._batch_bases = batch_reads, batch_bases

    @classmethod
    def build(cls, filter, output_filename, interval=DEFAULT_INTERVAL, verbose=False):
        return cls(filter, output_filename, interval, verbose)

    def close(self):
        if self._out is not None and not self._out.closed:
            self._out.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _run(self, batch, names, quals, i, j):
        o = batch.offsets
        sub = ReadBatch(batch.seq[o[i]:o[j]], o[i:j + 1] - o[i])
        passed = self.filter.filter_sequences(sub)
        keep = np.nonzero(passed)[0]
        if len(keep):
            seq = batch.seq
            self._out.write(''.join(fastx_text(names[i + k], seq[o[i + k]:o[i + k + 1]].tobytes().decode(),
                                               quals[i + k].decode() if quals[i + k] else None) for k in keep))
        self._n_passed += len(keep)

    def advance(self, parser):
        """FileProcessor::advance (processors.hh:200-222):
        (n_sequences, time, remaining)."""
        while True:
            if self._pending is None:
                batch, names, quals = parser.next_batch(self._batch_reads, self._batch_bases, quality=True)
                if batch.n_reads == 0:
                    self._out.flush()
                    return self._n_sequences, self._timer.total(), False
                times = np.maximum(np.diff(batch.offsets) - self.K + 1, 0)
                self._pending = [batch, names, quals, times, 0]
            batch, names, quals, times, i = self._pending
            reach = self._timer.counter() + np.cumsum(times[i:])
            hit = np.nonzero(reach >= self._interval)[0]
            j = i + (int(hit[0]) + 1 if len(hit) else len(times) - i)
            self._run(batch, names, quals, i, j)
            crossed = self._timer.poll(int(times[i:j].sum()))
            self._n_sequences += j - i
            if j >= batch.n_reads:
                self._pending = None
            else:
                self._pending[4] = j
            if crossed:
                self._out.flush()
                return self._n_sequences, self._timer.total(), True

    def process(self, filename, strict=False, min_length=0):
        """FileProcessor::process (processors.hh:100-126): (n_sequences, time)."""
        parser = filename if isinstance(filename, FastxParser) else FastxParser(filename, strict, min_length)
        remaining = True
        n = t = 0
        while remaining:
            n, t, remaining = self.advance(parser)
        return n, t

    def chunked_process(self, file, right_file=None):
        """pythonize_processors.py:12-33 over one file (pairs of files go
        through InserterProcessor; the filters here read single files)."""
        if right_file is not None or isinstance(file, SplitPairedReader):
            raise N.UnsupportedError('FilterProcessor reads single files; paired files go through InserterProcessor')
        return InserterProcessor.chunked_process(self, file)

    def n_passed(self):
        return self._n_passed

    def n_sequences(self):
        return self._n_sequences

    def time_elapsed(self):
        return self._timer.total()

    def interval(self):
        return self._interval


# ------------------------------------------------------------ streaming driver
# goetia keeps its streaming driver (AsyncSequenceProcessor and its channels,
# goetia/processors.py:38-351) in this module; here it lives in streaming.py.

from .streaming import (AsyncSequenceProcessor, MessageHandler, QueueManager, RunState,  # noqa: E402,F401
                        UniversalQueue, every_n_intervals)

DEFAULT_INTERVAL = IntervalCounter.DEFAULT_INTERVAL
