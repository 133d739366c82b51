"""Consumers of the dBG path: goetia's StreamingSolidFilter
(solidifier.hh:20-80) and DiginormFilter (diginorm.hh:20-125), batched on the
GPU with exact per-read results."""
import numpy as np

from . import _native as N
from .dbg import _as_batch, _stream_ptr


class _FilterFamily(type):
    """Filter[graph_type] -> a namespace with .Filter and .Processor, as
    goetia's cppyy classes (DiginormFilter<G>::Filter, ::Processor)."""

    def __getitem__(cls, graph_type):
        from .processors import FilterProcessor
        name = getattr(graph_type, '__name__', str(graph_type))
        return type(f'{cls.__name__}[{name}]', (), {'Filter': cls, 'Processor': FilterProcessor,
                                                    'graph_type': graph_type})


class StreamingSolidFilter(metaclass=_FilterFamily):
    """goetia::StreamingSolidFilter<dBG>::Filter: a read passes unless the
    share of its k-mers whose count (just after their own insert) is below
    solid_threshold reaches 1 - min_prop_solid."""

    def __init__(self, graph, min_prop_solid=0.75, solid_threshold=1):
        self.dbg = graph
        self.K = graph.K
        self.min_prop_solid = float(min_prop_solid)
        self.solid_threshold = int(solid_threshold)

    @classmethod
    def build(cls, graph, min_prop_solid=0.75, solid_threshold=1):
        return cls(graph, min_prop_solid, solid_threshold)

    def __smartptr__(self):
        return self

    def filter_sequence(self, sequence):
        """solidifier.hh:58-73: (passed, n_kmers); SequenceLengthException
        below K, as the reference's insert_and_query_sequence raises."""
        s = sequence.encode() if isinstance(sequence, str) else bytes(sequence)
        if len(s) < self.K:
            raise N.SequenceLengthException('Sequence must have length >= K')
        return bool(self.filter_sequences([s])[0]), len(s) - self.K + 1

    def filter_sequences(self, reads):
        """One pass over a batch: passed flag per read (reads shorter than K
        do not pass), identical to filtering the reads one after another."""
        batch = _as_batch(reads)
        passed = np.zeros(max(batch.n_reads, 1), np.uint8)
        N.call('gt_solid_filter', self.dbg.S.handle, self.K, int(self.dbg.canonical), self.dbg.mode,
               N.ptr(batch.seq), N.ptr(batch.offsets), batch.n_reads, self.min_prop_solid, self.solid_threshold,
               N.ptr(passed))
        return passed[:batch.n_reads].astype(bool)

    def filter_device(self, reads, passed, stream=None):
        """Device batch (DeviceReads) into a uint8 device tensor, asynchronous."""
        N.call('gt_solid_filter_dev', self.dbg.S.handle, self.K, int(self.dbg.canonical), self.dbg.mode,
               N.ptr(reads.seq), N.ptr(reads.offsets), reads.n_reads, reads.n_bases, self.min_prop_solid,
               self.solid_threshold, N.ptr(passed), _stream_ptr(stream))


class DiginormFilter(metaclass=_FilterFamily):
    """goetia::DiginormFilter<dBG<ByteStorage, Sh>>::Filter (diginorm.hh:93-119):
    a read is dropped when at least half of its k-mers (min_req =
    int(0.5 + float32(n_kmers) / 2)) already have a count of at least the
    cutoff, and inserted otherwise, before the next read is looked at.
    Batches give the same decisions, tables and counters as filtering their
    reads one after another."""

    def __init__(self, graph, cutoff):
        self.dbg = graph
        self.K = graph.K
        self.cutoff = int(cutoff)

    @classmethod
    def build(cls, graph, cutoff):
        return cls(graph, cutoff)

    def __smartptr__(self):
        return self

    def filter_sequence(self, sequence):
        """diginorm.hh:105-119: (passed, n_kmers); SequenceLengthException
        below K, as the reference's k-mer iterator raises."""
        s = sequence.encode() if isinstance(sequence, str) else bytes(sequence)
        if len(s) < self.K:
            raise N.SequenceLengthException('Sequence must have length >= K')
        return bool(self.filter_sequences([s])[0]), len(s) - self.K + 1

    def filter_sequences(self, reads):
        """Passed flag per read, in read order (reads shorter than K do not
        pass and are not inserted)."""
        batch = _as_batch(reads)
        passed = np.zeros(max(batch.n_reads, 1), np.uint8)
        N.call('gt_diginorm', self.dbg.S.handle, self.K, int(self.dbg.canonical), self.dbg.mode,
               N.ptr(batch.seq), N.ptr(batch.offsets), batch.n_reads, self.cutoff, N.ptr(passed))
        return passed[:batch.n_reads].astype(bool)

    def filter_device(self, reads, passed, stream=None):
        """Device batch (DeviceReads) into a uint8 device tensor; returns
        after the stream has run it."""
        N.call('gt_diginorm_dev', self.dbg.S.handle, self.K, int(self.dbg.canonical), self.dbg.mode,
               N.ptr(reads.seq), N.ptr(reads.offsets), reads.n_reads, reads.n_bases, self.cutoff, N.ptr(passed),
               _stream_ptr(stream))

    @staticmethod
    def median_count_at_least(sequence, cutoff, graph):
        """diginorm.hh:39-76: whether at least min_req of the read's k-mers
        have a count of at least cutoff (no insert)."""
        counts = np.asarray(graph.query_sequence(sequence), np.int64)
        min_req = int(0.5 + float(np.float32(len(counts)) / np.float32(2)))
        return int(np.sum((counts >= 0) & (counts >= int(cutoff)))) >= min_req

    @staticmethod
    def median_in_range(sequence, low, high, graph):
        """diginorm.hh:78-94: (low < median <= high, hashes), the median being
        the element n // 2 of the sorted counts."""
        counts, hashes = [], []
        graph.query_sequence(sequence, counts, hashes)
        median = int(np.partition(np.asarray(counts, np.int64), len(counts) // 2)[len(counts) // 2])
        return low < median <= high, hashes
