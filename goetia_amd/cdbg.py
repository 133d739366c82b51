"""StreamingCompactor's segment finder over the GPU dBG.

``StreamingCompactor[dBG[S, Sh]].Compactor.build(graph).find_new_segments(seq)``
follows compactor.hh:321-534: it splits a read into the runs of k-mers that
are new to the graph (segments with anchors and flanks), cut at the k-mers
that would become decision nodes once the read is in the graph, with
compact_segment as compactor.hh:77-145 and its Python conveniences as
pythonize_cdbg.py:53-97.  The work is one library call, gt_find_new_segments:
three GPU launches per call (or per batch of reads) give every window's
query count and, for the windows new to the graph, the eight neighbours with
their presence (gt_dbg_neighbors and a gather), and
the library's host code runs the two passes over each read, sequential per
read as in goetia.

goetia builds this code for forward hashes only: find_new_segments assigns
a neighbour's value_type to a hash (compactor.hh:405, 428), which Canonical
does not accept, and StreamingCompactor is instantiated for FwdLemireShifter
graphs (compactor.cc, end).  A canonical graph is refused here as well.

The rest of the compactor (the cDBG and its updates) is not part of this
package.
"""
import ctypes as C

import numpy as np

from . import _native as N
from .dbg import Kmer, _as_batch
from .hashing import Hash, _bytes
from .reads import ReadBatch

SYMBOLS = 'ACGT'


class compact_segment:
    """StreamingCompactor<G>::compact_segment (compactor.hh:77-145)."""
    __slots__ = ('left_anchor', 'right_anchor', 'left_flank', 'right_flank', 'is_decision_kmer', 'start_pos',
                 'length', 'tags')

    def __init__(self, left_anchor=None, right_anchor=None, is_decision_kmer=False, start_pos=0, length=0):
        # the default constructor sets the four hashes from 0 (compactor.hh:94-103)
        self.left_anchor = Hash(0) if left_anchor is None else left_anchor
        self.right_anchor = Hash(0) if right_anchor is None else right_anchor
        self.left_flank = Hash(0)
        self.right_flank = Hash(0)
        self.is_decision_kmer = bool(is_decision_kmer)
        self.start_pos = int(start_pos)
        self.length = int(length)
        self.tags = []

    def is_null(self):
        """compactor.hh:129-131: the delimiter between runs of segments."""
        return self.length == 0 and self.left_anchor.value == self.right_anchor.value and not self.is_decision_kmer

    @property
    def NULL(self):
        return self.is_null()

    @property
    def start(self):
        return self.start_pos

    def sequence(self, sequence):
        """pythonize_cdbg.py:55-66: the segment's substring of its read."""
        if self.is_null():
            return ''
        return sequence[self.start_pos:self.start_pos + self.length]

    def as_tuple(self):
        return (self.left_anchor.value, self.right_anchor.value, self.left_flank.value, self.right_flank.value,
                int(self.is_decision_kmer), self.start_pos, self.length)

    def __repr__(self):
        return (f'<Segment start={self.start_pos} length={self.length} left_anchor={self.left_anchor.value} '
                f'right_anchor={self.right_anchor.value} is_decision_kmer={self.is_decision_kmer} '
                f'is_null={self.is_null()}>')


class SegmentList(list):
    """The std::vector<compact_segment> a call returns (front / back as
    cppyy gives them)."""

    def front(self):
        return self[0]

    def back(self):
        return self[-1]


SEG_WORDS, DEC_WORDS = 7, 11


def _values(items):
    return np.array([h.value if isinstance(h, Hash) else int(h) for h in items], np.uint64)


class _Outputs:
    """Output buffers of gt_find_new_segments / gt_segments_from_view, grown
    to the totals the library reports when a first guess is too small."""

    def __init__(self, n_reads, n_windows):
        self.seg_off = np.zeros(n_reads + 1, np.int64)
        self.dec_off = np.zeros(n_reads + 1, np.int64)
        self.seg_cap = 4 * n_reads + n_windows // 16 + 16
        self.dec_cap = n_windows // 16 + 16
        self.segs = np.zeros(self.seg_cap * SEG_WORDS, np.uint64)
        self.decs = np.zeros(self.dec_cap * DEC_WORDS, np.uint64)
        self.n_segs, self.n_decs = C.c_uint64(), C.c_uint64()

    def args(self):
        return (N.ptr(self.seg_off), N.ptr(self.segs), self.seg_cap, N.ptr(self.dec_off), N.ptr(self.decs),
                self.dec_cap, C.byref(self.n_segs), C.byref(self.n_decs))

    def call(self, name, *head):
        try:
            N.call(name, *head, *self.args())
        except ValueError:
            if self.n_segs.value <= self.seg_cap and self.n_decs.value <= self.dec_cap:
                raise
            self.seg_cap, self.dec_cap = max(self.n_segs.value, 1), max(self.n_decs.value, 1)
            self.segs = np.zeros(self.seg_cap * SEG_WORDS, np.uint64)
            self.decs = np.zeros(self.dec_cap * DEC_WORDS, np.uint64)
            N.call(name, *head, *self.args())

    def read(self, r, s, K, counts, hv, hashes, new_kmers, segments, new_decision_kmers, decision_neighbors):
        """Appends read r's results (s: its sequence; counts / hv: its
        windows) to the caller's containers, as goetia's six-argument
        overload does."""
        for v in hv.tolist():
            hashes.append(Hash(v))
        for v, c in zip(hv.tolist(), counts.tolist()):
            if c == 0:
                new_kmers.add(Hash(v))
        for j in range(int(self.seg_off[r]), int(self.seg_off[r + 1])):
            w = self.segs[SEG_WORDS * j:SEG_WORDS * j + SEG_WORDS].tolist()
            seg = compact_segment(Hash(w[0]), Hash(w[1]), bool(w[4]), w[5], w[6])
            seg.left_flank, seg.right_flank = Hash(w[2]), Hash(w[3])
            segments.append(seg)
        for j in range(int(self.dec_off[r]), int(self.dec_off[r + 1])):
            w = self.decs[DEC_WORDS * j:DEC_WORDS * j + DEC_WORDS].tolist()
            pos, h, m, nb = w[0], w[1], w[2], w[3:]
            root = s[pos:pos + K]
            lk = [Kmer(Hash(nb[c]), SYMBOLS[c] + root[:K - 1]) for c in range(4) if (m >> c) & 1]
            rk = [Kmer(Hash(nb[4 + c]), root[1:] + SYMBOLS[c]) for c in range(4) if (m >> (4 + c)) & 1]
            decision_neighbors.append((lk, rk))
            new_decision_kmers.add(Hash(h))


def _containers(hashes, new_kmers, segments, new_decision_kmers, decision_neighbors):
    if hashes is None:
        return True, ([], set(), SegmentList(), set(), [])
    return False, (hashes, new_kmers, segments, new_decision_kmers, decision_neighbors)


def segments_from_view(K, sequence, counts, hash_values, nb_values, nb_present, hashes=None, new_kmers=None,
                       segments=None, new_decision_kmers=None, decision_neighbors=None):
    """find_new_segments of one read from its graph view (per window the
    query count, the hash value, the eight neighbour values -- left A C G T,
    right A C G T -- and the presence mask, bit c left, 4 + c right, as
    dBG.window_neighbors gives them), through the library's host code
    (gt_segments_from_view; no device).  Arguments and result as
    Compactor.find_new_segments."""
    s = sequence if isinstance(sequence, str) else _bytes(sequence).decode()
    one_arg, (hs, nk, segs, dks, pairs) = _containers(hashes, new_kmers, segments, new_decision_kmers,
                                                       decision_neighbors)
    counts = np.ascontiguousarray(counts, np.int16)
    hv = np.ascontiguousarray(hash_values, np.uint64)
    vals = np.ascontiguousarray(nb_values, np.uint64).reshape(-1, 8)
    mask = np.ascontiguousarray(nb_present, np.uint8)
    off = np.array([0, len(s)], np.int64)
    known, prefix = _values(nk), _values(hs)
    out = _Outputs(1, len(hv))
    out.call('gt_segments_from_view', K, N.ptr(off), 1, N.ptr(counts), N.ptr(hv), N.ptr(vals), N.ptr(mask),
             N.ptr(known), len(known), 0, N.ptr(prefix), len(prefix))
    out.read(0, s, K, counts, hv, hs, nk, segs, dks, pairs)
    return segs if one_arg else None


class Compactor:
    """StreamingCompactor<G>::Compactor, its segment finder."""

    def __init__(self, dbg, minimizer_window_size=8):
        if dbg.canonical:
            raise TypeError('StreamingCompactor::find_new_segments is built for forward (FwdLemireShifter) '
                            'graphs only')
        self.dbg = dbg
        self.K = dbg.K
        self.minimizer_window_size = int(minimizer_window_size)

    @classmethod
    def build(cls, dbg, minimizer_window_size=8):
        return cls(dbg, minimizer_window_size)

    def __smartptr__(self):
        return self

    def _run(self, batch, known, shared, prefix):
        nk = int(batch.kmer_counts(self.K).sum())
        counts = np.zeros(max(nk, 1), np.int16)
        hv = np.zeros(max(nk, 1), np.uint64)
        out = _Outputs(batch.n_reads, nk)
        out.call('gt_find_new_segments', self.dbg.S.handle, self.K, self.dbg.mode, N.ptr(batch.seq),
                 N.ptr(batch.offsets), batch.n_reads, N.ptr(known), len(known), int(shared), N.ptr(prefix),
                 len(prefix), N.ptr(counts), N.ptr(hv))
        return out, counts, hv

    def find_new_segments(self, sequence, hashes=None, new_kmers=None, segments=None, new_decision_kmers=None,
                          decision_neighbors=None):
        """compactor.hh:321-534.  With the sequence alone, returns the
        segments (fresh sets inside, the one-argument overload); with all six
        arguments, appends to the caller's hashes, new-k-mer set, segments,
        new decision k-mers and decision neighbour pairs, as the six-argument
        overload does.  Raises SequenceLengthException below K."""
        s = sequence if isinstance(sequence, str) else _bytes(sequence).decode()
        if len(s) < self.K:
            raise N.SequenceLengthException('KmerIterator: sequence shorter than K')
        one_arg, (hs, nk, segs, dks, pairs) = _containers(hashes, new_kmers, segments, new_decision_kmers,
                                                           decision_neighbors)
        out, counts, hv = self._run(ReadBatch.from_strings([s]), _values(nk), 0, _values(hs))
        n = len(s) - self.K + 1
        out.read(0, s, self.K, counts[:n], hv[:n], hs, nk, segs, dks, pairs)
        return segs if one_arg else None

    def find_new_segments_batch(self, reads, new_kmers=None):
        """find_new_segments over every read of a batch against the graph as
        it is (nothing inserted in between): three GPU launches and one pass of
        the library's host code for the whole batch.  new_kmers: None for a
        fresh set per read (the one-argument overload), or one set shared by
        the reads in order.  Returns per read (segments, hashes, new k-mers,
        new decision k-mers, decision neighbour pairs); reads shorter than K
        give None."""
        batch = _as_batch(reads)
        shared = new_kmers is not None
        out, counts, hv = self._run(batch, _values(new_kmers) if shared else _values(()), shared, _values(()))
        K = self.K
        res = []
        k = 0
        for r in range(batch.n_reads):
            a, b = int(batch.offsets[r]), int(batch.offsets[r + 1])
            if b - a < K:
                res.append(None)
                continue
            n = b - a - K + 1
            s = bytes(batch.seq[a:b]).decode()
            hs, segs, dks, pairs = [], SegmentList(), set(), []
            nk = new_kmers if shared else set()
            out.read(r, s, K, counts[k:k + n], hv[k:k + n], hs, nk, segs, dks, pairs)
            res.append((segs, hs, nk, dks, pairs))
            k += n
        return res


class _CompactorFamily(type):
    """StreamingCompactor[graph_type] -> a namespace with .Compactor and
    .compact_segment, as goetia's cppyy class template."""
    _cache = {}

    def __getitem__(cls, graph_type):
        sh = getattr(graph_type, 'shifter_type', None)
        if sh is not None and getattr(sh, 'canonical', False):
            raise TypeError('StreamingCompactor is built for forward (FwdLemireShifter) graphs only')
        if graph_type not in cls._cache:
            name = getattr(graph_type, '__name__', str(graph_type))
            cls._cache[graph_type] = type(f'StreamingCompactor[{name}]', (), {
                'Compactor': Compactor, 'compact_segment': compact_segment, 'graph_type': graph_type})
        return cls._cache[graph_type]


class StreamingCompactor(metaclass=_CompactorFamily):
    """goetia::StreamingCompactor<GraphType> (compactor.hh:35-60)."""
    Compactor = Compactor
    compact_segment = compact_segment
