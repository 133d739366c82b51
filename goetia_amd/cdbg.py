"""StreamingCompactor's segment finder over the GPU dBG.

``StreamingCompactor[dBG[S, Sh]].Compactor.build(graph).find_new_segments(seq)``
follows compactor.hh:321-534: it splits a read into the runs of k-mers that
are new to the graph (segments with anchors and flanks), cut at the k-mers
that would become decision nodes once the read is in the graph, with
compact_segment as compactor.hh:77-145 and its Python conveniences as
pythonize_cdbg.py:53-97.  The graph work runs on the GPU, two launches per
call (or per batch of reads): one query of every window, and one pass of the
neighbour kernel giving every window's eight neighbours and their presence
(gt_dbg_neighbors).  The segment bookkeeping is per read and sequential, as in
goetia, and stays on the host.

goetia builds this code for forward hashes only: find_new_segments assigns
a neighbour's value_type to a hash (compactor.hh:405, 428), which Canonical
does not accept, and StreamingCompactor is instantiated for FwdLemireShifter
graphs (compactor.cc, end).  A canonical graph is refused here as well.

The rest of the compactor (the cDBG and its updates) is not part of this
package.
"""

import numpy as np

from . import _native as N
from .dbg import Kmer, Shift, _as_batch
from .hashing import Hash, _bytes

SYMBOLS = 'ACGT'
_POPCOUNT4 = np.array([bin(i).count('1') for i in range(16)], np.uint8)


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


def _hash_at(hashes, i):
    # compactor.hh:471-525 index the hash vector at start_pos +- 1 even where
    # that runs off the read; the segment built from such an index is always
    # shorter than K and never kept, so its value is never seen.
    return hashes[i] if 0 <= i < len(hashes) else Hash()


class _SegmentFinder:
    """The host side of find_new_segments: the two passes over one read
    given its graph view."""

    def __init__(self, K):
        self.K = K

    def run(self, s, counts, hv, vals, mask, hashes, new_kmers, segments, new_decision_kmers, decision_neighbors):
        K = self.K
        one_arg = hashes is None
        if one_arg:
            hashes, new_kmers, segments = [], set(), SegmentList()
            new_decision_kmers, decision_neighbors = set(), []
        base = len(hashes)
        known = {h.value if isinstance(h, Hash) else int(h) for h in new_kmers}

        def present(i, lo):
            """Shifts of window i (lo 0: left, 4: right) present in the graph
            or among the new k-mers (filter_nodes with extras,
            unitig_walker.hh:457-468), in ACGT order."""
            m = int(mask[i])
            out = []
            for c in range(4):
                v = int(vals[i, lo + c])
                if (m >> (lo + c)) & 1 or v in known:
                    out.append(Shift(Hash(v), SYMBOLS[c], lo == 4))
            return out

        # first pass: runs of new, not yet seen k-mers (compactor.hh:352-400)
        preprocess = []
        current = compact_segment()
        prev_hash = Hash()
        prev_new = prev_seen = cur_new = cur_seen = False
        cur_hash = None
        n = len(hv)
        for pos in range(n):
            cur_hash = Hash(int(hv[pos]))
            cur_new = int(counts[pos]) == 0
            cur_seen = cur_hash.value in known
            hashes.append(cur_hash)
            if cur_new and not cur_seen:
                if not prev_new or prev_seen:
                    preprocess.append(current)
                    current = self._init_segment(cur_hash, prev_hash, pos)
                new_kmers.add(cur_hash)
                known.add(cur_hash.value)
            elif not current.is_null() and (not cur_new or cur_seen):
                self._finish_segment(current, pos - 1, prev_hash, cur_hash, preprocess)
                current = compact_segment()
            prev_hash, prev_new, prev_seen = cur_hash, cur_new, cur_seen

        if cur_new and not cur_seen:
            # the read ends inside a segment: its right flank is the unique
            # right neighbour, if there is one (compactor.hh:401-417)
            right_flank = cur_hash
            rn = present(n - 1, 4)
            if len(rn) == 1:
                right_flank = rn[0].hash
            self._finish_segment(current, n - 1, cur_hash, right_flank, preprocess)
        preprocess.append(compact_segment())
        if len(preprocess) < 3:
            return segments if one_arg else None

        if preprocess[1].start_pos == 0:
            # compactor.hh:425-437
            ln = present(0, 0)
            preprocess[1].left_flank = ln[0].hash if len(ln) == 1 else preprocess[1].left_anchor

        # second pass: cut each segment at its decision k-mers
        # (compactor.hh:439-532).  The new-k-mer set is final here, so the
        # presence of every window's neighbours is computed at once.
        vals = np.asarray(vals, np.uint64)
        in_set = np.isin(vals, np.fromiter(known, np.uint64, len(known))) if known else np.zeros(vals.shape, bool)
        union = np.asarray(mask, np.uint8) | (in_set.astype(np.uint8) << np.arange(8, dtype=np.uint8)).sum(
            axis=1, dtype=np.uint8)
        nl, nr = _POPCOUNT4[union & 15], _POPCOUNT4[union >> 4]
        is_decision = (nl > 1) | (nr > 1)
        for seg in preprocess:
            if seg.is_null():
                segments.append(seg)
                continue
            decisions = []
            a = seg.start_pos
            for pos in (a + np.flatnonzero(is_decision[a:a + seg.length - K + 1])).tolist():
                left, right = present(pos, 0), present(pos, 4)
                if len(left) > 1 or len(right) > 1:
                    root = s[pos:pos + K]
                    lk = [Kmer(x.hash, x.symbol + root[:K - 1]) for x in left]
                    rk = [Kmer(x.hash, root[1:] + x.symbol) for x in right]
                    decision_neighbors.append((lk, rk))
                    h = hashes[base + pos]
                    new_decision_kmers.add(h)
                    d = self._init_segment(h, h, pos)
                    d.length = K
                    d.right_anchor = d.left_anchor
                    d.right_flank = d.left_flank
                    d.is_decision_kmer = True
                    decisions.append(d)
            self._split(seg, decisions, hashes, segments)
        return segments if one_arg else None

    @staticmethod
    def _init_segment(left_anchor, left_flank, start_pos):
        seg = compact_segment()
        seg.start_pos = start_pos
        seg.left_anchor = left_anchor
        seg.left_flank = left_flank
        seg.is_decision_kmer = False
        return seg

    def _finish_segment(self, seg, end, right_anchor, right_flank, out):
        seg.length = end - seg.start_pos + self.K
        seg.right_anchor = right_anchor
        seg.right_flank = right_flank
        out.append(seg)

    def _split(self, seg, decisions, hashes, segments):
        """The cuts of compactor.hh:470-531 (hash indexes as the reference
        takes them, from the front of the caller's hash vector)."""
        K = self.K
        if not decisions:
            segments.append(seg)
            return
        if len(decisions) == 1:
            d = decisions[0]
            right = compact_segment()
            right.left_flank = d.right_anchor
            right.left_anchor = _hash_at(hashes, d.start_pos + 1)
            right.right_anchor = seg.right_anchor
            right.right_flank = seg.right_flank
            right.start_pos = d.start_pos + 1
            right.length = seg.length - (right.start_pos - seg.start_pos)
            seg.right_anchor = _hash_at(hashes, d.start_pos - 1)
            seg.right_flank = d.left_anchor
            seg.length = d.start_pos + K - 1 - seg.start_pos
            if seg.length >= K:
                segments.append(seg)
            segments.append(d)
            if right.length >= K:
                segments.append(right)
            return
        first_d, last_d = decisions[0], decisions[-1]
        first = compact_segment()
        first.left_flank = seg.left_flank
        first.left_anchor = seg.left_anchor
        first.right_anchor = _hash_at(hashes, first_d.start_pos - 1)
        first.right_flank = _hash_at(hashes, first_d.start_pos)
        first.start_pos = seg.start_pos
        first.length = first_d.start_pos + K - 1 - seg.start_pos
        if first.length >= K:
            segments.append(first)
        segments.append(first_d)
        for u, v in zip(decisions, decisions[1:]):
            mid = compact_segment()
            mid.left_flank = u.left_anchor
            mid.left_anchor = _hash_at(hashes, u.start_pos + 1)
            mid.right_anchor = _hash_at(hashes, v.start_pos - 1)
            mid.right_flank = v.left_anchor
            mid.start_pos = u.start_pos + 1
            mid.length = v.start_pos + K - 1 - mid.start_pos
            if mid.length >= K:
                segments.append(mid)
            segments.append(v)
        last = compact_segment()
        last.right_anchor = seg.right_anchor
        last.right_flank = seg.right_flank
        last.left_flank = last_d.right_anchor
        last.left_anchor = _hash_at(hashes, last_d.start_pos + 1)
        last.start_pos = last_d.start_pos + 1
        last.length = seg.length - (last.start_pos - seg.start_pos)
        if last.length >= K:
            segments.append(last)



def segments_from_view(K, sequence, counts, hash_values, nb_values, nb_present, hashes=None, new_kmers=None,
                       segments=None, new_decision_kmers=None, decision_neighbors=None):
    """find_new_segments of one read from its graph view: per window the
    query count, the hash value, the eight neighbour values (left A C G T,
    right A C G T) and the presence mask (bit c left, 4 + c right), as
    dBG.window_neighbors gives them.  Arguments and result as
    Compactor.find_new_segments."""
    return _SegmentFinder(K).run(sequence, counts, hash_values, nb_values, nb_present, hashes, new_kmers, segments,
                                 new_decision_kmers, decision_neighbors)


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

    # ------------------------------------------------------------ graph side
    def _graph_view(self, reads):
        """Per window of `reads`: the query count, the hash value, the eight
        neighbour values and their presence mask (two GPU launches)."""
        counts, hv = self.dbg.query_sequences(reads, with_hashes=True)
        _, _, vals, mask = self.dbg.window_neighbors(reads)
        return counts, hv, vals, mask

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
        counts, hv, vals, mask = self._graph_view([s])
        return segments_from_view(self.K, s, counts, hv, vals, mask, hashes, new_kmers, segments,
                                  new_decision_kmers, decision_neighbors)

    def find_new_segments_batch(self, reads, new_kmers=None):
        """find_new_segments over every read of a batch against the graph as
        it is (nothing inserted in between), with two GPU launches for the
        whole batch.  new_kmers: None for a fresh set per read (the
        one-argument overload), or one set shared by the reads in order.
        Returns per read (segments, hashes, new k-mers, new decision k-mers,
        decision neighbour pairs); reads shorter than K give None."""
        batch = _as_batch(reads)
        counts, hv, vals, mask = self._graph_view(batch)
        K = self.K
        out = []
        k = 0
        for r in range(batch.n_reads):
            a, b = int(batch.offsets[r]), int(batch.offsets[r + 1])
            if b - a < K:
                out.append(None)
                continue
            n = b - a - K + 1
            s = bytes(batch.seq[a:b]).decode()
            hs, segs, dks, pairs = [], SegmentList(), set(), []
            nk = set() if new_kmers is None else new_kmers
            segments_from_view(K, s, counts[k:k + n], hv[k:k + n], vals[k:k + n], mask[k:k + n], hs, nk, segs, dks,
                               pairs)
            out.append((segs, hs, nk, dks, pairs))
            k += n
        return out

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
