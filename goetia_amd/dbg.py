"""dBG<Storage, Shifter>: goetia's de Bruijn graph over a device storage.

Mirrors dbg.hh:39-438 (sequence API), the neighbour/degree functions of
unitig_walker.hh:322-418,1003-1117 and the pythonizor of
pythonize_dbg.py:4-63 (``add``, ``get``, ``left_degree``, ``right_degree``,
``hashes``, ``shallow_clone``).  Every sequence-level and storage operation
runs on the GPU through libgoetia_amd; batched forms (``insert_sequences``,
``query_sequences``, ``degrees`` and their ``*_device`` variants over reads
resident in HBM) sit beside the per-read calls.
"""
import ctypes as C

import numpy as np

from . import _native as N
from .hashing import CanLemireShifter, DNA_SIMPLE, Hash, KmerIterator, _LemireShifter, _bytes
from .reads import DeviceReads, ReadBatch
from .storage import _Storage


class Shift:
    """goetia::Shift<hash_type, Direction> (canonical.hh:357-412)."""
    __slots__ = ('hash', 'symbol', 'direction')

    def __init__(self, hash, symbol, direction):
        self.hash = hash
        self.symbol = symbol
        self.direction = direction

    @property
    def value(self):
        return self.hash.value

    def __eq__(self, other):
        return self.value == (other.value if hasattr(other, 'value') else int(other))

    def __hash__(self):
        return hash(self.value)

    def __repr__(self):
        d = 'R' if self.direction else 'L'
        return f'<Shift {d} {self.symbol} {self.hash!r}>'


class Kmer:
    """goetia::Kmer<hash_type> (hash + k-mer string)."""
    __slots__ = ('hash', 'kmer')

    def __init__(self, hash, kmer):
        self.hash = hash
        self.kmer = kmer

    @property
    def value(self):
        return self.hash.value

    def __repr__(self):
        return f'<Kmer {self.kmer} {self.hash!r}>'


def _as_batch(reads):
    if isinstance(reads, (ReadBatch, DeviceReads)):
        return reads
    if isinstance(reads, (str, bytes)):
        return ReadBatch.from_strings([reads])
    return ReadBatch.from_strings(list(reads))


def _stream_ptr(stream):
    if stream is None:
        try:
            import torch
            return torch.cuda.current_stream().cuda_stream
        except Exception:
            return None
    return getattr(stream, 'cuda_stream', stream)


class _dBGMeta(type):
    _cache = {}

    def __getitem__(cls, params):
        storage_type, shifter_type = params
        key = (storage_type, shifter_type)
        if key not in cls._cache:
            name = f'dBG[{storage_type.__name__},{shifter_type.__name__}]'
            cls._cache[key] = type(name, (cls,), {'storage_type': storage_type,
                                                  'shifter_type': shifter_type})
        return cls._cache[key]


class dBG(metaclass=_dBGMeta):
    storage_type = None
    shifter_type = None

    def __init__(self, storage, K_or_shifter, mode=N.MODE_RAW):
        if not isinstance(storage, _Storage):
            raise TypeError('storage must be a goetia_amd storage')
        shifter_type = self.shifter_type
        if isinstance(K_or_shifter, _LemireShifter):
            shifter_type = shifter_type or type(K_or_shifter)
            K = K_or_shifter.K
        else:
            K = int(K_or_shifter)
        self.shifter_type = shifter_type or CanLemireShifter
        self.storage_type = self.storage_type or type(storage)
        if not 1 <= K <= 255:
            raise ValueError('K must be in 1..255')
        self.K = K
        self.S = storage
        self.mode = mode
        self._shifter = self.shifter_type(K)
        self._cursor = None
        self.alphabet = DNA_SIMPLE
        self.hash_type = self.shifter_type.hash_type

    # ------------------------------------------------------------ build
    @classmethod
    def build(cls, storage, hasher, *args):
        """dbg.hh:111-121 (shared_ptr build)."""
        return cls(storage, hasher)

    def clone(self):
        """dbg.hh:128-131: same K, new empty storage of the same sizes."""
        return type(self)(self.S.clone(), self.K, self.mode)

    def shallow_clone(self):
        return self.clone()

    @property
    def canonical(self):
        return bool(self.shifter_type.canonical)

    # ------------------------------------------------------------ hashing
    def hash(self, kmer):
        return self._shifter.hash(kmer)

    def get_hasher(self):
        return self.shifter_type(self.K)

    def get_hash_iter(self, sequence):
        return KmerIterator(sequence, self.K, self.shifter_type)

    def hashes(self, sequence):
        """pythonize_dbg.py:24-28"""
        it = self.get_hash_iter(sequence)
        while not it.done():
            yield it.next()

    def _value(self, item):
        if isinstance(item, (str, bytes)):
            if len(item) < self.K:
                raise ValueError('k-mer shorter than K')
            return self.hash(item).value
        if isinstance(item, (Hash, Shift, Kmer)):
            return item.value
        return int(item)

    # ------------------------------------------------------------ per k-mer
    def insert(self, item):
        """dbg.hh:140-146: True if the k-mer was new."""
        return self.S.insert(self._value(item))

    def insert_and_query(self, item):
        """dbg.hh:156-162: post-insertion count."""
        return self.S.insert_and_query(self._value(item))

    def query(self, item):
        """dbg.hh:171-177."""
        return self.S.query(self._value(item))

    def add(self, item):
        """pythonize_dbg.py:8-14."""
        if not isinstance(item, int) and not isinstance(item, Hash) and len(item) < self.K:
            raise ValueError()
        if isinstance(item, (int, Hash)) or len(item) == self.K:
            return self.insert(item)
        return self.insert_sequence(item)

    def get(self, item):
        return self.query(item)

    def n_unique(self):
        return self.S.n_unique_kmers()

    def n_occupied(self):
        return self.S.n_occupied()

    def get_raw(self):
        return self.S.get_raw_tables()

    def estimated_fp(self):
        if not self.S.is_probabilistic:
            raise TypeError('estimated_fp is defined for probabilistic storages')
        return self.S.estimated_fp()

    def reset(self):
        self.S.reset()

    def suffix(self, kmer):
        return kmer[len(kmer) - self.K + 1:]

    def prefix(self, kmer):
        return kmer[:self.K - 1]

    def save(self, filename):
        self.S.save(filename, self.K)

    def load(self, filename):
        self.S.load(filename)

    # ------------------------------------------------------------ per read
    def _check_len(self, sequence):
        if len(sequence) < self.K:
            raise N.SequenceLengthException('Sequence must have length >= K')

    def insert_sequence(self, sequence, *outs):
        """dbg.hh:249-318.  Returns len - K + 1.  Overloads by extra
        arguments: (hashes, counts) lists get the hashes and post-insert
        counts appended; a set gets the new k-mers; a list gets the hashes."""
        self._check_len(sequence)
        n = len(sequence) - self.K + 1
        if len(outs) == 2:
            hashes, counts = outs
            cnt = self.insert_and_query_sequence(sequence)
            hs = list(self.hashes(sequence))
            hashes.extend(hs)
            counts.extend(int(c) for c in cnt)
            return n
        batch = ReadBatch.from_strings([sequence])
        want_flags = len(outs) == 1
        n_new, flags = self._insert_batch(batch, want_flags)
        if len(outs) == 1:
            out = outs[0]
            hs = list(self.hashes(sequence))
            if isinstance(out, set):
                for h, f in zip(hs, flags):
                    if f:
                        out.add(h)
            else:
                out.extend(hs)
        return n

    def insert_sequence_n_new(self, sequence):
        """dbg.hh:307-318 insert_sequence(seq, n_new): (len - K + 1, n_new)."""
        self._check_len(sequence)
        n_new, _ = self._insert_batch(ReadBatch.from_strings([sequence]), False)
        return len(sequence) - self.K + 1, int(n_new[0])

    def insert_and_query_sequence(self, sequence):
        """dbg.hh:327-340."""
        self._check_len(sequence)
        return list(self.insert_and_query_sequences([sequence]))

    def query_sequence(self, sequence, counts=None, hashes=None, new_hashes=None):
        """dbg.hh:349-394."""
        self._check_len(sequence)
        c, h = self.query_sequences([sequence], with_hashes=True)
        if counts is None and hashes is None:
            return [int(v) for v in c]
        hs = list(self.hashes(sequence))
        if counts is not None:
            counts.extend(int(v) for v in c)
        if hashes is not None:
            hashes.extend(hs)
        if new_hashes is not None:
            for v, hh in zip(c, hs):
                if v == 0:
                    new_hashes.add(hh)
        return None

    # ------------------------------------------------------------ batches
    def _insert_batch(self, batch, want_flags):
        nk = int(batch.kmer_counts(self.K).sum())
        n_new = np.zeros(max(batch.n_reads, 1), np.uint64)
        flags = np.zeros(max(nk, 1), np.uint8) if want_flags else None
        n_kmers = C.c_uint64()
        N.call('gt_dbg_insert', self.S.handle, self.K, int(self.canonical), self.mode,
               N.ptr(batch.seq), N.ptr(batch.offsets), batch.n_reads, N.ptr(n_new),
               N.ptr(flags), C.byref(n_kmers))
        return n_new[:batch.n_reads], (flags[:nk].astype(bool) if want_flags else None)

    def insert_sequences(self, reads, new_flags=False):
        """Batched insert_sequence(seq, n_new) over reads in order: per-read
        n_new (and per-k-mer new flags when asked), exactly as the sequential
        reference."""
        batch = _as_batch(reads)
        n_new, flags = self._insert_batch(batch, new_flags)
        return (n_new, flags) if new_flags else n_new

    def query_sequences(self, reads, with_hashes=False):
        """Batched query_sequence: counts packed per read (-1 = skipped)."""
        batch = _as_batch(reads)
        nk = int(batch.kmer_counts(self.K).sum())
        counts = np.zeros(max(nk, 1), np.int16)
        hashes = np.zeros(max(nk, 1), np.uint64) if with_hashes else None
        N.call('gt_dbg_query', self.S.handle, self.K, int(self.canonical), self.mode,
               N.ptr(batch.seq), N.ptr(batch.offsets), batch.n_reads, N.ptr(counts), N.ptr(hashes))
        if with_hashes:
            return counts[:nk], hashes[:nk]
        return counts[:nk]

    def insert_and_query_sequences(self, reads):
        batch = _as_batch(reads)
        nk = int(batch.kmer_counts(self.K).sum())
        counts = np.zeros(max(nk, 1), np.int16)
        N.call('gt_dbg_insert_and_query', self.S.handle, self.K, int(self.canonical), self.mode,
               N.ptr(batch.seq), N.ptr(batch.offsets), batch.n_reads, N.ptr(counts))
        return counts[:nk]

    def degrees(self, reads, min_count=1, only=None):
        """In/out degree of every window (0xFF where skipped or not in
        `only`)."""
        batch = _as_batch(reads)
        nk = int(batch.kmer_counts(self.K).sum())
        ind = np.zeros(max(nk, 1), np.uint8)
        outd = np.zeros(max(nk, 1), np.uint8)
        only_arr = None if only is None else np.ascontiguousarray(only, np.uint8)
        N.call('gt_dbg_degrees', self.S.handle, self.K, int(self.canonical), self.mode,
               int(min_count), N.ptr(batch.seq), N.ptr(batch.offsets), batch.n_reads,
               N.ptr(only_arr), N.ptr(ind), N.ptr(outd))
        return ind[:nk], outd[:nk]

    def window_neighbors(self, reads, min_count=1):
        """(in degree, out degree, neighbour values, presence mask) of every
        window: values[i, c] is the left neighbour with symbol "ACGT"[c],
        values[i, 4 + c] the right one (left_extensions / right_extensions,
        hashextender.hh:352-364, 386-396); bit c / 4 + c of mask[i] is set
        when its query is >= min_count."""
        batch = _as_batch(reads)
        nk = int(batch.kmer_counts(self.K).sum())
        ind = np.zeros(max(nk, 1), np.uint8)
        outd = np.zeros(max(nk, 1), np.uint8)
        vals = np.zeros((max(nk, 1), 8), np.uint64)
        mask = np.zeros(max(nk, 1), np.uint8)
        N.call('gt_dbg_neighbors', self.S.handle, self.K, int(self.canonical), self.mode,
               int(min_count), N.ptr(batch.seq), N.ptr(batch.offsets), batch.n_reads,
               N.ptr(ind), N.ptr(outd), N.ptr(vals), N.ptr(mask))
        return ind[:nk], outd[:nk], vals[:nk], mask[:nk]

    # device-resident batches (torch tensors; asynchronous on `stream`)
    def insert_device(self, reads, n_new=None, new_flags=None, stream=None):
        N.call('gt_dbg_insert_dev', self.S.handle, self.K, int(self.canonical), self.mode,
               N.ptr(reads.seq), N.ptr(reads.offsets), reads.n_reads, reads.n_bases,
               N.ptr(n_new), N.ptr(new_flags), _stream_ptr(stream))

    def query_device(self, reads, counts, hashes=None, stream=None):
        N.call('gt_dbg_query_dev', self.S.handle, self.K, int(self.canonical), self.mode,
               N.ptr(reads.seq), N.ptr(reads.offsets), reads.n_reads, reads.n_bases,
               N.ptr(counts), N.ptr(hashes), _stream_ptr(stream))

    def insert_and_query_device(self, reads, counts, stream=None):
        N.call('gt_dbg_insert_and_query_dev', self.S.handle, self.K, int(self.canonical),
               self.mode, N.ptr(reads.seq), N.ptr(reads.offsets), reads.n_reads, reads.n_bases,
               N.ptr(counts), _stream_ptr(stream))

    def degrees_device(self, reads, in_deg, out_deg, min_count=1, only=None, stream=None):
        N.call('gt_dbg_degrees_dev', self.S.handle, self.K, int(self.canonical), self.mode,
               int(min_count), N.ptr(reads.seq), N.ptr(reads.offsets), reads.n_reads,
               reads.n_bases, N.ptr(only), N.ptr(in_deg), N.ptr(out_deg), _stream_ptr(stream))

    # ------------------------------------------------------------ neighbours
    def set_cursor(self, kmer):
        """hashextender.hh:198-215."""
        if len(kmer) < self.K:
            raise N.InvalidSequenceException('Sequence must at least length K')
        self._cursor = kmer[:self.K] if isinstance(kmer, str) else _bytes(kmer)[:self.K].decode()
        return self._shifter.hash_base(self._cursor)

    def get_cursor(self):
        if self._cursor is None:
            raise N.UninitializedShifterException('cursor not set')
        return self._cursor

    def get_hash(self):
        return self._shifter.get()

    def shift_right(self, c):
        """hashextender.hh:349-353: the cursor moves one base right (c
        enters, the first base leaves); the new hash."""
        cur = self.get_cursor()
        h = self._shifter.shift_right(cur[0], c)
        self._cursor = cur[1:] + c
        return h

    def shift_left(self, c):
        """hashextender.hh:327-331: the cursor moves one base left."""
        cur = self.get_cursor()
        h = self._shifter.shift_left(c, cur[-1])
        self._cursor = c + cur[:-1]
        return h

    def _extensions(self, right):
        cur = self.get_cursor()
        sh = self._shifter
        out = []
        for c in DNA_SIMPLE.SYMBOLS:
            if right:   # hashextender.hh:386-396
                h = sh.shift_right(cur[0], c)
                out.append(Shift(h, c, True))
                sh.shift_left(cur[0], c)
            else:       # hashextender.hh:352-364
                h = sh.shift_left(c, cur[-1])
                out.append(Shift(h, c, False))
                sh.shift_right(c, cur[-1])
        return out

    def left_extensions(self):
        return self._extensions(False)

    def right_extensions(self):
        return self._extensions(True)

    def _present(self, exts, min_count=1, extras=None):
        counts = self.S.query_hashes([e.value for e in exts])
        known = () if extras is None else {x.value if hasattr(x, 'value') else int(x) for x in extras}
        return [e for e, c in zip(exts, counts) if c >= min_count or e.value in known]

    def count_nodes(self, exts, extras=None):
        """unitig_walker.hh:361-397; with extras, a node is also present when
        it is among them."""
        return len(self._present(exts, extras=extras))

    def filter_nodes(self, exts, extras=None):
        """unitig_walker.hh:407-418, 457-468."""
        return self._present(exts, extras=extras)

    def in_degree(self, extras=None):
        return self.count_nodes(self.left_extensions(), extras)

    def out_degree(self, extras=None):
        return self.count_nodes(self.right_extensions(), extras)

    def degree(self, extras=None):
        return self.in_degree(extras) + self.out_degree(extras)

    def left_degree(self, kmer):
        """pythonize_dbg.py:30-32."""
        self.set_cursor(kmer)
        return self.in_degree()

    def right_degree(self, kmer):
        self.set_cursor(kmer)
        return self.out_degree()

    def in_neighbors(self, seed=None):
        if seed is not None:
            self.set_cursor(seed)
        return self.filter_nodes(self.left_extensions())

    def out_neighbors(self, seed=None):
        if seed is not None:
            self.set_cursor(seed)
        return self.filter_nodes(self.right_extensions())

    def neighbors(self, seed=None):
        if seed is not None:
            self.set_cursor(seed)
        return self.in_neighbors(), self.out_neighbors()

    def is_decision_kmer(self, node=None):
        """unitig_walker.hh:1003-1023."""
        if node is not None:
            self.set_cursor(node)
        return self.in_degree() > 1 or self.out_degree() > 1

    def build_left_kmers(self, nodes, root):
        """unitig_walker.hh:612-621: the left neighbours' k-mers,
        symbol + prefix(root)."""
        prefix = self.prefix(root)
        return [Kmer(n.hash, n.symbol + prefix) for n in nodes]

    def build_right_kmers(self, nodes, root):
        """unitig_walker.hh:631-640: suffix(root) + symbol."""
        suffix = self.suffix(root)
        return [Kmer(n.hash, suffix + n.symbol) for n in nodes]

    def get_decision_neighbors(self, root=None, union_nodes=None):
        """unitig_walker.hh:1047-1117: (is decision, (left k-mers, right
        k-mers)) of the cursor (or root), with union_nodes counted as present
        when given; the pair is empty unless it is a decision k-mer, as the
        reference leaves `result` untouched then."""
        if root is not None:
            self.set_cursor(root)
        root = self.get_cursor()
        left = self.build_left_kmers(self.filter_nodes(self.left_extensions(), union_nodes), root)
        right = self.build_right_kmers(self.filter_nodes(self.right_extensions(), union_nodes), root)
        if len(left) > 1 or len(right) > 1:
            return True, (left, right)
        return False, ([], [])

    def find_decision_kmers(self, sequence, min_count=1, neighbors=False):
        """unitig_walker.hh:1025-1045: positions and hashes of the windows of
        `sequence` with in or out degree > 1, plus their (in, out) degrees
        (from the GPU degree kernel); with neighbors=True also each decision
        k-mer's (left k-mers, right k-mers), as get_decision_neighbors gives
        them (count >= 1)."""
        self._check_len(sequence)
        ind, outd = self.degrees([sequence], min_count=min_count)
        hs = list(self.hashes(sequence))
        pos = [i for i in range(len(ind)) if ind[i] != 0xFF and (ind[i] > 1 or outd[i] > 1)]
        out = (pos, [hs[i] for i in pos], [(int(ind[i]), int(outd[i])) for i in pos])
        if not neighbors:
            return out
        s = sequence if isinstance(sequence, str) else _bytes(sequence).decode()
        return out + ([self.get_decision_neighbors(s[i:i + self.K])[1] for i in pos],)


