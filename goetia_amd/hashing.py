"""Hash models and Lemire shifters with goetia's names.

``Hash`` / ``Canonical`` follow canonical.hh:30-147 (value() = fw, or
min(fw, rc) for canonical); ``FwdLemireShifter`` / ``CanLemireShifter`` follow
hashshifter.hh:71-205 over rollinghashshifter.hh:52-222, with the Python
conveniences of pythonize_hashing.py:49-83 (``.value`` property, ordering and
``__hash__`` by value).  The shifter state is host scalar code in
libgoetia_amd (per-call API, as in goetia); sequence-level hashing
(``KmerIterator``, ``hash_sequence``) runs on the GPU.
"""
import ctypes as C

import numpy as np

from . import _native as N
from .reads import ReadBatch

UINT64_MAX = (1 << 64) - 1


class Hash:
    """goetia::Hash<uint64_t> (canonical.hh:30-75)."""
    __slots__ = ('hash',)

    def __init__(self, hash=UINT64_MAX):
        self.hash = int(hash)

    @property
    def value(self):
        return self.hash

    def __int__(self):
        return self.value

    def __index__(self):
        return self.value

    def __eq__(self, other):
        return self.value == _value(other)

    def __ne__(self, other):
        return self.value != _value(other)

    def __lt__(self, other):
        return self.value < _value(other)

    def __le__(self, other):
        return self.value <= _value(other)

    def __gt__(self, other):
        return self.value > _value(other)

    def __ge__(self, other):
        return self.value >= _value(other)

    def __hash__(self):
        return hash(self.value)

    def __repr__(self):
        return f'<Hash h={self.hash}>'


class Canonical(Hash):
    """goetia::Canonical<uint64_t> (canonical.hh:94-147)."""
    __slots__ = ('fw_hash', 'rc_hash')

    def __init__(self, fw_hash=UINT64_MAX, rc_hash=UINT64_MAX):
        self.fw_hash = int(fw_hash)
        self.rc_hash = int(rc_hash)

    @property
    def hash(self):
        return self.value

    @property
    def value(self):
        return self.fw_hash if self.fw_hash < self.rc_hash else self.rc_hash

    @property
    def sign(self):
        return self.fw_hash < self.rc_hash

    def __repr__(self):
        return f'<Canonical fw={self.fw_hash} rc={self.rc_hash} sign={int(self.sign)}>'


def _value(h):
    return h.value if isinstance(h, Hash) else int(h)


class DNA_SIMPLE:
    """alphabets.hh:107-161."""
    SYMBOLS = 'ACGT'
    COMPLEMENTS = 'TGCA'
    _TABLE = str.maketrans('ACGTacgt', 'TGCAtgca')

    @staticmethod
    def complement(c):
        return {'A': 'T', 'T': 'A', 'C': 'G', 'G': 'C'}.get(c, '\0')

    @staticmethod
    def reverse_complement(seq):
        return ''.join(DNA_SIMPLE.complement(c) for c in reversed(seq))

    @staticmethod
    def validate(seq):
        """DNA_SIMPLE::validate: uppercases acgt; any other byte is invalid."""
        out = seq.upper()
        if any(c not in 'ACGT' for c in out):
            raise N.InvalidSequenceException('invalid character in sequence')
        return out


def _bytes(seq):
    return seq.encode() if isinstance(seq, str) else bytes(seq)


class _StaticOrBound:
    """Lets ``Shifter.hash(seq, K)`` (static, hashshifter.hh:177-194) and
    ``shifter.hash(seq)`` (bound, :171-175) share one name."""

    def __init__(self, bound, static):
        self.bound = bound
        self.static = static

    def __get__(self, obj, objtype=None):
        if obj is None:
            return self.static.__get__(objtype, objtype)
        return self.bound.__get__(obj, objtype)


class _LemireShifter:
    canonical = False
    hash_type = Hash
    alphabet = DNA_SIMPLE

    def __init__(self, K, start=None):
        if isinstance(K, _LemireShifter):
            K = K.K
        if not 1 <= int(K) <= 255:
            raise ValueError('K must be in 1..255')
        self.K = int(K)
        self._fw = C.c_uint64(0)
        self._rc = C.c_uint64(0)
        self._initialized = False
        if start is not None:
            self.hash_base(start)

    @classmethod
    def build(cls, K, *args):
        return cls(K)

    def _make(self):
        if self.canonical:
            return Canonical(self._fw.value, self._rc.value)
        return Hash(self._fw.value)

    def is_initialized(self):
        return self._initialized

    def hash_base(self, sequence, end=None):
        """hashshifter.hh:147-170; raises on len(sequence) < K."""
        b = _bytes(sequence)
        if len(b) < self.K:
            raise N.InvalidSequenceException('HashShifter::hash_base: Sequence must at least length K')
        N.call('gt_shifter_hash_base', self.K, b, C.byref(self._fw), C.byref(self._rc))
        self._initialized = True
        return self._make()

    def _hash_bound(self, sequence):
        other = type(self)(self.K)
        return other.hash_base(sequence)

    @classmethod
    def _hash_static(cls, sequence, K):
        return cls(K).hash_base(sequence)

    hash = _StaticOrBound(_hash_bound, _hash_static)

    def get(self):
        return self._make()

    def shift_right(self, out, in_):
        """rollinghashshifter.hh:200-208 shift_right(out, in)."""
        if not self._initialized:
            raise N.UninitializedShifterException('shifter used before hash_base')
        N.call('gt_shifter_shift_right', self.K, C.byref(self._fw), C.byref(self._rc),
               _bytes(out)[0], _bytes(in_)[0])
        return self._make()

    def shift_left(self, in_, out):
        """rollinghashshifter.hh:211-219 shift_left(in, out)."""
        if not self._initialized:
            raise N.UninitializedShifterException('shifter used before hash_base')
        N.call('gt_shifter_shift_left', self.K, C.byref(self._fw), C.byref(self._rc),
               _bytes(in_)[0], _bytes(out)[0])
        return self._make()

    def __repr__(self):
        return f'<{type(self).__name__} K={self.K}>'


class FwdLemireShifter(_LemireShifter):
    """HashShifter<LemireShifterPolicy<Hash<uint64_t>>> (hashshifter.hh:204)."""
    canonical = False
    hash_type = Hash


class CanLemireShifter(_LemireShifter):
    """HashShifter<LemireShifterPolicy<Canonical<uint64_t>>> (hashshifter.hh:205)."""
    canonical = True
    hash_type = Canonical


def hash_batch(batch, K, mode=N.MODE_RAW):
    """fw, rc and validity of every window of a host ReadBatch, hashed on the
    GPU (packed per read)."""
    n = int(batch.kmer_counts(K).sum())
    fw = np.zeros(max(n, 1), np.uint64)
    rc = np.zeros(max(n, 1), np.uint64)
    valid = np.zeros(max(n, 1), np.uint8)
    N.call('gt_hash_reads', K, mode, N.ptr(batch.seq), N.ptr(batch.offsets), batch.n_reads,
           N.ptr(fw), N.ptr(rc), N.ptr(valid))
    return fw[:n], rc[:n], valid[:n].astype(bool)


class _KmerIteratorType:
    """KmerIterator<Shifter> (kmeriterator.hh:32-141); the hashes of the
    whole sequence are computed on the GPU when the iterator is built."""

    def __init__(self, shifter_type):
        self.shifter_type = shifter_type

    def __call__(self, seq, K_or_shifter):
        return KmerIterator(seq, K_or_shifter, self.shifter_type)


class KmerIterator:
    def __init__(self, seq, K_or_shifter, shifter_type=None):
        if isinstance(K_or_shifter, _LemireShifter):
            K = K_or_shifter.K
            shifter_type = type(K_or_shifter)
        else:
            K = int(K_or_shifter)
        self.shifter_type = shifter_type or CanLemireShifter
        self.K = K
        b = _bytes(seq)
        if len(b) < K:
            raise N.SequenceLengthException('Sequence must have length >= K')
        fw, rc, _ = hash_batch(ReadBatch.from_strings([b]), K)
        self._fw = fw
        self._rc = rc
        self.index = 0
        self.length = len(b)

    def _at(self, i):
        if self.shifter_type.canonical:
            return Canonical(int(self._fw[i]), int(self._rc[i]))
        return Hash(int(self._fw[i]))

    def first(self):
        self.index = 1
        return self._at(0)

    def next(self):
        if self.done():
            return self._at(self.index - 1)
        h = self._at(self.index)
        self.index += 1
        return h

    def done(self):
        return self.index + self.K > self.length

    def get_start_pos(self):
        return max(self.index - 1, 0)

    def get_end_pos(self):
        return self.index + self.K - 1 if self.index else self.K

    def __iter__(self):
        while not self.done():
            yield self.next()


class _KmerIteratorFactory:
    """``KmerIterator[ShifterType](seq, K)`` as in the cppyy surface."""

    def __getitem__(self, shifter_type):
        return _KmerIteratorType(shifter_type)

    def __call__(self, seq, K_or_shifter):
        return KmerIterator(seq, K_or_shifter)


KmerIteratorT = _KmerIteratorFactory()

typenames = [(FwdLemireShifter, 'FwdLemireShifter'), (CanLemireShifter, 'CanLemireShifter')]
types = [t for t, _ in typenames]
StrandAware = Hash
