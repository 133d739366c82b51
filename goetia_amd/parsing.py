"""FASTA / FASTQ input: goetia's FastxParser<DNA_SIMPLE> (readers.hh:95-232)
over the native parser (gt_fastx_*).

Records come out validated and uppercased, exactly as goetia's parser hands
them to a dBG; whole batches come out in the ReadBatch layout every batched
dBG call takes, so a file goes to the GPU without a per-read Python step.
"""
import ctypes as C

import numpy as np

from . import _native as N
from .reads import ReadBatch


class NoMoreReadsAvailable(N.GoetiaException):
    """goetia::NoMoreReadsAvailable (readers.hh:150-152)."""


class Record:
    """goetia::Record (parsing.hh:25-57): name, validated sequence and
    quality (None for FASTA)."""
    __slots__ = ('name', 'sequence', 'quality')

    def __init__(self, name, sequence, quality=None):
        self.name = name
        self.sequence = sequence
        self.quality = quality

    def write_fastx(self, out):
        """parsing.hh:38-49: FASTQ when the record has a quality, else FASTA."""
        out.write(fastx_text(self.name, self.sequence, self.quality))

    def __repr__(self):
        return f'Record(name={self.name!r}, sequence={self.sequence!r})'


def fastx_text(name, sequence, quality=None):
    """The text Record::write_fastx writes (parsing.hh:38-49)."""
    if quality:
        return f'@{name}\n{sequence}\n+\n{quality}\n'
    return f'>{name}\n{sequence}\n'


class FastxParser:
    """goetia::FastxParser<DNA_SIMPLE>.  Skipped records (invalid characters,
    shorter than min_length, FASTQ quality of another length) are counted in
    n_skipped and never returned; a strict parser raises
    InvalidCharacterException at the first invalid record, after the records
    before it."""

    def __init__(self, filename, strict=False, min_length=0, batch_reads=4096, batch_bases=1 << 20):
        self.filename = str(filename)
        self._h = N.vp()
        N.call('gt_fastx_open', self.filename.encode(), int(bool(strict)), int(min_length), C.byref(self._h))
        self._batch_reads = batch_reads
        self._batch_bases = batch_bases
        self._records = []
        self._pos = 0
        self._error = None
        self._done = False

    @classmethod
    def build(cls, filename, strict=False, min_length=0):
        return cls(filename, strict, min_length)

    def close(self):
        if getattr(self, '_h', None) and self._h.value:
            N.load().gt_fastx_close(self._h)
            self._h = N.vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------ batches
    def next_batch(self, max_reads=1 << 16, max_bases=64 << 20, names=False, quality=False):
        """Up to max_reads records / max_bases bases as a ReadBatch (empty at
        the end of the input); with names=True also their names, with
        quality=True their names and qualities (bytes, or None for a record
        without one)."""
        if self._error is not None:
            err, self._error = self._error, None
            raise err
        names = names or quality
        seq = np.empty(max(max_bases, 1), np.uint8)
        off = np.zeros(max_reads + 1, np.int64)
        cap = max_reads * 256 if names else 0
        nb = C.create_string_buffer(cap) if names else None
        noff = np.zeros(max_reads + 1, np.int64) if names else None
        qual = np.empty(max(max_bases, 1), np.uint8) if quality else None
        hq = np.zeros(max(max_reads, 1), np.uint8) if quality else None
        n = C.c_uint64()
        rc = N.load().gt_fastx_next_batch_qual(self._h, max_reads, max_bases, N.ptr(seq), N.ptr(off),
                                               N.ptr(qual) if quality else None, N.ptr(hq) if quality else None,
                                               nb, cap, N.ptr(noff), C.byref(n))
        k = n.value
        if rc != N.GT_OK:
            if k == 0:
                N.check(rc, 'gt_fastx_next_batch')
            try:
                N.check(rc, 'gt_fastx_next_batch')
            except N.GoetiaException as e:
                self._error = e
        batch = ReadBatch(seq[:int(off[k])].copy(), off[:k + 1].copy())
        if names:
            raw = nb.raw
            nm = [raw[noff[i]:noff[i + 1]].decode() for i in range(k)]
            if quality:
                qs = [qual[off[i]:off[i + 1]].tobytes() if hq[i] else None for i in range(k)]
                return batch, nm, qs
            return batch, nm
        return batch

    # ------------------------------------------------------------ records
    def next(self):
        """The next record, or None once the input is exhausted (after which
        another call raises NoMoreReadsAvailable, as goetia's does)."""
        if self._pos >= len(self._records):
            if self._done:
                raise NoMoreReadsAvailable('No more reads available')
            batch, nm, qs = self.next_batch(self._batch_reads, self._batch_bases, quality=True)
            if batch.n_reads == 0:
                self._done = True
                return None
            self._records = [Record(nm[i], batch.read(i).decode(), qs[i].decode() if qs[i] else None)
                             for i in range(batch.n_reads)]
            self._pos = 0
        rec = self._records[self._pos]
        self._pos += 1
        return rec

    def __iter__(self):
        while True:
            if self._done and self._pos >= len(self._records):
                return
            rec = self.next()
            if rec is None:
                return
            yield rec

    # ------------------------------------------------------------ counters
    def _stats(self):
        parsed, skipped, invalid = C.c_uint64(), C.c_uint64(), C.c_uint64()
        complete = C.c_int()
        N.call('gt_fastx_stats', self._h, C.byref(parsed), C.byref(skipped), C.byref(invalid), C.byref(complete))
        return parsed.value, skipped.value, invalid.value, bool(complete.value)

    def n_parsed(self):
        return self._stats()[0]

    def n_skipped(self):
        return self._stats()[1]

    def n_invalid_reads(self):
        return self._stats()[2]

    def is_complete(self):
        return self._stats()[3] and self._pos >= len(self._records)


class SplitPairedReader:
    """goetia::SplitPairedReader<FastxParser<DNA_SIMPLE>> (readers.hh:235-334)
    over the native paired reader (gt_pairs_*): next() gives (left, right)
    Records, None for an empty side; a pair that a strict reader rejects
    raises InvalidCharacterException; unpaired names (force_name_match) give
    (None, None) and count two skipped, or raise GoetiaException when strict;
    files of different lengths raise GoetiaException."""

    def __init__(self, left, right, strict=False, min_length=0, force_name_match=False):
        self.left_filename, self.right_filename = str(left), str(right)
        self._h = N.vp()
        N.call('gt_pairs_open', self.left_filename.encode(), self.right_filename.encode(), int(bool(strict)),
               int(min_length), int(bool(force_name_match)), C.byref(self._h))
        self._buf = np.empty(1 << 20, np.uint8)
        self._names = C.create_string_buffer(1 << 16)

    @classmethod
    def build(cls, left, right, strict=False, min_length=0, force_name_match=False):
        return cls(left, right, strict, min_length, force_name_match)

    def close(self):
        if getattr(self, '_h', None) and self._h.value:
            N.load().gt_pairs_close(self._h)
            self._h = N.vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def next(self):
        if self.is_complete():
            raise NoMoreReadsAvailable('No more reads available')
        lens = np.zeros(2, np.int64)
        nlens = np.zeros(2, np.int64)
        done = C.c_int()
        N.call('gt_pairs_next', self._h, N.ptr(self._buf), len(self._buf), N.ptr(lens), self._names,
               len(self._names), N.ptr(nlens), C.byref(done))
        if done.value:
            return None, None
        out, at, nat = [], 0, 0
        raw = self._names.raw
        for k in range(2):
            if lens[k] < 0:
                out.append(None)
                continue
            seq = self._buf[at:at + lens[k]].tobytes().decode()
            name = raw[nat:nat + nlens[k]].decode()
            at += int(lens[k])
            nat += int(nlens[k])
            out.append(Record(name, seq))
        return tuple(out)

    def _stats(self):
        parsed, skipped, complete = C.c_uint64(), C.c_uint64(), C.c_int()
        N.call('gt_pairs_stats', self._h, C.byref(parsed), C.byref(skipped), C.byref(complete))
        return parsed.value, skipped.value, bool(complete.value)

    def n_parsed(self):
        return self._stats()[0]

    def n_skipped(self):
        return self._stats()[1]

    def is_complete(self):
        return self._stats()[2]
