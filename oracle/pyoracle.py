"""ctypes binding of oracle.c (the C restatement of goetia's dBG hot path).

TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'liboracle.so')

MODE_RAW = 0
MODE_SKIP_N = 1
BIT, BYTE, EXACT = 1, 2, 3
MAX_TABLES = 16


class _Storage(C.Structure):
    _fields_ = [('kind', C.c_int), ('n_tables', C.c_uint32),
                ('sizes', C.c_uint64 * MAX_TABLES),
                ('tables', C.POINTER(C.c_uint8) * MAX_TABLES),
                ('n_unique', C.c_uint64), ('n_occupied', C.c_uint64),
                ('keys', C.POINTER(C.c_uint64)), ('capacity', C.c_uint64),
                ('count', C.c_uint64), ('has_ones_key', C.c_int),
                ('use_bigcount', C.c_int), ('big_keys', C.POINTER(C.c_uint64)),
                ('big_vals', C.POINTER(C.c_int16)), ('big_cap', C.c_uint64), ('big_count', C.c_uint64),
                ('big_has_ones', C.c_int), ('big_ones_val', C.c_int16)]


def build():
    subprocess.run(['make', '-s', '-C', HERE, 'all'], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        u64p = C.POINTER(C.c_uint64)
        i64p = C.POINTER(C.c_int64)
        i16p = C.POINTER(C.c_int16)
        u8p = C.POINTER(C.c_uint8)
        sp = C.POINTER(_Storage)
        sigs = {
            'or_table_value': (C.c_uint64, [C.c_int]),
            'or_hash_kmer': (None, [C.c_char_p, C.c_int, u64p, u64p]),
            'or_hash_read': (C.c_size_t, [C.c_char_p, C.c_size_t, C.c_int, u64p, u64p]),
            'or_primes_near_x': (C.c_uint32, [C.c_uint32, C.c_uint64, u64p]),
            'or_storage_init': (C.c_int, [sp, C.c_int, u64p, C.c_uint32]),
            'or_storage_free': (None, [sp]),
            'or_table_bytes': (C.c_uint64, [sp, C.c_uint32]),
            'or_insert': (C.c_int, [sp, C.c_uint64]),
            'or_query': (C.c_int16, [sp, C.c_uint64]),
            'or_insert_and_query': (C.c_int16, [sp, C.c_uint64]),
            'or_insert_sequence': (C.c_uint64, [sp, C.c_int, C.c_int, C.c_int, C.c_char_p,
                                                C.c_size_t, u64p, u8p]),
            'or_insert_and_query_sequence': (C.c_uint64, [sp, C.c_int, C.c_int, C.c_int,
                                                          C.c_char_p, C.c_size_t, i16p]),
            'or_query_sequence': (C.c_uint64, [sp, C.c_int, C.c_int, C.c_int, C.c_char_p,
                                               C.c_size_t, i16p, u64p]),
            'or_degrees_sequence': (C.c_uint64, [sp, C.c_int, C.c_int, C.c_int, C.c_int,
                                                 C.c_char_p, C.c_size_t, u8p, u8p]),
            'or_kmer_degrees': (None, [sp, C.c_int, C.c_int, C.c_int, C.c_char_p, u8p, u8p,
                                       u64p, u64p]),
            'or_insert_reads': (C.c_uint64, [sp, C.c_int, C.c_int, C.c_int, C.c_char_p, i64p,
                                             C.c_uint64, u64p]),
            'or_query_reads': (C.c_uint64, [sp, C.c_int, C.c_int, C.c_int, C.c_char_p, i64p,
                                            C.c_uint64, i16p]),
            'or_exact_keys': (C.c_uint64, [sp, u64p, C.c_uint64]),
            'or_set_use_bigcount': (None, [sp, C.c_int]),
            'or_bigcount_get': (C.c_int, [sp, C.c_uint64, i16p]),
            'or_bigcount_set': (None, [sp, C.c_uint64, C.c_int16]),
            'or_bigcount_pairs': (C.c_uint64, [sp, u64p, i16p, C.c_uint64]),
            'or_insert_reads_mt': (C.c_uint64, [sp, C.c_int, C.c_int, C.c_int, C.c_char_p, i64p, C.c_uint64,
                                                C.c_int]),
            'or_query_reads_mt': (C.c_uint64, [sp, C.c_int, C.c_int, C.c_int, C.c_char_p, i64p, C.c_uint64,
                                               i16p, C.c_int]),
            'or_diginorm_filter': (C.c_int, [sp, C.c_int, C.c_int, C.c_int, C.c_uint, C.c_char_p,
                                             C.c_size_t]),
            'or_diginorm_reads': (C.c_uint64, [sp, C.c_int, C.c_int, C.c_int, C.c_uint, C.c_char_p,
                                               i64p, C.c_uint64, u8p]),
            'or_insert_reads_flags': (C.c_uint64, [sp, C.c_int, C.c_int, C.c_int, C.c_char_p, i64p,
                                                   C.c_uint64, u64p, u8p]),
            'or_insert_and_query_reads': (C.c_uint64, [sp, C.c_int, C.c_int, C.c_int, C.c_char_p, i64p,
                                                       C.c_uint64, i16p]),
            'or_hash_reads': (C.c_uint64, [C.c_int, C.c_int, C.c_int, C.c_char_p, i64p, C.c_uint64, u64p]),
            'or_degrees_reads_mt': (C.c_uint64, [sp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_char_p, i64p,
                                                 C.c_uint64, u8p, u8p, u8p, C.c_int]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(arr, ctype):
    return arr.ctypes.data_as(C.POINTER(ctype))


def table_value(c):
    return lib().or_table_value(c)


def hash_kmer(kmer):
    b = kmer.encode() if isinstance(kmer, str) else bytes(kmer)
    fw, rc = C.c_uint64(), C.c_uint64()
    lib().or_hash_kmer(b, len(b), C.byref(fw), C.byref(rc))
    return fw.value, rc.value


def hash_read(seq, K):
    b = seq.encode() if isinstance(seq, str) else bytes(seq)
    n = max(0, len(b) - K + 1)
    fw = np.zeros(max(n, 1), np.uint64)
    rc = np.zeros(max(n, 1), np.uint64)
    lib().or_hash_read(b, len(b), K, _p(fw, C.c_uint64), _p(rc, C.c_uint64))
    return fw[:n], rc[:n]


def _n_windows(offsets, K):
    return int(np.maximum(np.diff(offsets) - K + 1, 0).sum())


def hash_reads(seq, offsets, K, canonical=True, mode=MODE_RAW):
    """The value of every window, packed per read (0 where skipped)."""
    seq = np.ascontiguousarray(seq, np.uint8)
    offsets = np.ascontiguousarray(offsets, np.int64)
    n = _n_windows(offsets, K)
    out = np.zeros(max(n, 1), np.uint64)
    lib().or_hash_reads(K, int(canonical), mode, seq.ctypes.data_as(C.c_char_p), _p(offsets, C.c_int64),
                        len(offsets) - 1, _p(out, C.c_uint64))
    return out[:n]


def canonical(fw, rc):
    return np.minimum(fw, rc)


def primes_near_x(n, x):
    out = np.zeros(max(n, 1), np.uint64)
    k = lib().or_primes_near_x(n, x, _p(out, C.c_uint64))
    return [int(v) for v in out[:k]]


class Storage:
    """A sequential BitStorage / ByteStorage / exact-set restatement."""

    def __init__(self, kind, tablesizes=()):
        self.kind = kind
        self._st = _Storage()
        sizes = np.array(list(tablesizes) or [1], np.uint64)
        rc = lib().or_storage_init(C.byref(self._st), kind, _p(sizes, C.c_uint64),
                                   len(tablesizes))
        if rc != 0:
            raise MemoryError('oracle storage init failed')

    @classmethod
    def build(cls, kind, max_table=1_000_000, n_tables=4):
        if kind == EXACT:
            return cls(EXACT)
        return cls(kind, primes_near_x(n_tables, max_table))

    def __del__(self):
        try:
            lib().or_storage_free(C.byref(self._st))
        except Exception:
            pass

    @property
    def ref(self):
        return C.byref(self._st)

    def tablesizes(self):
        return [int(self._st.sizes[i]) for i in range(self._st.n_tables)]

    def table(self, i):
        n = lib().or_table_bytes(self.ref, i)
        return np.ctypeslib.as_array(self._st.tables[i], shape=(n,)).copy()

    def tables(self):
        return [self.table(i) for i in range(self._st.n_tables)]

    @property
    def n_unique(self):
        return int(self._st.n_unique)

    @property
    def n_occupied(self):
        return int(self._st.n_occupied)

    def insert(self, h):
        return bool(lib().or_insert(self.ref, int(h)))

    def query(self, h):
        return int(lib().or_query(self.ref, int(h)))

    def insert_and_query(self, h):
        return int(lib().or_insert_and_query(self.ref, int(h)))

    def set_use_bigcount(self, b=True):
        lib().or_set_use_bigcount(self.ref, int(bool(b)))

    def bigcount_pairs(self):
        """The bigcount map as (keys, values), sorted by key."""
        n = lib().or_bigcount_pairs(self.ref, None, None, 0)
        keys = np.zeros(max(n, 1), np.uint64)
        vals = np.zeros(max(n, 1), np.int16)
        lib().or_bigcount_pairs(self.ref, _p(keys, C.c_uint64), _p(vals, C.c_int16), n)
        return keys[:n], vals[:n]

    def bigcount_set(self, h, v):
        lib().or_bigcount_set(self.ref, int(h), int(v))

    def exact_keys(self):
        n = lib().or_exact_keys(self.ref, None, 0)
        out = np.zeros(max(n, 1), np.uint64)
        lib().or_exact_keys(self.ref, _p(out, C.c_uint64), n)
        return np.sort(out[:n])

    # ---- sequence level ----
    def insert_sequence(self, seq, K, canonical=True, mode=MODE_RAW, flags=False):
        b = seq.encode() if isinstance(seq, str) else bytes(seq)
        nn = C.c_uint64()
        n = max(0, len(b) - K + 1)
        fl = np.zeros(max(n, 1), np.uint8)
        hashed = lib().or_insert_sequence(self.ref, K, int(canonical), mode, b, len(b),
                                          C.byref(nn), _p(fl, C.c_uint8))
        if flags:
            return hashed, nn.value, fl[:n]
        return hashed, nn.value

    def insert_and_query_sequence(self, seq, K, canonical=True, mode=MODE_RAW):
        b = seq.encode() if isinstance(seq, str) else bytes(seq)
        n = max(0, len(b) - K + 1)
        out = np.zeros(max(n, 1), np.int16)
        lib().or_insert_and_query_sequence(self.ref, K, int(canonical), mode, b, len(b),
                                           _p(out, C.c_int16))
        return out[:n]

    def query_sequence(self, seq, K, canonical=True, mode=MODE_RAW):
        b = seq.encode() if isinstance(seq, str) else bytes(seq)
        n = max(0, len(b) - K + 1)
        out = np.zeros(max(n, 1), np.int16)
        hs = np.zeros(max(n, 1), np.uint64)
        lib().or_query_sequence(self.ref, K, int(canonical), mode, b, len(b),
                                _p(out, C.c_int16), _p(hs, C.c_uint64))
        return out[:n], hs[:n]

    def degrees_sequence(self, seq, K, canonical=True, mode=MODE_RAW, min_count=1):
        b = seq.encode() if isinstance(seq, str) else bytes(seq)
        n = max(0, len(b) - K + 1)
        ind = np.zeros(max(n, 1), np.uint8)
        outd = np.zeros(max(n, 1), np.uint8)
        lib().or_degrees_sequence(self.ref, K, int(canonical), mode, min_count, b, len(b),
                                  _p(ind, C.c_uint8), _p(outd, C.c_uint8))
        return ind[:n], outd[:n]

    def kmer_degrees(self, kmer, K, canonical=True, min_count=1):
        b = kmer.encode() if isinstance(kmer, str) else bytes(kmer)
        i, o = C.c_uint8(), C.c_uint8()
        left = (C.c_uint64 * 4)()
        right = (C.c_uint64 * 4)()
        lib().or_kmer_degrees(self.ref, K, int(canonical), min_count, b, C.byref(i),
                              C.byref(o), left, right)
        return i.value, o.value, list(left), list(right)

    def insert_reads(self, seq, offsets, K, canonical=True, mode=MODE_RAW):
        seq = np.ascontiguousarray(seq, np.uint8)
        offsets = np.ascontiguousarray(offsets, np.int64)
        n_reads = len(offsets) - 1
        nn = np.zeros(max(n_reads, 1), np.uint64)
        total = lib().or_insert_reads(self.ref, K, int(canonical), mode,
                                      seq.ctypes.data_as(C.c_char_p), _p(offsets, C.c_int64),
                                      n_reads, _p(nn, C.c_uint64))
        return total, nn[:n_reads]

    def insert_reads_flags(self, seq, offsets, K, canonical=True, mode=MODE_RAW):
        """insert_reads plus each window's first-touch flag (packed)."""
        seq = np.ascontiguousarray(seq, np.uint8)
        offsets = np.ascontiguousarray(offsets, np.int64)
        n_reads = len(offsets) - 1
        nn = np.zeros(max(n_reads, 1), np.uint64)
        n = _n_windows(offsets, K)
        fl = np.zeros(max(n, 1), np.uint8)
        total = lib().or_insert_reads_flags(self.ref, K, int(canonical), mode, seq.ctypes.data_as(C.c_char_p),
                                            _p(offsets, C.c_int64), n_reads, _p(nn, C.c_uint64),
                                            _p(fl, C.c_uint8))
        return total, nn[:n_reads], fl[:n]

    def insert_and_query_reads(self, seq, offsets, K, canonical=True, mode=MODE_RAW):
        """insert_and_query_sequence over reads in order, counts packed."""
        seq = np.ascontiguousarray(seq, np.uint8)
        offsets = np.ascontiguousarray(offsets, np.int64)
        n = _n_windows(offsets, K)
        out = np.zeros(max(n, 1), np.int16)
        lib().or_insert_and_query_reads(self.ref, K, int(canonical), mode, seq.ctypes.data_as(C.c_char_p),
                                        _p(offsets, C.c_int64), len(offsets) - 1, _p(out, C.c_int16))
        return out[:n]

    def degrees_reads_mt(self, seq, offsets, K, canonical=True, mode=MODE_RAW, min_count=1, only=None,
                         threads=1):
        """In / out degrees of the windows selected by `only` (every window
        when None), packed; 0xFF elsewhere."""
        seq = np.ascontiguousarray(seq, np.uint8)
        offsets = np.ascontiguousarray(offsets, np.int64)
        n = _n_windows(offsets, K)
        ind = np.zeros(max(n, 1), np.uint8)
        outd = np.zeros(max(n, 1), np.uint8)
        o = None if only is None else _p(np.ascontiguousarray(only, np.uint8), C.c_uint8)
        lib().or_degrees_reads_mt(self.ref, K, int(canonical), mode, int(min_count),
                                  seq.ctypes.data_as(C.c_char_p), _p(offsets, C.c_int64), len(offsets) - 1,
                                  o, _p(ind, C.c_uint8), _p(outd, C.c_uint8), int(threads))
        return ind[:n], outd[:n]

    def diginorm_filter(self, seq, K, cutoff, canonical=True, mode=MODE_RAW):
        """DiginormFilter::Filter::filter_sequence: 1 inserted, 0 dropped,
        -1 shorter than K."""
        s = seq.encode() if isinstance(seq, str) else bytes(seq)
        return lib().or_diginorm_filter(self.ref, K, int(canonical), mode, cutoff, s, len(s))

    def diginorm_reads(self, seq, offsets, K, cutoff, canonical=True, mode=MODE_RAW):
        seq = np.ascontiguousarray(seq, np.uint8)
        offsets = np.ascontiguousarray(offsets, np.int64)
        n_reads = len(offsets) - 1
        passed = np.zeros(max(n_reads, 1), np.uint8)
        lib().or_diginorm_reads(self.ref, K, int(canonical), mode, cutoff, seq.ctypes.data_as(C.c_char_p),
                                _p(offsets, C.c_int64), n_reads, _p(passed, C.c_uint8))
        return passed[:n_reads].astype(bool)

    def insert_reads_mt(self, seq, offsets, K, canonical=True, mode=MODE_RAW, threads=1):
        """insert_sequence over reads on `threads` host threads (Bloom /
        count-min; the counters are not updated)."""
        seq = np.ascontiguousarray(seq, np.uint8)
        offsets = np.ascontiguousarray(offsets, np.int64)
        return lib().or_insert_reads_mt(self.ref, K, int(canonical), mode, seq.ctypes.data_as(C.c_char_p),
                                        _p(offsets, C.c_int64), len(offsets) - 1, int(threads))

    def query_reads_mt(self, seq, offsets, K, canonical=True, mode=MODE_RAW, threads=1):
        seq = np.ascontiguousarray(seq, np.uint8)
        offsets = np.ascontiguousarray(offsets, np.int64)
        n = int(np.maximum(np.diff(offsets) - K + 1, 0).sum())
        out = np.zeros(max(n, 1), np.int16)
        lib().or_query_reads_mt(self.ref, K, int(canonical), mode, seq.ctypes.data_as(C.c_char_p),
                                _p(offsets, C.c_int64), len(offsets) - 1, _p(out, C.c_int16), int(threads))
        return out[:n]

    def query_reads(self, seq, offsets, K, canonical=True, mode=MODE_RAW):
        seq = np.ascontiguousarray(seq, np.uint8)
        offsets = np.ascontiguousarray(offsets, np.int64)
        n_reads = len(offsets) - 1
        lens = np.diff(offsets)
        n = int(np.maximum(lens - K + 1, 0).sum())
        out = np.zeros(max(n, 1), np.int16)
        lib().or_query_reads(self.ref, K, int(canonical), mode,
                             seq.ctypes.data_as(C.c_char_p), _p(offsets, C.c_int64),
                             n_reads, _p(out, C.c_int16))
        return out[:n]
