"""Device storages with goetia's names: BitStorage (Bloom filter),
ByteStorage (count-min sketch) and the exact k-mer set (SparseppSetStorage and
its same-semantics siblings PHMapStorage / BTreeStorage).

Tables live in HBM and are updated by the HIP kernels of libgoetia_amd; the
methods mirror storage.hh:98-143 and the per-storage headers, and ``build``
mirrors the pythonizor of pythonize_storage.py:26-38.
"""
import ctypes as C

import numpy as np

from . import _native as N

count_t = np.int16        # storage.hh:80
MAX_KCOUNT = 255          # bytestorage.hh:59


def get_n_primes_near_x(n, x):
    """storage.hh:166-190: up to n odd primes strictly below x, descending."""
    out = np.zeros(max(int(n), 1), np.uint64)
    got = C.c_uint32()
    N.call('gt_primes_near_x', int(n), int(x), out.ctypes.data_as(N.u64p), C.byref(got))
    return [int(v) for v in out[:got.value]]


class StorageTraits:
    def __init__(self, is_probabilistic, is_counting, default_params):
        self.is_probabilistic = is_probabilistic
        self.is_counting = is_counting
        self.default_params = default_params
        self.params_type = tuple


class _Storage:
    kind = None
    Traits = None

    def set_use_bigcount(self, b):
        """Storage::set_use_bigcount (storage.hh:129-135): only the count-min
        storage supports bigcount."""
        if b:
            raise N.GoetiaException('bigcount is not supported for this storage.')

    def get_use_bigcount(self):
        return False

    def __init__(self, *args, capacity_hint=0):
        self._h = C.c_void_p()
        self.build_params = args
        tablesizes = self._tablesizes_from(args)
        arr = np.array(tablesizes or [1], np.uint64)
        N.call('gt_storage_new', self.kind, arr.ctypes.data_as(N.u64p), len(tablesizes),
               int(capacity_hint), C.byref(self._h))

    @staticmethod
    def _tablesizes_from(args):
        if len(args) == 1 and isinstance(args[0], (list, tuple, np.ndarray)):
            return [int(v) for v in args[0]]
        if len(args) == 2:
            max_table, n = args
            return get_n_primes_near_x(int(n), int(max_table))
        raise TypeError('expected (max_table, n_tables) or (tablesizes,)')

    # goetia pythonizor: build() -> default params, build(*params)
    @classmethod
    def build(cls, *args, **kw):
        if len(args) == 0:
            args = cls.Traits.default_params
        elif len(args) == 1 and isinstance(args[0], tuple):
            args = args[0]
        obj = cls(*args, **kw)
        obj.build_params = args
        return obj

    @classmethod
    def make_params(cls, *args):
        return tuple(args)

    @property
    def handle(self):
        return self._h

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            try:
                N.load().gt_storage_free(h)
            except Exception:
                pass
            self._h = C.c_void_p()

    # ---- counters and shape
    def _counters(self):
        u, o = C.c_uint64(), C.c_uint64()
        N.call('gt_storage_counters', self._h, C.byref(u), C.byref(o))
        return u.value, o.value

    def n_unique_kmers(self):
        return self._counters()[0]

    def n_occupied(self):
        return self._counters()[1]

    def n_tables(self):
        n = C.c_uint32()
        N.call('gt_storage_n_tables', self._h, C.byref(n))
        return n.value

    def get_tablesizes(self):
        n = self.n_tables()
        out = np.zeros(max(n, 1), np.uint64)
        N.call('gt_storage_tablesizes', self._h, out.ctypes.data_as(N.u64p))
        return [int(v) for v in out[:n]]

    def estimated_fp(self):
        """bitstorage.hh:183-188: (occupied / size_0) ^ n_tables."""
        sizes = self.get_tablesizes()
        return (self.n_occupied() / sizes[0]) ** len(sizes)

    def table_bytes(self, i):
        n = C.c_uint64()
        N.call('gt_storage_table_bytes', self._h, int(i), C.byref(n))
        return n.value

    def get_raw_tables(self):
        """get_raw_tables() copied to host, one uint8 array per table in the
        reference layout."""
        out = []
        for i in range(self.n_tables()):
            buf = np.zeros(self.table_bytes(i), np.uint8)
            N.call('gt_storage_read_table', self._h, i, N.ptr(buf))
            out.append(buf)
        return out

    def set_raw_table(self, i, data):
        data = np.ascontiguousarray(data, np.uint8)
        if len(data) != self.table_bytes(i):
            raise ValueError('table size mismatch')
        N.call('gt_storage_write_table', self._h, int(i), N.ptr(data))

    def reset(self):
        N.call('gt_storage_reset', self._h)

    def clone(self):
        obj = type(self).__new__(type(self))
        obj._h = C.c_void_p()
        obj.build_params = self.build_params
        N.call('gt_storage_clone', self._h, C.byref(obj._h))
        return obj

    # ---- per-hash operations (storage.hh:120-122)
    def insert(self, h):
        arr = np.array([int(h)], np.uint64)
        out = np.zeros(1, np.uint8)
        N.call('gt_insert_hashes', self._h, N.ptr(arr), 1, N.ptr(out))
        return bool(out[0])

    def insert_and_query(self, h):
        arr = np.array([int(h)], np.uint64)
        out = np.zeros(1, np.int16)
        N.call('gt_insert_and_query_hashes', self._h, N.ptr(arr), 1, N.ptr(out))
        return int(out[0])

    def query(self, h):
        arr = np.array([int(h)], np.uint64)
        out = np.zeros(1, np.int16)
        N.call('gt_query_hashes', self._h, N.ptr(arr), 1, N.ptr(out))
        return int(out[0])

    def insert_hashes(self, hashes):
        arr = np.ascontiguousarray(hashes, np.uint64)
        out = np.zeros(max(len(arr), 1), np.uint8)
        N.call('gt_insert_hashes', self._h, N.ptr(arr), len(arr), N.ptr(out))
        return out[:len(arr)].astype(bool)

    def query_hashes(self, hashes):
        arr = np.ascontiguousarray(hashes, np.uint64)
        out = np.zeros(max(len(arr), 1), np.int16)
        N.call('gt_query_hashes', self._h, N.ptr(arr), len(arr), N.ptr(out))
        return out[:len(arr)]

    def insert_and_query_hashes(self, hashes):
        arr = np.ascontiguousarray(hashes, np.uint64)
        out = np.zeros(max(len(arr), 1), np.int16)
        N.call('gt_insert_and_query_hashes', self._h, N.ptr(arr), len(arr), N.ptr(out))
        return out[:len(arr)]

    def describe(self):
        """pythonize_storage.py:40-46"""
        return (f'{type(self).__name__}\n'
                f'- Params: {self.build_params}\n'
                f'- Probabilistic: {bool(self.is_probabilistic)}\n'
                f'- Counting:      {bool(self.is_counting)}')

    def save(self, filename, ksize):
        from .oxli import save_storage
        save_storage(self, filename, ksize)

    def load(self, filename):
        from .oxli import load_storage
        return load_storage(self, filename)

    def __repr__(self):
        return f'<{type(self).__name__} tables={self.get_tablesizes()}>'


class BitStorage(_Storage):
    """Bloom filter (bitstorage.hh:93-243) in HBM."""
    kind = N.BIT
    Traits = StorageTraits(True, False, (1_000_000, 4))
    is_probabilistic = True
    is_counting = False
    default_params = Traits.default_params
    params_type = tuple

    def update_from(self, other):
        """bitstorage.cc:103-137: OR-merge of another Bloom filter."""
        N.call('gt_storage_update_from', self._h, other._h)


class ByteStorage(_Storage):
    """Count-min sketch of saturating 8-bit counters (bytestorage.hh:96-243)
    in HBM."""
    kind = N.BYTE
    Traits = StorageTraits(True, True, (1_000_000, 4))
    is_probabilistic = True
    is_counting = True
    default_params = Traits.default_params
    params_type = tuple

    def set_use_bigcount(self, b):
        """storage.hh:129-135: with bigcount on, counts past 255 go to a map of
        k-mer -> int16 count on the device (bytestorage.cc:95-106, 130-137)."""
        N.call('gt_storage_set_use_bigcount', self._h, int(bool(b)))

    def get_use_bigcount(self):
        v = C.c_int()
        N.call('gt_storage_get_use_bigcount', self._h, C.byref(v))
        return bool(v.value)

    def bigcounts(self):
        """The bigcount map as (k-mers, counts) numpy arrays, sorted by k-mer."""
        n = C.c_uint64()
        N.call('gt_storage_bigcounts', self._h, None, None, 0, C.byref(n))
        keys = np.zeros(max(n.value, 1), np.uint64)
        vals = np.zeros(max(n.value, 1), np.int16)
        N.call('gt_storage_bigcounts', self._h, N.ptr(keys), N.ptr(vals), n.value, C.byref(n))
        return keys[:n.value], vals[:n.value]

    def set_bigcounts(self, keys, vals):
        """Replaces the bigcount map (a later pair of one k-mer wins)."""
        keys = np.ascontiguousarray(keys, np.uint64)
        vals = np.ascontiguousarray(vals, np.int16)
        N.call('gt_storage_set_bigcounts', self._h, N.ptr(keys), N.ptr(vals), len(keys))


class SparseppSetStorage(_Storage):
    """Exact k-mer set (sparseppstorage.hh:38-108) as a GPU open-addressing
    table of 64-bit keys."""
    kind = N.EXACT
    Traits = StorageTraits(False, False, ())
    is_probabilistic = False
    is_counting = False
    default_params = ()
    params_type = tuple

    def __init__(self, *args, capacity_hint=0):
        self._h = C.c_void_p()
        self.build_params = args
        N.call('gt_storage_new', self.kind, None, 0, int(capacity_hint), C.byref(self._h))

    def n_buckets(self):
        return self.n_occupied()

    def get_maxsize(self):
        return 2 ** 63

    def estimated_fp(self):
        return 0.0

    def get_tablesizes(self):
        return []

    def n_tables(self):
        return 0

    def get_raw_tables(self):
        return None

    def keys(self):
        """The set contents as a sorted uint64 array."""
        n = C.c_uint64()
        N.call('gt_storage_exact_keys', self._h, None, 0, C.byref(n))
        out = np.zeros(max(n.value, 1), np.uint64)
        N.call('gt_storage_exact_keys', self._h, N.ptr(out), n.value, C.byref(n))
        return np.sort(out[:n.value])


class PHMapStorage(SparseppSetStorage):
    """phmapstorage.hh:43,90-95: same set semantics as SparseppSetStorage.
    n_occupied() is sparsepp's bucket count here, not that of phmap's
    parallel set (a sum over 16 hash-selected sub-tables)."""


class BTreeStorage(SparseppSetStorage):
    """btreestorage.hh:43,90-95: same set semantics as SparseppSetStorage."""

    def n_buckets(self):
        """btreestorage.hh:78-80: the btree's max_size(), a constant
        (btree.h:1872, the largest signed 64-bit value as goetia's build
        reports it)."""
        return 2 ** 63 - 1

    def n_occupied(self):
        return self.n_buckets()


typenames = [(SparseppSetStorage, 'SparseppSetStorage'), (PHMapStorage, 'PHMapStorage'),
             (BitStorage, 'BitStorage'), (ByteStorage, 'ByteStorage'),
             (BTreeStorage, 'BTreeStorage')]
types = [t for t, _ in typenames]


def get_storage_args(parser, default='SparseppSetStorage', group_name='storage'):
    """goetia/storage.py:32-49."""
    group = parser.add_argument_group(group_name)
    group.add_argument('-S', '--storage', choices=[n for _, n in typenames], default=default)
    group.add_argument('-N', '--n_tables', default=4, type=int)
    group.add_argument('-x', '--max-tablesize', default=1e8, type=float)
    group.add_argument('-E', '--error-rate', default=0.02, type=float)
    return group


def process_storage_args(args):
    """goetia/storage.py:52-73."""
    args.storage = dict((n, t) for t, n in typenames)[args.storage]
    if args.storage in (BitStorage, ByteStorage):
        args.max_tablesize = int(args.max_tablesize)
        args.storage_args = (args.max_tablesize, args.n_tables)
    else:
        args.storage_args = tuple()
    return args

