"""khmer "OXLI" v4 table files for BitStorage and ByteStorage, as goetia
writes and reads them (BitStorage::save / load, bitstorage.cc:151-307;
ByteStorage::save / load through ByteStorageFile, bytestorage.cc:153-667,
gzip-compressed when a count-min file's name ends in ".gz").

The files are written and read by libgoetia_amd (gt_storage_save /
gt_storage_load, and the host-only gt_oxli_* for tables in host memory), so
the bytes -- compressed ones included -- are goetia's.  A load replaces the
storage's tables, sizes included, and restores n_occupied (n_unique is not
in the format).  Count-min files carry the bigcount flag and pairs: a
storage's save writes its map (pairs sorted by k-mer) and a load takes the
file's; the host-only readers skip the pairs.
"""
import ctypes as C

import numpy as np

from . import _native as N

SIGNATURE = b'OXLI'
VERSION = 4
SAVED_COUNTING_HT = 1
SAVED_HASHBITS = 2


def _kind(store):
    from .storage import BitStorage, ByteStorage
    if isinstance(store, BitStorage):
        return N.BIT
    if isinstance(store, ByteStorage):
        return N.BYTE
    raise N.UnsupportedError('OXLI files hold Bloom and count-min tables only '
                             '(the exact set has no file format, sparseppstorage.cc:52-58)')


def save_storage(store, filename, ksize):
    _kind(store)
    N.call('gt_storage_save', store.handle, str(filename).encode(), int(ksize))


def load_storage(store, filename):
    """Loads an OXLI file into `store` (its tables take the file's sizes);
    returns the file's ksize."""
    _kind(store)
    k = C.c_uint16()
    N.call('gt_storage_load', store.handle, str(filename).encode(), C.byref(k))
    return k.value


def read_file(filename, kind):
    """Host-only read of a file of `kind` (N.BIT / N.BYTE): (ksize,
    tablesizes, occupied, tables as uint8 arrays)."""
    p = str(filename).encode()
    k, n, occ = C.c_uint16(), C.c_uint32(), C.c_uint64()
    sizes = np.zeros(16, np.uint64)
    N.call('gt_oxli_read_header', p, int(kind), C.byref(k), C.byref(n), sizes.ctypes.data_as(N.u64p),
           C.byref(occ))
    sizes = [int(s) for s in sizes[:n.value]]
    tables = [np.zeros(s // 8 + 1 if kind == N.BIT else s, np.uint8) for s in sizes]
    ptrs = (C.c_void_p * len(tables))(*[t.ctypes.data for t in tables])
    N.call('gt_oxli_read', p, int(kind), ptrs)
    return k.value, sizes, occ.value, tables


def write_file(filename, kind, ksize, tablesizes, occupied, tables):
    """Host-only write of a file of `kind` from tables in host memory."""
    tables = [np.ascontiguousarray(t, np.uint8) for t in tables]
    sizes = np.array(tablesizes, np.uint64)
    for s, t in zip(tablesizes, tables):
        if len(t) != (s // 8 + 1 if kind == N.BIT else s):
            raise ValueError('table bytes do not match the table size')
    ptrs = (C.c_void_p * len(tables))(*[t.ctypes.data for t in tables])
    N.call('gt_oxli_write', str(filename).encode(), int(kind), int(ksize), len(tables),
           sizes.ctypes.data_as(N.u64p), int(occupied), ptrs)
