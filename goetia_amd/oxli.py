"""khmer "OXLI" v4 table files for BitStorage and ByteStorage.

Layout (little-endian), as written by BitStorage::save (bitstorage.cc:151-189)
and ByteStorageFileWriter (bytestorage.cc:480-536):

  "OXLI" | version u8 = 4 | type u8 (2 = hashbits, 1 = counting)
  [counting only: use_bigcount u8]
  ksize u32 | n_tables u8 | occupied_bins u64
  n_tables x ( tablesize u64 | table bytes: tablesize/8+1 (Bloom) or tablesize )
  [counting only: n_bigcounts u64 (+ pairs, always 0 here)]

The bytes are the device tables copied to host; a load writes them back and
restores n_occupied (n_unique is not stored by the format).
"""
import struct

import numpy as np

from . import _native as N

SIGNATURE = b'OXLI'
VERSION = 4
SAVED_COUNTING_HT = 1
SAVED_HASHBITS = 2


def save_storage(store, filename, ksize):
    from .storage import BitStorage, ByteStorage
    if isinstance(store, BitStorage):
        kind = SAVED_HASHBITS
    elif isinstance(store, ByteStorage):
        kind = SAVED_COUNTING_HT
    else:
        raise N.UnsupportedError('OXLI files hold Bloom and count-min tables only')
    sizes = store.get_tablesizes()
    tables = store.get_raw_tables()
    with open(filename, 'wb') as f:
        f.write(SIGNATURE)
        f.write(struct.pack('<BB', VERSION, kind))
        if kind == SAVED_COUNTING_HT:
            f.write(struct.pack('<B', 0))
        f.write(struct.pack('<IBQ', int(ksize), len(sizes), store.n_occupied()))
        for size, tab in zip(sizes, tables):
            f.write(struct.pack('<Q', size))
            f.write(tab.tobytes())
        if kind == SAVED_COUNTING_HT:
            f.write(struct.pack('<Q', 0))


def load_storage(store, filename):
    """Loads an OXLI file into `store` (same table sizes); returns ksize."""
    from .storage import BitStorage, ByteStorage
    with open(filename, 'rb') as f:
        data = f.read()
    if data[:4] != SIGNATURE:
        raise OSError(f'{filename}: not an OXLI file')
    version, kind = struct.unpack_from('<BB', data, 4)
    if version != VERSION:
        raise OSError(f'{filename}: OXLI version {version} is not supported')
    pos = 6
    if kind == SAVED_COUNTING_HT:
        if not isinstance(store, ByteStorage):
            raise OSError(f'{filename}: counting table loaded into {type(store).__name__}')
        (use_big,) = struct.unpack_from('<B', data, pos)
        pos += 1
    elif kind == SAVED_HASHBITS:
        if not isinstance(store, BitStorage):
            raise OSError(f'{filename}: Bloom table loaded into {type(store).__name__}')
    else:
        raise OSError(f'{filename}: unknown OXLI table type {kind}')
    ksize, n_tables, occupied = struct.unpack_from('<IBQ', data, pos)
    pos += 13
    if n_tables != store.n_tables():
        raise OSError(f'{filename}: {n_tables} tables, storage has {store.n_tables()}')
    sizes = store.get_tablesizes()
    for i in range(n_tables):
        (size,) = struct.unpack_from('<Q', data, pos)
        pos += 8
        if size != sizes[i]:
            raise OSError(f'{filename}: table {i} has size {size}, storage has {sizes[i]}')
        nb = size // 8 + 1 if kind == SAVED_HASHBITS else size
        store.set_raw_table(i, np.frombuffer(data, np.uint8, nb, pos))
        pos += nb
    N.call('gt_storage_set_counters', store.handle, store.n_unique_kmers(), int(occupied))
    return ksize
