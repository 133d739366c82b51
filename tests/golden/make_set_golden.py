"""Writes tests/golden/set_trace_ref.npz: what goetia's compiled exact-set
storages (SparseppSetStorage, BTreeStorage) report after every
single insert of scripted value sequences -- insert result, n_unique_kmers()
and n_occupied() (their containers' bucket_count() / max_size()) -- through
ref_set_trace in oracle/ref_driver.cc.

Run in the container where the goetia checkout is (needs oracle/_ref built
by `make -C oracle ref`):  python tests/golden/make_set_golden.py

Traces:
  boundary  16 distinct values (half of sparsepp's 32 starting buckets), then
            a value already present, then the 17th distinct value; the same
            at 32, 64 and 128 distinct values; then, after a reset, the
            duplicate one value before half
  resets    values from a pool of 300 (with 0 and all-ones) and resets after
            growth, twice in a row, and on an empty set
  long      24,000 inserts drawn from a pool of 12,000 values (duplicates
            at every size), a reset, then 3,000 more
The values are splitmix64 of pool indices (set_value); the fixture stores the
indices.
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_SO = os.path.join(ROOT, 'oracle', '_ref', 'libgoetia_ref.so')
OUT = os.path.join(HERE, 'set_trace_ref.npz')
FLAVORS = ('sparsepp', 'phmap', 'btree')
KEEP = ('sparsepp', 'btree')   # phmap's bucket count (16 hash-selected submaps) is not reproduced


def set_value(idx):
    """Value of pool index `idx` (0 and all-ones for indices 0 and 1, else
    splitmix64 of the index): the fixture stores indices, not values."""
    idx = np.asarray(idx, np.uint64)
    with np.errstate(over='ignore'):
        z = idx + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    z = np.where(idx == 0, np.uint64(0), z)
    return np.where(idx == 1, np.uint64(2**64 - 1), z).astype(np.uint64)


def traces():
    rng = np.random.default_rng(20240611)
    out = {}
    # boundary: distinct values, then a duplicate exactly when the set holds
    # half the buckets (or one value before that)
    idx = []
    nxt = 2
    for half in (16, 32, 64, 128):
        while nxt - 2 < half:
            idx.append(nxt); nxt += 1
        idx.append(2)                        # present: size + 1 > half -> grows
        idx.append(nxt); nxt += 1
    n1 = len(idx)
    nxt2 = 1000
    for half in (16, 32, 64):
        while nxt2 - 1000 < half - 1:
            idx.append(nxt2); nxt2 += 1
        idx.append(1000)                     # present, one value below half
        idx.append(nxt2); nxt2 += 1
        idx.append(nxt2); nxt2 += 1
    r = np.zeros(len(idx), np.uint8)
    r[n1] = 1
    out['boundary'] = (np.array(idx, np.uint16), r)
    # resets, with the zero and all-ones values in the pool
    idx = rng.integers(0, 300, size=3000).astype(np.uint16)
    r = np.zeros(len(idx), np.uint8)
    for i in (700, 701, 1500, 2200):
        r[i] = 1
    out['resets'] = (idx, r)
    # long: many duplicates at every size
    idx = np.concatenate([rng.integers(2, 12002, size=24000), rng.integers(2, 1000, size=3000)]).astype(np.uint16)
    r = np.zeros(len(idx), np.uint8)
    r[24000] = 1
    out['long'] = (idx, r)
    return out


def main():
    if not os.path.exists(REF_SO):
        sys.exit(f'{REF_SO} is not built: make -C oracle ref REF=/root/reference')
    L = C.CDLL(REF_SO)
    L.ref_set_trace.argtypes = [C.c_int, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.ref_set_trace.restype = None
    arrays = {}
    for name, (idx, r) in traces().items():
        v = np.ascontiguousarray(set_value(idx), np.uint64)
        r = np.ascontiguousarray(r, np.uint8)
        arrays[f'{name}__idx'] = idx
        arrays[f'{name}__reset_before'] = r
        for fi, fl in enumerate(FLAVORS):
            if fl not in KEEP:
                continue
            is_new = np.zeros(len(v), np.uint8)
            nu = np.zeros(len(v), np.uint64)
            no = np.zeros(len(v), np.uint64)
            L.ref_set_trace(fi, v.ctypes.data, len(v), r.ctypes.data, is_new.ctypes.data, nu.ctypes.data,
                            no.ctypes.data)
            arrays[f'{name}__{fl}__is_new'] = is_new
            arrays[f'{name}__{fl}__n_unique'] = nu
            arrays[f'{name}__{fl}__n_occupied'] = no
    np.savez_compressed(OUT, **arrays)
    print(f'wrote {OUT} ({os.path.getsize(OUT)} bytes)')


if __name__ == '__main__':
    main()
