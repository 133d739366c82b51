"""Seeded synthetic workloads (the shapes of the benchmark configurations).

Reads come from gt_gen_genome / gt_gen_reads of libgoetia_amd (host code,
xoshiro256** seeded through splitmix64; each read draws its own generator
from (seed, read index), so the bytes are identical on every machine).
"""
import ctypes as C

import numpy as np

from . import _native as N
from .reads import ReadBatch


def genome(seed, length):
    out = np.zeros(int(length), np.uint8)
    N.call('gt_gen_genome', int(seed), int(length), N.ptr(out))
    return out


def reads(seed, genome_arr, n_reads, len_min, len_max, sub_rate=0.0, n_rate=0.0, lengths=None):
    genome_arr = np.ascontiguousarray(genome_arr, np.uint8)
    if lengths is not None:
        lengths = np.ascontiguousarray(lengths, np.uint32)
        cap = int(np.minimum(lengths, len(genome_arr)).astype(np.int64).sum())
    else:
        cap = int(n_reads) * int(min(len_max, len(genome_arr)))
    seq = np.zeros(max(cap, 1), np.uint8)
    off = np.zeros(int(n_reads) + 1, np.int64)
    nb = C.c_uint64()
    N.call('gt_gen_reads', int(seed), N.ptr(genome_arr), len(genome_arr), int(n_reads),
           N.ptr(lengths), int(len_min), int(len_max), float(sub_rate), float(n_rate),
           N.ptr(seq), N.ptr(off), C.byref(nb))
    return ReadBatch(seq[:nb.value], off)


def lognormal_lengths(seed, n, median, sigma, lo, hi):
    """Read lengths clip(round(lognormal(ln median, sigma)), lo, hi)."""
    rng = np.random.default_rng(seed)
    v = np.rint(rng.lognormal(np.log(median), sigma, int(n)))
    return np.clip(v, lo, hi).astype(np.uint32)


# The benchmark configurations (sizes, rates, K, storage shapes).
CONFIGS = {
    'bloom_100k': dict(seed=1, genome=1_000_000, n_reads=100_000, len=(150, 150), sub=0.01,
                       n_rate=0.0, K=21, storage='BitStorage', max_table=10_000_000, n_tables=4),
    'cms_20m': dict(seed=2, genome=50_000_000, n_reads=20_000_000, len=(150, 150), sub=0.005,
                    n_rate=0.0001, K=31, storage='ByteStorage', max_table=2 ** 28, n_tables=4),
}


def make(config, scale=1.0):
    c = dict(CONFIGS[config])
    n = max(1, int(c['n_reads'] * scale))
    g = genome(c['seed'], c['genome'])
    b = reads(c['seed'] + 1000, g, n, c['len'][0], c['len'][1], c['sub'], c['n_rate'])
    return c, b
