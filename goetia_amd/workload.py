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


# ---------------------------------------------------------------- configurations

def _scaled(n, scale, lo=1):
    return max(lo, int(round(n * scale)))


def bloom_100k(scale=1.0):
    """1 Mbp genome, 100k x 150 bp reads at 1% substitutions, K=21,
    BitStorage(10^7, 4), per-read n_new."""
    g = genome(1, _scaled(1_000_000, scale, 10_000))
    b = reads(1001, g, _scaled(100_000, scale), 150, 150, 0.01, 0.0)
    return dict(K=21, storage='BitStorage', max_table=10_000_000, n_tables=4, mode=0), b


def exact_mixed(scale=1.0):
    """Uneven reads: 200k lognormal(median 8 kb, sigma 1.0) reads clipped to
    150 bp..100 kb and 2M x 150 bp reads, interleaved in a seeded random
    order, 5% substitutions, from a 50 Mbp genome; K=25 into the exact set."""
    g = genome(3, _scaled(50_000_000, scale, 100_000))
    n_long = _scaled(200_000, scale)
    n_short = _scaled(2_000_000, scale)
    lens = np.concatenate([lognormal_lengths(3, n_long, 8000, 1.0, 150, 100_000),
                           np.full(n_short, 150, np.uint32)])
    lens = lens[np.random.default_rng(33).permutation(len(lens))]
    b = reads(1003, g, len(lens), 0, 0, 0.05, 0.0, lengths=lens)
    return dict(K=25, storage='SparseppSetStorage', mode=0, n_long=n_long, n_short=n_short), b


def repeat_genome(seed, length, unit_len, copies, snp_rate):
    """A genome with `copies` copies of one repeat unit at random
    non-overlapping sites, and a second haplotype with `snp_rate`
    substitutions (SNP bubbles)."""
    g = genome(seed, length).copy()
    unit = genome(seed + 1, unit_len)
    rng = np.random.default_rng(seed)
    slots = length // unit_len
    for site in rng.choice(slots, size=min(copies, slots), replace=False):
        g[site * unit_len:(site + 1) * unit_len] = unit
    hap2 = g.copy()
    m = rng.random(length) < snp_rate
    acgt = np.frombuffer(b'ACGT', np.uint8)
    idx = np.searchsorted(acgt, hap2[m])
    hap2[m] = acgt[(idx + rng.integers(1, 4, int(m.sum()))) % 4]
    return g, hap2


def neighbors_5m(scale=1.0):
    """20 Mbp genome with 1,000 copies of a 2 kb repeat and a 0.1% SNP
    haplotype; 5M x 150 bp reads (half from each haplotype) at 1% error;
    K=31, ByteStorage(2^28, 4); decision nodes at count threshold 2."""
    L = _scaled(20_000_000, scale, 200_000)
    h1, h2 = repeat_genome(4, L, 2000, _scaled(1000, scale, 10), 0.001)
    n = _scaled(5_000_000, scale)
    a = reads(1004, h1, n // 2, 150, 150, 0.01, 0.0)
    b = reads(2004, h2, n - n // 2, 150, 150, 0.01, 0.0)
    seq = np.concatenate([a.seq, b.seq])
    off = np.concatenate([a.offsets, b.offsets[1:] + a.offsets[-1]])
    max_table = 2 ** 28 if scale >= 1.0 else max(100_000, int(2 ** 28 * scale))
    return dict(K=31, storage='ByteStorage', max_table=max_table, n_tables=4, mode=0,
                min_count=2), ReadBatch(seq, off)


def segments_5m(scale=1.0):
    """find_new_segments against a fixed graph: a 5 Mbp genome with 100
    copies of a 2 kb repeat and a 0.1% SNP haplotype; 400k x 150 bp reads of
    the first haplotype at 0.5% error make the graph (K=31, forward hashes,
    the exact set, as goetia instantiates its compactor), then 400k x 150 bp
    reads of the second haplotype at 1% error are cut into segments."""
    L = _scaled(5_000_000, scale, 100_000)
    h1, h2 = repeat_genome(6, L, 2000, _scaled(100, scale, 5), 0.001)
    n = _scaled(400_000, scale)
    bg = reads(1006, h1, n, 150, 150, 0.005, 0.0)
    q = reads(2006, h2, n, 150, 150, 0.01, 0.0)
    return dict(K=31, storage='SparseppSetStorage', mode=0), bg, q
