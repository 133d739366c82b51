"""Writes tests/golden/segments_ref.npz from the compiled goetia reference's
StreamingCompactor<dBG<S, FwdLemireShifter>>::Compactor::find_new_segments
(compactor.hh:321-534), built by `make -C oracle ref` from the reference
sources.  Run in a container that has the reference checkout:

    python tests/golden/make_segments_golden.py

Each case stores the graph (storage kind, K, max_table, n_tables), reads
inserted before (bg_seq / bg_off), the reads given to find_new_segments in
order (seq / off), whether one new-k-mer set is shared by all the reads
(persist) and whether each read is inserted after its call (insert_after),
and the reference's output stream (`out`, u64 words, format in
oracle/ref_driver.cc ref_find_new_segments) with the neighbour k-mer strings
(`kmers`).  Inputs: forks, merges and decision k-mers at read ends built as
goetia's tests/test_cdbg.py builds them, and seeded reads over a genome with
repeats and SNPs (errors, N, reads shorter than K, duplicates).
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from make_golden import pack, ref  # noqa: E402

BIT, BYTE, SET = 1, 2, 3


def rand_seq(rng, n):
    return ''.join(rng.choice(list('ACGT'), n))


def other(rng, c):
    return rng.choice([x for x in 'ACGT' if x != c])


def right_fork(rng, K, length=100, branch_len=40):
    """core, and a branch leaving it after the k-mer at `pivot` (that k-mer
    gets two right neighbours)."""
    core = rand_seq(rng, length)
    pivot = length // 2
    nxt = other(rng, core[pivot + K])
    branch = core[pivot + 1:pivot + K] + nxt + rand_seq(rng, branch_len - K)
    return core, branch, pivot


def left_fork(rng, K, length=100, branch_len=40):
    """core, and a branch joining it just before the k-mer at `pivot` (that
    k-mer gets two left neighbours)."""
    core = rand_seq(rng, length)
    pivot = length // 2
    prv = other(rng, core[pivot - 1])
    branch = rand_seq(rng, branch_len - K) + prv + core[pivot:pivot + K - 1]
    return core, branch, pivot


def fork_cases(seed, K):
    rng = np.random.default_rng(seed)
    cases = {}
    core, branch, pivot = right_fork(rng, K)
    cases['core_first'] = [core, core[:pivot + 1] + branch]
    core, branch, pivot = right_fork(rng, K)
    cases['right_decision_split'] = [branch, core]
    seq = rand_seq(rng, 100)
    left, right = seq[:50], seq[50:]
    cases['merge'] = [left, right, left + right]
    core, branch, pivot = right_fork(rng, K)
    cases['right_decision_on_end'] = [branch, core[pivot + 1:], core[:pivot + K]]
    core, branch, pivot = left_fork(rng, K)
    cases['left_decision_on_end'] = [branch, core[:pivot + K - 1], core[pivot:]]
    core, branch, pivot = left_fork(rng, K)
    cases['left_decision_split'] = [branch, core]
    # two decision k-mers in one new segment, a bubble, a self-loop
    core = rand_seq(rng, 120)
    b1 = core[30:30 + K - 1] + other(rng, core[30 + K - 1]) + rand_seq(rng, 20)
    b2 = rand_seq(rng, 20) + other(rng, core[70 - 1]) + core[70:70 + K - 1]
    cases['two_decisions'] = [b1, b2, core]
    bub = core[:60] + other(rng, core[60]) + core[61:]
    cases['bubble'] = [core, bub]
    unit = rand_seq(rng, K + 3)
    cases['cycle'] = [unit * 4, unit * 2 + rand_seq(rng, 30)]
    return cases


def genome_reads(seed, n, K):
    """Reads over a 4 kbp genome holding a 150 bp repeat at three sites and
    a second haplotype with SNPs; 1% errors, some N, some reads below K, some
    duplicates."""
    rng = np.random.default_rng(seed)
    g = list(rand_seq(rng, 4000))
    rep = rand_seq(rng, 150)
    for site in (500, 1800, 3100):
        g[site:site + 150] = list(rep)
    g = ''.join(g)
    h = list(g)
    for p in rng.choice(len(h), 12, replace=False):
        h[p] = other(rng, h[p])
    h = ''.join(h)
    out = []
    for i in range(n):
        src = g if rng.random() < 0.6 else h
        L = int(rng.integers(K - 3, 220))
        s = int(rng.integers(0, len(src) - L))
        r = list(src[s:s + L])
        for j in range(L):
            if rng.random() < 0.01:
                r[j] = other(rng, r[j])
        if i % 37 == 5:
            r[int(rng.integers(0, L))] = 'N'
        out.append(''.join(r))
        if i % 23 == 7:
            out.append(out[-1])
    return out


def run_case(L, kind, K, max_table, n_tables, bg, reads, persist, insert_after):
    h = L.ref_new(kind, 0, K, max_table, n_tables)
    if bg:
        bseq, boff = pack(bg)
        L.ref_insert_reads(h, K, 0, bseq.ctypes.data_as(C.c_char_p), boff.ctypes.data_as(C.POINTER(C.c_int64)),
                           len(bg), None)
    seq, off = pack(reads)
    cap, kcap = 1 << 22, 1 << 24
    out = np.zeros(cap, np.uint64)
    km = np.zeros(kcap, np.uint8)
    used = C.c_uint64()
    n = L.ref_find_new_segments(h, seq.ctypes.data_as(C.c_char_p), off.ctypes.data_as(C.POINTER(C.c_int64)),
                                len(reads), persist, insert_after, out.ctypes.data_as(C.POINTER(C.c_uint64)), cap,
                                km.ctypes.data_as(C.c_char_p), kcap, C.byref(used))
    assert n <= cap and used.value <= kcap, (n, used.value)
    L.ref_free(h)
    return out[:n].copy(), km[:used.value].copy()


def main():
    L = ref()
    L.ref_find_new_segments.restype = C.c_uint64
    L.ref_find_new_segments.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_int64), C.c_uint64, C.c_int, C.c_int,
                                        C.POINTER(C.c_uint64), C.c_uint64, C.c_char_p, C.c_uint64,
                                        C.POINTER(C.c_uint64)]
    L.ref_free.restype = None
    L.ref_free.argtypes = [C.c_void_p]
    L.ref_insert_reads.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_char_p, C.POINTER(C.c_int64), C.c_uint64,
                                   C.c_void_p]
    arrays = {}

    def add(name, kind, K, max_table, n_tables, bg, reads, persist, insert_after):
        out, km = run_case(L, kind, K, max_table, n_tables, bg, reads, persist, insert_after)
        seq, off = pack(reads)
        bseq, boff = pack(bg)
        arrays[f'{name}__params'] = np.array([kind, K, max_table, n_tables, persist, insert_after], np.int64)
        arrays[f'{name}__seq'], arrays[f'{name}__off'] = seq, off
        arrays[f'{name}__bg_seq'], arrays[f'{name}__bg_off'] = bseq, boff
        arrays[f'{name}__out'], arrays[f'{name}__kmers'] = out, km
        print(name, len(reads), 'reads', len(out), 'words')

    for K in (15, 21):
        for cname, reads in fork_cases(10 + K, K).items():
            add(f'fork_{cname}_set_k{K}', SET, K, 0, 0, [], reads, 0, 1)
        for cname, reads in fork_cases(50 + K, K).items():
            add(f'fork_{cname}_byte_k{K}', BYTE, K, 100_000, 4, [], reads, 0, 1)
    reads = genome_reads(7, 260, 21)
    add('genome_set_k21', SET, 21, 0, 0, [], reads, 0, 1)
    add('genome_set_k21_persist', SET, 21, 0, 0, [], reads, 1, 1)
    add('genome_byte_k21_small_tables', BYTE, 21, 3_000, 3, [], reads, 0, 1)
    add('genome_bit_k21', BIT, 21, 200_000, 4, [], reads, 0, 1)
    add('genome_set_k31', SET, 31, 0, 0, [], genome_reads(8, 200, 31), 0, 1)
    # the graph fixed: background reads, then calls that insert nothing
    bg = genome_reads(9, 300, 25)
    add('fixed_set_k25', SET, 25, 0, 0, bg, genome_reads(10, 200, 25), 0, 0)
    add('fixed_set_k25_persist', SET, 25, 0, 0, bg, genome_reads(10, 200, 25), 1, 0)
    add('fixed_byte_k25', BYTE, 25, 50_000, 4, bg, genome_reads(11, 200, 25), 0, 0)
    np.savez_compressed(os.path.join(HERE, 'segments_ref.npz'), **arrays)


if __name__ == '__main__':
    main()
