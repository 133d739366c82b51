"""Writes tests/golden/diginorm_ref.npz from the compiled goetia reference's
DiginormFilter<dBG<ByteStorage, Sh>>::Filter (diginorm.hh:105-119), built by
`make -C oracle ref` from the reference sources.  Run in a container that has
the reference checkout:

    python tests/golden/make_diginorm_golden.py

Each case stores its reads (seq, off), the graph (K, canonical, max_table,
n_tables), the cutoff, and what the reference left: the pass flag of every
read, n_unique / n_occupied, and the tables (SHA-256 of each; the bytes
themselves when small).  Inputs: the reference's own test data and seeded
synthetic reads, including exact duplicates, tandem repeats (one k-mer many
times in a read), reads shorter than K and reads holding N (hashed raw, as the
reference's k-mer iterator does).
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from make_golden import DATA, P, pack, read_fasta, read_fastq, ref, sha, synthetic  # noqa: E402


def repeats(seed, n):
    """Reads with tandem repeats, duplicates, short and N-holding reads."""
    rng = np.random.default_rng(seed)
    base = synthetic(seed + 1, 3000, 200, 40, 200, 0.01)
    out = []
    for i in range(n):
        kind = i % 7
        if kind == 0:
            out.append('A' * int(rng.integers(21, 700)))                   # one k-mer throughout
        elif kind == 1:
            unit = ''.join(rng.choice(list('ACGT'), int(rng.integers(1, 5))))
            out.append((unit * 400)[:int(rng.integers(30, 900))])           # short-period repeat
        elif kind == 2:
            out.append(base[int(rng.integers(0, 8))])                       # duplicates of a few reads
        elif kind == 3:
            out.append(''.join(rng.choice(list('ACGT'), int(rng.integers(0, 21)))))   # shorter than K
        elif kind == 4:
            r = list(base[int(rng.integers(0, len(base)))])
            for _ in range(int(rng.integers(1, 4))):
                r[int(rng.integers(0, len(r)))] = 'N'
            out.append(''.join(r))                                          # N hashed raw
        else:
            out.append(base[int(rng.integers(0, len(base)))])
    return out


def main():
    L = ref()
    L.ref_diginorm_reads.restype = C.c_uint64
    L.ref_diginorm_reads.argtypes = [C.c_void_p, C.c_uint, C.c_char_p, C.POINTER(C.c_int64), C.c_uint64,
                                     C.POINTER(C.c_uint8)]
    L.ref_free.restype = None
    L.ref_free.argtypes = [C.c_void_p]
    inputs = {
        'synth': synthetic(31, 4000, 3000, 60, 150, 0.01),
        'repeats': repeats(32, 700),
        'rand20': read_fasta(os.path.join(DATA, 'random-20-a.fa')),
        'fastq': read_fastq(os.path.join(DATA, 'test-fastq-reads.fq')),
        'long': synthetic(33, 30000, 60, 2000, 9000, 0.02),
    }
    cases = [
        # name, input, K, canonical, max_table, n_tables, cutoff
        ('synth_c5', 'synth', 21, 1, 20000, 4, 5),
        ('synth_c10', 'synth', 21, 1, 20000, 4, 10),
        ('synth_fwd_c3', 'synth', 25, 0, 20000, 3, 3),
        ('synth_small_c4', 'synth', 21, 1, 3000, 2, 4),
        ('repeats_c3', 'repeats', 21, 1, 5000, 2, 3),
        ('repeats_c1', 'repeats', 21, 1, 20000, 4, 1),
        ('rand20_c1', 'rand20', 21, 1, 2000, 4, 1),
        ('fastq_c2', 'fastq', 25, 1, 1000, 4, 2),
        ('fastq_fwd_c2', 'fastq', 21, 0, 5000, 4, 2),
        ('long_c2', 'long', 31, 1, 50000, 4, 2),
    ]
    out = {}
    for name, reads in inputs.items():
        out[f'in_{name}_seq'], out[f'in_{name}_off'] = pack(reads)
    for name, inp, K, can, mt, nt, cutoff in cases:
        seq, off = out[f'in_{inp}_seq'], out[f'in_{inp}_off']
        n = len(off) - 1
        h = L.ref_new(2, can, K, mt, nt)
        passed = np.zeros(n, np.uint8)
        L.ref_diginorm_reads(h, cutoff, seq.ctypes.data_as(C.c_char_p), P(off, C.c_int64), n, P(passed, C.c_uint8))
        u, o = C.c_uint64(), C.c_uint64()
        L.ref_counters(h, C.byref(u), C.byref(o))
        primes = np.zeros(nt, np.uint64)
        L.ref_primes(nt, mt, P(primes, C.c_uint64))
        out[f'{name}__params'] = np.array([K, can, mt, nt, cutoff], np.int64)
        out[f'{name}__input'] = np.frombuffer(inp.encode(), np.uint8)
        out[f'{name}__passed'] = passed
        out[f'{name}__counters'] = np.array([u.value, o.value], np.uint64)
        out[f'{name}__tablesizes'] = primes
        for i in range(nt):
            t = np.zeros(int(primes[i]), np.uint8)
            L.ref_table(h, i, int(primes[i]), P(t, C.c_uint8))
            out[f'{name}__table{i}_sha'] = np.frombuffer(bytes.fromhex(sha(t)), np.uint8)
            if mt <= 5000:
                out[f'{name}__table{i}'] = t
        print(name, 'passed', int(passed.sum()), 'of', n, 'n_unique', u.value)
        L.ref_free(h)
    np.savez_compressed(os.path.join(HERE, 'diginorm_ref.npz'), **out)


if __name__ == '__main__':
    sys.exit(main())
