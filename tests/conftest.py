import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN_DIR = os.path.join(ROOT, 'tests', 'golden')


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an AMD GPU (gfx950) and the built HIP library')
    config.addinivalue_line('markers', 'slow: large inputs')


@pytest.fixture(scope='session')
def golden():
    with np.load(os.path.join(GOLDEN_DIR, 'golden_ref.npz'), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope='session')
def golden_meta():
    import json
    with open(os.path.join(GOLDEN_DIR, 'golden_ref.json')) as f:
        return json.load(f)


@pytest.fixture(scope='session')
def oracle():
    from oracle import pyoracle
    pyoracle.lib()
    return pyoracle


def read_fasta(path):
    seqs, cur = [], []
    for line in open(path):
        line = line.strip()
        if line.startswith('>'):
            if cur:
                seqs.append(''.join(cur))
            cur = []
        elif line:
            cur.append(line)
    if cur:
        seqs.append(''.join(cur))
    return seqs


def read_fastq(path):
    lines = [l.rstrip('\n') for l in open(path)]
    return [lines[i + 1] for i in range(0, len(lines) - 3, 4)]


def golden_input(name, golden):
    """Reads of a golden run, as a list of str."""
    prefix = name.split('_')[0]
    if prefix == 'rand20':
        return read_fasta(os.path.join(GOLDEN_DIR, 'data', 'random-20-a.fa'))
    if prefix == 'fastq':
        return read_fastq(os.path.join(GOLDEN_DIR, 'data', 'test-fastq-reads.fq'))
    seq, off = golden[f'{prefix}_seq'], golden[f'{prefix}_off']
    return [seq[off[i]:off[i + 1]].tobytes().decode() for i in range(len(off) - 1)]


def random_dna(rng, n):
    return ''.join(rng.choice(list('ACGT'), n))


def revcomp(s):
    return s[::-1].translate(str.maketrans('ACGT', 'TGCA'))
