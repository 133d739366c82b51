"""Paired-end input against goetia's compiled SplitPairedReader and
InserterProcessor (tests/golden/pairs_ref.json, written by
tests/golden/make_pairs_golden.py on goetia's left.fq / right.fq and crafted
edge-case and many-pair files).

Host (CPU): SplitPairedReader.next() gives goetia's pairs step by step for
every reader setting (strict, min_length, force_name_match): the same
sequence lengths, empty sides, rejected pairs and the point where mismatched
files or unpaired names raise; check_is_pair's name rules.
GPU: InserterProcessor.advance over the pairs returns goetia's
(n_sequences, time, remaining) at every interval, raises where goetia's does,
and leaves goetia's tables and counters; chunked_process(left, right) too."""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN_DIR

GOLD = json.load(open(os.path.join(GOLDEN_DIR, 'pairs_ref.json')))
DATA = os.path.join(GOLDEN_DIR, 'data')
FILES = {'goetia': ('left.fq', 'right.fq'), 'edge': ('pairs/edge_1.fq', 'pairs/edge_2.fq'),
         'many': ('pairs/many_1.fa', 'pairs/many_2.fa')}


def paths(key):
    lf, rf = FILES[key]
    return os.path.join(DATA, lf), os.path.join(DATA, rf)


@pytest.mark.parametrize('key', sorted(GOLD['walks']))
def test_pair_walk_vs_reference(key):
    from goetia_amd import _native as N
    from goetia_amd.parsing import SplitPairedReader
    files, strict, minlen, force = key.split('|')
    reader = SplitPairedReader(*paths(files), bool(int(strict)), int(minlen), bool(int(force)))
    got = []
    while not reader.is_complete():
        try:
            left, right = reader.next()
            got += [len(left.sequence) if left else -1, len(right.sequence) if right else -1]
        except N.InvalidCharacterException:
            got += [-2, -2]
        except N.GoetiaException:
            got += [-3, -3]
            break
    assert got == GOLD['walks'][key]


def test_pair_records_are_validated():
    from goetia_amd.parsing import SplitPairedReader
    reader = SplitPairedReader(*paths('edge'))
    pairs = [reader.next() for _ in range(6)]
    assert pairs[0][0].name == 'r1/1' and pairs[0][1].name == 'r1/2'
    assert pairs[1][0] is None and pairs[1][1].name == 'r2/2'            # N on the left
    assert pairs[5][0].sequence == pairs[5][0].sequence.upper()           # lowercase made upper
    assert set(pairs[5][0].sequence) <= set('ACGT')


@pytest.mark.gpu
@pytest.mark.parametrize('case', range(len(GOLD['inserts'])))
def test_paired_inserter_vs_reference(case):
    from goetia_amd import _native as N
    from goetia_amd import libgoetia as lg
    from goetia_amd.parsing import SplitPairedReader
    from goetia_amd.processors import InserterProcessor
    c = GOLD['inserts'][case]
    cls = lg.BitStorage if c['kind'] == 1 else lg.ByteStorage
    sh = lg.CanLemireShifter if c['canonical'] else lg.FwdLemireShifter
    g = lg.dBG[cls, sh].build(cls.build(c['max_table'], c['n_tables']), c['K'])
    proc = InserterProcessor(g, c['interval'])
    reader = SplitPairedReader(*paths(c['files']), bool(c['strict']), c['min_length'], bool(c['force_name_match']))
    steps, err = [], 0
    try:
        remaining = True
        while remaining:
            n, t, remaining = proc.advance(reader)
            steps += [n, t, int(remaining)]
    except N.GoetiaException:
        err = 1
    assert (steps, err) == (c['steps'], c['error'])
    assert (reader.n_skipped(), reader.n_parsed()) == (c['n_skipped'], c['n_parsed'])
    assert [g.n_unique(), g.n_occupied()] == c['counters']
    assert [hashlib.sha256(t.tobytes()).hexdigest() for t in g.get_raw()] == c['table_sha256']


@pytest.mark.gpu
def test_chunked_process_pairs():
    """chunked_process(left, right) (pythonize_processors.py:12-33) over
    goetia's paired files yields after every advance that consumed reads, and
    ends with goetia's totals and tables."""
    from goetia_amd import libgoetia as lg
    from goetia_amd.processors import InserterProcessor
    c = GOLD['inserts'][1]          # goetia's files, count-min, interval 500
    g = lg.dBG[lg.ByteStorage, lg.CanLemireShifter].build(lg.ByteStorage.build(c['max_table'], c['n_tables']),
                                                          c['K'])
    out = list(InserterProcessor(g, c['interval']).chunked_process(*paths('goetia')))
    steps = np.array(c['steps']).reshape(-1, 3)
    want, prev = [], 0
    for n, t, _ in steps:               # yielded when the advance consumed reads
        if n > prev:
            want.append((int(n), int(t)))
        prev = n
    assert [(n, t) for n, t, _ in out] == want
    assert [hashlib.sha256(t.tobytes()).hexdigest() for t in g.get_raw()] == c['table_sha256']
