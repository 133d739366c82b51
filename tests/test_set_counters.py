"""The exact set's counters against goetia's compiled storages
(tests/golden/set_trace_ref.npz, written by tests/golden/make_set_golden.py):
after every insert, n_unique_kmers() and n_occupied() -- sparsepp's
bucket_count(), which grows by doubling before an insert whose size + 1 would
pass half the buckets, also when the value is already present, and returns to
32 buckets on reset() -- and BTreeStorage's constant max_size().  The oracle
replays every trace insert by insert (CPU); the device storage replays them
in chunks of uneven size through the batch API and must agree at every chunk
end (GPU)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN_DIR

GOLD = np.load(os.path.join(GOLDEN_DIR, 'set_trace_ref.npz'), allow_pickle=False)
TRACES = ('boundary', 'resets', 'long')


def set_value(idx):
    """splitmix64 of pool indices, 0 and all-ones for indices 0 and 1 (the
    fixture's value encoding)."""
    idx = np.asarray(idx, np.uint64)
    with np.errstate(over='ignore'):
        z = idx + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    z = np.where(idx == 0, np.uint64(0), z)
    return np.where(idx == 1, np.uint64(2**64 - 1), z).astype(np.uint64)


def test_fixture_values():
    """The encoding reproduces splitmix64's published first output for seed 0
    (0xE220A8397B1DCDAF) and keeps the two special values."""
    assert int(set_value([0x0])[0]) == 0
    v = set_value([2, 1])
    assert int(v[1]) == 2**64 - 1
    # splitmix64 with state 0: the first output adds the increment once
    with np.errstate(over='ignore'):
        z = np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    assert int(z) == 0xE220A8397B1DCDAF


def test_sparsepp_growth_rule_in_fixture():
    """The compiled reference grew at the duplicate: 16 distinct values and
    one insert of a present value give 64 buckets."""
    no = GOLD['boundary__sparsepp__n_occupied']
    nu = GOLD['boundary__sparsepp__n_unique']
    assert int(nu[15]) == 16 and int(no[15]) == 32
    assert int(nu[16]) == 16 and int(no[16]) == 64


@pytest.mark.parametrize('name', TRACES)
def test_oracle_set_trace(oracle, name):
    vals = set_value(GOLD[f'{name}__idx'])
    reset = GOLD[f'{name}__reset_before']
    st = oracle.Storage(oracle.EXACT)
    got_new, got_u, got_o = [], [], []
    for v, r in zip(vals, reset):
        if r:
            st = oracle.Storage(oracle.EXACT)
        got_new.append(st.insert(int(v)))
        got_u.append(st.n_unique)
        got_o.append(st.n_occupied)
    np.testing.assert_array_equal(np.array(got_new, np.uint8), GOLD[f'{name}__sparsepp__is_new'])
    np.testing.assert_array_equal(np.array(got_u, np.uint64), GOLD[f'{name}__sparsepp__n_unique'])
    np.testing.assert_array_equal(np.array(got_o, np.uint64), GOLD[f'{name}__sparsepp__n_occupied'])


def test_btree_max_size_constant():
    for name in TRACES:
        assert set(GOLD[f'{name}__btree__n_occupied'].tolist()) == {2**63 - 1}
        np.testing.assert_array_equal(GOLD[f'{name}__btree__n_unique'], GOLD[f'{name}__sparsepp__n_unique'])


CHUNKS = (1, 3, 17, 1, 250, 2, 1000, 1, 64, 5)


@pytest.mark.gpu
@pytest.mark.parametrize('name', TRACES)
@pytest.mark.parametrize('cls', ['SparseppSetStorage', 'BTreeStorage'])
def test_device_set_trace(name, cls):
    from goetia_amd import libgoetia as lg
    flavor = 'sparsepp' if cls == 'SparseppSetStorage' else 'btree'
    vals = set_value(GOLD[f'{name}__idx'])
    reset = GOLD[f'{name}__reset_before']
    exp_new = GOLD[f'{name}__{flavor}__is_new'].astype(bool)
    exp_u = GOLD[f'{name}__{flavor}__n_unique']
    exp_o = GOLD[f'{name}__{flavor}__n_occupied']
    st = getattr(lg, cls).build()
    assert st.n_occupied() == (32 if flavor == 'sparsepp' else 2**63 - 1)
    i, c = 0, 0
    while i < len(vals):
        n = CHUNKS[c % len(CHUNKS)]
        c += 1
        j = min(len(vals), i + n)
        cut = np.nonzero(reset[i + 1:j])[0]      # a chunk never spans a reset
        if len(cut):
            j = i + 1 + int(cut[0])
        if reset[i]:
            st.reset()
        got = st.insert_hashes(vals[i:j])
        np.testing.assert_array_equal(got, exp_new[i:j], err_msg=f'{name} [{i},{j})')
        assert (st.n_unique_kmers(), st.n_occupied()) == (int(exp_u[j - 1]), int(exp_o[j - 1])), (name, i, j)
        i = j
