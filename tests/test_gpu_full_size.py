"""The HIP path at full size against the oracle's digests
(tests/golden/config_digests.json, written by
tests/golden/make_config_digests.py on the host, where the sequential oracle
takes about half an hour):

cms_20m       20M x 150 bp reads, K=31 canonical, skip-N, ByteStorage(2^28, 4):
              per-read n_new, the four tables, n_unique / n_occupied, the
              query counts of every window, and the one-pass insert_and_query
              counts (plus its tables) on fresh tables.
neighbors_5m  5M x 150 bp reads with repeats and SNP bubbles, K=31,
              ByteStorage(2^28, 4): the first-touch flag of every window, in /
              out degrees at count >= 2 of every new window, and the
              decision-node set.

Everything is integer work, so every digest must match exactly."""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN_DIR

pytestmark = pytest.mark.gpu

DIGESTS = json.load(open(os.path.join(GOLDEN_DIR, 'config_digests.json')))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def input_digest(batch):
    h = hashlib.sha256(np.ascontiguousarray(batch.seq, np.uint8).tobytes())
    h.update(np.ascontiguousarray(batch.offsets, '<i8').tobytes())
    return h.hexdigest()


def test_cms_20m_full_size():
    import torch
    from goetia_amd import _native as N
    from goetia_amd import libgoetia as lg
    from goetia_amd import workload
    exp = DIGESTS['cms_20m']
    cfg, batch = workload.make('cms_20m')
    print(f'cms_20m: {batch.n_reads} reads generated', flush=True)
    assert input_digest(batch) == exp['input_sha256'], 'generated reads differ from the digested ones'
    K = cfg['K']
    dev = torch.device('cuda:0')
    reads = batch.to_device(dev)
    nk = int(batch.kmer_counts(K).sum())
    assert nk == exp['windows']
    store = lg.ByteStorage.build(cfg['max_table'], cfg['n_tables'])
    assert store.get_tablesizes() == exp['tablesizes']
    g = lg.dBG[lg.ByteStorage, lg.CanLemireShifter].build(store, K)
    g.mode = N.MODE_SKIP_N
    n_new = torch.zeros(batch.n_reads, dtype=torch.int64, device=dev)
    g.insert_device(reads, n_new=n_new)
    torch.cuda.synchronize()
    nn = n_new.cpu().numpy()
    assert int(nn.sum()) == exp['n_new_sum']
    assert sha(nn.astype('<u8')) == exp['n_new_sha256']
    assert [g.n_unique(), g.n_occupied()] == [exp['n_unique'], exp['n_occupied']]
    assert [sha(t) for t in g.get_raw()] == exp['table_sha256']
    counts = torch.empty(nk, dtype=torch.int16, device=dev)
    g.query_device(reads, counts)
    torch.cuda.synchronize()
    q = counts.cpu().numpy()
    assert np.bincount(q[q >= 0].astype(np.int64), minlength=256).tolist() == exp['query_hist']
    assert sha(q.astype('<i2')) == exp['query_sha256']
    del q
    store.reset()
    g.insert_and_query_device(reads, counts)
    torch.cuda.synchronize()
    assert sha(counts.cpu().numpy().astype('<i2')) == exp['insert_and_query_sha256']
    assert [sha(t) for t in g.get_raw()] == exp['insert_and_query_table_sha256']


def test_neighbors_5m_full_size():
    import torch
    from goetia_amd import libgoetia as lg
    from goetia_amd import workload
    exp = DIGESTS['neighbors_5m']
    cfg, batch = workload.neighbors_5m()
    print(f'neighbors_5m: {batch.n_reads} reads generated', flush=True)
    assert input_digest(batch) == exp['input_sha256'], 'generated reads differ from the digested ones'
    K = cfg['K']
    dev = torch.device('cuda:0')
    reads = batch.to_device(dev)
    nk = int(batch.kmer_counts(K).sum())
    store = lg.ByteStorage.build(cfg['max_table'], cfg['n_tables'])
    g = lg.dBG[lg.ByteStorage, lg.CanLemireShifter].build(store, K)
    n_new = torch.zeros(batch.n_reads, dtype=torch.int64, device=dev)
    flags = torch.zeros(nk, dtype=torch.uint8, device=dev)
    g.insert_device(reads, n_new=n_new, new_flags=flags)
    ind = torch.empty(nk, dtype=torch.uint8, device=dev)
    outd = torch.empty(nk, dtype=torch.uint8, device=dev)
    g.degrees_device(reads, ind, outd, min_count=cfg['min_count'], only=flags)
    hashes = torch.empty(nk, dtype=torch.int64, device=dev)
    counts = torch.empty(nk, dtype=torch.int16, device=dev)
    g.query_device(reads, counts, hashes=hashes)
    torch.cuda.synchronize()
    fl, i, o = flags.cpu().numpy(), ind.cpu().numpy(), outd.cpu().numpy()
    assert sha(n_new.cpu().numpy().astype('<u8')) == exp['n_new_sha256']
    assert [g.n_unique(), g.n_occupied()] == [exp['n_unique'], exp['n_occupied']]
    assert [sha(t) for t in g.get_raw()] == exp['table_sha256']
    assert int(fl.sum()) == exp['new_windows'] and sha(fl) == exp['new_flags_sha256']
    assert sha(i) == exp['in_deg_sha256'] and sha(o) == exp['out_deg_sha256']
    dec = fl.astype(bool) & ((i > 1) | (o > 1))
    nodes = np.unique(hashes.cpu().numpy().view(np.uint64)[dec])
    assert (int(dec.sum()), len(nodes)) == (exp['decision_windows'], exp['decision_nodes'])
    assert sha(nodes.astype('<u8')) == exp['decision_nodes_sha256']
