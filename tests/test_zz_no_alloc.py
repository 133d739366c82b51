"""gt_dbg_insert_noalloc_dev (the allocation-free insert) against
gt_dbg_insert_dev, through the C ABI: per-read n_new, per-k-mer new flags,
tables and counters are identical, on two calls in a row (the second with a
smaller input, so the shared scratch holds the first call's contents), for
count-min and Bloom storages, with batches spanning several sub-batches of
the fixed claim map; the device's free memory does not change across the
calls once the first call on the GPU has set its array up; oversized
inputs and the exact set are refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _insert(entry, store, K, d, n_new, flags, stream):
    from goetia_amd import _native as N
    N.call(entry, store.handle, K, 1, N.MODE_SKIP_N, N.ptr(d.seq), N.ptr(d.offsets), d.n_reads, d.n_bases,
           N.ptr(n_new), N.ptr(flags), stream)


@pytest.mark.parametrize('kind', ['ByteStorage', 'BitStorage'])
def test_noalloc_insert_equals_insert(kind):
    import torch
    from goetia_amd import libgoetia as lg
    from goetia_amd import workload
    K = 31
    genome = workload.genome(501, 400_000)
    big = workload.reads(502, genome, 40_000, 100, 250, 0.01, 0.0005)      # ~7M bases: 14 sub-batches
    small = workload.reads(503, genome, 3_000, 50, 200, 0.01, 0.0005)
    cls = getattr(lg, kind)
    sizes = lg.get_n_primes_near_x(4, 3_000_000)
    ref, na = cls(sizes), cls(sizes)
    stream = torch.cuda.current_stream().cuda_stream
    # The first call on a GPU takes the fixed array's address, which is when
    # the runtime backs it with memory (at most its 98 MiB, in 2 MiB pages);
    # after that no call may change the device's free memory.
    tiny = workload.reads(506, genome, 10, 100, 100, 0.0, 0.0).to_device()
    warm = cls(sizes)
    torch.cuda.synchronize()
    free_a = torch.cuda.mem_get_info()[0]
    _insert('gt_dbg_insert_noalloc_dev', warm, K, tiny, None, None, stream)
    torch.cuda.synchronize()
    assert free_a - torch.cuda.mem_get_info()[0] <= 104 << 20
    for b in (big, small):
        d = b.to_device()
        nk = int(b.kmer_counts(K).sum())
        out = {}
        for name, st in (('gt_dbg_insert_dev', ref), ('gt_dbg_insert_noalloc_dev', na)):
            n_new = torch.full((b.n_reads,), 7, dtype=torch.int64, device='cuda')
            flags = torch.full((nk,), 3, dtype=torch.uint8, device='cuda')
            torch.cuda.synchronize()
            free0 = torch.cuda.mem_get_info()[0]
            _insert(name, st, K, d, n_new, flags, stream)
            torch.cuda.synchronize()
            if name == 'gt_dbg_insert_noalloc_dev':
                assert torch.cuda.mem_get_info()[0] == free0, 'the allocation-free insert allocated device memory'
            out[name] = (n_new.cpu().numpy(), flags.cpu().numpy())
        a, b2 = out['gt_dbg_insert_dev'], out['gt_dbg_insert_noalloc_dev']
        np.testing.assert_array_equal(a[0], b2[0])
        np.testing.assert_array_equal(a[1], b2[1])
        assert a[0].sum() > 0
        assert ref.n_unique_kmers() == na.n_unique_kmers() and ref.n_occupied() == na.n_occupied()
        for x, y in zip(ref.get_raw_tables(), na.get_raw_tables()):
            np.testing.assert_array_equal(x, y)


def test_noalloc_refusals():
    import torch
    from goetia_amd import _native as N
    from goetia_amd import libgoetia as lg
    from goetia_amd import workload
    b = workload.reads(504, workload.genome(505, 10_000), 100, 150, 150, 0.0, 0.0)
    d = b.to_device()
    stream = torch.cuda.current_stream().cuda_stream
    byte = lg.ByteStorage(lg.get_n_primes_near_x(4, 100_000))
    with pytest.raises(ValueError, match='2\\^30'):
        N.call('gt_dbg_insert_noalloc_dev', byte.handle, 31, 1, 0, N.ptr(d.seq), N.ptr(d.offsets), d.n_reads,
               (1 << 30) + 1, None, None, stream)
    with pytest.raises(ValueError, match='2\\^22'):
        N.call('gt_dbg_insert_noalloc_dev', byte.handle, 31, 1, 0, N.ptr(d.seq), N.ptr(d.offsets), (1 << 22) + 1,
               d.n_bases, None, C.c_void_p(1), stream)
    exact = lg.SparseppSetStorage.build()
    with pytest.raises(N.UnsupportedError):
        N.call('gt_dbg_insert_noalloc_dev', exact.handle, 31, 1, 0, N.ptr(d.seq), N.ptr(d.offsets), d.n_reads,
               d.n_bases, None, None, stream)
