"""Serving engine (gt_server_* of libgoetia_amd) and the 64-stream serving
benchmark.

Each sample stream is its own dBG over a Bloom or count-min storage.  The
engine keeps floor(budget / tables bytes) storages in HBM, evicts the least
recently used idle stream to pinned host memory, and runs chunks of many
streams concurrently on its execution slots (continuous batching), each
stream's chunks in submission order.  ``n_slots=1, policy=0`` is the plain
one-chunk-at-a-time path.
"""
import ctypes as C
import time

import numpy as np

from . import _native as N
from .reads import ReadBatch


class ServerStats(C.Structure):
    _fields_ = [('submitted', C.c_uint64), ('completed', C.c_uint64), ('loads', C.c_uint64),
                ('evictions', C.c_uint64), ('kmers', C.c_uint64), ('resident_slots', C.c_uint64),
                ('bytes_per_stream', C.c_uint64), ('latency_sum_s', C.c_double), ('elapsed_s', C.c_double)]


class Server:
    def __init__(self, storage_type, tablesizes, K, n_streams, hbm_budget, n_slots=4, policy=1,
                 canonical=True, mode=N.MODE_RAW):
        self._h = C.c_void_p()
        sizes = np.array(tablesizes, np.uint64)
        self.table_bytes = [s // 8 + 1 if storage_type.kind == N.BIT else s for s in tablesizes]
        N.call('gt_server_new', storage_type.kind, sizes.ctypes.data_as(N.u64p), len(sizes), K,
               int(canonical), mode, n_streams, int(hbm_budget), n_slots, policy, C.byref(self._h))
        self.K = K
        self._outs = {}

    def submit(self, stream, batch, n_new=None):
        if not isinstance(batch, ReadBatch):
            batch = ReadBatch.from_strings(list(batch))
        out = n_new if n_new is not None else np.zeros(max(batch.n_reads, 1), np.uint64)
        t = C.c_uint64()
        N.call('gt_server_submit', self._h, int(stream), N.ptr(batch.seq), N.ptr(batch.offsets),
               batch.n_reads, N.ptr(out), C.byref(t))
        self._outs[t.value] = out
        return t.value, out

    def wait(self, ticket):
        lat = C.c_double()
        N.call('gt_server_wait', self._h, int(ticket), C.byref(lat))
        self._outs.pop(ticket, None)
        return lat.value

    def drain(self):
        N.call('gt_server_drain', self._h)

    def stats(self):
        st = ServerStats()
        N.call('gt_server_get_stats', self._h, C.byref(st))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def stream_tables(self, stream):
        ctr = np.zeros(2, np.uint64)
        tabs = []
        for i, nb in enumerate(self.table_bytes):
            t = np.zeros(nb, np.uint8)
            N.call('gt_server_stream_table', self._h, int(stream), i, N.ptr(t), N.ptr(ctr))
            tabs.append(t)
        return tabs, [int(ctr[0]), int(ctr[1])]

    def close(self):
        if self._h and self._h.value:
            N.load().gt_server_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- benchmark

def make_streams(n_streams, seed=100, min_mbp=5, max_mbp=50, scale=1.0):
    from . import workload
    rng = np.random.default_rng(seed)
    sizes = rng.integers(min_mbp, max_mbp + 1, n_streams) * 1_000_000
    return [workload.genome(seed + i, max(50_000, int(sz * scale))) for i, sz in enumerate(sizes)]


def make_chunks(genomes, n_chunks, seed=7, lo=1000, hi=200_000, median=10_000, sigma=1.2,
                len_min=100, len_max=250, sub=0.01, scale=1.0):
    """(stream, ReadBatch) chunks: stream uniform, size clip(lognormal) reads."""
    from . import workload
    rng = np.random.default_rng(seed)
    streams = rng.integers(0, len(genomes), n_chunks)
    sizes = np.clip(np.rint(rng.lognormal(np.log(median), sigma, n_chunks)), lo, hi).astype(int)
    sizes = np.maximum(1, (sizes * scale).astype(int))
    out = []
    for i, (s, n) in enumerate(zip(streams, sizes)):
        out.append((int(s), workload.reads(seed * 100_003 + i, genomes[s], int(n), len_min, len_max, sub, 0.0)))
    return out


def run(server, chunks, arrivals=None):
    """Submits chunks at their arrival times (seconds from start; None = all
    at once), waits for all; returns latencies (s) and wall time."""
    tickets = []
    t0 = time.perf_counter()
    for i, (s, b) in enumerate(chunks):
        if arrivals is not None:
            delay = arrivals[i] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
        tickets.append(server.submit(s, b)[0])
    lat = [server.wait(t) for t in tickets]
    server.drain()
    return np.array(lat), time.perf_counter() - t0


def bench(args):
    return serving_64(args.scale)


def serving_64(scale=1.0, n_chunks=None):
    """The serving configuration: 64 streams of BitStorage(2^30, 4), K=31, a
    16 GiB HBM budget (32 resident streams), lognormal chunks; the plain
    one-chunk-at-a-time path against the served path (4 slots, resident
    first) on the same chunks, saturated and with Poisson arrivals at 80% of
    the plain path's capacity.  `n_chunks` (default 400) trims the run."""
    from . import libgoetia as lg
    n_streams = 64
    K = 31
    sizes = lg.get_n_primes_near_x(4, 2 ** 30 if scale >= 1 else max(2 ** 20, int(2 ** 30 * scale)))
    per_stream = sum(s // 8 + 1 for s in sizes)
    budget = int(16 * 2 ** 30 * min(1.0, scale)) if scale < 1 else 16 * 2 ** 30
    t0 = time.perf_counter()
    genomes = make_streams(n_streams, scale=scale)
    if n_chunks is None:
        n_chunks = max(16, int(400 * min(1.0, scale * 10)))
    chunks = make_chunks(genomes, n_chunks, scale=min(1.0, scale * 10))
    kmers = int(sum(int(b.kmer_counts(K).sum()) for _, b in chunks))
    gen_s = time.perf_counter() - t0

    def server(slots, policy):
        return Server(lg.BitStorage, sizes, K, n_streams, budget, n_slots=slots, policy=policy)

    # saturated: every chunk queued at t=0
    sv = server(1, 0)
    lat_p, wall_p = run(sv, chunks)
    st_p = sv.stats()
    plain_tables = [sv.stream_tables(s) for s in range(n_streams)]
    sv.close()
    sv = server(4, 1)
    lat_s, wall_s = run(sv, chunks)
    st_s = sv.stats()
    same = all(all(np.array_equal(a, b) for a, b in zip(sv.stream_tables(s)[0], plain_tables[s][0]))
               and sv.stream_tables(s)[1] == plain_tables[s][1] for s in range(n_streams))
    sv.close()
    # Poisson arrivals at 80% of the plain path's capacity
    rate = 0.8 * len(chunks) / wall_p
    arrivals = np.cumsum(np.random.default_rng(11).exponential(1.0 / rate, len(chunks)))
    sv = server(1, 0)
    lat_pp, wall_pp = run(sv, chunks, arrivals)
    sv.close()
    sv = server(4, 1)
    lat_sp, wall_sp = run(sv, chunks, arrivals)
    sv.close()

    def pct(x, q):
        return round(float(np.percentile(x, q)) * 1e3, 2)

    return {
        'metric': 'canonical k-mers/s insert, 64 sample streams, served vs plain (1 GPU)',
        'workload': 'serving_64', 'value': round(kmers / wall_s, 1), 'unit': 'kmers/s',
        'plain_kmers_per_s': round(kmers / wall_p, 1), 'served_over_plain': round(wall_p / wall_s, 3),
        'chunks': len(chunks), 'kmers': kmers, 'streams': n_streams,
        'storage': f'BitStorage({sizes}) per stream, {per_stream} bytes',
        'hbm_budget_bytes': budget, 'resident_streams': st_s['resident_slots'],
        'saturated': {'plain': {'wall_s': round(wall_p, 3), 'loads': st_p['loads'], 'evictions': st_p['evictions']},
                      'served': {'wall_s': round(wall_s, 3), 'loads': st_s['loads'], 'evictions': st_s['evictions'],
                                 'slots': 4, 'policy': 'resident-first, age cap 0.25 s'}},
        'poisson': {'rate_chunks_per_s': round(rate, 2),
                    'plain': {'p50_ms': pct(lat_pp, 50), 'p99_ms': pct(lat_pp, 99), 'wall_s': round(wall_pp, 3)},
                    'served': {'p50_ms': pct(lat_sp, 50), 'p99_ms': pct(lat_sp, 99), 'wall_s': round(wall_sp, 3)}},
        'served_equals_plain': bool(same), 'generation_s': round(gen_s, 1),
    }
