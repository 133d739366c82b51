// gt_serving.cpp -- serving engine of goetia_amd (gt_server_*, declared in
// include/goetia_amd.h).  A runtime layer above the C ABI: it only uses the
// public gt_* entry points plus HIP streams, events and copies.
//
// Many sample streams, each its own dBG storage, send chunks of reads.  The
// engine keeps a fixed number of storages resident in HBM (the memory
// budget divided by one storage's tables), binds streams to them on demand
// and evicts the least recently used idle stream to pinned host memory when
// a new one needs a place (evictions on a D2H stream, loads on an H2D
// stream, so both link directions work at once).  A scheduler thread runs
// continuous batching: whenever one of its execution slots (a HIP stream
// plus a gt_workspace) is free, it starts a pending chunk whose stream is
// idle -- the oldest one, or under the resident-first policy the oldest one
// whose tables are already in HBM -- so chunks of different streams overlap
// on the device and each stream's chunks keep their order.  Results of a
// stream are therefore the same whatever else is in flight.
#include "goetia_amd.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;

// A runnable chunk older than this is started before younger chunks of
// resident streams (resident-first policy).
constexpr double kAgeCapS = 0.25;

thread_local std::string g_server_err;

int sfail(int code, const std::string& msg) {
    g_server_err = msg;
    return code;
}

#define SV_HIP(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) return sfail(GT_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define SV_TRY(call)                                              \
    do {                                                          \
        int rc_ = (call);                                         \
        if (rc_ != GT_OK) return sfail(rc_, gt_last_error());     \
    } while (0)

struct Chunk {
    uint64_t ticket = 0;
    uint32_t stream = 0;
    std::vector<uint8_t> seq;
    std::vector<int64_t> off;
    uint64_t n_reads = 0;
    uint64_t* user_out = nullptr;
    Clock::time_point t_submit, t_start, t_done;
    bool done = false;
    int error = GT_OK;
};

struct StreamState {
    int tslot = -1;                    // table slot while resident
    bool has_host = false;             // host copy holds its tables
    hipEvent_t saved = nullptr;        // its last eviction copy (save stream)
    std::vector<uint8_t*> host;        // pinned copies of the tables
    uint64_t host_ctr[2] = {0, 0};
    std::deque<Chunk*> queue;          // its pending chunks, in order
    bool busy = false;                 // a chunk is in flight
    uint64_t last_use = 0;
};

struct TableSlot {
    gt_storage* st = nullptr;
    int owner = -1;
    hipEvent_t ready = nullptr;        // tables loaded
};

struct ExecSlot {
    hipStream_t stream = nullptr;
    gt_workspace* ws = nullptr;
    uint8_t* d_seq = nullptr;
    int64_t* d_off = nullptr;
    uint64_t* d_nnew = nullptr;
    uint8_t* h_seq = nullptr;
    int64_t* h_off = nullptr;
    uint64_t* h_nnew = nullptr;
    uint64_t cap_bases = 0, cap_reads = 0;
    hipEvent_t done = nullptr;
    Chunk* chunk = nullptr;
};

}  // namespace

struct gt_server {
    int kind = 0;
    std::vector<uint64_t> sizes;
    std::vector<uint64_t> table_bytes;
    uint16_t K = 0;
    int canonical = 1, mode = 0;
    std::vector<StreamState> streams;
    std::vector<TableSlot> tslots;
    std::vector<ExecSlot> xslots;
    hipStream_t save = nullptr, load = nullptr;   // D2H evictions, H2D loads
    int policy = 0;                    // bit 0: resident streams first
    std::deque<Chunk*> pending;        // all pending chunks, arrival order
    std::unordered_map<uint64_t, Chunk*> chunks;
    std::mutex mu;
    std::condition_variable cv;
    std::thread worker;
    int device = 0;                    // the device the server was created on
    bool stop = false;
    uint64_t next_ticket = 1;
    uint64_t clock = 0;
    gt_server_stats stats{};
    Clock::time_point t0;
    int fatal = GT_OK;
    std::string fatal_msg;
};

namespace {

int grow_exec(ExecSlot& x, uint64_t bases, uint64_t reads) {
    if (bases > x.cap_bases) {
        if (x.d_seq) (void)hipFree(x.d_seq);
        if (x.h_seq) (void)hipHostFree(x.h_seq);
        x.d_seq = nullptr;
        x.h_seq = nullptr;
        const uint64_t c = std::max<uint64_t>(bases, 1 << 20);
        SV_HIP(hipMalloc((void**)&x.d_seq, c));
        SV_HIP(hipHostMalloc((void**)&x.h_seq, c));
        x.cap_bases = c;
    }
    if (reads > x.cap_reads) {
        if (x.d_off) (void)hipFree(x.d_off);
        if (x.d_nnew) (void)hipFree(x.d_nnew);
        if (x.h_off) (void)hipHostFree(x.h_off);
        if (x.h_nnew) (void)hipHostFree(x.h_nnew);
        x.d_off = nullptr;
        x.d_nnew = nullptr;
        x.h_off = nullptr;
        x.h_nnew = nullptr;
        const uint64_t c = std::max<uint64_t>(reads, 1 << 14);
        SV_HIP(hipMalloc((void**)&x.d_off, (c + 1) * sizeof(int64_t)));
        SV_HIP(hipMalloc((void**)&x.d_nnew, c * sizeof(uint64_t)));
        SV_HIP(hipHostMalloc((void**)&x.h_off, (c + 1) * sizeof(int64_t)));
        SV_HIP(hipHostMalloc((void**)&x.h_nnew, c * sizeof(uint64_t)));
        x.cap_reads = c;
    }
    return GT_OK;
}

// Binds stream `sid` to a table slot, evicting the least recently used idle
// stream when all are taken.  Returns false when every slot is in use by a
// busy stream (the chunk waits).
bool make_resident(gt_server* S, uint32_t sid, int* rc) {
    StreamState& ss = S->streams[sid];
    if (ss.tslot >= 0) return true;
    int pick = -1;
    for (size_t i = 0; i < S->tslots.size(); ++i) {
        if (S->tslots[i].owner < 0) {
            pick = (int)i;
            break;
        }
    }
    if (pick < 0) {
        uint64_t best = ~0ULL;
        for (size_t i = 0; i < S->tslots.size(); ++i) {
            const StreamState& o = S->streams[S->tslots[i].owner];
            if (!o.busy && o.last_use < best) {
                best = o.last_use;
                pick = (int)i;
            }
        }
        if (pick < 0) return false;
        // evict: its tables go to pinned host memory on the save stream
        TableSlot& ts = S->tslots[pick];
        StreamState& o = S->streams[ts.owner];
        *rc = gt_storage_save_async(ts.st, o.host.data(), o.host_ctr, S->save);
        if (*rc != GT_OK) return false;
        if (hipEventRecord(o.saved, S->save) != hipSuccess ||
            hipStreamWaitEvent(S->load, o.saved, 0) != hipSuccess) {   // the slot is overwritten after the save
            *rc = GT_ERR_DEVICE;
            return false;
        }
        o.has_host = true;
        o.tslot = -1;
        ts.owner = -1;
        S->stats.evictions++;
    }
    TableSlot& ts = S->tslots[pick];
    if (ss.has_host && hipStreamWaitEvent(S->load, ss.saved, 0) != hipSuccess) {   // its own last save
        *rc = GT_ERR_DEVICE;
        return false;
    }
    *rc = gt_storage_load_async(ts.st, ss.has_host ? (const uint8_t* const*)ss.host.data() : nullptr, ss.host_ctr,
                                S->load);
    if (*rc != GT_OK) return false;
    if (hipEventRecord(ts.ready, S->load) != hipSuccess) {
        *rc = GT_ERR_DEVICE;
        return false;
    }
    if (ss.has_host) S->stats.loads++;
    ts.owner = (int)sid;
    ss.tslot = pick;
    return true;
}

int dispatch(gt_server* S, ExecSlot& x, Chunk* c) {
    StreamState& ss = S->streams[c->stream];
    TableSlot& ts = S->tslots[ss.tslot];
    const uint64_t n_bases = c->off.empty() ? 0 : (uint64_t)c->off.back();
    SV_TRY(grow_exec(x, n_bases, c->n_reads));
    std::memcpy(x.h_seq, c->seq.data(), n_bases);
    std::memcpy(x.h_off, c->off.data(), (c->n_reads + 1) * sizeof(int64_t));
    SV_HIP(hipStreamWaitEvent(x.stream, ts.ready, 0));
    SV_HIP(hipMemcpyAsync(x.d_seq, x.h_seq, n_bases, hipMemcpyHostToDevice, x.stream));
    SV_HIP(hipMemcpyAsync(x.d_off, x.h_off, (c->n_reads + 1) * sizeof(int64_t), hipMemcpyHostToDevice, x.stream));
    SV_TRY(gt_dbg_insert_ws_dev(ts.st, x.ws, S->K, S->canonical, S->mode, x.d_seq, x.d_off, c->n_reads, n_bases,
                                x.d_nnew, nullptr, x.stream));
    SV_HIP(hipMemcpyAsync(x.h_nnew, x.d_nnew, c->n_reads * sizeof(uint64_t), hipMemcpyDeviceToHost, x.stream));
    SV_HIP(hipEventRecord(x.done, x.stream));
    c->t_start = Clock::now();
    x.chunk = c;
    ss.busy = true;
    ss.last_use = ++S->clock;
    return GT_OK;
}

void finish(gt_server* S, ExecSlot& x) {
    Chunk* c = x.chunk;
    if (c->user_out) std::memcpy(c->user_out, x.h_nnew, c->n_reads * sizeof(uint64_t));
    c->t_done = Clock::now();
    c->done = true;
    c->seq.clear();
    c->seq.shrink_to_fit();
    c->off.clear();
    c->off.shrink_to_fit();
    S->streams[c->stream].busy = false;
    S->stats.completed++;
    S->stats.latency_sum_s += std::chrono::duration<double>(c->t_done - c->t_submit).count();
    x.chunk = nullptr;
}

void worker_loop(gt_server* S) {
    (void)hipSetDevice(S->device);     // checked when the server was created
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        bool progress = false;
        bool inflight = false;
        for (auto& x : S->xslots) {
            if (!x.chunk) continue;
            const hipError_t q = hipEventQuery(x.done);
            if (q == hipSuccess) {
                finish(S, x);
                progress = true;
            } else if (q != hipErrorNotReady) {
                S->fatal = GT_ERR_DEVICE;
                S->fatal_msg = std::string("chunk failed on the device: ") + hipGetErrorString(q);
                x.chunk->error = GT_ERR_DEVICE;
                finish(S, x);
                progress = true;
            } else {
                inflight = true;
            }
        }
        for (auto& x : S->xslots) {
            if (x.chunk || S->fatal != GT_OK) continue;
            Chunk* pick = nullptr;
            if (S->policy & 1) {
                // resident streams first (no table copy), unless the oldest
                // runnable chunk has waited past the age cap
                const auto now = Clock::now();
                for (Chunk* c : S->pending) {
                    StreamState& ss = S->streams[c->stream];
                    if (ss.busy || ss.queue.front() != c) continue;
                    if (std::chrono::duration<double>(now - c->t_submit).count() > kAgeCapS) break;
                    if (ss.tslot >= 0) {
                        pick = c;
                        break;
                    }
                }
            }
            for (Chunk* c : S->pending) {
                if (pick) break;
                StreamState& ss = S->streams[c->stream];
                if (ss.busy || ss.queue.front() != c) continue;
                int rc = GT_OK;
                if (!make_resident(S, c->stream, &rc)) {
                    if (rc != GT_OK) {
                        S->fatal = rc;
                        S->fatal_msg = gt_last_error();
                    }
                    continue;
                }
                pick = c;
                break;
            }
            if (!pick) break;
            S->pending.erase(std::find(S->pending.begin(), S->pending.end(), pick));
            S->streams[pick->stream].queue.pop_front();
            const int rc = dispatch(S, x, pick);
            if (rc != GT_OK) {
                S->fatal = rc;
                S->fatal_msg = g_server_err;
                pick->error = rc;
                pick->done = true;
            }
            inflight = true;
            progress = true;
        }
        if (progress) S->cv.notify_all();
        if (S->stop && S->pending.empty() && !inflight) break;
        if (!progress) S->cv.wait_for(lk, std::chrono::microseconds(inflight ? 20 : 500));
    }
    S->cv.notify_all();
}

}  // namespace

extern "C" {

const char* gt_server_last_error(void) { return g_server_err.c_str(); }

int gt_server_new(int kind, const uint64_t* tablesizes, uint32_t n_tables, uint16_t K, int canonical, int mode,
                  uint32_t n_streams, uint64_t hbm_budget, uint32_t n_slots, int policy, gt_server** out) {
    if (!out) return sfail(GT_ERR_INVALID_ARG, "null output");
    *out = nullptr;
    if (kind != GT_BIT_STORAGE && kind != GT_BYTE_STORAGE)
        return sfail(GT_ERR_INVALID_ARG, "the server keeps Bloom or count-min storages");
    if (n_streams == 0 || n_slots == 0 || n_slots > 64) return sfail(GT_ERR_INVALID_ARG, "invalid stream or slot count");
    SV_TRY(gt_device_init(0));
    gt_server* S = new (std::nothrow) gt_server();
    if (!S) return sfail(GT_ERR_OUT_OF_MEMORY, "host allocation failed");
    S->kind = kind;
    S->sizes.assign(tablesizes, tablesizes + n_tables);
    S->K = K;
    S->canonical = canonical ? 1 : 0;
    S->mode = mode;
    S->policy = policy;
    uint64_t per_stream = 0;
    for (uint32_t i = 0; i < n_tables; ++i) {
        const uint64_t nb = kind == GT_BIT_STORAGE ? tablesizes[i] / 8 + 1 : tablesizes[i];
        S->table_bytes.push_back(nb);
        per_stream += nb;
    }
    const uint64_t n_resident = std::max<uint64_t>(1, std::min<uint64_t>(n_streams, hbm_budget / per_stream));
    int rc = GT_OK;
    auto hip = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && rc == GT_OK) rc = sfail(GT_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
    };
    hip(hipStreamCreateWithFlags(&S->save, hipStreamNonBlocking), "hipStreamCreate");
    hip(hipStreamCreateWithFlags(&S->load, hipStreamNonBlocking), "hipStreamCreate");
    S->tslots.resize(n_resident);
    for (auto& ts : S->tslots) {
        if (rc != GT_OK) break;
        if (gt_storage_new(kind, tablesizes, n_tables, 0, &ts.st) != GT_OK) rc = sfail(GT_ERR_OUT_OF_MEMORY, gt_last_error());
        hip(hipEventCreateWithFlags(&ts.ready, hipEventDisableTiming), "hipEventCreate");
    }
    S->streams.resize(n_streams);
    for (auto& ss : S->streams) {
        if (rc != GT_OK) break;
        for (uint64_t nb : S->table_bytes) {
            uint8_t* p = nullptr;
            hip(hipHostMalloc((void**)&p, nb), "hipHostMalloc");
            ss.host.push_back(p);
        }
        hip(hipEventCreateWithFlags(&ss.saved, hipEventDisableTiming), "hipEventCreate");
    }
    S->xslots.resize(n_slots);
    for (auto& x : S->xslots) {
        if (rc != GT_OK) break;
        hip(hipStreamCreateWithFlags(&x.stream, hipStreamNonBlocking), "hipStreamCreate");
        hip(hipEventCreateWithFlags(&x.done, hipEventDisableTiming), "hipEventCreate");
        if (rc == GT_OK && gt_workspace_new(&x.ws) != GT_OK) rc = sfail(GT_ERR_OUT_OF_MEMORY, gt_last_error());
    }
    hip(hipGetDevice(&S->device), "hipGetDevice");
    S->stats.resident_slots = n_resident;
    S->stats.bytes_per_stream = per_stream;
    if (rc != GT_OK) {
        gt_server_free(S);
        return rc;
    }
    S->t0 = Clock::now();
    S->worker = std::thread(worker_loop, S);
    *out = S;
    return GT_OK;
}

int gt_server_submit(gt_server* S, uint32_t stream, const uint8_t* seq, const int64_t* offsets, uint64_t n_reads,
                     uint64_t* n_new, uint64_t* ticket) {
    if (!S || stream >= S->streams.size()) return sfail(GT_ERR_INVALID_ARG, "invalid server or stream");
    if (!offsets || offsets[0] != 0) return sfail(GT_ERR_INVALID_ARG, "offsets must start at 0");
    for (uint64_t r = 0; r < n_reads; ++r) {
        if (offsets[r + 1] < offsets[r]) return sfail(GT_ERR_INVALID_ARG, "offsets must be non-decreasing");
    }
    Chunk* c = new (std::nothrow) Chunk();
    if (!c) return sfail(GT_ERR_OUT_OF_MEMORY, "host allocation failed");
    const uint64_t n_bases = (uint64_t)offsets[n_reads];
    c->seq.assign(seq, seq + n_bases);
    c->off.assign(offsets, offsets + n_reads + 1);
    c->n_reads = n_reads;
    c->stream = stream;
    c->user_out = n_new;
    c->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->fatal != GT_OK) {
            delete c;
            return sfail(S->fatal, S->fatal_msg);
        }
        c->ticket = S->next_ticket++;
        S->chunks[c->ticket] = c;
        S->pending.push_back(c);
        S->streams[stream].queue.push_back(c);
        S->stats.submitted++;
        uint64_t nk = 0;
        for (uint64_t r = 0; r < n_reads; ++r) {
            const int64_t len = offsets[r + 1] - offsets[r];
            if (len >= S->K) nk += (uint64_t)(len - S->K + 1);
        }
        S->stats.kmers += nk;
    }
    S->cv.notify_all();
    if (ticket) *ticket = c->ticket;
    return GT_OK;
}

int gt_server_wait(gt_server* S, uint64_t ticket, double* latency_s) {
    if (!S) return sfail(GT_ERR_INVALID_ARG, "invalid server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->chunks.find(ticket);
    if (it == S->chunks.end()) return sfail(GT_ERR_INVALID_ARG, "unknown ticket");
    Chunk* c = it->second;
    S->cv.wait(lk, [&] { return c->done || S->fatal != GT_OK; });
    if (!c->done) return sfail(S->fatal, S->fatal_msg);
    const int err = c->error;
    if (latency_s) *latency_s = std::chrono::duration<double>(c->t_done - c->t_submit).count();
    S->chunks.erase(it);
    delete c;
    return err == GT_OK ? GT_OK : sfail(err, S->fatal_msg);
}

int gt_server_drain(gt_server* S) {
    if (!S) return sfail(GT_ERR_INVALID_ARG, "invalid server");
    std::unique_lock<std::mutex> lk(S->mu);
    S->cv.wait(lk, [&] {
        if (S->fatal != GT_OK) return true;
        if (!S->pending.empty()) return false;
        for (auto& x : S->xslots) {
            if (x.chunk) return false;
        }
        return true;
    });
    if (S->fatal != GT_OK) return sfail(S->fatal, S->fatal_msg);
    SV_HIP(hipStreamSynchronize(S->save));
    SV_HIP(hipStreamSynchronize(S->load));
    return GT_OK;
}

int gt_server_get_stats(gt_server* S, gt_server_stats* out) {
    if (!S || !out) return sfail(GT_ERR_INVALID_ARG, "invalid arguments");
    std::lock_guard<std::mutex> lk(S->mu);
    *out = S->stats;
    out->elapsed_s = std::chrono::duration<double>(Clock::now() - S->t0).count();
    return GT_OK;
}

int gt_server_stream_table(gt_server* S, uint32_t stream, uint32_t i, uint8_t* host_dst, uint64_t* counters) {
    if (!S || stream >= S->streams.size() || i >= S->sizes.size() || !host_dst)
        return sfail(GT_ERR_INVALID_ARG, "invalid arguments");
    int rc = gt_server_drain(S);
    if (rc != GT_OK) return rc;
    std::lock_guard<std::mutex> lk(S->mu);
    StreamState& ss = S->streams[stream];
    if (ss.tslot >= 0) {
        gt_storage* st = S->tslots[ss.tslot].st;
        SV_TRY(gt_storage_read_table(st, i, host_dst));
        if (counters) SV_TRY(gt_storage_counters(st, &counters[0], &counters[1]));
    } else if (ss.has_host) {
        std::memcpy(host_dst, ss.host[i], S->table_bytes[i]);
        if (counters) std::memcpy(counters, ss.host_ctr, sizeof(ss.host_ctr));
    } else {
        std::memset(host_dst, 0, S->table_bytes[i]);
        if (counters) counters[0] = counters[1] = 0;
    }
    return GT_OK;
}

int gt_server_free(gt_server* S) {
    if (!S) return GT_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv.notify_all();
    if (S->worker.joinable()) S->worker.join();
    (void)hipDeviceSynchronize();
    for (auto& x : S->xslots) {
        if (x.ws) gt_workspace_free(x.ws);
        if (x.d_seq) (void)hipFree(x.d_seq);
        if (x.d_off) (void)hipFree(x.d_off);
        if (x.d_nnew) (void)hipFree(x.d_nnew);
        if (x.h_seq) (void)hipHostFree(x.h_seq);
        if (x.h_off) (void)hipHostFree(x.h_off);
        if (x.h_nnew) (void)hipHostFree(x.h_nnew);
        if (x.done) (void)hipEventDestroy(x.done);
        if (x.stream) (void)hipStreamDestroy(x.stream);
    }
    for (auto& ts : S->tslots) {
        if (ts.st) gt_storage_free(ts.st);
        if (ts.ready) (void)hipEventDestroy(ts.ready);
    }
    for (auto& ss : S->streams) {
        for (uint8_t* p : ss.host) {
            if (p) (void)hipHostFree(p);
        }
        if (ss.saved) (void)hipEventDestroy(ss.saved);
    }
    for (auto& kv : S->chunks) delete kv.second;
    if (S->save) (void)hipStreamDestroy(S->save);
    if (S->load) (void)hipStreamDestroy(S->load);
    delete S;
    return GT_OK;
}

}  // extern "C"
