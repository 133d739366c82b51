// gt_capi.cpp -- C ABI of goetia_amd (declared in include/goetia_amd.h).
//
// Owns device storages (tables in HBM), per-storage scratch (claim map,
// fresh masks, read layout), the sub-batch loop that keeps the first-touch
// semantics of goetia's sequential inserts, and the host-pointer wrappers.
#include "goetia_amd.h"

#include "gt_hash.h"
#include "gt_internal.h"
#include "gt_launch.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

using namespace gt;

// ------------------------------------------------------------ errors

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define GT_HIP(call)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            return fail(GT_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
        }                                                                                   \
    } while (0)

#define GT_TRY(call)             \
    do {                         \
        int rc_ = (call);        \
        if (rc_ != GT_OK) return rc_; \
    } while (0)

// ------------------------------------------------------------ device

struct DeviceState {
    std::mutex mu;
    bool ready = false;
    int device = 0;
    int cus = 0;
    int grid = 0;
};

DeviceState g_dev;

int ensure_device(int device = -1) {
    std::lock_guard<std::mutex> lk(g_dev.mu);
    if (g_dev.ready && (device < 0 || device == g_dev.device)) {
        hipError_t e = hipSetDevice(g_dev.device);
        if (e != hipSuccess) return fail(GT_ERR_DEVICE, std::string("hipSetDevice: ") + hipGetErrorString(e));
        return GT_OK;
    }
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        return fail(GT_ERR_NO_DEVICE, "goetia_amd needs an AMD GPU (gfx950); none is visible");
    }
    if (device < 0) device = 0;
    if (device >= n) return fail(GT_ERR_INVALID_ARG, "device index out of range");
    GT_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    GT_HIP(hipGetDeviceProperties(&prop, device));
    g_dev.device = device;
    g_dev.cus = prop.multiProcessorCount;
    // Persistent tile grids: 8 workgroups of 4 waves per CU.
    g_dev.grid = std::max(1, prop.multiProcessorCount * 8);
    g_dev.ready = true;
    return GT_OK;
}

// ------------------------------------------------------------ device buffers

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    bool borrowed = false;             // a view of memory the buffer does not own
    ~DevBuf() { release(); }
    void release() {
        if (p && !borrowed) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        borrowed = false;
    }
    // Makes the buffer a view of n bytes at ptr; it never allocates or frees.
    void borrow(void* ptr, size_t n) {
        release();
        p = ptr;
        bytes = n;
        borrowed = true;
    }
    // Grows (never shrinks); contents are not kept.  A borrowed buffer only
    // checks that it is large enough.
    int ensure(size_t need) {
        if (need <= bytes && p) return GT_OK;
        if (borrowed)
            return fail(GT_ERR_INVALID_ARG, "the call needs " + std::to_string(need) + " bytes of a " +
                                                std::to_string(bytes) + "-byte fixed scratch region");
        release();
        size_t sz = std::max<size_t>(need, 256);
        hipError_t e = hipMalloc(&p, sz);
        if (e != hipSuccess) {
            p = nullptr;
            return fail(GT_ERR_OUT_OF_MEMORY, "hipMalloc of " + std::to_string(sz) + " bytes failed: " +
                                                  hipGetErrorString(e));
        }
        bytes = sz;
        return GT_OK;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// ------------------------------------------------------------ kernel timing

struct ProfRecord {
    const char* name;
    hipEvent_t start, stop;
};

struct ProfState {
    std::mutex mu;
    std::atomic<bool> on{false};
    std::vector<ProfRecord> pending;
    std::vector<std::pair<std::string, std::pair<uint64_t, double>>> totals;
};

ProfState g_prof;

// Brackets the kernel launches of its scope with events on `s` while timing
// is enabled.
struct ProfScope {
    const char* name;
    hipStream_t s;
    hipEvent_t start = nullptr;
    ProfScope(const char* n, hipStream_t st) : name(n), s(st) {
        if (!g_prof.on.load(std::memory_order_relaxed)) return;
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
        if (hipEventCreate(&start) != hipSuccess || hipEventRecord(start, s) != hipSuccess) start = nullptr;
    }
    ~ProfScope() {
        if (!start) return;
        hipEvent_t stop = nullptr;
        if (hipEventCreate(&stop) != hipSuccess || hipEventRecord(stop, s) != hipSuccess) {
            (void)hipEventDestroy(start);
            if (stop) (void)hipEventDestroy(stop);
            return;
        }
        std::lock_guard<std::mutex> lk(g_prof.mu);
        g_prof.pending.push_back({name, start, stop});
    }
};

void prof_resolve() {
    std::lock_guard<std::mutex> lk(g_prof.mu);
    for (auto& r : g_prof.pending) {
        float ms = 0.f;
        (void)hipEventSynchronize(r.stop);
        (void)hipEventElapsedTime(&ms, r.start, r.stop);
        bool found = false;
        for (auto& t : g_prof.totals) {
            if (t.first == r.name) {
                t.second.first += 1;
                t.second.second += ms;
                found = true;
            }
        }
        if (!found) g_prof.totals.push_back({r.name, {1, (double)ms}});
        (void)hipEventDestroy(r.start);
        (void)hipEventDestroy(r.stop);
    }
    g_prof.pending.clear();
}

uint64_t next_pow2(uint64_t x) {
    uint64_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

uint32_t log2u(uint64_t p) {
    uint32_t l = 0;
    while ((1ULL << l) < p) ++l;
    return l;
}

}  // namespace

// ------------------------------------------------------------ scratch

// Per-call device scratch: the first-touch claim map with its own
// generation counter, the fresh masks of one sub-batch, and the read layout
// (tile_read, kmer_off).  Owned by a storage (its default), by a pipeline, or
// by a gt_workspace that many storages share (the serving engine's slots).
struct Scratch {
    DevBuf claim_keys, claim_vals, claim_gen;
    uint64_t claim_cap = 0;
    uint64_t claim_gen_issued = 0;     // bumps enqueued so far (the device counter follows it)
    DevBuf fresh, tile_read, kmer_off, scan_blocks;
    // binned insert: events of the two partition passes, first-pass digit
    // offsets, second-pass counts / starts and their scan sums, new-window
    // flags and the sub-batch's first packed window index
    DevBuf bin_ent1, bin_ent2, bin_dig2, bin_off1, bin_cnt2, bin_sums, bin_flags, bin_kbase, bin_vq;
    bool fixed = false;                // buffers preallocated and never moved (captured graphs)
};

// ------------------------------------------------------------ storage

struct gt_storage {
    int kind = 0;
    uint32_t n_tables = 0;
    std::vector<uint64_t> sizes;       // bins per table
    std::vector<uint64_t> nbytes;      // reference layout bytes per table
    std::vector<uint64_t> table_off;   // offsets in d_tables (256-aligned)
    DevBuf tables;
    DevBuf counters;                   // kNCounters u64
    // exact set
    DevBuf ekv, eones;                     // exact set: key / stamp word pairs, all-ones flag
    uint64_t ecap = 0;
    uint64_t ecount_ub = 0;            // host upper bound on the set size
    Scratch own;                       // default scratch of calls on this storage
    DevBuf h_seq, h_off, h_out1, h_out2, h_out3, h_out4;   // staging of host-pointer calls
    DevBuf seg_nb, seg_pres, seg_idx, seg_nbc, seg_pc;     // find_new_segments' window view
    DevBuf solid_counts;               // insert_and_query counts of a solid-filter batch
    // batched diginorm: read of every routed event and window, windows with
    // events, windows failing the cutoff, per-read state, iteration control
    DevBuf dn_rid, dn_wread, dn_hasev, dn_fails, dn_state, dn_ctl;
    uint64_t gen_issued = 0;           // generation bumps enqueued so far (device kGen follows it)
    uint64_t epoch = 0;                // bumped when device table pointers change (exact-set growth)
    std::vector<struct gt_pipeline*> pipelines;   // attached pipelines
    // Work another stream queued on this storage without waiting for it (the
    // file inserter's): later device calls on the storage wait for this event.
    hipEvent_t fence = nullptr;
    bool fence_set = false;
    // Reads a file inserter has taken but not yet queued (it returned at an
    // interval and queues them with its next batch): every call on the
    // storage queues them first (StorageLock), so each call sees the storage
    // as if they had been inserted when the inserter returned.
    int (*flush_fn)(void*) = nullptr;
    void* flush_ctx = nullptr;
    // count-min bigcount (storage.hh:129-135): off by default; the map lives
    // on the device once bigcount has been used
    bool use_bigcount = false;
    BigMapDev big;
    std::recursive_mutex mu;

    Tables tables_view() const {
        Tables T{};
        T.n = n_tables;
        T.kind = kind;
        for (uint32_t i = 0; i < n_tables; ++i) {
            T.p[i] = sizes[i];
            T.m[i] = gt_mod_magic(sizes[i]);
            T.base[i] = tables.as<uint8_t>() + table_off[i];
        }
        return T;
    }
    ExactSet exact_view() const {
        ExactSet E{};
        E.kv = ekv.as<uint64_t>();
        E.ones = eones.as<uint32_t>();
        E.cap = ecap;
        E.log2cap = log2u(ecap);
        return E;
    }
    // Exact-set view usable by kernels of Bloom / count-min storages (never
    // dereferenced there).
    ExactSet exact_or_empty() const {
        if (kind == kExact) return exact_view();
        ExactSet E{};
        E.cap = 1;
        return E;
    }
};

// Batched insert engine (gt_pipeline_*): two slots of pinned + device
// buffers; each slot replays its own captured hipGraph.
struct gt_pipeline {
    struct Slot {
        uint8_t* h_seq = nullptr;
        int64_t* h_off = nullptr;
        int64_t* h_dyn = nullptr;
        uint64_t* h_nnew = nullptr;
        DevBuf d_seq, d_off, d_dyn, d_nnew;
        hipEvent_t copied = nullptr, done = nullptr;
        hipGraphExec_t exec = nullptr;
        uint64_t epoch = 0;
        uint64_t* user_out = nullptr;
        uint64_t n_reads = 0;
        bool busy = false;
    };
    gt_storage* st = nullptr;
    uint16_t K = 0;
    int canonical = 1, mode = 0;
    uint64_t max_reads = 0, max_bases = 0;
    hipStream_t compute = nullptr, copy = nullptr;
    Slot slot[2];
    int next = 0;
    Scratch sc;
};

// Shared scratch for storages that are used one call at a time (no
// counterpart in goetia): see gt_workspace_new.
struct gt_workspace {
    Scratch sc;
    std::recursive_mutex mu;
};

namespace {

bool valid_storage(const gt_storage* st) { return st && st->kind >= kBit && st->kind <= kExact; }

int alloc_exact(gt_storage* st, uint64_t cap, hipStream_t s) {
    DevBuf kv;
    GT_TRY(kv.ensure(2 * (cap + 1) * sizeof(uint64_t)));
    launch_fill_u64(kv.as<uint64_t>(), 2 * (cap + 1), ~0ULL, s);
    GT_HIP(hipGetLastError());
    if (st->ecap) {
        ExactSet from = st->exact_view();
        ExactSet to = from;
        to.kv = kv.as<uint64_t>();
        to.cap = cap;
        to.log2cap = log2u(cap);
        launch_exact_rehash(from, to, st->counters.as<uint64_t>(), s);
        GT_HIP(hipGetLastError());
        GT_HIP(hipStreamSynchronize(s));
    }
    std::swap(st->ekv.p, kv.p);
    std::swap(st->ekv.bytes, kv.bytes);
    st->ecap = cap;
    st->epoch++;
    return GT_OK;
}

int zero_tables(gt_storage* st, hipStream_t s) {
    if (st->kind == kExact) {
        launch_fill_u64(st->ekv.as<uint64_t>(), 2 * (st->ecap + 1), ~0ULL, s);
        launch_fill_u64(st->eones.as<uint64_t>(), 1, 0, s);
    } else {
        launch_fill_u64(st->tables.as<uint64_t>(), st->tables.bytes / 8, 0, s);
    }
    GT_HIP(hipGetLastError());
    return GT_OK;
}

// Queues the reads a file inserter holds for this storage (see flush_fn); the
// caller holds st->mu.
int flush_pending(gt_storage* st) {
    if (!st->flush_fn) return GT_OK;
    int (*fn)(void*) = st->flush_fn;
    void* ctx = st->flush_ctx;
    st->flush_fn = nullptr;            // the inserter's own calls below take the lock again
    st->flush_ctx = nullptr;
    return fn(ctx);
}

// The lock every call on a storage takes; it queues a file inserter's held
// reads first.  rc is that flush's status.
struct StorageLock {
    std::lock_guard<std::recursive_mutex> g;
    int rc;
    explicit StorageLock(gt_storage* st) : g(st->mu), rc(flush_pending(st)) {}
};

// Orders stream s after work another stream queued on the storage without
// waiting (storage_set_fence).
int order_after_fence(gt_storage* st, hipStream_t s) {
    if (st->fence_set) GT_HIP(hipStreamWaitEvent(s, st->fence, 0));
    return GT_OK;
}

int read_counters(gt_storage* st, uint64_t* out, hipStream_t s) {
    GT_HIP(hipMemcpyAsync(out, st->counters.p, kNCounters * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    GT_HIP(hipStreamSynchronize(s));
    return GT_OK;
}

int check_kernel_errors(gt_storage* st, hipStream_t s) {
    uint64_t c[kNCounters];
    GT_TRY(read_counters(st, c, s));
    if (c[kErrors]) {
        return fail(GT_ERR_INTERNAL, "device table probe limit reached (error bits " + std::to_string(c[kErrors]) +
                                         "); results of this storage are incomplete");
    }
    return GT_OK;
}

// Prepares tile_read (and kmer_off when wanted) for one call.
int prepare_reads(Scratch& sc, Reads& R, bool want_kmer_off, hipStream_t s) {
    const int64_t n_tiles = (R.n_bases + kTileGrain - 1) / kTileGrain;
    GT_TRY(sc.tile_read.ensure((size_t)(n_tiles + 1) * sizeof(uint32_t)));
    // tile_read entries are all written for well-formed offsets; zero first so
    // malformed offsets stay in bounds.
    launch_fill_u64(sc.tile_read.as<uint64_t>(), (sc.tile_read.bytes) / 8, 0, s);
    launch_tile_read(R.off, R.n_reads, R.n_bases, sc.tile_read.as<uint32_t>(), n_tiles, R.dyn, s);
    R.tile_read = sc.tile_read.as<uint32_t>();
    R.kmer_off = nullptr;
    if (want_kmer_off) {
        GT_TRY(sc.kmer_off.ensure((R.n_reads + 1) * sizeof(int64_t)));
        GT_TRY(sc.scan_blocks.ensure((scan_block_count(R.n_reads) + 1) * sizeof(int64_t)));
        launch_kmer_offsets(R.off, R.n_reads, R.K, sc.kmer_off.as<int64_t>(), sc.scan_blocks.as<int64_t>(), R.dyn,
                            s);
        R.kmer_off = sc.kmer_off.as<int64_t>();
    }
    GT_HIP(hipGetLastError());
    return GT_OK;
}

int prepare_reads(gt_storage* st, Reads& R, bool want_kmer_off, hipStream_t s) {
    return prepare_reads(st->own, R, want_kmer_off, s);
}

int check_read_args(uint16_t K, int mode, const void* seq, const int64_t* offsets, uint64_t n_reads,
                    uint64_t n_bases) {
    if (K < 1 || K >= kMaxK) return fail(GT_ERR_INVALID_ARG, "K must be in 1..255");
    if (mode != kRaw && mode != kSkipN) return fail(GT_ERR_INVALID_ARG, "unknown input mode");
    if (n_reads == 0) return GT_OK;
    if (!offsets || (!seq && n_bases)) return fail(GT_ERR_INVALID_ARG, "null reads");
    if (n_reads >= 0xFFFFFFFFULL) return fail(GT_ERR_INVALID_ARG, "too many reads in one call (max 2^32-2)");
    if (n_bases >= (1ULL << 62)) return fail(GT_ERR_INVALID_ARG, "n_bases out of range");
    return GT_OK;
}

Reads make_reads(const uint8_t* seq, const int64_t* offsets, uint64_t n_reads, uint64_t n_bases, uint16_t K,
                 int mode, int canonical) {
    Reads R{};
    R.seq = seq;
    R.off = offsets;
    R.n_reads = n_reads;
    R.n_bases = (int64_t)n_bases;
    R.b0 = 0;
    R.b1 = (int64_t)n_bases;
    R.K = K;
    R.mode = mode;
    R.canonical = canonical ? 1 : 0;
    return R;
}

// Optional cap on the sub-batch size in start positions (environment variable
// GOETIA_AMD_SUB_BATCH), used to exercise the multi-sub-batch path on small
// inputs; 0 = no cap.
uint64_t sub_batch_cap() {
    const char* v = std::getenv("GOETIA_AMD_SUB_BATCH");
    if (!v || !*v) return 0;
    const unsigned long long x = std::strtoull(v, nullptr, 10);
    return x ? std::max<uint64_t>((x + kTileMax - 1) / kTileMax * kTileMax, kTileMax) : 0;
}

// Allocates a claim map of `want` slots (zeroed) with its generation counter.
int alloc_claims(Scratch& sc, uint64_t want, hipStream_t s) {
    GT_TRY(sc.claim_keys.ensure(want * sizeof(uint64_t)));
    GT_TRY(sc.claim_vals.ensure(want * sizeof(uint64_t)));
    GT_TRY(sc.claim_gen.ensure(sizeof(uint64_t)));
    launch_fill_u64(sc.claim_keys.as<uint64_t>(), want, 0, s);
    launch_fill_u64(sc.claim_vals.as<uint64_t>(), want, 0, s);
    launch_fill_u64(sc.claim_gen.as<uint64_t>(), 1, 0, s);
    GT_HIP(hipGetLastError());
    sc.claim_cap = want;
    sc.claim_gen_issued = 0;
    return GT_OK;
}

// Largest sub-batch (in start positions, a multiple of kTileMax) whose claims fit
// the claim map at load factor <= 1/2; grows the claim map and fresh buffer
// as needed for this call (fixed scratch only checks).
int prepare_claims(Scratch& sc, int kind, uint32_t n_tables, uint64_t n_starts, uint64_t* sub, hipStream_t s) {
    const uint64_t kMaxClaimCap = 1ULL << 29;      // 8 GiB of claim slots
    const uint64_t kMinClaimCap = 1ULL << 16;
    const uint64_t per = std::max<uint32_t>(n_tables, 1);
    const uint64_t cap = sub_batch_cap();
    const uint64_t starts = (n_starts + kTileMax - 1) / kTileMax * kTileMax;
    if (kind == kExact) {
        uint64_t sb = std::min<uint64_t>(starts, 1ULL << 31);
        if (cap) sb = std::min(sb, cap);
        if (sc.fixed) {
            if (sc.fresh.bytes < sb) sb = sc.fresh.bytes / kTileMax * kTileMax;
        } else {
            GT_TRY(sc.fresh.ensure(sb));
        }
        if (sb == 0) return fail(GT_ERR_INTERNAL, "scratch too small");
        if (!sc.claim_cap) GT_TRY(alloc_claims(sc, 2, s));   // unused by the exact set; a valid view
        *sub = sb;
        return GT_OK;
    }
    uint64_t want = std::min(kMaxClaimCap, std::max(kMinClaimCap, next_pow2(2 * per * (n_starts + kTileMax))));
    if (want > sc.claim_cap && !sc.fixed) GT_TRY(alloc_claims(sc, want, s));
    uint64_t sb = sc.claim_cap / (2 * per);
    sb = std::min<uint64_t>(sb, 1ULL << 31) / kTileMax * kTileMax;
    if (cap) sb = std::min(sb, cap);
    if (sc.fixed) {
        sb = std::min<uint64_t>(sb, sc.fresh.bytes / kTileMax * kTileMax);
    } else {
        GT_TRY(sc.fresh.ensure(std::min<uint64_t>(sb, starts)));
    }
    if (sb == 0) return fail(GT_ERR_INTERNAL, "claim map too small");
    *sub = sb;
    return GT_OK;
}

int prepare_claims(gt_storage* st, uint64_t n_starts, uint64_t* sub, hipStream_t s) {
    return prepare_claims(st->own, st->kind, st->n_tables, n_starts, sub, s);
}

// Enqueues the next first-touch generation of the storage (exact-set stamps,
// 32 bits) and of the claim map.  Claim-map keys keep 28 bits of the map's
// generation, so the map is cleared whenever that field would wrap, and
// values that are 0 in those bits are skipped (cleared keys read as gen 0).
int next_gen(gt_storage* st, Scratch& sc, hipStream_t s) {
    if (st->gen_issued + 2 >= 0xFFFFFFFFULL || sc.claim_gen_issued + 2 >= 0xFFFFFFFFULL)
        return fail(GT_ERR_INTERNAL, "batch generation counter exhausted");
    st->gen_issued++;
    sc.claim_gen_issued++;
    launch_bump_gen(st->counters.as<uint64_t>() + kGen, sc.claim_gen.as<uint64_t>(), s);
    if ((sc.claim_gen_issued & kGenKeyMask) == 0) {
        launch_fill_u64(sc.claim_keys.as<uint64_t>(), sc.claim_cap, 0, s);
        launch_fill_u64(sc.claim_vals.as<uint64_t>(), sc.claim_cap, 0, s);
        sc.claim_gen_issued++;
        launch_bump_gen(nullptr, sc.claim_gen.as<uint64_t>(), s);
    }
    GT_HIP(hipGetLastError());
    return GT_OK;
}

ClaimMap claim_view(const Scratch& sc) {
    ClaimMap cm{};
    cm.keys = sc.claim_keys.as<uint64_t>();
    cm.vals = sc.claim_vals.as<uint64_t>();
    cm.gen = sc.claim_gen.as<uint64_t>();
    cm.cap = sc.claim_cap;
    cm.log2cap = log2u(sc.claim_cap);
    return cm;
}

// Grows the exact set so that `more` new keys keep its load factor <= 1/2.
int exact_reserve(gt_storage* st, uint64_t more, hipStream_t s) {
    if (2 * (st->ecount_ub + more) <= st->ecap) return GT_OK;
    uint64_t c[kNCounters];
    GT_TRY(read_counters(st, c, s));
    st->ecount_ub = c[kNUnique];
    if (2 * (st->ecount_ub + more) <= st->ecap) return GT_OK;
    return alloc_exact(st, next_pow2(2 * (st->ecount_ub + more)), s);
}

// Tile shape per storage kind (index 0: hashing without a storage):
// windows per thread and persistent workgroups per CU.
struct Tune {
    int per = 16;
    int blocks_per_cu = 8;
};

std::mutex g_tune_mu;
Tune g_tune[4];

Launch make_launch(int kind, hipStream_t s) {
    Tune t;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        t = g_tune[(kind >= 0 && kind < 4) ? kind : 0];
    }
    return Launch{std::max(1, g_dev.cus * t.blocks_per_cu), s, t.per};
}

Launch make_launch(hipStream_t s) { return make_launch(0, s); }
Launch make_launch(const gt_storage* st, hipStream_t s) { return make_launch(st->kind, s); }

// Device work of one first-touch sub-batch [Rs.b0, Rs.b1) with the given
// claim map and fresh buffer.  For the exact set, `growth` allows a
// synchronising capacity check (never while a graph is being captured).
int enqueue_subbatch(gt_storage* st, Scratch& sc, const Launch& L, const Reads& Rs, uint64_t* n_new,
                     uint8_t* new_flags, hipStream_t s, bool growth) {
    uint64_t* ctr = st->counters.as<uint64_t>();
    uint8_t* fresh = sc.fresh.as<uint8_t>();
    GT_TRY(next_gen(st, sc, s));
    const ClaimMap cm = claim_view(sc);
    if (st->kind == kExact) {
        const uint64_t nstarts = (uint64_t)(Rs.b1 - Rs.b0);
        if (growth && 2 * (st->ecount_ub + nstarts) > st->ecap) {
            // Bound the growth by this sub-batch's fresh windows.
            launch_fill_u64(ctr + kFresh, 1, 0, s);
            {
                ProfScope ps("exact_lookup", s);
                launch_exact_lookup(L, Rs, st->exact_view(), fresh, ctr);
            }
            GT_HIP(hipGetLastError());
            uint64_t c[kNCounters];
            GT_TRY(read_counters(st, c, s));
            st->ecount_ub = c[kNUnique];
            GT_TRY(exact_reserve(st, c[kFresh], s));
            st->ecount_ub += c[kFresh];
        } else {
            ProfScope ps("exact_lookup", s);
            launch_exact_lookup(L, Rs, st->exact_view(), fresh, ctr);
            st->ecount_ub += nstarts;
        }
        const ExactSet E = st->exact_view();
        {
            ProfScope ps("exact_insert", s);
            launch_exact_insert(L, Rs, E, fresh, ctr);
        }
        {
            ProfScope ps("exact_resolve", s);
            launch_exact_resolve(L, Rs, E, fresh, n_new, new_flags, ctr);
        }
        launch_exact_tail(ctr, s);
    } else {
        const Tables T = st->tables_view();
        {
            ProfScope ps("claim", s);
            launch_claim(L, Rs, T, cm, fresh, ctr);
        }
        {
            ProfScope ps("apply", s);
            launch_apply(L, Rs, T, cm, fresh, n_new, new_flags, ctr);
        }
    }
    GT_HIP(hipGetLastError());
    return GT_OK;
}

// ------------------------------------------------------------ binned insert

// Buckets of the binned insert over a storage's tables (0: not eligible).
uint32_t bin_buckets(const gt_storage* st) {
    if ((st->kind != kByte && st->kind != kBit) || st->n_tables == 0 || st->n_tables > kBinMaxTables) return 0;
    uint64_t nb = 0;
    for (uint32_t i = 0; i < st->n_tables; ++i) nb += (st->sizes[i] + kBucketBins - 1) / kBucketBins;
    return nb <= kMaxBuckets ? (uint32_t)nb : 0;
}

// Whether an insert goes through the binned path.  GOETIA_AMD_BINNED=0 / 1
// forces it off / on (where the tables allow it); by default it is used for
// batches of at least 2^20 bases with at least one base per 256 table bins,
// where replaying buckets in LDS beats random atomics on the tables (on
// 4 x 2^28 count-min bins, 4.5M-base batches insert 3.7e9 k-mers/s binned
// against 2.2e9 with atomics, 1.2M-base batches 1.4e9 against 1.8e9).
bool use_binned(const gt_storage* st, const Scratch& sc, const Reads& R) {
    if (R.dyn || sc.fixed || !bin_buckets(st)) return false;
    const char* v = std::getenv("GOETIA_AMD_BINNED");
    if (v && *v) return v[0] != '0';
    uint64_t bins = 0;
    for (uint64_t p : st->sizes) bins += p;
    return (uint64_t)R.n_bases >= (1ULL << 20) && (uint64_t)R.n_bases * 256 >= bins;
}

// Binned sub-batch in start positions, a multiple of kTileMax: 2^27 for an
// insert, 2^26 for insert_and_query (its four result bytes per window are
// written scattered, and 2^26 windows keep them in a 256 MiB range), or
// GOETIA_AMD_BIN_SUB (at most 2^29); capped by GOETIA_AMD_SUB_BATCH.
uint64_t bin_sub_batch(uint64_t n_bases, bool iq) {
    uint64_t sb = iq ? 1ULL << 26 : 1ULL << 27;
    if (const char* v = std::getenv("GOETIA_AMD_BIN_SUB")) {
        const unsigned long long x = std::strtoull(v, nullptr, 10);
        if (x) sb = std::min<uint64_t>(x, 1ULL << 29);
    }
    if (const uint64_t cap = sub_batch_cap()) sb = std::min(sb, cap);
    sb = std::max<uint64_t>((sb + kTileMax - 1) / kTileMax * kTileMax, kTileMax);
    return std::min<uint64_t>(sb, (n_bases + kTileMax - 1) / kTileMax * kTileMax);
}

// Plan of the binned passes over this storage's tables (buckets, digits) and
// scratch for sub-batches of up to tiles_max first-pass tiles.
int bin_prepare(gt_storage* st, Scratch& sc, uint64_t tiles_max, BinPlan& P) {
    const uint32_t nt = st->n_tables;
    P = BinPlan{};
    P.n_tables = nt;
    uint32_t nb = 0;
    for (uint32_t i = 0; i < nt; ++i) {
        P.bfirst[i] = nb;
        P.gbase[i] = nb << kBinBits;
        nb += (uint32_t)((st->sizes[i] + kBucketBins - 1) / kBucketBins);
    }
    P.bfirst[nt] = nb;
    P.n_buckets = nb;
    uint32_t bits = 0;
    while ((1u << bits) < nb) ++bits;
    P.shift2 = bits > (uint32_t)kDigitBits1 ? bits - (uint32_t)kDigitBits1 : 0u;
    P.n_d1 = (nb + (1u << P.shift2) - 1) >> P.shift2;

    const uint64_t groups_max = (tiles_max + kBinGroup - 1) / kBinGroup;
    const uint64_t n_ent = tiles_max * nt * kBinTile;
    const uint64_t cnt_max = ((uint64_t)P.n_d1 << P.shift2) * groups_max;
    int rc = sc.bin_ent1.ensure(n_ent * sizeof(uint64_t));
    if (rc == GT_OK) rc = sc.bin_ent2.ensure(n_ent * sizeof(uint64_t));
    if (rc == GT_OK) rc = sc.bin_dig2.ensure(n_ent * sizeof(uint16_t));
    if (rc == GT_OK) rc = sc.bin_off1.ensure((kDigits1 + 1) * tiles_max * sizeof(uint16_t));
    if (rc == GT_OK) rc = sc.bin_cnt2.ensure((cnt_max + 1) * sizeof(uint32_t));
    if (rc == GT_OK) rc = sc.bin_sums.ensure((scan_u32_block_count(cnt_max) + 1) * sizeof(uint32_t));
    if (rc == GT_OK) rc = sc.bin_kbase.ensure(sizeof(uint64_t));
    if (rc != GT_OK) return rc;
    P.ent1 = sc.bin_ent1.as<uint64_t>();
    P.ent2 = sc.bin_ent2.as<uint64_t>();
    P.dig2 = sc.bin_dig2.as<uint16_t>();
    P.off1 = sc.bin_off1.as<uint16_t>();
    P.cnt2 = sc.bin_cnt2.as<uint32_t>();
    P.kbase = sc.bin_kbase.as<uint64_t>();
    return GT_OK;
}

// Routes the (window, table) events of the windows starting in [b0, b1) of
// prepared reads (at most the prepared tiles_max tiles): by bucket, in
// window order within a bucket (first pass, second-pass count and scan,
// second-pass scatter).
void bin_route(const Reads& R, const Tables& T, BinPlan& P, Scratch& sc, int64_t b0, int64_t b1, hipStream_t s) {
    Reads Rs = R;
    Rs.b0 = b0;
    Rs.b1 = b1;
    P.tile0 = b0 / kBinTile;
    P.tiles1 = (uint32_t)((b1 + kBinTile - 1) / kBinTile - P.tile0);
    P.n_groups = (P.tiles1 + kBinGroup - 1) / kBinGroup;
    const uint64_t cnt_len = ((uint64_t)P.n_d1 << P.shift2) * P.n_groups;
    const int grid_l1 = std::max(1, g_dev.cus * 2);     // 70 KB of LDS per workgroup
    const int grid = std::max(1, g_dev.cus * 8);
    launch_bin_kbase(R, b0, sc.bin_kbase.as<uint64_t>(), s);
    {
        ProfScope ps("bin_hash", s);
        launch_bin_l1(Rs, T, P, grid_l1, s);
    }
    {
        ProfScope ps("bin_count", s);
        launch_bin_l2_count(P, grid, s);
        launch_scan_u32(P.cnt2, cnt_len, sc.bin_sums.as<uint32_t>(), s);
    }
    {
        // two workgroups per CU: fewer partly written output lines in each L2
        // at a time (67.5 ms per full-size step against 72.8 with three)
        ProfScope ps("bin_route", s);
        launch_bin_l2_scatter(P, std::max(1, g_dev.cus * 2), s);
    }
}

// Binned insert of prepared reads (tile_read and kmer_off set): per
// sub-batch, route the (window, table) events by bin and replay each bucket
// in LDS; then per-read new counts from the window flags, or, with iq_counts
// (count-min only), the insert_and_query count of every window.
int binned_insert_impl(gt_storage* st, Scratch& sc, const Reads& R, uint64_t* n_new, uint8_t* new_flags,
                       int16_t* iq_counts, hipStream_t s) {
    const uint32_t nt = st->n_tables;
    const bool iq = iq_counts != nullptr;
    const uint64_t sub = bin_sub_batch((uint64_t)R.n_bases, iq);
    BinPlan P;
    GT_TRY(bin_prepare(st, sc, sub / kBinTile, P));
    uint8_t* flags = new_flags;
    if (iq) {
        GT_TRY(sc.bin_vq.ensure((size_t)R.n_bases * kBinMaxTables));
        launch_zero_words(sc.bin_vq.as<uint32_t>(), R.kmer_off + R.n_reads, (uint64_t)R.n_bases, s);
        flags = nullptr;
    } else {
        if (!flags) {
            GT_TRY(sc.bin_flags.ensure((size_t)R.n_bases));
            flags = sc.bin_flags.as<uint8_t>();
        }
        launch_zero_bytes(flags, R.kmer_off + R.n_reads, (uint64_t)R.n_bases, s);
    }
    P.flags = flags;
    P.vq = iq ? sc.bin_vq.as<uint8_t>() : nullptr;
    P.unordered = iq ? 0u : 1u;        // a plain insert's replay does not depend on event order

    const Tables T = st->tables_view();
    uint64_t* ctr = st->counters.as<uint64_t>();
    const int grid = std::max(1, g_dev.cus * 8);
    for (int64_t b0 = 0; b0 < R.n_bases; b0 += (int64_t)sub) {
        bin_route(R, T, P, sc, b0, std::min<int64_t>(R.n_bases, b0 + (int64_t)sub), s);
        {
            // the order-free replay launches as many workgroups as fit at once (two
            // per CU, by LDS), each walking its buckets with the next one prefetched
            // (35.9 ms per full step against 36.6 with eight launched per CU)
            ProfScope ps("bin_apply", s);
            launch_bin_apply(P, T, ctr, P.unordered && !P.vq ? std::max(1, g_dev.cus * 2) : grid, s);
        }
        GT_HIP(hipGetLastError());
    }
    if (iq) {
        ProfScope ps("bin_iq", s);
        launch_bin_iq_combine(sc.bin_vq.as<uint32_t>(), R.kmer_off + R.n_reads, (uint64_t)R.n_bases, nt, iq_counts,
                              ctr, s);
    } else {
        ProfScope ps("bin_new", s);
        launch_bin_read_new(R.kmer_off, R.n_reads, flags, n_new, ctr, s);
    }
    GT_HIP(hipGetLastError());
    return GT_OK;
}

// Count-min insert (or insert_and_query, with iq_counts) with bigcount on:
// the sort-based exact insert gives every window its tables' counts just
// before it, so the windows that found every table full are known in window
// order; insert_and_query takes its bigcount values from them before the map
// takes the batch (bytestorage.cc:60-150).  Synchronises the stream.
// The bigcount steps of insert_and_query over n windows whose table counts
// and full flags are known: answered in window chunks, in order, a chunk's
// full windows joining the map before the next chunk is answered (goetia's
// sequential order).  The selection and sort inside take int counts, so a
// chunk stays below 2^31 windows (GOETIA_AMD_SUB_BATCH caps it further).
int bigcount_iq_chunks(gt_storage* st, const uint64_t* fw, const uint64_t* rc, const uint8_t* valid, int canonical,
                       const uint8_t* full, int16_t* counts, uint64_t n, hipStream_t s) {
    uint64_t chunk = 1ULL << 28;
    if (const uint64_t cap = sub_batch_cap()) chunk = std::min(chunk, cap);
    for (uint64_t w0 = 0; w0 < n; w0 += chunk) {
        const uint64_t m = std::min(chunk, n - w0);
        GT_TRY(big_insert_query(st->big, fw + w0, rc ? rc + w0 : nullptr, valid ? valid + w0 : nullptr, canonical,
                                full + w0, counts + w0, m, g_dev.cus, s));
        GT_TRY(big_insert(st->big, fw + w0, rc ? rc + w0 : nullptr, canonical, full + w0, m, g_dev.cus, s));
    }
    return GT_OK;
}

int bigcount_insert_impl(gt_storage* st, Reads R, uint64_t* n_new, uint8_t* new_flags, int16_t* iq_counts,
                         hipStream_t s) {
    if (n_new && R.n_reads) launch_fill_u64(n_new, R.n_reads, 0, s);
    if (R.n_reads == 0 || R.n_bases == 0) return GT_OK;
    GT_TRY(prepare_reads(st, R, true, s));
    int64_t nk = 0;
    GT_HIP(hipMemcpyAsync(&nk, st->own.kmer_off.as<int64_t>() + R.n_reads, sizeof(nk), hipMemcpyDeviceToHost, s));
    GT_HIP(hipStreamSynchronize(s));
    if (nk <= 0) return GT_OK;
    DevBuf fw, rc, valid, fresh, full, junk;
    GT_TRY(fw.ensure((size_t)nk * 8));
    GT_TRY(rc.ensure((size_t)nk * 8));
    GT_TRY(valid.ensure((size_t)nk));
    GT_TRY(full.ensure((size_t)nk));
    uint8_t* flags = new_flags;
    if (!flags && n_new) {
        GT_TRY(fresh.ensure((size_t)nk));
        flags = fresh.as<uint8_t>();
    }
    launch_hash(make_launch(st, s), R, fw.as<uint64_t>(), rc.as<uint64_t>(), valid.as<uint8_t>());
    GT_HIP(hipGetLastError());
    GT_TRY(sorted_insert_query(st->tables_view(), fw.as<uint64_t>(), rc.as<uint64_t>(), valid.as<uint8_t>(),
                               (uint64_t)nk, R.canonical, iq_counts, flags, st->counters.as<uint64_t>(), g_dev.cus, s,
                               full.as<uint8_t>()));
    if (iq_counts)
        GT_TRY(bigcount_iq_chunks(st, fw.as<uint64_t>(), rc.as<uint64_t>(), valid.as<uint8_t>(), R.canonical,
                                  full.as<uint8_t>(), iq_counts, (uint64_t)nk, s));
    else
        GT_TRY(big_insert(st->big, fw.as<uint64_t>(), rc.as<uint64_t>(), R.canonical, full.as<uint8_t>(),
                          (uint64_t)nk, g_dev.cus, s));
    if (n_new) {
        // per-read sums of the flags (n_unique was counted by the sort-based pass)
        GT_TRY(junk.ensure(kNCounters * sizeof(uint64_t)));
        launch_bin_read_new(R.kmer_off, R.n_reads, flags, n_new, junk.as<uint64_t>(), s);
        GT_HIP(hipGetLastError());
    }
    GT_HIP(hipStreamSynchronize(s));
    return GT_OK;
}

// The entry points that cannot keep the bigcount map refuse such storages.
int refuse_bigcount(const gt_storage* st, const char* what) {
    if (st->use_bigcount)
        return fail(GT_ERR_UNSUPPORTED, std::string(what) + " is not available on a storage with bigcount on");
    return GT_OK;
}

int dbg_insert_impl(gt_storage* st, Scratch& sc, Reads R, uint64_t* n_new, uint8_t* new_flags, hipStream_t s) {
    if (st->use_bigcount) {
        if (&sc != &st->own) return refuse_bigcount(st, "an insert with outside scratch");
        return bigcount_insert_impl(st, R, n_new, new_flags, nullptr, s);
    }
    if (R.n_reads == 0 || R.n_bases == 0) {
        if (n_new && R.n_reads) launch_fill_u64(n_new, R.n_reads, 0, s);
        return GT_OK;
    }
    const bool binned = use_binned(st, sc, R);
    GT_TRY(prepare_reads(sc, R, new_flags != nullptr || binned, s));
    if (n_new) launch_fill_u64(n_new, R.n_reads, 0, s);
    if (binned) {
        const int rc = binned_insert_impl(st, sc, R, n_new, new_flags, nullptr, s);
        if (rc != GT_ERR_OUT_OF_MEMORY) return rc;
        // not enough device memory for the event buffers: release them and
        // insert with atomics on the tables instead (same results)
        sc.bin_ent1.release();
        sc.bin_ent2.release();
        sc.bin_dig2.release();
        sc.bin_flags.release();
        sc.bin_vq.release();
    }
    uint64_t sub = 0;
    GT_TRY(prepare_claims(sc, st->kind, st->n_tables, (uint64_t)R.n_bases, &sub, s));
    const Launch L = make_launch(st, s);
    for (int64_t b0 = 0; b0 < R.n_bases; b0 += (int64_t)sub) {
        Reads Rs = R;
        Rs.b0 = b0;
        Rs.b1 = std::min<int64_t>(R.n_bases, b0 + (int64_t)sub);
        GT_TRY(enqueue_subbatch(st, sc, L, Rs, n_new, new_flags, s, true));
    }
    return GT_OK;
}

int dbg_insert_impl(gt_storage* st, Reads R, uint64_t* n_new, uint8_t* new_flags, hipStream_t s) {
    return dbg_insert_impl(st, st->own, R, n_new, new_flags, s);
}

int hashes_insert_impl(gt_storage* st, const uint64_t* h, uint64_t n, uint8_t* is_new, hipStream_t s) {
    if (n == 0) return GT_OK;
    if (st->use_bigcount) {
        DevBuf full;
        GT_TRY(full.ensure(n));
        GT_TRY(sorted_insert_query(st->tables_view(), h, nullptr, nullptr, n, 0, nullptr, is_new,
                                   st->counters.as<uint64_t>(), g_dev.cus, s, full.as<uint8_t>()));
        return big_insert(st->big, h, nullptr, 0, full.as<uint8_t>(), n, g_dev.cus, s);
    }
    const Launch L = make_launch(st, s);
    uint64_t* ctr = st->counters.as<uint64_t>();
    uint64_t sub = 0;
    GT_TRY(prepare_claims(st, n, &sub, s));
    for (uint64_t i0 = 0; i0 < n; i0 += sub) {
        const uint64_t m = std::min<uint64_t>(sub, n - i0);
        GT_TRY(next_gen(st, st->own, s));
        uint8_t* fresh = st->own.fresh.as<uint8_t>();
        if (st->kind == kExact) {
            GT_TRY(exact_reserve(st, m, s));
            st->ecount_ub += m;
            const ExactSet E = st->exact_view();
            launch_hash_exact_lookup(L, h + i0, m, E, fresh, ctr);
            launch_hash_exact_insert(L, h + i0, m, E, fresh, ctr);
            launch_hash_exact_resolve(L, h + i0, m, E, fresh, is_new ? is_new + i0 : nullptr, ctr);
            launch_exact_tail(ctr, s);
        } else {
            const Tables T = st->tables_view();
            const ClaimMap cm = claim_view(st->own);
            launch_hash_claim(L, h + i0, m, T, cm, fresh, ctr);
            launch_hash_apply(L, h + i0, m, T, cm, fresh, is_new ? is_new + i0 : nullptr, ctr);
        }
        GT_HIP(hipGetLastError());
    }
    return GT_OK;
}

// Host staging: copies host arrays into the storage's staging buffers.
template <class T>
int stage_in(DevBuf& buf, const T* host, size_t n, hipStream_t s) {
    GT_TRY(buf.ensure(std::max<size_t>(n * sizeof(T), 16)));
    if (n) GT_HIP(hipMemcpyAsync(buf.p, host, n * sizeof(T), hipMemcpyHostToDevice, s));
    return GT_OK;
}

template <class T>
int stage_out(T* host, const DevBuf& buf, size_t n, hipStream_t s) {
    if (host && n) GT_HIP(hipMemcpyAsync(host, buf.p, n * sizeof(T), hipMemcpyDeviceToHost, s));
    return GT_OK;
}

uint64_t packed_kmers(const int64_t* offsets, uint64_t n_reads, int K) {
    uint64_t n = 0;
    for (uint64_t r = 0; r < n_reads; ++r) {
        const int64_t len = offsets[r + 1] - offsets[r];
        if (len >= K) n += (uint64_t)(len - K + 1);
    }
    return n;
}

int check_host_offsets(const int64_t* offsets, uint64_t n_reads) {
    if (n_reads == 0) return GT_OK;
    if (!offsets) return fail(GT_ERR_INVALID_ARG, "null offsets");
    if (offsets[0] != 0) return fail(GT_ERR_INVALID_ARG, "offsets[0] must be 0");
    for (uint64_t r = 0; r < n_reads; ++r) {
        if (offsets[r + 1] < offsets[r]) return fail(GT_ERR_INVALID_ARG, "offsets must be non-decreasing");
    }
    return GT_OK;
}

}  // namespace

// ============================================================ C ABI

namespace gt {
// Sets this thread's gt_last_error() message from the library's other
// translation units; returns code.
int set_error(int code, const std::string& msg) { return fail(code, msg); }

// Marks the end of work queued on stream s for this storage by a caller that
// does not wait for it (the file inserter); the storage's later device calls
// wait for it (order_after_fence).
int storage_set_fence(gt_storage* st, hipStream_t s) {
    StorageLock lk(st);
    GT_TRY(lk.rc);
    if (!st->fence) GT_HIP(hipEventCreateWithFlags(&st->fence, hipEventDisableTiming));
    GT_HIP(hipEventRecord(st->fence, s));
    st->fence_set = true;
    return GT_OK;
}

// Makes fn(ctx) the storage's holder of reads taken but not yet queued (a
// file inserter); every later call on the storage runs it first
// (StorageLock).  Another holder's reads are queued first.  fn = nullptr
// ends ctx's hold.
int storage_hold(gt_storage* st, int (*fn)(void*), void* ctx) {
    std::lock_guard<std::recursive_mutex> g(st->mu);
    if (!fn) {
        if (st->flush_ctx == ctx) {
            st->flush_fn = nullptr;
            st->flush_ctx = nullptr;
        }
        return GT_OK;
    }
    if (st->flush_fn && st->flush_ctx != ctx) GT_TRY(flush_pending(st));
    st->flush_fn = fn;
    st->flush_ctx = ctx;
    return GT_OK;
}

// The storage's lock, for a holder that changes what it holds.
std::recursive_mutex& storage_mutex(gt_storage* st) { return st->mu; }

// Replaces a Bloom / count-min storage's tables with zeroed tables of the
// given sizes (the file loaders: goetia's load() takes the file's table
// sizes, bitstorage.cc:239-263, bytestorage.cc:196-240).  Counters are kept;
// captured pipelines see a new epoch and capture again.
int storage_reshape(gt_storage* st, const uint64_t* sizes, uint32_t n_tables) {
    if (!valid_storage(st) || st->kind == kExact || !sizes || n_tables == 0 || n_tables > (uint32_t)kMaxTables)
        return fail(GT_ERR_INVALID_ARG, "invalid arguments");
    for (uint32_t i = 0; i < n_tables; ++i)
        if (sizes[i] == 0 || sizes[i] >= kMaxBins) return fail(GT_ERR_INVALID_ARG, "table sizes must be in 1..2^32-1");
    GT_TRY(ensure_device());
    StorageLock lk(st);
    GT_TRY(lk.rc);
    if (st->n_tables == n_tables && std::equal(sizes, sizes + n_tables, st->sizes.begin())) return GT_OK;
    GT_HIP(hipDeviceSynchronize());
    std::vector<uint64_t> sz(sizes, sizes + n_tables), nb, off;
    uint64_t total = 0;
    for (uint32_t i = 0; i < n_tables; ++i) {
        nb.push_back(st->kind == kBit ? sz[i] / 8 + 1 : sz[i]);
        off.push_back(total);
        total += (nb[i] + 255) / 256 * 256;
    }
    DevBuf t;
    GT_TRY(t.ensure(total));
    std::swap(st->tables.p, t.p);
    std::swap(st->tables.bytes, t.bytes);
    st->n_tables = n_tables;
    st->sizes = sz;
    st->nbytes = nb;
    st->table_off = off;
    GT_TRY(zero_tables(st, nullptr));
    GT_HIP(hipStreamSynchronize(nullptr));
    st->epoch++;
    return GT_OK;
}
}  // namespace gt

extern "C" {

const char* gt_last_error(void) { return g_err.c_str(); }


int gt_abi_version(void) { return GT_ABI_VERSION; }

int gt_device_init(int device) { return ensure_device(device); }

int gt_device_synchronize(void) {
    GT_TRY(ensure_device());
    GT_HIP(hipDeviceSynchronize());
    return GT_OK;
}

// storage.hh:146-190 (is_prime by trial division, odd candidates below x).
int gt_primes_near_x(uint32_t n, uint64_t x, uint64_t* out, uint32_t* n_out) {
    if (!out || !n_out) return fail(GT_ERR_INVALID_ARG, "null output");
    auto is_prime = [](uint64_t v) {
        if (v < 2) return false;
        if (v == 2) return true;
        if (v % 2 == 0) return false;
        for (unsigned long long i = 3; i < std::sqrt((double)v) + 1; i += 2) {
            if (v % i == 0) return false;
        }
        return true;
    };
    uint32_t found = 0;
    if (x == 0) {
        *n_out = 0;
        return GT_OK;
    }
    if (x == 1) {
        if (n) out[found++] = 1;
        *n_out = found;
        return GT_OK;
    }
    uint64_t i = x - 1;
    if (i % 2 == 0) i--;
    while (found != n) {
        if (is_prime(i)) out[found++] = i;
        if (i == 1) break;
        i -= 2;
    }
    *n_out = found;
    return GT_OK;
}

int gt_storage_new(int kind, const uint64_t* tablesizes, uint32_t n_tables, uint64_t capacity_hint,
                   gt_storage** out) {
    if (!out) return fail(GT_ERR_INVALID_ARG, "null output handle");
    *out = nullptr;
    if (kind < kBit || kind > kExact) return fail(GT_ERR_INVALID_ARG, "unknown storage kind");
    if (kind != kExact) {
        if (n_tables == 0 || n_tables > (uint32_t)kMaxTables)
            return fail(GT_ERR_INVALID_ARG, "n_tables must be in 1..16");
        if (!tablesizes) return fail(GT_ERR_INVALID_ARG, "null tablesizes");
        for (uint32_t i = 0; i < n_tables; ++i) {
            if (tablesizes[i] == 0 || tablesizes[i] >= kMaxBins)
                return fail(GT_ERR_INVALID_ARG, "table sizes must be in 1..2^32-1");
        }
    }
    GT_TRY(ensure_device());
    gt_storage* st = new (std::nothrow) gt_storage();
    if (!st) return fail(GT_ERR_OUT_OF_MEMORY, "host allocation failed");
    st->kind = kind;
    hipStream_t s = nullptr;
    int rc = st->counters.ensure(kNCounters * sizeof(uint64_t));
    if (rc == GT_OK) launch_fill_u64(st->counters.as<uint64_t>(), kNCounters, 0, s);
    if (rc == GT_OK && kind != kExact) {
        st->n_tables = n_tables;
        uint64_t total = 0;
        for (uint32_t i = 0; i < n_tables; ++i) {
            st->sizes.push_back(tablesizes[i]);
            const uint64_t nb = kind == kBit ? tablesizes[i] / 8 + 1 : tablesizes[i];
            st->nbytes.push_back(nb);
            st->table_off.push_back(total);
            total += (nb + 255) / 256 * 256;
        }
        rc = st->tables.ensure(total);
        if (rc == GT_OK) rc = zero_tables(st, s);
    } else if (rc == GT_OK) {
        rc = st->eones.ensure(256);
        if (rc == GT_OK) launch_fill_u64(st->eones.as<uint64_t>(), 1, 0, s);
        if (rc == GT_OK) rc = alloc_exact(st, next_pow2(std::max<uint64_t>(1024, 2 * capacity_hint)), s);
    }
    if (rc == GT_OK) {
        hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = fail(GT_ERR_DEVICE, std::string("storage init: ") + hipGetErrorString(e));
    }
    if (rc != GT_OK) {
        delete st;
        return rc;
    }
    *out = st;
    return GT_OK;
}

int gt_storage_free(gt_storage* st) {
    if (!st) return GT_OK;
    if (!valid_storage(st)) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    {
        StorageLock lk(st);            // a file inserter's held reads go in before the storage goes
    }
    (void)hipDeviceSynchronize();
    for (auto* p : st->pipelines) p->st = nullptr;   // their later calls fail cleanly
    if (st->fence) (void)hipEventDestroy(st->fence);
    big_free(st->big);
    delete st;
    return GT_OK;
}

int gt_storage_clone(const gt_storage* st, gt_storage** out) {
    if (!valid_storage(st)) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    return gt_storage_new(st->kind, st->sizes.data(), st->n_tables, st->kind == kExact ? st->ecap / 2 : 0, out);
}

int gt_storage_reset(gt_storage* st) {
    if (!valid_storage(st)) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    GT_TRY(ensure_device());
    StorageLock lk(st);
    GT_TRY(lk.rc);
    GT_TRY(zero_tables(st, nullptr));
    if (st->kind == kExact) {
        launch_fill_u64(st->counters.as<uint64_t>() + kNUnique, 2, 0, nullptr);   // n_unique, last-was-present
        st->ecount_ub = 0;
    }
    GT_HIP(hipStreamSynchronize(nullptr));
    return GT_OK;
}

int gt_storage_kind(const gt_storage* st, int* kind) {
    if (!valid_storage(st) || !kind) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    *kind = st->kind;
    return GT_OK;
}

int gt_storage_n_tables(const gt_storage* st, uint32_t* n) {
    if (!valid_storage(st) || !n) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    *n = st->n_tables;
    return GT_OK;
}

int gt_storage_tablesizes(const gt_storage* st, uint64_t* out) {
    if (!valid_storage(st) || (!out && st->n_tables)) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    for (uint32_t i = 0; i < st->n_tables; ++i) out[i] = st->sizes[i];
    return GT_OK;
}

int gt_storage_table_bytes(const gt_storage* st, uint32_t i, uint64_t* nbytes) {
    if (!valid_storage(st) || !nbytes) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    if (i >= st->n_tables) return fail(GT_ERR_INVALID_ARG, "table index out of range");
    *nbytes = st->nbytes[i];
    return GT_OK;
}

// The exact set's n_occupied is its container's bucket count
// (sparseppstorage.hh:80-86).  sparsepp starts at 32 buckets, grows before an
// insert whose size + 1 would pass half the buckets, by doubling
// (spp.h:2691-2747, HT_OCCUPANCY_PCT = 50), and clear() of a non-empty set
// returns to 32 (spp.h:2980-2989).  Every insert call checks, also for a value
// already present, so after n distinct values the count covers n + 1 when the
// last value inserted was already there.
static uint64_t set_bucket_count(uint64_t n_unique, bool last_was_present) {
    const uint64_t need = n_unique + (n_unique && last_was_present ? 1 : 0);
    uint64_t b = 32;
    while (need > b / 2) b *= 2;
    return b;
}

// Storage::set_use_bigcount (storage.hh:129-135): count-min only.
int gt_storage_set_use_bigcount(gt_storage* st, int use) {
    if (!valid_storage(st)) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    if (st->kind != kByte) return fail(GT_ERR_UNSUPPORTED, "bigcount is not supported for this storage.");
    StorageLock lk(st);
    GT_TRY(lk.rc);
    st->use_bigcount = use != 0;
    return GT_OK;
}

int gt_storage_get_use_bigcount(gt_storage* st, int* use) {
    if (!valid_storage(st) || !use) return fail(GT_ERR_INVALID_ARG, "invalid arguments");
    *use = st->use_bigcount ? 1 : 0;
    return GT_OK;
}

// The bigcount map's entries, sorted by k-mer; *n gets their number, and
// keys / vals (cap entries each) get them when they fit.
int gt_storage_bigcounts(gt_storage* st, uint64_t* keys, int16_t* vals, uint64_t cap, uint64_t* n) {
    if (!valid_storage(st) || !n) return fail(GT_ERR_INVALID_ARG, "invalid arguments");
    if (st->kind != kByte) return fail(GT_ERR_UNSUPPORTED, "bigcount is not supported for this storage.");
    GT_TRY(ensure_device());
    StorageLock lk(st);
    GT_TRY(lk.rc);
    hipStream_t s = nullptr;
    std::vector<uint64_t> k;
    std::vector<int16_t> v;
    GT_TRY(big_pairs(st->big, k, v, s));
    *n = k.size();
    if (keys && vals && cap >= k.size()) {
        std::copy(k.begin(), k.end(), keys);
        std::copy(v.begin(), v.end(), vals);
    }
    return GT_OK;
}

// Replaces the bigcount map's entries (the count-min loader,
// bytestorage.cc:271-285); a later pair of one k-mer wins.
int gt_storage_set_bigcounts(gt_storage* st, const uint64_t* keys, const int16_t* vals, uint64_t n) {
    if (!valid_storage(st) || (n && (!keys || !vals))) return fail(GT_ERR_INVALID_ARG, "invalid arguments");
    if (st->kind != kByte) return fail(GT_ERR_UNSUPPORTED, "bigcount is not supported for this storage.");
    GT_TRY(ensure_device());
    StorageLock lk(st);
    GT_TRY(lk.rc);
    GT_HIP(hipDeviceSynchronize());
    return big_set_pairs(st->big, keys, vals, n, nullptr);
}

int gt_storage_counters(gt_storage* st, uint64_t* n_unique, uint64_t* n_occupied) {
    if (!valid_storage(st)) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    GT_TRY(ensure_device());
    StorageLock lk(st);
    GT_TRY(lk.rc);
    uint64_t c[kNCounters];
    GT_HIP(hipDeviceSynchronize());
    GT_TRY(read_counters(st, c, nullptr));
    if (c[kErrors]) return fail(GT_ERR_INTERNAL, "device table probe limit was reached; counters are not exact");
    if (n_unique) *n_unique = c[kNUnique];
    if (n_occupied) *n_occupied = st->kind == kExact ? set_bucket_count(c[kNUnique], c[kNOccupied] != 0)
                                                     : c[kNOccupied];
    return GT_OK;
}

int gt_storage_set_counters(gt_storage* st, uint64_t n_unique, uint64_t n_occupied) {
    if (!valid_storage(st)) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    GT_TRY(ensure_device());
    StorageLock lk(st);
    GT_TRY(lk.rc);
    uint64_t c[2] = {n_unique, n_occupied};
    GT_HIP(hipMemcpy(st->counters.p, c, sizeof(c), hipMemcpyHostToDevice));
    return GT_OK;
}

int gt_storage_read_table(gt_storage* st, uint32_t i, uint8_t* host_dst) {
    if (!valid_storage(st) || !host_dst) return fail(GT_ERR_INVALID_ARG, "invalid arguments");
    if (st->kind == kExact) return fail(GT_ERR_UNSUPPORTED, "the exact set has no raw tables (get_raw_tables() is null)");
    if (i >= st->n_tables) return fail(GT_ERR_INVALID_ARG, "table index out of range");
    GT_TRY(ensure_device());
    StorageLock lk(st);
    GT_TRY(lk.rc);
    GT_HIP(hipDeviceSynchronize());
    GT_HIP(hipMemcpy(host_dst, st->tables.as<uint8_t>() + st->table_off[i], st->nbytes[i], hipMemcpyDeviceToHost));
    return GT_OK;
}

int gt_storage_write_table(gt_storage* st, uint32_t i, const uint8_t* host_src) {
    if (!valid_storage(st) || !host_src) return fail(GT_ERR_INVALID_ARG, "invalid arguments");
    if (st->kind == kExact) return fail(GT_ERR_UNSUPPORTED, "the exact set has no raw tables");
    if (i >= st->n_tables) return fail(GT_ERR_INVALID_ARG, "table index out of range");
    GT_TRY(ensure_device());
    StorageLock lk(st);
    GT_TRY(lk.rc);
    GT_HIP(hipDeviceSynchronize());
    GT_HIP(hipMemcpy(st->tables.as<uint8_t>() + st->table_off[i], host_src, st->nbytes[i], hipMemcpyHostToDevice));
    return GT_OK;
}

int gt_storage_update_from(gt_storage* st, const gt_storage* other) {
    if (!valid_storage(st) || !valid_storage(other)) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    if (st->kind != kBit || other->kind != kBit) return fail(GT_ERR_UNSUPPORTED, "update_from is defined for BitStorage");
    if (st->sizes != other->sizes) return fail(GT_ERR_INVALID_ARG, "both nodegraphs must have same table sizes");
    GT_TRY(ensure_device());
    StorageLock lk(st);
    GT_TRY(lk.rc);
    GT_TRY(refuse_bigcount(st, "update_from"));
    GT_TRY(refuse_bigcount(other, "update_from"));
    GT_HIP(hipDeviceSynchronize());
    for (uint32_t i = 0; i < st->n_tables; ++i) {
        const uint64_t words = (st->nbytes[i] + 3) / 4;
        launch_bit_union(reinterpret_cast<uint32_t*>(st->tables.as<uint8_t>() + st->table_off[i]),
                         reinterpret_cast<const uint32_t*>(other->tables.as<uint8_t>() + other->table_off[i]), words,
                         i == 0, st->counters.as<uint64_t>(), nullptr);
    }
    GT_HIP(hipGetLastError());
    GT_HIP(hipStreamSynchronize(nullptr));
    return GT_OK;
}

int gt_storage_exact_keys(gt_storage* st, uint64_t* host_out, uint64_t cap, uint64_t* n) {
    if (!valid_storage(st) || !n) return fail(GT_ERR_INVALID_ARG, "invalid arguments");
    if (st->kind != kExact) return fail(GT_ERR_UNSUPPORTED, "exact_keys is defined for the exact set");
    GT_TRY(ensure_device());
    StorageLock lk(st);
    GT_TRY(lk.rc);
    GT_HIP(hipDeviceSynchronize());
    DevBuf out, cnt;
    GT_TRY(out.ensure(std::max<uint64_t>(cap, 1) * sizeof(uint64_t)));
    GT_TRY(cnt.ensure(sizeof(unsigned long long)));
    launch_fill_u64(cnt.as<uint64_t>(), 1, 0, nullptr);
    launch_exact_keys(st->exact_view(), out.as<uint64_t>(), cap, cnt.as<unsigned long long>(), nullptr);
    GT_HIP(hipGetLastError());
    uint64_t total = 0;
    uint32_t ones = 0;
    GT_HIP(hipMemcpy(&total, cnt.p, sizeof(total), hipMemcpyDeviceToHost));
    GT_HIP(hipMemcpy(&ones, st->eones.p, sizeof(ones), hipMemcpyDeviceToHost));
    const uint64_t take = std::min(total, cap);
    if (take) GT_HIP(hipMemcpy(host_out, out.p, take * sizeof(uint64_t), hipMemcpyDeviceToHost));
    if (ones) {
        if (total < cap) host_out[total] = ~0ULL;
        total++;
    }
    *n = total;
    return GT_OK;
}

int gt_storage_save_async(gt_storage* st, uint8_t* const* host_tables, uint64_t* counters, void* stream) {
    if (!valid_storage(st) || st->kind == kExact) return fail(GT_ERR_INVALID_ARG, "a Bloom or count-min storage is needed");
    if (!host_tables || !counters) return fail(GT_ERR_INVALID_ARG, "null host buffers");
    GT_TRY(ensure_device());
    StorageLock lk(st);
    GT_TRY(lk.rc);
    GT_TRY(order_after_fence(st, (hipStream_t)stream));
    hipStream_t s = (hipStream_t)stream;
    for (uint32_t i = 0; i < st->n_tables; ++i) {
        GT_HIP(hipMemcpyAsync(host_tables[i], st->tables.as<uint8_t>() + st->table_off[i], st->nbytes[i],
                              hipMemcpyDeviceToHost, s));
    }
    GT_HIP(hipMemcpyAsync(counters, st->counters.p, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    return GT_OK;
}

int gt_storage_load_async(gt_storage* st, const uint8_t* const* host_tables, const uint64_t* counters,
                          void* stream) {
    if (!valid_storage(st) || st->kind == kExact) return fail(GT_ERR_INVALID_ARG, "a Bloom or count-min storage is needed");
    GT_TRY(ensure_device());
    StorageLock lk(st);
    GT_TRY(lk.rc);
    GT_TRY(order_after_fence(st, (hipStream_t)stream));
    hipStream_t s = (hipStream_t)stream;
    if (!host_tables) {
        launch_fill_u64(st->tables.as<uint64_t>(), st->tables.bytes / 8, 0, s);
        launch_fill_u64(st->counters.as<uint64_t>(), 2, 0, s);
        GT_HIP(hipGetLastError());
        return GT_OK;
    }
    if (!counters) return fail(GT_ERR_INVALID_ARG, "null counters");
    for (uint32_t i = 0; i < st->n_tables; ++i) {
        GT_HIP(hipMemcpyAsync(st->tables.as<uint8_t>() + st->table_off[i], host_tables[i], st->nbytes[i],
                              hipMemcpyHostToDevice, s));
    }
    GT_HIP(hipMemcpyAsync(st->counters.p, counters, 2 * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    return GT_OK;
}

// ------------------------------------------------------------ hashing

int gt_kmer_offsets_dev(uint16_t K, const int64_t* offsets, uint64_t n_reads, int64_t* kmer_offsets, void* stream) {
    if (K < 1 || K >= kMaxK) return fail(GT_ERR_INVALID_ARG, "K must be in 1..255");
    if (!kmer_offsets || (n_reads && !offsets)) return fail(GT_ERR_INVALID_ARG, "null pointer");
    GT_TRY(ensure_device());
    hipStream_t s = (hipStream_t)stream;
    DevBuf blocks;
    GT_TRY(blocks.ensure((scan_block_count(n_reads) + 1) * sizeof(int64_t)));
    launch_kmer_offsets(offsets, n_reads, K, kmer_offsets, blocks.as<int64_t>(), nullptr, s);
    GT_HIP(hipGetLastError());
    GT_HIP(hipStreamSynchronize(s));   // `blocks` is released on return
    return GT_OK;
}

int gt_hash_reads_dev(uint16_t K, int mode, const uint8_t* seq, const int64_t* offsets, uint64_t n_reads,
                      uint64_t n_bases, uint64_t* fw, uint64_t* rc, uint8_t* valid, void* stream) {
    GT_TRY(check_read_args(K, mode, seq, offsets, n_reads, n_bases));
    GT_TRY(ensure_device());
    if (n_reads == 0 || n_bases == 0) return GT_OK;
    hipStream_t s = (hipStream_t)stream;
    // A private scratch holder: hashing needs no storage.
    static std::mutex mu;
    static gt_storage* scratch = nullptr;
    std::lock_guard<std::mutex> lk(mu);
    if (!scratch) scratch = new gt_storage();
    scratch->kind = kBit;
    Reads R = make_reads(seq, offsets, n_reads, n_bases, K, mode, 1);
    GT_TRY(prepare_reads(scratch, R, true, s));
    {
        ProfScope ps("hash", s);
        launch_hash(make_launch(s), R, fw, rc, valid);
    }
    GT_HIP(hipGetLastError());
    GT_HIP(hipStreamSynchronize(s));
    return GT_OK;
}

int gt_hash_reads(uint16_t K, int mode, const uint8_t* seq, const int64_t* offsets, uint64_t n_reads, uint64_t* fw,
                  uint64_t* rc, uint8_t* valid) {
    GT_TRY(check_host_offsets(offsets, n_reads));
    const uint64_t n_bases = n_reads ? (uint64_t)offsets[n_reads] : 0;
    GT_TRY(check_read_args(K, mode, seq, offsets, n_reads, n_bases));
    GT_TRY(ensure_device());
    const uint64_t nk = packed_kmers(offsets, n_reads, K);
    if (nk == 0) return GT_OK;
    DevBuf dseq, doff, dfw, drc, dval;
    hipStream_t s = nullptr;
    GT_TRY(stage_in(dseq, seq, n_bases, s));
    GT_TRY(stage_in(doff, offsets, n_reads + 1, s));
    GT_TRY(dfw.ensure(nk * 8));
    GT_TRY(drc.ensure(nk * 8));
    GT_TRY(dval.ensure(nk));
    GT_TRY(gt_hash_reads_dev(K, mode, dseq.as<uint8_t>(), doff.as<int64_t>(), n_reads, n_bases, dfw.as<uint64_t>(),
                             drc.as<uint64_t>(), dval.as<uint8_t>(), s));
    GT_TRY(stage_out(fw, dfw, nk, s));
    GT_TRY(stage_out(rc, drc, nk, s));
    GT_TRY(stage_out(valid, dval, nk, s));
    GT_HIP(hipStreamSynchronize(s));
    return GT_OK;
}

// ------------------------------------------------------------ shifter

int gt_shifter_hash_base(uint16_t K, const uint8_t* kmer, uint64_t* fw, uint64_t* rc) {
    if (K < 1 || K >= kMaxK || !kmer || !fw || !rc) return fail(GT_ERR_INVALID_ARG, "invalid shifter arguments");
    uint64_t f = 0, r = 0;
    for (int i = 0; i < K; ++i) {
        f = gt_rotl(f, 1) ^ GT_CHAR_HASH[kmer[i]];
        r = gt_rotl(r, 1) ^ GT_CHAR_HASH[gt_complement(kmer[K - 1 - i])];
    }
    *fw = f;
    *rc = r;
    return GT_OK;
}

int gt_shifter_shift_right(uint16_t K, uint64_t* fw, uint64_t* rc, uint8_t out, uint8_t in) {
    if (K < 1 || K >= kMaxK || !fw || !rc) return fail(GT_ERR_INVALID_ARG, "invalid shifter arguments");
    const unsigned Kr = K & 63u;
    *fw = gt_rotl(*fw, 1) ^ gt_rotl(GT_CHAR_HASH[out], Kr) ^ GT_CHAR_HASH[in];
    *rc = gt_rotr(*rc ^ gt_rotl(GT_CHAR_HASH[gt_complement(in)], Kr) ^ GT_CHAR_HASH[gt_complement(out)], 1);
    return GT_OK;
}

int gt_shifter_shift_left(uint16_t K, uint64_t* fw, uint64_t* rc, uint8_t in, uint8_t out) {
    if (K < 1 || K >= kMaxK || !fw || !rc) return fail(GT_ERR_INVALID_ARG, "invalid shifter arguments");
    const unsigned Kr = K & 63u;
    *fw = gt_rotr(*fw ^ gt_rotl(GT_CHAR_HASH[in], Kr) ^ GT_CHAR_HASH[out], 1);
    *rc = gt_rotl(*rc, 1) ^ gt_rotl(GT_CHAR_HASH[gt_complement(out)], Kr) ^ GT_CHAR_HASH[gt_complement(in)];
    return GT_OK;
}

// ------------------------------------------------------------ dBG batches

int gt_dbg_insert_dev(gt_storage* st, uint16_t K, int canonical, int mode, const uint8_t* seq, const int64_t* offsets,
                      uint64_t n_reads, uint64_t n_bases, uint64_t* n_new, uint8_t* new_flags, void* stream) {
    if (!valid_storage(st)) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    GT_TRY(check_read_args(K, mode, seq, offsets, n_reads, n_bases));
    GT_TRY(ensure_device());
    StorageLock lk(st);
    GT_TRY(lk.rc);
    GT_TRY(order_after_fence(st, (hipStream_t)stream));
    return dbg_insert_impl(st, make_reads(seq, offsets, n_reads, n_bases, K, mode, canonical), n_new, new_flags,
                           (hipStream_t)stream);
}

int gt_dbg_insert(gt_storage* st, uint16_t K, int canonical, int mode, const uint8_t* seq, const int64_t* offsets,
                  uint64_t n_reads, uint64_t* n_new, uint8_t* new_flags, uint64_t* n_kmers) {
    if (!valid_storage(st)) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    GT_TRY(check_host_offsets(offsets, n_reads));
    const uint64_t n_bases = n_reads ? (uint64_t)offsets[n_reads] : 0;
    GT_TRY(check_read_args(K, mode, seq, offsets, n_reads, n_bases));
    GT_TRY(ensure_device());
    StorageLock lk(st);
    GT_TRY(lk.rc);
    hipStream_t s = nullptr;
    const uint64_t nk = packed_kmers(offsets, n_reads, K);
    GT_TRY(stage_in(st->h_seq, seq, n_bases, s));
    GT_TRY(stage_in(st->h_off, offsets, n_reads + 1, s));
    GT_TRY(st->h_out1.ensure(std::max<uint64_t>(n_reads, 1) * 8));
    GT_TRY(st->h_out2.ensure(std::max<uint64_t>(nk, 1)));
    GT_TRY(dbg_insert_impl(st, make_reads(st->h_seq.as<uint8_t>(), st->h_off.as<int64_t>(), n_reads, n_bases, K, mode,
                                          canonical),
                           n_new ? st->h_out1.as<uint64_t>() : nullptr, new_flags ? st->h_out2.as<uint8_t>() : nullptr,
                           s));
    GT_TRY(stage_out(n_new, st->h_out1, n_reads, s));
    GT_TRY(stage_out(new_flags, st->h_out2, nk, s));
    GT_HIP(hipStreamSynchronize(s));
    GT_TRY(check_kernel_errors(st, s));
    if (n_kmers) {
        if (mode == kRaw) {
            *n_kmers = nk;
        } else {
            uint64_t c = 0;
            for (uint64_t r = 0; r < n_reads; ++r) {
                int run = 0;
                for (int64_t j = offsets[r]; j < offsets[r + 1]; ++j) {
                    run = gt_is_acgt(seq[j]) ? run + 1 : 0;
                    if (run >= K) c++;
                }
            }
            *n_kmers = c;
        }
    }
    return GT_OK;
}

int gt_dbg_query_dev(gt_storage* st, uint16_t K, int canonical, int mode, const uint8_t* seq, const int64_t* offsets,
                     uint64_t n_reads, uint64_t n_bases, int16_t* counts, uint64_t* hashes, void* stream) {
    if (!valid_storage(st)) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    GT_TRY(check_read_args(K, mode, seq, offsets, n_reads, n_bases));
    GT_TRY(ensure_device());
    if (n_reads == 0 || n_bases == 0) return GT_OK;
    StorageLock lk(st);
    GT_TRY(lk.rc);
    GT_TRY(order_after_fence(st, (hipStream_t)stream));
    hipStream_t s = (hipStream_t)stream;
    Reads R = make_reads(seq, offsets, n_reads, n_bases, K, mode, canonical);
    GT_TRY(prepare_reads(st, R, true, s));
    {
        ProfScope ps("query", s);
        launch_query(make_launch(st, s), R, st->tables_view(), st->exact_or_empty(), counts, hashes);
    }
    GT_HIP(hipGetLastError());
    if (st->use_bigcount && st->big.aux) {
        // ByteStorage::query's bigcount step for counts of 255 (bytestorage.cc:130-137)
        int64_t nk = 0;
        GT_HIP(hipMemcpyAsync(&nk, st->own.kmer_off.as<int64_t>() + n_reads, sizeof(nk), hipMemcpyDeviceToHost, s));
        GT_HIP(hipStreamSynchronize(s));
        if (nk > 0) {
            DevBuf fw, rc, valid;
            GT_TRY(fw.ensure((size_t)nk * 8));
            GT_TRY(rc.ensure((size_t)nk * 8));
            GT_TRY(valid.ensure((size_t)nk));
            launch_hash(make_launch(st, s), R, fw.as<uint64_t>(), rc.as<uint64_t>(), valid.as<uint8_t>());
            GT_TRY(big_query(st->big, fw.as<uint64_t>(), rc.as<uint64_t>(), valid.as<uint8_t>(), R.canonical, counts,
                             (uint64_t)nk, g_dev.cus, s));
            GT_HIP(hipStreamSynchronize(s));
        }
    }
    return GT_OK;
}

int gt_dbg_query(gt_storage* st, uint16_t K, int canonical, int mode, const uint8_t* seq, const int64_t* offsets,
                 uint64_t n_reads, int16_t* counts, uint64_t* hashes) {
    if (!valid_storage(st)) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    GT_TRY(check_host_offsets(offsets, n_reads));
    const uint64_t n_bases = n_reads ? (uint64_t)offsets[n_reads] : 0;
    GT_TRY(check_read_args(K, mode, seq, offsets, n_reads, n_bases));
    GT_TRY(ensure_device());
    const uint64_t nk = packed_kmers(offsets, n_reads, K);
    if (nk == 0) return GT_OK;
    hipStream_t s = nullptr;
    StorageLock lk(st);
    GT_TRY(lk.rc);
    GT_TRY(stage_in(st->h_seq, seq, n_bases, s));
    GT_TRY(stage_in(st->h_off, offsets, n_reads + 1, s));
    GT_TRY(st->h_out1.ensure(nk * sizeof(int16_t)));
    GT_TRY(st->h_out3.ensure(nk * sizeof(uint64_t)));
    GT_TRY(gt_dbg_query_dev(st, K, canonical, mode, st->h_seq.as<uint8_t>(), st->h_off.as<int64_t>(), n_reads, n_bases,
                            counts ? st->h_out1.as<int16_t>() : nullptr, hashes ? st->h_out3.as<uint64_t>() : nullptr,
                            s));
    GT_TRY(stage_out(counts, st->h_out1, nk, s));
    GT_TRY(stage_out(hashes, st->h_out3, nk, s));
    GT_HIP(hipStreamSynchronize(s));
    return GT_OK;
}

int gt_dbg_insert_and_query_dev(gt_storage* st, uint16_t K, int canonical, int mode, const uint8_t* seq,
                                const int64_t* offsets, uint64_t n_reads, uint64_t n_bases, int16_t* counts,
                                void* stream) {
    if (!valid_storage(st)) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    GT_TRY(check_read_args(K, mode, seq, offsets, n_reads, n_bases));
    GT_TRY(ensure_device());
    if (n_reads == 0 || n_bases == 0) return GT_OK;
    StorageLock lk(st);
    GT_TRY(lk.rc);
    GT_TRY(order_after_fence(st, (hipStream_t)stream));
    hipStream_t s = (hipStream_t)stream;
    Reads R = make_reads(seq, offsets, n_reads, n_bases, K, mode, canonical);
    if (st->use_bigcount) return bigcount_insert_impl(st, R, nullptr, nullptr, counts, s);
    if (st->kind == kByte && bin_buckets(st)) {
        // Count-min results depend on the exact order of every earlier insert
        // of the same bins: the binned replay gives each event its bins'
        // counts just before it.
        GT_TRY(prepare_reads(st, R, true, s));
        const int rc = binned_insert_impl(st, st->own, R, nullptr, nullptr, counts, s);
        if (rc != GT_ERR_OUT_OF_MEMORY) return rc;
        st->own.bin_ent1.release();     // no room for the event buffers: strict order below
        st->own.bin_ent2.release();
        st->own.bin_dig2.release();
        st->own.bin_vq.release();
    }
    if (st->kind == kByte) {
        // More than four tables or 2^30 bins: hash the batch, then apply it
        // with the sort-based exact insert_and_query.
        GT_TRY(prepare_reads(st, R, true, s));
        int64_t nk = 0;
        GT_HIP(hipMemcpyAsync(&nk, st->own.kmer_off.as<int64_t>() + n_reads, sizeof(nk), hipMemcpyDeviceToHost, s));
        GT_HIP(hipStreamSynchronize(s));
        if (nk <= 0) return GT_OK;
        DevBuf fw, rc, valid;
        GT_TRY(fw.ensure((size_t)nk * 8));
        GT_TRY(rc.ensure((size_t)nk * 8));
        GT_TRY(valid.ensure((size_t)nk));
        const Launch L = make_launch(st, s);
        launch_hash(L, R, fw.as<uint64_t>(), rc.as<uint64_t>(), valid.as<uint8_t>());
        GT_HIP(hipGetLastError());
        GT_TRY(sorted_insert_query(st->tables_view(), fw.as<uint64_t>(), rc.as<uint64_t>(), valid.as<uint8_t>(),
                                   (uint64_t)nk, R.canonical, counts, nullptr, st->counters.as<uint64_t>(), g_dev.cus,
                                   s));
        return GT_OK;   // (synchronised: the hash buffers are released on return)
    }
    GT_TRY(dbg_insert_impl(st, R, nullptr, nullptr, s));
    // Presence storages return 1 after every insert (bitstorage.cc:78-84,
    // sparseppstorage.cc:21-25); a query after the batch gives exactly that
    // (and -1 for skipped windows).
    GT_TRY(prepare_reads(st, R, true, s));
    launch_query(make_launch(st, s), R, st->tables_view(), st->exact_or_empty(), counts, nullptr);
    GT_HIP(hipGetLastError());
    return GT_OK;
}

int gt_dbg_insert_and_query(gt_storage* st, uint16_t K, int canonical, int mode, const uint8_t* seq,
                            const int64_t* offsets, uint64_t n_reads, int16_t* counts) {
    if (!valid_storage(st)) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    GT_TRY(check_host_offsets(offsets, n_reads));
    const uint64_t n_bases = n_reads ? (uint64_t)offsets[n_reads] : 0;
    GT_TRY(check_read_args(K, mode, seq, offsets, n_reads, n_bases));
    GT_TRY(ensure_device());
    const uint64_t nk = packed_kmers(offsets, n_reads, K);
    DevBuf dseq, doff, dout;
    hipStream_t s = nullptr;
    GT_TRY(stage_in(dseq, seq, n_bases, s));
    GT_TRY(stage_in(doff, offsets, n_reads + 1, s));
    GT_TRY(dout.ensure(std::max<uint64_t>(nk, 1) * sizeof(int16_t)));
    GT_TRY(gt_dbg_insert_and_query_dev(st, K, canonical, mode, dseq.as<uint8_t>(), doff.as<int64_t>(), n_reads,
                                       n_bases, dout.as<int16_t>(), s));
    GT_TRY(stage_out(counts, dout, nk, s));
    GT_HIP(hipStreamSynchronize(s));
    return check_kernel_errors(st, s);
}

int gt_solid_filter_dev(gt_storage* st, uint16_t K, int canonical, int mode, const uint8_t* seq, const int64_t* offsets,
                        uint64_t n_reads, uint64_t n_bases, float min_prop_solid, uint32_t solid_threshold,
                        uint8_t* passed, void* stream) {
    if (!valid_storage(st)) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    GT_TRY(check_read_args(K, mode, seq, offsets, n_reads, n_bases));
    GT_TRY(ensure_device());
    if (n_reads == 0) return GT_OK;
    if (!passed) return fail(GT_ERR_INVALID_ARG, "null passed");
    StorageLock lk(st);
    GT_TRY(lk.rc);
    GT_TRY(refuse_bigcount(st, "the batched solid filter"));
    GT_TRY(order_after_fence(st, (hipStream_t)stream));
    hipStream_t s = (hipStream_t)stream;
    GT_TRY(st->solid_counts.ensure(std::max<uint64_t>(n_bases, 1) * sizeof(int16_t)));
    int16_t* counts = st->solid_counts.as<int16_t>();
    GT_TRY(gt_dbg_insert_and_query_dev(st, K, canonical, mode, seq, offsets, n_reads, n_bases, counts, stream));
    // insert_and_query left this batch's packed k-mer offsets in the storage's scratch
    launch_solid_reads(offsets, st->own.kmer_off.as<int64_t>(), n_reads, K, counts, solid_threshold, min_prop_solid,
                       passed, s);
    GT_HIP(hipGetLastError());
    return GT_OK;
}

int gt_solid_filter(gt_storage* st, uint16_t K, int canonical, int mode, const uint8_t* seq, const int64_t* offsets,
                    uint64_t n_reads, float min_prop_solid, uint32_t solid_threshold, uint8_t* passed) {
    if (!valid_storage(st)) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    GT_TRY(check_host_offsets(offsets, n_reads));
    const uint64_t n_bases = n_reads ? (uint64_t)offsets[n_reads] : 0;
    GT_TRY(check_read_args(K, mode, seq, offsets, n_reads, n_bases));
    GT_TRY(ensure_device());
    if (n_reads == 0) return GT_OK;
    DevBuf dseq, doff, dpass;
    hipStream_t s = nullptr;
    GT_TRY(stage_in(dseq, seq, n_bases, s));
    GT_TRY(stage_in(doff, offsets, n_reads + 1, s));
    GT_TRY(dpass.ensure(n_reads));
    GT_TRY(gt_solid_filter_dev(st, K, canonical, mode, dseq.as<uint8_t>(), doff.as<int64_t>(), n_reads, n_bases,
                               min_prop_solid, solid_threshold, dpass.as<uint8_t>(), s));
    GT_TRY(stage_out(passed, dpass, n_reads, s));
    GT_HIP(hipStreamSynchronize(s));
    return check_kernel_errors(st, s);
}

namespace {

// Start positions per diginorm chunk (GOETIA_AMD_DIGINORM_CHUNK, default
// 2^25, at least one first-pass tile) and the number of guesses after which
// a chunk is cut at its first undecided read (GOETIA_AMD_DIGINORM_GUESSES,
// default 32, at least 2).
uint64_t dn_chunk_starts() {
    uint64_t v = 1ULL << 25;
    if (const char* e = std::getenv("GOETIA_AMD_DIGINORM_CHUNK")) {
        const unsigned long long x = std::strtoull(e, nullptr, 10);
        if (x) v = std::min<uint64_t>(x, 1ULL << 28);
    }
    return std::max<uint64_t>(v, kBinTile);
}

int dn_guess_cap() {
    int v = 32;
    if (const char* e = std::getenv("GOETIA_AMD_DIGINORM_GUESSES")) v = std::atoi(e);
    return std::max(v, 2);
}

}  // namespace

int gt_diginorm_dev(gt_storage* st, uint16_t K, int canonical, int mode, const uint8_t* seq, const int64_t* offsets,
                    uint64_t n_reads, uint64_t n_bases, uint32_t cutoff, uint8_t* passed, void* stream) {
    if (!valid_storage(st)) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    if (st->kind != kByte) return fail(GT_ERR_UNSUPPORTED, "DiginormFilter needs a counting (ByteStorage) graph");
    if (!bin_buckets(st))
        return fail(GT_ERR_UNSUPPORTED, "DiginormFilter needs at most 4 tables of at most 2^30 bins in all");
    GT_TRY(check_read_args(K, mode, seq, offsets, n_reads, n_bases));
    GT_TRY(ensure_device());
    if (n_reads == 0) return GT_OK;
    if (!passed) return fail(GT_ERR_INVALID_ARG, "null passed");
    StorageLock lk(st);
    GT_TRY(lk.rc);
    GT_TRY(refuse_bigcount(st, "the batched diginorm filter"));
    GT_TRY(order_after_fence(st, (hipStream_t)stream));
    hipStream_t s = (hipStream_t)stream;
    // chunks end at read boundaries: the offsets come to the host once
    std::vector<int64_t> off(n_reads + 1);
    GT_HIP(hipMemcpyAsync(off.data(), offsets, (n_reads + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    GT_HIP(hipStreamSynchronize(s));
    GT_TRY(check_host_offsets(off.data(), n_reads));
    if ((uint64_t)off[n_reads] != n_bases) return fail(GT_ERR_INVALID_ARG, "offsets[n_reads] must equal n_bases");
    std::vector<int64_t> kof(n_reads + 1, 0);      // packed windows before each read (as kmer_off)
    for (uint64_t r = 0; r < n_reads; ++r) kof[r + 1] = kof[r] + std::max<int64_t>(off[r + 1] - off[r] - K + 1, 0);
    const int64_t Tl = kBinTile;
    auto tiles = [&](uint64_t a, uint64_t b) { return (uint64_t)((off[b] + Tl - 1) / Tl - off[a] / Tl); };
    uint64_t tiles_max = dn_chunk_starts() / kBinTile;
    for (uint64_t r = 0; r < n_reads; ++r) tiles_max = std::max(tiles_max, tiles(r, r + 1));

    Scratch& sc = st->own;
    Reads R = make_reads(seq, offsets, n_reads, n_bases, K, mode, canonical);
    GT_TRY(prepare_reads(sc, R, true, s));
    BinPlan P;
    GT_TRY(bin_prepare(st, sc, tiles_max, P));
    const uint64_t ev_max = tiles_max * st->n_tables * kBinTile;
    const uint64_t win_max = tiles_max * kBinTile;
    GT_TRY(st->dn_rid.ensure(ev_max * sizeof(uint32_t)));
    GT_TRY(st->dn_wread.ensure(win_max * sizeof(uint32_t)));
    GT_TRY(st->dn_hasev.ensure(win_max));
    GT_TRY(st->dn_fails.ensure(win_max));
    GT_TRY(st->dn_state.ensure(n_reads));
    GT_TRY(st->dn_ctl.ensure(2 * sizeof(uint32_t)));
    GT_TRY(sc.bin_flags.ensure(std::max<uint64_t>((uint64_t)kof[n_reads], 1)));
    uint32_t* rid = st->dn_rid.as<uint32_t>();
    uint32_t* wread = st->dn_wread.as<uint32_t>();
    uint8_t* hasev = st->dn_hasev.as<uint8_t>();
    uint8_t* fails = st->dn_fails.as<uint8_t>();
    uint8_t* state = st->dn_state.as<uint8_t>();
    uint32_t* ctl = st->dn_ctl.as<uint32_t>();
    uint8_t* flags = sc.bin_flags.as<uint8_t>();
    P.flags = flags;
    P.vq = nullptr;
    P.rid = rid;
    P.keep = state;
    const Tables T = st->tables_view();
    uint64_t* ctr = st->counters.as<uint64_t>();
    const int grid = std::max(1, g_dev.cus * 8);
    const int cap = dn_guess_cap();

    uint64_t r0 = 0;
    while (r0 < n_reads) {
        uint64_t r1 = r0 + 1;
        while (r1 < n_reads && tiles(r0, r1 + 1) <= tiles_max) ++r1;
        launch_dn_init(R.off, R.kmer_off, (uint32_t)r0, (uint32_t)r1, K, wread, state, s);
        uint64_t r_end = r1;
        const uint64_t n_win = (uint64_t)(kof[r1] - kof[r0]);
        if (n_win) {
            P.hasev = hasev;
            bin_route(R, T, P, sc, off[r0], off[r1], s);
            P.hasev = nullptr;
            launch_dn_event_reads(P, wread, rid, n_win * st->n_tables, s);
            bool settled = false;
            uint32_t h_ctl[2] = {0, 0};
            for (int g = 1; g <= cap && !settled; ++g) {
                {
                    ProfScope ps("dn_replay", s);
                    launch_fill_u8(fails, n_win, 0, s);
                    launch_dn_replay(P, T, rid, state, cutoff, fails, grid, s);
                }
                {
                    ProfScope ps("dn_decide", s);
                    launch_dn_ctl_reset(ctl, s);
                    launch_dn_decide(R.off, R.kmer_off, (uint32_t)r0, (uint32_t)r1, K, hasev, fails, g & 1, state,
                                     ctl, s);
                }
                GT_HIP(hipGetLastError());
                GT_HIP(hipMemcpyAsync(h_ctl, ctl, sizeof(h_ctl), hipMemcpyDeviceToHost, s));
                GT_HIP(hipStreamSynchronize(s));
                settled = h_ctl[0] == 0;
            }
            if (!settled) {
                // reads before the first undecided one are final; the rest
                // start the next chunk
                if (h_ctl[1] <= r0 || h_ctl[1] > r1) return fail(GT_ERR_INTERNAL, "diginorm: no read was decided");
                r_end = h_ctl[1];
            }
            // insert the kept reads of [r0, r_end) in read order
            ProfScope ps("dn_apply", s);
            launch_fill_u8(flags + kof[r0], (uint64_t)(kof[r_end] - kof[r0]), 0, s);
            P.r_end = (uint32_t)r_end;
            launch_dn_apply(P, T, ctr, grid, s);
            launch_bin_read_new(R.kmer_off + r0, r_end - r0, flags, nullptr, ctr, s);
        }
        launch_dn_passed(state, (uint32_t)r0, (uint32_t)r_end, passed, s);
        GT_HIP(hipGetLastError());
        r0 = r_end;
    }
    return GT_OK;
}

int gt_diginorm(gt_storage* st, uint16_t K, int canonical, int mode, const uint8_t* seq, const int64_t* offsets,
                uint64_t n_reads, uint32_t cutoff, uint8_t* passed) {
    if (!valid_storage(st)) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    GT_TRY(check_host_offsets(offsets, n_reads));
    const uint64_t n_bases = n_reads ? (uint64_t)offsets[n_reads] : 0;
    GT_TRY(check_read_args(K, mode, seq, offsets, n_reads, n_bases));
    GT_TRY(ensure_device());
    if (n_reads == 0) return GT_OK;
    DevBuf dseq, doff, dpass;
    hipStream_t s = nullptr;
    GT_TRY(stage_in(dseq, seq, n_bases, s));
    GT_TRY(stage_in(doff, offsets, n_reads + 1, s));
    GT_TRY(dpass.ensure(n_reads));
    GT_TRY(gt_diginorm_dev(st, K, canonical, mode, dseq.as<uint8_t>(), doff.as<int64_t>(), n_reads, n_bases, cutoff,
                           dpass.as<uint8_t>(), s));
    GT_TRY(stage_out(passed, dpass, n_reads, s));
    GT_HIP(hipStreamSynchronize(s));
    return check_kernel_errors(st, s);
}

namespace {

int degrees_impl(gt_storage* st, uint16_t K, int canonical, int mode, int min_count, const uint8_t* seq,
                 const int64_t* offsets, uint64_t n_reads, uint64_t n_bases, const uint8_t* only, uint8_t* in_deg,
                 uint8_t* out_deg, uint64_t* nb_val, uint8_t* nb_mask, void* stream) {
    if (!valid_storage(st)) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    if (!in_deg || !out_deg) return fail(GT_ERR_INVALID_ARG, "null degree outputs");
    if (min_count < 1 || min_count > 255) return fail(GT_ERR_INVALID_ARG, "min_count must be in 1..255");
    GT_TRY(check_read_args(K, mode, seq, offsets, n_reads, n_bases));
    GT_TRY(ensure_device());
    if (n_reads == 0 || n_bases == 0) return GT_OK;
    StorageLock lk(st);
    GT_TRY(lk.rc);
    GT_TRY(refuse_bigcount(st, "the neighbour / degree query"));
    GT_TRY(order_after_fence(st, (hipStream_t)stream));
    hipStream_t s = (hipStream_t)stream;
    Reads R = make_reads(seq, offsets, n_reads, n_bases, K, mode, canonical);
    GT_TRY(prepare_reads(st, R, true, s));
    {
        ProfScope ps("degrees", s);
        launch_degrees(make_launch(st, s), R, st->tables_view(), st->exact_or_empty(), min_count, only, in_deg,
                       out_deg, nb_val, nb_mask);
    }
    GT_HIP(hipGetLastError());
    return GT_OK;
}

}  // namespace

int gt_dbg_degrees_dev(gt_storage* st, uint16_t K, int canonical, int mode, int min_count, const uint8_t* seq,
                       const int64_t* offsets, uint64_t n_reads, uint64_t n_bases, const uint8_t* only,
                       uint8_t* in_deg, uint8_t* out_deg, void* stream) {
    return degrees_impl(st, K, canonical, mode, min_count, seq, offsets, n_reads, n_bases, only, in_deg, out_deg,
                        nullptr, nullptr, stream);
}

int gt_dbg_neighbors_dev(gt_storage* st, uint16_t K, int canonical, int mode, int min_count, const uint8_t* seq,
                         const int64_t* offsets, uint64_t n_reads, uint64_t n_bases, uint8_t* in_deg,
                         uint8_t* out_deg, uint64_t* nb_values, uint8_t* nb_present, void* stream) {
    if (!nb_values || !nb_present) return fail(GT_ERR_INVALID_ARG, "null neighbour outputs");
    return degrees_impl(st, K, canonical, mode, min_count, seq, offsets, n_reads, n_bases, nullptr, in_deg, out_deg,
                        nb_values, nb_present, stream);
}

int gt_dbg_neighbors(gt_storage* st, uint16_t K, int canonical, int mode, int min_count, const uint8_t* seq,
                     const int64_t* offsets, uint64_t n_reads, uint8_t* in_deg, uint8_t* out_deg,
                     uint64_t* nb_values, uint8_t* nb_present) {
    if (!valid_storage(st)) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    if (!in_deg || !out_deg || !nb_values || !nb_present) return fail(GT_ERR_INVALID_ARG, "null outputs");
    GT_TRY(check_host_offsets(offsets, n_reads));
    const uint64_t n_bases = n_reads ? (uint64_t)offsets[n_reads] : 0;
    GT_TRY(check_read_args(K, mode, seq, offsets, n_reads, n_bases));
    GT_TRY(ensure_device());
    const uint64_t nk = packed_kmers(offsets, n_reads, K);
    if (nk == 0) return GT_OK;
    DevBuf dseq, doff, din, dout, dval, dpres;
    hipStream_t s = nullptr;
    GT_TRY(stage_in(dseq, seq, n_bases, s));
    GT_TRY(stage_in(doff, offsets, n_reads + 1, s));
    GT_TRY(din.ensure(nk));
    GT_TRY(dout.ensure(nk));
    GT_TRY(dval.ensure(nk * 8 * sizeof(uint64_t)));
    GT_TRY(dpres.ensure(nk));
    GT_TRY(gt_dbg_neighbors_dev(st, K, canonical, mode, min_count, dseq.as<uint8_t>(), doff.as<int64_t>(), n_reads,
                                n_bases, din.as<uint8_t>(), dout.as<uint8_t>(), dval.as<uint64_t>(),
                                dpres.as<uint8_t>(), s));
    GT_TRY(stage_out(in_deg, din, nk, s));
    GT_TRY(stage_out(out_deg, dout, nk, s));
    GT_TRY(stage_out(nb_values, dval, nk * 8, s));
    GT_TRY(stage_out(nb_present, dpres, nk, s));
    GT_HIP(hipStreamSynchronize(s));
    return GT_OK;
}

int gt_dbg_degrees(gt_storage* st, uint16_t K, int canonical, int mode, int min_count, const uint8_t* seq,
                   const int64_t* offsets, uint64_t n_reads, const uint8_t* only, uint8_t* in_deg,
                   uint8_t* out_deg) {
    if (!valid_storage(st)) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    GT_TRY(check_host_offsets(offsets, n_reads));
    const uint64_t n_bases = n_reads ? (uint64_t)offsets[n_reads] : 0;
    GT_TRY(check_read_args(K, mode, seq, offsets, n_reads, n_bases));
    GT_TRY(ensure_device());
    const uint64_t nk = packed_kmers(offsets, n_reads, K);
    if (nk == 0) return GT_OK;
    DevBuf dseq, doff, donly, din, dout;
    hipStream_t s = nullptr;
    GT_TRY(stage_in(dseq, seq, n_bases, s));
    GT_TRY(stage_in(doff, offsets, n_reads + 1, s));
    if (only) GT_TRY(stage_in(donly, only, nk, s));
    GT_TRY(din.ensure(nk));
    GT_TRY(dout.ensure(nk));
    GT_TRY(gt_dbg_degrees_dev(st, K, canonical, mode, min_count, dseq.as<uint8_t>(), doff.as<int64_t>(), n_reads,
                              n_bases, only ? donly.as<uint8_t>() : nullptr, din.as<uint8_t>(), dout.as<uint8_t>(), s));
    GT_TRY(stage_out(in_deg, din, nk, s));
    GT_TRY(stage_out(out_deg, dout, nk, s));
    GT_HIP(hipStreamSynchronize(s));
    return GT_OK;
}

}  // extern "C"

// find_new_segments' graph view (gt_segments.cpp): every window's count and
// hash (forward, raw bytes), and the eight neighbour values and presence mask
// of the windows whose count is 0 -- the only ones its passes look at --
// gathered on the device so that only they cross to the host.
namespace gt {
int segment_view(gt_storage* st, uint16_t K, const uint8_t* seq, const int64_t* offsets, uint64_t n_reads,
                 int16_t* counts, uint64_t* hashes, std::vector<uint64_t>& nbc, std::vector<uint8_t>& pc) {
    if (!valid_storage(st)) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    GT_TRY(check_host_offsets(offsets, n_reads));
    const uint64_t n_bases = n_reads ? (uint64_t)offsets[n_reads] : 0;
    GT_TRY(check_read_args(K, kRaw, seq, offsets, n_reads, n_bases));
    GT_TRY(ensure_device());
    nbc.clear();
    pc.clear();
    const uint64_t nk = packed_kmers(offsets, n_reads, K);
    if (nk == 0) return GT_OK;
    hipStream_t s = nullptr;
    StorageLock lk(st);
    GT_TRY(lk.rc);
    GT_TRY(stage_in(st->h_seq, seq, n_bases, s));
    GT_TRY(stage_in(st->h_off, offsets, n_reads + 1, s));
    GT_TRY(st->h_out1.ensure(nk * sizeof(int16_t)));
    GT_TRY(st->h_out3.ensure(nk * sizeof(uint64_t)));
    GT_TRY(st->h_out2.ensure(2 * nk));
    GT_TRY(st->seg_nb.ensure(nk * 8 * sizeof(uint64_t)));
    GT_TRY(st->seg_pres.ensure(nk));
    const uint8_t* dseq = st->h_seq.as<uint8_t>();
    const int64_t* doff = st->h_off.as<int64_t>();
    GT_TRY(gt_dbg_query_dev(st, K, 0, kRaw, dseq, doff, n_reads, n_bases, st->h_out1.as<int16_t>(),
                            st->h_out3.as<uint64_t>(), s));
    GT_TRY(gt_dbg_neighbors_dev(st, K, 0, kRaw, 1, dseq, doff, n_reads, n_bases, st->h_out2.as<uint8_t>(),
                                st->h_out2.as<uint8_t>() + nk, st->seg_nb.as<uint64_t>(), st->seg_pres.as<uint8_t>(),
                                s));
    GT_TRY(stage_out(counts, st->h_out1, nk, s));
    GT_TRY(stage_out(hashes, st->h_out3, nk, s));
    GT_HIP(hipStreamSynchronize(s));
    std::vector<uint64_t> idx;
    for (uint64_t w = 0; w < nk; ++w)
        if (counts[w] == 0) idx.push_back(w);
    const uint64_t n_new = idx.size();
    if (n_new == 0) return GT_OK;
    nbc.resize(n_new * 8);
    pc.resize(n_new);
    GT_TRY(stage_in(st->seg_idx, idx.data(), n_new, s));
    GT_TRY(st->seg_nbc.ensure(n_new * 8 * sizeof(uint64_t)));
    GT_TRY(st->seg_pc.ensure(n_new));
    launch_seg_gather(st->seg_nb.as<uint64_t>(), st->seg_pres.as<uint8_t>(), st->seg_idx.as<uint64_t>(), n_new,
                      st->seg_nbc.as<uint64_t>(), st->seg_pc.as<uint8_t>(), s);
    GT_HIP(hipGetLastError());
    GT_TRY(stage_out(nbc.data(), st->seg_nbc, n_new * 8, s));
    GT_TRY(stage_out(pc.data(), st->seg_pc, n_new, s));
    GT_HIP(hipStreamSynchronize(s));
    return GT_OK;
}
}  // namespace gt

extern "C" {

// ------------------------------------------------------------ per-hash

int gt_insert_hashes_dev(gt_storage* st, const uint64_t* hashes, uint64_t n, uint8_t* is_new, void* stream) {
    if (!valid_storage(st)) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    if (n && !hashes) return fail(GT_ERR_INVALID_ARG, "null hashes");
    GT_TRY(ensure_device());
    StorageLock lk(st);
    GT_TRY(lk.rc);
    GT_TRY(order_after_fence(st, (hipStream_t)stream));
    return hashes_insert_impl(st, hashes, n, is_new, (hipStream_t)stream);
}

int gt_insert_hashes(gt_storage* st, const uint64_t* hashes, uint64_t n, uint8_t* is_new) {
    if (!valid_storage(st)) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    if (n && !hashes) return fail(GT_ERR_INVALID_ARG, "null hashes");
    GT_TRY(ensure_device());
    if (n == 0) return GT_OK;
    StorageLock lk(st);
    GT_TRY(lk.rc);
    hipStream_t s = nullptr;
    GT_TRY(stage_in(st->h_out3, hashes, n, s));
    GT_TRY(st->h_out2.ensure(n));
    GT_TRY(hashes_insert_impl(st, st->h_out3.as<uint64_t>(), n, is_new ? st->h_out2.as<uint8_t>() : nullptr, s));
    GT_TRY(stage_out(is_new, st->h_out2, n, s));
    GT_HIP(hipStreamSynchronize(s));
    return check_kernel_errors(st, s);
}

int gt_query_hashes_dev(gt_storage* st, const uint64_t* hashes, uint64_t n, int16_t* counts, void* stream) {
    if (!valid_storage(st)) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    if (n && (!hashes || !counts)) return fail(GT_ERR_INVALID_ARG, "null pointer");
    GT_TRY(ensure_device());
    if (n == 0) return GT_OK;
    StorageLock lk(st);
    GT_TRY(lk.rc);
    GT_TRY(order_after_fence(st, (hipStream_t)stream));
    hipStream_t s = (hipStream_t)stream;
    launch_hash_query(make_launch(st, s), hashes, n, st->tables_view(), st->exact_or_empty(), counts);
    GT_HIP(hipGetLastError());
    if (st->use_bigcount) GT_TRY(big_query(st->big, hashes, nullptr, nullptr, 0, counts, n, g_dev.cus, s));
    return GT_OK;
}

int gt_query_hashes(gt_storage* st, const uint64_t* hashes, uint64_t n, int16_t* counts) {
    if (!valid_storage(st)) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    if (n && (!hashes || !counts)) return fail(GT_ERR_INVALID_ARG, "null pointer");
    GT_TRY(ensure_device());
    if (n == 0) return GT_OK;
    DevBuf dh, dc;
    hipStream_t s = nullptr;
    GT_TRY(stage_in(dh, hashes, n, s));
    GT_TRY(dc.ensure(n * sizeof(int16_t)));
    GT_TRY(gt_query_hashes_dev(st, dh.as<uint64_t>(), n, dc.as<int16_t>(), s));
    GT_TRY(stage_out(counts, dc, n, s));
    GT_HIP(hipStreamSynchronize(s));
    return GT_OK;
}

// Storage::insert_and_query per hash in index order.  Count-min results
// depend on every earlier insert of the same bins: the sort-based exact
// insert_and_query gives each hash its bins' counts in index order.
int gt_insert_and_query_hashes(gt_storage* st, const uint64_t* hashes, uint64_t n, int16_t* counts) {
    if (!valid_storage(st)) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    if (n && (!hashes || !counts)) return fail(GT_ERR_INVALID_ARG, "null pointer");
    GT_TRY(ensure_device());
    if (n == 0) return GT_OK;
    StorageLock lk(st);
    GT_TRY(lk.rc);
    hipStream_t s = nullptr;
    GT_TRY(stage_in(st->h_out3, hashes, n, s));
    GT_TRY(st->h_out2.ensure(n));
    GT_TRY(st->h_out1.ensure(n * sizeof(int16_t)));
    if (st->kind == kByte) {
        DevBuf full;
        uint8_t* fl = nullptr;
        if (st->use_bigcount) {
            GT_TRY(full.ensure(n));
            fl = full.as<uint8_t>();
        }
        GT_TRY(sorted_insert_query(st->tables_view(), st->h_out3.as<uint64_t>(), nullptr, nullptr, n, 0,
                                   st->h_out1.as<int16_t>(), nullptr, st->counters.as<uint64_t>(), g_dev.cus, s, fl));
        if (fl)
            GT_TRY(bigcount_iq_chunks(st, st->h_out3.as<uint64_t>(), nullptr, nullptr, 0, fl, st->h_out1.as<int16_t>(),
                                      n, s));
        GT_TRY(stage_out(counts, st->h_out1, n, s));
        GT_HIP(hipStreamSynchronize(s));
    } else {
        GT_TRY(hashes_insert_impl(st, st->h_out3.as<uint64_t>(), n, nullptr, s));
        GT_HIP(hipStreamSynchronize(s));
        for (uint64_t i = 0; i < n; ++i) counts[i] = 1;
    }
    return check_kernel_errors(st, s);
}

// ------------------------------------------------------------ tuning

int gt_tune_set(int kind, int per, int blocks_per_cu) {
    if (kind < 0 || kind > 3) return fail(GT_ERR_INVALID_ARG, "kind must be 0..3");
    if (per != 8 && per != 16 && per != 32) return fail(GT_ERR_INVALID_ARG, "per must be 8, 16 or 32");
    if (blocks_per_cu < 1 || blocks_per_cu > 32) return fail(GT_ERR_INVALID_ARG, "blocks_per_cu must be 1..32");
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[kind].per = per;
    g_tune[kind].blocks_per_cu = blocks_per_cu;
    return GT_OK;
}

int gt_tune_get(int kind, int* per, int* blocks_per_cu) {
    if (kind < 0 || kind > 3 || !per || !blocks_per_cu) return fail(GT_ERR_INVALID_ARG, "invalid arguments");
    std::lock_guard<std::mutex> lk(g_tune_mu);
    *per = g_tune[kind].per;
    *blocks_per_cu = g_tune[kind].blocks_per_cu;
    return GT_OK;
}

int gt_autotune(gt_storage* st, uint16_t K, int canonical, int mode, const uint8_t* seq, const int64_t* offsets,
                uint64_t n_reads, uint64_t n_bases, int* per_out, int* bpc_out, void* stream) {
    if (!valid_storage(st)) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    GT_TRY(check_read_args(K, mode, seq, offsets, n_reads, n_bases));
    GT_TRY(ensure_device());
    if (n_reads == 0 || n_bases == 0) return fail(GT_ERR_INVALID_ARG, "autotune needs a non-empty batch");
    StorageLock lk(st);
    GT_TRY(lk.rc);
    hipStream_t s = (hipStream_t)stream;
    Reads R = make_reads(seq, offsets, n_reads, n_bases, K, mode, canonical);
    GT_TRY(prepare_reads(st, R, true, s));
    uint64_t sub = 0;
    GT_TRY(prepare_claims(st, n_bases, &sub, s));
    R.b1 = std::min<int64_t>(R.n_bases, (int64_t)sub);
    DevBuf counts;
    GT_TRY(counts.ensure((size_t)std::max<uint64_t>(n_bases, 1) * sizeof(int16_t)));
    hipEvent_t e0, e1;
    GT_HIP(hipEventCreate(&e0));
    GT_HIP(hipEventCreate(&e1));
    const int pers[3] = {8, 16, 32};
    const int bpcs[3] = {4, 8, 16};
    float best = 1e30f;
    Tune best_t;
    Tune saved;
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        saved = g_tune[st->kind];
    }
    uint64_t* ctr = st->counters.as<uint64_t>();
    for (int pi = 0; pi < 3; ++pi) {
        for (int bi = 0; bi < 3; ++bi) {
            {
                std::lock_guard<std::mutex> g(g_tune_mu);
                g_tune[st->kind].per = pers[pi];
                g_tune[st->kind].blocks_per_cu = bpcs[bi];
            }
            const Launch L = make_launch(st, s);
            float total = 0.f;
            for (int rep = 0; rep < 2; ++rep) {
                GT_HIP(hipEventRecord(e0, s));
                // read-only passes: the claim (or lookup) pass writes only
                // scratch; a later insert starts a new generation first.
                if (st->kind == kExact) {
                    launch_exact_lookup(L, R, st->exact_view(), st->own.fresh.as<uint8_t>(), ctr);
                } else {
                    launch_claim(L, R, st->tables_view(), claim_view(st->own), st->own.fresh.as<uint8_t>(), ctr);
                }
                launch_query(L, R, st->tables_view(), st->exact_or_empty(), counts.as<int16_t>(), nullptr);
                GT_HIP(hipEventRecord(e1, s));
                GT_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                GT_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (rep) total += ms;   // the first repetition warms up
            }
            if (total < best) {
                best = total;
                best_t = Tune{pers[pi], bpcs[bi]};
            }
        }
    }
    (void)saved;
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        g_tune[st->kind] = best_t;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (per_out) *per_out = best_t.per;
    if (bpc_out) *bpc_out = best_t.blocks_per_cu;
    return GT_OK;
}

// ------------------------------------------------------------ pipeline

namespace {

void pipeline_release(gt_pipeline* p) {
    for (auto& sl : p->slot) {
        if (sl.exec) (void)hipGraphExecDestroy(sl.exec);
        if (sl.copied) (void)hipEventDestroy(sl.copied);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.h_seq) (void)hipHostFree(sl.h_seq);
        if (sl.h_off) (void)hipHostFree(sl.h_off);
        if (sl.h_dyn) (void)hipHostFree(sl.h_dyn);
        if (sl.h_nnew) (void)hipHostFree(sl.h_nnew);
        sl = gt_pipeline::Slot();
    }
    if (p->compute) (void)hipStreamDestroy(p->compute);
    if (p->copy) (void)hipStreamDestroy(p->copy);
    p->compute = p->copy = nullptr;
}

// Enqueues (while the compute stream is being captured) the device work of
// one batch of slot `sl`.
int pipeline_enqueue(gt_pipeline* p, gt_pipeline::Slot& sl) {
    gt_storage* st = p->st;
    hipStream_t s = p->compute;
    Reads R = make_reads(sl.d_seq.as<uint8_t>(), sl.d_off.as<int64_t>(), p->max_reads, p->max_bases, p->K, p->mode,
                         p->canonical);
    R.dyn = sl.d_dyn.as<int64_t>();
    const int64_t n_grains = ((int64_t)p->max_bases + kTileGrain - 1) / kTileGrain;
    launch_fill_u64(p->sc.tile_read.as<uint64_t>(), p->sc.tile_read.bytes / 8, 0, s);
    launch_tile_read(R.off, R.n_reads, R.n_bases, p->sc.tile_read.as<uint32_t>(), n_grains, R.dyn, s);
    R.tile_read = p->sc.tile_read.as<uint32_t>();
    launch_fill_u64(sl.d_nnew.as<uint64_t>(), p->max_reads, 0, s);
    return enqueue_subbatch(st, p->sc, make_launch(st, s), R, sl.d_nnew.as<uint64_t>(), nullptr, s, false);
}

int pipeline_capture(gt_pipeline* p, gt_pipeline::Slot& sl) {
    if (sl.exec) {
        (void)hipGraphExecDestroy(sl.exec);
        sl.exec = nullptr;
    }
    const uint64_t gen_before = p->st->gen_issued, claim_before = p->sc.claim_gen_issued;
    const uint64_t ub_before = p->st->ecount_ub;
    GT_HIP(hipStreamBeginCapture(p->compute, hipStreamCaptureModeThreadLocal));
    int rc = pipeline_enqueue(p, sl);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(p->compute, &graph);
    // Capturing ran nothing: the generation bumps and the set growth happen
    // on replay and are counted there.
    p->st->gen_issued = gen_before;
    p->sc.claim_gen_issued = claim_before;
    p->st->ecount_ub = ub_before;
    if (rc != GT_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail(GT_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&sl.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(GT_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    sl.epoch = p->st->epoch;
    return GT_OK;
}

int pipeline_deliver(gt_pipeline::Slot& sl) {
    if (!sl.busy) return GT_OK;
    GT_HIP(hipEventSynchronize(sl.done));
    if (sl.user_out) std::memcpy(sl.user_out, sl.h_nnew, sl.n_reads * sizeof(uint64_t));
    sl.busy = false;
    sl.user_out = nullptr;
    return GT_OK;
}

}  // namespace

int gt_pipeline_new(gt_storage* st, uint16_t K, int canonical, int mode, uint64_t max_reads, uint64_t max_bases,
                    gt_pipeline** out) {
    if (!out) return fail(GT_ERR_INVALID_ARG, "null output");
    *out = nullptr;
    if (!valid_storage(st)) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    if (K < 1 || K >= kMaxK) return fail(GT_ERR_INVALID_ARG, "K must be in 1..255");
    if (mode != kRaw && mode != kSkipN) return fail(GT_ERR_INVALID_ARG, "unknown input mode");
    if (max_reads == 0 || max_bases == 0 || max_reads >= 0xFFFFFFFFULL)
        return fail(GT_ERR_INVALID_ARG, "invalid batch bounds");
    GT_TRY(ensure_device());
    StorageLock lk(st);
    GT_TRY(lk.rc);
    GT_TRY(refuse_bigcount(st, "the captured pipeline"));
    const uint64_t per = std::max<uint32_t>(st->n_tables, 1);
    const uint64_t cap = next_pow2(2 * per * (max_bases + kTileMax));
    if (st->kind != kExact && cap > (1ULL << 29))
        return fail(GT_ERR_INVALID_ARG, "max_bases exceeds one first-touch sub-batch");
    if (max_bases >= (1ULL << 31)) return fail(GT_ERR_INVALID_ARG, "max_bases must be below 2^31");
    gt_pipeline* p = new (std::nothrow) gt_pipeline();
    if (!p) return fail(GT_ERR_OUT_OF_MEMORY, "host allocation failed");
    p->st = st;
    p->K = K;
    p->canonical = canonical ? 1 : 0;
    p->mode = mode;
    p->max_reads = max_reads;
    p->max_bases = max_bases;
    int rc = GT_OK;
    auto hip_ok = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && rc == GT_OK) rc = fail(GT_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
        return rc == GT_OK;
    };
    hip_ok(hipStreamCreateWithFlags(&p->compute, hipStreamNonBlocking), "hipStreamCreate");
    hip_ok(hipStreamCreateWithFlags(&p->copy, hipStreamNonBlocking), "hipStreamCreate");
    for (auto& sl : p->slot) {
        if (rc != GT_OK) break;
        hip_ok(hipHostMalloc((void**)&sl.h_seq, max_bases), "hipHostMalloc");
        hip_ok(hipHostMalloc((void**)&sl.h_off, (max_reads + 1) * sizeof(int64_t)), "hipHostMalloc");
        hip_ok(hipHostMalloc((void**)&sl.h_dyn, 2 * sizeof(int64_t)), "hipHostMalloc");
        hip_ok(hipHostMalloc((void**)&sl.h_nnew, max_reads * sizeof(uint64_t)), "hipHostMalloc");
        hip_ok(hipEventCreateWithFlags(&sl.copied, hipEventDisableTiming), "hipEventCreate");
        hip_ok(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming), "hipEventCreate");
        if (rc == GT_OK) rc = sl.d_seq.ensure(max_bases);
        if (rc == GT_OK) rc = sl.d_off.ensure((max_reads + 1) * sizeof(int64_t));
        if (rc == GT_OK) rc = sl.d_dyn.ensure(2 * sizeof(int64_t));
        if (rc == GT_OK) rc = sl.d_nnew.ensure(max_reads * sizeof(uint64_t));
    }
    if (rc == GT_OK) rc = p->sc.tile_read.ensure(((max_bases + kTileGrain - 1) / kTileGrain + 1) * sizeof(uint32_t));
    if (rc == GT_OK) rc = p->sc.fresh.ensure((max_bases + kTileMax - 1) / kTileMax * kTileMax);
    if (rc == GT_OK) rc = alloc_claims(p->sc, st->kind != kExact ? cap : 2, p->compute);
    p->sc.fixed = true;
    if (rc == GT_OK) hip_ok(hipStreamSynchronize(p->compute), "hipStreamSynchronize");
    if (rc != GT_OK) {
        pipeline_release(p);
        delete p;
        return rc;
    }
    st->pipelines.push_back(p);
    *out = p;
    return GT_OK;
}

int gt_pipeline_submit(gt_pipeline* p, const uint8_t* seq, const int64_t* offsets, uint64_t n_reads, uint64_t* n_new) {
    if (!p || !p->st) return fail(GT_ERR_INVALID_ARG, "invalid pipeline");
    GT_TRY(check_host_offsets(offsets, n_reads));
    const uint64_t n_bases = n_reads ? (uint64_t)offsets[n_reads] : 0;
    if (n_reads > p->max_reads || n_bases > p->max_bases) return fail(GT_ERR_INVALID_ARG, "batch exceeds the pipeline bounds");
    if (n_bases && !seq) return fail(GT_ERR_INVALID_ARG, "null reads");
    GT_TRY(ensure_device());
    gt_storage* st = p->st;
    StorageLock lk(st);
    GT_TRY(lk.rc);
    GT_TRY(refuse_bigcount(st, "the captured pipeline"));
    GT_TRY(order_after_fence(st, p->compute));
    gt_pipeline::Slot& sl = p->slot[p->next];
    p->next ^= 1;
    GT_TRY(pipeline_deliver(sl));
    if (n_reads == 0) return GT_OK;
    if (st->kind == kExact && 2 * (st->ecount_ub + n_bases) > st->ecap) {
        // Grow the set between batches: finish in-flight work, read the exact
        // size, rehash; graphs holding the old table are captured again.
        GT_TRY(pipeline_deliver(p->slot[0]));
        GT_TRY(pipeline_deliver(p->slot[1]));
        GT_HIP(hipStreamSynchronize(p->compute));
        uint64_t c[kNCounters];
        GT_TRY(read_counters(st, c, p->compute));
        st->ecount_ub = c[kNUnique];
        GT_TRY(exact_reserve(st, n_bases, p->compute));
    }
    std::memcpy(sl.h_seq, seq, n_bases);
    std::memcpy(sl.h_off, offsets, (n_reads + 1) * sizeof(int64_t));
    sl.h_dyn[0] = (int64_t)n_reads;
    sl.h_dyn[1] = (int64_t)n_bases;
    GT_HIP(hipMemcpyAsync(sl.d_seq.p, sl.h_seq, n_bases, hipMemcpyHostToDevice, p->copy));
    GT_HIP(hipMemcpyAsync(sl.d_off.p, sl.h_off, (n_reads + 1) * sizeof(int64_t), hipMemcpyHostToDevice, p->copy));
    GT_HIP(hipMemcpyAsync(sl.d_dyn.p, sl.h_dyn, 2 * sizeof(int64_t), hipMemcpyHostToDevice, p->copy));
    GT_HIP(hipEventRecord(sl.copied, p->copy));
    GT_HIP(hipStreamWaitEvent(p->compute, sl.copied, 0));
    // The replay bumps both generations once; skip the claim-map value that
    // would wrap its key field (next_gen clears the map there).
    if (((p->sc.claim_gen_issued + 1) & kGenKeyMask) == 0) GT_TRY(next_gen(st, p->sc, p->compute));
    if (!sl.exec || sl.epoch != st->epoch) GT_TRY(pipeline_capture(p, sl));
    if (st->gen_issued + 2 >= 0xFFFFFFFFULL) return fail(GT_ERR_INTERNAL, "batch generation counter exhausted");
    st->gen_issued++;
    p->sc.claim_gen_issued++;
    if (st->kind == kExact) st->ecount_ub += n_bases;
    GT_HIP(hipGraphLaunch(sl.exec, p->compute));
    GT_HIP(hipMemcpyAsync(sl.h_nnew, sl.d_nnew.p, n_reads * sizeof(uint64_t), hipMemcpyDeviceToHost, p->compute));
    GT_HIP(hipEventRecord(sl.done, p->compute));
    sl.busy = true;
    sl.user_out = n_new;
    sl.n_reads = n_reads;
    return GT_OK;
}

int gt_pipeline_flush(gt_pipeline* p) {
    if (!p || !p->st) return fail(GT_ERR_INVALID_ARG, "invalid pipeline");
    StorageLock lk(p->st);
    GT_TRY(lk.rc);
    // deliver in submission order
    GT_TRY(pipeline_deliver(p->slot[p->next]));
    GT_TRY(pipeline_deliver(p->slot[p->next ^ 1]));
    GT_HIP(hipStreamSynchronize(p->compute));
    return check_kernel_errors(p->st, p->compute);
}

int gt_pipeline_free(gt_pipeline* p) {
    if (!p) return GT_OK;
    if (p->st) {
        StorageLock lk(p->st);   // a failed flush of held reads does not keep the pipeline alive
        (void)pipeline_deliver(p->slot[0]);
        (void)pipeline_deliver(p->slot[1]);
        auto& v = p->st->pipelines;
        v.erase(std::remove(v.begin(), v.end(), p), v.end());
    }
    pipeline_release(p);
    delete p;
    return GT_OK;
}

// ------------------------------------------------------------ workspaces

int gt_workspace_new(gt_workspace** out) {
    if (!out) return fail(GT_ERR_INVALID_ARG, "null output");
    GT_TRY(ensure_device());
    *out = new (std::nothrow) gt_workspace();
    if (!*out) return fail(GT_ERR_OUT_OF_MEMORY, "host allocation failed");
    return GT_OK;
}

int gt_workspace_free(gt_workspace* ws) {
    if (!ws) return GT_OK;
    (void)hipDeviceSynchronize();
    delete ws;
    return GT_OK;
}

}  // extern "C"

// ------------------------------------------------------------ allocation-free insert
//
// gt_dbg_insert_noalloc_dev runs the atomic first-touch insert with its
// scratch (claim map, per-start fresh flags, tile -> read table, packed k-mer
// offsets) in one fixed device array instead of library allocations.  Each GPU
// holds its own copy of the array for the life of the process; host code
// reaches it only through the address hipGetSymbolAddress gives on that GPU.
// Calls share it: they must not overlap on one GPU.

constexpr uint64_t kNaClaimSlots = 1ULL << 22;          // 2^19 starts per sub-batch at 4 tables
constexpr uint64_t kNaFreshBytes = 1ULL << 20;          // >= the largest sub-batch (2^21 starts, 1 table)
constexpr uint64_t kNaMaxBases = 1ULL << 30;
constexpr uint64_t kNaMaxReadsFlags = 1ULL << 22;       // reads of a call that asks for new_flags
constexpr uint64_t kNaKeysOff = 0;
constexpr uint64_t kNaValsOff = kNaKeysOff + kNaClaimSlots * 8;
constexpr uint64_t kNaGenOff = kNaValsOff + kNaClaimSlots * 8;
constexpr uint64_t kNaFreshOff = kNaGenOff + 256;
constexpr uint64_t kNaTileOff = kNaFreshOff + kNaFreshBytes;
constexpr uint64_t kNaTileBytes = ((kNaMaxBases + kTileGrain - 1) / kTileGrain + 1) * 4 + 256;
constexpr uint64_t kNaKoffOff = (kNaTileOff + kNaTileBytes + 255) / 256 * 256;
constexpr uint64_t kNaKoffBytes = (kNaMaxReadsFlags + 1) * 8;
constexpr uint64_t kNaScanOff = (kNaKoffOff + kNaKoffBytes + 255) / 256 * 256;
constexpr uint64_t kNaScanBytes = 1ULL << 20;
constexpr uint64_t kNaBytes = kNaScanOff + kNaScanBytes;   // about 98 MiB

static __device__ __align__(256) uint8_t scratch[kNaBytes];

namespace {

struct NoAllocScratch {
    std::mutex mu;
    bool ready = false;
    uint8_t* base = nullptr;   // this GPU's address of `scratch`
    Scratch sc;
};
constexpr int kNaDevices = 64;
NoAllocScratch g_noalloc[kNaDevices];

// Once per GPU, with that GPU current: take the array's device address and
// lay the scratch out over it; the claim map starts cleared.
int noalloc_setup(NoAllocScratch& na) {
    if (na.ready) return GT_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(GT_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    uint8_t* b = static_cast<uint8_t*>(ptr);
    Scratch& sc = na.sc;
    sc.claim_keys.borrow(b + kNaKeysOff, kNaClaimSlots * 8);
    sc.claim_vals.borrow(b + kNaValsOff, kNaClaimSlots * 8);
    sc.claim_gen.borrow(b + kNaGenOff, 256);
    sc.fresh.borrow(b + kNaFreshOff, kNaFreshBytes);
    sc.tile_read.borrow(b + kNaTileOff, kNaTileBytes);
    sc.kmer_off.borrow(b + kNaKoffOff, kNaKoffBytes);
    sc.scan_blocks.borrow(b + kNaScanOff, kNaScanBytes);
    sc.fixed = true;
    launch_fill_u64(sc.claim_keys.as<uint64_t>(), kNaClaimSlots, 0, nullptr);
    launch_fill_u64(sc.claim_vals.as<uint64_t>(), kNaClaimSlots, 0, nullptr);
    launch_fill_u64(sc.claim_gen.as<uint64_t>(), 1, 0, nullptr);
    GT_HIP(hipGetLastError());
    GT_HIP(hipStreamSynchronize(nullptr));
    sc.claim_cap = kNaClaimSlots;
    sc.claim_gen_issued = 0;
    na.base = b;
    na.ready = true;
    return GT_OK;
}

}  // namespace

extern "C" {

int gt_dbg_insert_noalloc_dev(gt_storage* st, uint16_t K, int canonical, int mode, const uint8_t* seq,
                              const int64_t* offsets, uint64_t n_reads, uint64_t n_bases, uint64_t* n_new,
                              uint8_t* new_flags, void* stream) {
    if (st && st->use_bigcount) return refuse_bigcount(st, "the allocation-free insert");
    if (!valid_storage(st)) return fail(GT_ERR_INVALID_ARG, "invalid storage");
    GT_TRY(check_read_args(K, mode, seq, offsets, n_reads, n_bases));
    if (st->kind == kExact)
        return fail(GT_ERR_UNSUPPORTED, "gt_dbg_insert_noalloc_dev takes Bloom and count-min storages (the exact set "
                                        "grows by allocating)");
    if (n_bases > kNaMaxBases)
        return fail(GT_ERR_INVALID_ARG, "gt_dbg_insert_noalloc_dev: n_bases " + std::to_string(n_bases) +
                                            " exceeds its fixed scratch (at most 2^30 bases per call)");
    if (new_flags && n_reads > kNaMaxReadsFlags)
        return fail(GT_ERR_INVALID_ARG, "gt_dbg_insert_noalloc_dev: with new_flags at most 2^22 reads per call");
    GT_TRY(ensure_device());
    if (g_dev.device < 0 || g_dev.device >= kNaDevices) return fail(GT_ERR_DEVICE, "device index out of range");
    NoAllocScratch& na = g_noalloc[g_dev.device];
    std::lock_guard<std::mutex> ln(na.mu);
    StorageLock lk(st);
    GT_TRY(lk.rc);
    GT_TRY(order_after_fence(st, (hipStream_t)stream));
    GT_TRY(noalloc_setup(na));
    return dbg_insert_impl(st, na.sc, make_reads(seq, offsets, n_reads, n_bases, K, mode, canonical), n_new,
                           new_flags, (hipStream_t)stream);
}

int gt_dbg_insert_ws_dev(gt_storage* st, gt_workspace* ws, uint16_t K, int canonical, int mode, const uint8_t* seq,
                         const int64_t* offsets, uint64_t n_reads, uint64_t n_bases, uint64_t* n_new,
                         uint8_t* new_flags, void* stream) {
    if (!valid_storage(st) || !ws) return fail(GT_ERR_INVALID_ARG, "invalid storage or workspace");
    GT_TRY(check_read_args(K, mode, seq, offsets, n_reads, n_bases));
    GT_TRY(ensure_device());
    std::lock_guard<std::recursive_mutex> lw(ws->mu);
    StorageLock lk(st);
    GT_TRY(lk.rc);
    GT_TRY(order_after_fence(st, (hipStream_t)stream));
    return dbg_insert_impl(st, ws->sc, make_reads(seq, offsets, n_reads, n_bases, K, mode, canonical), n_new,
                           new_flags, (hipStream_t)stream);
}

int gt_dbg_query_ws_dev(gt_storage* st, gt_workspace* ws, uint16_t K, int canonical, int mode, const uint8_t* seq,
                        const int64_t* offsets, uint64_t n_reads, uint64_t n_bases, int16_t* counts,
                        uint64_t* hashes, void* stream) {
    if (!valid_storage(st) || !ws) return fail(GT_ERR_INVALID_ARG, "invalid storage or workspace");
    GT_TRY(check_read_args(K, mode, seq, offsets, n_reads, n_bases));
    GT_TRY(ensure_device());
    if (n_reads == 0 || n_bases == 0) return GT_OK;
    std::lock_guard<std::recursive_mutex> lw(ws->mu);
    StorageLock lk(st);
    GT_TRY(lk.rc);
    GT_TRY(refuse_bigcount(st, "a query with outside scratch"));
    GT_TRY(order_after_fence(st, (hipStream_t)stream));
    hipStream_t s = (hipStream_t)stream;
    Reads R = make_reads(seq, offsets, n_reads, n_bases, K, mode, canonical);
    GT_TRY(prepare_reads(ws->sc, R, true, s));
    {
        ProfScope ps("query", s);
        launch_query(make_launch(st, s), R, st->tables_view(), st->exact_or_empty(), counts, hashes);
    }
    GT_HIP(hipGetLastError());
    return GT_OK;
}

// ------------------------------------------------------------ timing

int gt_profile_enable(int on) {
    g_prof.on.store(on != 0);
    return GT_OK;
}

int gt_profile_reset(void) {
    prof_resolve();
    std::lock_guard<std::mutex> lk(g_prof.mu);
    g_prof.totals.clear();
    return GT_OK;
}

int gt_profile_get(const char* kernel, uint64_t* launches, double* total_ms) {
    if (!kernel || !launches || !total_ms) return fail(GT_ERR_INVALID_ARG, "null argument");
    prof_resolve();
    std::lock_guard<std::mutex> lk(g_prof.mu);
    *launches = 0;
    *total_ms = 0.0;
    for (auto& t : g_prof.totals) {
        if (t.first == kernel) {
            *launches = t.second.first;
            *total_ms = t.second.second;
        }
    }
    return GT_OK;
}

// ------------------------------------------------------------ workload

namespace {

uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

struct Xoshiro256ss {
    uint64_t s[4];
    explicit Xoshiro256ss(uint64_t seed) {
        uint64_t x = seed;
        for (auto& v : s) v = splitmix64(x);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t result = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return result;
    }
    // Uniform in [0, n) by multiply-shift (n >= 1).
    uint64_t below(uint64_t n) { return (uint64_t)(((unsigned __int128)next() * n) >> 64); }
};

uint64_t rate_threshold(double p) {
    if (!(p > 0.0)) return 0;
    if (p >= 1.0) return ~0ULL;
    return (uint64_t)std::ldexp(p, 64);
}

const char kBases[4] = {'A', 'C', 'G', 'T'};

int base_index(uint8_t c) { return c == 'A' ? 0 : c == 'C' ? 1 : c == 'G' ? 2 : 3; }

uint64_t read_seed(uint64_t seed, uint64_t r) { return seed * 0x9E3779B97F4A7C15ULL + (r + 1) * 0xD1B54A32D192ED03ULL; }

}  // namespace

int gt_gen_genome(uint64_t seed, uint64_t length, uint8_t* out) {
    if (length && !out) return fail(GT_ERR_INVALID_ARG, "null output");
    Xoshiro256ss rng(seed);
    for (uint64_t i = 0; i < length; ++i) out[i] = (uint8_t)kBases[rng.next() >> 62];
    return GT_OK;
}

int gt_gen_reads(uint64_t seed, const uint8_t* genome, uint64_t genome_len, uint64_t n_reads, const uint32_t* lengths,
                 uint32_t len_min, uint32_t len_max, double sub_rate, double n_rate, uint8_t* seq, int64_t* offsets,
                 uint64_t* n_bases) {
    if (!genome || genome_len == 0 || !offsets || (n_reads && !seq))
        return fail(GT_ERR_INVALID_ARG, "null generator arguments");
    if (!lengths && len_min > len_max) return fail(GT_ERR_INVALID_ARG, "len_min > len_max");
    // Pass 1: lengths (each read's generator is private, so this is cheap).
    offsets[0] = 0;
    for (uint64_t r = 0; r < n_reads; ++r) {
        uint64_t len;
        Xoshiro256ss rng(read_seed(seed, r));
        if (lengths) {
            len = lengths[r];
        } else {
            len = len_min + rng.below((uint64_t)len_max - len_min + 1);
        }
        if (len > genome_len) len = genome_len;
        offsets[r + 1] = offsets[r] + (int64_t)len;
    }
    const uint64_t sub_thr = rate_threshold(sub_rate), n_thr = rate_threshold(n_rate);
    auto fill = [&](uint64_t r0, uint64_t r1) {
        for (uint64_t r = r0; r < r1; ++r) {
            Xoshiro256ss rng(read_seed(seed, r));
            if (!lengths) (void)rng.next();   // the length draw of pass 1
            const uint64_t len = (uint64_t)(offsets[r + 1] - offsets[r]);
            const uint64_t start = rng.below(genome_len - len + 1);
            const bool rev = rng.next() >> 63;
            uint8_t* out = seq + offsets[r];
            for (uint64_t j = 0; j < len; ++j) {
                uint8_t b = rev ? (uint8_t)gt_complement(genome[start + len - 1 - j]) : genome[start + j];
                if (rng.next() < sub_thr) b = (uint8_t)kBases[(base_index(b) + 1 + rng.below(3)) & 3];
                if (n_thr && rng.next() < n_thr) b = 'N';
                out[j] = b;
            }
        }
    };
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n_reads < 4096) nt = 1;
    std::vector<std::thread> th;
    const uint64_t per = (n_reads + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        const uint64_t a = t * per, b = std::min(n_reads, a + per);
        if (a < b) th.emplace_back(fill, a, b);
    }
    for (auto& t : th) t.join();
    if (n_bases) *n_bases = (uint64_t)offsets[n_reads];
    return GT_OK;
}

}  // extern "C"
