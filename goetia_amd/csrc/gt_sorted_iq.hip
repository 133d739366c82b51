// gt_sorted_iq.hip -- count-min insert_and_query in goetia's sequential order
// (ByteStorage::insert / query / insert_and_query, bytestorage.cc:60-150) for
// any number of tables and any table size, by sorting instead of replaying.
//
// A k-mer's result is the minimum over tables of its bin's count just after
// its own insert.  The count just before it is min(255, c0 + r): c0 the bin's
// count before the batch and r the number of earlier events of the batch on
// the same (table, bin).  So every (window, table) event becomes a key
// table << 32 | bin with its window index as value; a stable radix sort puts
// each bin's events together in window order, a max-scan of run starts gives
// r, every event reads c0 and writes its count just before, and then the
// last event of each run stores one more (at most 255) in the table.  A
// k-mer is new exactly when some table was 0 just before it, table 0's bin
// became occupied exactly when table 0 was 0, and every table was full (the
// bigcount condition, bytestorage.cc:95-96) when the smallest was 255.  Windows go through in
// chunks, each applied before the next, which keeps the order.
//
// Used where the binned replay does not apply: more than four tables, tables
// over 2^30 bins, per-hash calls.

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <string>

#include "goetia_amd.h"
#include "gt_hash.h"
#include "gt_internal.h"

namespace gt {

int set_error(int code, const std::string& msg);

namespace {

constexpr uint64_t kSentinel = ~0ULL;
constexpr int kSqBlock = 256;

__device__ __forceinline__ uint64_t window_value(const uint64_t* fw, const uint64_t* rc, uint64_t k, int canonical) {
    if (!rc) return fw[k];
    return canonical && rc[k] < fw[k] ? rc[k] : fw[k];
}

__global__ __launch_bounds__(kSqBlock) void sq_keys_kernel(const uint64_t* fw, const uint64_t* rc, const uint8_t* valid,
                                                           uint64_t w0, uint32_t m, int canonical, Tables T,
                                                           uint64_t* keys, uint32_t* vals) {
    const uint32_t nt = T.n;
    for (uint64_t i = blockIdx.x * (uint64_t)kSqBlock + threadIdx.x; i < m; i += (uint64_t)gridDim.x * kSqBlock) {
        const uint64_t k = w0 + i;
        const bool ok = !valid || valid[k];
        const uint64_t v = ok ? window_value(fw, rc, k, canonical) : 0;
#pragma unroll
        for (uint32_t t = 0; t < (uint32_t)kMaxTables; ++t) {   // constant indexes keep T in registers
            if (t >= nt) break;
            keys[i * nt + t] = ok ? ((uint64_t)t << 32) | gt_fastmod(v, T.p[t], T.m[t]) : kSentinel;
            vals[i * nt + t] = (uint32_t)i;
        }
    }
}

// start[j] = j where a run of equal keys begins, else 0 (a max-scan then gives
// every event its run's first index).
__global__ __launch_bounds__(kSqBlock) void sq_starts_kernel(const uint64_t* keys, uint32_t n, uint32_t* start) {
    for (uint64_t j = blockIdx.x * (uint64_t)kSqBlock + threadIdx.x; j < n; j += (uint64_t)gridDim.x * kSqBlock)
        start[j] = (j == 0 || keys[j] != keys[j - 1]) ? (uint32_t)j : 0u;
}

__global__ __launch_bounds__(kSqBlock) void sq_post_kernel(const uint64_t* keys, const uint32_t* vals,
                                                           const uint32_t* start, uint32_t n, Tables T, uint8_t* vq) {
    for (uint64_t j = blockIdx.x * (uint64_t)kSqBlock + threadIdx.x; j < n; j += (uint64_t)gridDim.x * kSqBlock) {
        const uint64_t key = keys[j];
        if (key == kSentinel) continue;
        const uint32_t t = (uint32_t)(key >> 32), bin = (uint32_t)key;
        const uint32_t c = (uint32_t)T.base[t][bin] + (uint32_t)(j - start[j]);   // count just before
        vq[(uint64_t)vals[j] * T.n + t] = (uint8_t)(c < 255u ? c : 255u);
    }
}

__global__ __launch_bounds__(kSqBlock) void sq_store_kernel(const uint64_t* keys, const uint32_t* vals, uint32_t n,
                                                            Tables T, const uint8_t* vq) {
    for (uint64_t j = blockIdx.x * (uint64_t)kSqBlock + threadIdx.x; j < n; j += (uint64_t)gridDim.x * kSqBlock) {
        const uint64_t key = keys[j];
        if (key == kSentinel || (j + 1 < n && keys[j + 1] == key)) continue;   // last event of its run
        const uint32_t t = (uint32_t)(key >> 32);
        const uint32_t before = vq[(uint64_t)vals[j] * T.n + t];
        T.base[t][(uint32_t)key] = (uint8_t)(before < 255u ? before + 1u : 255u);
    }
}

// Per window, from its tables' counts just before it: the minimum just
// after it (its insert_and_query result), whether it is new (a table was 0),
// whether table 0's bin became occupied, and whether every table was full.
__global__ __launch_bounds__(kSqBlock) void sq_combine_kernel(const uint8_t* valid, uint64_t w0, uint32_t m,
                                                              uint32_t nt, const uint8_t* vq, int16_t* counts,
                                                              uint8_t* is_new, uint8_t* full, uint64_t* counters) {
    uint64_t nn = 0, occ = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)kSqBlock + threadIdx.x; i < m; i += (uint64_t)gridDim.x * kSqBlock) {
        const uint64_t k = w0 + i;
        if (valid && !valid[k]) {
            if (counts) counts[k] = -1;
            if (is_new) is_new[k] = 0;
            if (full) full[k] = 0;
            continue;
        }
        uint32_t lo = 255;
        for (uint32_t t = 0; t < nt; ++t) {
            const uint32_t c = vq[i * nt + t];
            lo = c < lo ? c : lo;
        }
        const uint32_t after = lo < 255u ? lo + 1u : 255u;
        nn += lo == 0;
        occ += vq[i * nt] == 0;
        if (counts) counts[k] = (int16_t)after;
        if (is_new) is_new[k] = lo == 0;
        if (full) full[k] = lo == 255u;
    }
    for (int o = 32; o > 0; o >>= 1) {
        nn += __shfl_xor(nn, o, 64);
        occ += __shfl_xor(occ, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        if (nn) atomicAdd((unsigned long long*)(counters + kNUnique), (unsigned long long)nn);
        if (occ) atomicAdd((unsigned long long*)(counters + kNOccupied), (unsigned long long)occ);
    }
}

struct Buf {
    void* p = nullptr;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t n) {
        if (hipMalloc(&p, n ? n : 1) != hipSuccess) {
            p = nullptr;
            return set_error(GT_ERR_OUT_OF_MEMORY, "hipMalloc of " + std::to_string(n) + " bytes failed");
        }
        return GT_OK;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

int hip_check(hipError_t e, const char* what) {
    return e == hipSuccess ? GT_OK : set_error(GT_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define SQ_TRY(x)                        \
    do {                                 \
        const int rc_ = (x);             \
        if (rc_ != GT_OK) return rc_;    \
    } while (0)

unsigned grid_for(uint64_t n, int cus) {
    const uint64_t want = (n + kSqBlock - 1) / kSqBlock;
    const uint64_t cap = (uint64_t)cus * 16;
    return (unsigned)(want < 1 ? 1 : (want < cap ? want : cap));
}

}  // namespace

int sorted_insert_query(const Tables& T, const uint64_t* fw, const uint64_t* rc, const uint8_t* valid, uint64_t n,
                        int canonical, int16_t* counts, uint8_t* is_new, uint64_t* counters, int cus,
                        hipStream_t s, uint8_t* full) {
    if (n == 0) return GT_OK;
    if (T.n == 0 || T.n > (uint32_t)kMaxTables) return set_error(GT_ERR_INVALID_ARG, "1..16 tables");
    const uint32_t nt = T.n;
    // windows per chunk: at most 2^26 events
    const uint64_t chunk = std::max<uint64_t>(1, (1ULL << 26) / nt);
    const uint64_t m_max = std::min<uint64_t>(chunk, n);
    const uint64_t pairs = m_max * nt;
    // Sort on the bits a key can use: table < nt above a 32-bit bin.  Bins are
    // below 2^32 - 1, so the sentinel's low end_bit bits (all ones) exceed every
    // real key and sentinels sort last.
    const int end_bit = 32 + (nt > 1 ? 32 - __builtin_clz(nt - 1) : 0);
    Buf keys_in, keys_out, vals_in, vals_out, starts, runs, vq, tmp;
    SQ_TRY(keys_in.alloc(pairs * 8));
    SQ_TRY(keys_out.alloc(pairs * 8));
    SQ_TRY(vals_in.alloc(pairs * 4));
    SQ_TRY(vals_out.alloc(pairs * 4));
    SQ_TRY(starts.alloc(pairs * 4));
    SQ_TRY(runs.alloc(pairs * 4));
    SQ_TRY(vq.alloc(pairs));
    size_t sort_bytes = 0, scan_bytes = 0;
    SQ_TRY(hip_check(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, keys_in.as<uint64_t>(),
                                                        keys_out.as<uint64_t>(), vals_in.as<uint32_t>(),
                                                        vals_out.as<uint32_t>(), (int)pairs, 0, end_bit, s),
                     "radix sort sizing"));
    SQ_TRY(hip_check(hipcub::DeviceScan::InclusiveScan(nullptr, scan_bytes, starts.as<uint32_t>(), runs.as<uint32_t>(),
                                                       hipcub::Max(), (int)pairs, s),
                     "scan sizing"));
    SQ_TRY(tmp.alloc(std::max(sort_bytes, scan_bytes)));
    for (uint64_t w0 = 0; w0 < n; w0 += chunk) {
        const uint32_t m = (uint32_t)std::min<uint64_t>(chunk, n - w0);
        const uint32_t np = m * nt;
        sq_keys_kernel<<<grid_for(m, cus), kSqBlock, 0, s>>>(fw, rc, valid, w0, m, canonical, T, keys_in.as<uint64_t>(),
                                                             vals_in.as<uint32_t>());
        size_t b = sort_bytes;
        SQ_TRY(hip_check(hipcub::DeviceRadixSort::SortPairs(tmp.p, b, keys_in.as<uint64_t>(), keys_out.as<uint64_t>(),
                                                            vals_in.as<uint32_t>(), vals_out.as<uint32_t>(), (int)np,
                                                            0, end_bit, s),
                         "radix sort"));
        sq_starts_kernel<<<grid_for(np, cus), kSqBlock, 0, s>>>(keys_out.as<uint64_t>(), np, starts.as<uint32_t>());
        b = scan_bytes;
        SQ_TRY(hip_check(hipcub::DeviceScan::InclusiveScan(tmp.p, b, starts.as<uint32_t>(), runs.as<uint32_t>(),
                                                           hipcub::Max(), (int)np, s),
                         "scan"));
        sq_post_kernel<<<grid_for(np, cus), kSqBlock, 0, s>>>(keys_out.as<uint64_t>(), vals_out.as<uint32_t>(),
                                                              runs.as<uint32_t>(), np, T, vq.as<uint8_t>());
        sq_store_kernel<<<grid_for(np, cus), kSqBlock, 0, s>>>(keys_out.as<uint64_t>(), vals_out.as<uint32_t>(), np, T,
                                                               vq.as<uint8_t>());
        sq_combine_kernel<<<grid_for(m, cus), kSqBlock, 0, s>>>(valid, w0, m, nt, vq.as<uint8_t>(), counts, is_new,
                                                                full, counters);
        SQ_TRY(hip_check(hipGetLastError(), "sorted insert_and_query launch"));
    }
    // the buffers are released on return
    return hip_check(hipStreamSynchronize(s), "sorted insert_and_query");
}

}  // namespace gt
