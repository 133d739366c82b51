// gt_bigcount.hip -- the count-min bigcount map (bytestorage.cc:95-106,
// 130-137; storage.hh:129-135) on the device.
//
// goetia keeps, when bigcount is on, a map from k-mer hash to an int16 count
// for k-mers inserted while every table already held 255.  Such an insert
// makes the entry (0 when absent) 256 if it was 0, else one more while the
// int16 value, compared as an unsigned int, is below 65535: 1..32766 step
// up, 32767 wraps to -32768, and a negative value stays.  A query whose
// table minimum is 255 returns the entry when there is one.
//
// Here the map is an open-addressing table of 64-bit keys (linear probing
// from big_mix(key); ~0 marks a free slot and the all-ones key has its own
// slot at index cap) with the int16 value in the low half of a u32.  A batch
// first counts, per key, the windows that found every table full (their
// order does not matter: n such inserts take a value v to big_step(v, n)),
// then applies the counts in one pass.  insert_and_query needs each window's
// value just after its own insert, which depends on the full windows of the
// same k-mer before it: those windows are sorted by k-mer (stably, so in
// window order) and counted per run.

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <string>
#include <vector>

#include "goetia_amd.h"
#include "gt_hash.h"
#include "gt_internal.h"
#include "gt_launch.h"

namespace gt {

int set_error(int code, const std::string& msg);

GT_HD uint64_t big_mix(uint64_t k) {
    k ^= k >> 31;
    k *= 0x7FB5D329728EA185ULL;
    k ^= k >> 27;
    k *= 0x81DADEF4BC2DD44DULL;
    k ^= k >> 33;
    return k;
}

// n inserts that find every table full, from value v (see above).
GT_HD int16_t big_step(int16_t v, uint64_t n) {
    if (n == 0 || v < 0) return v;
    if (v == 0) {
        v = 256;
        if (--n == 0) return v;
    }
    const uint64_t room = (uint64_t)(32767 - v);
    return n <= room ? (int16_t)(v + (int16_t)n) : (int16_t)-32768;
}

namespace {

constexpr uint64_t kEmpty = ~0ULL;
constexpr int kBigBlock = 256;

__device__ __forceinline__ uint64_t value_of(const uint64_t* fw, const uint64_t* rc, uint64_t i, int canonical) {
    if (!rc) return fw[i];
    return canonical && rc[i] < fw[i] ? rc[i] : fw[i];
}

// Slot of key h, or -1 when absent.
__device__ __forceinline__ int64_t big_find(const BigMapDev& M, uint64_t h) {
    if (h == kEmpty) return M.aux[kBigOnes] ? (int64_t)M.cap : -1;
    if (M.cap == 0) return -1;
    const uint64_t mask = M.cap - 1;
    uint64_t s = big_mix(h) & mask;
    for (uint64_t p = 0; p < M.cap; ++p) {
        const uint64_t k = M.keys[s];
        if (k == h) return (int64_t)s;
        if (k == kEmpty) return -1;
        s = (s + 1) & mask;
    }
    return -1;
}

__global__ __launch_bounds__(kBigBlock) void big_count_kernel(const uint8_t* full, uint64_t n, unsigned long long* out) {
    unsigned long long c = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)kBigBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBigBlock)
        c += full[i] != 0;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(out, c);
}

// Adds one pending insert per full window to its key, making the key when
// absent.  The table has room for every window (reserved by the caller).
__global__ __launch_bounds__(kBigBlock) void big_add_kernel(const uint64_t* fw, const uint64_t* rc, int canonical,
                                                            const uint8_t* full, uint64_t n, BigMapDev M) {
    const uint64_t mask = M.cap - 1;
    for (uint64_t i = blockIdx.x * (uint64_t)kBigBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBigBlock) {
        if (!full[i]) continue;
        const uint64_t h = value_of(fw, rc, i, canonical);
        if (h == kEmpty) {
            atomicOr(&M.aux[kBigOnes], 1u);
            atomicAdd(&M.pend[M.cap], 1u);
            continue;
        }
        uint64_t s = big_mix(h) & mask;
        for (uint64_t p = 0; p < M.cap; ++p) {
            uint64_t k = M.keys[s];
            if (k == kEmpty) {
                k = atomicCAS((unsigned long long*)&M.keys[s], (unsigned long long)kEmpty, (unsigned long long)h);
                if (k == kEmpty) {
                    atomicAdd(&M.aux[kBigKeys], 1u);
                    k = h;
                }
            }
            if (k == h) {
                atomicAdd(&M.pend[s], 1u);
                break;
            }
            s = (s + 1) & mask;
        }
    }
}

// Applies each key's pending inserts to its value.
__global__ __launch_bounds__(kBigBlock) void big_settle_kernel(BigMapDev M) {
    for (uint64_t s = blockIdx.x * (uint64_t)kBigBlock + threadIdx.x; s <= M.cap; s += (uint64_t)gridDim.x * kBigBlock) {
        const uint32_t p = M.pend[s];
        if (!p) continue;
        M.vals[s] = (uint16_t)big_step((int16_t)(uint16_t)M.vals[s], p);
        M.pend[s] = 0;
    }
}

__global__ __launch_bounds__(kBigBlock) void big_rehash_kernel(const uint64_t* keys, const uint32_t* vals, uint64_t old_cap,
                                                               BigMapDev M) {
    const uint64_t mask = M.cap - 1;
    for (uint64_t j = blockIdx.x * (uint64_t)kBigBlock + threadIdx.x; j < old_cap; j += (uint64_t)gridDim.x * kBigBlock) {
        const uint64_t h = keys[j];
        if (h == kEmpty) continue;
        uint64_t s = big_mix(h) & mask;
        while (atomicCAS((unsigned long long*)&M.keys[s], (unsigned long long)kEmpty, (unsigned long long)h) !=
               (unsigned long long)kEmpty)
            s = (s + 1) & mask;
        M.vals[s] = vals[j];
    }
}

__global__ __launch_bounds__(kBigBlock) void big_fill_kernel(BigMapDev M) {
    for (uint64_t s = blockIdx.x * (uint64_t)kBigBlock + threadIdx.x; s <= M.cap; s += (uint64_t)gridDim.x * kBigBlock) {
        if (s < M.cap) M.keys[s] = kEmpty;
        M.vals[s] = 0;
        M.pend[s] = 0;
    }
}

// ByteStorage::query's bigcount step: a window whose count is 255 takes its
// k-mer's entry when there is one.
__global__ __launch_bounds__(kBigBlock) void big_query_kernel(const uint64_t* fw, const uint64_t* rc, const uint8_t* valid,
                                                              int canonical, uint64_t n, int16_t* counts, BigMapDev M) {
    for (uint64_t i = blockIdx.x * (uint64_t)kBigBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBigBlock) {
        if (counts[i] != 255 || (valid && !valid[i])) continue;
        const int64_t s = big_find(M, value_of(fw, rc, i, canonical));
        if (s >= 0) counts[i] = (int16_t)(uint16_t)M.vals[s];
    }
}

__global__ __launch_bounds__(kBigBlock) void big_mark_kernel(const int16_t* counts, const uint8_t* valid, uint64_t n,
                                                             uint8_t* sel) {
    for (uint64_t i = blockIdx.x * (uint64_t)kBigBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBigBlock)
        sel[i] = counts[i] == 255 && !(valid && !valid[i]);
}

__global__ __launch_bounds__(kBigBlock) void big_keys_kernel(const uint64_t* fw, const uint64_t* rc, int canonical,
                                                             const uint32_t* idx, uint32_t m, uint64_t* keys) {
    for (uint64_t j = blockIdx.x * (uint64_t)kBigBlock + threadIdx.x; j < m; j += (uint64_t)gridDim.x * kBigBlock)
        keys[j] = value_of(fw, rc, idx[j], canonical);
}

// Per sorted entry: its full flag and, where a run of equal keys starts, its
// index (else 0; a max-scan then gives every entry its run's start).
__global__ __launch_bounds__(kBigBlock) void big_runs_kernel(const uint64_t* keys, const uint32_t* idx, uint32_t m,
                                                             const uint8_t* full, uint32_t* f, uint32_t* start) {
    for (uint64_t j = blockIdx.x * (uint64_t)kBigBlock + threadIdx.x; j < m; j += (uint64_t)gridDim.x * kBigBlock) {
        f[j] = full[idx[j]] ? 1u : 0u;
        start[j] = (j == 0 || keys[j] != keys[j - 1]) ? (uint32_t)j : 0u;
    }
}

// insert_and_query's value for a window whose count after its insert is
// 255: the entry after the full windows of its k-mer up to and including
// it (csum - the sum before its run), or 255 when there is no entry.
__global__ __launch_bounds__(kBigBlock) void big_iq_kernel(const uint64_t* keys, const uint32_t* idx, uint32_t m,
                                                           const uint32_t* csum, const uint32_t* run0, BigMapDev M,
                                                           int16_t* counts) {
    for (uint64_t j = blockIdx.x * (uint64_t)kBigBlock + threadIdx.x; j < m; j += (uint64_t)gridDim.x * kBigBlock) {
        const uint32_t r = run0[j];
        const uint32_t before = r ? csum[r - 1] : 0u;
        const uint32_t k = csum[j] - before;
        const int64_t s = big_find(M, keys[j]);
        if (s < 0 && k == 0) continue;                   // no entry: the table's 255 stands
        const int16_t v0 = s >= 0 ? (int16_t)(uint16_t)M.vals[s] : (int16_t)0;
        counts[idx[j]] = big_step(v0, k);
    }
}

unsigned grid_for(uint64_t n, int cus) {
    const uint64_t want = (n + kBigBlock - 1) / kBigBlock;
    const uint64_t cap = (uint64_t)std::max(cus, 1) * 16;
    return (unsigned)std::max<uint64_t>(1, std::min(want, cap));
}

int hip_check(hipError_t e, const char* what) {
    return e == hipSuccess ? GT_OK : set_error(GT_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define BIG_TRY(x)                       \
    do {                                 \
        const int rc_ = (x);             \
        if (rc_ != GT_OK) return rc_;    \
    } while (0)

struct Tmp {
    void* p = nullptr;
    ~Tmp() {
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

// Allocates an empty map of `cap` slots (a power of two) into M.
int big_alloc(BigMapDev& M, uint64_t cap, hipStream_t s) {
    void* k = nullptr;
    void* v = nullptr;
    void* p = nullptr;
    if (hipMalloc(&k, cap * 8) != hipSuccess || hipMalloc(&v, (cap + 1) * 4) != hipSuccess ||
        hipMalloc(&p, (cap + 1) * 4) != hipSuccess) {
        if (k) (void)hipFree(k);
        if (v) (void)hipFree(v);
        if (p) (void)hipFree(p);
        return set_error(GT_ERR_OUT_OF_MEMORY, "bigcount map of " + std::to_string(cap) + " slots");
    }
    M.keys = static_cast<uint64_t*>(k);
    M.vals = static_cast<uint32_t*>(v);
    M.pend = static_cast<uint32_t*>(p);
    M.cap = cap;
    big_fill_kernel<<<grid_for(cap + 1, 64), kBigBlock, 0, s>>>(M);
    return hip_check(hipGetLastError(), "bigcount map fill");
}

int read_aux(const BigMapDev& M, uint32_t* host, hipStream_t s) {
    BIG_TRY(hip_check(hipMemcpyAsync(host, M.aux, kBigAux * sizeof(uint32_t), hipMemcpyDeviceToHost, s), "bigcount aux"));
    return hip_check(hipStreamSynchronize(s), "bigcount aux");
}

}  // namespace

int big_init(BigMapDev& M, hipStream_t s) {
    if (M.aux) return GT_OK;
    void* a = nullptr;
    if (hipMalloc(&a, kBigAux * sizeof(uint32_t)) != hipSuccess)
        return set_error(GT_ERR_OUT_OF_MEMORY, "bigcount map counters");
    M.aux = static_cast<uint32_t*>(a);
    BIG_TRY(hip_check(hipMemsetAsync(M.aux, 0, kBigAux * sizeof(uint32_t), s), "bigcount map counters"));
    return big_alloc(M, 64, s);
}

void big_free(BigMapDev& M) {
    if (M.keys) (void)hipFree(M.keys);
    if (M.vals) (void)hipFree(M.vals);
    if (M.pend) (void)hipFree(M.pend);
    if (M.aux) (void)hipFree(M.aux);
    M = BigMapDev{};
}

// Grows the map (rehashing) so that `more` new keys keep it at most half full.
int big_reserve(BigMapDev& M, uint64_t more, hipStream_t s) {
    BIG_TRY(big_init(M, s));
    uint32_t aux[kBigAux];
    BIG_TRY(read_aux(M, aux, s));
    const uint64_t want = 2 * ((uint64_t)aux[kBigKeys] + more);
    if (want <= M.cap) return GT_OK;
    uint64_t cap = M.cap;
    while (cap < want) cap *= 2;
    if (cap > (1ULL << 31)) return set_error(GT_ERR_OUT_OF_MEMORY, "bigcount map over 2^31 slots");
    BigMapDev old = M;
    BIG_TRY(big_alloc(M, cap, s));
    // the all-ones key keeps its slot at index cap
    BIG_TRY(hip_check(hipMemcpyAsync(M.vals + cap, old.vals + old.cap, 4, hipMemcpyDeviceToDevice, s), "bigcount rehash"));
    big_rehash_kernel<<<grid_for(old.cap, 64), kBigBlock, 0, s>>>(old.keys, old.vals, old.cap, M);
    BIG_TRY(hip_check(hipGetLastError(), "bigcount rehash"));
    BIG_TRY(hip_check(hipStreamSynchronize(s), "bigcount rehash"));
    (void)hipFree(old.keys);
    (void)hipFree(old.vals);
    (void)hipFree(old.pend);
    return GT_OK;
}

// The bigcount step of a batch of inserts: every window flagged full adds
// one to its k-mer's entry.  Window values are fw (or the canonical of fw
// and rc when rc is given).
int big_insert(BigMapDev& M, const uint64_t* fw, const uint64_t* rc, int canonical, const uint8_t* full, uint64_t n,
               int cus, hipStream_t s) {
    if (n == 0) return GT_OK;
    BIG_TRY(big_init(M, s));
    Tmp cnt;
    BIG_TRY(cnt.alloc(8));
    BIG_TRY(hip_check(hipMemsetAsync(cnt.p, 0, 8, s), "bigcount count"));
    big_count_kernel<<<grid_for(n, cus), kBigBlock, 0, s>>>(full, n, cnt.as<unsigned long long>());
    unsigned long long n_full = 0;
    BIG_TRY(hip_check(hipMemcpyAsync(&n_full, cnt.p, 8, hipMemcpyDeviceToHost, s), "bigcount count"));
    BIG_TRY(hip_check(hipStreamSynchronize(s), "bigcount count"));
    if (n_full == 0) return GT_OK;
    BIG_TRY(big_reserve(M, n_full, s));
    big_add_kernel<<<grid_for(n, cus), kBigBlock, 0, s>>>(fw, rc, canonical, full, n, M);
    big_settle_kernel<<<grid_for(M.cap + 1, cus), kBigBlock, 0, s>>>(M);
    BIG_TRY(hip_check(hipGetLastError(), "bigcount insert"));
    return hip_check(hipStreamSynchronize(s), "bigcount insert");
}

// The bigcount step of queries: counts of 255 take their k-mer's entry.
int big_query(const BigMapDev& M, const uint64_t* fw, const uint64_t* rc, const uint8_t* valid, int canonical,
              int16_t* counts, uint64_t n, int cus, hipStream_t s) {
    if (n == 0 || !M.aux) return GT_OK;
    big_query_kernel<<<grid_for(n, cus), kBigBlock, 0, s>>>(fw, rc, valid, canonical, n, counts, M);
    return hip_check(hipGetLastError(), "bigcount query");
}

// insert_and_query's bigcount step, run before big_insert applies the
// batch: counts (the minimum over tables just after each window's insert)
// of 255 become the k-mer's entry just after that insert.
int big_insert_query(const BigMapDev& M, const uint64_t* fw, const uint64_t* rc, const uint8_t* valid, int canonical,
                     const uint8_t* full, int16_t* counts, uint64_t n, int cus, hipStream_t s) {
    if (n == 0) return GT_OK;
    if (n > 0x7FFFFFFFULL)   // the selection, sort and scans take int counts
        return set_error(GT_ERR_INVALID_ARG, "bigcount insert_and_query: more than 2^31 - 1 windows in one call");
    Tmp sel, idx, nsel;
    BIG_TRY(sel.alloc(n));
    BIG_TRY(idx.alloc(n * 4));
    BIG_TRY(nsel.alloc(8));
    big_mark_kernel<<<grid_for(n, cus), kBigBlock, 0, s>>>(counts, valid, n, sel.as<uint8_t>());
    hipcub::CountingInputIterator<uint32_t> it(0);
    size_t tb = 0;
    BIG_TRY(hip_check(hipcub::DeviceSelect::Flagged(nullptr, tb, it, sel.as<uint8_t>(), idx.as<uint32_t>(),
                                                    nsel.as<uint32_t>(), (int)n, s), "bigcount select sizing"));
    Tmp t0;
    BIG_TRY(t0.alloc(tb));
    BIG_TRY(hip_check(hipcub::DeviceSelect::Flagged(t0.p, tb, it, sel.as<uint8_t>(), idx.as<uint32_t>(),
                                                    nsel.as<uint32_t>(), (int)n, s), "bigcount select"));
    uint32_t m = 0;
    BIG_TRY(hip_check(hipMemcpyAsync(&m, nsel.p, 4, hipMemcpyDeviceToHost, s), "bigcount select"));
    BIG_TRY(hip_check(hipStreamSynchronize(s), "bigcount select"));
    if (m == 0) return GT_OK;
    Tmp keys_in, keys_out, idx_out, f, start, run0, csum;
    BIG_TRY(keys_in.alloc((size_t)m * 8));
    BIG_TRY(keys_out.alloc((size_t)m * 8));
    BIG_TRY(idx_out.alloc((size_t)m * 4));
    BIG_TRY(f.alloc((size_t)m * 4));
    BIG_TRY(start.alloc((size_t)m * 4));
    BIG_TRY(run0.alloc((size_t)m * 4));
    BIG_TRY(csum.alloc((size_t)m * 4));
    big_keys_kernel<<<grid_for(m, cus), kBigBlock, 0, s>>>(fw, rc, canonical, idx.as<uint32_t>(), m,
                                                           keys_in.as<uint64_t>());
    size_t sb = 0, s1 = 0, s2 = 0;
    BIG_TRY(hip_check(hipcub::DeviceRadixSort::SortPairs(nullptr, sb, keys_in.as<uint64_t>(), keys_out.as<uint64_t>(),
                                                        idx.as<uint32_t>(), idx_out.as<uint32_t>(), (int)m, 0, 64, s),
                      "bigcount sort sizing"));
    BIG_TRY(hip_check(hipcub::DeviceScan::InclusiveScan(nullptr, s1, start.as<uint32_t>(), run0.as<uint32_t>(),
                                                       hipcub::Max(), (int)m, s), "bigcount scan sizing"));
    BIG_TRY(hip_check(hipcub::DeviceScan::InclusiveSum(nullptr, s2, f.as<uint32_t>(), csum.as<uint32_t>(), (int)m, s),
                      "bigcount sum sizing"));
    Tmp t1;
    BIG_TRY(t1.alloc(std::max(sb, std::max(s1, s2))));
    BIG_TRY(hip_check(hipcub::DeviceRadixSort::SortPairs(t1.p, sb, keys_in.as<uint64_t>(), keys_out.as<uint64_t>(),
                                                        idx.as<uint32_t>(), idx_out.as<uint32_t>(), (int)m, 0, 64, s),
                      "bigcount sort"));
    big_runs_kernel<<<grid_for(m, cus), kBigBlock, 0, s>>>(keys_out.as<uint64_t>(), idx_out.as<uint32_t>(), m, full,
                                                           f.as<uint32_t>(), start.as<uint32_t>());
    BIG_TRY(hip_check(hipcub::DeviceScan::InclusiveScan(t1.p, s1, start.as<uint32_t>(), run0.as<uint32_t>(),
                                                       hipcub::Max(), (int)m, s), "bigcount scan"));
    BIG_TRY(hip_check(hipcub::DeviceScan::InclusiveSum(t1.p, s2, f.as<uint32_t>(), csum.as<uint32_t>(), (int)m, s),
                      "bigcount sum"));
    BigMapDev view = M;
    big_iq_kernel<<<grid_for(m, cus), kBigBlock, 0, s>>>(keys_out.as<uint64_t>(), idx_out.as<uint32_t>(), m,
                                                         csum.as<uint32_t>(), run0.as<uint32_t>(), view, counts);
    BIG_TRY(hip_check(hipGetLastError(), "bigcount insert_and_query"));
    return hip_check(hipStreamSynchronize(s), "bigcount insert_and_query");   // temporaries are released on return
}

// The map's entries, sorted by key (goetia writes its unordered_map's order).
int big_pairs(const BigMapDev& M, std::vector<uint64_t>& keys, std::vector<int16_t>& vals, hipStream_t s) {
    keys.clear();
    vals.clear();
    if (!M.aux) return GT_OK;
    uint32_t aux[kBigAux];
    BIG_TRY(read_aux(M, aux, s));
    std::vector<uint64_t> k(M.cap);
    std::vector<uint32_t> v(M.cap + 1);
    BIG_TRY(hip_check(hipMemcpyAsync(k.data(), M.keys, M.cap * 8, hipMemcpyDeviceToHost, s), "bigcount read"));
    BIG_TRY(hip_check(hipMemcpyAsync(v.data(), M.vals, (M.cap + 1) * 4, hipMemcpyDeviceToHost, s), "bigcount read"));
    BIG_TRY(hip_check(hipStreamSynchronize(s), "bigcount read"));
    std::vector<std::pair<uint64_t, int16_t>> p;
    for (uint64_t j = 0; j < M.cap; ++j)
        if (k[j] != kEmpty) p.emplace_back(k[j], (int16_t)(uint16_t)v[j]);
    if (aux[kBigOnes]) p.emplace_back(kEmpty, (int16_t)(uint16_t)v[M.cap]);
    std::sort(p.begin(), p.end());
    for (auto& e : p) {
        keys.push_back(e.first);
        vals.push_back(e.second);
    }
    return GT_OK;
}

// Replaces the map's entries (ByteStorage's loader, bytestorage.cc:271-285:
// later pairs of one key overwrite earlier ones).
int big_set_pairs(BigMapDev& M, const uint64_t* keys, const int16_t* vals, uint64_t n, hipStream_t s) {
    big_free(M);
    BIG_TRY(big_init(M, s));
    uint64_t cap = 64;
    while (cap < 2 * (n + 1)) cap *= 2;
    std::vector<uint64_t> k(cap, kEmpty);
    std::vector<uint32_t> v(cap + 1, 0);
    uint32_t aux[kBigAux] = {};
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t h = keys[i];
        const uint32_t val = (uint16_t)vals[i];
        if (h == kEmpty) {
            aux[kBigOnes] = 1;
            v[cap] = val;
            continue;
        }
        uint64_t j = big_mix(h) & (cap - 1);
        while (k[j] != kEmpty && k[j] != h) j = (j + 1) & (cap - 1);
        if (k[j] == kEmpty) ++aux[kBigKeys];
        k[j] = h;
        v[j] = val;
    }
    {
        // the empty map big_init made is replaced by one of `cap` slots
        BigMapDev old = M;
        BIG_TRY(big_alloc(M, cap, s));
        (void)hipFree(old.keys);
        (void)hipFree(old.vals);
        (void)hipFree(old.pend);
    }
    BIG_TRY(hip_check(hipMemcpyAsync(M.keys, k.data(), cap * 8, hipMemcpyHostToDevice, s), "bigcount load"));
    BIG_TRY(hip_check(hipMemcpyAsync(M.vals, v.data(), (cap + 1) * 4, hipMemcpyHostToDevice, s), "bigcount load"));
    BIG_TRY(hip_check(hipMemcpyAsync(M.aux, aux, kBigAux * 4, hipMemcpyHostToDevice, s), "bigcount load"));
    return hip_check(hipStreamSynchronize(s), "bigcount load");
}

}  // namespace gt
