// Random-access roofline of the device (no counterpart in goetia): uniform
// random 32-bit reads, atomicOr or atomicCAS over a caller's buffer from every
// CU.  The k-mer table kernels are bound by exactly these accesses (a random
// table byte per k-mer and table), so the rate measured here is the
// denominator their achieved rates are compared with.

#include "goetia_amd.h"

#include <hip/hip_runtime.h>

#include <string>

namespace gt {
int set_error(int code, const std::string& msg);
}

namespace {

constexpr int kMbBlock = 256;
constexpr int kMbUnroll = 8;

// splitmix64 step
__device__ __forceinline__ uint64_t mb_next(uint64_t& x) {
    x += 0x9E3779B97F4A7C15ULL;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Word index in [0, n_words) from the high bits of a random word.
__device__ __forceinline__ uint64_t mb_index(uint64_t r, uint64_t n_words) { return __umul64hi(r, n_words); }

// kNonTemporal: loads with the non-temporal cache policy (global_load ... nt).
template <bool kNonTemporal>
__global__ __launch_bounds__(kMbBlock) void mb_read_kernel(const uint32_t* buf, uint64_t n_words, uint64_t rounds,
                                                           uint64_t seed, uint32_t* sink) {
    uint64_t x = seed ^ ((blockIdx.x * (uint64_t)kMbBlock + threadIdx.x) * 0xD1B54A32D192ED03ULL);
    uint32_t acc = 0;
    for (uint64_t k = 0; k < rounds; ++k) {
        uint64_t idx[kMbUnroll];
#pragma unroll
        for (int u = 0; u < kMbUnroll; ++u) idx[u] = mb_index(mb_next(x), n_words);
        uint32_t v[kMbUnroll];
#pragma unroll
        for (int u = 0; u < kMbUnroll; ++u) {
            if constexpr (kNonTemporal)
                v[u] = __builtin_nontemporal_load(buf + idx[u]);
            else
                v[u] = buf[idx[u]];
        }
#pragma unroll
        for (int u = 0; u < kMbUnroll; ++u) acc += v[u];
    }
    if (acc == 0x9E3779B9u) sink[0] = acc;       // keeps the loads; practically never taken
}

__global__ __launch_bounds__(kMbBlock) void mb_or_kernel(uint32_t* buf, uint64_t n_words, uint64_t rounds,
                                                         uint64_t seed) {
    uint64_t x = seed ^ ((blockIdx.x * (uint64_t)kMbBlock + threadIdx.x) * 0xD1B54A32D192ED03ULL);
    for (uint64_t k = 0; k < rounds; ++k) {
#pragma unroll
        for (int u = 0; u < kMbUnroll; ++u) {
            const uint64_t r = mb_next(x);
            atomicOr(&buf[mb_index(r, n_words)], 1u << (r & 31));
        }
    }
}

__global__ __launch_bounds__(kMbBlock) void mb_cas_kernel(uint32_t* buf, uint64_t n_words, uint64_t rounds,
                                                          uint64_t seed) {
    uint64_t x = seed ^ ((blockIdx.x * (uint64_t)kMbBlock + threadIdx.x) * 0xD1B54A32D192ED03ULL);
    for (uint64_t k = 0; k < rounds; ++k) {
#pragma unroll
        for (int u = 0; u < kMbUnroll; ++u) {
            const uint64_t r = mb_next(x);
            uint32_t* p = &buf[mb_index(r, n_words)];
            uint32_t old = *p;
            // one saturating increment of the low byte, as the count-min tables do
            for (int tries = 0; tries < 64; ++tries) {
                const uint32_t lo = old & 0xFFu;
                if (lo == 0xFFu) break;
                const uint32_t seen = atomicCAS(p, old, old + 1u);
                if (seen == old) break;
                old = seen;
            }
        }
    }
}

}  // namespace

extern "C" int gt_membench_random(void* buf, uint64_t n_words, uint64_t n_ops, int kind, uint64_t seed,
                                  void* stream, uint64_t* ops_done) {
    if (!buf || n_words == 0 || !ops_done) return gt::set_error(GT_ERR_INVALID_ARG, "null buffer or empty range");
    if (kind < 0 || kind > 3)
        return gt::set_error(GT_ERR_INVALID_ARG, "kind must be 0 (read), 1 (or), 2 (cas), 3 (non-temporal read)");
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                 hipSuccess)
        return gt::set_error(GT_ERR_NO_DEVICE, "no HIP device");
    const uint64_t blocks = (uint64_t)cus * 8;
    const uint64_t per_round = blocks * kMbBlock * kMbUnroll;
    const uint64_t rounds = (n_ops + per_round - 1) / per_round;
    hipStream_t s = (hipStream_t)stream;
    uint32_t* b = static_cast<uint32_t*>(buf);
    if (kind == 0)
        mb_read_kernel<false><<<(unsigned)blocks, kMbBlock, 0, s>>>(b, n_words, rounds, seed, b);
    else if (kind == 3)
        mb_read_kernel<true><<<(unsigned)blocks, kMbBlock, 0, s>>>(b, n_words, rounds, seed, b);
    else if (kind == 1)
        mb_or_kernel<<<(unsigned)blocks, kMbBlock, 0, s>>>(b, n_words, rounds, seed);
    else
        mb_cas_kernel<<<(unsigned)blocks, kMbBlock, 0, s>>>(b, n_words, rounds, seed);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return gt::set_error(GT_ERR_DEVICE, hipGetErrorString(e));
    *ops_done = rounds * per_round;
    return GT_OK;
}
