// gt_kernels.hip -- gfx950 kernels of goetia_amd's streaming dBG hot path.
//
// All arithmetic is integer.  The k-mer walkers stage a tile of read bytes in
// LDS with 16-byte loads, roll goetia's canonical cyclic hash
// (rollinghashshifter.hh:180-219, cyclichash.h:85-108) over PER consecutive
// windows per thread, and hand every window to an operation functor:
//
//   Bloom / count-min insert  = claim pass + apply pass (first-touch exact)
//   exact-set insert          = lookup pass + insert pass + resolve pass
//   query, degrees            = one read-only pass
//
// "First touch": goetia inserts k-mers one at a time in read order, so a
// k-mer is new iff, for some table, its bin was empty before it
// (bitstorage.hh:195-219, bytestorage.cc:60-113).  A batch reproduces that
// exactly: a bin that was empty at batch start is claimed by the batch event
// with the smallest ordinal (start position), found with a 64-bit atomicMax on
// gen:32 | ~ordinal:32 stamps, and that event is the one that counts as new.
#include "gt_internal.h"
#include "gt_hash.h"

#include <algorithm>
#include <type_traits>

namespace gt {

__constant__ uint64_t c_char_hash[256] = GT_CHAR_HASH_INIT;

// Table loops run over the compile-time maximum with a uniform guard so that
// every index into the by-value Tables argument is a constant (the argument
// then stays in scalar registers instead of being copied to scratch).
#define GT_FOR_TABLES(i, T) _Pragma("unroll") for (uint32_t i = 0; i < (uint32_t)kMaxTables; ++i) if (i < (T).n)

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ULL;
constexpr uint64_t kEmpty = ~0ULL;

__device__ __forceinline__ uint64_t mix64(uint64_t k) {
    k ^= k >> 33; k *= 0xff51afd7ed558ccdULL;
    k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ULL;
    k ^= k >> 33;
    return k;
}

__device__ __forceinline__ uint64_t load_relaxed(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t wave_sum(uint64_t v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Adds a per-thread count to a device counter with one atomic per wave.
__device__ __forceinline__ void add_counter(uint64_t* ctr, uint64_t v) {
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd((unsigned long long*)ctr, (unsigned long long)v);
}

__device__ __forceinline__ uint64_t wave_max(uint64_t v) {
    for (int o = 32; o > 0; o >>= 1) {
        const uint64_t w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

// Raises a device counter to the wave's largest value, one atomic per wave.
__device__ __forceinline__ void max_counter(uint64_t* ctr, uint64_t v) {
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0 && v) atomicMax((unsigned long long*)ctr, (unsigned long long)v);
}

// ------------------------------------------------------------ claim map

__device__ __forceinline__ uint64_t claim_key(uint32_t gen, uint32_t table, uint64_t bin) {
    return ((uint64_t)(gen & kGenKeyMask) << 36) | ((uint64_t)table << 32) | bin;
}

__device__ __forceinline__ uint64_t claim_stamp(uint32_t gen, uint32_t ordinal) {
    return ((uint64_t)gen << 32) | (uint64_t)(0xFFFFFFFFu - ordinal);
}

__device__ __forceinline__ uint64_t claim_slot(const ClaimMap& cm, uint64_t key) {
    return ((key & ((1ULL << 36) - 1)) * kGolden) >> (64 - cm.log2cap);
}

// The generation of the running sub-batch (uniform across the grid): the
// storage's (exact-set stamps) or a claim map's.
__device__ __forceinline__ uint32_t current_gen(const uint64_t* counters) {
    return (uint32_t)__hip_atomic_load(counters + kGen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t claim_gen(const ClaimMap& cm) {
    return (uint32_t)__hip_atomic_load(cm.gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ void claim_insert(const ClaimMap& cm, uint64_t key, uint64_t stamp, uint64_t* errors) {
    const uint64_t mask = cm.cap - 1;
    const uint64_t gen = key >> 36;
    uint64_t slot = claim_slot(cm, key);
    for (uint64_t probes = 0; probes <= cm.cap; ++probes) {
        uint64_t cur = load_relaxed(&cm.keys[slot]);
        if (cur == key) {
            atomicMax((unsigned long long*)&cm.vals[slot], (unsigned long long)stamp);
            return;
        }
        if ((cur >> 36) != gen) {
            // a slot left by an earlier generation is free
            uint64_t prev = atomicCAS((unsigned long long*)&cm.keys[slot],
                                      (unsigned long long)cur, (unsigned long long)key);
            if (prev == cur || prev == key) {
                atomicMax((unsigned long long*)&cm.vals[slot], (unsigned long long)stamp);
                return;
            }
            if ((prev >> 36) != gen) continue;  // changed but still stale: look again
        }
        slot = (slot + 1) & mask;
    }
    atomicOr((unsigned long long*)errors, (unsigned long long)kErrClaim);
}

__device__ uint64_t claim_find(const ClaimMap& cm, uint64_t key) {
    const uint64_t mask = cm.cap - 1;
    const uint64_t gen = key >> 36;
    uint64_t slot = claim_slot(cm, key);
    for (uint64_t probes = 0; probes <= cm.cap; ++probes) {
        uint64_t cur = cm.keys[slot];
        if (cur == key) return cm.vals[slot];
        if ((cur >> 36) != gen) return 0;
        slot = (slot + 1) & mask;
    }
    return 0;
}

// ------------------------------------------------------------ exact set

__device__ __forceinline__ uint64_t exact_home(const ExactSet& es, uint64_t key) {
    return mix64(key) >> (64 - es.log2cap);
}

// Slot of key or ~0 when absent (read-only tables).
__device__ uint64_t exact_find(const ExactSet& es, uint64_t key) {
    if (key == kEmpty) return (*es.ones) ? es.cap : kEmpty;
    const uint64_t mask = es.cap - 1;
    uint64_t slot = exact_home(es, key);
    for (uint64_t probes = 0; probes < es.cap; ++probes) {
        uint64_t cur = *exact_key(es, slot);
        if (cur == key) return slot;
        if (cur == kEmpty) return kEmpty;
        slot = (slot + 1) & mask;
    }
    return kEmpty;
}

// Inserts key (idempotent) and returns its slot.
__device__ uint64_t exact_insert(const ExactSet& es, uint64_t key, uint64_t* errors) {
    if (key == kEmpty) {
        atomicOr(es.ones, 1u);
        return es.cap;
    }
    const uint64_t mask = es.cap - 1;
    uint64_t slot = exact_home(es, key);
    for (uint64_t probes = 0; probes < es.cap; ++probes) {
        uint64_t cur = load_relaxed(exact_key(es, slot));
        if (cur == key) return slot;
        if (cur == kEmpty) {
            uint64_t prev = atomicCAS((unsigned long long*)exact_key(es, slot),
                                      (unsigned long long)kEmpty, (unsigned long long)key);
            if (prev == kEmpty || prev == key) return slot;
        }
        slot = (slot + 1) & mask;
    }
    atomicOr((unsigned long long*)errors, (unsigned long long)kErrProbe);
    return kEmpty;
}

// ------------------------------------------------------------ storage reads

// Bloom: presence of bin b of table i (bit b%8 of byte b/8, which is bit b%32
// of the little-endian 32-bit word b/32).
__device__ __forceinline__ bool bit_get(const uint8_t* t, uint64_t b) {
    const uint32_t w = reinterpret_cast<const uint32_t*>(t)[b >> 5];
    return (w >> (b & 31)) & 1u;
}

// BitStorage::query (bitstorage.cc:87-100), ByteStorage::query
// (bytestorage.cc:116-139: min over tables, bounded by 255), exact-set count
// (sparseppstorage.cc:28-31).
__device__ int16_t storage_query(const Tables& T, const ExactSet& E, uint64_t h) {
    if (T.kind == kExact) return exact_find(E, h) != kEmpty ? 1 : 0;
    if (T.kind == kBit) {
        GT_FOR_TABLES(i, T) {
            if (!bit_get(T.base[i], gt_fastmod(h, T.p[i], T.m[i]))) return 0;
        }
        return 1;
    }
    int16_t m = 255;
    GT_FOR_TABLES(i, T) {
        int16_t c = T.base[i][gt_fastmod(h, T.p[i], T.m[i])];
        m = c < m ? c : m;
    }
    return m;
}

// ------------------------------------------------------------ tile walker

// LDS of the tile walker for tiles of up to TILE starts.
template <int TILE>
struct TileSharedT {
    uint64_t tf[256];                 // T[c]
    uint64_t tr[256];                 // T[complement(c)]
    int64_t roff[kReadCache + 1];
    alignas(8) uint8_t bytes[TILE + kMaxK];
    uint32_t r_lo, r_hi;
    int cached;
};
using TileShared = TileSharedT<kTileMax>;

template <class Sh>
__device__ __forceinline__ void load_char_tables(Sh& sh) {
    for (int i = threadIdx.x; i < 256; i += kBlock) {
        sh.tf[i] = c_char_hash[i];
        sh.tr[i] = c_char_hash[gt_complement((unsigned char)i)];
    }
}

// Actual batch sizes: the bounds in R, or the run-time sizes in R.dyn.
__device__ __forceinline__ void batch_sizes(const Reads& R, uint64_t& n_reads, int64_t& n_bases) {
    n_reads = R.n_reads;
    n_bases = R.n_bases;
    if (R.dyn) {
        const uint64_t dr = (uint64_t)R.dyn[0];
        const int64_t db = R.dyn[1];
        n_reads = dr < n_reads ? dr : n_reads;
        n_bases = db < n_bases ? db : n_bases;
    }
}

// Visits every k-mer window whose start lies in [R.b0, R.b1).  For windows
// inside a read, op(s, read, read_start, fw, rc, valid) is called in start
// order per thread; valid is false for skip-N windows holding a non-ACGT byte.
// Ops may define end_tile(tile): called by every thread of the workgroup after
// the tile's windows (uniform control flow, so it may synchronise).
template <class Op>
__device__ __forceinline__ auto call_end_tile(Op& op, int64_t tile, int) -> decltype(op.end_tile(tile), void()) {
    op.end_tile(tile);
}
template <class Op>
__device__ __forceinline__ void call_end_tile(Op&, int64_t, long) {}

// Window starts stitched from segments.  Segment u is bytes [u*PER, u*PER+PER)
// of the tile.  With m = ceil(K / PER) and L = K - (m-1)*PER, the window at
// u*PER covers segments u .. u+m-2 whole and the first L bytes of u+m-1, and
// the cyclic hash is linear over XOR, so
//   fw = XOR_j<m-1 rotl(pf[u+j], K - (j+1)*PER)  ^  qf[u+m-1]
//   rc = XOR_j<m-1 rotl(pr[u+j], j*PER)          ^  rotl(qr[u+m-1], (m-1)*PER)
// with pf / qf the forward hash of a whole segment / of its first L bytes
// (CyclicHash::hash over them) and pr / qr the reverse-complement terms
// (XOR_i rotl(T[comp(b_i)], i) over them).  Every thread computes its own
// segment (the first m-1 threads also the segments past the tile), the
// values meet in LDS, and each thread combines m of them instead of rolling
// K bytes twice.  kSegExtra bounds m - 1 (K <= 255, PER >= 4).
constexpr int kSegExtra = 64;
template <int N>
struct SegScratchT {
    static constexpr int kSegs = N;
    uint64_t pf[N], qf[N];
    uint64_t pr[N], qr[N];
    uint8_t bf[N], bq[N];   // non-ACGT bytes in the segment / its prefix
};
using SegScratch = SegScratchT<kBlock + kSegExtra>;

// The op's LDS for stitched starts (free while the windows are visited), or
// null for ops without it (those roll K bytes per start).
template <class Op>
__device__ __forceinline__ auto seg_scratch(Op& op, int) -> decltype(op.seg_scratch()) {
    return op.seg_scratch();
}
template <class Op>
__device__ __forceinline__ SegScratch* seg_scratch(Op&, long) {
    return nullptr;
}

template <int PER, class Sh, class Sg>
__device__ __forceinline__ void seg_values(const Sh& sh, Sg* sg, int u, int L) {
    static_assert(PER % 4 == 0, "segments are read 4 or 8 bytes at a time");
    constexpr int kW = PER % 8 == 0 ? 8 : 4;
    uint64_t pf = 0, pr = 0, qf = 0, qr = 0;
    uint32_t bf = 0, bq = 0;
#pragma unroll
    for (int i0 = 0; i0 < PER; i0 += kW) {
        const uint64_t w = kW == 8 ? *reinterpret_cast<const uint64_t*>(sh.bytes + u * PER + i0)
                                   : (uint64_t)*reinterpret_cast<const uint32_t*>(sh.bytes + u * PER + i0);
#pragma unroll
        for (int k = 0; k < kW; ++k) {
            const int i = i0 + k;
            const uint32_t c = (uint32_t)(w >> (8 * k)) & 0xFFu;
            pf = gt_rotl(pf, 1) ^ sh.tf[c];
            pr ^= gt_rotl(sh.tr[c], (unsigned)i);
            bf += (uint32_t)!gt_is_acgt((unsigned char)c);
            if (i == L - 1) {
                qf = pf;
                qr = pr;
                bq = bf;
            }
        }
    }
    sg->pf[u] = pf;
    sg->qf[u] = qf;
    sg->pr[u] = pr;
    sg->qr[u] = qr;
    sg->bf[u] = (uint8_t)bf;
    sg->bq[u] = (uint8_t)bq;
}

template <int PER, int BLK = kBlock, class Op, class Sh>
__device__ void walk_tiles(const Reads& R, Sh& sh, Op& op) {
    constexpr int kTile = BLK * PER;
    static_assert(sizeof(sh.bytes) >= (size_t)(kTile + kMaxK), "tile LDS too small for PER");
    const int tid = threadIdx.x;
    const int K = R.K;
    const unsigned Kr = (unsigned)K & 63u;
    uint64_t n_reads;
    int64_t n_bases;
    batch_sizes(R, n_reads, n_bases);
    const int64_t b1 = R.b1 < n_bases ? R.b1 : n_bases;
    const int64_t n_grains = (n_bases + kTileGrain - 1) / kTileGrain;
    const int64_t tile_first = R.b0 / kTile;
    const int64_t tile_end = n_reads ? (b1 + kTile - 1) / kTile : 0;
    const bool aligned16 = ((uintptr_t)R.seq & 15u) == 0;

    load_char_tables(sh);
    for (int64_t tile = tile_first + blockIdx.x; tile < tile_end; tile += gridDim.x) {
        __syncthreads();
        const int64_t t0 = tile * kTile;
        const int span = kTile + K - 1;
        const int64_t avail = n_bases - t0;
        const int nload = (int)(avail < span ? avail : span);
        if (aligned16) {
            const int nvec = nload >> 4;
            const uint4* src = reinterpret_cast<const uint4*>(R.seq + t0);
            uint4* dst = reinterpret_cast<uint4*>(sh.bytes);
            for (int i = tid; i < nvec; i += BLK) dst[i] = src[i];
            for (int i = (nvec << 4) + tid; i < span; i += BLK) sh.bytes[i] = i < nload ? R.seq[t0 + i] : 0;
        } else {
            for (int i = tid; i < span; i += BLK) sh.bytes[i] = i < nload ? R.seq[t0 + i] : 0;
        }
        if (tid == 0) {
            const int64_t g_end = (t0 + kTile) / kTileGrain;
            uint32_t rl = R.tile_read[t0 / kTileGrain];
            uint32_t rh = g_end < n_grains ? R.tile_read[g_end] : (uint32_t)(n_reads - 1);
            if (rl >= n_reads) rl = (uint32_t)(n_reads - 1);
            if (rh >= n_reads) rh = (uint32_t)(n_reads - 1);
            if (rh < rl) rh = rl;
            sh.r_lo = rl;
            sh.r_hi = rh;
            sh.cached = (rh - rl + 2) <= (kReadCache + 1);
        }
        __syncthreads();
        const uint32_t rl = sh.r_lo, rh = sh.r_hi;
        const bool cached = sh.cached;
        if (cached) {
            for (uint32_t i = tid; i < rh - rl + 2; i += BLK) sh.roff[i] = R.off[rl + i];
        }
        __syncthreads();

        op.begin_tile(t0);
        auto* sg = seg_scratch(op, 0);
        const int n_seg = (K + PER - 1) / PER;
        if (sg) {
            static_assert(std::remove_pointer_t<decltype(sg)>::kSegs >= BLK + kSegExtra, "segment scratch too small");
            if (n_seg - 1 > kSegExtra) sg = nullptr;       // uniform: K too long to stitch at this PER (none here)
        }
        if (sg) {
            const int L = K - (n_seg - 1) * PER;
            seg_values<PER>(sh, sg, tid, L);
            if (tid < n_seg - 1) seg_values<PER>(sh, sg, BLK + tid, L);
            __syncthreads();
        }
        const int64_t s0 = t0 + (int64_t)tid * PER;
        int64_t s_end = s0 + PER;
        if (s_end > b1) s_end = b1;
        if (s0 < s_end) {
            // read offsets from the LDS cache or from memory; the two cases are
            // kept apart (uniform branch) so each compiles to its own loads
            uint32_t lo = rl, hi = rh;
            if (cached) {
                while (lo < hi) {
                    const uint32_t mid = (lo + hi + 1) >> 1;
                    if (sh.roff[mid - rl] <= s0) lo = mid; else hi = mid - 1;
                }
            } else {
                while (lo < hi) {
                    const uint32_t mid = (lo + hi + 1) >> 1;
                    if (R.off[mid] <= s0) lo = mid; else hi = mid - 1;
                }
            }
            uint32_t r = lo;
            int64_t rstart, rend;
            if (cached) {
                rstart = sh.roff[r - rl];
                rend = sh.roff[r + 1 - rl];
            } else {
                rstart = R.off[r];
                rend = R.off[r + 1];
            }

            const uint8_t* B = sh.bytes;
            const int li = (int)(s0 - t0);
            uint64_t fw = 0, rc = 0;
            int bad = 0;
            if (sg) {
                for (int j = 0; j + 1 < n_seg; ++j) {
                    fw ^= gt_rotl(sg->pf[tid + j], (unsigned)(K - (j + 1) * PER));
                    rc ^= gt_rotl(sg->pr[tid + j], (unsigned)(j * PER));
                    bad += sg->bf[tid + j];
                }
                fw ^= sg->qf[tid + n_seg - 1];
                rc ^= gt_rotl(sg->qr[tid + n_seg - 1], (unsigned)((n_seg - 1) * PER));
                bad += sg->bq[tid + n_seg - 1];
            } else {
                for (int k = 0; k < K; ++k) {
                    const uint8_t c = B[li + k];
                    fw = gt_rotl(fw, 1) ^ sh.tf[c];
                    bad += !gt_is_acgt(c);
                }
                for (int k = K - 1; k >= 0; --k) rc = gt_rotl(rc, 1) ^ sh.tr[B[li + k]];
            }

            for (int j = 0; j < PER; ++j) {
                const int64_t s = s0 + j;
                if (s >= s_end) break;
                if (j > 0) {
                    const uint8_t out = B[li + j - 1], in = B[li + j + K - 1];
                    // CyclicHash::update (fw) and reverse_update on complements (rc)
                    fw = gt_rotl(fw, 1) ^ gt_rotl(sh.tf[out], Kr) ^ sh.tf[in];
                    rc = gt_rotr(rc ^ gt_rotl(sh.tr[in], Kr) ^ sh.tr[out], 1);
                    bad += (int)!gt_is_acgt(in) - (int)!gt_is_acgt(out);
                }
                if (cached) {
                    while (s >= rend && r < rh) {
                        ++r;
                        rstart = rend;
                        rend = sh.roff[r + 1 - rl];
                    }
                } else {
                    while (s >= rend && r < rh) {
                        ++r;
                        rstart = rend;
                        rend = R.off[r + 1];
                    }
                }
                if (s < rstart || s + K > rend) continue;
                op(s, r, rstart, fw, rc, R.mode == kRaw || bad == 0);
            }
            op.flush();
        }
        call_end_tile(op, tile, 0);
    }
}

__device__ __forceinline__ uint64_t kmer_value(uint64_t fw, uint64_t rc, int canonical) {
    return canonical ? (fw < rc ? fw : rc) : fw;
}

// Packed index of the first window starting at or after base x (x a multiple
// of kTileGrain, or the end of the batch).  Packed indices of the windows
// starting in [x, y) are then exactly [first_kidx(x), first_kidx(y)).
__device__ int64_t first_kidx(const Reads& R, uint64_t n_reads, int64_t n_bases, int64_t x) {
    if (x >= n_bases) return R.kmer_off[n_reads];
    uint32_t r = R.tile_read[x / kTileGrain];
    if (r >= n_reads) r = (uint32_t)(n_reads - 1);
    while (r > 0 && R.off[r] > x) --r;
    while (r + 1 < n_reads && R.off[r + 1] <= x) ++r;
    const int64_t len = R.off[r + 1] - R.off[r];
    const int64_t wins = len >= R.K ? len - R.K + 1 : 0;
    int64_t d = x - R.off[r];
    d = d < 0 ? 0 : (d > wins ? wins : d);
    return R.kmer_off[r] + d;
}

// Per-read accumulation of new k-mers: one atomic per run of the same read.
struct ReadAcc {
    uint64_t* out;
    uint32_t cur = 0xFFFFFFFFu;
    uint64_t cnt = 0;
    __device__ __forceinline__ void add(uint32_t r, uint64_t v) {
        if (!out) return;
        if (r != cur) {
            flush();
            cur = r;
        }
        cnt += v;
    }
    __device__ __forceinline__ void flush() {
        if (out && cnt) atomicAdd((unsigned long long*)&out[cur], (unsigned long long)cnt);
        cnt = 0;
    }
};

// ------------------------------------------------------------ Bloom / CMS ops

// Claim pass: reads the tables only.  For each table whose bin is empty at
// batch start, records the bin in the claim map with this event's stamp and
// sets bit i of the fresh mask (bit 7 stands for every table >= 7).
struct ClaimOp {
    Tables T;
    ClaimMap cm;
    uint32_t gen;
    uint8_t* fresh;
    int64_t b0;
    int canonical;
    uint64_t* errors;
    uint64_t n_fresh = 0;

    __device__ void operator()(int64_t s, uint32_t, int64_t, uint64_t fw, uint64_t rc, bool valid) {
        if (!valid) return;
        const uint64_t h = kmer_value(fw, rc, canonical);
        const uint32_t ord = (uint32_t)(s - b0);
        uint8_t mask = 0;
        GT_FOR_TABLES(i, T) {
            const uint64_t b = gt_fastmod(h, T.p[i], T.m[i]);
            const bool empty = T.kind == kBit ? !bit_get(T.base[i], b) : T.base[i][b] == 0;
            if (empty) {
                mask |= (uint8_t)(1u << (i < 7 ? i : 7));
                claim_insert(cm, claim_key(gen, i, b), claim_stamp(gen, ord), errors);
            }
        }
        fresh[s - b0] = mask;
        n_fresh += mask != 0;
    }
    __device__ void begin_tile(int64_t) {}
    __device__ void flush() {}
};

// Saturating byte increment (bytestorage.cc:88-93: +1 while below 255) with
// a 32-bit CAS on the enclosing word; returns the count before the insert.
__device__ __forceinline__ uint32_t byte_sat_inc(uint8_t* t, uint64_t b) {
    uint32_t* w = reinterpret_cast<uint32_t*>(t + (b & ~3ULL));
    const uint32_t sh = (uint32_t)(b & 3) * 8u;
    uint32_t old = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
        const uint32_t c = (old >> sh) & 0xFFu;
        if (c == 255u) return c;
        const uint32_t prev = atomicCAS(w, old, old + (1u << sh));
        if (prev == old) return c;
        old = prev;
    }
}

// Apply pass: updates every table (atomicOr for Bloom, saturating CAS for
// count-min), counts table-0 bins that go from empty to occupied, and
// decides newness from the claim map.
struct ApplyOp {
    Tables T;
    ClaimMap cm;
    uint32_t gen;
    const uint8_t* fresh;
    int64_t b0;
    int canonical;
    const int64_t* kmer_off;
    uint8_t* new_flags;
    ReadAcc acc;
    uint64_t n_new = 0, n_occ = 0;

    __device__ void operator()(int64_t s, uint32_t r, int64_t rstart, uint64_t fw, uint64_t rc, bool valid) {
        if (!valid) {
            if (new_flags) new_flags[kmer_off[r] + (s - rstart)] = 0;
            return;
        }
        const uint64_t h = kmer_value(fw, rc, canonical);
        const uint8_t mask = fresh[s - b0];
        const uint64_t stamp = claim_stamp(gen, (uint32_t)(s - b0));
        bool is_new = false;
        GT_FOR_TABLES(i, T) {
            const uint64_t b = gt_fastmod(h, T.p[i], T.m[i]);
            if (T.kind == kBit) {
                const uint32_t bit = 1u << (b & 31);
                const uint32_t old = atomicOr(reinterpret_cast<uint32_t*>(T.base[i]) + (b >> 5), bit);
                if (i == 0 && !(old & bit)) n_occ++;
            } else {
                const uint32_t before = byte_sat_inc(T.base[i], b);
                if (i == 0 && before == 0) n_occ++;
            }
            if (!is_new && (mask & (1u << (i < 7 ? i : 7)))) {
                is_new = claim_find(cm, claim_key(gen, i, b)) == stamp;
            }
        }
        n_new += is_new;
        acc.add(r, is_new);
        if (new_flags) new_flags[kmer_off[r] + (s - rstart)] = is_new;
    }
    __device__ void begin_tile(int64_t) {}
    __device__ void flush() { acc.flush(); }
};

// ------------------------------------------------------------ exact-set ops

// Lookup pass: a window is fresh iff its value is absent at batch start.
struct ExactLookupOp {
    ExactSet E;
    uint8_t* fresh;
    int64_t b0;
    int canonical;
    uint64_t n_fresh = 0;
    __device__ void operator()(int64_t s, uint32_t, int64_t, uint64_t fw, uint64_t rc, bool valid) {
        if (!valid) return;
        const uint8_t f = exact_find(E, kmer_value(fw, rc, canonical)) == kEmpty;
        fresh[s - b0] = f;
        n_fresh += f;
    }
    __device__ void begin_tile(int64_t) {}
    __device__ void flush() {}
};

// Insert pass: fresh windows insert their value and stamp its slot.  The
// slot's stamp word is kept inverted (a new table is all ~0): an atomicMin of
// ~(gen:32 | ~ordinal:32) keeps the largest stamp, which is the current
// batch's smallest ordinal.
struct ExactInsertOp {
    ExactSet E;
    const uint8_t* fresh;
    int64_t b0;
    int canonical;
    uint32_t gen;
    uint64_t* errors;
    __device__ void operator()(int64_t s, uint32_t, int64_t, uint64_t fw, uint64_t rc, bool valid) {
        if (!valid || !fresh[s - b0]) return;
        const uint64_t slot = exact_insert(E, kmer_value(fw, rc, canonical), errors);
        if (slot != kEmpty)
            atomicMin((unsigned long long*)exact_stamp(E, slot),
                      (unsigned long long)~claim_stamp(gen, (uint32_t)(s - b0)));
    }
    __device__ void begin_tile(int64_t) {}
    __device__ void flush() {}
};

// Resolve pass: the fresh window whose stamp its slot kept (the batch's
// smallest ordinal on that key) is the one goetia's sequential insert reports
// as new (sparseppstorage.hh:92-97).
struct ExactResolveOp {
    ExactSet E;
    const uint8_t* fresh;
    int64_t b0;
    int canonical;
    uint32_t gen;
    const int64_t* kmer_off;
    uint8_t* new_flags;
    ReadAcc acc;
    uint64_t n_new = 0;
    uint64_t last_ins = 0, last_new = 0;   // position + 1 of the last inserted / new window
    __device__ void operator()(int64_t s, uint32_t r, int64_t rstart, uint64_t fw, uint64_t rc, bool valid) {
        bool is_new = false;
        if (valid && fresh[s - b0]) {
            const uint64_t slot = exact_find(E, kmer_value(fw, rc, canonical));
            is_new = slot != kEmpty && *exact_stamp(E, slot) == ~claim_stamp(gen, (uint32_t)(s - b0));
        }
        if (valid) last_ins = (uint64_t)(s - b0) + 1;
        if (is_new) last_new = (uint64_t)(s - b0) + 1;
        n_new += is_new;
        acc.add(r, is_new);
        if (new_flags) new_flags[kmer_off[r] + (s - rstart)] = is_new;
    }
    __device__ void begin_tile(int64_t) {}
    __device__ void flush() { acc.flush(); }
};

// ------------------------------------------------------------ query / degrees

// query_sequence (dbg.hh:364-376): per window count (or -1 when skipped) and
// hash value.
// Counts are staged in LDS per tile and written out as one contiguous run
// (the tile's windows have consecutive packed indices).
template <int TILE>
struct QueryStage {
    int16_t out[TILE];
    int64_t k0, k1;                   // packed indices of the tile's windows
};

template <int TILE>
struct QueryOp {
    Tables T;
    ExactSet E;
    Reads R;
    int tile_len;
    int16_t* counts;
    uint64_t* hashes;
    QueryStage<TILE>* st;
    int64_t k0 = 0;
    __device__ void operator()(int64_t s, uint32_t r, int64_t rstart, uint64_t fw, uint64_t rc, bool valid) {
        const int64_t k = R.kmer_off[r] + (s - rstart);
        if (!valid) {
            st->out[k - k0] = -1;
            if (hashes) hashes[k] = 0;
            return;
        }
        const uint64_t h = kmer_value(fw, rc, R.canonical);
        st->out[k - k0] = storage_query(T, E, h);
        if (hashes) hashes[k] = h;
    }
    __device__ void begin_tile(int64_t t0) {
        if (threadIdx.x == 0) {
            uint64_t n_reads;
            int64_t n_bases;
            batch_sizes(R, n_reads, n_bases);
            int64_t t1 = t0 + tile_len;
            if (t1 > R.b1) t1 = R.b1;
            st->k0 = first_kidx(R, n_reads, n_bases, t0);
            st->k1 = first_kidx(R, n_reads, n_bases, t1);
        }
        __syncthreads();
        k0 = st->k0;
    }
    __device__ void flush() {}
    __device__ void end_tile(int64_t) {
        __syncthreads();
        if (counts) {
            const int n = (int)(st->k1 - k0);
            for (int i = threadIdx.x; i < n; i += kBlock) counts[k0 + i] = st->out[i];
        }
    }
};

// Hashing only: fw / rc / validity per window.
struct HashOp {
    const int64_t* kmer_off;
    uint64_t* fw_out;
    uint64_t* rc_out;
    uint8_t* valid_out;
    __device__ void operator()(int64_t s, uint32_t r, int64_t rstart, uint64_t fw, uint64_t rc, bool valid) {
        const int64_t k = kmer_off[r] + (s - rstart);
        if (fw_out) fw_out[k] = fw;
        if (rc_out) rc_out[k] = rc;
        if (valid_out) valid_out[k] = valid;
    }
    __device__ void begin_tile(int64_t) {}
    __device__ void flush() {}
};

// In/out degree of each window: the 8 one-symbol extensions in closed form
// (equal to hashextender.hh:352-364,386-396 shift-and-undo), counted as present
// when their query is >= min_count (unitig_walker.hh:361-371 uses != 0, which
// is min_count 1).  Windows not selected by `only` get 0xFF.
// The `only` mask is read and the degrees are written per tile through LDS,
// one contiguous run each (the tile's windows have consecutive packed
// indices).
template <int TILE>
struct DegreeStage {
    uint8_t only[TILE];
    uint8_t in[TILE];
    uint8_t out[TILE];
    int64_t k0, k1;
};

template <int TILE>
struct DegreeOp {
    Tables T;
    ExactSet E;
    Reads R;
    int tile_len;
    int min_count;
    const uint8_t* only;
    uint8_t* in_deg;
    uint8_t* out_deg;
    uint64_t* nb_val;       // optional: 8 neighbour values per window (left A C G T, right A C G T)
    uint8_t* nb_mask;       // optional: bit c left c present, bit 4 + c right c present
    const uint8_t* bytes;   // LDS tile
    const uint64_t* tf;
    const uint64_t* tr;
    DegreeStage<TILE>* st;
    int64_t t0 = 0, k0 = 0;
    __device__ void operator()(int64_t s, uint32_t r, int64_t rstart, uint64_t fw, uint64_t rc, bool valid) {
        const int64_t i = R.kmer_off[r] + (s - rstart) - k0;
        if (!valid || (only && !st->only[i])) {
            st->in[i] = 0xFF;
            st->out[i] = 0xFF;
            if (nb_mask) {
                nb_mask[k0 + i] = 0;
#pragma unroll
                for (int c = 0; c < 8; ++c) nb_val[(k0 + i) * 8 + c] = 0;
            }
            return;
        }
        const int K = R.K;
        const unsigned Kr = (unsigned)K & 63u;
        const uint8_t first = bytes[s - t0];
        const uint8_t last = bytes[s - t0 + K - 1];
        uint8_t nin = 0, nout = 0, mask = 0;
        const uint64_t rf = gt_rotl(fw, 1) ^ gt_rotl(tf[first], Kr);
        const uint64_t lr = gt_rotl(rc, 1) ^ gt_rotl(tr[last], Kr);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const uint8_t x = (uint8_t)"ACGT"[c];
            // right extension: shift_right(first, x)
            const uint64_t F1 = rf ^ tf[x];
            const uint64_t R1 = gt_rotr(rc ^ gt_rotl(tr[x], Kr) ^ tr[first], 1);
            const uint64_t v1 = kmer_value(F1, R1, R.canonical);
            const bool p1 = storage_query(T, E, v1) >= min_count;
            nout += p1;
            // left extension: shift_left(x, last)
            const uint64_t F2 = gt_rotr(fw ^ gt_rotl(tf[x], Kr) ^ tf[last], 1);
            const uint64_t R2 = lr ^ tr[x];
            const uint64_t v2 = kmer_value(F2, R2, R.canonical);
            const bool p2 = storage_query(T, E, v2) >= min_count;
            nin += p2;
            if (nb_val) {
                nb_val[(k0 + i) * 8 + c] = v2;
                nb_val[(k0 + i) * 8 + 4 + c] = v1;
            }
            mask |= (uint8_t)((p2 ? 1u : 0u) << c) | (uint8_t)((p1 ? 1u : 0u) << (4 + c));
        }
        st->in[i] = nin;
        st->out[i] = nout;
        if (nb_mask) nb_mask[k0 + i] = mask;
    }
    __device__ void begin_tile(int64_t tile_base) {
        t0 = tile_base;
        if (threadIdx.x == 0) {
            uint64_t n_reads;
            int64_t n_bases;
            batch_sizes(R, n_reads, n_bases);
            int64_t t1 = t0 + tile_len;
            if (t1 > R.b1) t1 = R.b1;
            st->k0 = first_kidx(R, n_reads, n_bases, t0);
            st->k1 = first_kidx(R, n_reads, n_bases, t1);
        }
        __syncthreads();
        k0 = st->k0;
        if (only) {
            const int n = (int)(st->k1 - k0);
            for (int j = threadIdx.x; j < n; j += kBlock) st->only[j] = only[k0 + j];
            __syncthreads();
        }
    }
    __device__ void flush() {}
    __device__ void end_tile(int64_t) {
        __syncthreads();
        const int n = (int)(st->k1 - k0);
        for (int j = threadIdx.x; j < n; j += kBlock) {
            in_deg[k0 + j] = st->in[j];
            out_deg[k0 + j] = st->out[j];
        }
    }
};

// ------------------------------------------------------------ kernels

template <int PER>
__global__ __launch_bounds__(kBlock) void claim_kernel(Reads R, Tables T, ClaimMap cm, uint8_t* fresh,
                                                     uint64_t* counters) {
    __shared__ TileShared sh;
    ClaimOp op{T, cm, claim_gen(cm), fresh, R.b0, R.canonical, counters + kErrors};
    walk_tiles<PER>(R, sh, op);
    add_counter(counters + kFresh, op.n_fresh);
}

template <int PER>
__global__ __launch_bounds__(kBlock) void apply_kernel(Reads R, Tables T, ClaimMap cm, const uint8_t* fresh,
                                                     uint64_t* n_new, uint8_t* new_flags, uint64_t* counters) {
    __shared__ TileShared sh;
    ApplyOp op{T, cm, claim_gen(cm), fresh, R.b0, R.canonical, R.kmer_off, new_flags, ReadAcc{n_new}};
    walk_tiles<PER>(R, sh, op);
    add_counter(counters + kNUnique, op.n_new);
    add_counter(counters + kNOccupied, op.n_occ);
}

template <int PER>
__global__ __launch_bounds__(kBlock) void exact_lookup_kernel(Reads R, ExactSet E, uint8_t* fresh,
                                                            uint64_t* counters) {
    __shared__ TileShared sh;
    ExactLookupOp op{E, fresh, R.b0, R.canonical};
    walk_tiles<PER>(R, sh, op);
    add_counter(counters + kFresh, op.n_fresh);
}

template <int PER>
__global__ __launch_bounds__(kBlock) void exact_insert_kernel(Reads R, ExactSet E, const uint8_t* fresh,
                                                            uint64_t* counters) {
    __shared__ TileShared sh;
    ExactInsertOp op{E, fresh, R.b0, R.canonical, current_gen(counters), counters + kErrors};
    walk_tiles<PER>(R, sh, op);
}

template <int PER>
__global__ __launch_bounds__(kBlock) void exact_resolve_kernel(Reads R, ExactSet E, const uint8_t* fresh,
                                                             uint64_t* n_new, uint8_t* new_flags,
                                                             uint64_t* counters) {
    __shared__ TileShared sh;
    ExactResolveOp op{E, fresh, R.b0, R.canonical, current_gen(counters), R.kmer_off, new_flags, ReadAcc{n_new}};
    walk_tiles<PER>(R, sh, op);
    add_counter(counters + kNUnique, op.n_new);
    max_counter(counters + kLastIns, op.last_ins);
    max_counter(counters + kLastNew, op.last_new);
}

template <int PER>
__global__ __launch_bounds__(kBlock) void query_kernel(Reads R, Tables T, ExactSet E, int16_t* counts,
                                                     uint64_t* hashes) {
    __shared__ TileSharedT<kBlock * PER> sh;
    __shared__ QueryStage<kBlock * PER> st;
    QueryOp<kBlock * PER> op{T, E, R, kBlock * PER, counts, hashes, &st};
    walk_tiles<PER>(R, sh, op);
}

template <int PER>
__global__ __launch_bounds__(kBlock) void hash_kernel(Reads R, uint64_t* fw, uint64_t* rc, uint8_t* valid) {
    __shared__ TileShared sh;
    HashOp op{R.kmer_off, fw, rc, valid};
    walk_tiles<PER>(R, sh, op);
}

template <int PER>
__global__ __launch_bounds__(kBlock) void degree_kernel(Reads R, Tables T, ExactSet E, int min_count,
                                                      const uint8_t* only, uint8_t* in_deg, uint8_t* out_deg,
                                                      uint64_t* nb_val, uint8_t* nb_mask) {
    __shared__ TileSharedT<kBlock * PER> sh;
    __shared__ DegreeStage<kBlock * PER> st;
    DegreeOp<kBlock * PER> op{T,       E,       R,        kBlock * PER, min_count, only, in_deg,
                              out_deg, nb_val,  nb_mask,  sh.bytes,     sh.tf,     sh.tr, &st};
    walk_tiles<PER>(R, sh, op);
}

// Per-hash operations (dbg.hh:140-177 insert / query / insert_and_query on a
// hash value), over an array in index order.
__global__ __launch_bounds__(kBlock) void hash_claim_kernel(const uint64_t* h, uint64_t n, Tables T, ClaimMap cm,
                                                          uint8_t* fresh, uint64_t* counters) {
    uint64_t nf = 0;
    const uint32_t gen = claim_gen(cm);
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) {
        uint8_t mask = 0;
        GT_FOR_TABLES(t, T) {
            const uint64_t b = gt_fastmod(h[i], T.p[t], T.m[t]);
            const bool empty = T.kind == kBit ? !bit_get(T.base[t], b) : T.base[t][b] == 0;
            if (empty) {
                mask |= (uint8_t)(1u << (t < 7 ? t : 7));
                claim_insert(cm, claim_key(gen, t, b), claim_stamp(gen, (uint32_t)i), counters + kErrors);
            }
        }
        fresh[i] = mask;
        nf += mask != 0;
    }
    add_counter(counters + kFresh, nf);
}

__global__ __launch_bounds__(kBlock) void hash_apply_kernel(const uint64_t* h, uint64_t n, Tables T, ClaimMap cm,
                                                          const uint8_t* fresh, uint8_t* is_new,
                                                          uint64_t* counters) {
    uint64_t nn = 0, occ = 0;
    const uint32_t gen = claim_gen(cm);
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) {
        const uint8_t mask = fresh[i];
        const uint64_t stamp = claim_stamp(gen, (uint32_t)i);
        bool nw = false;
        GT_FOR_TABLES(t, T) {
            const uint64_t b = gt_fastmod(h[i], T.p[t], T.m[t]);
            if (T.kind == kBit) {
                const uint32_t bit = 1u << (b & 31);
                const uint32_t old = atomicOr(reinterpret_cast<uint32_t*>(T.base[t]) + (b >> 5), bit);
                if (t == 0 && !(old & bit)) occ++;
            } else {
                if (byte_sat_inc(T.base[t], b) == 0 && t == 0) occ++;
            }
            if (!nw && (mask & (1u << (t < 7 ? t : 7)))) nw = claim_find(cm, claim_key(gen, t, b)) == stamp;
        }
        nn += nw;
        if (is_new) is_new[i] = nw;
    }
    add_counter(counters + kNUnique, nn);
    add_counter(counters + kNOccupied, occ);
}

__global__ __launch_bounds__(kBlock) void hash_exact_lookup_kernel(const uint64_t* h, uint64_t n, ExactSet E,
                                                                 uint8_t* fresh, uint64_t* counters) {
    uint64_t nf = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) {
        const uint8_t f = exact_find(E, h[i]) == kEmpty;
        fresh[i] = f;
        nf += f;
    }
    add_counter(counters + kFresh, nf);
}

__global__ __launch_bounds__(kBlock) void hash_exact_insert_kernel(const uint64_t* h, uint64_t n, ExactSet E,
                                                                 const uint8_t* fresh, uint64_t* counters) {
    const uint32_t gen = current_gen(counters);
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) {
        if (!fresh[i]) continue;
        const uint64_t slot = exact_insert(E, h[i], counters + kErrors);
        if (slot != kEmpty)
            atomicMin((unsigned long long*)exact_stamp(E, slot), (unsigned long long)~claim_stamp(gen, (uint32_t)i));
    }
}

__global__ __launch_bounds__(kBlock) void hash_exact_resolve_kernel(const uint64_t* h, uint64_t n, ExactSet E,
                                                                  const uint8_t* fresh, uint8_t* is_new,
                                                                  uint64_t* counters) {
    uint64_t nn = 0, last_ins = 0, last_new = 0;
    const uint32_t gen = current_gen(counters);
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) {
        bool nw = false;
        if (fresh[i]) {
            const uint64_t slot = exact_find(E, h[i]);
            nw = slot != kEmpty && *exact_stamp(E, slot) == ~claim_stamp(gen, (uint32_t)i);
        }
        nn += nw;
        last_ins = i + 1;
        if (nw) last_new = i + 1;
        if (is_new) is_new[i] = nw;
    }
    add_counter(counters + kNUnique, nn);
    max_counter(counters + kLastIns, last_ins);
    max_counter(counters + kLastNew, last_new);
}

__global__ __launch_bounds__(kBlock) void hash_query_kernel(const uint64_t* h, uint64_t n, Tables T, ExactSet E,
                                                          int16_t* counts) {
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) {
        counts[i] = storage_query(T, E, h[i]);
    }
}

// ------------------------------------------------------------ read layout

// tile_read[t] = read holding base t*kTileGrain (reads are [off[r], off[r+1])).
// With dyn set, the sizes are read from device memory (bounded by the args).
__global__ void tile_read_kernel(const int64_t* off, uint64_t n_reads, int64_t n_bases, uint32_t* tile_read,
                                 int64_t n_tiles, const int64_t* dyn) {
    if (dyn) {
        n_reads = (uint64_t)dyn[0] < n_reads ? (uint64_t)dyn[0] : n_reads;
        n_bases = dyn[1] < n_bases ? dyn[1] : n_bases;
        const int64_t t = (n_bases + kTileGrain - 1) / kTileGrain;
        n_tiles = t < n_tiles ? t : n_tiles;
    }
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < n_reads;
         r += (uint64_t)gridDim.x * blockDim.x) {
        int64_t a = off[r], b = off[r + 1];
        a = a < 0 ? 0 : (a > n_bases ? n_bases : a);
        b = b < a ? a : (b > n_bases ? n_bases : b);
        for (int64_t t = (a + kTileGrain - 1) / kTileGrain; t * kTileGrain < b && t < n_tiles; ++t)
            tile_read[t] = (uint32_t)r;
    }
}

// Exclusive scan of per-read k-mer counts max(0, len - K + 1) in three steps:
// per-block scan, scan of block sums, add-back.
constexpr int kScanBlock = 256;
constexpr int kScanPer = 4;
constexpr int kScanChunk = kScanBlock * kScanPer;

__device__ __forceinline__ int64_t read_kmers(const int64_t* off, uint64_t r, int K, uint64_t live) {
    if (r >= live) return 0;
    const int64_t len = off[r + 1] - off[r];
    return len >= K ? len - K + 1 : 0;
}

__global__ __launch_bounds__(kScanBlock) void scan_local_kernel(const int64_t* off, uint64_t n_reads, int K,
                                                              int64_t* kmer_off, int64_t* block_sums,
                                                              const int64_t* dyn) {
    const uint64_t live = dyn && (uint64_t)dyn[0] < n_reads ? (uint64_t)dyn[0] : n_reads;
    __shared__ int64_t part[kScanBlock];
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk + (uint64_t)threadIdx.x * kScanPer;
    int64_t v[kScanPer];
    int64_t sum = 0;
#pragma unroll
    for (int j = 0; j < kScanPer; ++j) {
        v[j] = base + j < n_reads ? read_kmers(off, base + j, K, live) : 0;
        sum += v[j];
    }
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int o = 1; o < kScanBlock; o <<= 1) {
        int64_t add = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0;
        __syncthreads();
        part[threadIdx.x] += add;
        __syncthreads();
    }
    int64_t run = part[threadIdx.x] - sum;
#pragma unroll
    for (int j = 0; j < kScanPer; ++j) {
        if (base + j < n_reads) kmer_off[base + j] = run;
        run += v[j];
    }
    if (threadIdx.x == kScanBlock - 1) block_sums[blockIdx.x] = part[threadIdx.x];
}

__global__ __launch_bounds__(kScanBlock) void scan_blocks_kernel(int64_t* block_sums, uint64_t n_blocks,
                                                               int64_t* total) {
    __shared__ int64_t part[kScanBlock];
    __shared__ int64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint64_t base = 0; base < n_blocks; base += kScanBlock) {
        const uint64_t i = base + threadIdx.x;
        const int64_t v = i < n_blocks ? block_sums[i] : 0;
        part[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < kScanBlock; o <<= 1) {
            int64_t add = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0;
            __syncthreads();
            part[threadIdx.x] += add;
            __syncthreads();
        }
        if (i < n_blocks) block_sums[i] = carry + part[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == kScanBlock - 1) carry += part[threadIdx.x];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(kScanBlock) void scan_add_kernel(int64_t* kmer_off, uint64_t n_reads,
                                                            const int64_t* block_sums) {
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk;
    const int64_t add = block_sums[blockIdx.x];
    for (int j = threadIdx.x; j < kScanChunk; j += kScanBlock) {
        if (base + j < n_reads) kmer_off[base + j] += add;
    }
}

// Starts a new first-touch generation of a storage and of a claim map (one
// lane; stream-ordered before the sub-batch that uses them).
__global__ void bump_gen_kernel(uint64_t* storage_gen, uint64_t* claim_gen) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (storage_gen) *storage_gen += 1;
        if (claim_gen) *claim_gen += 1;
    }
}

// After an exact-set sub-batch: if it inserted anything, record whether its
// last inserted value was already present (kNOccupied), and clear the
// positions for the next sub-batch.
__global__ void exact_tail_kernel(uint64_t* counters) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const uint64_t li = counters[kLastIns], ln = counters[kLastNew];
        if (li) counters[kNOccupied] = li != ln;
        counters[kLastIns] = 0;
        counters[kLastNew] = 0;
    }
}

__global__ void fill_u64_kernel(uint64_t* p, uint64_t n, uint64_t v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Bloom union (bitstorage.cc:103-137): me |= other, counting table-0 bits
// that become set.
__global__ void bit_union_kernel(uint32_t* me, const uint32_t* other, uint64_t n_words, int count_bits,
                                 uint64_t* counters) {
    uint64_t added = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n_words;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t a = me[i], b = other[i];
        const uint32_t c = a | b;
        if (count_bits) added += __popc(c ^ a);
        me[i] = c;
    }
    add_counter(counters + kNOccupied, added);
}

// Rehash of an exact set into a larger one (stamps are per batch and not kept).
__global__ void exact_rehash_kernel(ExactSet from, ExactSet to, uint64_t* counters) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < from.cap;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = *exact_key(from, i);
        if (k != kEmpty) exact_insert(to, k, counters + kErrors);
    }
}

// Compacts the keys of an exact set into out (count in *n).
__global__ void exact_keys_kernel(ExactSet E, uint64_t* out, uint64_t cap_out, unsigned long long* n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < E.cap;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = *exact_key(E, i);
        if (k != kEmpty) {
            const unsigned long long j = atomicAdd(n, 1ULL);
            if (j < cap_out) out[j] = k;
        }
    }
}

// ------------------------------------------------------------ binned insert
//
// Bloom / count-min insert of a large sub-batch without global atomics on the
// tables.  Every (window, table) event is routed by its bin:
//
//   pass 1  hash the tile's windows, sort the tile's events by first digit in
//           LDS and write them to the tile's own region (coalesced);
//   pass 2  per (first digit, group of kBinGroup tiles): count, then scatter
//           the events by second digit into bucket order (stable);
//   apply   one wavefront per bucket of kBucketBins bins loads the table slice
//           into LDS and replays the bucket's events in event order.
//
// Both partitions are stable, so a bucket holds its events in window order.
// The replay takes 64 events at a time: lanes that hit the same bin are found
// with one ballot per bin bit, and a lane's rank among them is the number of
// earlier events of the chunk on that bin, so the value the bin held just
// before the event is (slice value at chunk start) + rank.  That is goetia's
// sequential insert, event by event (bytestorage.cc:60-113,
// bitstorage.hh:195-219): the count becomes min(255, count + hits), an event
// is new iff some table's bin held 0 just before it, and table-0 bins going
// from 0 to nonzero count as occupied.
//
// A plain insert does not need event order inside a bucket: a bin ends at
// min(255, before + hits) (set, for Bloom) and the k-mer new through it is
// its event with the smallest window ordinal, when the bin was empty before
// the batch.  Plain inserts therefore take order-free variants: both passes
// rank an event within its digit by an LDS atomicAdd (bin_l1_kernel<true>,
// bin_l2_scatter_u_kernel), and bin_apply_u_kernel replays a bucket per
// workgroup with an LDS atomicMin of the ordinal and an atomicAdd of one hit
// per bin, then settles each bin from its smallest-ordinal event.
// insert_and_query and diginorm keep the ordered kernels.

// Mask of the lanes whose v equals this lane's (all 64 lanes must be active).
template <int BITS>
__device__ __forceinline__ uint64_t match_mask(uint32_t v) {
    uint64_t m = ~0ULL;
#pragma unroll
    for (int k = 0; k < BITS; ++k) {
        const bool bit = (v >> k) & 1u;
        const uint64_t b = __ballot(bit);
        m &= bit ? b : ~b;
    }
    return m;
}

__device__ __forceinline__ uint64_t lanes_below() { return (1ULL << (threadIdx.x & 63)) - 1ULL; }

// Mask of the active lanes of this wave whose key equals this lane's, built
// in a per-wave LDS slot array indexed by key: clear, OR in the lane bit, read
// back.  A wave's LDS instructions execute in issue order, so the three steps
// need no barrier beyond keeping the compiler from moving them.
__device__ __forceinline__ uint64_t match_lds(uint64_t* slots, uint32_t key, bool active) {
    if (active) slots[key] = 0;
    __builtin_amdgcn_wave_barrier();
    if (active)
        __hip_atomic_fetch_or(&slots[key], 1ULL << (threadIdx.x & 63), __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_WORKGROUP);
    __builtin_amdgcn_wave_barrier();
    const uint64_t m = active ? slots[key] : 0ULL;
    __builtin_amdgcn_wave_barrier();
    return m;
}

// Packed index of the first window starting at or after b0 (one lane).
__global__ void bin_kbase_kernel(const int64_t* off, const int64_t* kmer_off, uint64_t n_reads, int K, int64_t b0,
                                 uint64_t* out) {
    if (blockIdx.x || threadIdx.x) return;
    uint64_t lo = 0, hi = n_reads - 1;
    while (lo < hi) {
        const uint64_t mid = (lo + hi + 1) >> 1;
        if (off[mid] <= b0) lo = mid; else hi = mid - 1;
    }
    const int64_t len = off[lo + 1] - off[lo];
    const int64_t wins = len >= K ? len - K + 1 : 0;
    int64_t d = b0 - off[lo];
    d = d < 0 ? 0 : (d > wins ? wins : d);
    *out = (uint64_t)(kmer_off[lo] + d);
}

// Exclusive scan over the workgroup (one value per thread): a shuffle scan
// per wave, then the wave totals through `part` (BLK / 64 entries); the
// workgroup total goes to *total.  Ends with a barrier, so `part` can be
// reused at once.
template <int BLK = kBlock>
__device__ uint32_t block_exclusive_u32(uint32_t v, uint32_t* part, uint32_t* total = nullptr) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) part[wave] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < BLK / 64; ++w) {
        const uint32_t p = part[w];
        off += w < wave ? p : 0u;
        tot += p;
    }
    __syncthreads();
    if (total) *total = tot;
    return off + x - v;
}

// Waves that rank a first-pass tile's events in the stable sort (each a
// contiguous quarter, with LDS match slots of kDigits1 words each).
constexpr int kRankWaves = 4;

template <int BLK, bool kU>
struct BinL1Shared {
    uint32_t g[kBinMaxTables * kBinTile];   // table-major: [table][window]; kNoBin = no event
    uint32_t ord[kBinTile];
    union {
        uint16_t perm[kBinMaxTables * kBinTile];  // region position -> event index in g
        SegScratchT<BLK + kSegExtra> seg;         // stitched window starts (before end_tile)
    };
    uint32_t cnt[kU ? kDigits1 : 1];                 // order-free: events per digit, then starts
    uint32_t wpos[kU ? 1 : kRankWaves][kDigits1 + 1];  // stable: per ranking wave, digit counts then positions
    uint32_t part[BLK / 64];
};
static_assert(sizeof(int64_t) * (kReadCache + 1) >= sizeof(uint64_t) * kRankWaves * kDigits1,
              "the stable sort borrows the walker's read-offset cache for its match slots");

__device__ __forceinline__ uint32_t bin_d1(uint32_t g, uint32_t shift2) {
    return g == kNoBin ? (uint32_t)kDigits1 : (g >> kBinBits) >> shift2;
}

// First pass over one tile of kBinTile starts, BLK threads of PER starts
// each.  kU: the events of a digit may leave the tile in any order (plain
// inserts, whose replay depends only on each bin's event count and first
// window); otherwise they keep window order (insert_and_query and diginorm
// replay events in order).
template <bool kU, int BLK, int PER>
struct BinL1Op {
    static_assert(BLK * PER == kBinTile, "a first-pass tile is kBinTile starts");
    static_assert(BLK >= kDigits1, "one thread per first digit");
    static_assert(kU || BLK >= kRankWaves * 64, "the stable sort ranks with kRankWaves waves");
    Tables T;
    BinPlan P;
    const int64_t* kmer_off;
    int canonical;
    uint64_t kbase;
    BinL1Shared<BLK, kU>* sh;
    uint64_t* slots;          // kRankWaves x kDigits1 match slots (the walker's read-offset cache, idle here)
    int64_t b_lo;             // first start of the sub-batch (tiles may begin before it)
    int64_t t0 = 0;
    uint32_t kr = 0xFFFFFFFFu;  // read whose packed-index offset is cached in kd
    int64_t kd = 0;             // kmer_off[kr] - start of read kr - kbase

    __device__ void begin_tile(int64_t tile_base) {
        t0 = tile_base;
        const int w0 = threadIdx.x * PER;
        for (uint32_t t = 0; t < P.n_tables; ++t)
            for (int j = 0; j < PER; ++j) sh->g[t * kBinTile + w0 + j] = kNoBin;
        if (kU && threadIdx.x < kDigits1) sh->cnt[threadIdx.x] = 0;      // digit counters
    }

    __device__ void operator()(int64_t s, uint32_t r, int64_t rstart, uint64_t fw, uint64_t rc, bool valid) {
        if (s < b_lo) return;
        if (r != kr) {
            kr = r;
            kd = kmer_off[r] - rstart - (int64_t)kbase;
        }
        const uint32_t ord = (uint32_t)(s + kd);
        if (P.hasev) P.hasev[ord] = valid ? 1 : 0;
        if (!valid) return;
        const int w = (int)(s - t0);
        const uint64_t h = kmer_value(fw, rc, canonical);
        sh->ord[w] = ord;
#pragma unroll
        for (uint32_t i = 0; i < kBinMaxTables; ++i)
            if (i < P.n_tables) sh->g[i * kBinTile + w] = P.gbase[i] + gt_fastmod30(h, (uint32_t)T.p[i], T.m[i]);
    }

    __device__ void flush() {}

    __device__ SegScratchT<BLK + kSegExtra>* seg_scratch() { return &sh->seg; }

    // Sort of the tile's events by first digit, built in LDS and written out
    // as one contiguous region.  Order-free: an LDS counter per digit gives
    // each event its slot.
    __device__ void end_tile_unordered(int64_t tile) {
        __syncthreads();
        constexpr int kQ = (int)(kBinMaxTables * kBinTile / BLK);
        const uint32_t ne = P.n_tables * kBinTile;
        uint32_t* cnt = sh->cnt;
        uint32_t di[kQ];                                   // digit << 16 | slot in the digit, or ~0
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
            const uint32_t e = (uint32_t)q * BLK + threadIdx.x;
            di[q] = 0xFFFFFFFFu;
            if (e < ne) {
                const uint32_t d = bin_d1(sh->g[e], P.shift2);
                if (d != (uint32_t)kDigits1) di[q] = (d << 16) | atomicAdd(&cnt[d], 1u);
            }
        }
        __syncthreads();
        const uint32_t lt1 = (uint32_t)(tile - P.tile0);
        const int d = threadIdx.x;                         // threads below kDigits1 own a digit
        uint32_t n_ev;
        const uint32_t start = block_exclusive_u32<BLK>(d < kDigits1 ? cnt[d] : 0u, sh->part, &n_ev);
        if (d < kDigits1) {
            P.off1[(size_t)d * P.tiles1 + lt1] = (uint16_t)start;
            cnt[d] = start;
        }
        if (d == kDigits1 - 1) P.off1[(size_t)kDigits1 * P.tiles1 + lt1] = (uint16_t)n_ev;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kQ; ++q)
            if (di[q] != 0xFFFFFFFFu)
                sh->perm[cnt[di[q] >> 16] + (di[q] & 0xFFFFu)] = (uint16_t)((uint32_t)q * BLK + threadIdx.x);
        __syncthreads();
        write_sorted(lt1, n_ev);
    }

    // Writes the tile's events in perm order to its region of ent1, and their
    // second digits to the same positions of dig2 (the second-pass count
    // reads only those).
    __device__ void write_sorted(uint32_t lt1, uint32_t n_ev) {
        const size_t r0 = (size_t)lt1 * P.n_tables * kBinTile;
        uint64_t* dst = P.ent1 + r0;
        uint16_t* dst2 = P.dig2 + r0;
        const uint32_t mask2 = (1u << P.shift2) - 1u;
        for (uint32_t i0 = 0; i0 < n_ev; i0 += 4 * BLK) {
            uint32_t e[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t i = i0 + q * BLK + threadIdx.x;
                e[q] = i < n_ev ? sh->perm[i] : 0u;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t i = i0 + q * BLK + threadIdx.x;
                if (i < n_ev) {
                    const uint32_t g = sh->g[e[q]];
                    dst[i] = ((uint64_t)g << 32) | sh->ord[e[q] & (kBinTile - 1)];
                    dst2[i] = (uint16_t)((g >> kBinBits) & mask2);
                }
            }
        }
    }

    __device__ void end_tile(int64_t tile) {
        if constexpr (kU)
            end_tile_unordered(tile);
        else
            end_tile_stable(tile);
    }

    // Stable sort of the tile's events by first digit: waves 0..kRankWaves-1
    // each rank a contiguous quarter of the events (the other waves wait at
    // the barriers), the sorted order is built in LDS and the region is
    // written out contiguously by every thread.
    __device__ void end_tile_stable(int64_t tile) {
        __syncthreads();
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const bool ranks = wave < kRankWaves;
        const uint64_t lt = lanes_below();
        const int per_wave = (int)(P.n_tables * kBinTile) / kRankWaves;   // a multiple of 512
        const int e0 = wave * per_wave;
        uint32_t* wc = sh->wpos[ranks ? wave : 0];
        uint64_t* sl = slots + (ranks ? wave : 0) * kDigits1;
        if (ranks) {
            for (int d = lane; d <= kDigits1; d += 64) wc[d] = 0;
            __builtin_amdgcn_wave_barrier();
            for (int c = 0; c < per_wave; c += 4 * 64) {
                uint32_t d[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) d[q] = bin_d1(sh->g[e0 + c + q * 64 + lane], P.shift2);
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (d[q] != (uint32_t)kDigits1) atomicAdd(&wc[d[q]], 1u);
            }
        }
        __syncthreads();
        const uint32_t lt1 = (uint32_t)(tile - P.tile0);
        const int d = threadIdx.x;                         // threads below kDigits1 own a digit
        uint32_t tot = 0;
        if (d < kDigits1)
            for (int w = 0; w < kRankWaves; ++w) tot += sh->wpos[w][d];
        uint32_t n_ev;
        const uint32_t start = block_exclusive_u32<BLK>(tot, sh->part, &n_ev);
        if (d < kDigits1) {
            P.off1[(size_t)d * P.tiles1 + lt1] = (uint16_t)start;
            uint32_t p = start;
            for (int w = 0; w < kRankWaves; ++w) {
                const uint32_t c = sh->wpos[w][d];
                sh->wpos[w][d] = p;
                p += c;
            }
        }
        if (d == kDigits1 - 1) P.off1[(size_t)kDigits1 * P.tiles1 + lt1] = (uint16_t)n_ev;
        __syncthreads();
        if (ranks) {
            for (int c = 0; c < per_wave; c += 64) {
                const int e = e0 + c + lane;
                const uint32_t dd = bin_d1(sh->g[e], P.shift2);
                const bool act = dd != (uint32_t)kDigits1;
                const uint64_t m = match_lds(sl, dd, act);
                if (act) sh->perm[wc[dd] + (uint32_t)__popcll(m & lt)] = (uint16_t)e;
                __builtin_amdgcn_wave_barrier();
                if (act && !(m & lt)) wc[dd] += (uint32_t)__popcll(m);
                __builtin_amdgcn_wave_barrier();
            }
        }
        __syncthreads();
        write_sorted(lt1, n_ev);
    }
};

// 512 threads of 4 starts per tile: two workgroups (their LDS) and four
// waves per SIMD.
constexpr int kBinL1Block = 2 * kBlock;

template <bool kU>
__global__ __launch_bounds__(kBinL1Block) void bin_l1_kernel(Reads R, Tables T, BinPlan P) {
    constexpr int BLK = kBinL1Block;
    constexpr int PER = kBinTile / BLK;
    __shared__ TileSharedT<kBinTile> sh;
    __shared__ BinL1Shared<BLK, kU> bs;
    BinL1Op<kU, BLK, PER> op{T, P, R.kmer_off, R.canonical, *P.kbase, &bs, reinterpret_cast<uint64_t*>(sh.roff), R.b0};
    walk_tiles<PER, BLK>(R, sh, op);
}

// Second pass: work item (d1, group) = the d1 runs of kBinGroup consecutive
// first-pass tiles, concatenated in tile order.
struct BinRuns {
    uint32_t rstart[kBinGroup];       // position in ent1 of each tile's run
    uint32_t rpre[kBinGroup + 1];     // exclusive prefix of the run lengths
    uint32_t part[kBlock];
};

// Loads the runs of item (d1, grp) with BLK threads; returns its event count.
template <int BLK = kBlock>
__device__ uint32_t bin_load_runs(const BinPlan& P, uint32_t d1, uint32_t grp, BinRuns& sh) {
    constexpr int kRunsPerThread = kBinGroup / BLK;
    static_assert(BLK / 64 <= kBlock, "wave totals fit the runs' part array");
    uint32_t len[kRunsPerThread];
    uint32_t tot = 0;
#pragma unroll
    for (int k = 0; k < kRunsPerThread; ++k) {
        const uint32_t j = threadIdx.x * kRunsPerThread + k;
        const uint32_t lt = grp * kBinGroup + j;
        uint32_t st = 0;
        len[k] = 0;
        if (lt < P.tiles1) {
            const uint32_t a = P.off1[(size_t)d1 * P.tiles1 + lt];
            const uint32_t b = P.off1[(size_t)(d1 + 1) * P.tiles1 + lt];
            st = lt * P.n_tables * kBinTile + a;
            len[k] = b - a;
        }
        sh.rstart[j] = st;
        tot += len[k];
    }
    uint32_t run = block_exclusive_u32<BLK>(tot, sh.part);
#pragma unroll
    for (int k = 0; k < kRunsPerThread; ++k) {
        sh.rpre[threadIdx.x * kRunsPerThread + k] = run;
        run += len[k];
    }
    if (threadIdx.x == BLK - 1) sh.rpre[kBinGroup] = run;
    __syncthreads();
    return sh.rpre[kBinGroup];
}

// Run holding item event i: the last k with rpre[k] <= i.
__device__ __forceinline__ uint32_t bin_run_of(const BinRuns& sh, uint32_t i) {
    uint32_t lo = 0, hi = kBinGroup - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (sh.rpre[mid] <= i) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Events [c0, c0 + m) of the loaded item, m <= kBinChunk, are found with one
// binary search per 64 events (into `first`, by threads 0..63) and a forward
// step from there per lane; lanes read consecutive events.
template <int CH = kBinChunk, int BLK = kBlock>
__device__ __forceinline__ void bin_chunk_first(const BinRuns& sh, uint32_t* first, uint32_t c0, uint32_t m) {
    static_assert(CH / 64 <= BLK, "one thread per 64 events of a chunk");
    if (threadIdx.x < (unsigned)(CH / 64) && threadIdx.x * 64u < m) first[threadIdx.x] = bin_run_of(sh, c0 + threadIdx.x * 64u);
}

// Issues a thread's kBinChunk / kBlock loads of the chunk into v (first set
// and visible); none is used here, so they stay in flight.
template <int CH = kBinChunk>
__device__ __forceinline__ void bin_chunk_load(const BinPlan& P, const BinRuns& sh, const uint32_t* first, uint32_t c0,
                                               uint32_t m, uint64_t* v) {
#pragma unroll
    for (int q = 0; q < CH / kBlock; ++q) {
        const uint32_t j = threadIdx.x + (uint32_t)q * kBlock;
        v[q] = 0;
        if (j < m) {
            const uint32_t i = c0 + j;
            uint32_t k = first[j >> 6];
            while (sh.rpre[k + 1] <= i) ++k;
            v[q] = P.ent1[sh.rstart[k] + (i - sh.rpre[k])];
        }
    }
}

__device__ __forceinline__ uint32_t bin_d2(uint64_t e, uint32_t mask2) {
    return ((uint32_t)(e >> 32) >> kBinBits) & mask2;
}

// Second digits of chunk events [c0, c0 + m) into d (kNoDigit past m): the
// loads are issued together and stay in flight until d is read.
constexpr uint32_t kNoDigit = 0xFFFFFFFFu;

template <int BLK = kBlock>
__device__ __forceinline__ void bin_chunk_load_d2(const BinPlan& P, const BinRuns& sh, const uint32_t* first,
                                                  uint32_t c0, uint32_t m, uint32_t* d) {
#pragma unroll
    for (int q = 0; q < kBinChunk / BLK; ++q) {
        const uint32_t j = threadIdx.x + (uint32_t)q * BLK;
        d[q] = kNoDigit;
        if (j < m) {
            const uint32_t i = c0 + j;
            uint32_t k = first[j >> 6];
            while (sh.rpre[k + 1] <= i) ++k;
            d[q] = P.dig2[sh.rstart[k] + (i - sh.rpre[k])];
        }
    }
}

// For each (first digit, group) item: events per second digit, from the
// first pass's digit array; a chunk's loads are issued two chunks ahead.  A
// workgroup counts kCountGroups consecutive groups of one first digit, so
// each second digit's counts for them (adjacent in cnt2) go out as one write.
constexpr uint32_t kCountGroups = 8;

// 512 threads: the count's chains of dependent loads (run offsets, then
// digits) overlap across eight waves per workgroup.
constexpr int kCountBlock = 2 * kBlock;

__global__ __launch_bounds__(kCountBlock) void bin_l2_count_kernel(BinPlan P) {
    constexpr int BLK = kCountBlock;
    __shared__ BinRuns sh;
    __shared__ uint32_t hist[kCountGroups][kDigits2];
    __shared__ uint32_t first[kBinChunk / 64];
    constexpr int kQ = kBinChunk / BLK;
    const uint32_t g_blocks = (P.n_groups + kCountGroups - 1) / kCountGroups;
    const uint32_t n_blocks = P.n_d1 * g_blocks;
    const uint32_t mask2 = (1u << P.shift2) - 1u;
    uint32_t da[kQ], db[kQ];
    for (uint32_t blk = blockIdx.x; blk < n_blocks; blk += gridDim.x) {
        const uint32_t d1 = blk / g_blocks, grp0 = (blk % g_blocks) * kCountGroups;
        const uint32_t ng = P.n_groups - grp0 < kCountGroups ? P.n_groups - grp0 : kCountGroups;
        __syncthreads();
        for (uint32_t x = threadIdx.x; x < kCountGroups * (uint32_t)kDigits2; x += BLK) (&hist[0][0])[x] = 0;
        for (uint32_t gi = 0; gi < ng; ++gi) {
            uint32_t* h = hist[gi];
            const uint32_t n = bin_load_runs<BLK>(P, d1, grp0 + gi, sh);
            auto chunk_len = [&](uint32_t c) {
                return c < n ? (n - c < (uint32_t)kBinChunk ? n - c : (uint32_t)kBinChunk) : 0u;
            };
            auto prefetch = [&](uint32_t c, uint32_t (&d)[kQ]) {
                bin_chunk_first<kBinChunk, BLK>(sh, first, c, chunk_len(c));
                __syncthreads();
                bin_chunk_load_d2<BLK>(P, sh, first, c, chunk_len(c), d);
                __syncthreads();                           // `first` is free again
            };
            prefetch(0, da);
            prefetch(kBinChunk, db);
            auto step = [&](uint32_t c0, uint32_t (&d)[kQ]) {
#pragma unroll
                for (int q = 0; q < kQ; ++q)
                    if (d[q] != kNoDigit) atomicAdd(&h[d[q]], 1u);
                prefetch(c0 + 2 * kBinChunk, d);
            };
            for (uint32_t c0 = 0; c0 < n; c0 += 2 * kBinChunk) {
                step(c0, da);
                if (c0 + kBinChunk < n) step(c0 + kBinChunk, db);
            }
            __syncthreads();                               // bin_load_runs of the next group reuses sh
        }
        // lanes take (digit, group) pairs group-fastest: a digit's groups are adjacent words of cnt2
        for (uint32_t x = threadIdx.x; x < kCountGroups * (mask2 + 1u); x += BLK) {
            const uint32_t d2 = x / kCountGroups, gi = x % kCountGroups;
            if (gi < ng) P.cnt2[(size_t)((d1 << P.shift2) | d2) * P.n_groups + grp0 + gi] = hist[gi][d2];
        }
    }
}

struct BinL2Shared {
    BinRuns runs;
    uint64_t ev[kBinChunk];                 // the chunk's events, item order
    uint16_t perm[kBinChunk];               // chunk position by second digit -> index in ev
    uint32_t wpos[4][kDigits2 + 1];
    uint64_t slots[4][kDigits2];            // per-wave match slots
    uint32_t lstart[kDigits2 + 1];          // chunk-local start of each second digit
    uint32_t cur[kDigits2];                 // next ent2 position of each second digit for this item
    uint32_t first[kBinChunk / 64];         // run of the first event of each 64 in the chunk
};

// Stable scatter of each item by second digit, kBinChunk events at a time:
// the chunk is staged in LDS, ranked per wave, sorted in LDS and written out
// as one contiguous run per digit.
__global__ __launch_bounds__(kBlock) void bin_l2_scatter_kernel(BinPlan P) {
    __shared__ BinL2Shared sh;
    const uint32_t n_items = P.n_d1 * P.n_groups;
    const uint32_t mask2 = (1u << P.shift2) - 1u;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t lt = lanes_below();
    uint32_t* wc = sh.wpos[wave];
    uint64_t* sl = sh.slots[wave];
    for (uint32_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        const uint32_t d1 = item / P.n_groups, grp = item % P.n_groups;
        __syncthreads();
        for (uint32_t d2 = threadIdx.x; d2 < (uint32_t)kDigits2; d2 += kBlock)
            sh.cur[d2] = d2 <= mask2 ? P.cnt2[(size_t)((d1 << P.shift2) | d2) * P.n_groups + grp] : 0u;
        const uint32_t n = bin_load_runs(P, d1, grp, sh.runs);
        // chunk c's events are loaded while chunk c - 1 is sorted
        uint64_t v[kBinChunk / kBlock];
        {
            const uint32_t m0 = n < (uint32_t)kBinChunk ? n : (uint32_t)kBinChunk;
            bin_chunk_first(sh.runs, sh.first, 0, m0);
            __syncthreads();
            bin_chunk_load(P, sh.runs, sh.first, 0, m0, v);
        }
        for (uint32_t c0 = 0; c0 < n; c0 += kBinChunk) {
            const uint32_t m = n - c0 < (uint32_t)kBinChunk ? n - c0 : (uint32_t)kBinChunk;
#pragma unroll
            for (int q = 0; q < kBinChunk / kBlock; ++q) {
                const uint32_t j = threadIdx.x + (uint32_t)q * kBlock;
                if (j < m) sh.ev[j] = v[q];
            }
            for (int d = lane; d <= kDigits2; d += 64) wc[d] = 0;
            const uint32_t c1 = c0 + kBinChunk;
            const uint32_t m1 = c1 < n ? (n - c1 < (uint32_t)kBinChunk ? n - c1 : (uint32_t)kBinChunk) : 0u;
            bin_chunk_first(sh.runs, sh.first, c1, m1);
            __syncthreads();
            bin_chunk_load(P, sh.runs, sh.first, c1, m1, v);
            const uint32_t per_wave = (m + 255u) / 256u * 64u;
            const uint32_t i0 = (uint32_t)wave * per_wave;
            for (uint32_t c = 0; c < per_wave; c += 64) {
                const uint32_t j = i0 + c + (uint32_t)lane;
                if (j < m) atomicAdd(&wc[bin_d2(sh.ev[j], mask2)], 1u);
            }
            __syncthreads();
            // chunk-local digit starts (two digits per thread)
            uint32_t a[2], tot = 0;
            for (int k = 0; k < 2; ++k) {
                const int dd = 2 * threadIdx.x + k;
                a[k] = (uint32_t)sh.wpos[0][dd] + sh.wpos[1][dd] + sh.wpos[2][dd] + sh.wpos[3][dd];
                tot += a[k];
            }
            uint32_t run = block_exclusive_u32(tot, sh.runs.part);
            for (int k = 0; k < 2; ++k) {
                const int dd = 2 * threadIdx.x + k;
                sh.lstart[dd] = run;
                uint32_t p = run;
                for (int w = 0; w < 4; ++w) {
                    const uint32_t c = sh.wpos[w][dd];
                    sh.wpos[w][dd] = p;
                    p += c;
                }
                run += a[k];
            }
            if (threadIdx.x == kBlock - 1) sh.lstart[kDigits2] = run;
            __syncthreads();
            for (uint32_t c = 0; c < per_wave; c += 64) {
                const uint32_t j = i0 + c + (uint32_t)lane;
                const bool act = j < m;
                const uint32_t d = act ? bin_d2(sh.ev[j], mask2) : 0u;
                const uint64_t mk = match_lds(sl, d, act);
                if (act) sh.perm[wc[d] + (uint32_t)__popcll(mk & lt)] = (uint16_t)j;
                __builtin_amdgcn_wave_barrier();
                if (act && !(mk & lt)) wc[d] += (uint32_t)__popcll(mk);
                __builtin_amdgcn_wave_barrier();
            }
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < m; i += kBlock) {
                const uint64_t e = sh.ev[sh.perm[i]];
                const uint32_t d = bin_d2(e, mask2);
                P.ent2[sh.cur[d] + (i - sh.lstart[d])] = e;
            }
            __syncthreads();
            for (uint32_t dd = threadIdx.x; dd < (uint32_t)kDigits2; dd += kBlock)
                sh.cur[dd] += sh.lstart[dd + 1] - sh.lstart[dd];
        }
    }
}

// Events per chunk of the order-free scatter: 8192 (64 KB in LDS, two
// workgroups per CU) puts about 16 events of a digit in each run it writes.
constexpr int kScatChunk = 8192;

struct BinL2SharedU {
    BinRuns runs;
    uint64_t sorted[kScatChunk];            // the chunk's events by second digit
    uint32_t cnt[kDigits2];                 // chunk-local events per digit, then their starts
    uint32_t lstart[kDigits2 + 1];          // chunk-local start of each second digit
    uint32_t cur[kDigits2];                 // next ent2 position of each second digit for this item
    uint32_t first[kScatChunk / 64];
};

// Scatter of each item by second digit for order-free replays, kScatChunk
// events at a time: an LDS counter per digit gives each event of the chunk
// its slot in the sorted chunk, which is written out as one run per digit.
// The next chunk's loads are issued as soon as this chunk is in LDS.
__global__ __launch_bounds__(kBlock) void bin_l2_scatter_u_kernel(BinPlan P) {
    __shared__ BinL2SharedU sh;
    constexpr int kQ = kScatChunk / kBlock;
    const uint32_t n_items = P.n_d1 * P.n_groups;
    const uint32_t mask2 = (1u << P.shift2) - 1u;
    uint64_t v[kQ];
    for (uint32_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        const uint32_t d1 = item / P.n_groups, grp = item % P.n_groups;
        __syncthreads();
        for (uint32_t d2 = threadIdx.x; d2 < (uint32_t)kDigits2; d2 += kBlock)
            sh.cur[d2] = d2 <= mask2 ? P.cnt2[(size_t)((d1 << P.shift2) | d2) * P.n_groups + grp] : 0u;
        const uint32_t n = bin_load_runs(P, d1, grp, sh.runs);
        auto chunk_len = [&](uint32_t c) {
            return c < n ? (n - c < (uint32_t)kScatChunk ? n - c : (uint32_t)kScatChunk) : 0u;
        };
        bin_chunk_first<kScatChunk>(sh.runs, sh.first, 0, chunk_len(0));
        __syncthreads();
        bin_chunk_load<kScatChunk>(P, sh.runs, sh.first, 0, chunk_len(0), v);
        for (uint32_t c0 = 0; c0 < n; c0 += kScatChunk) {
            const uint32_t m = chunk_len(c0);
            for (uint32_t d = threadIdx.x; d < (uint32_t)kDigits2; d += kBlock) sh.cnt[d] = 0;
            __syncthreads();
            uint32_t di[kQ];                                 // digit << 16 | slot in the digit, or ~0
#pragma unroll
            for (int q = 0; q < kQ; ++q) {
                const uint32_t j = threadIdx.x + (uint32_t)q * kBlock;
                di[q] = 0xFFFFFFFFu;
                if (j < m) {
                    const uint32_t d = bin_d2(v[q], mask2);
                    di[q] = (d << 16) | atomicAdd(&sh.cnt[d], 1u);
                }
            }
            __syncthreads();
            // chunk-local digit starts (two digits per thread)
            uint32_t a[2], tot = 0;
            for (int k = 0; k < 2; ++k) {
                a[k] = sh.cnt[2 * threadIdx.x + k];
                tot += a[k];
            }
            uint32_t run = block_exclusive_u32(tot, sh.runs.part);
            for (int k = 0; k < 2; ++k) {
                sh.lstart[2 * threadIdx.x + k] = run;
                run += a[k];
            }
            if (threadIdx.x == kBlock - 1) sh.lstart[kDigits2] = run;
            __syncthreads();
#pragma unroll
            for (int q = 0; q < kQ; ++q)
                if (di[q] != 0xFFFFFFFFu) sh.sorted[sh.lstart[di[q] >> 16] + (di[q] & 0xFFFFu)] = v[q];
            // the next chunk's loads stay in flight while this one is written out
            const uint32_t c1 = c0 + kScatChunk;
            bin_chunk_first<kScatChunk>(sh.runs, sh.first, c1, chunk_len(c1));
            __syncthreads();
            bin_chunk_load<kScatChunk>(P, sh.runs, sh.first, c1, chunk_len(c1), v);
            for (uint32_t i = threadIdx.x; i < m; i += kBlock) {
                const uint64_t e = sh.sorted[i];
                const uint32_t d = bin_d2(e, mask2);
                P.ent2[sh.cur[d] + (i - sh.lstart[d])] = e;
            }
            __syncthreads();
            for (uint32_t dd = threadIdx.x; dd < (uint32_t)kDigits2; dd += kBlock)
                sh.cur[dd] += sh.lstart[dd + 1] - sh.lstart[dd];
        }
    }
}

// Lanes of the wavefront on the same bin of a bucket (key < kBucketBins):
// the low 8 bits through LDS slots, the rest by ballot.
__device__ __forceinline__ uint64_t bin_match(uint64_t* slots, uint32_t key, bool active) {
    uint64_t m = match_lds(slots, key & 255u, active);
#pragma unroll
    for (int k = 8; k < kBinBits; ++k) {
        const bool bit = (key >> k) & 1u;
        const uint64_t b = __ballot(bit);
        m &= bit ? b : ~b;
    }
    return m;
}

// Stages the bins of bucket b of byte tables in LDS (sw: kBucketBins / 4
// words); returns the table of the bucket and the staged words' source.
__device__ __forceinline__ uint32_t* bin_stage_slice(const BinPlan& P, const Tables& T, uint32_t b, uint32_t* sw,
                                                     uint32_t* table, uint32_t* n_words) {
    const int lane = threadIdx.x & 63;
    uint32_t t = 0;
    uint64_t pt = T.p[0];
    uint8_t* base = T.base[0];
#pragma unroll
    for (uint32_t i = 1; i < kBinMaxTables; ++i) {
        if (i < P.n_tables && b >= P.bfirst[i]) {
            t = i;
            pt = T.p[i];
            base = T.base[i];
        }
    }
    const uint64_t bin0 = (uint64_t)(b - P.bfirst[t]) << kBinBits;
    const uint64_t left = pt - bin0;
    const uint32_t nb = (uint32_t)(left < (uint64_t)kBucketBins ? left : (uint64_t)kBucketBins);
    const uint32_t nw = (nb + 3u) / 4u;
    uint32_t* src = reinterpret_cast<uint32_t*>(base + bin0);
    for (uint32_t w0 = 0; w0 < nw; w0 += 8 * 64) {
        uint32_t x[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const uint32_t w = w0 + (uint32_t)q * 64 + (uint32_t)lane;
            x[q] = w < nw ? src[w] : 0u;
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const uint32_t w = w0 + (uint32_t)q * 64 + (uint32_t)lane;
            if (w < nw) sw[w] = x[q];
        }
    }
    *table = t;
    *n_words = nw;
    return src;
}

// Replay: one wavefront per bucket (four per workgroup).  Masked: only the
// events of kept reads before P.r_end take part (batched diginorm).
template <bool kMasked>
__device__ __forceinline__ void bin_apply_body(const BinPlan& P, const Tables& T, uint64_t* counters,
                                               uint32_t (*slice)[kBucketBins / 4], uint64_t (*slots)[256]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t lt = lanes_below();
    const uint64_t kb = *P.kbase;
    const bool bytes = T.kind == kByte;
    uint32_t* sw = slice[wave];
    uint8_t* sb = reinterpret_cast<uint8_t*>(sw);
    uint64_t occ = 0;
    for (uint32_t b = blockIdx.x * 4u + (uint32_t)wave; b < P.n_buckets; b += gridDim.x * 4u) {
        const uint32_t lo_e = P.cnt2[(size_t)b * P.n_groups];
        const uint32_t hi_e = P.cnt2[(size_t)(b + 1) * P.n_groups];
        if (lo_e == hi_e) continue;
        uint32_t t = 0;
        uint64_t pt = T.p[0];
        uint8_t* base = T.base[0];
#pragma unroll
        for (uint32_t i = 1; i < kBinMaxTables; ++i) {
            if (i < P.n_tables && b >= P.bfirst[i]) {
                t = i;
                pt = T.p[i];
                base = T.base[i];
            }
        }
        const uint64_t bin0 = (uint64_t)(b - P.bfirst[t]) << kBinBits;
        const uint64_t left = pt - bin0;
        const uint32_t nb = (uint32_t)(left < (uint64_t)kBucketBins ? left : (uint64_t)kBucketBins);
        const uint32_t nw = bytes ? (nb + 3u) / 4u : (nb + 31u) / 32u;
        uint32_t* src = reinterpret_cast<uint32_t*>(bytes ? base + bin0 : base + bin0 / 8);
        for (uint32_t w0 = 0; w0 < nw; w0 += 8 * 64) {
            uint32_t x[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const uint32_t w = w0 + (uint32_t)q * 64 + (uint32_t)lane;
                x[q] = w < nw ? src[w] : 0u;
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const uint32_t w = w0 + (uint32_t)q * 64 + (uint32_t)lane;
                if (w < nw) sw[w] = x[q];
            }
        }
        __builtin_amdgcn_wave_barrier();
        // events are loaded four chunks ahead
        constexpr int kAhead = 4;
        uint64_t pf[kAhead];
#pragma unroll
        for (int q = 0; q < kAhead; ++q) {
            const uint32_t i = lo_e + (uint32_t)q * 64 + (uint32_t)lane;
            pf[q] = i < hi_e ? P.ent2[i] : 0;
        }
        for (uint32_t c = lo_e; c < hi_e; c += 64) {
            const uint32_t i = c + (uint32_t)lane;
            bool v = i < hi_e;
            if (kMasked && v) {
                const uint32_t r = P.rid[i];
                v = r < P.r_end && (P.keep[r] & 1u);
            }
            const uint64_t e = pf[0];
#pragma unroll
            for (int q = 0; q + 1 < kAhead; ++q) pf[q] = pf[q + 1];
            {
                const uint32_t nx = c + (uint32_t)kAhead * 64 + (uint32_t)lane;
                pf[kAhead - 1] = nx < hi_e ? P.ent2[nx] : 0;
            }
            const uint32_t key = v ? ((uint32_t)(e >> 32) & (kBucketBins - 1)) : 0u;
            const uint64_t m = bin_match(slots[wave], key, v);
            const uint32_t rank = (uint32_t)__popcll(m & lt);
            if (v) {
                if (bytes) {
                    const uint32_t pre = sb[key];
                    if (P.vq) {
                        // insert_and_query (bytestorage.cc:142-150): this table's
                        // count just after the event
                        const uint32_t after = pre + rank + 1;
                        P.vq[(kb + (uint32_t)e) * kBinMaxTables + t] = (uint8_t)(after < 255u ? after : 255u);
                    } else if (pre + rank == 0) {
                        P.flags[kb + (uint32_t)e] = 1;
                    }
                    if (rank == 0) {
                        const uint32_t post = pre + (uint32_t)__popcll(m);
                        sb[key] = (uint8_t)(post < 255u ? post : 255u);
                        occ += (t == 0 && pre == 0);
                    }
                } else {
                    const uint32_t bit = 1u << (key & 31u);
                    const bool pre = (sw[key >> 5] & bit) != 0;
                    if (!pre && rank == 0) {
                        P.flags[kb + (uint32_t)e] = 1;
                        atomicOr(&sw[key >> 5], bit);
                        occ += (t == 0);
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
        for (uint32_t w = lane; w < nw; w += 64) src[w] = sw[w];
        __builtin_amdgcn_wave_barrier();
    }
    add_counter(counters + kNOccupied, occ);
}

__global__ __launch_bounds__(kBlock) void bin_apply_kernel(BinPlan P, Tables T, uint64_t* counters) {
    __shared__ uint32_t slice[4][kBucketBins / 4];
    __shared__ uint64_t slots[4][256];
    bin_apply_body<false>(P, T, counters, slice, slots);
}

// Order-free replay of plain inserts (bitstorage.hh:195-219, bytestorage.cc:60-90
// in batch order), one workgroup per bucket.  Within a batch a bin's final
// value depends only on how many events hit it (count-min: min(255, before +
// hits); Bloom: set when hit), and a k-mer is new through a bin exactly when
// the bin was empty before the batch and the k-mer is the bin's first event,
// i.e. its smallest window ordinal (a bucket belongs to one table, so its
// events have distinct ordinals).  So the events only need, per bin, a hit
// count (count-min) and the smallest ordinal; the event holding that ordinal
// then settles the bin and resets both for the next bucket.  The next
// bucket's slice and first events are loaded while this one is replayed.
template <bool kBytes, int BLK>
struct ApplyNext {
    static constexpr int kQ = 4096 / BLK;                                 // events per thread in registers
    static constexpr int kWords = kBytes ? kBucketBins / 4 : kBucketBins / 32;         // slice words
    static constexpr int kW = (kWords + BLK - 1) / BLK;                   // slice words per thread
    uint32_t b, lo, hi, t, nw;
    uint32_t* src;
    uint32_t w[kW];
    uint64_t ev[kQ];
};

// The first bucket at or after b (stepping by `step`) with events, or n_buckets.
__device__ __forceinline__ uint32_t next_busy_bucket(const BinPlan& P, uint32_t b, uint32_t step) {
    for (; b < P.n_buckets; b += step)
        if (P.cnt2[(size_t)b * P.n_groups] != P.cnt2[(size_t)(b + 1) * P.n_groups]) break;
    return b < P.n_buckets ? b : P.n_buckets;
}

template <bool kBytes, int BLK>
__device__ __forceinline__ void apply_fetch(const BinPlan& P, const Tables& T, uint32_t b, ApplyNext<kBytes, BLK>& x) {
    x.b = b;
    if (b >= P.n_buckets) return;
    x.lo = P.cnt2[(size_t)b * P.n_groups];
    x.hi = P.cnt2[(size_t)(b + 1) * P.n_groups];
    uint32_t t = 0;
    uint64_t pt = T.p[0];
    uint8_t* base = T.base[0];
#pragma unroll
    for (uint32_t i = 1; i < kBinMaxTables; ++i) {
        if (i < P.n_tables && b >= P.bfirst[i]) {
            t = i;
            pt = T.p[i];
            base = T.base[i];
        }
    }
    const uint64_t bin0 = (uint64_t)(b - P.bfirst[t]) << kBinBits;
    const uint64_t left = pt - bin0;
    const uint32_t nb = (uint32_t)(left < (uint64_t)kBucketBins ? left : (uint64_t)kBucketBins);
    x.t = t;
    x.nw = kBytes ? (nb + 3u) / 4u : (nb + 31u) / 32u;
    x.src = reinterpret_cast<uint32_t*>(kBytes ? base + bin0 : base + bin0 / 8);
#pragma unroll
    for (int k = 0; k < ApplyNext<kBytes, BLK>::kW; ++k) {
        const uint32_t w = (uint32_t)k * BLK + threadIdx.x;
        x.w[k] = w < x.nw ? x.src[w] : 0u;
    }
#pragma unroll
    for (int q = 0; q < ApplyNext<kBytes, BLK>::kQ; ++q) {
        const uint32_t i = x.lo + (uint32_t)q * BLK + threadIdx.x;
        x.ev[q] = i < x.hi ? P.ent2[i] : 0;
    }
}

// Order-free replay, one bucket per workgroup at a time; 512 threads (four
// waves per SIMD at two workgroups per CU).
constexpr int kApplyBlock = 2 * kBlock;

template <bool kBytes>
__global__ __launch_bounds__(kApplyBlock) void bin_apply_u_kernel(BinPlan P, Tables T, uint64_t* counters) {
    constexpr int BLK = kApplyBlock;
    using Next = ApplyNext<kBytes, BLK>;
    constexpr int kQ = Next::kQ;
    __shared__ uint32_t slice[kBytes ? kBucketBins / 4 : kBucketBins / 32];
    __shared__ uint32_t first[kBucketBins];               // smallest ordinal per bin, ~0 = none
    __shared__ uint32_t hits[kBytes ? kBucketBins : 1];   // events per bin
    const uint64_t kb = *P.kbase;
    uint8_t* sb = reinterpret_cast<uint8_t*>(slice);
    for (uint32_t i = threadIdx.x; i < (uint32_t)kBucketBins; i += BLK) {
        first[i] = 0xFFFFFFFFu;
        if (kBytes) hits[i] = 0;
    }
    uint64_t occ = 0;
    Next cur, nxt;
    apply_fetch<kBytes, BLK>(P, T, next_busy_bucket(P, blockIdx.x, gridDim.x), cur);
    while (cur.b < P.n_buckets) {
        apply_fetch<kBytes, BLK>(P, T, next_busy_bucket(P, cur.b + gridDim.x, gridDim.x), nxt);
        __syncthreads();                                  // the previous bucket is written back
#pragma unroll
        for (int k = 0; k < Next::kW; ++k) {
            const uint32_t w = (uint32_t)k * BLK + threadIdx.x;
            if (w < cur.nw) slice[w] = cur.w[k];
        }
        auto hit = [&](uint64_t e) {
            const uint32_t key = (uint32_t)(e >> 32) & (kBucketBins - 1);
            atomicMin(&first[key], (uint32_t)e);
            if (kBytes) atomicAdd(&hits[key], 1u);
        };
#pragma unroll
        for (int q = 0; q < kQ; ++q)
            if (cur.lo + (uint32_t)q * BLK + threadIdx.x < cur.hi) hit(cur.ev[q]);
        for (uint32_t i = cur.lo + (uint32_t)kQ * BLK + threadIdx.x; i < cur.hi; i += BLK) hit(P.ent2[i]);
        __syncthreads();
        auto settle = [&](uint64_t e) {
            const uint32_t key = (uint32_t)(e >> 32) & (kBucketBins - 1);
            const uint32_t o = (uint32_t)e;
            if (first[key] != o) return;                  // not the bin's first event
            first[key] = 0xFFFFFFFFu;
            bool fresh;
            if (kBytes) {
                const uint32_t h = atomicExch(&hits[key], 0u);
                const uint32_t pre = sb[key];
                const uint32_t post = pre + (h < 255u ? h : 255u);
                sb[key] = (uint8_t)(post < 255u ? post : 255u);
                fresh = pre == 0;
            } else {
                const uint32_t bit = 1u << (key & 31u);
                fresh = (atomicOr(&slice[key >> 5], bit) & bit) == 0;
            }
            if (fresh) {
                P.flags[kb + o] = 1;
                occ += cur.t == 0;
            }
        };
#pragma unroll
        for (int q = 0; q < kQ; ++q)
            if (cur.lo + (uint32_t)q * BLK + threadIdx.x < cur.hi) settle(cur.ev[q]);
        for (uint32_t i = cur.lo + (uint32_t)kQ * BLK + threadIdx.x; i < cur.hi; i += BLK) settle(P.ent2[i]);
        __syncthreads();
        for (uint32_t w = threadIdx.x; w < cur.nw; w += BLK) cur.src[w] = slice[w];
        cur = nxt;
    }
    add_counter(counters + kNOccupied, occ);
}

__global__ __launch_bounds__(kBlock) void dn_apply_kernel(BinPlan P, Tables T, uint64_t* counters) {
    __shared__ uint32_t slice[4][kBucketBins / 4];
    __shared__ uint64_t slots[4][256];
    bin_apply_body<true>(P, T, counters, slice, slots);
}

// Per-read new k-mers from the window flags; adds the total to n_unique.  A
// workgroup takes kBlock consecutive reads; their flags are one contiguous
// range, staged in LDS with coalesced loads when it fits.
constexpr int kReadNewStage = 32768;

__global__ __launch_bounds__(kBlock) void bin_read_new_kernel(const int64_t* kmer_off, uint64_t n_reads,
                                                            const uint8_t* flags, uint64_t* n_new,
                                                            uint64_t* counters) {
    __shared__ uint8_t fl[kReadNewStage];
    uint64_t tot = 0;
    for (uint64_t r0 = (uint64_t)blockIdx.x * kBlock; r0 < n_reads; r0 += (uint64_t)gridDim.x * kBlock) {
        const uint64_t r1 = r0 + kBlock < n_reads ? r0 + kBlock : n_reads;
        const int64_t k0 = kmer_off[r0], k1 = kmer_off[r1];
        const bool staged = k1 - k0 <= kReadNewStage;
        __syncthreads();
        if (staged)
            for (int64_t i = threadIdx.x; i < k1 - k0; i += kBlock) fl[i] = flags[k0 + i];
        __syncthreads();
        const uint64_t r = r0 + threadIdx.x;
        if (r < r1) {
            uint64_t c = 0;
            const int64_t a = kmer_off[r], b = kmer_off[r + 1];
            if (staged) {
                for (int64_t k = a; k < b; ++k) c += fl[k - k0];
            } else {
                for (int64_t k = a; k < b; ++k) c += flags[k];
            }
            if (n_new) n_new[r] += c;
            tot += c;
        }
    }
    add_counter(counters + kNUnique, tot);
}

// out[j] = the eight neighbour values and the presence mask of window idx[j]
// (every idx[j] is below the view's window count).
__global__ void seg_gather_kernel(const uint64_t* nb, const uint8_t* mask, const uint64_t* idx, uint64_t n,
                                  uint64_t* out_nb, uint8_t* out_mask) {
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t w = idx[j];
#pragma unroll
        for (int c = 0; c < 8; ++c) out_nb[8 * j + c] = nb[8 * w + c];
        out_mask[j] = mask[w];
    }
}

// insert_and_query results from the per-table counts after each event:
// count = min over tables (the reference returns 1 for a new k-mer, and a
// k-mer is new exactly when that minimum is 1); windows without events (N
// holding, skipped) read -1.  Adds the new k-mers to n_unique.
__global__ void bin_iq_combine_kernel(const uint32_t* vq, const int64_t* nk, uint32_t n_tables, int16_t* counts,
                                      uint64_t* counters) {
    const uint64_t n = (uint64_t)*nk;
    const uint32_t mask = n_tables >= 4 ? 0xFFFFFFFFu : ((1u << (8 * n_tables)) - 1u);
    uint64_t n_new = 0;
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t w = vq[k] & mask;
        int16_t c = -1;
        if (w) {
            uint32_t mn = 255;
            for (uint32_t t = 0; t < n_tables; ++t) {
                const uint32_t b = (w >> (8 * t)) & 0xFFu;
                mn = b < mn ? b : mn;
            }
            c = (int16_t)mn;
            n_new += mn == 1;
        }
        if (counts) counts[k] = c;
    }
    add_counter(counters + kNUnique, n_new);
}

// StreamingSolidFilter::Filter::filter_sequence (solidifier.hh:58-73) per read
// from its insert_and_query counts, with the reference's types: count_t
// against a uint32_t threshold, a float ratio compared in double with
// 1.0 - min_prop_solid.  Reads shorter than K do not pass.
__global__ void solid_reads_kernel(const int64_t* off, const int64_t* kmer_off, uint64_t n_reads, int K,
                                   const int16_t* counts, uint32_t threshold, float min_prop_solid,
                                   uint8_t* passed) {
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < n_reads;
         r += (uint64_t)gridDim.x * blockDim.x) {
        const int64_t len = off[r + 1] - off[r];
        if (len < K) {
            passed[r] = 0;
            continue;
        }
        uint32_t n_not_solid = 0;
        for (int64_t k = kmer_off[r]; k < kmer_off[r + 1]; ++k)
            n_not_solid += (uint32_t)(int)counts[k] < threshold ? 1u : 0u;
        const uint32_t n_kmers = (uint32_t)(len - K + 1);
        const bool fails = (double)((float)n_not_solid / (float)n_kmers) >= (1.0 - (double)min_prop_solid);
        passed[r] = fails ? 0 : 1;
    }
}

// ------------------------------------------------------------ batched diginorm
//
// DiginormFilter::Filter::filter_sequence (diginorm.hh:105-119) drops a read
// when median_count_at_least holds (diginorm.hh:39-76): at least
// min_req = (unsigned)(0.5 + float(n_kmers) / 2) of its k-mers have a count of
// at least the cutoff; every other read is inserted before the next read is
// looked at.  A chunk of reads is resolved together: a state byte per read
// holds the current guess of the kept set (kDnKept) and whether the read's
// decision is final (kDnSettled).  One replay of the chunk's events gives every
// k-mer of every unsettled read the count it would see under the guess (the
// tables before the chunk plus the guessed-kept reads before it), and the
// decision pass makes the next guess from those counts.

constexpr uint8_t kDnKept = 1, kDnSettled = 2;

// Read of every window of reads [r0, r1) (wread, indexed by window - the
// chunk's first window) and the reads' starting state: reads shorter than K
// are final and not kept (the reference's k-mer iterator raises
// SequenceLengthException, kmeriterator.hh:72-74, and FilterProcessor skips
// the read, processors.hh:394-401); the others start not kept.  One wavefront
// per read.
__global__ void dn_init_kernel(const int64_t* off, const int64_t* kmer_off, uint32_t r0, uint32_t r1, int K,
                               uint32_t* wread, uint8_t* state) {
    const int64_t kb = kmer_off[r0];
    const int lane = threadIdx.x & 63;
    const uint64_t wpb = blockDim.x / 64;
    for (uint64_t r = r0 + (uint64_t)blockIdx.x * wpb + threadIdx.x / 64; r < r1; r += (uint64_t)gridDim.x * wpb) {
        const int64_t a = kmer_off[r] - kb, b = kmer_off[r + 1] - kb;
        for (int64_t w = a + lane; w < b; w += 64) wread[w] = (uint32_t)r;
        if (lane == 0) state[r] = off[r + 1] - off[r] < K ? kDnSettled : 0;
    }
}

// Read of every event of the chunk (rid, aligned with ent2).
__global__ void dn_event_reads_kernel(BinPlan P, const uint32_t* wread, uint32_t* rid) {
    const uint32_t n = P.cnt2[(size_t)P.n_buckets * P.n_groups];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        rid[i] = wread[(uint32_t)P.ent2[i]];
}

// The events [pos, ...) of one read rr that runs past a 64-event step: each
// sees its bin as it was before the read; a kept read then raises its bins.
// Returns the end of the read's events.
__device__ uint32_t dn_long_read(const BinPlan& P, const uint32_t* rid, uint8_t st, uint32_t cutoff, uint8_t* fails,
                                 uint8_t* sb, uint64_t* sl, uint32_t pos, uint32_t hi_e, uint32_t rr) {
    const int lane = threadIdx.x & 63;
    const uint64_t lt = lanes_below();
    uint32_t q = pos;
    for (;;) {
        const uint32_t i = q + (uint32_t)lane;
        const bool in = i < hi_e && rid[i] == rr;          // a prefix of the lanes
        const uint32_t cnt = (uint32_t)__popcll(__ballot(in));
        if (in && !(st & kDnSettled)) {
            const uint64_t e = P.ent2[i];
            if ((uint32_t)sb[(uint32_t)(e >> 32) & (kBucketBins - 1)] < cutoff) fails[(uint32_t)e] = 1;
        }
        q += cnt;
        if (cnt < 64) break;
    }
    if (st & kDnKept) {
        for (uint32_t c = pos; c < q; c += 64) {
            const uint32_t i = c + (uint32_t)lane;
            const bool act = i < q;
            const uint32_t key = act ? ((uint32_t)(P.ent2[i] >> 32) & (kBucketBins - 1)) : 0u;
            const uint64_t m = bin_match(sl, key, act);
            if (act && !(m & lt)) {
                const uint32_t post = sb[key] + (uint32_t)__popcll(m);
                sb[key] = (uint8_t)(post < 255u ? post : 255u);
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    return q;
}

// One replay of the chunk (one wavefront per bucket, the bucket's slice of the
// tables as they were before the chunk staged in LDS, never written back).
// Events are taken in read order, whole reads at a time: all k-mers of a read
// are queried before any of them is inserted (diginorm.hh:105-116), so an
// event sees its bin's count before its read; an event of an unsettled read
// whose count is below the cutoff marks its window in `fails`; the events of a
// guessed-kept read then raise their bins, saturating at 255
// (bytestorage.cc:60-113).  The events of one read are consecutive within a
// bucket; a step takes the whole reads among the next 64 events.
__global__ __launch_bounds__(kBlock) void dn_replay_kernel(BinPlan P, Tables T, const uint32_t* rid,
                                                           const uint8_t* state, uint32_t cutoff, uint8_t* fails) {
    __shared__ uint32_t slice[4][kBucketBins / 4];
    __shared__ uint64_t slots[4][256];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t lt = lanes_below();
    const uint64_t le = lt | (1ULL << lane);
    uint32_t* sw = slice[wave];
    uint8_t* sb = reinterpret_cast<uint8_t*>(sw);
    for (uint32_t b = blockIdx.x * 4u + (uint32_t)wave; b < P.n_buckets; b += gridDim.x * 4u) {
        const uint32_t lo_e = P.cnt2[(size_t)b * P.n_groups];
        const uint32_t hi_e = P.cnt2[(size_t)(b + 1) * P.n_groups];
        if (lo_e == hi_e) continue;
        uint32_t t, nw;
        bin_stage_slice(P, T, b, sw, &t, &nw);
        __builtin_amdgcn_wave_barrier();
        // the events of the next two steps are loaded ahead: [pos, pos + 64)
        // in (e, r), [pos + 64, pos + 128) in (eb, rb)
        auto load = [&](uint32_t at, uint64_t& ex, uint32_t& rx) {
            const uint32_t i = at + (uint32_t)lane;
            ex = i < hi_e ? P.ent2[i] : 0;
            rx = i < hi_e ? rid[i] : 0xFFFFFFFFu;
        };
        uint32_t pos = lo_e;
        uint64_t e, eb;
        uint32_t r, rb;
        load(pos, e, r);
        load(pos + 64u, eb, rb);
        while (pos < hi_e) {
            uint64_t ec;
            uint32_t rc;
            load(pos + 128u, ec, rc);
            const bool v = pos + (uint32_t)lane < hi_e;
            const uint32_t rn = __shfl(rb, 0);                              // read of event pos + 64
            const uint32_t rp = __shfl_up(r, 1);
            const uint64_t bmask = __ballot(v && (lane == 0 || r != rp));   // first event of each read
            uint32_t np = hi_e - pos < 64u ? hi_e - pos : 64u;
            if (np == 64u && __shfl(r, 63) == rn) {
                // the last read of the step goes on past it
                const uint32_t s63 = 63u - (uint32_t)__clzll(bmask);
                if (s63 == 0) {
                    pos = dn_long_read(P, rid, state[__shfl(r, 0)], cutoff, fails, sb, slots[wave], pos, hi_e,
                                       __shfl(r, 0));
                    load(pos, e, r);
                    load(pos + 64u, eb, rb);
                    continue;
                }
                np = s63;
            }
            const bool act = (uint32_t)lane < np;
            const uint32_t key = act ? ((uint32_t)(e >> 32) & (kBucketBins - 1)) : 0u;
            const uint64_t m = bin_match(slots[wave], key, act);
            const uint8_t st = act ? state[r] : (uint8_t)0;
            const uint64_t km = __ballot(act && (st & kDnKept));
            const uint32_t first = 63u - (uint32_t)__clzll(bmask & le);      // first lane of my read
            if (act) {
                uint32_t seen = sb[key] + (uint32_t)__popcll(m & km & ((1ULL << first) - 1ULL));
                seen = seen < 255u ? seen : 255u;
                if (!(st & kDnSettled) && seen < cutoff) fails[(uint32_t)e] = 1;
            }
            __builtin_amdgcn_wave_barrier();
            if (act && (st & kDnKept) && !(m & km & lt)) {
                const uint32_t post = sb[key] + (uint32_t)__popcll(m & km);
                sb[key] = (uint8_t)(post < 255u ? post : 255u);
            }
            __builtin_amdgcn_wave_barrier();
            pos += np;
            if (np == 64u) {
                e = eb;
                r = rb;
                eb = ec;
                rb = rc;
            } else {
                load(pos, e, r);
                load(pos + 64u, eb, rb);
            }
        }
    }
}

// The next guess for every unsettled read of [r0, r1): C = its windows that
// have events and no mark in `fails`; the read is dropped when C >= min_req,
// computed as the reference does (diginorm.hh:43).  With `upper` set the new
// guess is an upper bound of the kept set (the one replayed was a lower
// bound), so its drops are final; otherwise its keeps are.  ctl[0] counts
// reads whose guess changed, ctl[1] is the first read not yet final.  A
// thread per read; a workgroup's kBlock reads have one contiguous window
// range, staged in LDS when it fits.
__global__ __launch_bounds__(kBlock) void dn_decide_kernel(const int64_t* off, const int64_t* kmer_off, uint32_t r0,
                                                           uint32_t r1, int K, const uint8_t* hasev,
                                                           const uint8_t* fails, int upper, uint8_t* state,
                                                           uint32_t* ctl) {
    __shared__ uint8_t ok[kReadNewStage];
    const int64_t kb = kmer_off[r0];
    uint32_t changed = 0, first_open = 0xFFFFFFFFu;
    for (uint64_t q0 = r0 + (uint64_t)blockIdx.x * kBlock; q0 < r1; q0 += (uint64_t)gridDim.x * kBlock) {
        const uint64_t q1 = q0 + kBlock < r1 ? q0 + kBlock : r1;
        const int64_t w0 = kmer_off[q0] - kb, w1 = kmer_off[q1] - kb;
        const bool staged = w1 - w0 <= kReadNewStage;
        __syncthreads();
        if (staged)
            for (int64_t i = threadIdx.x; i < w1 - w0; i += kBlock) ok[i] = hasev[w0 + i] && !fails[w0 + i];
        __syncthreads();
        const uint64_t r = q0 + threadIdx.x;
        if (r >= q1) continue;
        const uint8_t st = state[r];
        if (st & kDnSettled) continue;
        const int64_t a = kmer_off[r] - kb, b = kmer_off[r + 1] - kb;
        uint32_t c = 0;
        if (staged) {
            for (int64_t w = a; w < b; ++w) c += ok[w - w0];
        } else {
            for (int64_t w = a; w < b; ++w) c += (hasev[w] && !fails[w]) ? 1u : 0u;
        }
        const uint64_t n_kmers = (uint64_t)(off[r + 1] - off[r]) - (uint64_t)K + 1u;
        const uint32_t min_req = (uint32_t)(0.5 + (double)((float)n_kmers / 2));
        const bool keep = c < min_req;
        const bool settled = upper ? !keep : keep;
        changed += keep != ((st & kDnKept) != 0);
        state[r] = (uint8_t)((keep ? kDnKept : 0) | (settled ? kDnSettled : 0));
        if (!settled && (uint32_t)r < first_open) first_open = (uint32_t)r;
    }
    // one atomic per wavefront
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        changed += __shfl_xor(changed, d);
        const uint32_t o = __shfl_xor(first_open, d);
        first_open = o < first_open ? o : first_open;
    }
    if ((threadIdx.x & 63) == 0) {
        if (changed) atomicAdd(&ctl[0], changed);
        if (first_open != 0xFFFFFFFFu) atomicMin(&ctl[1], first_open);
    }
}

__global__ void dn_ctl_reset_kernel(uint32_t* ctl) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl[0] = 0;
        ctl[1] = 0xFFFFFFFFu;
    }
}

__global__ void dn_passed_kernel(const uint8_t* state, uint32_t r0, uint32_t r1, uint8_t* passed) {
    for (uint64_t r = r0 + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < r1; r += (uint64_t)gridDim.x * blockDim.x)
        passed[r] = state[r] & kDnKept;
}

// p[0, n) = v.
__global__ void fill_u8_kernel(uint8_t* p, uint64_t n, uint8_t v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Zeroes p[0, *n) (32-bit words).
__global__ void zero_words_kernel(uint32_t* p, const int64_t* n) {
    const uint64_t len = (uint64_t)*n;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < len; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// Zeroes p[0, *n).
__global__ void zero_bytes_kernel(uint8_t* p, const int64_t* n) {
    const uint64_t len = (uint64_t)*n;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < len; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// Exclusive scan of a u32 array in place, with the total stored at a[n].
constexpr int kScanU32Per = 16;
constexpr int kScanU32Chunk = kBlock * kScanU32Per;


__global__ __launch_bounds__(kBlock) void scan_u32_sums_kernel(const uint32_t* a, uint64_t n, uint32_t* sums) {
    __shared__ uint32_t part[kBlock];
    const uint64_t base = (uint64_t)blockIdx.x * kScanU32Chunk + (uint64_t)threadIdx.x * kScanU32Per;
    uint32_t s = 0;
    for (int j = 0; j < kScanU32Per; ++j) s += base + j < n ? a[base + j] : 0u;
    uint32_t tot;
    block_exclusive_u32(s, part, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kBlock) void scan_u32_blocks_kernel(uint32_t* sums, uint64_t nb, uint32_t* total) {
    __shared__ uint32_t part[kBlock];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint64_t base = 0; base < nb; base += kBlock) {
        const uint64_t i = base + threadIdx.x;
        const uint32_t v = i < nb ? sums[i] : 0u;
        uint32_t tot;
        const uint32_t x = block_exclusive_u32(v, part, &tot);
        if (i < nb) sums[i] = carry + x;
        __syncthreads();
        if (threadIdx.x == 0) carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(kBlock) void scan_u32_apply_kernel(uint32_t* a, uint64_t n, const uint32_t* sums) {
    __shared__ uint32_t part[kBlock];
    const uint64_t base = (uint64_t)blockIdx.x * kScanU32Chunk + (uint64_t)threadIdx.x * kScanU32Per;
    uint32_t v[kScanU32Per];
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanU32Per; ++j) {
        v[j] = base + j < n ? a[base + j] : 0u;
        s += v[j];
    }
    uint32_t run = block_exclusive_u32(s, part) + sums[blockIdx.x];
#pragma unroll
    for (int j = 0; j < kScanU32Per; ++j) {
        if (base + j < n) a[base + j] = run;
        run += v[j];
    }
}

}  // namespace gt

// ------------------------------------------------------------ launchers

#include "gt_launch.h"

namespace gt {

static int tile_grid(const Launch& L, const Reads& R) {
    const int64_t tile = (int64_t)kBlock * L.per;
    const int64_t tiles = (R.b1 + tile - 1) / tile - R.b0 / tile;
    const int64_t g = tiles < L.grid ? tiles : L.grid;
    return g > 0 ? (int)g : 1;
}

static int flat_grid(const Launch& L, uint64_t n) {
    const uint64_t need = (n + kBlock - 1) / kBlock;
    const uint64_t g = need < (uint64_t)L.grid * 4 ? need : (uint64_t)L.grid * 4;
    return g > 0 ? (int)g : 1;
}

// Instantiates a tile kernel for the launch's PER.
#define GT_TILE_LAUNCH(L, kernel, ...)                                                        \
    do {                                                                                      \
        switch ((L).per) {                                                                    \
            case 8: kernel<8><<<tile_grid(L, R), kBlock, 0, (L).stream>>>(__VA_ARGS__); break;   \
            case 32: kernel<32><<<tile_grid(L, R), kBlock, 0, (L).stream>>>(__VA_ARGS__); break; \
            default: kernel<16><<<tile_grid(L, R), kBlock, 0, (L).stream>>>(__VA_ARGS__); break; \
        }                                                                                     \
    } while (0)

void launch_tile_read(const int64_t* off, uint64_t n_reads, int64_t n_bases, uint32_t* tile_read,
                      int64_t n_tiles, const int64_t* dyn, hipStream_t s) {
    const uint64_t g = (n_reads + 255) / 256;
    tile_read_kernel<<<(unsigned)(g < 65536 ? (g ? g : 1) : 65536), 256, 0, s>>>(off, n_reads, n_bases, tile_read,
                                                                              n_tiles, dyn);
}

uint64_t scan_block_count(uint64_t n_reads) { return (n_reads + kScanChunk - 1) / kScanChunk; }

void launch_kmer_offsets(const int64_t* off, uint64_t n_reads, int K, int64_t* kmer_off, int64_t* block_sums,
                         const int64_t* dyn, hipStream_t s) {
    const uint64_t nb = scan_block_count(n_reads);
    if (nb) scan_local_kernel<<<(unsigned)nb, kScanBlock, 0, s>>>(off, n_reads, K, kmer_off, block_sums, dyn);
    scan_blocks_kernel<<<1, kScanBlock, 0, s>>>(block_sums, nb, kmer_off + n_reads);
    if (nb) scan_add_kernel<<<(unsigned)nb, kScanBlock, 0, s>>>(kmer_off, n_reads, block_sums);
}

void launch_fill_u64(uint64_t* p, uint64_t n, uint64_t v, hipStream_t s) {
    if (!n) return;
    const uint64_t g = (n + 255) / 256;
    fill_u64_kernel<<<(unsigned)(g < 16384 ? g : 16384), 256, 0, s>>>(p, n, v);
}

void launch_claim(const Launch& L, const Reads& R, const Tables& T, const ClaimMap& cm, uint8_t* fresh,
                  uint64_t* counters) {
    GT_TILE_LAUNCH(L, claim_kernel, R, T, cm, fresh, counters);
}

void launch_apply(const Launch& L, const Reads& R, const Tables& T, const ClaimMap& cm, const uint8_t* fresh,
                  uint64_t* n_new, uint8_t* new_flags, uint64_t* counters) {
    GT_TILE_LAUNCH(L, apply_kernel, R, T, cm, fresh, n_new, new_flags, counters);
}

void launch_exact_lookup(const Launch& L, const Reads& R, const ExactSet& E, uint8_t* fresh, uint64_t* counters) {
    GT_TILE_LAUNCH(L, exact_lookup_kernel, R, E, fresh, counters);
}

void launch_exact_insert(const Launch& L, const Reads& R, const ExactSet& E, const uint8_t* fresh,
                         uint64_t* counters) {
    GT_TILE_LAUNCH(L, exact_insert_kernel, R, E, fresh, counters);
}

void launch_exact_resolve(const Launch& L, const Reads& R, const ExactSet& E, const uint8_t* fresh,
                          uint64_t* n_new, uint8_t* new_flags, uint64_t* counters) {
    GT_TILE_LAUNCH(L, exact_resolve_kernel, R, E, fresh, n_new, new_flags, counters);
}

void launch_bump_gen(uint64_t* storage_gen, uint64_t* claim_gen, hipStream_t s) {
    bump_gen_kernel<<<1, 64, 0, s>>>(storage_gen, claim_gen);
}

void launch_exact_tail(uint64_t* counters, hipStream_t s) {
    exact_tail_kernel<<<1, 64, 0, s>>>(counters);
}

void launch_query(const Launch& L, const Reads& R, const Tables& T, const ExactSet& E, int16_t* counts,
                  uint64_t* hashes) {
    GT_TILE_LAUNCH(L, query_kernel, R, T, E, counts, hashes);
}

void launch_hash(const Launch& L, const Reads& R, uint64_t* fw, uint64_t* rc, uint8_t* valid) {
    GT_TILE_LAUNCH(L, hash_kernel, R, fw, rc, valid);
}

void launch_degrees(const Launch& L, const Reads& R, const Tables& T, const ExactSet& E, int min_count,
                    const uint8_t* only, uint8_t* in_deg, uint8_t* out_deg, uint64_t* nb_val, uint8_t* nb_mask) {
    GT_TILE_LAUNCH(L, degree_kernel, R, T, E, min_count, only, in_deg, out_deg, nb_val, nb_mask);
}

void launch_hash_claim(const Launch& L, const uint64_t* h, uint64_t n, const Tables& T, const ClaimMap& cm,
                       uint8_t* fresh, uint64_t* counters) {
    hash_claim_kernel<<<flat_grid(L, n), kBlock, 0, L.stream>>>(h, n, T, cm, fresh, counters);
}

void launch_hash_apply(const Launch& L, const uint64_t* h, uint64_t n, const Tables& T, const ClaimMap& cm,
                       const uint8_t* fresh, uint8_t* is_new, uint64_t* counters) {
    hash_apply_kernel<<<flat_grid(L, n), kBlock, 0, L.stream>>>(h, n, T, cm, fresh, is_new, counters);
}

void launch_hash_exact_lookup(const Launch& L, const uint64_t* h, uint64_t n, const ExactSet& E, uint8_t* fresh,
                              uint64_t* counters) {
    hash_exact_lookup_kernel<<<flat_grid(L, n), kBlock, 0, L.stream>>>(h, n, E, fresh, counters);
}

void launch_hash_exact_insert(const Launch& L, const uint64_t* h, uint64_t n, const ExactSet& E,
                              const uint8_t* fresh, uint64_t* counters) {
    hash_exact_insert_kernel<<<flat_grid(L, n), kBlock, 0, L.stream>>>(h, n, E, fresh, counters);
}

void launch_hash_exact_resolve(const Launch& L, const uint64_t* h, uint64_t n, const ExactSet& E,
                               const uint8_t* fresh, uint8_t* is_new, uint64_t* counters) {
    hash_exact_resolve_kernel<<<flat_grid(L, n), kBlock, 0, L.stream>>>(h, n, E, fresh, is_new, counters);
}

void launch_hash_query(const Launch& L, const uint64_t* h, uint64_t n, const Tables& T, const ExactSet& E,
                       int16_t* counts) {
    hash_query_kernel<<<flat_grid(L, n), kBlock, 0, L.stream>>>(h, n, T, E, counts);
}

void launch_bit_union(uint32_t* me, const uint32_t* other, uint64_t n_words, int count_bits, uint64_t* counters,
                      hipStream_t s) {
    const uint64_t g = (n_words + 255) / 256;
    bit_union_kernel<<<(unsigned)(g < 16384 ? (g ? g : 1) : 16384), 256, 0, s>>>(me, other, n_words, count_bits,
                                                                              counters);
}

void launch_exact_rehash(const ExactSet& from, const ExactSet& to, uint64_t* counters, hipStream_t s) {
    const uint64_t g = (from.cap + 255) / 256;
    exact_rehash_kernel<<<(unsigned)(g < 16384 ? g : 16384), 256, 0, s>>>(from, to, counters);
}

void launch_exact_keys(const ExactSet& E, uint64_t* out, uint64_t cap_out, unsigned long long* n, hipStream_t s) {
    const uint64_t g = (E.cap + 255) / 256;
    exact_keys_kernel<<<(unsigned)(g < 16384 ? g : 16384), 256, 0, s>>>(E, out, cap_out, n);
}

}  // namespace gt

// ------------------------------------------------------------ binned insert launchers

namespace gt {

void launch_bin_kbase(const Reads& R, int64_t b0, uint64_t* out, hipStream_t s) {
    bin_kbase_kernel<<<1, 64, 0, s>>>(R.off, R.kmer_off, R.n_reads, R.K, b0, out);
}

void launch_bin_l1(const Reads& R, const Tables& T, const BinPlan& P, int grid, hipStream_t s) {
    const int g = (int)std::min<int64_t>(std::max<int64_t>(P.tiles1, 1), grid);
    if (P.unordered)
        bin_l1_kernel<true><<<g, kBinL1Block, 0, s>>>(R, T, P);
    else
        bin_l1_kernel<false><<<g, kBinL1Block, 0, s>>>(R, T, P);
}

void launch_bin_l2_count(const BinPlan& P, int grid, hipStream_t s) {
    const int64_t blocks = (int64_t)P.n_d1 * ((P.n_groups + kCountGroups - 1) / kCountGroups);
    const int g = (int)std::min<int64_t>(std::max<int64_t>(blocks, 1), grid);
    bin_l2_count_kernel<<<g, kCountBlock, 0, s>>>(P);
}

void launch_bin_l2_scatter(const BinPlan& P, int grid, hipStream_t s) {
    const int g = (int)std::min<int64_t>(std::max<int64_t>((int64_t)P.n_d1 * P.n_groups, 1), grid);
    if (P.unordered)
        bin_l2_scatter_u_kernel<<<g, kBlock, 0, s>>>(P);
    else
        bin_l2_scatter_kernel<<<g, kBlock, 0, s>>>(P);
}

void launch_bin_apply(const BinPlan& P, const Tables& T, uint64_t* counters, int grid, hipStream_t s) {
    if (P.unordered && !P.vq) {
        const int g = (int)std::min<int64_t>(std::max<int64_t>(P.n_buckets, 1), grid);
        if (T.kind == kByte)
            bin_apply_u_kernel<true><<<g, kApplyBlock, 0, s>>>(P, T, counters);
        else
            bin_apply_u_kernel<false><<<g, kApplyBlock, 0, s>>>(P, T, counters);
        return;
    }
    const int g = (int)std::min<int64_t>(std::max<int64_t>((P.n_buckets + 3) / 4, 1), grid);
    bin_apply_kernel<<<g, kBlock, 0, s>>>(P, T, counters);
}

void launch_bin_read_new(const int64_t* kmer_off, uint64_t n_reads, const uint8_t* flags, uint64_t* n_new,
                         uint64_t* counters, hipStream_t s) {
    const uint64_t g = (n_reads + kBlock - 1) / kBlock;
    bin_read_new_kernel<<<(unsigned)(g < 4096 ? (g ? g : 1) : 4096), kBlock, 0, s>>>(kmer_off, n_reads, flags, n_new,
                                                                                  counters);
}

static unsigned grid_of(uint64_t units, uint64_t per_block, uint64_t cap) {
    const uint64_t g = (units + per_block - 1) / per_block;
    return (unsigned)(g < cap ? (g ? g : 1) : cap);
}

void launch_fill_u8(uint8_t* p, uint64_t n, uint8_t v, hipStream_t s) {
    if (n) fill_u8_kernel<<<grid_of(n, 256 * 16, 16384), 256, 0, s>>>(p, n, v);
}

void launch_dn_init(const int64_t* off, const int64_t* kmer_off, uint32_t r0, uint32_t r1, int K, uint32_t* wread,
                    uint8_t* state, hipStream_t s) {
    dn_init_kernel<<<grid_of(r1 - r0, 4, 65535), 256, 0, s>>>(off, kmer_off, r0, r1, K, wread, state);
}

void launch_dn_event_reads(const BinPlan& P, const uint32_t* wread, uint32_t* rid, uint64_t max_events,
                           hipStream_t s) {
    dn_event_reads_kernel<<<grid_of(max_events, 256 * 4, 16384), 256, 0, s>>>(P, wread, rid);
}

void launch_dn_replay(const BinPlan& P, const Tables& T, const uint32_t* rid, const uint8_t* state, uint32_t cutoff,
                      uint8_t* fails, int grid, hipStream_t s) {
    const int g = (int)std::min<int64_t>(std::max<int64_t>((P.n_buckets + 3) / 4, 1), grid);
    dn_replay_kernel<<<g, kBlock, 0, s>>>(P, T, rid, state, cutoff, fails);
}

void launch_dn_decide(const int64_t* off, const int64_t* kmer_off, uint32_t r0, uint32_t r1, int K,
                      const uint8_t* hasev, const uint8_t* fails, int upper, uint8_t* state, uint32_t* ctl,
                      hipStream_t s) {
    dn_decide_kernel<<<grid_of(r1 - r0, kBlock, 16384), kBlock, 0, s>>>(off, kmer_off, r0, r1, K, hasev, fails,
                                                                          upper, state, ctl);
}

void launch_dn_ctl_reset(uint32_t* ctl, hipStream_t s) { dn_ctl_reset_kernel<<<1, 64, 0, s>>>(ctl); }

void launch_dn_passed(const uint8_t* state, uint32_t r0, uint32_t r1, uint8_t* passed, hipStream_t s) {
    if (r1 > r0) dn_passed_kernel<<<grid_of(r1 - r0, 256, 16384), 256, 0, s>>>(state, r0, r1, passed);
}

void launch_dn_apply(const BinPlan& P, const Tables& T, uint64_t* counters, int grid, hipStream_t s) {
    const int g = (int)std::min<int64_t>(std::max<int64_t>((P.n_buckets + 3) / 4, 1), grid);
    dn_apply_kernel<<<g, kBlock, 0, s>>>(P, T, counters);
}

void launch_zero_bytes(uint8_t* p, const int64_t* n, uint64_t max_n, hipStream_t s) {
    const uint64_t g = (max_n + 255) / 256;
    zero_bytes_kernel<<<(unsigned)(g < 16384 ? (g ? g : 1) : 16384), 256, 0, s>>>(p, n);
}

void launch_seg_gather(const uint64_t* nb, const uint8_t* mask, const uint64_t* idx, uint64_t n, uint64_t* out_nb,
                       uint8_t* out_mask, hipStream_t s) {
    const uint64_t g = (n + 255) / 256;
    seg_gather_kernel<<<(unsigned)(g < 16384 ? (g ? g : 1) : 16384), 256, 0, s>>>(nb, mask, idx, n, out_nb,
                                                                               out_mask);
}

void launch_bin_iq_combine(const uint32_t* vq, const int64_t* nk, uint64_t max_n, uint32_t n_tables, int16_t* counts,
                           uint64_t* counters, hipStream_t s) {
    const uint64_t g = (max_n + 255) / 256;
    bin_iq_combine_kernel<<<(unsigned)(g < 16384 ? (g ? g : 1) : 16384), 256, 0, s>>>(vq, nk, n_tables, counts,
                                                                                   counters);
}

void launch_zero_words(uint32_t* p, const int64_t* n, uint64_t max_n, hipStream_t s) {
    const uint64_t g = (max_n + 255) / 256;
    zero_words_kernel<<<(unsigned)(g < 16384 ? (g ? g : 1) : 16384), 256, 0, s>>>(p, n);
}

void launch_solid_reads(const int64_t* off, const int64_t* kmer_off, uint64_t n_reads, int K, const int16_t* counts,
                        uint32_t threshold, float min_prop_solid, uint8_t* passed, hipStream_t s) {
    const uint64_t g = (n_reads + 255) / 256;
    solid_reads_kernel<<<(unsigned)(g < 16384 ? (g ? g : 1) : 16384), 256, 0, s>>>(off, kmer_off, n_reads, K, counts,
                                                                                threshold, min_prop_solid, passed);
}

uint64_t scan_u32_block_count(uint64_t n) { return (n + kScanU32Chunk - 1) / kScanU32Chunk; }

void launch_scan_u32(uint32_t* a, uint64_t n, uint32_t* sums, hipStream_t s) {
    const uint64_t nb = scan_u32_block_count(n);
    if (nb) scan_u32_sums_kernel<<<(unsigned)nb, kBlock, 0, s>>>(a, n, sums);
    scan_u32_blocks_kernel<<<1, kBlock, 0, s>>>(sums, nb, a + n);
    if (nb) scan_u32_apply_kernel<<<(unsigned)nb, kBlock, 0, s>>>(a, n, sums);
}

}  // namespace gt
