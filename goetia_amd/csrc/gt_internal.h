// gt_internal.h -- structures shared by the HIP kernels and the C-ABI layer
// of goetia_amd.  Not part of the public interface (include/goetia_amd.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gt {

constexpr int kMaxTables = 16;

// Work decomposition of the k-mer walkers: a tile is kBlock * PER consecutive
// k-mer start positions of the concatenated read bytes; a workgroup of kBlock
// threads owns one tile at a time and each thread rolls the hash over PER
// consecutive starts.  PER is 8, 16 or 32 (picked per storage kind by the
// autotuner); tiles are multiples of kTileGrain, and sub-batch boundaries are
// multiples of kTileMax.
constexpr int kBlock = 256;
constexpr int kPerMin = 8;
constexpr int kPerMax = 32;
constexpr int kTileGrain = kBlock * kPerMin;
constexpr int kTileMax = kBlock * kPerMax;
constexpr int kMaxK = 256;          // K <= 255 (goetia's K is a uint16_t)
constexpr int kReadCache = 1024;    // read offsets cached in LDS per tile

enum Kind : int { kBit = 1, kByte = 2, kExact = 3 };
enum Mode : int { kRaw = 0, kSkipN = 1 };

// Device counters of one storage (u64 each).  kGen is the storage's
// first-touch generation (exact-set stamps): bumped on the device once per
// sub-batch, so replays of a captured batch never reuse a generation.
// For the exact set, kNOccupied holds 1 when the last value inserted was
// already present (its bucket count depends on it), and kLastIns / kLastNew
// the position + 1 of a sub-batch's last inserted and last new window.
enum Counter : int {
    kNUnique = 0, kNOccupied = 1, kFresh = 2, kErrors = 3, kGen = 4, kLastIns = 5, kLastNew = 6, kNCounters = 7
};

// Error bits OR-ed into counters[kErrors] by kernels.
enum ErrBit : uint64_t { kErrProbe = 1, kErrClaim = 2 };

// Bloom / count-min tables: table i holds bins [0, p[i]) at base[i].
struct Tables {
    uint32_t n;
    int kind;
    uint64_t p[kMaxTables];
    uint64_t m[kMaxTables];      // gt_mod_magic(p[i])
    uint8_t* base[kMaxTables];   // 4-byte aligned, padded to a multiple of 4 bytes
};

// Exact k-mer set: open addressing with linear probing over cap = 2^log2cap
// slots.  keys[cap] is the slot of the all-ones key (which is also the empty
// marker); flag[0] says whether that key is present.  stamp[] carries the
// first-touch stamp of keys created in the current batch.
// Exact set: open addressing with linear probing over cap slots (a power of
// two); slot i is the word pair kv[2i] (its key, empty = ~0) and kv[2i + 1]
// (the inverted first-touch stamp of the batch being inserted, ~0 = none), so
// a key and its stamp share a cache line.  Slot cap holds the all-ones key's
// stamp.  A new table is all ~0.
struct ExactSet {
    uint64_t* kv;      // 2 (cap + 1) words
    uint32_t* ones;    // presence flag of the all-ones key
    uint64_t cap;
    uint32_t log2cap;
};
__host__ __device__ inline uint64_t* exact_key(const ExactSet& es, uint64_t slot) { return es.kv + 2 * slot; }
__host__ __device__ inline uint64_t* exact_stamp(const ExactSet& es, uint64_t slot) { return es.kv + 2 * slot + 1; }

// First-touch claim map for Bloom / count-min batches: key word =
// gen:28 | table:4 | bin:32, value word = gen:32 | ~ordinal:32 (atomicMax).
// The generation is a device counter owned with the map (a map may serve
// several storages, so it cannot use a storage's generation).
struct ClaimMap {
    uint64_t* keys;
    uint64_t* vals;
    const uint64_t* gen;
    uint64_t cap;
    uint32_t log2cap;
};

constexpr uint64_t kMaxBins = 1ULL << 32;       // bins per table
constexpr uint64_t kGenKeyMask = (1ULL << 28) - 1;

// Binned insert of Bloom / count-min tables: the (table, bin) events of a
// sub-batch are partitioned by bin in two stable passes (256-way, then up to
// 512-way) into buckets of kBucketBins consecutive bins of one table, then
// each bucket is replayed in event order against an LDS copy of its slice.
constexpr int kBinBits = 13;
constexpr int kBucketBins = 1 << kBinBits;        // bins per bucket
constexpr int kDigitBits1 = 8;
constexpr int kDigits1 = 1 << kDigitBits1;        // first-pass fan-out
constexpr int kDigitBits2 = 9;
constexpr int kDigits2 = 1 << kDigitBits2;        // largest second-pass fan-out
constexpr int kBinTile = 2048;                    // starts per first-pass tile (512 threads x 4)
constexpr uint32_t kBinMaxTables = 4;
constexpr uint32_t kBinGroup = 1024;              // first-pass tiles per second-pass work item
constexpr int kBinChunk = 4096;                   // events staged in LDS per second-pass step
constexpr uint32_t kMaxBuckets = 1u << (kDigitBits1 + kDigitBits2);
constexpr uint32_t kNoBin = 0xFFFFFFFFu;

// One binned sub-batch.  An event is the u64 (global bin << 32 | ordinal):
// global bin = gbase[table] + bin, ordinal = packed window index - *kbase.
// Bucket = global bin >> kBinBits = d1 << shift2 | d2 (first / second pass
// digits).
struct BinPlan {
    uint32_t n_tables;
    uint32_t shift2;
    uint32_t gbase[kBinMaxTables];
    uint32_t bfirst[kBinMaxTables + 1];   // first bucket of each table; [n_tables] = n_buckets
    uint32_t n_buckets;
    uint32_t n_d1;                        // first-pass digits in use
    int64_t tile0;                        // first first-pass tile (absolute) of the sub-batch
    uint32_t tiles1;                      // first-pass tiles of the sub-batch
    uint32_t n_groups;                    // ceil(tiles1 / kBinGroup)
    uint64_t* ent1;                       // per tile a region of n_tables * kBinTile events, by d1
    uint16_t* dig2;                       // second digit of each ent1 event (same index)
    uint64_t* ent2;                       // events by bucket, event order within a bucket
    uint16_t* off1;                       // (kDigits1 + 1) x tiles1, digit-major: start of d1 in its tile
    uint32_t* cnt2;                       // (n_d1 << shift2) x n_groups (+1): counts, then exclusive starts
    const uint64_t* kbase;                // packed index of the sub-batch's first window
    uint8_t* flags;                       // per packed window: 1 = new k-mer
    uint8_t* vq;                          // insert_and_query: per packed window, kBinMaxTables
                                          // counts just after the event (flags unused then)
    // masked replay (batched diginorm): only events of reads r < r_end with
    // keep[r] & 1 are applied; rid[i] is the read of ent2 event i
    const uint32_t* rid;
    const uint8_t* keep;
    uint32_t r_end;
    uint8_t* hasev;                       // optional, per window of the sub-batch: the first pass
                                          // writes 1 for windows with events, 0 for skipped ones
    uint32_t unordered;                   // 1: a bucket's events may be in any order (plain
                                          // inserts); 0: window order (insert_and_query, diginorm)
};

// One call's view of the reads.  Starts s in [b0, b1) are processed; the
// ordinal of the k-mer at s is s - b0.  With `dyn` set, n_reads and n_bases
// are upper bounds and the actual sizes are read from device memory, so one
// captured launch sequence serves batches of any size up to the bounds.
struct Reads {
    const uint8_t* seq;
    const int64_t* off;        // n_reads + 1 entries, off[0] = 0
    const int64_t* kmer_off;   // n_reads + 1 entries (may be null if unused)
    const uint32_t* tile_read; // read holding the first base of each tile
    const int64_t* dyn;        // optional device {n_reads, n_bases}: sizes read at run time
    uint64_t n_reads;
    int64_t n_bases;
    int64_t b0, b1;
    int K;
    int mode;
    int canonical;
};

// Count-min bigcount map (bytestorage.cc:95-106, 130-137) on the device:
// open addressing over 64-bit keys, ~0 = free, the all-ones key at index
// cap; values are int16 in the low half of a u32.
enum BigAux : int { kBigOnes = 0, kBigKeys = 1, kBigAux = 2 };
struct BigMapDev {
    uint64_t* keys = nullptr;   // cap entries
    uint32_t* vals = nullptr;   // cap + 1 entries
    uint32_t* pend = nullptr;   // cap + 1: inserts of the current batch that found every table full
    uint32_t* aux = nullptr;    // [kBigOnes] the all-ones key is present, [kBigKeys] keys in the table
    uint64_t cap = 0;           // a power of two
};

}  // namespace gt
