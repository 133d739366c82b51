// gt_launch.h -- host-side launchers of the kernels in gt_kernels.hip.
#pragma once

#include <vector>

#include "gt_internal.h"

namespace gt {

struct Launch {
    int grid;              // workgroups for tile walkers (persistent grid)
    hipStream_t stream;
    int per;               // windows per thread: 8, 16 or 32
};

void launch_tile_read(const int64_t* off, uint64_t n_reads, int64_t n_bases, uint32_t* tile_read,
                      int64_t n_tiles, const int64_t* dyn, hipStream_t s);
void launch_kmer_offsets(const int64_t* off, uint64_t n_reads, int K, int64_t* kmer_off,
                         int64_t* block_sums, const int64_t* dyn, hipStream_t s);
void launch_bump_gen(uint64_t* storage_gen, uint64_t* claim_gen, hipStream_t s);
uint64_t scan_block_count(uint64_t n_reads);
void launch_fill_u64(uint64_t* p, uint64_t n, uint64_t v, hipStream_t s);

void launch_claim(const Launch& L, const Reads& R, const Tables& T, const ClaimMap& cm, uint8_t* fresh,
                  uint64_t* counters);
void launch_apply(const Launch& L, const Reads& R, const Tables& T, const ClaimMap& cm, const uint8_t* fresh,
                  uint64_t* n_new, uint8_t* new_flags, uint64_t* counters);
void launch_exact_lookup(const Launch& L, const Reads& R, const ExactSet& E, uint8_t* fresh, uint64_t* counters);
void launch_exact_insert(const Launch& L, const Reads& R, const ExactSet& E, const uint8_t* fresh,
                         uint64_t* counters);
void launch_exact_tail(uint64_t* counters, hipStream_t s);
void launch_exact_resolve(const Launch& L, const Reads& R, const ExactSet& E, const uint8_t* fresh,
                          uint64_t* n_new, uint8_t* new_flags, uint64_t* counters);
void launch_query(const Launch& L, const Reads& R, const Tables& T, const ExactSet& E, int16_t* counts,
                  uint64_t* hashes);
void launch_hash(const Launch& L, const Reads& R, uint64_t* fw, uint64_t* rc, uint8_t* valid);
void launch_degrees(const Launch& L, const Reads& R, const Tables& T, const ExactSet& E, int min_count,
                    const uint8_t* only, uint8_t* in_deg, uint8_t* out_deg, uint64_t* nb_val, uint8_t* nb_mask);

void launch_hash_claim(const Launch& L, const uint64_t* h, uint64_t n, const Tables& T, const ClaimMap& cm,
                       uint8_t* fresh, uint64_t* counters);
void launch_hash_apply(const Launch& L, const uint64_t* h, uint64_t n, const Tables& T, const ClaimMap& cm,
                       const uint8_t* fresh, uint8_t* is_new, uint64_t* counters);
void launch_hash_exact_lookup(const Launch& L, const uint64_t* h, uint64_t n, const ExactSet& E, uint8_t* fresh,
                              uint64_t* counters);
void launch_hash_exact_insert(const Launch& L, const uint64_t* h, uint64_t n, const ExactSet& E,
                              const uint8_t* fresh, uint64_t* counters);
void launch_hash_exact_resolve(const Launch& L, const uint64_t* h, uint64_t n, const ExactSet& E,
                               const uint8_t* fresh, uint8_t* is_new, uint64_t* counters);
void launch_hash_query(const Launch& L, const uint64_t* h, uint64_t n, const Tables& T, const ExactSet& E,
                       int16_t* counts);

void launch_bit_union(uint32_t* me, const uint32_t* other, uint64_t n_words, int count_bits, uint64_t* counters,
                      hipStream_t s);
void launch_exact_rehash(const ExactSet& from, const ExactSet& to, uint64_t* counters, hipStream_t s);
// Count-min insert_and_query of n values in index order (fw[k], or
// min(fw, rc) when rc is set and canonical), sort-based (gt_sorted_iq.hip);
// full (optional) gets 1 for the values that found every table at 255.
// Synchronises the stream before it returns.
int sorted_insert_query(const Tables& T, const uint64_t* fw, const uint64_t* rc, const uint8_t* valid, uint64_t n,
                        int canonical, int16_t* counts, uint8_t* is_new, uint64_t* counters, int cus,
                        hipStream_t s, uint8_t* full = nullptr);

// The count-min bigcount map (gt_bigcount.hip).  Window values as above.
int big_init(BigMapDev& M, hipStream_t s);
void big_free(BigMapDev& M);
int big_insert(BigMapDev& M, const uint64_t* fw, const uint64_t* rc, int canonical, const uint8_t* full, uint64_t n,
               int cus, hipStream_t s);
int big_query(const BigMapDev& M, const uint64_t* fw, const uint64_t* rc, const uint8_t* valid, int canonical,
              int16_t* counts, uint64_t n, int cus, hipStream_t s);
int big_insert_query(const BigMapDev& M, const uint64_t* fw, const uint64_t* rc, const uint8_t* valid, int canonical,
                     const uint8_t* full, int16_t* counts, uint64_t n, int cus, hipStream_t s);
int big_pairs(const BigMapDev& M, std::vector<uint64_t>& keys, std::vector<int16_t>& vals, hipStream_t s);
int big_set_pairs(BigMapDev& M, const uint64_t* keys, const int16_t* vals, uint64_t n, hipStream_t s);
void launch_exact_keys(const ExactSet& E, uint64_t* out, uint64_t cap_out, unsigned long long* n, hipStream_t s);

// Binned insert of Bloom / count-min tables (one sub-batch = kbase, l1,
// l2_count, scan, l2_scatter, apply; read_new once per call).
void launch_bin_kbase(const Reads& R, int64_t b0, uint64_t* out, hipStream_t s);
void launch_bin_l1(const Reads& R, const Tables& T, const BinPlan& P, int grid, hipStream_t s);
void launch_bin_l2_count(const BinPlan& P, int grid, hipStream_t s);
void launch_bin_l2_scatter(const BinPlan& P, int grid, hipStream_t s);
void launch_bin_apply(const BinPlan& P, const Tables& T, uint64_t* counters, int grid, hipStream_t s);
void launch_bin_read_new(const int64_t* kmer_off, uint64_t n_reads, const uint8_t* flags, uint64_t* n_new,
                         uint64_t* counters, hipStream_t s);
void launch_zero_bytes(uint8_t* p, const int64_t* n, uint64_t max_n, hipStream_t s);
void launch_zero_words(uint32_t* p, const int64_t* n, uint64_t max_n, hipStream_t s);
void launch_solid_reads(const int64_t* off, const int64_t* kmer_off, uint64_t n_reads, int K, const int16_t* counts,
                        uint32_t threshold, float min_prop_solid, uint8_t* passed, hipStream_t s);
void launch_seg_gather(const uint64_t* nb, const uint8_t* mask, const uint64_t* idx, uint64_t n, uint64_t* out_nb,
                       uint8_t* out_mask, hipStream_t s);
void launch_bin_iq_combine(const uint32_t* vq, const int64_t* nk, uint64_t max_n, uint32_t n_tables, int16_t* counts,
                           uint64_t* counters, hipStream_t s);
void launch_fill_u8(uint8_t* p, uint64_t n, uint8_t v, hipStream_t s);

// Batched diginorm over one chunk of reads [r0, r1) whose events are routed
// (init, event_reads once; replay, ctl_reset, decide per guess; apply of the
// kept reads and passed at the end).
void launch_dn_init(const int64_t* off, const int64_t* kmer_off, uint32_t r0, uint32_t r1, int K, uint32_t* wread,
                    uint8_t* state, hipStream_t s);
void launch_dn_event_reads(const BinPlan& P, const uint32_t* wread, uint32_t* rid, uint64_t max_events,
                           hipStream_t s);
void launch_dn_replay(const BinPlan& P, const Tables& T, const uint32_t* rid, const uint8_t* state, uint32_t cutoff,
                      uint8_t* fails, int grid, hipStream_t s);
void launch_dn_decide(const int64_t* off, const int64_t* kmer_off, uint32_t r0, uint32_t r1, int K,
                      const uint8_t* hasev, const uint8_t* fails, int upper, uint8_t* state, uint32_t* ctl,
                      hipStream_t s);
void launch_dn_ctl_reset(uint32_t* ctl, hipStream_t s);
void launch_dn_passed(const uint8_t* state, uint32_t r0, uint32_t r1, uint8_t* passed, hipStream_t s);
void launch_dn_apply(const BinPlan& P, const Tables& T, uint64_t* counters, int grid, hipStream_t s);
uint64_t scan_u32_block_count(uint64_t n);
void launch_scan_u32(uint32_t* a, uint64_t n, uint32_t* sums, hipStream_t s);

}  // namespace gt
