/*
 * oracle.h -- CPU restatement of goetia's streaming dBG hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker that the HIP product
 * path is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product library never links
 * or calls it.
 *
 * Every function restates a reference function sequentially, in the order the
 * reference executes it, citing the reference file:line it follows (paths are
 * relative to the goetia v1.0 checkout).
 *
 * Parity pins (what checks each function besides retyping its formula):
 *   or_hash_kmer / or_hash_read : known-answer vector of tests/test_hashing.py:15-18,
 *       rolling == direct on every window, canonical == min(fw(x), fw(revcomp x)),
 *       and golden vectors written by the compiled reference (oracle/_ref) for all
 *       256 byte values.
 *   or_primes_near_x            : brute-force primality on small x, golden values.
 *   Bit/Byte/exact storages     : brute-force closed forms on small inputs
 *       (Bloom bit = OR of events, CMS byte = min(255, #events)), n_unique vs
 *       python set size, golden table digests/n_new written by oracle/_ref.
 *   or_degrees                  : neighbour hashes vs direct hashing of the
 *       neighbour strings, linear-path and fork degrees, golden decision sets
 *       from oracle/_ref.
 *   or_diginorm_*               : golden pass flags, tables and counters written
 *       by oracle/_ref's DiginormFilter (tests/golden/diginorm_ref.npz); identical
 *       reads pass exactly until cutoff copies are in; the early return of
 *       median_count_at_least equals the plain count of k-mers at the cutoff.
 */
#ifndef GOETIA_ORACLE_H
#define GOETIA_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Input policies for sequence-level calls. */
#define OR_MODE_RAW    0  /* hash every window of the raw bytes (dbg.hh:296-305) */
#define OR_MODE_SKIP_N 1  /* windows holding a byte outside {A,C,G,T} are skipped */

#define OR_BIT   1
#define OR_BYTE  2
#define OR_EXACT 3

#define OR_MAX_TABLES 16

typedef struct or_storage {
    int       kind;
    uint32_t  n_tables;
    uint64_t  sizes[OR_MAX_TABLES];
    uint8_t*  tables[OR_MAX_TABLES];
    uint64_t  n_unique;
    uint64_t  n_occupied;
    /* exact set: open addressing over keys, with a separate flag for the
     * all-ones key (which doubles as the empty marker). */
    uint64_t* keys;
    uint64_t  capacity;
    uint64_t  count;
    int       has_ones_key;
    /* count-min bigcount map (bytestorage.cc:95-106, 130-137): k-mer hash ->
     * int16 count, open addressing like the exact set's keys. */
    int       use_bigcount;
    uint64_t* big_keys;
    int16_t*  big_vals;
    uint64_t  big_cap;
    uint64_t  big_count;
    int       big_has_ones;
    int16_t   big_ones_val;
} or_storage;

uint64_t or_table_value(int c);
uint64_t or_rotl(uint64_t x, unsigned r);
uint64_t or_rotr(uint64_t x, unsigned r);
char     or_complement(char c);

void     or_hash_kmer(const char* s, int K, uint64_t* fw, uint64_t* rc);
size_t   or_hash_read(const char* s, size_t len, int K, uint64_t* fw, uint64_t* rc);

uint32_t or_primes_near_x(uint32_t n, uint64_t x, uint64_t* out);

int      or_storage_init(or_storage* st, int kind, const uint64_t* sizes, uint32_t n_tables);
void     or_storage_free(or_storage* st);
uint64_t or_table_bytes(const or_storage* st, uint32_t i);

int      or_insert(or_storage* st, uint64_t h);
void     or_set_use_bigcount(or_storage* st, int b);
int      or_bigcount_get(const or_storage* st, uint64_t h, int16_t* v);
void     or_bigcount_set(or_storage* st, uint64_t h, int16_t v);
uint64_t or_bigcount_pairs(const or_storage* st, uint64_t* keys, int16_t* vals, uint64_t cap);
int16_t  or_query(const or_storage* st, uint64_t h);
int16_t  or_insert_and_query(or_storage* st, uint64_t h);

/* Sequence level: canonical != 0 selects CanLemireShifter, else FwdLemireShifter.
 * Per-position outputs have len-K+1 entries; skipped windows get -1 counts
 * and 0 hashes / flags.  Return the number of k-mers hashed (windows not
 * skipped). */
uint64_t or_insert_sequence(or_storage* st, int K, int canonical, int mode,
                            const char* s, size_t len, uint64_t* n_new, uint8_t* new_flags);
uint64_t or_insert_and_query_sequence(or_storage* st, int K, int canonical, int mode,
                                      const char* s, size_t len, int16_t* counts);
uint64_t or_query_sequence(const or_storage* st, int K, int canonical, int mode,
                           const char* s, size_t len, int16_t* counts, uint64_t* hashes);
uint64_t or_degrees_sequence(const or_storage* st, int K, int canonical, int mode,
                             int min_count, const char* s, size_t len,
                             uint8_t* in_deg, uint8_t* out_deg);
void     or_kmer_degrees(const or_storage* st, int K, int canonical, int min_count,
                         const char* kmer, uint8_t* in_deg, uint8_t* out_deg,
                         uint64_t left[4], uint64_t right[4]);

/* Batched drivers over concatenated reads (offsets has n_reads+1 entries). */
uint64_t or_insert_reads(or_storage* st, int K, int canonical, int mode,
                         const char* seq, const int64_t* offsets, uint64_t n_reads,
                         uint64_t* n_new);
uint64_t or_query_reads(const or_storage* st, int K, int canonical, int mode,
                        const char* seq, const int64_t* offsets, uint64_t n_reads,
                        int16_t* counts);

uint64_t or_exact_keys(const or_storage* st, uint64_t* out, uint64_t cap);

/* The same drivers on n_threads host threads with atomic table updates
 * (Bloom / count-min; counters not updated by the insert). */
uint64_t or_insert_reads_mt(or_storage* st, int K, int canonical, int mode, const char* seq,
                            const int64_t* offsets, uint64_t n_reads, int n_threads);
uint64_t or_query_reads_mt(const or_storage* st, int K, int canonical, int mode, const char* seq,
                           const int64_t* offsets, uint64_t n_reads, int16_t* counts, int n_threads);

/* DiginormFilter (count-min graphs): one read, then reads in order. */
int      or_diginorm_filter(or_storage* st, int K, int canonical, int mode, unsigned cutoff,
                            const char* s, size_t len);
uint64_t or_diginorm_reads(or_storage* st, int K, int canonical, int mode, unsigned cutoff,
                           const char* seq, const int64_t* offsets, uint64_t n_reads, uint8_t* passed);

uint64_t or_insert_reads_flags(or_storage* st, int K, int canonical, int mode,
                               const char* seq, const int64_t* offsets, uint64_t n_reads,
                               uint64_t* n_new, uint8_t* new_flags);
uint64_t or_insert_and_query_reads(or_storage* st, int K, int canonical, int mode,
                                   const char* seq, const int64_t* offsets, uint64_t n_reads,
                                   int16_t* counts);
uint64_t or_hash_reads(int K, int canonical, int mode, const char* seq, const int64_t* offsets,
                       uint64_t n_reads, uint64_t* out);
uint64_t or_degrees_reads_mt(const or_storage* st, int K, int canonical, int mode, int min_count,
                             const char* seq, const int64_t* offsets, uint64_t n_reads, const uint8_t* only,
                             uint8_t* in_deg, uint8_t* out_deg, int n_threads);

#ifdef __cplusplus
}
#endif

#endif
