/*
 * goetia_amd.h -- C ABI of the MI355X-native streaming de Bruijn graph hot
 * path (canonical rolling k-mer hashing feeding Bloom, count-min and exact
 * k-mer set storages).
 *
 * Every entry point is extern "C" with plain pointers and sizes.  Each one
 * names the goetia (v1.0) interface it replaces; paths are relative to the
 * goetia checkout.  The C++ facade in include/goetia/ and the Python package
 * goetia_amd bind to this header only.
 *
 * Conventions
 *   - Return value: GT_OK (0) or a GT_ERR_* code; gt_last_error() gives a
 *     thread-local message for the last failure on the calling thread.
 *   - "_dev" entry points take DEVICE pointers and enqueue work on `stream`
 *     (a hipStream_t, NULL = the null stream); they return once the work is
 *     queued.  Calls on one storage must be ordered on one stream.
 *   - Entry points without "_dev" take HOST pointers and are synchronous.
 *   - Reads are concatenated bytes `seq` with `offsets[n_reads + 1]`
 *     (offsets[0] = 0, non-decreasing, offsets[n_reads] = n_bases); read r is
 *     seq[offsets[r] .. offsets[r+1]).
 *   - Per-k-mer outputs are packed per read: read r owns max(0, len_r - K + 1)
 *     consecutive entries in window order (dbg.hh:349-394 returns the same
 *     count per call).
 *   - Input policy `mode`: GT_MODE_RAW hashes every window of the raw bytes,
 *     exactly as dBG::insert_sequence does (dbg.hh:296-305; goetia does no
 *     validation there, tests/test_dbg.py:117-123).  GT_MODE_SKIP_N skips
 *     windows holding a byte outside {A,C,G,T}, as inserting each A/C/G/T run
 *     with insert_sequence in read order would; skipped windows report count -1,
 *     hash 0 and degree 0xFF.
 *   - K is 1..255 (HashShifter's K is a uint16_t, hashshifter.hh:95).
 */
#ifndef GOETIA_AMD_H
#define GOETIA_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GT_OK                  0
#define GT_ERR_INVALID_ARG     1
#define GT_ERR_SEQUENCE_LENGTH 2   /* SequenceLengthException, kmeriterator.hh:57-59 */
#define GT_ERR_OUT_OF_MEMORY   3
#define GT_ERR_DEVICE          4
#define GT_ERR_UNSUPPORTED     5
#define GT_ERR_NO_DEVICE       6
#define GT_ERR_IO              7   /* GoetiaFileException, bitstorage.cc:186 */
#define GT_ERR_INTERNAL        8
#define GT_ERR_INVALID_CHARACTER 9 /* InvalidCharacterException, alphabets.hh:41-55 */
#define GT_ERR_PAIRING         10  /* GoetiaException of SplitPairedReader: files of different
                                      lengths, unpaired names (readers.hh:265-269, 314-317) */

#define GT_BIT_STORAGE   1         /* BitStorage, bitstorage.hh:93 */
#define GT_BYTE_STORAGE  2         /* ByteStorage, bytestorage.hh:96 */
#define GT_EXACT_STORAGE 3         /* SparseppSetStorage, sparseppstorage.hh:38 */

#define GT_MODE_RAW    0
#define GT_MODE_SKIP_N 1

typedef struct gt_storage gt_storage;

/* ------------------------------------------------------------ library */

/* Thread-local message of the last failed call on this thread. */
const char* gt_last_error(void);
/* ABI version of this header (GT_ABI_VERSION). */
#define GT_ABI_VERSION 1
int gt_abi_version(void);
/* Selects and initialises the GPU; fails with GT_ERR_NO_DEVICE without one.
 * Called implicitly by every device entry point. */
int gt_device_init(int device);
/* Waits for all work queued by this library on the device. */
int gt_device_synchronize(void);

/* storage.hh:166-190 get_n_primes_near_x: up to n odd primes strictly below
 * x, descending.  Host only. */
int gt_primes_near_x(uint32_t n, uint64_t x, uint64_t* out, uint32_t* n_out);

/* ------------------------------------------------------------ storages */

/* BitStorage(tablesizes) bitstorage.hh:114-121, ByteStorage(tablesizes)
 * bytestorage.hh:144-154, SparseppSetStorage() sparseppstorage.hh:53-56.
 * For Bloom / count-min, tablesizes[n_tables] are the bin counts (the
 * reference's (max_table, N) constructor is tablesizes =
 * gt_primes_near_x(N, max_table)).  For the exact set, tablesizes is
 * ignored and capacity_hint is the expected number of distinct k-mers (the
 * set grows as needed; 0 = a small default).  Tables live in HBM. */
int gt_storage_new(int kind, const uint64_t* tablesizes, uint32_t n_tables,
                   uint64_t capacity_hint, gt_storage** out);
/* ~BitStorage / ~ByteStorage */
int gt_storage_free(gt_storage* st);
/* clone() bitstorage.hh:136-138: a new EMPTY storage of the same sizes. */
int gt_storage_clone(const gt_storage* st, gt_storage** out);
/* reset() bitstorage.cc:141-148, bytestorage.hh:171-177,
 * sparseppstorage.hh:60-62: zero the tables.  Like the reference, the
 * n_unique / n_occupied counters of Bloom and count-min storages are left
 * as they are; the exact set's size is its count and goes to 0. */
int gt_storage_reset(gt_storage* st);
int gt_storage_kind(const gt_storage* st, int* kind);
/* n_tables() / get_tablesizes() bitstorage.hh:159-167 */
int gt_storage_n_tables(const gt_storage* st, uint32_t* n_tables);
int gt_storage_tablesizes(const gt_storage* st, uint64_t* out);
/* Bytes of table i in the reference layout: p/8+1 (Bloom,
 * bitstorage.hh:151), p (count-min, bytestorage.hh:131). */
int gt_storage_table_bytes(const gt_storage* st, uint32_t i, uint64_t* nbytes);
/* n_unique_kmers() / n_occupied(): bitstorage.hh:173-181,
 * bytestorage.hh:199-209, sparseppstorage.hh:72-86.  For the exact set
 * n_occupied is sparsepp's bucket_count() after the same inserts: 32
 * buckets, doubled before any insert (of a new or a present value) that
 * would take the size + 1 past half the buckets (spp.h:2691-2747), back to
 * 32 on reset (spp.h:2980-2989). */
int gt_storage_counters(gt_storage* st, uint64_t* n_unique, uint64_t* n_occupied);
/* Restores counters after load() (bitstorage.cc:196-307). */
int gt_storage_set_counters(gt_storage* st, uint64_t n_unique, uint64_t n_occupied);
/* Storage::set_use_bigcount / get_use_bigcount (storage.hh:129-140); count-min
 * only, else GT_ERR_UNSUPPORTED ("bigcount is not supported for this
 * storage.").  With it on, a count-min insert that finds every table at 255
 * adds one to its k-mer's bigcount entry (bytestorage.cc:95-106: 0 -> 256,
 * then +1 while the int16 value is below 65535 as an unsigned int, so 32767
 * wraps to -32768 and stays), and a query whose minimum is 255 returns the
 * entry when there is one (:130-137); insert_and_query sees each k-mer's
 * entry just after its own insert.  The filters, degree queries, captured
 * pipelines, workspace and allocation-free entry points refuse a storage with
 * bigcount on (GT_ERR_UNSUPPORTED).  reset() keeps the map, as goetia's
 * does. */
int gt_storage_set_use_bigcount(gt_storage* st, int use);
int gt_storage_get_use_bigcount(gt_storage* st, int* use);
/* The bigcount map's entries sorted by k-mer: *n their number; keys and vals
 * (cap entries each) get them when they fit.  Count-min only. */
int gt_storage_bigcounts(gt_storage* st, uint64_t* keys, int16_t* vals, uint64_t cap, uint64_t* n);
/* Replaces the bigcount map (the count-min loader, bytestorage.cc:271-285;
 * a later pair of one k-mer wins).  Count-min only. */
int gt_storage_set_bigcounts(gt_storage* st, const uint64_t* keys, const int16_t* vals, uint64_t n);
/* get_raw_tables() bitstorage.hh:228-231: copies table i (reference layout)
 * to host memory of gt_storage_table_bytes() bytes. */
int gt_storage_read_table(gt_storage* st, uint32_t i, uint8_t* host_dst);
/* Overwrites table i from host memory (the load path, bitstorage.cc:196-307). */
int gt_storage_write_table(gt_storage* st, uint32_t i, const uint8_t* host_src);
/* update_from() bitstorage.cc:103-137: Bloom union, occupied bins of table 0
 * fixed by popcount.  Bloom only. */
int gt_storage_update_from(gt_storage* st, const gt_storage* other);
/* Exact set contents (unordered) into host memory: writes min(n, cap) keys
 * and sets *n to the set size. */
int gt_storage_exact_keys(gt_storage* st, uint64_t* host_out, uint64_t cap, uint64_t* n);
/* Asynchronous copy of a Bloom / count-min storage's tables (reference
 * layout, host_tables[i] of gt_storage_table_bytes(i) bytes) and its
 * n_unique / n_occupied counters (counters[2]) to host memory on `stream`,
 * and the reverse (host_tables NULL = empty tables and zero counters).  Used
 * to evict a storage from HBM and bring it back; pinned host memory lets the
 * copies overlap other work. */
int gt_storage_save_async(gt_storage* st, uint8_t* const* host_tables, uint64_t* counters, void* stream);
int gt_storage_load_async(gt_storage* st, const uint8_t* const* host_tables, const uint64_t* counters,
                          void* stream);

/* OXLI v4 table files, as goetia writes and reads them: BitStorage::save /
 * load (bitstorage.cc:151-307, always uncompressed) and ByteStorage::save /
 * load (bytestorage.cc:638-667 through ByteStorageFile::save / load,
 * :153-167, :620-636), which write and read gzip (bytestorage.cc:304-478,
 * 538-618) when the name's text after its last '.' is "gz" and plain files
 * otherwise (:170-302, 480-536).  load replaces the storage's tables, sizes
 * included, with the file's and restores n_occupied; n_unique is kept (the
 * format does not hold it).  Count-min files carry the bigcount flag and
 * pairs: save writes the storage's (pairs sorted by k-mer; goetia writes its
 * hash map's order), load sets the flag and, when the file holds pairs,
 * replaces the map with them (bytestorage.cc:247, 271-285).  The host-only
 * readers below skip the pairs.  I/O failures are GT_ERR_IO
 * (GoetiaFileException).  The exact set has no file format
 * (sparseppstorage.cc:52-58 are empty): GT_ERR_UNSUPPORTED. */
int gt_storage_save(gt_storage* st, const char* path, uint16_t ksize);
int gt_storage_load(gt_storage* st, const char* path, uint16_t* ksize);
/* The host halves of the two above, for tables already in host memory (no
 * device): write one file of `kind` GT_BIT_STORAGE or GT_BYTE_STORAGE from
 * tables[n_tables] in the reference layout; read a file's header
 * (tablesizes has room for 16) and then its tables into caller buffers of
 * the sizes the header gave. */
int gt_oxli_write(const char* path, int kind, uint16_t ksize, uint32_t n_tables, const uint64_t* tablesizes,
                  uint64_t occupied, const uint8_t* const* tables);
int gt_oxli_read_header(const char* path, int kind, uint16_t* ksize, uint32_t* n_tables, uint64_t* tablesizes,
                        uint64_t* occupied);
int gt_oxli_read(const char* path, int kind, uint8_t* const* tables);

/* ------------------------------------------------------------ hashing */

/* Prefix-summed k-mer offsets: kmer_offsets[r] = sum over reads < r of
 * max(0, len - K + 1); n_reads + 1 entries.  Device pointers. */
int gt_kmer_offsets_dev(uint16_t K, const int64_t* offsets, uint64_t n_reads,
                        int64_t* kmer_offsets, void* stream);

/* KmerIterator<CanLemireShifter> over every read (kmeriterator.hh:109-123,
 * rollinghashshifter.hh:180-219): forward and reverse-complement cyclic hash
 * of every window.  Canonical = min(fw, rc) (canonical.hh:124-126);
 * FwdLemireShifter's value is fw.  valid (optional) flags windows kept by
 * the input policy.  Device pointers, packed per read. */
int gt_hash_reads_dev(uint16_t K, int mode, const uint8_t* seq, const int64_t* offsets,
                      uint64_t n_reads, uint64_t n_bases, uint64_t* fw, uint64_t* rc,
                      uint8_t* valid, void* stream);
int gt_hash_reads(uint16_t K, int mode, const uint8_t* seq, const int64_t* offsets,
                  uint64_t n_reads, uint64_t* fw, uint64_t* rc, uint8_t* valid);

/* Single-k-mer shifter state of HashShifter<LemireShifterPolicy<...>>, kept
 * by the caller as (fw, rc) (host scalar code, as goetia's own per-call
 * shifter API):  hash_base hashshifter.hh:147-170 /
 * rollinghashshifter.hh:180-197 (kmer holds K bytes);  shift_right(out, in)
 * rollinghashshifter.hh:200-208;  shift_left(in, out) :211-219.  The forward
 * shifter uses fw only.  Sequence-level hashing runs on the GPU
 * (gt_hash_reads*). */
int gt_shifter_hash_base(uint16_t K, const uint8_t* kmer, uint64_t* fw, uint64_t* rc);
int gt_shifter_shift_right(uint16_t K, uint64_t* fw, uint64_t* rc, uint8_t out, uint8_t in);
int gt_shifter_shift_left(uint16_t K, uint64_t* fw, uint64_t* rc, uint8_t in, uint8_t out);

/* ------------------------------------------------------------ dBG batches */

/* dBG::insert_sequence(seq, n_new) (dbg.hh:307-318) over every read in order:
 * tables end byte-identical to the sequential reference, n_new[r] is the
 * number of k-mers of read r that were new when inserted in read order
 * (bitstorage.hh:195-219, bytestorage.cc:60-113, sparseppstorage.hh:92-97),
 * and new_flags (optional, per k-mer) marks them.  n_new (optional) has
 * n_reads entries.  canonical selects CanLemireShifter (else Fwd). */
int gt_dbg_insert_dev(gt_storage* st, uint16_t K, int canonical, int mode, const uint8_t* seq,
                      const int64_t* offsets, uint64_t n_reads, uint64_t n_bases,
                      uint64_t* n_new, uint8_t* new_flags, void* stream);
int gt_dbg_insert(gt_storage* st, uint16_t K, int canonical, int mode, const uint8_t* seq,
                  const int64_t* offsets, uint64_t n_reads, uint64_t* n_new, uint8_t* new_flags,
                  uint64_t* n_kmers);
/* gt_dbg_insert_dev without any allocation: no counterpart in goetia.  For
 * callers that cannot allocate device memory (code under hipGraph capture,
 * tight serving loops, hosts that forbid hipMalloc).  Same arguments and
 * results as gt_dbg_insert_dev, for Bloom and count-min storages
 * (GT_ERR_UNSUPPORTED for the exact set, which grows by allocating); the
 * insert's scratch is one fixed array of about 98 MiB per GPU, kept for the
 * life of the process, so a call takes at most 2^30 bases (and at most 2^22
 * reads when new_flags is asked for), GT_ERR_INVALID_ARG beyond.  All calls
 * on one GPU share that array: calls that overlap on one GPU (other streams,
 * other threads) race. */
int gt_dbg_insert_noalloc_dev(gt_storage* st, uint16_t K, int canonical, int mode, const uint8_t* seq,
                              const int64_t* offsets, uint64_t n_reads, uint64_t n_bases, uint64_t* n_new,
                              uint8_t* new_flags, void* stream);

/* dBG::query_sequence(seq, counts, hashes) (dbg.hh:364-376) over every read:
 * counts (count_t = int16, storage.hh:80) and hashes (optional) packed per
 * read. */
int gt_dbg_query_dev(gt_storage* st, uint16_t K, int canonical, int mode, const uint8_t* seq,
                     const int64_t* offsets, uint64_t n_reads, uint64_t n_bases, int16_t* counts,
                     uint64_t* hashes, void* stream);
int gt_dbg_query(gt_storage* st, uint16_t K, int canonical, int mode, const uint8_t* seq,
                 const int64_t* offsets, uint64_t n_reads, int16_t* counts, uint64_t* hashes);

/* dBG::insert_and_query_sequence (dbg.hh:327-340) over every read: the
 * post-insert count of each k-mer as the sequential reference returns it
 * (1 for Bloom and exact set, bitstorage.cc:78-84, sparseppstorage.cc:21-25;
 * count-min: 1 if new else the post-insert minimum, bytestorage.cc:142-150). */
int gt_dbg_insert_and_query_dev(gt_storage* st, uint16_t K, int canonical, int mode,
                                const uint8_t* seq, const int64_t* offsets, uint64_t n_reads,
                                uint64_t n_bases, int16_t* counts, void* stream);
int gt_dbg_insert_and_query(gt_storage* st, uint16_t K, int canonical, int mode, const uint8_t* seq,
                            const int64_t* offsets, uint64_t n_reads, int16_t* counts);

/* In/out degree of every window (UnitigWalker::in_degree / out_degree,
 * unitig_walker.hh:322-330, counted as in count_nodes :361-371 with the
 * presence test query >= min_count; the reference's != 0 is min_count 1).
 * A window is a decision k-mer when in > 1 or out > 1 (is_decision_kmer
 * :1016-1023, find_decision_kmers :1025-1045).  only (optional, per k-mer)
 * restricts the work to flagged windows; others get 0xFF. */
int gt_dbg_degrees_dev(gt_storage* st, uint16_t K, int canonical, int mode, int min_count,
                       const uint8_t* seq, const int64_t* offsets, uint64_t n_reads,
                       uint64_t n_bases, const uint8_t* only, uint8_t* in_deg, uint8_t* out_deg,
                       void* stream);
int gt_dbg_degrees(gt_storage* st, uint16_t K, int canonical, int mode, int min_count,
                   const uint8_t* seq, const int64_t* offsets, uint64_t n_reads,
                   const uint8_t* only, uint8_t* in_deg, uint8_t* out_deg);

/* gt_dbg_degrees plus every window's eight neighbours, as the walker's
 * left_extensions / right_extensions give them (hashextender.hh:352-364,
 * 386-396): nb_values[8 i + c] is the value of the left neighbour with
 * symbol "ACGT"[c] of window i, nb_values[8 i + 4 + c] the right one;
 * bit c / 4 + c of nb_present[i] is set when that neighbour's query is >=
 * min_count (filter_nodes, unitig_walker.hh:407-418).  Callers that add
 * nodes of their own to the graph (filter_nodes with extras, :457-468, as
 * StreamingCompactor::find_new_segments does, compactor.hh:346-534) test the
 * values against them.  Skipped windows get zero values and mask. */
int gt_dbg_neighbors_dev(gt_storage* st, uint16_t K, int canonical, int mode, int min_count,
                         const uint8_t* seq, const int64_t* offsets, uint64_t n_reads,
                         uint64_t n_bases, uint8_t* in_deg, uint8_t* out_deg,
                         uint64_t* nb_values, uint8_t* nb_present, void* stream);
int gt_dbg_neighbors(gt_storage* st, uint16_t K, int canonical, int mode, int min_count,
                     const uint8_t* seq, const int64_t* offsets, uint64_t n_reads,
                     uint8_t* in_deg, uint8_t* out_deg, uint64_t* nb_values,
                     uint8_t* nb_present);

/* StreamingCompactor<dBG<S, FwdLemireShifter>>::Compactor::find_new_segments
 * (compactor.hh:321-534) over a batch of reads against the graph as it is
 * (nothing is inserted).  Forward hashes of raw bytes (goetia builds the
 * compactor for Hash<> only and its k-mer iterator hashes every byte), so
 * mode must be GT_MODE_RAW (GT_ERR_UNSUPPORTED otherwise).  Each read
 * starts from the new-k-mer set `known` (the one-argument overload:
 * n_known 0), or with shared != 0 one set runs through
 * the reads in order (a caller-owned set across calls).  prefix / n_prefix:
 * hashes already in the caller's hash vector, which goetia indexes from its
 * front (one read only).  Per window: counts (query) and hashes (value) out.
 * Segments of read r are segs[7 j .. 7 j + 6] for j in [seg_off[r],
 * seg_off[r + 1]): left_anchor, right_anchor, left_flank, right_flank,
 * is_decision_kmer, start_pos, length (null delimiters included, none for a
 * read shorter than K or without new k-mers).  Decision k-mers of read r are
 * decs[11 j ..] for j in [dec_off[r], dec_off[r + 1]): window, hash, a mask
 * of the neighbours present in the graph or the set (bit c left "ACGT"[c],
 * 4 + c right; get_decision_neighbors with union nodes,
 * unitig_walker.hh:1047-1082) and the eight neighbour values.  *n_segs /
 * *n_decs receive the totals; GT_ERR_INVALID_ARG when seg_cap / dec_cap are
 * smaller (call again with them). */
int gt_find_new_segments(gt_storage* st, uint16_t K, int mode, const uint8_t* seq,
                         const int64_t* offsets, uint64_t n_reads, const uint64_t* known,
                         uint64_t n_known, int shared, const uint64_t* prefix, uint64_t n_prefix,
                         int16_t* counts, uint64_t* hashes, int64_t* seg_off, uint64_t* segs,
                         uint64_t seg_cap, int64_t* dec_off, uint64_t* decs, uint64_t dec_cap,
                         uint64_t* n_segs, uint64_t* n_decs);
/* The host part of gt_find_new_segments, given each window's graph view
 * (count, hash, neighbour values and presence as gt_dbg_neighbors gives
 * them); needs no device. */
int gt_segments_from_view(uint16_t K, const int64_t* offsets, uint64_t n_reads,
                          const int16_t* counts, const uint64_t* hashes,
                          const uint64_t* nb_values, const uint8_t* nb_present,
                          const uint64_t* known, uint64_t n_known, int shared,
                          const uint64_t* prefix, uint64_t n_prefix, int64_t* seg_off,
                          uint64_t* segs, uint64_t seg_cap, int64_t* dec_off, uint64_t* decs,
                          uint64_t dec_cap, uint64_t* n_segs, uint64_t* n_decs);

/* ------------------------------------------------------------ per-hash */

/* dBG::insert(hash) / Storage::insert (dbg.hh:140-146, storage.hh:120)
 * applied to hashes[0..n) in index order; is_new (optional) per hash. */
int gt_insert_hashes_dev(gt_storage* st, const uint64_t* hashes, uint64_t n, uint8_t* is_new,
                         void* stream);
int gt_insert_hashes(gt_storage* st, const uint64_t* hashes, uint64_t n, uint8_t* is_new);
/* Storage::query (dbg.hh:171-177). */
int gt_query_hashes_dev(gt_storage* st, const uint64_t* hashes, uint64_t n, int16_t* counts,
                        void* stream);
int gt_query_hashes(gt_storage* st, const uint64_t* hashes, uint64_t n, int16_t* counts);
/* Storage::insert_and_query (dbg.hh:156-162) in index order. */
int gt_insert_and_query_hashes(gt_storage* st, const uint64_t* hashes, uint64_t n,
                               int16_t* counts);

/* ------------------------------------------------------------ tuning */

/* Tile shape of the k-mer kernels per storage kind (0 = hashing without a
 * storage): windows rolled per thread (8, 16 or 32; a workgroup of 256
 * threads owns 256 x per start positions) and persistent workgroups per CU.
 * No counterpart in goetia. */
int gt_tune_set(int kind, int per, int blocks_per_cu);
int gt_tune_get(int kind, int* per, int* blocks_per_cu);
/* Autotuner: times the storage's read-only k-mer kernels (the first-touch
 * claim pass, or the exact-set lookup pass, plus query) on a batch of reads
 * in device memory for every tile shape, keeps the fastest for the storage's
 * kind and returns it.  The storage's contents are not changed. */
int gt_autotune(gt_storage* st, uint16_t K, int canonical, int mode, const uint8_t* seq,
                const int64_t* offsets, uint64_t n_reads, uint64_t n_bases, int* per,
                int* blocks_per_cu, void* stream);

/* ------------------------------------------------------------ pipeline */

/* Batched insert engine (no counterpart in goetia; it replaces calling
 * dBG::insert_sequence read by read, processors.hh:304-331).  Batches of at
 * most max_reads reads / max_bases bases arrive in host memory; each is
 * copied into one of two pinned + device buffer pairs on a copy stream while
 * the previous batch runs, and its device work (read layout, one first-touch
 * sub-batch, insert, per-read n_new) is a hipGraph captured once per buffer
 * and replayed for every batch (batch sizes are read from device memory).
 * Results equal gt_dbg_insert over the same reads in submission order.
 * gt_pipeline_submit returns once the batch is copied to pinned memory;
 * n_new (host, n_reads entries, may be NULL) is filled when the batch
 * completes, at the latest by gt_pipeline_flush.  max_bases must fit one
 * first-touch sub-batch (at most 2^26 start positions for 4 tables). */
typedef struct gt_pipeline gt_pipeline;
int gt_pipeline_new(gt_storage* st, uint16_t K, int canonical, int mode, uint64_t max_reads,
                    uint64_t max_bases, gt_pipeline** out);
int gt_pipeline_submit(gt_pipeline* p, const uint8_t* seq, const int64_t* offsets, uint64_t n_reads,
                       uint64_t* n_new);
int gt_pipeline_flush(gt_pipeline* p);
int gt_pipeline_free(gt_pipeline* p);

/* ------------------------------------------------------------ workspaces */

/* Device scratch (first-touch claim map, fresh masks, read layout) that many
 * storages can share, one call at a time (no counterpart in goetia; the
 * serving engine gives each of its concurrent slots one, so resident
 * storages hold only their tables).  Calls through one workspace are
 * serialised; use one workspace per concurrent stream. */
typedef struct gt_workspace gt_workspace;
int gt_workspace_new(gt_workspace** out);
int gt_workspace_free(gt_workspace* ws);
/* gt_dbg_insert_dev / gt_dbg_query_dev with the workspace's scratch. */
int gt_dbg_insert_ws_dev(gt_storage* st, gt_workspace* ws, uint16_t K, int canonical, int mode,
                         const uint8_t* seq, const int64_t* offsets, uint64_t n_reads, uint64_t n_bases,
                         uint64_t* n_new, uint8_t* new_flags, void* stream);
int gt_dbg_query_ws_dev(gt_storage* st, gt_workspace* ws, uint16_t K, int canonical, int mode,
                        const uint8_t* seq, const int64_t* offsets, uint64_t n_reads, uint64_t n_bases,
                        int16_t* counts, uint64_t* hashes, void* stream);

/* ------------------------------------------------------------ serving */

/* Serving engine (no counterpart in goetia; it serves many sample streams,
 * each its own dBG, as goetia's per-sample processors would one at a time).
 * All streams share one storage shape (Bloom or count-min tables of
 * `tablesizes`); floor(hbm_budget / tables bytes) storages stay resident in
 * HBM and streams are bound to them on demand, evicting the least recently
 * used idle stream to pinned host memory.  A scheduler thread keeps n_slots
 * chunks in flight (continuous batching): a chunk starts as soon as a slot is
 * free and its stream is idle and resident, chunks of one stream run in
 * submission order, and each stream's results equal inserting its chunks
 * alone.  policy bit 0 starts runnable chunks of resident streams before
 * older chunks that need a table load (a chunk waiting over 0.25 s goes
 * first); policy 0 with n_slots = 1 is the plain one-chunk-at-a-time path. */
typedef struct gt_server gt_server;
typedef struct gt_server_stats {
    uint64_t submitted, completed, loads, evictions, kmers, resident_slots, bytes_per_stream;
    double latency_sum_s, elapsed_s;
} gt_server_stats;
const char* gt_server_last_error(void);
int gt_server_new(int kind, const uint64_t* tablesizes, uint32_t n_tables, uint16_t K, int canonical, int mode,
                  uint32_t n_streams, uint64_t hbm_budget, uint32_t n_slots, int policy, gt_server** out);
/* Queues a chunk of host reads (copied before returning) for `stream`;
 * n_new (host, n_reads entries, may be NULL) receives dBG::insert_sequence's
 * per-read n_new when the chunk completes. */
int gt_server_submit(gt_server* S, uint32_t stream, const uint8_t* seq, const int64_t* offsets,
                     uint64_t n_reads, uint64_t* n_new, uint64_t* ticket);
/* Blocks until the chunk completes; latency = completion - submission. */
int gt_server_wait(gt_server* S, uint64_t ticket, double* latency_s);
int gt_server_drain(gt_server* S);
int gt_server_get_stats(gt_server* S, gt_server_stats* out);
/* Table i and the counters of a stream, wherever they live (drains first). */
int gt_server_stream_table(gt_server* S, uint32_t stream, uint32_t i, uint8_t* host_dst, uint64_t* counters);
int gt_server_free(gt_server* S);

/* ------------------------------------------------------------ timing */

/* Kernel timing (no counterpart in goetia): while enabled, every kernel this
 * library launches is bracketed by a pair of hipEvents recorded on the
 * kernel's own stream; gt_profile_get synchronizes and returns, for one
 * kernel name ("claim", "apply", "exact_lookup", "exact_insert",
 * "exact_resolve", "query", "hash", "degrees", "seq_insert_query", the binned
 * insert's "bin_hash", "bin_count", "bin_route", "bin_apply", "bin_new",
 * "bin_iq", diginorm's "dn_replay", "dn_decide", "dn_apply"), the number of
 * launches and their summed duration in milliseconds since the last reset. */
int gt_profile_enable(int on);
int gt_profile_reset(void);
int gt_profile_get(const char* kernel, uint64_t* launches, double* total_ms);

/* Random-access roofline (no counterpart in goetia; measurement only):
 * at least n_ops uniform random 32-bit accesses to the device buffer
 * buf[0, n_words) from every CU, asynchronously on `stream`: kind 0 reads,
 * 1 atomicOr of one bit, 2 atomicCAS saturating increment of the low byte,
 * 3 reads with the non-temporal cache policy.
 * *ops_done receives the number issued.  The buffer's contents change for
 * kinds 1 and 2. */
int gt_membench_random(void* buf, uint64_t n_words, uint64_t n_ops, int kind, uint64_t seed, void* stream,
                       uint64_t* ops_done);

/* ------------------------------------------------------------ consumers */

/* StreamingSolidFilter<dBG<S, Sh>>::Filter::filter_sequence (solidifier.hh:58-73)
 * over a batch: each read's k-mers are inserted and queried in read order,
 * exactly as insert_and_query_sequence (dbg.hh:327-340) would one read after
 * the other, and read r passes (passed[r] = 1) unless
 * (float)n_not_solid / (float)n_kmers >= 1.0 - min_prop_solid, where a
 * k-mer is not solid when its count is below solid_threshold.  Reads shorter
 * than K do not pass (FilterProcessor::process_sequence skips them,
 * processors.hh:389-401).  goetia's defaults: min_prop_solid 0.75,
 * solid_threshold 1. */
int gt_solid_filter_dev(gt_storage* st, uint16_t K, int canonical, int mode, const uint8_t* seq,
                        const int64_t* offsets, uint64_t n_reads, uint64_t n_bases, float min_prop_solid,
                        uint32_t solid_threshold, uint8_t* passed, void* stream);
int gt_solid_filter(gt_storage* st, uint16_t K, int canonical, int mode, const uint8_t* seq,
                    const int64_t* offsets, uint64_t n_reads, float min_prop_solid, uint32_t solid_threshold,
                    uint8_t* passed);

/* goetia's DiginormFilter<dBG<ByteStorage, Sh>>::Filter::filter_sequence
 * (diginorm.hh:105-119) over a batch, in read order: a read is dropped
 * (passed = 0) when median_count_at_least (diginorm.hh:39-76) holds -- at
 * least min_req = (unsigned)(0.5 + float(len - K + 1) / 2) of its k-mers have
 * a count >= cutoff -- and otherwise inserted (passed = 1) before the next read
 * is looked at.  Results, tables and counters equal filtering the reads one
 * after another.  Reads shorter than K do not pass and are not inserted
 * (FilterProcessor::process_sequence skips them, processors.hh:389-401).
 * Skip-N mode: windows holding a non-ACGT byte are neither counted nor
 * inserted (min_req still counts them).  Count-min tables only, at most 4
 * tables and 2^30 bins in all (GT_ERR_UNSUPPORTED otherwise).  The device
 * variant synchronises `stream` (the offsets come to the host once, and each
 * chunk of reads is resolved in a few passes). */
int gt_diginorm_dev(gt_storage* st, uint16_t K, int canonical, int mode, const uint8_t* seq, const int64_t* offsets,
                    uint64_t n_reads, uint64_t n_bases, uint32_t cutoff, uint8_t* passed, void* stream);
int gt_diginorm(gt_storage* st, uint16_t K, int canonical, int mode, const uint8_t* seq, const int64_t* offsets,
                uint64_t n_reads, uint32_t cutoff, uint8_t* passed);

/* ------------------------------------------------------------ FASTX input */

/* goetia's FastxParser<DNA_SIMPLE> (readers.hh:95-232; construction
 * readers.cc:14-33): FASTA / FASTQ, plain or gzip, parsed record for record
 * as the vendored kseq.h does, each sequence validated by DNA_SIMPLE
 * (alphabets.hh:112-130: a/c/g/t uppercased, any other byte outside A/C/G/T
 * rejects the record).  Rejected records are skipped and counted (n_skipped);
 * with strict set the first one ends the call with GT_ERR_INVALID_CHARACTER
 * (InvalidCharacterException) after the records before it.  Records shorter
 * than min_length are skipped.  A FASTQ record whose quality length differs
 * from its sequence (InvalidRead, readers.hh:205-208) is skipped and counted
 * in n_invalid_reads (and n_skipped, as goetia counts it).  Host only. */
typedef struct gt_fastx gt_fastx;
int gt_fastx_open(const char* path, int strict, uint32_t min_length, gt_fastx** out);
/* Fills one batch in the layout of every gt_dbg_* call: up to max_reads
 * records and max_bases bases (seq, offsets with n + 1 entries), optionally
 * their names (names / name_offsets, names_cap bytes).  *n_reads = 0 when the
 * input is exhausted.  A record larger than the buffers fails with
 * GT_ERR_INVALID_ARG and is kept for the next call (FastxParser::next,
 * readers.hh:148-219, one batch at a time). */
int gt_fastx_next_batch(gt_fastx* p, uint64_t max_reads, uint64_t max_bases, uint8_t* seq, int64_t* offsets,
                        char* names, uint64_t names_cap, int64_t* name_offsets, uint64_t* n_reads);
/* gt_fastx_next_batch with each record's quality (Record::quality,
 * readers.hh:186-191): qual has the layout of seq (a FASTQ record's quality
 * has its sequence's length), has_qual[i] = 1 when record i has one (FASTQ
 * with a non-empty sequence), 0 otherwise (FASTA).  qual and has_qual are
 * both given or both null. */
int gt_fastx_next_batch_qual(gt_fastx* p, uint64_t max_reads, uint64_t max_bases, uint8_t* seq, int64_t* offsets,
                             uint8_t* qual, uint8_t* has_qual, char* names, uint64_t names_cap,
                             int64_t* name_offsets, uint64_t* n_reads);
/* n_parsed / n_skipped (readers.hh:221-227), invalid reads, is_complete. */
int gt_fastx_stats(const gt_fastx* p, uint64_t* n_parsed, uint64_t* n_skipped, uint64_t* n_invalid_reads,
                   int* complete);
int gt_fastx_close(gt_fastx* p);

/* InserterProcessor<dBG<...>> (processors.hh:248-344) driving a FileProcessor
 * (processors.hh:45-236) with an IntervalCounter (metrics.hh:101-145):
 * gt_inserter_advance inserts parsed reads into the storage, batched through
 * gt_dbg_insert, and returns right after the read whose k-mers (len - K + 1,
 * 0 below K) bring the interval counter to `interval`, as FileProcessor::advance
 * does: *n_sequences and *time_total are the running totals and *remaining is
 * 0 once the parser is exhausted.  At an interval the call returns with the
 * reads taken since the last batch held by the inserter: they are inserted
 * with its next batch, or before any other call on the storage (every entry
 * point inserts them first), so what the storage shows is exactly the reads
 * up to the interval; at the end of the input every insert is finished.  The
 * storage must outlive the inserter (gt_inserter_free inserts held reads).  Records with invalid characters are
 * skipped even from a strict parser (FileProcessor::handle_next,
 * processors.hh:149-171).  batch_bases sizes the host batch (0: 128 MiB). */
typedef struct gt_inserter gt_inserter;
int gt_inserter_new(gt_storage* st, uint16_t K, int canonical, int mode, uint64_t interval, uint64_t batch_bases,
                    gt_inserter** out);
int gt_inserter_advance(gt_inserter* ins, gt_fastx* parser, uint64_t* n_sequences, uint64_t* time_total,
                        int* remaining);
/* Running n_sequences, time (k-mers) and the interval counter. */
int gt_inserter_stats(const gt_inserter* ins, uint64_t* n_sequences, uint64_t* time_total, uint64_t* counter);

/* SplitPairedReader<FastxParser<DNA_SIMPLE>> (readers.hh:235-334) over two
 * files read record by record: each step takes one record attempt from each
 * file.  A side whose record is rejected (invalid characters from a
 * non-strict parser, shorter than min_length, a quality of another length) is
 * empty; with strict set, a raising side (an invalid character, a quality of
 * another length) drops the whole pair (FileProcessor::handle_next,
 * processors.hh:149-171).  With force_name_match, two present records whose
 * names are not a pair (check_is_pair, parsing.cc:56-87) drop the pair and
 * count two skipped, or, strict, fail with GT_ERR_PAIRING ("Unpaired reads").
 * Files ending at different steps fail with GT_ERR_PAIRING ("Mismatched
 * split paired files.") at the next step.  Host only. */
typedef struct gt_pairs gt_pairs;
int gt_pairs_open(const char* left, const char* right, int strict, uint32_t min_length, int force_name_match,
                  gt_pairs** out);
/* SplitPairedReader::next: the pair's sequences, left then right, into seq
 * (seq_cap bytes), lens[2] their lengths (-1 for an empty side) and, when
 * names is given, their names likewise (name_lens[2]); *done = 1 once both
 * files are exhausted.  A pair dropped by a strict reader fails with
 * GT_ERR_INVALID_CHARACTER. */
int gt_pairs_next(gt_pairs* pr, uint8_t* seq, uint64_t seq_cap, int64_t* lens, char* names, uint64_t names_cap,
                  int64_t* name_lens, int* done);
/* n_parsed (both files), n_skipped (both files and unpaired reads),
 * is_complete (readers.hh:262-270, 327-333). */
int gt_pairs_stats(const gt_pairs* pr, uint64_t* n_parsed, uint64_t* n_skipped, int* complete);
int gt_pairs_close(gt_pairs* pr);
/* FileProcessor::advance over a SplitPairedReader for an InserterProcessor
 * (processors.hh:180-204, 304-306): each pair's present reads are inserted,
 * left then right; n_sequences counts present reads and the interval counter
 * is polled once per pair with both reads' k-mers, so the call returns right
 * after the pair that reaches `interval` (every insert finished), or at the
 * end of both files with *remaining = 0.  An inserter drives either single
 * files (gt_inserter_advance) or pairs. */
int gt_inserter_advance_pairs(gt_inserter* ins, gt_pairs* pr, uint64_t* n_sequences, uint64_t* time_total,
                              int* remaining);
int gt_inserter_free(gt_inserter* ins);

/* ------------------------------------------------------------ workload */

/* Seeded synthetic workload (host memory, deterministic for a seed; no
 * counterpart in goetia, whose tests draw random sequences from an external
 * fixture package).  gt_gen_genome: uniform A/C/G/T bases from xoshiro256**
 * seeded through splitmix64.  gt_gen_reads: read r draws its own generator
 * from (seed, r); its length is lengths[r] when lengths is not NULL, else
 * uniform in [len_min, len_max]; it starts uniformly in the genome, is
 * reverse complemented with probability 1/2, and each base is substituted
 * by a different base with probability sub_rate, then replaced by 'N' with
 * probability n_rate.  seq must hold the sum of the lengths (at most
 * n_reads * len_max); offsets has n_reads + 1 entries; *n_bases receives the
 * total.  Reads longer than the genome are clipped to it. */
int gt_gen_genome(uint64_t seed, uint64_t length, uint8_t* out);
int gt_gen_reads(uint64_t seed, const uint8_t* genome, uint64_t genome_len, uint64_t n_reads,
                 const uint32_t* lengths, uint32_t len_min, uint32_t len_max, double sub_rate,
                 double n_rate, uint8_t* seq, int64_t* offsets, uint64_t* n_bases);

#ifdef __cplusplus
}
#endif

#endif /* GOETIA_AMD_H */
