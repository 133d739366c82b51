// ref_driver.cc -- extern "C" entry points over the compiled goetia reference
// hot path, used only to cross-check the C restatement (oracle.c) and to write
// the golden fixtures under tests/golden/.  Built by oracle/Makefile from the
// reference sources where they lie; the output stays in oracle/_ref/.
//
// TEST INFRASTRUCTURE ONLY: nothing in the product loads this library.

#include "goetia/dbg.hh"
#include "goetia/storage/bitstorage.hh"
#include "goetia/storage/bytestorage.hh"
#include "goetia/storage/sparseppstorage.hh"
#include "goetia/storage/phmapstorage.hh"
#include "goetia/storage/btreestorage.hh"
#include "goetia/hashing/hashshifter.hh"
#include "goetia/hashing/kmeriterator.hh"
#include "goetia/parsing/readers.hh"
#include "goetia/diginorm.hh"
#include "goetia/solidifier.hh"
#include "goetia/cdbg/compactor.hh"

#include <cstdint>
#include <cstring>
#include <type_traits>
#include <memory>
#include <string>
#include <vector>

using namespace goetia;

namespace {

template <class G>
struct Holder {
    std::shared_ptr<G> g;
};

struct Handle {
    int kind;       // 1 bit, 2 byte, 3 sparsepp set
    int canonical;
    void* graph;    // Holder<dBG<...>>*
    std::shared_ptr<ByteStorage> bytes;   // the count-min storage (kind 2)
};

template <class S, class H>
using Graph = dBG<S, H>;

bool acgt(char c) { return c == 'A' || c == 'C' || c == 'G' || c == 'T'; }

// Splits a read into the pieces the input policy hands to insert_sequence:
// mode 0 = the whole read, mode 1 = every maximal A/C/G/T run of length >= K
// (with its start position).
std::vector<std::pair<size_t, std::string>> pieces(const char* s, size_t len, int K, int mode) {
    std::vector<std::pair<size_t, std::string>> out;
    if (mode == 0) {
        if (len >= (size_t)K) out.emplace_back(0, std::string(s, len));
        return out;
    }
    size_t i = 0;
    while (i < len) {
        while (i < len && !acgt(s[i])) ++i;
        size_t j = i;
        while (j < len && acgt(s[j])) ++j;
        if (j - i >= (size_t)K) out.emplace_back(i, std::string(s + i, j - i));
        i = j;
    }
    return out;
}

template <class G>
G* get(Handle* h) { return static_cast<Holder<G>*>(h->graph)->g.get(); }

template <class F>
auto dispatch(Handle* h, F&& f) {
    if (h->kind == 1) {
        if (h->canonical) return f(get<Graph<BitStorage, CanLemireShifter>>(h));
        return f(get<Graph<BitStorage, FwdLemireShifter>>(h));
    } else if (h->kind == 2) {
        if (h->canonical) return f(get<Graph<ByteStorage, CanLemireShifter>>(h));
        return f(get<Graph<ByteStorage, FwdLemireShifter>>(h));
    }
    if (h->canonical) return f(get<Graph<SparseppSetStorage, CanLemireShifter>>(h));
    return f(get<Graph<SparseppSetStorage, FwdLemireShifter>>(h));
}

template <class S, class H>
void* make_graph(std::shared_ptr<S> store, int K) {
    auto* holder = new Holder<Graph<S, H>>();
    holder->g = Graph<S, H>::build(store, (uint16_t)K);
    return holder;
}

// DiginormFilter<G>::Filter::filter_sequence over reads in order; a read
// shorter than K raises SequenceLengthException and does not pass, as in
// FilterProcessor::process_sequence (processors.hh:389-401).
template <class G>
uint64_t diginorm_run(void* graph, unsigned cutoff, const char* seq, const int64_t* off, uint64_t n_reads,
                      uint8_t* passed) {
    auto f = DiginormFilter<G>::Filter::build(static_cast<Holder<G>*>(graph)->g, cutoff);
    uint64_t n = 0;
    for (uint64_t r = 0; r < n_reads; ++r) {
        std::string s(seq + off[r], (size_t)(off[r + 1] - off[r]));
        passed[r] = 0;
        try {
            bool p = std::get<0>(f->filter_sequence(s));
            passed[r] = p ? 1 : 0;
            n += p;
        } catch (SequenceLengthException&) {
        }
    }
    return n;
}

// StreamingSolidFilter<G>::Filter::filter_sequence (solidifier.hh:58-73)
// over reads in order; a read shorter than K raises SequenceLengthException
// and does not pass, as in FilterProcessor::process_sequence.
template <class G>
uint64_t solid_run(void* graph, float min_prop, uint32_t threshold, const char* seq, const int64_t* off,
                   uint64_t n_reads, uint8_t* passed) {
    auto f = StreamingSolidFilter<G>::Filter::build(static_cast<Holder<G>*>(graph)->g, min_prop, threshold);
    uint64_t n = 0;
    for (uint64_t r = 0; r < n_reads; ++r) {
        std::string s(seq + off[r], (size_t)(off[r + 1] - off[r]));
        passed[r] = 0;
        try {
            bool p = std::get<0>(f->filter_sequence(s));
            passed[r] = p ? 1 : 0;
            n += p;
        } catch (SequenceLengthException&) {
        }
    }
    return n;
}

// InserterProcessor<G>::advance over a SplitPairedReader (processors.hh:180-204)
// until no pairs remain: every advance's (n_sequences, time, remaining) into
// steps (3 words each, up to cap steps); *err = 1 when the reader raised a
// GoetiaException (the advance then stops), 0 otherwise.
template <class G>
uint64_t paired_run(void* graph, const char* left, const char* right, int strict, uint32_t min_length, int force,
                    uint64_t interval, uint64_t* steps, uint64_t cap, int* err, uint64_t* n_skipped,
                    uint64_t* n_parsed) {
    auto g = static_cast<Holder<G>*>(graph)->g;
    auto proc = G::Processor::build(g, interval);
    auto reader = SplitPairedReader<FastxParser<DNA_SIMPLE>>::build(left, right, strict, min_length, force);
    uint64_t n = 0;
    *err = 0;
    try {
        bool remaining = true;
        while (remaining) {
            uint64_t ns, tt;
            std::tie(ns, tt, remaining) = proc->advance(reader);
            if (n < cap) {
                steps[3 * n] = ns;
                steps[3 * n + 1] = tt;
                steps[3 * n + 2] = remaining ? 1 : 0;
            }
            ++n;
        }
    } catch (GoetiaException&) {
        *err = 1;
    }
    *n_skipped = reader->n_skipped();
    *n_parsed = reader->n_parsed();
    return n;
}

// StreamingCompactor<G>::Compactor::find_new_segments (compactor.hh:321-534)
// over reads in order; see ref_find_new_segments.
template <class G>
uint64_t segments_run(std::shared_ptr<G> g, const char* seq, const int64_t* off, uint64_t n_reads, int persist,
                      int insert_after, uint64_t* out, uint64_t cap, char* kmers, uint64_t kcap, uint64_t* kused) {
    using C = typename StreamingCompactor<G>::Compactor;
    using hash_type = typename G::hash_type;
    C cp(g);
    uint64_t w = 0, kb = 0;
    auto put = [&](uint64_t v) {
        if (w < cap) out[w] = v;
        ++w;
    };
    auto put_kmer = [&](const std::string& k) {
        for (char c : k) {
            if (kb < kcap) kmers[kb] = c;
            ++kb;
        }
    };
    std::set<hash_type> shared_new;
    for (uint64_t r = 0; r < n_reads; ++r) {
        std::string s(seq + off[r], (size_t)(off[r + 1] - off[r]));
        std::vector<hash_type> hashes;
        std::set<hash_type> fresh_new, dec;
        std::set<hash_type>& nk = persist ? shared_new : fresh_new;
        std::vector<typename StreamingCompactor<G>::compact_segment> segs;
        std::deque<typename StreamingCompactor<G>::neighbor_pair_type> pairs;
        try {
            cp.find_new_segments(s, hashes, nk, segs, dec, pairs);
        } catch (SequenceLengthException&) {
            put(~0ULL);
            continue;
        }
        put(segs.size());
        for (auto& sg : segs) {
            put(sg.left_anchor.value());
            put(sg.right_anchor.value());
            put(sg.left_flank.value());
            put(sg.right_flank.value());
            put(sg.is_decision_kmer ? 1 : 0);
            put(sg.start_pos);
            put(sg.length);
        }
        put(nk.size());
        if (!persist || r + 1 == n_reads)
            for (auto& h : nk) put(h.value());
        put(dec.size());
        for (auto& h : dec) put(h.value());
        put(pairs.size());
        for (auto& pr : pairs) {
            put(pr.first.size());
            put(pr.second.size());
            for (auto& k : pr.first) {
                put(k.value());
                put_kmer(k.kmer);
            }
            for (auto& k : pr.second) {
                put(k.value());
                put_kmer(k.kmer);
            }
        }
        if (insert_after && s.size() >= (size_t)g->K) g->insert_sequence(s);
    }
    *kused = kb;
    return w;
}

}  // namespace

extern "C" {

// HashShifter::hash_base + KmerIterator over one read (kmeriterator.hh:109-123).
int ref_hash_read(int K, const char* seq, size_t len, uint64_t* fw, uint64_t* rc) {
    std::string s(seq, len);
    if (len < (size_t)K) return 0;
    KmerIterator<CanLemireShifter> it(s, (uint16_t)K);
    size_t i = 0;
    while (!it.done()) {
        auto h = it.next();
        fw[i] = h.fw_hash;
        rc[i] = h.rc_hash;
        ++i;
    }
    return (int)i;
}

uint64_t ref_fwd_hash(int K, const char* kmer) {
    return FwdLemireShifter::hash(std::string(kmer, (size_t)K), (uint16_t)K).value();
}

uint32_t ref_primes(uint32_t n, uint64_t x, uint64_t* out) {
    auto v = get_n_primes_near_x(n, x);
    for (size_t i = 0; i < v.size(); ++i) out[i] = v[i];
    return (uint32_t)v.size();
}

void* ref_new(int kind, int canonical, int K, uint64_t max_table, uint16_t n_tables) {
    auto* h = new Handle{kind, canonical, nullptr, nullptr};
    if (kind == 1) {
        auto s = BitStorage::build(max_table, n_tables);
        h->graph = canonical ? make_graph<BitStorage, CanLemireShifter>(s, K)
                             : make_graph<BitStorage, FwdLemireShifter>(s, K);
    } else if (kind == 2) {
        auto s = ByteStorage::build(max_table, n_tables);
        h->bytes = s;
        h->graph = canonical ? make_graph<ByteStorage, CanLemireShifter>(s, K)
                             : make_graph<ByteStorage, FwdLemireShifter>(s, K);
    } else {
        auto s = SparseppSetStorage::build();
        h->graph = canonical ? make_graph<SparseppSetStorage, CanLemireShifter>(s, K)
                             : make_graph<SparseppSetStorage, FwdLemireShifter>(s, K);
    }
    return h;
}

// dBG::insert_sequence(seq, n_new) per read piece, in read order.
uint64_t ref_insert_reads(void* hv, int K, int mode, const char* seq, const int64_t* off,
                          uint64_t n_reads, uint64_t* n_new) {
    Handle* h = (Handle*)hv;
    uint64_t total = 0;
    for (uint64_t r = 0; r < n_reads; ++r) {
        uint64_t nn_read = 0;
        for (auto& p : pieces(seq + off[r], (size_t)(off[r + 1] - off[r]), K, mode)) {
            uint64_t nn = 0;
            total += dispatch(h, [&](auto* g) { return g->insert_sequence(p.second, nn); });
            nn_read += nn;
        }
        if (n_new) n_new[r] = nn_read;
    }
    return total;
}

// dBG::insert_and_query_sequence per read (mode 0 only).
uint64_t ref_insert_and_query_reads(void* hv, int K, const char* seq, const int64_t* off,
                                    uint64_t n_reads, int16_t* counts) {
    Handle* h = (Handle*)hv;
    uint64_t pos = 0;
    for (uint64_t r = 0; r < n_reads; ++r) {
        size_t len = (size_t)(off[r + 1] - off[r]);
        if (len < (size_t)K) continue;
        std::string s(seq + off[r], len);
        auto c = dispatch(h, [&](auto* g) { return g->insert_and_query_sequence(s); });
        for (auto v : c) counts[pos++] = v;
    }
    return pos;
}

// dBG::query_sequence per read piece; counts packed len-K+1 per read with -1
// for skipped windows.
uint64_t ref_query_reads(void* hv, int K, int mode, const char* seq, const int64_t* off,
                         uint64_t n_reads, int16_t* counts) {
    Handle* h = (Handle*)hv;
    uint64_t pos = 0;
    for (uint64_t r = 0; r < n_reads; ++r) {
        size_t len = (size_t)(off[r + 1] - off[r]);
        if (len < (size_t)K) continue;
        size_t n = len - (size_t)K + 1;
        for (size_t j = 0; j < n; ++j) counts[pos + j] = -1;
        for (auto& p : pieces(seq + off[r], len, K, mode)) {
            auto c = dispatch(h, [&](auto* g) { return g->query_sequence(p.second); });
            for (size_t j = 0; j < c.size(); ++j) counts[pos + p.first + j] = c[j];
        }
        pos += n;
    }
    return pos;
}

int16_t ref_query_hash(void* hv, uint64_t v) {
    Handle* h = (Handle*)hv;
    return dispatch(h, [&](auto* g) -> int16_t {
        using HT = typename std::remove_pointer_t<decltype(g)>::hash_type;
        HT hh;
        if constexpr (std::is_same_v<HT, Canonical<uint64_t>>) hh = HT(v, v);
        else hh = HT(v);
        return g->query(hh);
    });
}

void ref_counters(void* hv, uint64_t* n_unique, uint64_t* n_occupied) {
    Handle* h = (Handle*)hv;
    *n_unique = dispatch(h, [&](auto* g) { return g->n_unique(); });
    *n_occupied = dispatch(h, [&](auto* g) { return g->n_occupied(); });
}

// Copies table i of get_raw() (bitstorage.hh:228-231 / bytestorage.hh).
void ref_table(void* hv, uint32_t i, uint64_t nbytes, uint8_t* out) {
    Handle* h = (Handle*)hv;
    uint8_t** t = dispatch(h, [&](auto* g) { return g->get_raw(); });
    std::memcpy(out, t[i], nbytes);
}

// Left/right extension hashes and degrees of one k-mer via the walker
// (set_cursor + in_degree/out_degree, pythonize_dbg.py:30-36).
void ref_kmer_extensions(void* hv, const char* kmer, int K, uint64_t* left, uint64_t* right,
                         uint8_t* in_deg, uint8_t* out_deg) {
    Handle* h = (Handle*)hv;
    dispatch(h, [&](auto* g) {
        std::string s(kmer, (size_t)K);
        g->set_cursor(s);
        auto l = g->left_extensions();
        auto r = g->right_extensions();
        for (int c = 0; c < 4; ++c) {
            left[c] = l[c].value();
            right[c] = r[c].value();
        }
        g->set_cursor(s);
        *in_deg = (uint8_t)g->in_degree();
        *out_deg = (uint8_t)g->out_degree();
        return 0;
    });
}

// UnitigWalker::find_decision_kmers (unitig_walker.hh:1025-1045).
uint64_t ref_find_decision_kmers(void* hv, const char* seq, size_t len, uint32_t* positions,
                                 uint64_t* hashes, uint64_t cap) {
    Handle* h = (Handle*)hv;
    return dispatch(h, [&](auto* g) -> uint64_t {
        using G = std::remove_pointer_t<decltype(g)>;
        std::vector<uint32_t> pos;
        std::vector<typename G::hash_type> hs;
        std::vector<typename G::neighbor_pair_type> nb;
        g->find_decision_kmers(std::string(seq, len), pos, hs, nb);
        for (size_t i = 0; i < pos.size() && i < cap; ++i) {
            positions[i] = pos[i];
            hashes[i] = hs[i].value();
        }
        return pos.size();
    });
}

// FastxParser<DNA_SIMPLE> (readers.hh:148-219) over a whole file, as a
// caller looping on next() sees it: the records it returns (validated,
// uppercased sequence; name), in order, then n_parsed / n_skipped and the
// number of InvalidRead exceptions.  In strict mode the first invalid
// character ends the parse and counts[3] is set.  Returns the record count
// (records past max_records or the buffers are counted, not stored).
uint64_t ref_fastx_parse(const char* path, int strict, uint32_t min_length, char* seq, uint64_t seq_cap,
                         int64_t* offsets, char* names, uint64_t names_cap, int64_t* name_off,
                         uint64_t max_records, uint64_t* counts) {
    auto parser = goetia::FastxParser<goetia::DNA_SIMPLE>::build(path, strict != 0, min_length);
    uint64_t n = 0, sl = 0, nl = 0;
    counts[0] = counts[1] = counts[2] = counts[3] = 0;
    offsets[0] = 0;
    name_off[0] = 0;
    while (!parser->is_complete()) {
        std::optional<goetia::Record> rec;
        try {
            rec = parser->next();
        } catch (goetia::InvalidRead&) {
            counts[2]++;
            continue;
        } catch (goetia::InvalidCharacterException&) {
            counts[3] = 1;
            break;
        }
        if (!rec) continue;
        if (n < max_records && sl + rec->sequence.size() <= seq_cap && nl + rec->name.size() <= names_cap) {
            std::memcpy(seq + sl, rec->sequence.data(), rec->sequence.size());
            sl += rec->sequence.size();
            std::memcpy(names + nl, rec->name.data(), rec->name.size());
            nl += rec->name.size();
            offsets[n + 1] = (int64_t)sl;
            name_off[n + 1] = (int64_t)nl;
        }
        n++;
    }
    counts[0] = parser->n_parsed();
    counts[1] = parser->n_skipped();
    return n;
}

// DiginormFilter over reads in order (count-min graphs only; returns the
// reads that passed, or UINT64_MAX for another storage).
uint64_t ref_diginorm_reads(void* hv, unsigned cutoff, const char* seq, const int64_t* off, uint64_t n_reads,
                            uint8_t* passed) {
    Handle* h = (Handle*)hv;
    if (h->kind != 2) return UINT64_MAX;
    if (h->canonical) return diginorm_run<Graph<ByteStorage, CanLemireShifter>>(h->graph, cutoff, seq, off, n_reads,
                                                                             passed);
    return diginorm_run<Graph<ByteStorage, FwdLemireShifter>>(h->graph, cutoff, seq, off, n_reads, passed);
}

// find_new_segments over reads in order.  persist != 0: one new-k-mer set
// shared by all the reads (the six-argument overload with a caller-owned
// set); otherwise a fresh set per read (the one-argument overload).
// insert_after != 0: insert_sequence(read) after its segments, as goetia's
// tests do.  Per read, into `out` (u64 words): the segment count (~0 when the
// read is shorter than K), 7 words per segment (left_anchor, right_anchor,
// left_flank, right_flank, is_decision_kmer, start_pos, length; hashes as
// value()), the new-k-mer set (count, values in set order; with persist, the
// values after the last read only), the new decision
// k-mers (count, values), the decision neighbour pairs (count; per pair the
// left and right counts, then each k-mer's value); the neighbour k-mers'
// strings go to `kmers`, K bytes each, in the same order.  Returns the words
// needed; *kused the bytes.  Forward (non-canonical) graphs only: goetia's
// find_new_segments does not compile for Canonical hashes.
uint64_t ref_find_new_segments(void* hv, const char* seq, const int64_t* off, uint64_t n_reads, int persist,
                               int insert_after, uint64_t* out, uint64_t cap, char* kmers, uint64_t kcap,
                               uint64_t* kused) {
    Handle* h = (Handle*)hv;
    auto run = [&](auto* holder) {
        return segments_run(holder->g, seq, off, n_reads, persist, insert_after, out, cap, kmers, kcap, kused);
    };
    if (h->canonical) return ~0ULL;   // compactor.hh:405,428 assign a value_type to the hash: Hash<> only
    if (h->kind == 1) return run(static_cast<Holder<Graph<BitStorage, FwdLemireShifter>>*>(h->graph));
    if (h->kind == 2) return run(static_cast<Holder<Graph<ByteStorage, FwdLemireShifter>>*>(h->graph));
    return run(static_cast<Holder<Graph<SparseppSetStorage, FwdLemireShifter>>*>(h->graph));
}

// StreamingSolidFilter over reads (Bloom or count-min graphs); returns the
// number of reads that pass.
uint64_t ref_solid_filter_reads(void* hv, float min_prop, uint32_t threshold, const char* seq, const int64_t* off,
                                uint64_t n_reads, uint8_t* passed) {
    Handle* h = (Handle*)hv;
    if (h->kind == 1) {
        if (h->canonical)
            return solid_run<Graph<BitStorage, CanLemireShifter>>(h->graph, min_prop, threshold, seq, off, n_reads, passed);
        return solid_run<Graph<BitStorage, FwdLemireShifter>>(h->graph, min_prop, threshold, seq, off, n_reads, passed);
    }
    if (h->kind == 2) {
        if (h->canonical)
            return solid_run<Graph<ByteStorage, CanLemireShifter>>(h->graph, min_prop, threshold, seq, off, n_reads, passed);
        return solid_run<Graph<ByteStorage, FwdLemireShifter>>(h->graph, min_prop, threshold, seq, off, n_reads, passed);
    }
    return ~0ULL;
}

// Paired-file insertion through goetia's own processor and reader (Bloom
// and count-min graphs); returns the number of advances.
uint64_t ref_paired_insert(void* hv, const char* left, const char* right, int strict, uint32_t min_length,
                           int force, uint64_t interval, uint64_t* steps, uint64_t cap, int* err,
                           uint64_t* n_skipped, uint64_t* n_parsed) {
    Handle* h = (Handle*)hv;
    auto run = [&](auto tag) {
        using G = typename decltype(tag)::type;
        return paired_run<G>(h->graph, left, right, strict, min_length, force, interval, steps, cap, err, n_skipped,
                             n_parsed);
    };
    if (h->kind == 1)
        return h->canonical ? run(std::common_type<Graph<BitStorage, CanLemireShifter>>{})
                            : run(std::common_type<Graph<BitStorage, FwdLemireShifter>>{});
    return h->canonical ? run(std::common_type<Graph<ByteStorage, CanLemireShifter>>{})
                        : run(std::common_type<Graph<ByteStorage, FwdLemireShifter>>{});
}

// SplitPairedReader::next over two files until both end: per step, the
// lengths of the left and right sequences (-1 for an empty side) or -2 for
// both when next() raised (strict rejection), and -3 when the reader raised a
// GoetiaException (the walk stops there).  Returns the steps.
uint64_t ref_paired_walk(const char* left, const char* right, int strict, uint32_t min_length, int force,
                         int64_t* lens, uint64_t cap) {
    auto reader = SplitPairedReader<FastxParser<DNA_SIMPLE>>::build(left, right, strict, min_length, force);
    uint64_t n = 0;
    auto put = [&](int64_t a, int64_t b) {
        if (n < cap) {
            lens[2 * n] = a;
            lens[2 * n + 1] = b;
        }
        ++n;
    };
    try {
        while (!reader->is_complete()) {
            try {
                auto pr = reader->next();
                put(pr.first ? (int64_t)pr.first->sequence.size() : -1,
                    pr.second ? (int64_t)pr.second->sequence.size() : -1);
            } catch (InvalidCharacterException&) {
                put(-2, -2);
            } catch (InvalidRead&) {
                put(-2, -2);
            }
        }
    } catch (GoetiaException&) {
        put(-3, -3);
    }
    return n;
}

// dBG::save(filename) (dbg.hh:396-399): the graph's storage writes its OXLI
// file (BitStorage::save, ByteStorage::save through ByteStorageFile).
// Storage::set_use_bigcount (storage.hh:129-135) on a count-min handle.
void ref_set_use_bigcount(void* hv, int b) {
    Handle* h = (Handle*)hv;
    if (h->bytes) h->bytes->set_use_bigcount(b != 0);
}

// dBG::load (dbg.hh:400-403) into a count-min handle's storage; 0, or -1 when
// the loader threw.
int ref_load_into(void* hv, const char* path) {
    Handle* h = (Handle*)hv;
    if (h->kind != 2) return -1;
    try {
        dispatch(h, [&](auto* g) {
            g->load(std::string(path));
            return 0;
        });
    } catch (std::exception&) {
        return -1;
    }
    return 0;
}

void ref_save(void* hv, const char* path) {
    Handle* h = (Handle*)hv;
    dispatch(h, [&](auto* g) {
        g->save(std::string(path));
        return 0;
    });
}

// A Bloom (kind 1) or count-min (kind 2) storage loaded from an OXLI file by
// the reference's loader: ksize, table count, sizes (room for 16), occupied
// bins, and the tables concatenated into out (cap bytes).  Returns the bytes
// the tables need, or ~0 when the loader threw.
uint64_t ref_load_file(int kind, const char* path, uint16_t* ksize, uint32_t* n_tables, uint64_t* sizes,
                       uint64_t* occupied, uint8_t* out, uint64_t cap) {
    auto run = [&](auto store) -> uint64_t {
        uint16_t k = 0;
        try {
            store->load(std::string(path), k);
        } catch (std::exception&) {
            return ~0ULL;
        }
        *ksize = k;
        auto ts = store->get_tablesizes();
        *n_tables = (uint32_t)ts.size();
        *occupied = store->n_occupied();
        byte_t** t = store->get_raw_tables();
        uint64_t w = 0;
        for (size_t i = 0; i < ts.size() && i < 16; ++i) {
            sizes[i] = ts[i];
            const uint64_t nb = kind == 1 ? ts[i] / 8 + 1 : ts[i];
            for (uint64_t j = 0; j < nb; ++j, ++w)
                if (w < cap) out[w] = t[i][j];
        }
        return w;
    };
    if (kind == 1) return run(BitStorage::build(1000, 1));
    return run(ByteStorage::build(1000, 1));
}

// Storage-level insert trace of an exact set (flavor 0 SparseppSetStorage,
// 1 PHMapStorage, 2 BTreeStorage): before value i, reset() when
// reset_before[i]; after it, the insert's result and the storage's
// n_unique_kmers() / n_occupied().
void ref_set_trace(int flavor, const uint64_t* h, uint64_t n, const uint8_t* reset_before, uint8_t* is_new,
                   uint64_t* n_unique, uint64_t* n_occupied) {
    auto run = [&](auto store) {
        for (uint64_t i = 0; i < n; ++i) {
            if (reset_before[i]) store->reset();
            is_new[i] = store->insert(h[i]) ? 1 : 0;
            n_unique[i] = store->n_unique_kmers();
            n_occupied[i] = store->n_occupied();
        }
    };
    if (flavor == 0) run(SparseppSetStorage::build());
    else if (flavor == 1) run(PHMapStorage::build());
    else run(BTreeStorage::build());
}

void ref_free(void* hv) {
    Handle* h = (Handle*)hv;
    // graphs are leaked deliberately: the driver lives for one short process.
    delete h;
}

}  // extern "C"
