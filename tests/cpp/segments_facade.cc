// C++ facade check of StreamingCompactor::find_new_segments: runs cases and
// writes each one's record stream in the format of the reference driver's
// ref_find_new_segments, for tests/test_cpp_facade.py to compare with the
// golden vectors.  Usage: segments_facade CASES WORDS KMERS.  CASES holds the
// number of cases, then per case "kind K max_table n_tables persist
// insert_after", the count and lines of the reads inserted first, then the
// count and lines of the reads.  WORDS gets per case its word count and
// words, KMERS per case its byte count on a line and the neighbour k-mers.
// For cases whose graph does not change between reads, the batch form must
// give what the per-read calls gave (checked here; exit status 1 if not).
#include <cstdio>
#include <cstdint>
#include <deque>
#include <fstream>
#include <set>
#include <string>
#include <vector>

#include "goetia/cdbg/compactor.hh"
#include "goetia/dbg.hh"

using namespace goetia;

static std::vector<std::string> read_lines(std::ifstream& in) {
    size_t n = 0;
    in >> n;
    std::string line;
    std::getline(in, line);
    std::vector<std::string> out;
    for (size_t i = 0; i < n; ++i) {
        std::getline(in, line);
        out.push_back(line);
    }
    return out;
}

template <class S>
int run(std::shared_ptr<S> store, int K, bool persist, bool insert_after, const std::vector<std::string>& bg,
        const std::vector<std::string>& reads, std::vector<uint64_t>& words, std::string& kmers) {
    typedef dBG<S, FwdLemireShifter> G;
    typedef StreamingCompactor<G> SC;
    auto g = G::build(store, (uint16_t)K);
    if (!bg.empty()) g->insert_sequences(ReadBatch(bg));
    auto cp = SC::Compactor::build(g);
    std::set<typename G::hash_type> shared;
    auto put = [&](uint64_t v) { words.push_back(v); };
    // per-read results, compared below with the batch form when the graph
    // does not change between reads
    std::vector<std::vector<typename SC::compact_segment>> all_segs(reads.size());
    std::vector<std::set<typename G::hash_type>> all_dec(reads.size());
    std::vector<std::deque<typename G::neighbor_pair_type>> all_pairs(reads.size());
    for (size_t r = 0; r < reads.size(); ++r) {
        std::vector<typename G::hash_type> hashes;
        std::set<typename G::hash_type> fresh, dec;
        auto& nk = persist ? shared : fresh;
        std::vector<typename SC::compact_segment> segs;
        std::deque<typename G::neighbor_pair_type> pairs;
        try {
            cp->find_new_segments(reads[r], hashes, nk, segs, dec, pairs);
        } catch (SequenceLengthException&) {
            put(~0ULL);
            continue;
        }
        put(segs.size());
        for (auto& s : segs) {
            put(s.left_anchor.value());
            put(s.right_anchor.value());
            put(s.left_flank.value());
            put(s.right_flank.value());
            put(s.is_decision_kmer ? 1 : 0);
            put(s.start_pos);
            put(s.length);
        }
        put(nk.size());
        if (!persist || r + 1 == reads.size())
            for (auto& h : nk) put(h.value());
        put(dec.size());
        for (auto& h : dec) put(h.value());
        put(pairs.size());
        for (auto& p : pairs) {
            put(p.first.size());
            put(p.second.size());
            for (auto& k : p.first) {
                put(k.value());
                kmers += k.kmer;
            }
            for (auto& k : p.second) {
                put(k.value());
                kmers += k.kmer;
            }
        }
        all_segs[r] = segs;
        all_dec[r] = dec;
        all_pairs[r] = pairs;
        if (insert_after && reads[r].size() >= (size_t)K) g->insert_sequence(reads[r]);
    }
    if (!insert_after) {
        std::vector<std::vector<typename SC::compact_segment>> bs;
        std::vector<std::set<typename G::hash_type>> bd;
        std::vector<std::deque<typename G::neighbor_pair_type>> bp;
        std::set<typename G::hash_type> bshared;
        cp->find_new_segments(ReadBatch(reads), bs, bd, bp, persist ? &bshared : nullptr);
        for (size_t r = 0; r < reads.size(); ++r) {
            bool same = bs[r].size() == all_segs[r].size() && bd[r] == all_dec[r] && bp[r].size() == all_pairs[r].size();
            for (size_t j = 0; same && j < bs[r].size(); ++j) {
                const auto &a = bs[r][j], &b = all_segs[r][j];
                same = a.left_anchor == b.left_anchor && a.right_anchor == b.right_anchor &&
                       a.left_flank == b.left_flank && a.right_flank == b.right_flank &&
                       a.is_decision_kmer == b.is_decision_kmer && a.start_pos == b.start_pos && a.length == b.length;
            }
            for (size_t j = 0; same && j < bp[r].size(); ++j) {
                const auto &a = bp[r][j], &b = all_pairs[r][j];
                same = a.first.size() == b.first.size() && a.second.size() == b.second.size();
                for (size_t k = 0; same && k < a.first.size(); ++k)
                    same = a.first[k].value() == b.first[k].value() && a.first[k].kmer == b.first[k].kmer;
                for (size_t k = 0; same && k < a.second.size(); ++k)
                    same = a.second[k].value() == b.second[k].value() && a.second[k].kmer == b.second[k].kmer;
            }
            if (!same) {
                std::fprintf(stderr, "batch form differs at read %zu\n", r);
                return 1;
            }
        }
        if (persist && bshared != shared) {
            std::fprintf(stderr, "batch form's shared set differs\n");
            return 1;
        }
    }
    return 0;
}

int main(int argc, char** argv) {
    if (argc != 4) return 2;
    std::ifstream in(argv[1]);
    std::FILE* wf = std::fopen(argv[2], "w");
    std::FILE* kf = std::fopen(argv[3], "w");
    if (!in || !wf || !kf) return 2;
    int n_cases = 0;
    in >> n_cases;
    for (int c = 0; c < n_cases; ++c) {
        int kind = 0, K = 0, persist = 0, insert_after = 0;
        unsigned long long max_table = 0, n_tables = 0;
        in >> kind >> K >> max_table >> n_tables >> persist >> insert_after;
        auto bg = read_lines(in);
        auto reads = read_lines(in);
        std::vector<uint64_t> words;
        std::string kmers;
        int rc;
        if (kind == 1)
            rc = run(BitStorage::build(max_table, (uint16_t)n_tables), K, persist, insert_after, bg, reads, words,
                     kmers);
        else if (kind == 2)
            rc = run(ByteStorage::build(max_table, (uint16_t)n_tables), K, persist, insert_after, bg, reads, words,
                     kmers);
        else
            rc = run(SparseppSetStorage::build(), K, persist, insert_after, bg, reads, words, kmers);
        if (rc) return rc;
        std::fprintf(wf, "%zu\n", words.size());
        for (uint64_t w : words) std::fprintf(wf, "%llu\n", (unsigned long long)w);
        std::fprintf(kf, "%zu\n%s\n", kmers.size(), kmers.c_str());
    }
    std::fclose(wf);
    std::fclose(kf);
    std::printf("SEGMENTS OK %d\n", n_cases);
    return 0;
}
