/**
 * compactor.hh -- StreamingCompactor<dBG<S, Sh>>::Compactor's segment finder
 * (goetia compactor.hh:35-534): find_new_segments cuts a read into runs of
 * k-mers new to the graph (compact_segment, :77-145), with anchors and
 * flanks, split at the k-mers that become decision nodes once the read is in
 * the graph.  One library call per read or per batch of reads
 * (gt_find_new_segments): three GPU launches give every window's count and
 * the new windows' neighbour values and presence bits, and the library's host code runs the
 * two passes over each read, as in goetia.
 *
 * goetia compiles find_new_segments for forward hashes only (it assigns a
 * neighbour's value_type to the hash, :405 and :428, which Canonical does not
 * accept); the same holds here.  The cDBG the reference's compactor updates
 * from the segments is not part of this facade.
 */
#ifndef GOETIA_AMD_CDBG_COMPACTOR_HH
#define GOETIA_AMD_CDBG_COMPACTOR_HH

#include <cstdint>
#include <deque>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "goetia/dbg.hh"

namespace goetia {

template <class T>
struct StreamingCompactor;

template <template <class, class> class GraphType, class StorageType, class ShifterType>
struct StreamingCompactor<GraphType<StorageType, ShifterType>> {
    typedef GraphType<StorageType, ShifterType> graph_type;
    typedef ShifterType shifter_type;
    typedef typename shifter_type::hash_type hash_type;
    typedef typename hash_type::value_type value_type;
    typedef typename shifter_type::kmer_type kmer_type;
    typedef typename graph_type::neighbor_pair_type neighbor_pair_type;

    static_assert(!ShifterType::is_canonical,
                  "StreamingCompactor::find_new_segments is built for forward hashes only");

    struct compact_segment {
        hash_type left_anchor;
        hash_type right_anchor;
        hash_type left_flank;
        hash_type right_flank;
        bool is_decision_kmer;
        size_t start_pos;
        size_t length;
        std::vector<hash_type> tags;

        // the default segment is the null delimiter (compactor.hh:94-103)
        compact_segment()
            : left_anchor(0), right_anchor(0), left_flank(0), right_flank(0), is_decision_kmer(false), start_pos(0),
              length(0) {}

        bool is_null() const { return length == 0 && (left_anchor == right_anchor) && !is_decision_kmer; }
    };

    class Compactor {
      public:
        const uint16_t K;
        std::shared_ptr<graph_type> dbg;

        explicit Compactor(std::shared_ptr<graph_type> dbg, uint64_t minimizer_window_size = 8)
            : K(dbg->K), dbg(dbg), _minimizer_window_size(minimizer_window_size) {}

        static std::shared_ptr<Compactor> build(std::shared_ptr<graph_type> dbg,
                                                uint64_t minimizer_window_size = 8) {
            return std::make_shared<Compactor>(dbg, minimizer_window_size);
        }

        std::vector<compact_segment> find_new_segments(const std::string& sequence) {
            std::vector<compact_segment> result;
            find_new_segments(sequence, result);
            return result;
        }

        void find_new_segments(const std::string& sequence, std::vector<compact_segment>& result) {
            std::vector<hash_type> hashes;
            std::set<hash_type> new_kmers, new_decision_kmers;
            std::deque<neighbor_pair_type> decision_neighbors;
            find_new_segments(sequence, hashes, new_kmers, result, new_decision_kmers, decision_neighbors);
        }

        // compactor.hh:346-534: appends to the caller's containers.
        void find_new_segments(const std::string& sequence, std::vector<hash_type>& hashes,
                               std::set<hash_type>& new_kmers, std::vector<compact_segment>& segments,
                               std::set<hash_type>& new_decision_kmers,
                               std::deque<neighbor_pair_type>& decision_neighbors) {
            if (sequence.size() < K) throw SequenceLengthException("Sequence must have length >= K");
            ReadBatch b;
            b.push_back(sequence);
            std::vector<uint64_t> known, prefix;
            for (const auto& h : new_kmers) known.push_back(h.value());
            for (const auto& h : hashes) prefix.push_back(h.value());
            Result res = run(b, known, false, prefix);
            res.append(0, sequence, K, hashes, new_kmers, segments, new_decision_kmers, decision_neighbors);
        }

        // find_new_segments over every read of a batch against the graph as
        // it is (nothing inserted in between): three GPU launches and one pass
        // of the library's host code for the whole batch.  With `shared`,
        // one new-k-mer set (new_kmers) runs through the reads in order;
        // otherwise each read starts from an empty one.  Results per read are
        // appended to the vectors' entries for that read (reads shorter than
        // K get none).
        void find_new_segments(const ReadBatch& reads, std::vector<std::vector<compact_segment>>& segments,
                               std::vector<std::set<hash_type>>& new_decision_kmers,
                               std::vector<std::deque<neighbor_pair_type>>& decision_neighbors,
                               std::set<hash_type>* new_kmers = nullptr) {
            std::vector<uint64_t> known;
            if (new_kmers)
                for (const auto& h : *new_kmers) known.push_back(h.value());
            Result res = run(reads, known, new_kmers != nullptr, {});
            segments.resize(reads.size());
            new_decision_kmers.resize(reads.size());
            decision_neighbors.resize(reads.size());
            uint64_t w = 0;
            for (uint64_t r = 0; r < reads.size(); ++r) {
                const std::string s = reads.read(r);
                if (s.size() < K) continue;
                std::vector<hash_type> hs;
                std::set<hash_type> own;
                res.append(r, s, K, hs, new_kmers ? *new_kmers : own, segments[r], new_decision_kmers[r],
                           decision_neighbors[r], w);
                w += s.size() - K + 1;
            }
        }

      private:
        uint64_t _minimizer_window_size;

        struct Result {
            std::vector<int16_t> counts;
            std::vector<uint64_t> hv, segs, decs;
            std::vector<int64_t> seg_off, dec_off;

            // Appends read r's results (its windows start at w in counts / hv).
            void append(uint64_t r, const std::string& s, uint16_t K, std::vector<hash_type>& hashes,
                        std::set<hash_type>& new_kmers, std::vector<compact_segment>& segments,
                        std::set<hash_type>& new_decision_kmers, std::deque<neighbor_pair_type>& decision_neighbors,
                        uint64_t w = 0) const {
                const uint64_t n = s.size() - K + 1;
                for (uint64_t i = 0; i < n; ++i) {
                    hashes.push_back(hash_type(hv[w + i]));
                    if (counts[w + i] == 0) new_kmers.insert(hash_type(hv[w + i]));
                }
                for (int64_t j = seg_off[r]; j < seg_off[r + 1]; ++j) {
                    const uint64_t* o = segs.data() + 7 * j;
                    compact_segment sg;
                    sg.left_anchor = hash_type(o[0]);
                    sg.right_anchor = hash_type(o[1]);
                    sg.left_flank = hash_type(o[2]);
                    sg.right_flank = hash_type(o[3]);
                    sg.is_decision_kmer = o[4] != 0;
                    sg.start_pos = o[5];
                    sg.length = o[6];
                    segments.push_back(sg);
                }
                for (int64_t j = dec_off[r]; j < dec_off[r + 1]; ++j) {
                    const uint64_t* o = decs.data() + 11 * j;
                    const std::string root = s.substr(o[0], K);
                    std::vector<kmer_type> lk, rk;
                    for (int c = 0; c < 4; ++c) {
                        if ((o[2] >> c) & 1) lk.push_back(kmer_type(hash_type(o[3 + c]), "ACGT"[c] + root.substr(0, K - 1)));
                    }
                    for (int c = 0; c < 4; ++c) {
                        if ((o[2] >> (4 + c)) & 1) rk.push_back(kmer_type(hash_type(o[7 + c]), root.substr(1) + "ACGT"[c]));
                    }
                    decision_neighbors.push_back(std::make_pair(lk, rk));
                    new_decision_kmers.insert(hash_type(o[1]));
                }
            }
        };

        Result run(const ReadBatch& reads, const std::vector<uint64_t>& known, bool shared,
                   const std::vector<uint64_t>& prefix) {
            Result res;
            const uint64_t nk = reads.n_kmers(K);
            res.counts.assign(nk ? nk : 1, 0);
            res.hv.assign(nk ? nk : 1, 0);
            res.seg_off.assign(reads.size() + 1, 0);
            res.dec_off.assign(reads.size() + 1, 0);
            uint64_t seg_cap = 4 * reads.size() + nk / 16 + 16, dec_cap = nk / 16 + 16;
            for (int attempt = 0; attempt < 2; ++attempt) {
                res.segs.assign(seg_cap * 7, 0);
                res.decs.assign(dec_cap * 11, 0);
                uint64_t n_segs = 0, n_decs = 0;
                const int rc = gt_find_new_segments(
                    dbg->storage()->handle(), K, dbg->input_mode(), reads.data(), reads.offsets.data(), reads.size(),
                    known.data(), known.size(), shared ? 1 : 0, prefix.data(), prefix.size(), res.counts.data(),
                    res.hv.data(), res.seg_off.data(), res.segs.data(), seg_cap, res.dec_off.data(), res.decs.data(),
                    dec_cap, &n_segs, &n_decs);
                if (rc == GT_OK) return res;
                if (rc != GT_ERR_INVALID_ARG || (n_segs <= seg_cap && n_decs <= dec_cap) || attempt)
                    detail::check(rc, "find_new_segments");
                seg_cap = n_segs;
                dec_cap = n_decs;
            }
            return res;
        }
    };
};

}  // namespace goetia

#endif
