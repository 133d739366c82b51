/**
 * compactor.hh -- StreamingCompactor<dBG<S, Sh>>::Compactor's segment finder
 * (goetia compactor.hh:35-534): find_new_segments cuts a read into runs of
 * k-mers new to the graph (compact_segment, :77-145), with anchors and
 * flanks, split at the k-mers that become decision nodes once the read is in
 * the graph.  The graph work is two GPU calls per read (a query of every
 * window and the neighbour kernel's values and presence bits); the two
 * passes over the read are host code, as in goetia.
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
            std::vector<uint64_t> hv, nb;
            std::vector<uint8_t> in, out, present;
            const std::vector<count_t> counts = dbg->query_sequences(b, hv);
            dbg->window_neighbors(b, in, out, nb, present);
            const size_t n = counts.size();

            // Shifts of window i (lo 0 left, 4 right) in the graph or among
            // the new k-mers (filter_nodes with extras).
            auto shifts = [&](size_t i, int lo, std::vector<std::pair<hash_type, char>>& res) {
                res.clear();
                for (int c = 0; c < 4; ++c) {
                    const hash_type h(nb[8 * i + lo + c]);
                    if (((present[i] >> (lo + c)) & 1) || new_kmers.count(h)) res.emplace_back(h, "ACGT"[c]);
                }
            };
            std::vector<std::pair<hash_type, char>> sl;

            // runs of new, not yet seen k-mers (compactor.hh:352-400)
            std::vector<compact_segment> pre;
            compact_segment current;
            hash_type prev_hash, cur_hash;
            bool cur_new = false, prev_new = false, cur_seen = false, prev_seen = false;
            for (size_t pos = 0; pos < n; ++pos) {
                cur_hash = hash_type(hv[pos]);
                cur_new = counts[pos] == 0;
                cur_seen = new_kmers.count(cur_hash) != 0;
                hashes.push_back(cur_hash);
                if (cur_new && !cur_seen) {
                    if (!prev_new || prev_seen) {
                        pre.push_back(current);
                        current = init_segment(cur_hash, prev_hash, pos);
                    }
                    new_kmers.insert(cur_hash);
                } else if (!current.is_null() && (!cur_new || cur_seen)) {
                    finish_segment(current, pos - 1, prev_hash, cur_hash, pre);
                    current = compact_segment();
                }
                prev_hash = cur_hash;
                prev_new = cur_new;
                prev_seen = cur_seen;
            }
            if (cur_new && !cur_seen) {
                hash_type right_flank = cur_hash;
                shifts(n - 1, 4, sl);
                if (sl.size() == 1) right_flank = sl.front().first;
                finish_segment(current, n - 1, cur_hash, right_flank, pre);
            }
            pre.push_back(compact_segment());
            if (pre.size() < 3) return;
            if (pre[1].start_pos == 0) {   // compactor.hh:425-437
                shifts(0, 0, sl);
                pre[1].left_flank = sl.size() == 1 ? sl.front().first : pre[1].left_anchor;
            }

            // cuts at decision k-mers (compactor.hh:439-532)
            std::vector<std::pair<hash_type, char>> sr;
            for (auto segment : pre) {
                if (segment.is_null()) {
                    segments.push_back(segment);
                    continue;
                }
                std::vector<compact_segment> decisions;
                for (size_t pos = segment.start_pos; pos + K <= segment.start_pos + segment.length; ++pos) {
                    shifts(pos, 0, sl);
                    shifts(pos, 4, sr);
                    if (sl.size() > 1 || sr.size() > 1) {
                        const std::string root = sequence.substr(pos, K);
                        std::vector<kmer_type> lk, rk;
                        for (auto& x : sl) lk.push_back(kmer_type(x.first, x.second + root.substr(0, K - 1)));
                        for (auto& x : sr) rk.push_back(kmer_type(x.first, root.substr(1) + x.second));
                        decision_neighbors.push_back(std::make_pair(lk, rk));
                        const hash_type h(hv[pos]);
                        new_decision_kmers.insert(h);
                        compact_segment d = init_segment(h, h, pos);
                        d.length = K;
                        d.right_anchor = d.left_anchor;
                        d.right_flank = d.left_flank;
                        d.is_decision_kmer = true;
                        decisions.push_back(d);
                    }
                }
                split(segment, decisions, hashes, segments);
            }
        }

      private:
        uint64_t _minimizer_window_size;

        static compact_segment init_segment(hash_type left_anchor, hash_type left_flank, size_t start_pos) {
            compact_segment s;
            s.start_pos = start_pos;
            s.left_anchor = left_anchor;
            s.left_flank = left_flank;
            s.is_decision_kmer = false;
            return s;
        }

        void finish_segment(compact_segment& s, size_t end, hash_type right_anchor, hash_type right_flank,
                            std::vector<compact_segment>& out) {
            s.length = end - s.start_pos + K;
            s.right_anchor = right_anchor;
            s.right_flank = right_flank;
            out.push_back(s);
        }

        // compactor.hh:471-525 take hashes[start_pos +- 1] even where that is
        // off the read; the segment built from it is then shorter than K and
        // dropped, so the value never shows.
        static hash_type at(const std::vector<hash_type>& hashes, size_t i, bool ok) {
            return ok && i < hashes.size() ? hashes[i] : hash_type();
        }

        void split(compact_segment& seg, const std::vector<compact_segment>& decisions,
                   const std::vector<hash_type>& hashes, std::vector<compact_segment>& segments) {
            if (decisions.empty()) {
                segments.push_back(seg);
                return;
            }
            if (decisions.size() == 1) {
                const compact_segment& d = decisions.front();
                compact_segment right;
                right.left_flank = d.right_anchor;
                right.left_anchor = at(hashes, d.start_pos + 1, true);
                right.right_anchor = seg.right_anchor;
                right.right_flank = seg.right_flank;
                right.start_pos = d.start_pos + 1;
                right.length = seg.length - (right.start_pos - seg.start_pos);
                seg.right_anchor = at(hashes, d.start_pos - 1, d.start_pos > 0);
                seg.right_flank = d.left_anchor;
                seg.length = d.start_pos + K - 1 - seg.start_pos;
                if (seg.length >= K) segments.push_back(seg);
                segments.push_back(d);
                if (right.length >= K) segments.push_back(right);
                return;
            }
            const compact_segment& fd = decisions.front();
            compact_segment first;
            first.left_flank = seg.left_flank;
            first.left_anchor = seg.left_anchor;
            first.right_anchor = at(hashes, fd.start_pos - 1, fd.start_pos > 0);
            first.right_flank = at(hashes, fd.start_pos, true);
            first.start_pos = seg.start_pos;
            first.length = fd.start_pos + K - 1 - seg.start_pos;
            if (first.length >= K) segments.push_back(first);
            segments.push_back(fd);
            for (size_t j = 1; j < decisions.size(); ++j) {
                const compact_segment& u = decisions[j - 1];
                const compact_segment& v = decisions[j];
                compact_segment mid;
                mid.left_flank = u.left_anchor;
                mid.left_anchor = at(hashes, u.start_pos + 1, true);
                mid.right_anchor = at(hashes, v.start_pos - 1, true);
                mid.right_flank = v.left_anchor;
                mid.start_pos = u.start_pos + 1;
                mid.length = v.start_pos + K - 1 - mid.start_pos;
                if (mid.length >= K) segments.push_back(mid);
                segments.push_back(v);
            }
            const compact_segment& ld = decisions.back();
            compact_segment last;
            last.right_anchor = seg.right_anchor;
            last.right_flank = seg.right_flank;
            last.left_flank = ld.right_anchor;
            last.left_anchor = at(hashes, ld.start_pos + 1, true);
            last.start_pos = ld.start_pos + 1;
            last.length = seg.length - (last.start_pos - seg.start_pos);
            if (last.length >= K) segments.push_back(last);
        }
    };
};

}  // namespace goetia

#endif
