/**
 * diginorm.hh -- goetia's DiginormFilter<dBG<ByteStorage, Sh>> (diginorm.hh:20-125)
 * over gt_diginorm: filter_sequence for one read and filter_sequences for a
 * batch, both with the reference's sequential result (a read is dropped when
 * at least half of its k-mers already reach the cutoff, else inserted before
 * the next read), and the static median_count_at_least / median_in_range
 * helpers over query_sequence.
 */
#ifndef GOETIA_AMD_DIGINORM_HH
#define GOETIA_AMD_DIGINORM_HH

#include <algorithm>
#include <cstdint>
#include <memory>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "goetia/dbg.hh"
#include "goetia/goetia_amd.hh"
#include "goetia/processors.hh"

namespace goetia {

template <class GraphType>
struct DiginormFilter {
    typedef GraphType graph_type;
    typedef typename graph_type::hash_type hash_type;

    // diginorm.hh:39-76: at least min_req = 0.5 + float(n_kmers) / 2 of the
    // read's k-mers have a count of at least cutoff (nothing is inserted).
    static bool median_count_at_least(const std::string& sequence, const unsigned int cutoff, graph_type& graph) {
        const auto counts = graph.query_sequence(sequence);
        const unsigned int min_req = 0.5 + float(sequence.size() - graph.K + 1) / 2;
        unsigned int n = 0;
        for (auto c : counts) n += c >= 0 && (unsigned int)c >= cutoff;
        return n >= min_req;
    }

    // diginorm.hh:78-94: whether the median count (element n / 2 after
    // nth_element) lies in (low, high], and the read's hashes.
    static std::pair<bool, std::vector<hash_type>> median_in_range(const std::string& sequence,
                                                                   const unsigned int low,
                                                                   const unsigned int high, graph_type& graph) {
        std::vector<count_t> counts;
        std::vector<hash_type> hashes;
        graph.query_sequence(sequence, counts, hashes);
        const size_t n = counts.size() / 2;
        std::nth_element(counts.begin(), counts.begin() + n, counts.end());
        const auto median = counts[n];
        return {median > (count_t)low && median <= (count_t)high, hashes};
    }

    class Filter {
        std::shared_ptr<graph_type> graph;

      public:
        const uint16_t K;
        const unsigned int cutoff;

        Filter(std::shared_ptr<graph_type> graph, unsigned int cutoff) : graph(graph), K(graph->K), cutoff(cutoff) {}

        static std::shared_ptr<Filter> build(std::shared_ptr<graph_type> graph, unsigned int cutoff) {
            return std::make_shared<Filter>(graph, cutoff);
        }

        // Passed flag per read, in read order (reads shorter than K do not pass).
        std::vector<uint8_t> filter_sequences(const ReadBatch& reads) {
            std::vector<uint8_t> passed(reads.size());
            if (reads.size())
                detail::check(gt_diginorm(graph->storage()->handle(), K, graph_type::is_canonical(),
                                          graph->input_mode(), reinterpret_cast<const uint8_t*>(reads.seq.data()),
                                          reads.offsets.data(), reads.size(), cutoff, passed.data()),
                              "DiginormFilter::filter_sequences");
            return passed;
        }

        // diginorm.hh:105-119: (passed, len - K + 1).
        std::tuple<bool, uint64_t> filter_sequence(const std::string& sequence) {
            if (sequence.length() < K) throw SequenceLengthException("Sequence must have length >= K");
            ReadBatch b;
            b.push_back(sequence);
            return {filter_sequences(b)[0] != 0, sequence.length() - K + 1};
        }
    };

    using Processor = FilterProcessor<Filter>;
};

}  // namespace goetia

#endif
