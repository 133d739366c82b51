/**
 * solidifier.hh -- goetia's StreamingSolidFilter<dBG<S, Sh>>::Filter
 * (solidifier.hh:20-80) over gt_solid_filter: filter_sequence for one read,
 * filter_sequences for a batch, both with the reference's per-read result
 * (each read's k-mers inserted and queried in read order).
 */
#ifndef GOETIA_AMD_SOLIDIFIER_HH
#define GOETIA_AMD_SOLIDIFIER_HH

#include <cstdint>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "goetia/dbg.hh"
#include "goetia/goetia_amd.hh"
#include "goetia/processors.hh"

namespace goetia {

template <class GraphType>
struct StreamingSolidFilter {
    typedef GraphType graph_type;

    class Filter {
        std::shared_ptr<graph_type> dbg;

      public:
        const uint16_t K;
        const float min_prop_solid;
        const uint32_t solid_threshold;

        Filter(std::shared_ptr<graph_type> dbg, const float min_prop_solid = 0.75, const uint32_t solid_threshold = 1)
            : dbg(dbg), K(dbg->K), min_prop_solid(min_prop_solid), solid_threshold(solid_threshold) {}

        static std::shared_ptr<Filter> build(std::shared_ptr<graph_type> dbg, const float min_prop_solid = 0.75,
                                             const uint32_t solid_threshold = 1) {
            return std::make_shared<Filter>(dbg, min_prop_solid, solid_threshold);
        }

        std::vector<uint8_t> filter_sequences(const ReadBatch& reads) {
            std::vector<uint8_t> passed(reads.size());
            if (reads.size())
                detail::check(gt_solid_filter(dbg->storage()->handle(), K, graph_type::is_canonical(),
                                              dbg->input_mode(), reinterpret_cast<const uint8_t*>(reads.seq.data()),
                                              reads.offsets.data(), reads.size(), min_prop_solid, solid_threshold,
                                              passed.data()),
                              "StreamingSolidFilter::filter_sequences");
            return passed;
        }

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
