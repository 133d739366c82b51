/**
 * processors.hh -- goetia's InserterProcessor<dBG<S, Sh>> (processors.hh:248-344)
 * over FileProcessor::advance / process (processors.hh:45-236), on the native
 * chunk driver (gt_inserter_*): reads are parsed and inserted in batches on
 * the GPU, and advance() returns right after the read whose k-mers bring the
 * IntervalCounter to the interval, with the running (n_sequences, time,
 * remaining) goetia returns.
 */
#ifndef GOETIA_AMD_PROCESSORS_HH
#define GOETIA_AMD_PROCESSORS_HH

#include <cstdint>
#include <memory>
#include <string>
#include <tuple>

#include "goetia/goetia_amd.hh"
#include "goetia/metrics.hh"
#include "goetia/parsing/readers.hh"

namespace goetia {

template <class InserterType, class ParserType = FastxParser<>>
class InserterProcessor {
    gt_inserter* _h = nullptr;
    uint64_t _interval;

  public:
    std::shared_ptr<InserterType> inserter;

    InserterProcessor(std::shared_ptr<InserterType> inserter, uint64_t interval = IntervalCounter::DEFAULT_INTERVAL,
                      bool verbose = false)
        : _interval(interval), inserter(inserter) {
        (void)verbose;
        detail::check(gt_inserter_new(inserter->storage()->handle(), inserter->K, InserterType::is_canonical(),
                                      inserter->input_mode(), interval, 0, &_h),
                      "InserterProcessor");
    }
    InserterProcessor(const InserterProcessor&) = delete;
    InserterProcessor& operator=(const InserterProcessor&) = delete;
    ~InserterProcessor() { gt_inserter_free(_h); }

    static std::shared_ptr<InserterProcessor> build(std::shared_ptr<InserterType> inserter,
                                                    uint64_t interval = IntervalCounter::DEFAULT_INTERVAL,
                                                    bool verbose = false) {
        return std::make_shared<InserterProcessor>(inserter, interval, verbose);
    }

    std::tuple<uint64_t, uint64_t, bool> advance(std::shared_ptr<ParserType>& parser) {
        uint64_t n = 0, t = 0;
        int remaining = 0;
        detail::check(gt_inserter_advance(_h, parser->handle(), &n, &t, &remaining), "InserterProcessor::advance");
        return {n, t, remaining != 0};
    }

    std::tuple<uint64_t, uint64_t> process(std::shared_ptr<ParserType>& parser) {
        uint64_t n = 0, t = 0;
        bool remaining = true;
        while (remaining) std::tie(n, t, remaining) = advance(parser);
        return {n, t};
    }

    std::tuple<uint64_t, uint64_t> process(const std::string& filename, bool strict = false,
                                           uint32_t min_length = 0) {
        auto parser = ParserType::build(filename, strict, min_length);
        return process(parser);
    }

    uint64_t n_sequences() const { return stat(0); }
    uint64_t time_elapsed() const { return stat(1); }
    uint64_t interval() const { return _interval; }

  private:
    uint64_t stat(int which) const {
        uint64_t v[3] = {0, 0, 0};
        detail::check(gt_inserter_stats(_h, &v[0], &v[1], &v[2]), "InserterProcessor");
        return v[which];
    }
};

}  // namespace goetia

#endif
