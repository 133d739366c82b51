/**
 * processors.hh -- goetia's InserterProcessor<dBG<S, Sh>> (processors.hh:248-344)
 * on the native chunk driver (gt_inserter_*) and FilterProcessor<Filter>
 * (processors.hh:355-440), both over FileProcessor::advance / process
 * (processors.hh:45-236): reads are parsed and processed in batches on the
 * GPU, and advance() returns right after the read whose k-mers bring the
 * IntervalCounter to the interval, with the running (n_sequences, time,
 * remaining) goetia returns.
 */
#ifndef GOETIA_AMD_PROCESSORS_HH
#define GOETIA_AMD_PROCESSORS_HH

#include <cstdint>
#include <fstream>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

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

    // FileProcessor::advance over a split pair of files (processors.hh:180-204).
    std::tuple<uint64_t, uint64_t, bool> advance(std::shared_ptr<SplitPairedReader<ParserType>>& reader) {
        uint64_t n = 0, t = 0;
        int remaining = 0;
        detail::check(gt_inserter_advance_pairs(_h, reader->handle(), &n, &t, &remaining),
                      "InserterProcessor::advance");
        return {n, t, remaining != 0};
    }

    std::tuple<uint64_t, uint64_t> process(std::shared_ptr<ParserType>& parser) {
        uint64_t n = 0, t = 0;
        bool remaining = true;
        while (remaining) std::tie(n, t, remaining) = advance(parser);
        return {n, t};
    }

    std::tuple<uint64_t, uint64_t> process(std::shared_ptr<SplitPairedReader<ParserType>>& reader) {
        uint64_t n = 0, t = 0;
        bool remaining = true;
        while (remaining) std::tie(n, t, remaining) = advance(reader);
        return {n, t};
    }

    std::tuple<uint64_t, uint64_t> process(const std::string& filename, bool strict = false,
                                           uint32_t min_length = 0) {
        auto parser = ParserType::build(filename, strict, min_length);
        return process(parser);
    }

    // FileProcessor::process(left, right, ...) (processors.hh:69-94).
    std::tuple<uint64_t, uint64_t> process(const std::string& left_filename, const std::string& right_filename,
                                           bool strict = false, uint32_t min_length = 0,
                                           bool force_name_match = false) {
        auto reader =
            SplitPairedReader<ParserType>::build(left_filename, right_filename, strict, min_length, force_name_match);
        return process(reader);
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

// FilterProcessor<FilterType> (processors.hh:355-440): every read goes
// through the filter (StreamingSolidFilter or DiginormFilter ::Filter) and the
// reads that pass are written with Record::write_fastx.  A read's time is its
// k-mer count (0 below K); the filter runs over the reads up to each interval
// return, so the graph at a return holds exactly the reads before it.
template <class FilterType, class ParserType = FastxParser<>>
class FilterProcessor {
    std::shared_ptr<FilterType> filter;
    std::ofstream _output_stream;
    IntervalCounter timer;
    uint64_t _n_sequences = 0, _n_passed = 0;
    bool _verbose;
    // parsed reads not yet consumed
    ReadBatch _batch;
    std::vector<std::string> _names, _quals;
    uint64_t _next = 0;

  public:
    FilterProcessor(std::shared_ptr<FilterType> filter, const std::string output_filename,
                    uint64_t interval = IntervalCounter::DEFAULT_INTERVAL, bool verbose = false)
        : filter(filter), _output_stream(output_filename.c_str()), timer(interval), _verbose(verbose) {}

    static std::shared_ptr<FilterProcessor> build(std::shared_ptr<FilterType> filter,
                                                  const std::string& output_filename,
                                                  uint64_t interval = IntervalCounter::DEFAULT_INTERVAL,
                                                  bool verbose = false) {
        return std::make_shared<FilterProcessor>(filter, output_filename, interval, verbose);
    }

    std::tuple<uint64_t, uint64_t, bool> advance(std::shared_ptr<ParserType>& parser) {
        const uint64_t K = filter->K;
        for (;;) {
            if (_next >= _batch.size()) {
                _batch = parser->next_batch(_names, _quals, 1 << 16, 16ULL << 20);
                _next = 0;
                if (_batch.size() == 0) {
                    _output_stream.flush();
                    return {_n_sequences, timer.total(), false};
                }
            }
            // reads [i, j): up to and including the one that reaches the interval
            const uint64_t i = _next;
            uint64_t j = i, t = 0, c = timer.counter();
            bool reach = false;
            while (j < _batch.size() && !reach) {
                const int64_t len = _batch.offsets[j + 1] - _batch.offsets[j];
                const uint64_t tj = len >= (int64_t)K ? (uint64_t)(len - (int64_t)K + 1) : 0;
                t += tj;
                c += tj;
                reach = c >= timer.interval;
                ++j;
            }
            ReadBatch sub;
            sub.seq = _batch.seq.substr(_batch.offsets[i], _batch.offsets[j] - _batch.offsets[i]);
            sub.offsets.assign(1, 0);
            for (uint64_t r = i; r < j; ++r) sub.offsets.push_back(_batch.offsets[r + 1] - _batch.offsets[i]);
            const auto passed = filter->filter_sequences(sub);
            for (uint64_t r = i; r < j; ++r) {
                if (!passed[r - i]) continue;
                ++_n_passed;
                Record{_names[r], _batch.read(r), _quals[r]}.write_fastx(_output_stream);
            }
            const bool crossed = timer.poll(t);
            _n_sequences += j - i;
            _next = j;
            if (crossed) {
                _output_stream.flush();
                return {_n_sequences, timer.total(), true};
            }
        }
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

    uint64_t n_passed() const { return _n_passed; }
    uint64_t n_sequences() const { return _n_sequences; }
    uint64_t time_elapsed() const { return timer.total(); }
};

}  // namespace goetia

#endif
