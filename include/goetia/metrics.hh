/**
 * metrics.hh -- goetia's IntervalCounter (metrics.hh:101-145): poll(incr)
 * adds incr and returns true, resetting the counter, once it reaches the
 * interval.  Host code.
 */
#ifndef GOETIA_AMD_METRICS_HH
#define GOETIA_AMD_METRICS_HH

#include <cstdint>

namespace goetia {

class IntervalCounter {
    uint64_t _counter = 0;
    uint64_t _total = 0;

  public:
    const uint64_t interval;
    static constexpr uint64_t DEFAULT_INTERVAL = 500000;

    explicit IntervalCounter(uint64_t interval = DEFAULT_INTERVAL) : interval(interval) {}

    uint64_t counter() const { return _counter; }
    uint64_t total() const { return _total; }

    bool poll(uint64_t incr = 1) {
        _counter += incr;
        _total += incr;
        if (_counter >= interval) {
            _counter = 0;
            return true;
        }
        return false;
    }
};

}  // namespace goetia

#endif
