/**
 * kmeriterator.hh -- KmerIterator<Shifter> (goetia kmeriterator.hh:32-141):
 * the ordered hashes of every length-K window of a sequence.  The whole
 * sequence is hashed on the GPU (gt_hash_reads) when the iterator is built.
 */
#ifndef GOETIA_AMD_KMERITERATOR_HH
#define GOETIA_AMD_KMERITERATOR_HH

#include <string>
#include <vector>

#include "goetia/hashing/hashshifter.hh"

namespace goetia {

template <class ShifterType>
class KmerIterator {
  public:
    typedef ShifterType shifter_type;
    typedef typename ShifterType::hash_type hash_type;
    const uint16_t K;

    KmerIterator(const std::string& seq, uint16_t K) : K(K), _seq(seq) { load(); }
    KmerIterator(const std::string& seq, ShifterType* shifter) : K(shifter->K), _seq(seq) { load(); }
    KmerIterator(const std::string& seq, ShifterType& proto) : K(proto.K), _seq(seq) { load(); }

    hash_type first() {
        _initialized = true;
        index = 1;
        return at(0);
    }

    hash_type next() {
        if (!_initialized) return first();
        if (done()) return at(index - 1);
        return at(index++);
    }

    bool done() const { return index + K > _seq.length(); }
    unsigned int get_start_pos() const { return _initialized ? index - 1 : 0; }
    unsigned int get_end_pos() const { return _initialized ? index + K - 1 : K; }

  private:
    void load() {
        if (_seq.length() < K) throw SequenceLengthException("Sequence must have length >= K");
        const size_t n = _seq.length() - K + 1;
        _fw.resize(n);
        _rc.resize(n);
        const int64_t off[2] = {0, (int64_t)_seq.length()};
        detail::check(gt_hash_reads(K, GT_MODE_RAW, reinterpret_cast<const uint8_t*>(_seq.data()), off, 1,
                                    _fw.data(), _rc.data(), nullptr),
                      "KmerIterator");
    }

    hash_type at(size_t i) const {
        if constexpr (ShifterType::is_canonical) {
            return hash_type(_fw[i], _rc[i]);
        } else {
            return hash_type(_fw[i]);
        }
    }

    const std::string _seq;
    std::vector<uint64_t> _fw, _rc;
    unsigned int index = 0;
    bool _initialized = false;
};

}  // namespace goetia

#endif
