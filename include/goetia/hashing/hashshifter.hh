/**
 * hashshifter.hh -- FwdLemireShifter / CanLemireShifter (goetia
 * hashshifter.hh:71-205 over rollinghashshifter.hh:52-222): the per-call
 * shifter state, kept on the host as (fw, rc) and advanced by the
 * gt_shifter_* entry points of libgoetia_amd.
 */
#ifndef GOETIA_AMD_HASHSHIFTER_HH
#define GOETIA_AMD_HASHSHIFTER_HH

#include <cstdint>
#include <string>

#include "goetia/goetia_amd.hh"
#include "goetia/hashing/canonical.hh"
#include "goetia/sequences/alphabets.hh"

namespace goetia {

template <class HashType>
class LemireShifter {
  public:
    typedef HashType hash_type;
    typedef typename hash_type::value_type value_type;
    typedef Kmer<hash_type> kmer_type;
    typedef DNA_SIMPLE alphabet;
    static constexpr bool is_canonical = std::is_same<HashType, Canonical<uint64_t>>::value;

    const uint16_t K;

    explicit LemireShifter(uint16_t K) : K(K) {}
    LemireShifter(const std::string& start, uint16_t K) : K(K) { hash_base(start); }
    LemireShifter(const LemireShifter& other) : K(other.K) {}

    hash_type get() const { return make(); }

    // hashshifter.hh:147-153: throws on sequences shorter than K; hashes the
    // first K bytes.
    hash_type hash_base(const std::string& sequence) {
        if (sequence.length() < K) {
            throw InvalidSequenceException("HashShifter::hash_base: Sequence must at least length K");
        }
        return hash_base(sequence.c_str());
    }

    hash_type hash_base(const char* sequence) {
        detail::check(gt_shifter_hash_base(K, reinterpret_cast<const uint8_t*>(sequence), &_fw, &_rc),
                      "hash_base");
        _initialized = true;
        return make();
    }

    // hashshifter.hh:171-175: hashes without touching this shifter's state.
    const hash_type hash(const std::string& sequence) const {
        LemireShifter other(*this);
        return other.hash_base(sequence);
    }

    static hash_type hash(const std::string& sequence, const uint16_t K) {
        LemireShifter s(K);
        return s.hash_base(sequence);
    }

    hash_type shift_right(const char& out, const char& in) {
        if (!_initialized) throw UninitializedShifterException();
        detail::check(gt_shifter_shift_right(K, &_fw, &_rc, (uint8_t)out, (uint8_t)in), "shift_right");
        return make();
    }

    hash_type shift_left(const char& in, const char& out) {
        if (!_initialized) throw UninitializedShifterException();
        detail::check(gt_shifter_shift_left(K, &_fw, &_rc, (uint8_t)in, (uint8_t)out), "shift_left");
        return make();
    }

    bool is_initialized() const { return _initialized; }

  private:
    hash_type make() const {
        if constexpr (is_canonical) {
            return hash_type(_fw, _rc);
        } else {
            return hash_type(_fw);
        }
    }

    uint64_t _fw = 0, _rc = 0;
    bool _initialized = false;
};

typedef LemireShifter<Hash<uint64_t>> FwdLemireShifter;        // hashshifter.hh:204
typedef LemireShifter<Canonical<uint64_t>> CanLemireShifter;   // hashshifter.hh:205

}  // namespace goetia

#endif
