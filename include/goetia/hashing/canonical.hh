/**
 * canonical.hh -- hash models (goetia canonical.hh:30-147, 357-412):
 * Hash (a single strand-specific value), Canonical (fw and rc, value =
 * min), Shift (a hash plus the symbol that produced it) and Kmer.
 */
#ifndef GOETIA_AMD_CANONICAL_HH
#define GOETIA_AMD_CANONICAL_HH

#include <cstdint>
#include <limits>
#include <ostream>
#include <string>

namespace goetia {

inline constexpr bool DIR_LEFT = false;
inline constexpr bool DIR_RIGHT = true;

template <class ValueType = uint64_t>
struct Hash {
    typedef ValueType value_type;
    value_type hash;

    Hash(const value_type h) : hash(h) {}
    Hash() : hash(std::numeric_limits<value_type>::max()) {}
    const value_type value() const { return hash; }
    operator value_type() const { return value(); }
    friend bool operator<(const Hash& a, const Hash& b) { return a.value() < b.value(); }
    friend bool operator>(const Hash& a, const Hash& b) { return a.value() > b.value(); }
    friend bool operator==(const Hash& a, const Hash& b) { return a.value() == b.value(); }
};

template <class ValueType = uint64_t>
struct Canonical {
    typedef ValueType value_type;
    value_type fw_hash, rc_hash;

    Canonical() : fw_hash(std::numeric_limits<value_type>::max()), rc_hash(std::numeric_limits<value_type>::max()) {}
    Canonical(const value_type fw, const value_type rc) : fw_hash(fw), rc_hash(rc) {}
    const bool sign() const { return fw_hash < rc_hash; }
    const value_type value() const { return fw_hash < rc_hash ? fw_hash : rc_hash; }
    operator value_type() const { return value(); }
    friend bool operator<(const Canonical& a, const Canonical& b) { return a.value() < b.value(); }
    friend bool operator>(const Canonical& a, const Canonical& b) { return a.value() > b.value(); }
    friend bool operator==(const Canonical& a, const Canonical& b) { return a.value() == b.value(); }
};

template <class V>
inline std::ostream& operator<<(std::ostream& os, const Hash<V>& h) {
    return os << "<Hash h=" << h.value() << ">";
}

template <class V>
inline std::ostream& operator<<(std::ostream& os, const Canonical<V>& h) {
    return os << "<Canonical fw=" << h.fw_hash << " rc=" << h.rc_hash << " sign=" << h.sign() << ">";
}

template <class HashType, bool Direction>
struct Shift {
    typedef HashType hash_type;
    typedef typename hash_type::value_type value_type;
    static const bool direction = Direction;
    hash_type hash;
    char symbol;

    Shift(const hash_type h, const char s) : hash(h), symbol(s) {}
    Shift() : hash(), symbol('\0') {}
    const value_type value() const { return hash.value(); }
    operator value_type() const { return hash.value(); }
    operator hash_type() const { return hash; }
    friend bool operator==(const Shift& a, const Shift& b) { return a.hash == b.hash; }
    friend bool operator<(const Shift& a, const Shift& b) { return a.hash < b.hash; }
};

template <class HashType>
struct Kmer {
    typedef HashType hash_type;
    hash_type hash;
    std::string kmer;
    Kmer() = default;
    Kmer(const hash_type& h, const std::string& k) : hash(h), kmer(k) {}
    const typename hash_type::value_type value() const { return hash.value(); }
};

}  // namespace goetia

#endif
