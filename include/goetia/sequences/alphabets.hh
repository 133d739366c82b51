/**
 * alphabets.hh -- DNA_SIMPLE (goetia alphabets.hh:107-161): symbols ACGT,
 * complement A<->T, C<->G, anything else -> '\0'; validate uppercases acgt.
 */
#ifndef GOETIA_AMD_ALPHABETS_HH
#define GOETIA_AMD_ALPHABETS_HH

#include <string>
#include <string_view>

namespace goetia {

struct DNA_SIMPLE {
    static constexpr auto SYMBOLS = std::string_view("ACGT");
    static constexpr auto COMPLEMENTS = std::string_view("TGCA");

    static char validate(const char c) {
        switch (c) {
            case 'A': case 'C': case 'G': case 'T': return c;
            case 'a': return 'A';
            case 'c': return 'C';
            case 'g': return 'G';
            case 't': return 'T';
            default: return '\0';
        }
    }

    static char complement(const char c) {
        switch (c) {
            case 'A': return 'T';
            case 'T': return 'A';
            case 'C': return 'G';
            case 'G': return 'C';
            default: return '\0';
        }
    }

    static std::string reverse_complement(const std::string& s) {
        std::string out(s.rbegin(), s.rend());
        for (auto& c : out) c = complement(c);
        return out;
    }
};

}  // namespace goetia

#endif
