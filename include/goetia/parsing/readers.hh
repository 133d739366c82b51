/**
 * readers.hh -- goetia's FastxParser<DNA_SIMPLE> (readers.hh:95-232) over the
 * native parser of libgoetia_amd (gt_fastx_*): FASTA / FASTQ, plain or gzip,
 * records validated and uppercased by DNA_SIMPLE.  Skipped records are
 * counted and never returned; a strict parser throws
 * InvalidCharacterException at the first invalid record.  next_batch() fills a
 * ReadBatch for the batched dBG calls.
 */
#ifndef GOETIA_AMD_READERS_HH
#define GOETIA_AMD_READERS_HH

#include <cstdint>
#include <deque>
#include <ostream>
#include <memory>
#include <optional>
#include <string>
#include <utility>
#include <vector>

#include "goetia/dbg.hh"
#include "goetia/goetia_amd.hh"
#include "goetia/sequences/alphabets.hh"

namespace goetia {

class NoMoreReadsAvailable : public GoetiaException {
  public:
    NoMoreReadsAvailable() : GoetiaException("No more reads available") {}
};

// parsing.hh:25-57
struct Record {
    std::string name;
    std::string sequence;
    std::string quality;

    // parsing.hh:38-49: FASTQ when the record has a quality, else FASTA.
    void write_fastx(std::ostream& output) const {
        if (quality.length() != 0)
            output << "@" << name << '\n' << sequence << '\n' << "+" << '\n' << quality << '\n';
        else
            output << ">" << name << '\n' << sequence << '\n';
    }
};

// parsing.hh:60: one record from each file of a pair, either may be absent.
typedef std::pair<std::optional<Record>, std::optional<Record>> RecordPair;

template <class Alphabet = DNA_SIMPLE>
class FastxParser {
    gt_fastx* _h = nullptr;
    std::deque<Record> _buffered;
    bool _done = false;

  public:
    typedef Alphabet alphabet;
    typedef Record value_type;

    FastxParser(const std::string& filename, bool strict = false, uint32_t min_length = 0) {
        detail::check(gt_fastx_open(filename.c_str(), strict ? 1 : 0, min_length, &_h), "FastxParser");
    }
    FastxParser(const FastxParser&) = delete;
    FastxParser& operator=(const FastxParser&) = delete;
    ~FastxParser() { gt_fastx_close(_h); }

    static std::shared_ptr<FastxParser> build(const std::string& filename, bool strict = false,
                                              uint32_t min_length = 0) {
        return std::make_shared<FastxParser>(filename, strict, min_length);
    }

    gt_fastx* handle() const { return _h; }

    // Up to max_reads records / max_bases bases (an empty batch at the end).
    ReadBatch next_batch(uint64_t max_reads = 1 << 16, uint64_t max_bases = 64ULL << 20) {
        ReadBatch b;
        b.seq.resize(max_bases);
        b.offsets.assign(max_reads + 1, 0);
        uint64_t n = 0;
        const int rc = gt_fastx_next_batch(_h, max_reads, max_bases, reinterpret_cast<uint8_t*>(&b.seq[0]),
                                           b.offsets.data(), nullptr, 0, nullptr, &n);
        b.offsets.resize(n + 1);
        b.seq.resize((size_t)b.offsets[n]);
        if (rc != GT_OK && n == 0) detail::check(rc, "FastxParser::next_batch");
        return b;
    }

    // The same with each record's name and quality (empty for FASTA).
    ReadBatch next_batch(std::vector<std::string>& names, std::vector<std::string>& qualities,
                         uint64_t max_reads = 1 << 16, uint64_t max_bases = 64ULL << 20) {
        ReadBatch b;
        b.seq.resize(max_bases);
        b.offsets.assign(max_reads + 1, 0);
        const uint64_t names_cap = max_reads * 256;
        std::string nm(names_cap, '\0'), q(max_bases, '\0');
        std::vector<int64_t> noff(max_reads + 1, 0);
        std::vector<uint8_t> hq(max_reads, 0);
        uint64_t n = 0;
        const int rc = gt_fastx_next_batch_qual(_h, max_reads, max_bases, reinterpret_cast<uint8_t*>(&b.seq[0]),
                                                b.offsets.data(), reinterpret_cast<uint8_t*>(&q[0]), hq.data(),
                                                &nm[0], names_cap, noff.data(), &n);
        b.offsets.resize(n + 1);
        b.seq.resize((size_t)b.offsets[n]);
        names.clear();
        qualities.clear();
        for (uint64_t i = 0; i < n; ++i) {
            names.push_back(nm.substr(noff[i], noff[i + 1] - noff[i]));
            qualities.push_back(hq[i] ? q.substr(b.offsets[i], b.offsets[i + 1] - b.offsets[i]) : std::string());
        }
        if (rc != GT_OK && n == 0) detail::check(rc, "FastxParser::next_batch");
        return b;
    }

    // FastxParser::next: the next record; empty once the input is exhausted,
    // NoMoreReadsAvailable after that.
    std::optional<Record> next() {
        if (_buffered.empty()) {
            if (_done) throw NoMoreReadsAvailable();
            std::vector<std::string> names, quals;
            ReadBatch b = next_batch(names, quals, 4096, 1 << 20);
            for (uint64_t i = 0; i < b.size(); ++i)
                _buffered.push_back(Record{names[i], b.read(i), quals[i]});
            if (_buffered.empty()) {
                _done = true;
                return {};
            }
        }
        Record r = std::move(_buffered.front());
        _buffered.pop_front();
        return r;
    }

    uint64_t n_parsed() const { return stat(0); }
    uint64_t n_skipped() const { return stat(1); }
    bool is_complete() const { return stat(3) && _buffered.empty(); }

  private:
    uint64_t stat(int which) const {
        uint64_t v[3] = {0, 0, 0};
        int complete = 0;
        detail::check(gt_fastx_stats(_h, &v[0], &v[1], &v[2], &complete), "FastxParser");
        return which == 3 ? (uint64_t)complete : v[which];
    }
};

// SplitPairedReader<FastxParser<DNA_SIMPLE>> (readers.hh:235-334) over the
// library's paired reader: next() gives (left, right), an empty side where
// that file's record was rejected; a pair a strict reader rejects throws
// InvalidCharacterException, files of different lengths and (strict,
// force_name_match) unpaired names throw GoetiaException.
template <class ParserType = FastxParser<>>
class SplitPairedReader {
    gt_pairs* _h = nullptr;

  public:
    typedef ParserType parser_type;
    typedef RecordPair value_type;

    SplitPairedReader(const std::string& left, const std::string& right, bool strict = false,
                      uint32_t min_length = 0, bool force_name_match = false) {
        detail::check(gt_pairs_open(left.c_str(), right.c_str(), strict ? 1 : 0, min_length, force_name_match ? 1 : 0,
                                    &_h),
                      "SplitPairedReader");
    }
    SplitPairedReader(const SplitPairedReader&) = delete;
    SplitPairedReader& operator=(const SplitPairedReader&) = delete;
    ~SplitPairedReader() { gt_pairs_close(_h); }

    static std::shared_ptr<SplitPairedReader> build(const std::string& left, const std::string& right,
                                                    bool strict = false, uint32_t min_length = 0,
                                                    bool force_name_match = false) {
        return std::make_shared<SplitPairedReader>(left, right, strict, min_length, force_name_match);
    }

    gt_pairs* handle() const { return _h; }

    RecordPair next() {
        if (is_complete()) throw NoMoreReadsAvailable();
        std::string seq(1 << 20, '\0'), names(1 << 16, '\0');
        int64_t lens[2], nlens[2];
        int done = 0;
        detail::check(gt_pairs_next(_h, reinterpret_cast<uint8_t*>(&seq[0]), seq.size(), lens, &names[0],
                                    names.size(), nlens, &done),
                      "SplitPairedReader::next");
        RecordPair out;
        if (done) return out;
        size_t at = 0, nat = 0;
        std::optional<Record>* side[2] = {&out.first, &out.second};
        for (int k = 0; k < 2; ++k) {
            if (lens[k] < 0) continue;
            *side[k] = Record{names.substr(nat, (size_t)nlens[k]), seq.substr(at, (size_t)lens[k]), std::string()};
            at += (size_t)lens[k];
            nat += (size_t)nlens[k];
        }
        return out;
    }

    bool is_complete() const { return stat(2) != 0; }
    uint64_t n_parsed() const { return stat(0); }
    uint64_t n_skipped() const { return stat(1); }

  private:
    uint64_t stat(int which) const {
        uint64_t v[2] = {0, 0};
        int complete = 0;
        detail::check(gt_pairs_stats(_h, &v[0], &v[1], &complete), "SplitPairedReader");
        return which == 2 ? (uint64_t)complete : v[which];
    }
};

}  // namespace goetia

#endif
