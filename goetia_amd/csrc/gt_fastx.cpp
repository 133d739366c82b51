// gt_fastx.cpp -- FASTA / FASTQ reading into read batches, and the chunk
// driver that feeds them to a dBG with goetia's interval semantics.
//
// Parsing follows goetia's FastxParser<DNA_SIMPLE> (readers.hh:148-219) over
// the vendored kseq.h reader record for record: the same record boundaries
// (headers start at '>' or '@', sequence lines run to a line starting with
// '>', '+' or '@', the quality of a FASTQ record is read line by line until it
// is at least as long as the sequence), the same line-end handling (a line's
// trailing '\r' is dropped), names up to the first white space, and the same
// treatment of bad records: DNA_SIMPLE uppercases a/c/g/t and rejects any
// other byte outside A/C/G/T (alphabets.hh:41-55, 112-130), which skips the
// record or, strict, stops the parse; records shorter than min_length are
// skipped; a quality of another length is an invalid read.  n_parsed and
// n_skipped count as goetia's do.  Host code only: no device is needed.
#include "goetia_amd.h"

#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cstring>
#include <string>
#include <vector>

namespace gt {
int set_error(int code, const std::string& msg);
}

namespace {

using gt::set_error;

constexpr size_t kReadBuf = 1 << 20;
constexpr int kSepSpace = 0;
constexpr int kSepLine = 2;

// Buffered reader with kseq's kstream semantics (ks_getc, ks_getuntil2).
struct Stream {
    gzFile f = nullptr;
    std::vector<unsigned char> buf = std::vector<unsigned char>(kReadBuf);
    int64_t begin = 0, end = 0;
    bool is_eof = false;

    bool err() const { return end == -1; }
    bool eof() const { return is_eof && begin >= end; }

    bool fill() {
        begin = 0;
        const int n = gzread(f, buf.data(), (unsigned)kReadBuf);
        end = n < 0 ? -1 : n;
        if (end <= 0) is_eof = true;
        return end > 0;
    }

    int getc() {
        if (err()) return -3;
        if (is_eof && begin >= end) return -1;
        if (begin >= end && !fill()) return err() ? -3 : -1;
        return buf[begin++];
    }

    // Appends (or sets) s up to the delimiter, which is consumed; *dret gets
    // the delimiter byte.  Returns the length, -1 at end of input with nothing
    // read, -3 on a read error.
    int64_t getuntil(int delim, std::string& s, int* dret, bool append) {
        bool gotany = false;
        if (dret) *dret = 0;
        if (!append) s.clear();
        for (;;) {
            if (err()) return -3;
            if (begin >= end) {
                if (is_eof) break;
                if (!fill()) {
                    if (err()) return -3;
                    break;
                }
            }
            int64_t i = begin;
            if (delim == kSepLine) {
                const void* p = std::memchr(buf.data() + begin, '\n', (size_t)(end - begin));
                i = p ? (const unsigned char*)p - buf.data() : end;
            } else {
                while (i < end && !std::isspace(buf[i])) ++i;
            }
            s.append((const char*)buf.data() + begin, (size_t)(i - begin));
            gotany = true;
            begin = i + 1;
            if (i < end) {
                if (dret) *dret = buf[i];
                break;
            }
        }
        if (!gotany && eof()) return -1;
        if (delim == kSepLine && s.size() > 1 && s.back() == '\r') s.pop_back();
        return (int64_t)s.size();
    }
};

// DNA_SIMPLE::_validate (alphabets.hh:112-130): A/C/G/T kept, a/c/g/t
// uppercased, anything else invalid (0).
struct Validator {
    unsigned char map[256];
    Validator() {
        std::memset(map, 0, sizeof(map));
        for (const char* c = "ACGT"; *c; ++c) {
            map[(unsigned char)*c] = (unsigned char)*c;
            map[(unsigned char)std::tolower(*c)] = (unsigned char)*c;
        }
    }
};
const Validator kValidator;

enum Next { kRecord, kSkipped, kDone, kInvalidRead, kIoError, kInvalidChar };

}  // namespace

struct gt_fastx {
    Stream ks;
    int last_char = 0;
    bool strict = false;
    uint32_t min_length = 0;
    uint64_t n_parsed = 0, n_skipped = 0, n_invalid_reads = 0;
    bool complete = false;
    std::string name, comment, seq, qual;
    bool pending = false;                  // seq / name hold a record not yet handed out
    std::string path;
    std::string last_bad;                  // sequence of the record that stopped a strict parse
};

namespace {

// kseq_read: >= 0 sequence length, -1 end of input, -2 truncated or
// mismatched quality, -3 read error.
int64_t read_record(gt_fastx& p) {
    Stream& ks = p.ks;
    int c;
    if (p.last_char == 0) {
        while ((c = ks.getc()) >= 0 && c != '>' && c != '@') {
        }
        if (c < 0) return c;
        p.last_char = c;
    }
    p.comment.clear();
    p.seq.clear();
    p.qual.clear();
    const int64_t r = ks.getuntil(kSepSpace, p.name, &c, false);
    if (r < 0) return r;
    if (c != '\n') ks.getuntil(kSepLine, p.comment, nullptr, false);
    while ((c = ks.getc()) >= 0 && c != '>' && c != '+' && c != '@') {
        if (c == '\n') continue;
        p.seq.push_back((char)c);
        ks.getuntil(kSepLine, p.seq, nullptr, true);
    }
    if (c == '>' || c == '@') p.last_char = c;
    if (c != '+') return (int64_t)p.seq.size();
    while ((c = ks.getc()) >= 0 && c != '\n') {
    }
    if (c == -1) return -2;
    while (ks.getuntil(kSepLine, p.qual, nullptr, true) >= 0 && p.qual.size() < p.seq.size()) {
    }
    p.last_char = 0;
    if (p.seq.size() != p.qual.size()) return -2;
    return (int64_t)p.seq.size();
}

// FastxParser::next without the exceptions.
Next next_record(gt_fastx& p) {
    if (p.complete) return kDone;
    int64_t stat = read_record(p);
    if (stat >= 0) {
        bool bad = false;
        for (char& ch : p.seq) {
            const unsigned char v = kValidator.map[(unsigned char)ch];
            if (!v) {
                bad = true;
                break;
            }
            ch = (char)v;
        }
        if (bad) {
            ++p.n_skipped;
            if (p.strict) {
                p.last_bad = p.seq;
                return kInvalidChar;
            }
            stat = -4;
        }
        if (p.seq.size() < p.min_length) {
            stat = -4;
            ++p.n_skipped;
        }
        ++p.n_parsed;
    }
    if (stat == -1) {
        p.complete = true;
        return kDone;
    }
    if (stat == -2) {
        ++p.n_skipped;
        ++p.n_invalid_reads;
        return kInvalidRead;
    }
    if (stat == -3) return kIoError;
    if (stat == -4) return kSkipped;
    return kRecord;
}

}  // namespace

extern "C" {

int gt_fastx_open(const char* path, int strict, uint32_t min_length, gt_fastx** out) {
    if (!path || !out) return set_error(GT_ERR_INVALID_ARG, "null path or output");
    *out = nullptr;
    gzFile f = gzopen(path, "r");
    if (!f) return set_error(GT_ERR_IO, std::string("cannot open ") + path);
    gzbuffer(f, 1 << 20);
    gt_fastx* p = new (std::nothrow) gt_fastx();
    if (!p) {
        gzclose(f);
        return set_error(GT_ERR_OUT_OF_MEMORY, "gt_fastx_open");
    }
    p->ks.f = f;
    p->strict = strict != 0;
    p->min_length = min_length;
    p->path = path;
    *out = p;
    return GT_OK;
}

int gt_fastx_next_batch(gt_fastx* p, uint64_t max_reads, uint64_t max_bases, uint8_t* seq, int64_t* offsets,
                        char* names, uint64_t names_cap, int64_t* name_offsets, uint64_t* n_reads) {
    if (!p || !n_reads || !offsets || (!seq && max_bases)) return set_error(GT_ERR_INVALID_ARG, "null argument");
    if (names && !name_offsets) return set_error(GT_ERR_INVALID_ARG, "names need name_offsets");
    uint64_t n = 0, bases = 0, nbytes = 0;
    offsets[0] = 0;
    if (names) name_offsets[0] = 0;
    *n_reads = 0;
    for (;;) {
        if (!p->pending) {
            const Next r = next_record(*p);
            if (r == kDone) break;
            if (r == kSkipped || r == kInvalidRead) continue;
            if (r == kIoError) {
                *n_reads = n;
                return set_error(GT_ERR_IO, "error reading " + p->path);
            }
            if (r == kInvalidChar) {
                *n_reads = n;
                return set_error(GT_ERR_INVALID_CHARACTER,
                                 "Alphabet: Invalid symbol in sequence " + p->last_bad + " (alphabet=ACGT).");
            }
            p->pending = true;
        }
        const uint64_t len = p->seq.size();
        const bool fits = n < max_reads && bases + len <= max_bases && (!names || nbytes + p->name.size() <= names_cap);
        if (!fits) {
            if (n == 0) {
                return set_error(GT_ERR_INVALID_ARG, "record " + p->name + " (" + std::to_string(len) +
                                                         " bases) does not fit the batch buffers");
            }
            break;
        }
        std::memcpy(seq + bases, p->seq.data(), len);
        bases += len;
        offsets[n + 1] = (int64_t)bases;
        if (names) {
            std::memcpy(names + nbytes, p->name.data(), p->name.size());
            nbytes += p->name.size();
            name_offsets[n + 1] = (int64_t)nbytes;
        }
        ++n;
        p->pending = false;
    }
    *n_reads = n;
    return GT_OK;
}

int gt_fastx_stats(const gt_fastx* p, uint64_t* n_parsed, uint64_t* n_skipped, uint64_t* n_invalid_reads,
                   int* complete) {
    if (!p) return set_error(GT_ERR_INVALID_ARG, "null parser");
    if (n_parsed) *n_parsed = p->n_parsed;
    if (n_skipped) *n_skipped = p->n_skipped;
    if (n_invalid_reads) *n_invalid_reads = p->n_invalid_reads;
    if (complete) *complete = p->complete && !p->pending;
    return GT_OK;
}

int gt_fastx_close(gt_fastx* p) {
    if (!p) return GT_OK;
    if (p->ks.f) gzclose(p->ks.f);
    delete p;
    return GT_OK;
}

}  // extern "C"

// ------------------------------------------------------------ chunk driver

struct gt_inserter {
    gt_storage* st = nullptr;
    uint16_t K = 0;
    int canonical = 1, mode = GT_MODE_RAW;
    uint64_t interval = 500000;
    uint64_t counter = 0, total = 0, n_sequences = 0;
    uint64_t max_reads = 0, max_bases = 0;
    std::vector<uint8_t> seq;
    std::vector<int64_t> off, sub_off;
    uint64_t n = 0, cur = 0;              // parsed reads in the buffer, next one to insert
};

extern "C" {

int gt_inserter_new(gt_storage* st, uint16_t K, int canonical, int mode, uint64_t interval, uint64_t batch_bases,
                    gt_inserter** out) {
    if (!st || !out) return set_error(GT_ERR_INVALID_ARG, "null storage or output");
    if (K < 1 || K > 255) return set_error(GT_ERR_INVALID_ARG, "K must be in 1..255");
    if (interval == 0) return set_error(GT_ERR_INVALID_ARG, "interval must be positive");
    gt_inserter* ins = new (std::nothrow) gt_inserter();
    if (!ins) return set_error(GT_ERR_OUT_OF_MEMORY, "gt_inserter_new");
    ins->st = st;
    ins->K = K;
    ins->canonical = canonical;
    ins->mode = mode;
    ins->interval = interval;
    ins->max_bases = batch_bases ? batch_bases : (64ULL << 20);
    ins->max_reads = std::max<uint64_t>(ins->max_bases / 64, 1024);
    try {
        ins->seq.resize(ins->max_bases);
        ins->off.resize(ins->max_reads + 1);
        ins->sub_off.resize(ins->max_reads + 1);
    } catch (...) {
        delete ins;
        return set_error(GT_ERR_OUT_OF_MEMORY, "gt_inserter_new buffers");
    }
    *out = ins;
    return GT_OK;
}

// InserterProcessor / FileProcessor::advance (processors.hh:206-231, 304-306)
// with IntervalCounter::poll (metrics.hh:126-135): reads are inserted until
// the k-mers of the reads taken since the last interval reach `interval`
// (each read counts len - K + 1, 0 when shorter than K); the call returns
// right after the read that reaches it, or when the input is exhausted.
int gt_inserter_advance(gt_inserter* ins, gt_fastx* parser, uint64_t* n_sequences, uint64_t* time_total,
                        int* remaining) {
    if (!ins || !parser) return set_error(GT_ERR_INVALID_ARG, "null inserter or parser");
    for (;;) {
        if (ins->cur >= ins->n) {
            ins->n = ins->cur = 0;
            uint64_t got = 0;
            const int rc = gt_fastx_next_batch(parser, ins->max_reads, ins->max_bases, ins->seq.data(),
                                               ins->off.data(), nullptr, 0, nullptr, &got);
            // FileProcessor::handle_next (processors.hh:149-171) skips a record
            // with an invalid character even from a strict parser
            if (rc != GT_OK && rc != GT_ERR_INVALID_CHARACTER) return rc;
            ins->n = got;
            if (got == 0 && rc == GT_ERR_INVALID_CHARACTER) continue;
            if (got == 0) {
                if (n_sequences) *n_sequences = ins->n_sequences;
                if (time_total) *time_total = ins->total;
                if (remaining) *remaining = 0;
                return GT_OK;
            }
        }
        uint64_t end = ins->cur;
        bool crossed = false;
        while (end < ins->n) {
            const int64_t len = ins->off[end + 1] - ins->off[end];
            const uint64_t t = len >= ins->K ? (uint64_t)(len - ins->K + 1) : 0;
            ++end;
            ++ins->n_sequences;
            ins->total += t;
            ins->counter += t;
            if (ins->counter >= ins->interval) {
                ins->counter = 0;
                crossed = true;
                break;
            }
        }
        const uint64_t nr = end - ins->cur;
        const int64_t b0 = ins->off[ins->cur];
        for (uint64_t i = 0; i <= nr; ++i) ins->sub_off[i] = ins->off[ins->cur + i] - b0;
        const int rc = gt_dbg_insert(ins->st, ins->K, ins->canonical, ins->mode, ins->seq.data() + b0,
                                     ins->sub_off.data(), nr, nullptr, nullptr, nullptr);
        if (rc != GT_OK) return rc;
        ins->cur = end;
        if (crossed) {
            if (n_sequences) *n_sequences = ins->n_sequences;
            if (time_total) *time_total = ins->total;
            if (remaining) *remaining = 1;
            return GT_OK;
        }
    }
}

int gt_inserter_stats(const gt_inserter* ins, uint64_t* n_sequences, uint64_t* time_total, uint64_t* counter) {
    if (!ins) return set_error(GT_ERR_INVALID_ARG, "null inserter");
    if (n_sequences) *n_sequences = ins->n_sequences;
    if (time_total) *time_total = ins->total;
    if (counter) *counter = ins->counter;
    return GT_OK;
}

int gt_inserter_free(gt_inserter* ins) {
    delete ins;
    return GT_OK;
}

}  // extern "C"
