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
//
// Plain (not gzip) files are read in large blocks and each block is parsed by
// several threads when it is well formed -- FASTA whose lines do not start
// with '@' or '+', or FASTQ of 4-line records with equal sequence and quality
// lengths, without '\r' -- because there record boundaries can be found from
// any point (a FASTA header is a line starting with '>'; a FASTQ header is a
// line starting with '@' two lines before a line starting with '+', which a
// quality line starting with '@' never is, since the next sequence line cannot
// start with '+').  On those inputs the kseq reading above gives exactly the
// same records.  A block that is not well formed, a gzip file or a strict
// parser is read record by record from that block on, as is everything when
// only one thread is available.  The next block is parsed in the background
// while the current one is handed out.
#include "goetia_amd.h"

#include <hip/hip_runtime.h>
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <future>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

namespace gt {
int set_error(int code, const std::string& msg);
int storage_set_fence(gt_storage* st, hipStream_t s);
int storage_hold(gt_storage* st, int (*fn)(void*), void* ctx);
std::recursive_mutex& storage_mutex(gt_storage* st);
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

namespace {

// Valid records of one parsed block, in file order.
struct Block {
    std::string seq, names, qual;
    std::vector<int64_t> off{0}, noff{0}, qoff{0};     // qualities: empty for FASTA records
    std::vector<uint64_t> cum_parsed, cum_skipped;   // counts up to and including each record
    uint64_t tot_parsed = 0, tot_skipped = 0;
    size_t next = 0;

    size_t count() const { return off.size() - 1; }
    void clear() {
        seq.clear();
        names.clear();
        qual.clear();
        off.assign(1, 0);
        noff.assign(1, 0);
        qoff.assign(1, 0);
        cum_parsed.clear();
        cum_skipped.clear();
        tot_parsed = tot_skipped = 0;
        next = 0;
    }
};

// One block of a mapped file parsed by several threads: the valid records of
// each thread's part in order, or ok = false when the block is not well formed.
struct ParsedBlock {
    std::vector<Block> parts;
    uint64_t start = 0, end = 0;           // file offsets
    uint64_t tot_parsed = 0, tot_skipped = 0;
    bool ok = false;
};

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
    // block-parallel reading of plain files (the file mapped read-only)
    bool par = false;
    const char* map = nullptr;
    size_t map_len = 0;
    char fmt = 0;                          // '>' or '@'
    int threads = 1;
    size_t block = 64u << 20;
    uint64_t next_pos = 0;                 // file offset where the next block starts
    ParsedBlock cur;                       // block being handed out
    size_t cur_part = 0, cur_rec = 0;
    uint64_t part_parsed = 0, part_skipped = 0;   // counts of cur's parts before cur_part
    std::future<ParsedBlock> ahead;        // the block after cur, parsed in the background
    uint64_t base_parsed = 0, base_skipped = 0;   // counts of the blocks before cur
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

// ------------------------------------------------------------ block-parallel reading

inline size_t line_end(const char* b, size_t i, size_t n) {
    const void* q = i < n ? std::memchr(b + i, '\n', n - i) : nullptr;
    return q ? (size_t)((const char*)q - b) : n;
}

inline bool has_cr(const char* b, size_t i, size_t e) { return e > i && std::memchr(b + i, '\r', e - i); }

// A well-formed FASTQ record starts at line start i: '@' line, a non-empty
// sequence line not starting with '>', '+' or '@', a '+' line and a quality
// line of the same length; every line ends in '\n' unless the input ends.
bool fq_header_at(const char* b, size_t i, size_t n, bool eof) {
    if (i >= n || b[i] != '@') return false;
    const size_t e0 = line_end(b, i, n);
    if (e0 >= n) return false;
    const size_t s1 = e0 + 1, e1 = line_end(b, s1, n);
    if (e1 >= n || e1 == s1 || b[s1] == '>' || b[s1] == '+' || b[s1] == '@') return false;
    const size_t s2 = e1 + 1;
    if (s2 >= n || b[s2] != '+') return false;
    const size_t e2 = line_end(b, s2, n);
    if (e2 >= n) return false;
    const size_t s3 = e2 + 1, e3 = line_end(b, s3, n);
    if (e3 >= n && !eof) return false;
    return e3 - s3 == e1 - s1;
}

// First record start at or after `from` (n if none): a line starting with
// '>' (FASTA) or a well-formed FASTQ record.
size_t next_header(const char* b, size_t from, size_t n, char fmt, bool eof) {
    size_t i = from;
    if (i > 0 && b[i - 1] != '\n') i = line_end(b, i, n) + 1;
    while (i < n) {
        if (fmt == '>' ? b[i] == '>' : fq_header_at(b, i, n, eof)) return i;
        i = line_end(b, i, n) + 1;
    }
    return n;
}

// Start of the last record whose header line is complete and verifiable
// (0 if none beyond the first): the block is parsed up to there and the rest
// kept for the next block.
size_t last_header(const char* b, size_t n, char fmt) {
    size_t i = n;
    while (i > 1) {
        const void* q = memrchr(b, '\n', i - 1);
        if (!q) return 0;
        const size_t ls = (size_t)((const char*)q - b) + 1;
        if (fmt == '>' ? b[ls] == '>' : fq_header_at(b, ls, n, false)) return ls;
        i = ls - 1;
    }
    return 0;
}

// Validated (uppercased) copy of [i, e) appended to out; false if a byte is
// outside A/C/G/T/a/c/g/t.
inline bool append_valid(std::string& out, const char* b, size_t i, size_t e) {
    const size_t o = out.size();
    out.resize(o + (e - i));
    char* d = &out[o];
    unsigned char ok = 0xFF;
    for (size_t k = i; k < e; ++k) {
        const unsigned char v = kValidator.map[(unsigned char)b[k]];
        ok &= (unsigned char)(v ? 0xFF : 0);
        *d++ = (char)v;
    }
    return ok != 0;
}

// FastxParser::next's bookkeeping for one record: keep it, or count it as
// skipped (twice when it is both invalid and short, as goetia counts).
inline void finish_record(Block& out, size_t seq_start, size_t qual_start, bool valid, uint32_t min_length,
                          const char* name, size_t name_len) {
    const size_t len = out.seq.size() - seq_start;
    out.tot_parsed++;
    if (!valid) out.tot_skipped++;
    if (len < min_length) out.tot_skipped++;
    if (!valid || len < min_length) {
        out.seq.resize(seq_start);
        out.qual.resize(qual_start);
        return;
    }
    out.off.push_back((int64_t)out.seq.size());
    out.qoff.push_back((int64_t)out.qual.size());
    out.names.append(name, name_len);
    out.noff.push_back((int64_t)out.names.size());
    out.cum_parsed.push_back(out.tot_parsed);
    out.cum_skipped.push_back(out.tot_skipped);
}

inline size_t name_end(const char* b, size_t i, size_t e) {
    while (i < e && !std::isspace((unsigned char)b[i])) ++i;
    return i;
}

// Records whose headers lie in [a, end) of a well-formed block; false when the
// range is not well formed.
bool parse_range(const char* b, size_t a, size_t end, size_t n, bool eof, char fmt, uint32_t min_length,
                 Block& out) {
    size_t pos = a;
    while (pos < end) {
        if (b[pos] != fmt) return false;
        const size_t he = line_end(b, pos, n);
        if (has_cr(b, pos, he)) return false;
        const char* name = b + pos + 1;
        const size_t name_len = name_end(b, pos + 1, he) - (pos + 1);
        const size_t start = out.seq.size(), qstart = out.qual.size();
        bool valid = true;
        if (fmt == '>') {
            pos = he + 1;
            while (pos < n && b[pos] != '>') {
                const size_t le = line_end(b, pos, n);
                if (le > pos) {
                    if (b[pos] == '@' || b[pos] == '+' || has_cr(b, pos, le)) return false;
                    valid &= append_valid(out.seq, b, pos, le);
                }
                pos = le + 1;
            }
        } else {
            if (!fq_header_at(b, pos, n, eof)) return false;
            const size_t s1 = he + 1, e1 = line_end(b, s1, n);
            const size_t e2 = line_end(b, e1 + 1, n);
            const size_t e3 = line_end(b, e2 + 1, n);
            if (has_cr(b, s1, e3)) return false;
            valid = append_valid(out.seq, b, s1, e1);
            out.qual.append(b + e2 + 1, e3 - (e2 + 1));
            pos = e3 + 1;
        }
        finish_record(out, start, qstart, valid, min_length, name, name_len);
    }
    return pos == end || (eof && end == n && pos >= n);
}

int block_threads() {
    if (const char* v = std::getenv("GOETIA_AMD_FASTX_THREADS")) {
        const int t = std::atoi(v);
        if (t > 0) return std::min(t, 64);
    }
    const unsigned hc = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(16u, hc ? hc : 1u));
}

size_t block_bytes() {
    if (const char* v = std::getenv("GOETIA_AMD_FASTX_BLOCK")) {
        const unsigned long long x = std::strtoull(v, nullptr, 10);
        if (x >= 64) return (size_t)x;
    }
    return 64u << 20;
}

// A run of consecutive kept records of one parsed part going into a batch:
// their bases in one piece, and their end offsets shifted by `delta`.
struct RunCopy {
    uint8_t* dst;
    const char* src;
    int64_t* dst_off;
    const int64_t* src_off;
    size_t count;
    int64_t delta;
};

void copy_run(const RunCopy& r) {
    const int64_t nb = r.src_off[r.count - 1] - (r.src_off[-1]);
    std::memcpy(r.dst, r.src, (size_t)nb);
    for (size_t k = 0; k < r.count; ++k) r.dst_off[k] = r.src_off[k] + r.delta;
}

// Copies the runs (on several threads when there is much to copy) and empties
// the list.
void copy_runs(std::vector<RunCopy>& runs) {
    if (runs.empty()) return;
    size_t total = 0;
    for (const RunCopy& r : runs) total += (size_t)(r.src_off[r.count - 1] - r.src_off[-1]);
    const int T = (int)std::min<size_t>({runs.size(), (size_t)std::max(1, block_threads() / 2), total >> 20});
    if (T <= 1) {
        for (const RunCopy& r : runs) copy_run(r);
    } else {
        std::vector<std::thread> pool;
        auto work = [&](int t) {
            for (size_t i = (size_t)t; i < runs.size(); i += (size_t)T) copy_run(runs[i]);
        };
        for (int t = 1; t < T; ++t) pool.emplace_back(work, t);
        work(0);
        for (auto& th : pool) th.join();
    }
    runs.clear();
}

// Parses the block of the mapping starting at `pos` (a record start): up to
// the last verifiable record start within `block` bytes, growing the block
// while it holds less than one record.
ParsedBlock parse_block_at(const char* map, size_t map_len, uint64_t pos, size_t block, int T, char fmt,
                           uint32_t min_length) {
    ParsedBlock out;
    out.start = pos;
    const char* b = map + pos;
    const size_t avail = map_len - pos;
    if (avail == 0 || b[0] != fmt) return out;
    size_t n = 0, cut = 0;
    for (;;) {
        n = std::min(block, avail);
        cut = n == avail ? n : last_header(b, n, fmt);
        if (cut > 0) break;
        block *= 2;
    }
    const bool eof = n == avail && cut == n;
    std::vector<size_t> st(T + 1);
    st[0] = 0;
    st[T] = cut;
    for (int t = 1; t < T; ++t) st[t] = std::max(st[t - 1], next_header(b, cut * (size_t)t / (size_t)T, cut, fmt, eof));
    out.parts.resize(T);
    std::vector<char> ok(T, 1);
    auto work = [&](int t) { ok[t] = parse_range(b, st[t], st[t + 1], n, eof, fmt, min_length, out.parts[t]) ? 1 : 0; };
    if (T == 1) {
        work(0);
    } else {
        std::vector<std::thread> pool;
        for (int t = 1; t < T; ++t) pool.emplace_back(work, t);
        work(0);
        for (auto& th : pool) th.join();
    }
    for (int t = 0; t < T; ++t) {
        if (!ok[t]) return out;
        out.tot_parsed += out.parts[t].tot_parsed;
        out.tot_skipped += out.parts[t].tot_skipped;
    }
    out.end = pos + cut;
    out.ok = true;
    return out;
}

void launch_ahead(gt_fastx& p) {
    p.ahead = std::async(std::launch::async, parse_block_at, p.map, p.map_len, p.next_pos, p.block, p.threads, p.fmt,
                         p.min_length);
}

// Switches a block-parallel parser to record-by-record reading from file
// offset pos (a record start, so kseq resumes exactly there).
int fall_back(gt_fastx& p, uint64_t pos) {
    p.par = false;
    p.cur = ParsedBlock();
    p.n_parsed = p.base_parsed;
    p.n_skipped = p.base_skipped;
    if (gzseek(p.ks.f, (z_off_t)pos, SEEK_SET) < 0) return -3;
    p.ks.begin = p.ks.end = 0;
    p.ks.is_eof = false;
    p.last_char = 0;
    return -1;
}

// Makes the next block current: 1 ready, 0 end of input, -1 switched to
// record-by-record reading, -3 read error.
int next_block(gt_fastx& p) {
    p.base_parsed += p.cur.tot_parsed;
    p.base_skipped += p.cur.tot_skipped;
    p.n_parsed = p.base_parsed;
    p.n_skipped = p.base_skipped;
    p.cur = ParsedBlock();
    p.cur_part = p.cur_rec = 0;
    p.part_parsed = p.part_skipped = 0;
    if (!p.ahead.valid()) {
        if (p.next_pos >= p.map_len) return 0;
        launch_ahead(p);
    }
    ParsedBlock blk = p.ahead.get();
    if (!blk.ok) return fall_back(p, blk.start);
    p.cur = std::move(blk);
    p.next_pos = p.cur.end;
    if (p.next_pos < p.map_len) launch_ahead(p);
    return 1;
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
    // block-parallel reading for plain FASTA / FASTQ files (not gzip, not strict)
    if (!p->strict && block_threads() > 1) {
        const int fd = ::open(path, O_RDONLY);
        struct stat sb;
        if (fd >= 0 && ::fstat(fd, &sb) == 0 && S_ISREG(sb.st_mode) && sb.st_size >= 2) {
            void* m = ::mmap(nullptr, (size_t)sb.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
            if (m != MAP_FAILED) {
                const unsigned char* h = (const unsigned char*)m;
                const bool gz = h[0] == 0x1f && h[1] == 0x8b;
                if (!gz && (h[0] == '>' || h[0] == '@')) {
                    ::madvise(m, (size_t)sb.st_size, MADV_SEQUENTIAL);
                    p->map = (const char*)m;
                    p->map_len = (size_t)sb.st_size;
                    p->par = true;
                    p->fmt = (char)h[0];
                    p->threads = block_threads();
                    p->block = block_bytes();
                } else {
                    ::munmap(m, (size_t)sb.st_size);
                }
            }
        }
        if (fd >= 0) ::close(fd);
    }
    *out = p;
    return GT_OK;
}

int gt_fastx_next_batch(gt_fastx* p, uint64_t max_reads, uint64_t max_bases, uint8_t* seq, int64_t* offsets,
                        char* names, uint64_t names_cap, int64_t* name_offsets, uint64_t* n_reads) {
    return gt_fastx_next_batch_qual(p, max_reads, max_bases, seq, offsets, nullptr, nullptr, names, names_cap,
                                    name_offsets, n_reads);
}

int gt_fastx_next_batch_qual(gt_fastx* p, uint64_t max_reads, uint64_t max_bases, uint8_t* seq, int64_t* offsets,
                             uint8_t* qual, uint8_t* has_qual, char* names, uint64_t names_cap,
                             int64_t* name_offsets, uint64_t* n_reads) {
    if (!p || !n_reads || !offsets || (!seq && max_bases)) return set_error(GT_ERR_INVALID_ARG, "null argument");
    if ((qual == nullptr) != (has_qual == nullptr)) return set_error(GT_ERR_INVALID_ARG, "qual needs has_qual");
    if (names && !name_offsets) return set_error(GT_ERR_INVALID_ARG, "names need name_offsets");
    uint64_t n = 0, bases = 0, nbytes = 0;
    offsets[0] = 0;
    if (names) name_offsets[0] = 0;
    *n_reads = 0;
    // Runs of records taken from the current block, copied together before
    // the block is released (next_block) and before returning.
    std::vector<RunCopy> runs;
    for (;;) {
        if (p->par) {
            if (p->cur_part >= p->cur.parts.size()) {
                copy_runs(runs);
                const int r = next_block(*p);
                if (r == 0) {
                    p->complete = true;
                    break;
                }
                if (r == -3) {
                    *n_reads = n;
                    return set_error(GT_ERR_IO, "error reading " + p->path);
                }
                continue;
            }
            const Block& B = p->cur.parts[p->cur_part];
            if (p->cur_rec >= B.count()) {
                p->part_parsed += B.tot_parsed;
                p->part_skipped += B.tot_skipped;
                p->n_parsed = p->base_parsed + p->part_parsed;
                p->n_skipped = p->base_skipped + p->part_skipped;
                ++p->cur_part;
                p->cur_rec = 0;
                continue;
            }
            if (!qual && !names) {
                // every record of the part from cur_rec that fits, as one run
                const size_t i = p->cur_rec;
                const size_t jmax = std::min<uint64_t>(B.count(), i + (max_reads - n));
                const int64_t lim = B.off[i] + (int64_t)(max_bases - bases);
                const size_t j =
                    (size_t)(std::upper_bound(B.off.begin() + i, B.off.begin() + jmax + 1, lim) - B.off.begin()) - 1;
                if (j > i) {
                    runs.push_back(RunCopy{seq + bases, B.seq.data() + B.off[i], offsets + n + 1, B.off.data() + i + 1,
                                           j - i, (int64_t)bases - B.off[i]});
                    bases += (uint64_t)(B.off[j] - B.off[i]);
                    n += j - i;
                    p->cur_rec = j;
                    p->n_parsed = p->base_parsed + p->part_parsed + B.cum_parsed[j - 1];
                    p->n_skipped = p->base_skipped + p->part_skipped + B.cum_skipped[j - 1];
                    continue;
                }
                // the record at cur_rec does not fit: the record-by-record step below stops
            }
            const size_t i = p->cur_rec;
            const uint64_t len = (uint64_t)(B.off[i + 1] - B.off[i]);
            const uint64_t nlen = (uint64_t)(B.noff[i + 1] - B.noff[i]);
            const bool fits = n < max_reads && bases + len <= max_bases && (!names || nbytes + nlen <= names_cap);
            p->n_parsed = p->base_parsed + p->part_parsed + B.cum_parsed[i];
            p->n_skipped = p->base_skipped + p->part_skipped + B.cum_skipped[i];
            if (!fits) {
                if (n == 0) {
                    return set_error(GT_ERR_INVALID_ARG, "record of " + std::to_string(len) +
                                                             " bases does not fit the batch buffers");
                }
                break;
            }
            std::memcpy(seq + bases, B.seq.data() + B.off[i], len);
            if (qual) {
                const bool q = B.qoff[i + 1] > B.qoff[i];
                if (q) std::memcpy(qual + bases, B.qual.data() + B.qoff[i], len);
                has_qual[n] = q ? 1 : 0;
            }
            bases += len;
            offsets[n + 1] = (int64_t)bases;
            if (names) {
                std::memcpy(names + nbytes, B.names.data() + B.noff[i], nlen);
                nbytes += nlen;
                name_offsets[n + 1] = (int64_t)nbytes;
            }
            ++n;
            ++p->cur_rec;
            continue;
        }
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
        if (qual) {
            const bool q = !p->qual.empty();
            if (q) std::memcpy(qual + bases, p->qual.data(), len);
            has_qual[n] = q ? 1 : 0;
        }
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
    copy_runs(runs);
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
    if (p->ahead.valid()) p->ahead.wait();
    if (p->map) ::munmap((void*)p->map, p->map_len);
    if (p->ks.f) gzclose(p->ks.f);
    delete p;
    return GT_OK;
}

}  // extern "C"

// ------------------------------------------------------------ chunk driver

namespace {

// One batch slot: reads in pinned host memory and their device copy.  Reads
// [0, cur) are queued on the device, [cur, end) are taken (counted by the
// interval) but held, [end, n) are parsed and not yet taken.
struct Slot {
    uint8_t* h_seq = nullptr;
    int64_t* h_off = nullptr;              // offsets of the slot's reads (h_off[0] = 0)
    int64_t* h_sub = nullptr;              // offsets of a range that starts past read 0, rebased
    uint8_t* d_seq = nullptr;
    int64_t* d_off = nullptr;
    int64_t* d_sub = nullptr;
    hipEvent_t copied = nullptr;           // the host buffers' last copy (copy stream)
    hipEvent_t inserted = nullptr;         // the device buffers' last insert (insert stream)
    uint64_t n = 0, cur = 0, end = 0;
};

}  // namespace

struct gt_inserter {
    gt_storage* st = nullptr;
    uint16_t K = 0;
    int canonical = 1, mode = GT_MODE_RAW;
    uint64_t interval = 500000;
    uint64_t counter = 0, total = 0, n_sequences = 0;
    uint64_t max_reads = 0, max_bases = 0;
    Slot slot[2];
    int active = 0;                        // the slot reads are taken from (or paired reads appended to)
    hipStream_t stream = nullptr;          // inserts
    hipStream_t copy = nullptr;            // host-to-device copies
    bool pending_work = false;             // inserts queued since the last synchronisation
    std::atomic<bool> holding{false};      // registered with the storage as holder of taken reads (set under its lock)

    ~gt_inserter() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (copy) (void)hipStreamSynchronize(copy);
        for (Slot& sl : slot) {
            if (sl.h_seq) (void)hipHostFree(sl.h_seq);
            if (sl.h_off) (void)hipHostFree(sl.h_off);
            if (sl.h_sub) (void)hipHostFree(sl.h_sub);
            if (sl.d_seq) (void)hipFree(sl.d_seq);
            if (sl.d_off) (void)hipFree(sl.d_off);
            if (sl.d_sub) (void)hipFree(sl.d_sub);
            if (sl.copied) (void)hipEventDestroy(sl.copied);
            if (sl.inserted) (void)hipEventDestroy(sl.inserted);
        }
        if (stream) (void)hipStreamDestroy(stream);
        if (copy) (void)hipStreamDestroy(copy);
    }
};

namespace {

#define GT_TRY_FX(expr)                 \
    do {                                \
        const int rc_ = (expr);         \
        if (rc_ != GT_OK) return rc_;   \
    } while (0)

int hip_fail(hipError_t e, const char* what) {
    return set_error(GT_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define GT_FX_HIP(call)                                       \
    do {                                                      \
        hipError_t e_ = (call);                               \
        if (e_ != hipSuccess) return hip_fail(e_, #call);     \
    } while (0)

using StorageGuard = std::lock_guard<std::recursive_mutex>;

int sync_inserts(gt_inserter* ins) {
    if (ins->pending_work) {
        GT_FX_HIP(hipStreamSynchronize(ins->stream));
        ins->pending_work = false;
    }
    return GT_OK;
}

// Queues the slot's taken reads [cur, end) as one insert: their bases and
// offsets go to the device on the copy stream (after the slot's previous
// insert has stopped reading the device buffers), the insert stream waits
// for the copy, and the storage's fence follows the insert.  The caller
// holds the storage's lock.
int queue_taken(gt_inserter* ins, Slot& sl) {
    if (ins->holding) {
        ins->holding = false;
        GT_TRY_FX(gt::storage_hold(ins->st, nullptr, ins));
    }
    if (sl.cur >= sl.end) return GT_OK;
    const uint64_t nr = sl.end - sl.cur;
    const int64_t b0 = sl.h_off[sl.cur];
    const uint64_t nb = (uint64_t)(sl.h_off[sl.end] - b0);
    GT_FX_HIP(hipStreamWaitEvent(ins->copy, sl.inserted, 0));
    const int64_t* d_offsets = sl.d_off;
    if (sl.cur == 0) {
        GT_FX_HIP(hipMemcpyAsync(sl.d_off, sl.h_off, (nr + 1) * sizeof(int64_t), hipMemcpyHostToDevice, ins->copy));
    } else {
        // a range after a flush: rebase its offsets (h_sub's previous copy must be done)
        GT_FX_HIP(hipEventSynchronize(sl.copied));
        for (uint64_t i = 0; i <= nr; ++i) sl.h_sub[i] = sl.h_off[sl.cur + i] - b0;
        GT_FX_HIP(hipMemcpyAsync(sl.d_sub, sl.h_sub, (nr + 1) * sizeof(int64_t), hipMemcpyHostToDevice, ins->copy));
        d_offsets = sl.d_sub;
    }
    if (nb) GT_FX_HIP(hipMemcpyAsync(sl.d_seq + b0, sl.h_seq + b0, nb, hipMemcpyHostToDevice, ins->copy));
    GT_FX_HIP(hipEventRecord(sl.copied, ins->copy));
    GT_FX_HIP(hipStreamWaitEvent(ins->stream, sl.copied, 0));
    sl.cur = sl.end;
    const int rc = gt_dbg_insert_dev(ins->st, ins->K, ins->canonical, ins->mode, sl.d_seq + b0, d_offsets, nr, nb,
                                     nullptr, nullptr, ins->stream);
    if (rc != GT_OK) return rc;
    GT_FX_HIP(hipEventRecord(sl.inserted, ins->stream));
    GT_TRY_FX(gt::storage_set_fence(ins->st, ins->stream));
    ins->pending_work = true;
    return GT_OK;
}

// The storage's flush of an inserter's held reads (they are in the active slot).
int flush_held(void* ctx) {
    gt_inserter* ins = static_cast<gt_inserter*>(ctx);
    ins->holding = false;              // the storage has already dropped the hold
    return queue_taken(ins, ins->slot[ins->active]);
}

// Holds the active slot's taken reads until the next batch or the storage's
// next call (caller holds the storage's lock).
int hold_taken(gt_inserter* ins) {
    const Slot& sl = ins->slot[ins->active];
    if (sl.cur >= sl.end || ins->holding) return GT_OK;
    GT_TRY_FX(gt::storage_hold(ins->st, flush_held, ins));
    ins->holding = true;
    return GT_OK;
}

// Queues the active slot's taken reads and makes the other slot active and
// empty (caller holds the storage's lock); the caller then waits for the new
// slot's host buffers (hipEventSynchronize of its `copied`) before filling it.
int rotate(gt_inserter* ins) {
    GT_TRY_FX(queue_taken(ins, ins->slot[ins->active]));
    ins->active ^= 1;
    Slot& nx = ins->slot[ins->active];
    nx.n = nx.cur = nx.end = 0;
    nx.h_off[0] = 0;
    return GT_OK;
}

}  // namespace

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
    ins->max_bases = batch_bases ? batch_bases : (128ULL << 20);
    ins->max_reads = std::max<uint64_t>(ins->max_bases / 64, 1024);
    hipError_t e = hipStreamCreate(&ins->stream);
    if (e == hipSuccess) e = hipStreamCreate(&ins->copy);
    for (Slot& sl : ins->slot) {
        const size_t ob = (ins->max_reads + 1) * sizeof(int64_t);
        if (e == hipSuccess) e = hipHostMalloc((void**)&sl.h_seq, ins->max_bases);
        if (e == hipSuccess) e = hipHostMalloc((void**)&sl.h_off, ob);
        if (e == hipSuccess) e = hipHostMalloc((void**)&sl.h_sub, ob);
        if (e == hipSuccess) e = hipMalloc((void**)&sl.d_seq, ins->max_bases);
        if (e == hipSuccess) e = hipMalloc((void**)&sl.d_off, ob);
        if (e == hipSuccess) e = hipMalloc((void**)&sl.d_sub, ob);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.copied, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.inserted, hipEventDisableTiming);
        // recorded once, so the first waits on them return at once
        if (e == hipSuccess) e = hipEventRecord(sl.copied, ins->copy);
        if (e == hipSuccess) e = hipEventRecord(sl.inserted, ins->stream);
        if (e == hipSuccess) sl.h_off[0] = 0;
    }
    if (e != hipSuccess) {
        delete ins;
        return set_error(GT_ERR_OUT_OF_MEMORY, std::string("gt_inserter_new: ") + hipGetErrorString(e));
    }
    *out = ins;
    return GT_OK;
}

// InserterProcessor / FileProcessor::advance (processors.hh:206-231, 304-306)
// with IntervalCounter::poll (metrics.hh:126-135): reads are taken until the
// k-mers of the reads taken since the last interval reach `interval` (each
// read counts len - K + 1, 0 when shorter than K); the call returns right
// after the read that reaches it, or, with every insert finished, at the end
// of the input.  Taken reads are inserted in batches of up to the slot size:
// those taken when the call returns are held, and go in with the next batch
// or before the storage's next call, whichever comes first, so every call on
// the storage sees every read taken so far.  While the GPU inserts one slot
// the next batch is parsed into the other.
int gt_inserter_advance(gt_inserter* ins, gt_fastx* parser, uint64_t* n_sequences, uint64_t* time_total,
                        int* remaining) {
    if (!ins || !parser) return set_error(GT_ERR_INVALID_ARG, "null inserter or parser");
    std::recursive_mutex& mu = gt::storage_mutex(ins->st);
    for (;;) {
        Slot& sl = ins->slot[ins->active];
        if (sl.end >= sl.n) {
            // every parsed read is taken: queue them, parse the next batch into the other slot
            {
                StorageGuard g(mu);
                GT_TRY_FX(rotate(ins));
            }
            Slot& nx = ins->slot[ins->active];
            GT_FX_HIP(hipEventSynchronize(nx.copied));
            uint64_t got = 0;
            const int rc = gt_fastx_next_batch(parser, ins->max_reads, ins->max_bases, nx.h_seq, nx.h_off, nullptr, 0,
                                               nullptr, &got);
            // FileProcessor::handle_next (processors.hh:149-171) skips a record
            // with an invalid character even from a strict parser
            if (rc != GT_OK && rc != GT_ERR_INVALID_CHARACTER) return rc;
            {
                StorageGuard g(mu);
                nx.n = got;
            }
            if (got == 0 && rc == GT_ERR_INVALID_CHARACTER) continue;
            if (got == 0) {
                GT_TRY_FX(sync_inserts(ins));
                if (n_sequences) *n_sequences = ins->n_sequences;
                if (time_total) *time_total = ins->total;
                if (remaining) *remaining = 0;
                return GT_OK;
            }
            continue;
        }
        uint64_t end = sl.end;
        bool crossed = false;
        while (end < sl.n) {
            const int64_t len = sl.h_off[end + 1] - sl.h_off[end];
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
        {
            StorageGuard g(mu);
            sl.end = end;
            if (crossed) GT_TRY_FX(hold_taken(ins));
        }
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

// Held reads go in before the inserter goes (the storage must outlive its
// inserters; a storage freed first has already taken them).
int gt_inserter_free(gt_inserter* ins) {
    if (!ins) return GT_OK;
    int rc = GT_OK;
    if (ins->holding) {
        // checked again under the storage's lock: another thread's call on
        // the storage may have flushed the held reads meanwhile (a storage
        // freed first has flushed them and cleared the flag)
        StorageGuard g(gt::storage_mutex(ins->st));
        if (ins->holding) rc = queue_taken(ins, ins->slot[ins->active]);
    }
    delete ins;
    return rc;
}

}  // extern "C"

// ------------------------------------------------------------ paired files
//
// SplitPairedReader<FastxParser<DNA_SIMPLE>> (readers.hh:235-334) and
// FileProcessor::advance over it (processors.hh:180-204).  Every step reads
// one record attempt from each file.  A side whose attempt is rejected
// (invalid characters, shorter than min_length) is empty; an attempt that
// raises (an invalid character from a strict parser, a quality of another
// length) makes that side empty too, unless the reader is strict, in which
// case the whole pair is dropped (handle_next, processors.hh:149-171).  With
// force_name_match, two present records whose names are not a pair
// (check_is_pair, parsing.cc:56-87) drop the pair and count two skipped, or,
// strict, stop with "Unpaired reads".  Files that end at different steps
// stop with "Mismatched split paired files." at the step after the shorter
// one ended.  A pair's reads are inserted left then right, and the interval
// is polled once per pair with both sides' k-mers.

struct gt_pairs {
    gt_fastx* side[2] = {nullptr, nullptr};
    bool strict = false;
    bool force_name_match = false;
    uint64_t n_unpaired = 0;               // SplitPairedReader::_n_skipped
    // the record of each side of the current pair (present = false when empty)
    std::string seq[2], name[2], qual[2];
    bool present[2] = {false, false};
    bool held = false;                     // the current pair was read but not yet taken
};

namespace {

// split_on_first (parsing.cc:31-52): the text before the first delimiter
// run, and the text after it.
std::pair<std::string, std::string> split_first(const std::string& s, const char* delims) {
    auto is_delim = [&](char c) { return std::strchr(delims, c) != nullptr && c != '\0'; };
    std::string left, right;
    bool have_left = false;
    for (size_t i = 0; i < s.size(); ++i) {
        if (!have_left || left.empty()) {
            if (is_delim(s[i])) {
                left = s.substr(0, i);
                have_left = !left.empty();
            }
        } else if (!is_delim(s[i])) {
            right = s.substr(i);
            break;
        }
    }
    if (left.empty()) left = s;
    return {left, right};
}

bool ends_with(const std::string& s, const char* suffix) {
    const size_t n = std::strlen(suffix);
    return s.size() >= n && s.compare(s.size() - n, n, suffix) == 0;
}

// check_is_pair (parsing.cc:56-87): "<id>/1" and "<id>/2"; or equal first
// words whose rest ends in "1:" / "2:"; or equal first words whose rest is
// "<x>/1" / "<x>/2" with the same non-empty <x>.
bool names_pair(const std::string& l, const std::string& r) {
    const auto ls = split_first(l, " \t"), rs = split_first(r, " \t");
    if (ends_with(ls.first, "/1") && ends_with(rs.first, "/2")) {
        const auto a = split_first(ls.first, "/"), b = split_first(rs.first, "/");
        return !a.first.empty() && a.first == b.first;
    }
    if (ls.first == rs.first && ends_with(ls.second, "1:") && ends_with(rs.second, "2:")) return true;
    if (ls.first == rs.first && ends_with(ls.second, "/1") && ends_with(rs.second, "/2")) {
        const auto a = split_first(ls.second, "/"), b = split_first(rs.second, "/");
        return !a.first.empty() && a.first == b.first;
    }
    return false;
}

// The record-at-a-time reader for a parser opened by gt_pairs_open.
void drop_blocks(gt_fastx* p) {
    if (!p->par) return;
    if (p->map) ::munmap((void*)p->map, p->map_len);
    p->map = nullptr;
    p->map_len = 0;
    p->par = false;
}

enum PairStep { kPairReady, kPairDropped, kPairsDone };

// One SplitPairedReader::next (with the advance loop's is_complete check and
// handle_next) into pr->seq / name / qual / present.
int next_pair(gt_pairs* pr, PairStep* step) {
    if (pr->held) {
        pr->held = false;
        *step = kPairReady;
        return GT_OK;
    }
    gt_fastx* L = pr->side[0];
    gt_fastx* R = pr->side[1];
    if (L->complete != R->complete) return set_error(GT_ERR_PAIRING, "Mismatched split paired files.");
    if (L->complete) {
        *step = kPairsDone;
        return GT_OK;
    }
    bool threw = false;
    for (int k = 0; k < 2; ++k) {
        gt_fastx* p = pr->side[k];
        pr->present[k] = false;
        const Next r = next_record(*p);
        if (r == kIoError) return set_error(GT_ERR_IO, "error reading " + p->path);
        if (r == kInvalidChar || r == kInvalidRead) threw = true;
        if (r == kRecord) {
            pr->present[k] = true;
            pr->seq[k].swap(p->seq);
            pr->name[k].swap(p->name);
            pr->qual[k].swap(p->qual);
        }
    }
    if (threw && pr->strict) {
        *step = kPairDropped;
        return GT_OK;
    }
    if (pr->force_name_match && pr->present[0] && pr->present[1] && !names_pair(pr->name[0], pr->name[1])) {
        if (pr->strict) return set_error(GT_ERR_PAIRING, "Unpaired reads");
        pr->n_unpaired += 2;
        pr->present[0] = pr->present[1] = false;
    }
    *step = kPairReady;
    return GT_OK;
}

}  // namespace

extern "C" {

int gt_pairs_open(const char* left, const char* right, int strict, uint32_t min_length, int force_name_match,
                  gt_pairs** out) {
    if (!left || !right || !out) return set_error(GT_ERR_INVALID_ARG, "null path or output");
    *out = nullptr;
    gt_pairs* pr = new (std::nothrow) gt_pairs();
    if (!pr) return set_error(GT_ERR_OUT_OF_MEMORY, "gt_pairs_open");
    int rc = gt_fastx_open(left, strict, min_length, &pr->side[0]);
    if (rc == GT_OK) rc = gt_fastx_open(right, strict, min_length, &pr->side[1]);
    if (rc != GT_OK) {
        gt_pairs_close(pr);
        return rc;
    }
    drop_blocks(pr->side[0]);
    drop_blocks(pr->side[1]);
    pr->strict = strict != 0;
    pr->force_name_match = force_name_match != 0;
    *out = pr;
    return GT_OK;
}

int gt_pairs_next(gt_pairs* pr, uint8_t* seq, uint64_t seq_cap, int64_t* lens, char* names, uint64_t names_cap,
                  int64_t* name_lens, int* done) {
    if (!pr || !lens || !done || (!seq && seq_cap)) return set_error(GT_ERR_INVALID_ARG, "null argument");
    for (;;) {
        PairStep step = kPairReady;
        const int rc = next_pair(pr, &step);
        if (rc != GT_OK) return rc;
        *done = step == kPairsDone;
        if (step == kPairsDone) return GT_OK;
        if (step == kPairDropped) return set_error(GT_ERR_INVALID_CHARACTER, "a record of the pair was rejected "
                                                                             "by a strict reader");
        uint64_t at = 0, nat = 0;
        for (int k = 0; k < 2; ++k) {
            lens[k] = -1;
            if (name_lens) name_lens[k] = -1;
            if (!pr->present[k]) continue;
            const uint64_t n = pr->seq[k].size();
            if (at + n > seq_cap) return set_error(GT_ERR_INVALID_ARG, "pair does not fit the buffer");
            std::memcpy(seq + at, pr->seq[k].data(), n);
            at += n;
            lens[k] = (int64_t)n;
            if (names && name_lens) {
                const uint64_t m = pr->name[k].size();
                if (nat + m > names_cap) return set_error(GT_ERR_INVALID_ARG, "names do not fit the buffer");
                std::memcpy(names + nat, pr->name[k].data(), m);
                nat += m;
                name_lens[k] = (int64_t)m;
            }
        }
        return GT_OK;
    }
}

int gt_pairs_stats(const gt_pairs* pr, uint64_t* n_parsed, uint64_t* n_skipped, int* complete) {
    if (!pr) return set_error(GT_ERR_INVALID_ARG, "null reader");
    if (n_parsed) *n_parsed = pr->side[0]->n_parsed + pr->side[1]->n_parsed;
    if (n_skipped) *n_skipped = pr->n_unpaired + pr->side[0]->n_skipped + pr->side[1]->n_skipped;
    if (complete) *complete = pr->side[0]->complete && pr->side[1]->complete;
    return GT_OK;
}

int gt_pairs_close(gt_pairs* pr) {
    if (!pr) return GT_OK;
    for (gt_fastx* p : pr->side)
        if (p) gt_fastx_close(p);
    delete pr;
    return GT_OK;
}

// FileProcessor::advance over a SplitPairedReader for an InserterProcessor
// (processors.hh:180-204, 304-306): pairs are appended to the active slot,
// left then right, and taken as they come; a full slot is queued and the
// other one filled.  The call returns when a pair brings the interval
// counter to `interval` (its reads held as in gt_inserter_advance), or, with
// every insert finished, at the end of the input or when the reader raises.
int gt_inserter_advance_pairs(gt_inserter* ins, gt_pairs* pr, uint64_t* n_sequences, uint64_t* time_total,
                              int* remaining) {
    if (!ins || !pr) return set_error(GT_ERR_INVALID_ARG, "null inserter or reader");
    for (const Slot& s : ins->slot)
        if (s.end < s.n) return set_error(GT_ERR_INVALID_ARG, "the inserter holds reads of a single file");
    std::recursive_mutex& mu = gt::storage_mutex(ins->st);
    for (;;) {
        Slot& sl = ins->slot[ins->active];
        uint64_t nr = sl.n;
        int64_t nb = sl.h_off[nr];
        bool crossed = false, done = false, full = false;
        int err = GT_OK;
        for (;;) {
            if (nr + 2 > ins->max_reads) {             // a pair adds at most two reads
                full = true;
                break;
            }
            PairStep step = kPairReady;
            err = next_pair(pr, &step);
            if (err != GT_OK) break;
            if (step == kPairsDone) {
                done = true;
                break;
            }
            if (step == kPairDropped) continue;
            uint64_t need = 0;
            for (int k = 0; k < 2; ++k)
                if (pr->present[k]) need += pr->seq[k].size();
            if ((uint64_t)nb + need > ins->max_bases) {
                if (nr == 0) {
                    err = set_error(GT_ERR_INVALID_ARG, "a read pair does not fit the batch buffers");
                    break;
                }
                pr->held = true;                     // first pair of the next slot
                full = true;
                break;
            }
            uint64_t t = 0;
            for (int k = 0; k < 2; ++k) {
                if (!pr->present[k]) continue;
                const uint64_t len = pr->seq[k].size();
                std::memcpy(sl.h_seq + nb, pr->seq[k].data(), len);
                nb += (int64_t)len;
                sl.h_off[++nr] = nb;
                t += len >= ins->K ? len - ins->K + 1 : 0;
                ++ins->n_sequences;
            }
            ins->total += t;
            ins->counter += t;
            if (ins->counter >= ins->interval) {
                ins->counter = 0;
                crossed = true;
                break;
            }
        }
        {
            StorageGuard g(mu);
            sl.n = sl.end = nr;
            if (full)
                GT_TRY_FX(rotate(ins));
            else if (done || err != GT_OK)
                GT_TRY_FX(queue_taken(ins, sl));
            else if (crossed)
                GT_TRY_FX(hold_taken(ins));
        }
        if (err != GT_OK) {
            GT_TRY_FX(sync_inserts(ins));
            return err;
        }
        if (crossed || done) {
            if (done) GT_TRY_FX(sync_inserts(ins));
            if (n_sequences) *n_sequences = ins->n_sequences;
            if (time_total) *time_total = ins->total;
            if (remaining) *remaining = crossed ? 1 : 0;
            return GT_OK;
        }
        // full: fill the other slot once its host buffers are free
        GT_FX_HIP(hipEventSynchronize(ins->slot[ins->active].copied));
    }
}

}  // extern "C"
