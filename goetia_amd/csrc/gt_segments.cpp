// StreamingCompactor<dBG<S, FwdLemireShifter>>::Compactor::find_new_segments
// (goetia compactor.hh:321-534) over a batch of reads.
//
// The graph view of a batch -- every window's query count and hash, and the
// eight neighbour values with their presence bits of every new window (count
// 0, the only ones the passes look at) -- comes from three GPU launches for
// the whole batch (query, neighbours, gather of the new windows; see
// gt::segment_view).  The two passes over each read (runs of new k-mers, then
// the cuts at decision k-mers with goetia's anchor and flank bookkeeping) are
// sequential per read and run here on the host, independent reads on several
// threads.  gt_segments_from_view takes a full per-window view from the
// caller (no device needed); gt_find_new_segments builds it on the GPU.

#include "goetia_amd.h"

#include <algorithm>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace gt {
int set_error(int code, const std::string& msg);
// gt_capi.cpp: counts and hashes of every window, neighbour values and
// presence of the new ones (count 0), in window order.
int segment_view(gt_storage* st, uint16_t K, const uint8_t* seq, const int64_t* offsets, uint64_t n_reads,
                 int16_t* counts, uint64_t* hashes, std::vector<uint64_t>& nbc, std::vector<uint8_t>& pc);
}

namespace {

// A set of hash values: open addressing with linear probing, cleared in O(1)
// by a generation stamp so one instance serves many reads.
class FlatSet {
  public:
    void clear() {
        if (++gen_ == 0) {
            std::fill(stamp_.begin(), stamp_.end(), 0u);
            gen_ = 1;
        }
        size_ = 0;
    }
    void reserve(size_t n) {
        if (2 * n > keys_.size()) rehash(2 * n);
    }
    bool contains(uint64_t k) const {
        if (keys_.empty()) return false;
        const size_t mask = keys_.size() - 1;
        for (size_t i = slot(k, mask);; i = (i + 1) & mask) {
            if (stamp_[i] != gen_) return false;
            if (keys_[i] == k) return true;
        }
    }
    void insert(uint64_t k) {
        if (2 * (size_ + 1) > keys_.size()) rehash(2 * (size_ + 1));
        const size_t mask = keys_.size() - 1;
        for (size_t i = slot(k, mask);; i = (i + 1) & mask) {
            if (stamp_[i] != gen_) {
                stamp_[i] = gen_;
                keys_[i] = k;
                ++size_;
                return;
            }
            if (keys_[i] == k) return;
        }
    }
    template <class F>
    void for_each(F&& f) const {
        for (size_t i = 0; i < keys_.size(); ++i)
            if (stamp_[i] == gen_) f(keys_[i]);
    }

  private:
    std::vector<uint64_t> keys_;
    std::vector<uint32_t> stamp_;
    uint32_t gen_ = 1;
    size_t size_ = 0;

    static size_t slot(uint64_t k, size_t mask) {
        k ^= k >> 33;
        k *= 0xff51afd7ed558ccdULL;
        k ^= k >> 33;
        return (size_t)k & mask;
    }
    void rehash(size_t want) {
        size_t cap = 64;
        while (cap < want) cap <<= 1;
        std::vector<uint64_t> old;
        for_each([&](uint64_t k) { old.push_back(k); });
        keys_.assign(cap, 0);
        stamp_.assign(cap, 0);
        gen_ = 1;
        size_ = 0;
        for (uint64_t k : old) insert(k);
    }
};

constexpr uint64_t kHashDefault = ~0ULL;   // Hash<>() (canonical.hh:42-45)
constexpr int kSegWords = 7;
constexpr int kDecWords = 11;

struct Seg {
    uint64_t left_anchor = 0, right_anchor = 0, left_flank = 0, right_flank = 0;   // compactor.hh:94-103
    bool decision = false;
    uint64_t start = 0, length = 0;
    bool is_null() const { return length == 0 && left_anchor == right_anchor && !decision; }
};

struct Dec {
    uint64_t pos, hash;
    uint8_t present;        // bit c left "ACGT"[c], 4 + c right, graph or new-k-mer set
    uint64_t nb[8];
};

class Finder {
  public:
    Finder(int K, FlatSet& known, const uint64_t* prefix, uint64_t n_prefix)
        : K_((uint64_t)K), known_(known), prefix_(prefix), n_prefix_(n_prefix) {}

    // One read of n >= 1 windows.  nb / mask: the neighbour values and
    // presence of every window, or with `compact` of its new windows (count
    // 0) only, in window order -- the only ones the passes look at.  Returns
    // the number of new windows.
    uint64_t run(const int16_t* counts, const uint64_t* hv, const uint64_t* nb, const uint8_t* mask, uint64_t n,
                 bool compact, std::vector<Seg>& out, std::vector<Dec>& decs) {
        hv_ = hv;
        n_ = n;
        uint64_t n_new = 0;
        if (compact) {
            cidx_.resize(n);
            for (uint64_t i = 0; i < n; ++i) cidx_[i] = counts[i] == 0 ? n_new++ : 0;
        } else {
            for (uint64_t i = 0; i < n; ++i) n_new += counts[i] == 0;
        }
        auto at = [&](uint64_t i) { return compact ? cidx_[i] : i; };
        // runs of new, not yet seen k-mers (compactor.hh:352-400)
        std::vector<Seg> pre;
        Seg cur;
        uint64_t prev_hash = kHashDefault, cur_hash = 0;
        bool cur_new = false, prev_new = false, cur_seen = false, prev_seen = false;
        for (uint64_t pos = 0; pos < n; ++pos) {
            cur_hash = hv[pos];
            cur_new = counts[pos] == 0;
            cur_seen = known_.contains(cur_hash);
            if (cur_new && !cur_seen) {
                if (!prev_new || prev_seen) {
                    pre.push_back(cur);
                    cur = init_segment(cur_hash, prev_hash, pos);
                }
                known_.insert(cur_hash);
            } else if (!cur.is_null() && (!cur_new || cur_seen)) {
                finish_segment(cur, pos - 1, prev_hash, cur_hash, pre);
                cur = Seg();
            }
            prev_hash = cur_hash;
            prev_new = cur_new;
            prev_seen = cur_seen;
        }
        // presence of a window's neighbours in the graph or the (now final) set
        auto present = [&](uint64_t i) {
            const uint64_t j = at(i);
            uint8_t m = mask[j];
            for (int c = 0; c < 8; ++c)
                if (!((m >> c) & 1) && known_.contains(nb[8 * j + c])) m |= (uint8_t)(1u << c);
            return m;
        };
        if (cur_new && !cur_seen) {   // compactor.hh:401-417
            uint64_t right_flank = cur_hash;
            const uint8_t r = present(n - 1) >> 4;
            if (__builtin_popcount(r) == 1) right_flank = nb[8 * at(n - 1) + 4 + __builtin_ctz(r)];
            finish_segment(cur, n - 1, cur_hash, right_flank, pre);
        }
        pre.push_back(Seg());
        if (pre.size() < 3) return n_new;
        if (pre[1].start == 0) {      // compactor.hh:425-437
            const uint8_t l = present(0) & 15u;
            pre[1].left_flank = __builtin_popcount(l) == 1 ? nb[8 * at(0) + __builtin_ctz(l)] : pre[1].left_anchor;
        }
        // cuts at decision k-mers (compactor.hh:439-532)
        for (Seg seg : pre) {
            if (seg.is_null()) {
                out.push_back(seg);
                continue;
            }
            std::vector<Seg> ds;
            for (uint64_t pos = seg.start; pos + K_ <= seg.start + seg.length; ++pos) {
                const uint8_t p = present(pos);
                if (__builtin_popcount(p & 15u) > 1 || __builtin_popcount(p >> 4) > 1) {
                    Dec d;
                    d.pos = pos;
                    d.hash = hv[pos];
                    d.present = p;
                    std::memcpy(d.nb, nb + 8 * at(pos), sizeof(d.nb));
                    decs.push_back(d);
                    Seg s = init_segment(hv[pos], hv[pos], pos);
                    s.length = K_;
                    s.right_anchor = s.left_anchor;
                    s.right_flank = s.left_flank;
                    s.decision = true;
                    ds.push_back(s);
                }
            }
            split(seg, ds, out);
        }
        return n_new;
    }

  private:
    uint64_t K_;
    FlatSet& known_;
    const uint64_t* prefix_;
    uint64_t n_prefix_;
    const uint64_t* hv_ = nullptr;
    uint64_t n_ = 0;
    std::vector<uint64_t> cidx_;

    // The caller's hash vector as goetia indexes it: its earlier entries,
    // then this read's; goetia reads past the end where a decision k-mer
    // sits at a read end, and builds a segment shorter than K from it that
    // is never kept, so any value does there.
    uint64_t hash_at(uint64_t i, bool ok) const {
        if (!ok) return kHashDefault;
        if (i < n_prefix_) return prefix_[i];
        return i - n_prefix_ < n_ ? hv_[i - n_prefix_] : kHashDefault;
    }

    static Seg init_segment(uint64_t left_anchor, uint64_t left_flank, uint64_t start) {
        Seg s;
        s.start = start;
        s.left_anchor = left_anchor;
        s.left_flank = left_flank;
        return s;
    }

    void finish_segment(Seg& s, uint64_t end, uint64_t right_anchor, uint64_t right_flank, std::vector<Seg>& out) {
        s.length = end - s.start + K_;
        s.right_anchor = right_anchor;
        s.right_flank = right_flank;
        out.push_back(s);
    }

    void split(Seg& seg, const std::vector<Seg>& ds, std::vector<Seg>& out) {
        if (ds.empty()) {
            out.push_back(seg);
            return;
        }
        if (ds.size() == 1) {
            const Seg& d = ds.front();
            Seg right;
            right.left_flank = d.right_anchor;
            right.left_anchor = hash_at(d.start + 1, true);
            right.right_anchor = seg.right_anchor;
            right.right_flank = seg.right_flank;
            right.start = d.start + 1;
            right.length = seg.length - (right.start - seg.start);
            seg.right_anchor = hash_at(d.start - 1, d.start > 0);
            seg.right_flank = d.left_anchor;
            seg.length = d.start + K_ - 1 - seg.start;
            if (seg.length >= K_) out.push_back(seg);
            out.push_back(d);
            if (right.length >= K_) out.push_back(right);
            return;
        }
        const Seg& fd = ds.front();
        Seg first;
        first.left_flank = seg.left_flank;
        first.left_anchor = seg.left_anchor;
        first.right_anchor = hash_at(fd.start - 1, fd.start > 0);
        first.right_flank = hash_at(fd.start, true);
        first.start = seg.start;
        first.length = fd.start + K_ - 1 - seg.start;
        if (first.length >= K_) out.push_back(first);
        out.push_back(fd);
        for (size_t j = 1; j < ds.size(); ++j) {
            const Seg& u = ds[j - 1];
            const Seg& v = ds[j];
            Seg mid;
            mid.left_flank = u.left_anchor;
            mid.left_anchor = hash_at(u.start + 1, true);
            mid.right_anchor = hash_at(v.start - 1, true);
            mid.right_flank = v.left_anchor;
            mid.start = u.start + 1;
            mid.length = v.start + K_ - 1 - mid.start;
            if (mid.length >= K_) out.push_back(mid);
            out.push_back(v);
        }
        const Seg& ld = ds.back();
        Seg last;
        last.right_anchor = seg.right_anchor;
        last.right_flank = seg.right_flank;
        last.left_flank = ld.right_anchor;
        last.left_anchor = hash_at(ld.start + 1, true);
        last.start = ld.start + 1;
        last.length = seg.length - (last.start - seg.start);
        if (last.length >= K_) out.push_back(last);
    }
};


// Results of a run of consecutive reads.
struct Part {
    std::vector<Seg> segs;
    std::vector<Dec> decs;
    std::vector<uint64_t> seg_end, dec_end;   // per read, within the part
};

// Reads [r0, r1) whose windows start at w0 (and, with `compact`, whose new
// windows start at c0 in nb_values / nb_present); `set` is their new-k-mer
// set (fresh per read unless shared).
void run_reads(int K, const int64_t* offsets, uint64_t r0, uint64_t r1, uint64_t w0, uint64_t c0,
               const int16_t* counts, const uint64_t* hashes, const uint64_t* nb_values, const uint8_t* nb_present,
               bool compact, const uint64_t* known, uint64_t n_known, bool shared, const uint64_t* prefix,
               uint64_t n_prefix, FlatSet& set, Part& part) {
    uint64_t w = w0, c = c0;
    for (uint64_t r = r0; r < r1; ++r) {
        const int64_t len = offsets[r + 1] - offsets[r];
        if (len >= K) {
            const uint64_t n = (uint64_t)(len - K + 1);
            if (!shared) {
                set.clear();
                set.reserve(n_known + n);
                for (uint64_t i = 0; i < n_known; ++i) set.insert(known[i]);
            }
            Finder f(K, set, prefix, n_prefix);
            const uint64_t v = compact ? c : w;
            c += f.run(counts + w, hashes + w, nb_values + 8 * v, nb_present + v, n, compact, part.segs, part.decs);
            w += n;
        }
        part.seg_end.push_back(part.segs.size());
        part.dec_end.push_back(part.decs.size());
    }
}

// gt_segments_from_view, with the neighbour view either per window or, with
// `compact`, for the new windows (count 0) only, in window order.
int segments_core(uint16_t K, const int64_t* offsets, uint64_t n_reads, const int16_t* counts,
                  const uint64_t* hashes, const uint64_t* nb_values, const uint8_t* nb_present, bool compact,
                  const uint64_t* known, uint64_t n_known, int shared, const uint64_t* prefix, uint64_t n_prefix,
                  int64_t* seg_off, uint64_t* segs, uint64_t seg_cap, int64_t* dec_off, uint64_t* decs,
                  uint64_t dec_cap, uint64_t* n_segs, uint64_t* n_decs) {
    if (K < 1 || K > 255) return gt::set_error(GT_ERR_INVALID_ARG, "K must be in 1..255");
    if (n_reads && (!offsets || !seg_off || !dec_off))
        return gt::set_error(GT_ERR_INVALID_ARG, "null offsets or outputs");
    if (!n_segs || !n_decs) return gt::set_error(GT_ERR_INVALID_ARG, "null totals");
    if (n_known && !known) return gt::set_error(GT_ERR_INVALID_ARG, "null known set");
    if (n_prefix && (!prefix || n_reads != 1))
        return gt::set_error(GT_ERR_INVALID_ARG, "a hash prefix needs exactly one read");
    std::vector<uint64_t> wstart(n_reads + 1, 0);   // packed window index of each read
    for (uint64_t r = 0; r < n_reads; ++r) {
        const int64_t len = offsets[r + 1] - offsets[r];
        if (len < 0) return gt::set_error(GT_ERR_INVALID_ARG, "offsets must not decrease");
        wstart[r + 1] = wstart[r] + (len >= K ? (uint64_t)(len - K + 1) : 0);
    }
    const uint64_t nk = wstart[n_reads];
    // (a compact view is empty when no window is new)
    if (nk && (!counts || !hashes || (!compact && (!nb_values || !nb_present))))
        return gt::set_error(GT_ERR_INVALID_ARG, "null graph view");

    // Reads are independent unless they share one set: split them over
    // threads by windows (at most 16, the host share of one GPU).
    unsigned n_parts = 1;
    if (!shared && nk >= (1u << 16)) {
        n_parts = std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency()));
        n_parts = (unsigned)std::min<uint64_t>(n_parts, n_reads);
    }
    std::vector<uint64_t> cut(n_parts + 1, n_reads);
    cut[0] = 0;
    for (unsigned p = 1; p < n_parts; ++p)
        cut[p] = (uint64_t)(std::lower_bound(wstart.begin(), wstart.end(), nk * p / n_parts) - wstart.begin());
    for (unsigned p = 1; p <= n_parts; ++p) cut[p] = std::max(cut[p], cut[p - 1]);
    // with a compact view, each part's first new window
    std::vector<uint64_t> cbase(n_parts, 0);
    if (compact) {
        uint64_t c = 0;
        for (unsigned p = 0; p < n_parts; ++p) {
            cbase[p] = c;
            for (uint64_t w = wstart[cut[p]]; w < wstart[cut[p + 1]]; ++w) c += counts[w] == 0;
        }
    }
    std::vector<Part> parts(n_parts);
    auto work = [&](unsigned p) {
        FlatSet set;
        if (shared)
            for (uint64_t i = 0; i < n_known; ++i) set.insert(known[i]);
        run_reads(K, offsets, cut[p], cut[p + 1], wstart[cut[p]], cbase[p], counts, hashes, nb_values, nb_present,
                  compact, known, n_known, shared != 0, prefix, n_prefix, set, parts[p]);
    };
    if (n_parts == 1) {
        work(0);
    } else {
        std::vector<std::thread> th;
        for (unsigned p = 0; p < n_parts; ++p) th.emplace_back(work, p);
        for (auto& t : th) t.join();
    }
    uint64_t total_s = 0, total_d = 0;
    for (const Part& part : parts) {
        total_s += part.segs.size();
        total_d += part.decs.size();
    }
    *n_segs = total_s;
    *n_decs = total_d;
    if (total_s > seg_cap || total_d > dec_cap || (total_s && !segs) || (total_d && !decs))
        return gt::set_error(GT_ERR_INVALID_ARG, "segment outputs too small: need " + std::to_string(total_s) +
                                                      " segments and " + std::to_string(total_d) + " decisions");
    uint64_t s0 = 0, d0 = 0, r = 0;
    if (seg_off) seg_off[0] = 0;
    if (dec_off) dec_off[0] = 0;
    for (const Part& part : parts) {
        for (size_t i = 0; i < part.segs.size(); ++i) {
            const Seg& sg = part.segs[i];
            uint64_t* o = segs + kSegWords * (s0 + i);
            o[0] = sg.left_anchor;
            o[1] = sg.right_anchor;
            o[2] = sg.left_flank;
            o[3] = sg.right_flank;
            o[4] = sg.decision ? 1 : 0;
            o[5] = sg.start;
            o[6] = sg.length;
        }
        for (size_t i = 0; i < part.decs.size(); ++i) {
            uint64_t* o = decs + kDecWords * (d0 + i);
            o[0] = part.decs[i].pos;
            o[1] = part.decs[i].hash;
            o[2] = part.decs[i].present;
            std::memcpy(o + 3, part.decs[i].nb, sizeof(part.decs[i].nb));
        }
        for (size_t k = 0; k < part.seg_end.size(); ++k, ++r) {
            seg_off[r + 1] = (int64_t)(s0 + part.seg_end[k]);
            dec_off[r + 1] = (int64_t)(d0 + part.dec_end[k]);
        }
        s0 += part.segs.size();
        d0 += part.decs.size();
    }
    return GT_OK;
}

}  // namespace

extern "C" int gt_segments_from_view(uint16_t K, const int64_t* offsets, uint64_t n_reads, const int16_t* counts,
                                     const uint64_t* hashes, const uint64_t* nb_values, const uint8_t* nb_present,
                                     const uint64_t* known, uint64_t n_known, int shared, const uint64_t* prefix,
                                     uint64_t n_prefix, int64_t* seg_off, uint64_t* segs, uint64_t seg_cap,
                                     int64_t* dec_off, uint64_t* decs, uint64_t dec_cap, uint64_t* n_segs,
                                     uint64_t* n_decs) {
    return segments_core(K, offsets, n_reads, counts, hashes, nb_values, nb_present, false, known, n_known, shared,
                         prefix, n_prefix, seg_off, segs, seg_cap, dec_off, decs, dec_cap, n_segs, n_decs);
}

extern "C" int gt_find_new_segments(gt_storage* st, uint16_t K, int mode, const uint8_t* seq, const int64_t* offsets,
                                    uint64_t n_reads, const uint64_t* known, uint64_t n_known, int shared,
                                    const uint64_t* prefix, uint64_t n_prefix, int16_t* counts, uint64_t* hashes,
                                    int64_t* seg_off, uint64_t* segs, uint64_t seg_cap, int64_t* dec_off,
                                    uint64_t* decs, uint64_t dec_cap, uint64_t* n_segs, uint64_t* n_decs) {
    if (K < 1 || K > 255) return gt::set_error(GT_ERR_INVALID_ARG, "K must be in 1..255");
    if (mode != GT_MODE_RAW)
        return gt::set_error(GT_ERR_UNSUPPORTED, "find_new_segments hashes raw bytes (GT_MODE_RAW), as goetia does");
    if (n_reads && !offsets) return gt::set_error(GT_ERR_INVALID_ARG, "null offsets");
    uint64_t nk = 0;
    for (uint64_t r = 0; r < n_reads; ++r) {
        const int64_t len = offsets[r + 1] - offsets[r];
        if (len < 0) return gt::set_error(GT_ERR_INVALID_ARG, "offsets must not decrease");
        if (len >= K) nk += (uint64_t)(len - K + 1);
    }
    if (nk && (!counts || !hashes || !seq)) return gt::set_error(GT_ERR_INVALID_ARG, "null reads or per-window outputs");
    std::vector<uint64_t> nbc;     // neighbour values of the new windows only
    std::vector<uint8_t> pc;       // and their presence masks
    if (nk) {
        const int rc = gt::segment_view(st, K, seq, offsets, n_reads, counts, hashes, nbc, pc);
        if (rc != GT_OK) return rc;
    }
    return segments_core(K, offsets, n_reads, counts, hashes, nbc.data(), pc.data(), true, known, n_known, shared,
                         prefix, n_prefix, seg_off, segs, seg_cap, dec_off, decs, dec_cap, n_segs, n_decs);
}
