// StreamingCompactor<dBG<S, FwdLemireShifter>>::Compactor::find_new_segments
// (goetia compactor.hh:321-534) over a batch of reads.
//
// The graph view of a read -- every window's query count, hash and eight
// neighbour values with their presence bits -- comes from two GPU launches
// (gt_dbg_query, gt_dbg_neighbors) for the whole batch; the two passes over
// each read (runs of new k-mers, then the cuts at decision k-mers with
// goetia's anchor and flank bookkeeping) are sequential per read and run
// here on the host.  gt_segments_from_view takes the view from the caller
// (no device needed); gt_find_new_segments builds it on the GPU first.

#include "goetia_amd.h"

#include <algorithm>
#include <cstring>
#include <string>
#include <unordered_set>
#include <vector>

namespace gt {
int set_error(int code, const std::string& msg);
}

namespace {

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
    Finder(int K, std::unordered_set<uint64_t>& known, const uint64_t* prefix, uint64_t n_prefix)
        : K_((uint64_t)K), known_(known), prefix_(prefix), n_prefix_(n_prefix) {}

    // One read of n >= 1 windows.
    void run(const int16_t* counts, const uint64_t* hv, const uint64_t* nb, const uint8_t* mask, uint64_t n,
             std::vector<Seg>& out, std::vector<Dec>& decs) {
        hv_ = hv;
        n_ = n;
        // runs of new, not yet seen k-mers (compactor.hh:352-400)
        std::vector<Seg> pre;
        Seg cur;
        uint64_t prev_hash = kHashDefault, cur_hash = 0;
        bool cur_new = false, prev_new = false, cur_seen = false, prev_seen = false;
        for (uint64_t pos = 0; pos < n; ++pos) {
            cur_hash = hv[pos];
            cur_new = counts[pos] == 0;
            cur_seen = known_.count(cur_hash) != 0;
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
        // presence of every neighbour in the graph or the (now final) set
        present_.resize(n);
        for (uint64_t i = 0; i < n; ++i) {
            uint8_t m = mask[i];
            for (int c = 0; c < 8; ++c)
                if (known_.count(nb[8 * i + c])) m |= (uint8_t)(1u << c);
            present_[i] = m;
        }
        if (cur_new && !cur_seen) {   // compactor.hh:401-417
            uint64_t right_flank = cur_hash;
            const uint8_t r = present_[n - 1] >> 4;
            if (__builtin_popcount(r) == 1) right_flank = nb[8 * (n - 1) + 4 + __builtin_ctz(r)];
            finish_segment(cur, n - 1, cur_hash, right_flank, pre);
        }
        pre.push_back(Seg());
        if (pre.size() < 3) return;
        if (pre[1].start == 0) {      // compactor.hh:425-437
            const uint8_t l = present_[0] & 15u;
            pre[1].left_flank = __builtin_popcount(l) == 1 ? nb[__builtin_ctz(l)] : pre[1].left_anchor;
        }
        // cuts at decision k-mers (compactor.hh:439-532)
        for (Seg seg : pre) {
            if (seg.is_null()) {
                out.push_back(seg);
                continue;
            }
            std::vector<Seg> ds;
            for (uint64_t pos = seg.start; pos + K_ <= seg.start + seg.length; ++pos) {
                const uint8_t p = present_[pos];
                if (__builtin_popcount(p & 15u) > 1 || __builtin_popcount(p >> 4) > 1) {
                    Dec d;
                    d.pos = pos;
                    d.hash = hv[pos];
                    d.present = p;
                    std::memcpy(d.nb, nb + 8 * pos, sizeof(d.nb));
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
    }

  private:
    uint64_t K_;
    std::unordered_set<uint64_t>& known_;
    const uint64_t* prefix_;
    uint64_t n_prefix_;
    const uint64_t* hv_ = nullptr;
    uint64_t n_ = 0;
    std::vector<uint8_t> present_;

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

}  // namespace

extern "C" int gt_segments_from_view(uint16_t K, const int64_t* offsets, uint64_t n_reads, const int16_t* counts,
                                     const uint64_t* hashes, const uint64_t* nb_values, const uint8_t* nb_present,
                                     const uint64_t* known, uint64_t n_known, int shared, const uint64_t* prefix,
                                     uint64_t n_prefix, int64_t* seg_off, uint64_t* segs, uint64_t seg_cap,
                                     int64_t* dec_off, uint64_t* decs, uint64_t dec_cap, uint64_t* n_segs,
                                     uint64_t* n_decs) {
    if (K < 1 || K > 255) return gt::set_error(GT_ERR_INVALID_ARG, "K must be in 1..255");
    if (n_reads && (!offsets || !seg_off || !dec_off))
        return gt::set_error(GT_ERR_INVALID_ARG, "null offsets or outputs");
    if (!n_segs || !n_decs) return gt::set_error(GT_ERR_INVALID_ARG, "null totals");
    if (n_known && !known) return gt::set_error(GT_ERR_INVALID_ARG, "null known set");
    if (n_prefix && (!prefix || n_reads != 1))
        return gt::set_error(GT_ERR_INVALID_ARG, "a hash prefix needs exactly one read");
    uint64_t nk = 0;
    for (uint64_t r = 0; r < n_reads; ++r) {
        const int64_t len = offsets[r + 1] - offsets[r];
        if (len < 0) return gt::set_error(GT_ERR_INVALID_ARG, "offsets must not decrease");
        if (len >= K) nk += (uint64_t)(len - K + 1);
    }
    if (nk && (!counts || !hashes || !nb_values || !nb_present))
        return gt::set_error(GT_ERR_INVALID_ARG, "null graph view");
    std::unordered_set<uint64_t> base_known(known, known + n_known);
    std::vector<Seg> out;
    std::vector<Dec> ds;
    uint64_t w = 0;
    seg_off[0] = 0;
    dec_off[0] = 0;
    for (uint64_t r = 0; r < n_reads; ++r) {
        const int64_t len = offsets[r + 1] - offsets[r];
        if (len >= K) {
            const uint64_t n = (uint64_t)(len - K + 1);
            std::unordered_set<uint64_t> fresh;
            if (!shared) fresh = base_known;
            std::unordered_set<uint64_t>& set = shared ? base_known : fresh;
            Finder f(K, set, prefix, n_prefix);
            f.run(counts + w, hashes + w, nb_values + 8 * w, nb_present + w, n, out, ds);
            w += n;
        }
        seg_off[r + 1] = (int64_t)out.size();
        dec_off[r + 1] = (int64_t)ds.size();
    }
    *n_segs = out.size();
    *n_decs = ds.size();
    if (out.size() > seg_cap || ds.size() > dec_cap || (out.size() && !segs) || (ds.size() && !decs))
        return gt::set_error(GT_ERR_INVALID_ARG, "segment outputs too small: need " + std::to_string(out.size()) +
                                                      " segments and " + std::to_string(ds.size()) + " decisions");
    for (size_t i = 0; i < out.size(); ++i) {
        const Seg& s = out[i];
        uint64_t* o = segs + kSegWords * i;
        o[0] = s.left_anchor;
        o[1] = s.right_anchor;
        o[2] = s.left_flank;
        o[3] = s.right_flank;
        o[4] = s.decision ? 1 : 0;
        o[5] = s.start;
        o[6] = s.length;
    }
    for (size_t i = 0; i < ds.size(); ++i) {
        uint64_t* o = decs + kDecWords * i;
        o[0] = ds[i].pos;
        o[1] = ds[i].hash;
        o[2] = ds[i].present;
        std::memcpy(o + 3, ds[i].nb, sizeof(ds[i].nb));
    }
    return GT_OK;
}

extern "C" int gt_find_new_segments(gt_storage* st, uint16_t K, int mode, const uint8_t* seq, const int64_t* offsets,
                                    uint64_t n_reads, const uint64_t* known, uint64_t n_known, int shared,
                                    const uint64_t* prefix, uint64_t n_prefix, int16_t* counts, uint64_t* hashes,
                                    int64_t* seg_off, uint64_t* segs, uint64_t seg_cap, int64_t* dec_off,
                                    uint64_t* decs, uint64_t dec_cap, uint64_t* n_segs, uint64_t* n_decs) {
    if (K < 1 || K > 255) return gt::set_error(GT_ERR_INVALID_ARG, "K must be in 1..255");
    if (n_reads && !offsets) return gt::set_error(GT_ERR_INVALID_ARG, "null offsets");
    uint64_t nk = 0;
    for (uint64_t r = 0; r < n_reads; ++r) {
        const int64_t len = offsets[r + 1] - offsets[r];
        if (len < 0) return gt::set_error(GT_ERR_INVALID_ARG, "offsets must not decrease");
        if (len >= K) nk += (uint64_t)(len - K + 1);
    }
    if (nk && (!counts || !hashes)) return gt::set_error(GT_ERR_INVALID_ARG, "null per-window outputs");
    std::vector<uint8_t> in(nk), out(nk), present(nk);
    std::vector<uint64_t> nb(nk * 8);
    if (nk) {
        // forward hashes (goetia builds the compactor for Hash<> only)
        int rc = gt_dbg_query(st, K, 0, mode, seq, offsets, n_reads, counts, hashes);
        if (rc != GT_OK) return rc;
        rc = gt_dbg_neighbors(st, K, 0, mode, 1, seq, offsets, n_reads, in.data(), out.data(), nb.data(),
                              present.data());
        if (rc != GT_OK) return rc;
    }
    return gt_segments_from_view(K, offsets, n_reads, counts, hashes, nb.data(), present.data(), known, n_known,
                                 shared, prefix, n_prefix, seg_off, segs, seg_cap, dec_off, decs, dec_cap, n_segs,
                                 n_decs);
}
