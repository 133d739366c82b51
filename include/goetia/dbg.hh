/**
 * dbg.hh -- dBG<StorageType, ShifterType> (goetia dbg.hh:39-438) over a
 * device storage, with the neighbour / degree functions of
 * unitig_walker.hh:322-418,1003-1117.  Per-k-mer and per-read calls keep
 * goetia's signatures and return conventions; batched calls over many reads
 * (insert_sequences, query_sequences, degrees) sit beside them.  Every
 * storage and sequence operation runs on the GPU through goetia_amd.h.
 */
#ifndef GOETIA_AMD_DBG_HH
#define GOETIA_AMD_DBG_HH

#include <cstdint>
#include <memory>
#include <set>
#include <string>
#include <utility>
#include <vector>

#include "goetia/hashing/canonical.hh"
#include "goetia/hashing/hashshifter.hh"
#include "goetia/hashing/kmeriterator.hh"
#include "goetia/storage/storage_types.hh"

namespace goetia {

// Concatenated reads for the batched calls (offsets has n + 1 entries).
struct ReadBatch {
    std::string seq;
    std::vector<int64_t> offsets{0};

    ReadBatch() = default;
    explicit ReadBatch(const std::vector<std::string>& reads) {
        for (const auto& r : reads) push_back(r);
    }
    void push_back(const std::string& r) {
        seq += r;
        offsets.push_back((int64_t)seq.size());
    }
    uint64_t size() const { return offsets.size() - 1; }
    std::string read(uint64_t i) const { return seq.substr(offsets[i], offsets[i + 1] - offsets[i]); }
    uint64_t n_kmers(uint16_t K) const {
        uint64_t n = 0;
        for (size_t i = 0; i + 1 < offsets.size(); ++i) {
            const int64_t len = offsets[i + 1] - offsets[i];
            if (len >= K) n += (uint64_t)(len - K + 1);
        }
        return n;
    }
    const uint8_t* data() const { return reinterpret_cast<const uint8_t*>(seq.data()); }
};

template <class StorageType, class ShifterType>
class dBG {
  public:
    typedef ShifterType shifter_type;
    typedef StorageType storage_type;
    typedef StorageTraits<storage_type> storage_traits;
    typedef typename ShifterType::hash_type hash_type;
    typedef typename ShifterType::kmer_type kmer_type;
    typedef typename ShifterType::alphabet alphabet;
    typedef KmerIterator<ShifterType> kmer_iter_type;
    template <bool Dir>
    using shift_type = Shift<hash_type, Dir>;
    typedef std::pair<std::vector<shift_type<DIR_LEFT>>, std::vector<shift_type<DIR_RIGHT>>> shift_pair_type;
    typedef std::pair<std::vector<kmer_type>, std::vector<kmer_type>> neighbor_pair_type;   // unitig_walker.hh:188-189

    const uint16_t K;

    dBG(std::shared_ptr<StorageType> S, ShifterType& shifter) : K(shifter.K), S(S), _shifter(shifter.K) {}
    dBG(std::shared_ptr<StorageType> S, uint16_t K) : K(K), S(S), _shifter(K) {}
    // dbg.hh:97-101: a reference copy sharing the storage.
    dBG(const dBG& other) : K(other.K), S(other.S), _shifter(other.K), _mode(other._mode) {}

    static std::shared_ptr<dBG> build(std::shared_ptr<StorageType> S, ShifterType& hasher) {
        return std::make_shared<dBG>(S, hasher);
    }
    static std::shared_ptr<dBG> build(std::shared_ptr<StorageType> S, uint16_t K) {
        return std::make_shared<dBG>(S, K);
    }
    // dbg.hh:128-131: new empty storage of the same sizes.
    std::shared_ptr<dBG> clone() { return std::make_shared<dBG>(S->clone(), K); }

    // GT_MODE_RAW (goetia's behaviour) or GT_MODE_SKIP_N.
    void set_input_mode(int mode) { _mode = mode; }
    int input_mode() const { return _mode; }
    static constexpr int is_canonical() { return ShifterType::is_canonical ? 1 : 0; }

    // ------------------------------------------------------------ per k-mer
    const hash_type hash(const std::string& kmer) const { return _shifter.hash(kmer); }
    const bool insert(const std::string& kmer) { return S->insert(hash(kmer).value()); }
    const bool insert(const hash_type& h) { return S->insert(h.value()); }
    const count_t insert_and_query(const std::string& kmer) { return S->insert_and_query(hash(kmer).value()); }
    const count_t insert_and_query(hash_type& h) { return S->insert_and_query(h.value()); }
    const count_t query(const std::string& kmer) { return query(hash(kmer)); }
    const count_t query(const hash_type& h) const { return S->query(h.value()); }
    uint64_t n_unique() const { return S->n_unique_kmers(); }
    uint64_t n_occupied() const { return S->n_occupied(); }
    uint8_t** get_raw() const { return S->get_raw_tables(); }
    double estimated_fp() { return S->estimated_fp(); }
    void reset() { S->reset(); }
    void save(std::string filename) { S->save(filename, K); }
    void load(std::string filename) {
        uint16_t k = K;
        S->load(filename, k);
    }
    const std::string suffix(const std::string& kmer) { return kmer.substr(kmer.length() - K + 1); }
    const std::string prefix(const std::string& kmer) { return kmer.substr(0, K - 1); }
    std::shared_ptr<kmer_iter_type> get_hash_iter(const std::string& sequence) {
        return std::make_shared<kmer_iter_type>(sequence, K);
    }
    ShifterType get_hasher() { return ShifterType(K); }

    // ------------------------------------------------------------ per read
    // dbg.hh:249-265
    uint64_t insert_sequence(const std::string& sequence, std::vector<hash_type>& kmer_hashes,
                             std::vector<count_t>& counts) {
        check_len(sequence);
        auto c = insert_and_query_sequence(sequence);
        kmer_iter_type it(sequence, K);
        while (!it.done()) kmer_hashes.push_back(it.next());
        counts.insert(counts.end(), c.begin(), c.end());
        return sequence.size() - K + 1;
    }

    // dbg.hh:267-280
    uint64_t insert_sequence(const std::string& sequence, std::set<hash_type>& new_kmers) {
        check_len(sequence);
        std::vector<uint8_t> flags;
        insert_one(sequence, nullptr, &flags);
        kmer_iter_type it(sequence, K);
        for (size_t i = 0; !it.done(); ++i) {
            hash_type h = it.next();
            if (flags[i]) new_kmers.insert(h);
        }
        return sequence.size() - K + 1;
    }

    // dbg.hh:282-294
    uint64_t insert_sequence(const std::string& sequence, std::vector<hash_type>& hashes) {
        check_len(sequence);
        insert_one(sequence, nullptr, nullptr);
        kmer_iter_type it(sequence, K);
        while (!it.done()) hashes.push_back(it.next());
        return sequence.size() - K + 1;
    }

    // dbg.hh:296-305
    uint64_t insert_sequence(const std::string& sequence) {
        check_len(sequence);
        insert_one(sequence, nullptr, nullptr);
        return sequence.size() - K + 1;
    }

    // dbg.hh:307-318
    uint64_t insert_sequence(const std::string& sequence, uint64_t& n_new) {
        check_len(sequence);
        insert_one(sequence, &n_new, nullptr);
        return sequence.size() - K + 1;
    }

    // dbg.hh:327-340
    std::vector<count_t> insert_and_query_sequence(const std::string& sequence) {
        check_len(sequence);
        const int64_t off[2] = {0, (int64_t)sequence.size()};
        std::vector<count_t> counts(sequence.size() - K + 1);
        detail::check(gt_dbg_insert_and_query(S->handle(), K, canonical(), _mode,
                                              reinterpret_cast<const uint8_t*>(sequence.data()), off, 1,
                                              counts.data()),
                      "insert_and_query_sequence");
        return counts;
    }

    // dbg.hh:349-362
    std::vector<count_t> query_sequence(const std::string& sequence) {
        check_len(sequence);
        std::vector<count_t> counts(sequence.size() - K + 1);
        query_one(sequence, counts.data(), nullptr);
        return counts;
    }

    // dbg.hh:364-376
    void query_sequence(const std::string& sequence, std::vector<count_t>& counts, std::vector<hash_type>& hashes) {
        auto c = query_sequence(sequence);
        kmer_iter_type it(sequence, K);
        while (!it.done()) hashes.push_back(it.next());
        counts.insert(counts.end(), c.begin(), c.end());
    }

    // dbg.hh:378-394
    void query_sequence(const std::string& sequence, std::vector<count_t>& counts, std::vector<hash_type>& hashes,
                        std::set<hash_type>& new_hashes) {
        auto c = query_sequence(sequence);
        kmer_iter_type it(sequence, K);
        for (size_t i = 0; !it.done(); ++i) {
            hash_type h = it.next();
            if (c[i] == 0) new_hashes.insert(h);
            hashes.push_back(h);
        }
        counts.insert(counts.end(), c.begin(), c.end());
    }

    // ------------------------------------------------------------ batches
    // insert_sequence(seq, n_new) over every read, in order; returns n_new.
    std::vector<uint64_t> insert_sequences(const ReadBatch& reads, std::vector<uint8_t>* new_flags = nullptr) {
        std::vector<uint64_t> n_new(reads.size());
        if (new_flags) new_flags->assign(reads.n_kmers(K), 0);
        uint64_t nk = 0;
        detail::check(gt_dbg_insert(S->handle(), K, canonical(), _mode, reads.data(), reads.offsets.data(),
                                    reads.size(), n_new.data(), new_flags ? new_flags->data() : nullptr, &nk),
                      "insert_sequences");
        return n_new;
    }

    // query_sequence over every read: counts packed per read (-1 = skipped).
    std::vector<count_t> query_sequences(const ReadBatch& reads) {
        std::vector<count_t> counts(reads.n_kmers(K));
        detail::check(gt_dbg_query(S->handle(), K, canonical(), _mode, reads.data(), reads.offsets.data(),
                                   reads.size(), counts.data(), nullptr),
                      "query_sequences");
        return counts;
    }

    // In/out degree of every window (0xFF where skipped).
    void degrees(const ReadBatch& reads, std::vector<uint8_t>& in_deg, std::vector<uint8_t>& out_deg,
                 int min_count = 1) {
        in_deg.assign(reads.n_kmers(K), 0);
        out_deg.assign(reads.n_kmers(K), 0);
        detail::check(gt_dbg_degrees(S->handle(), K, canonical(), _mode, min_count, reads.data(),
                                     reads.offsets.data(), reads.size(), nullptr, in_deg.data(), out_deg.data()),
                      "degrees");
    }

    // query_sequence over every read with the window hashes (value()).
    std::vector<count_t> query_sequences(const ReadBatch& reads, std::vector<uint64_t>& hashes) {
        std::vector<count_t> counts(reads.n_kmers(K));
        hashes.assign(counts.size(), 0);
        detail::check(gt_dbg_query(S->handle(), K, canonical(), _mode, reads.data(), reads.offsets.data(),
                                   reads.size(), counts.data(), hashes.data()),
                      "query_sequences");
        return counts;
    }

    // Degrees plus every window's eight neighbours: values[8 i + c] the left
    // neighbour with symbol "ACGT"[c], values[8 i + 4 + c] the right one;
    // bit c / 4 + c of present[i] set when that neighbour's query is >=
    // min_count (left_extensions / right_extensions, hashextender.hh:352-396).
    void window_neighbors(const ReadBatch& reads, std::vector<uint8_t>& in_deg, std::vector<uint8_t>& out_deg,
                          std::vector<uint64_t>& values, std::vector<uint8_t>& present, int min_count = 1) {
        const uint64_t nk = reads.n_kmers(K);
        in_deg.assign(nk, 0);
        out_deg.assign(nk, 0);
        values.assign(nk * 8, 0);
        present.assign(nk, 0);
        detail::check(gt_dbg_neighbors(S->handle(), K, canonical(), _mode, min_count, reads.data(),
                                       reads.offsets.data(), reads.size(), in_deg.data(), out_deg.data(),
                                       values.data(), present.data()),
                      "window_neighbors");
    }

    // ------------------------------------------------------------ neighbours
    hash_type set_cursor(const std::string& kmer) {
        if (kmer.length() < K) throw InvalidSequenceException("Sequence must at least length K");
        _cursor = kmer.substr(0, K);
        return _cursor_shifter().hash_base(_cursor);
    }
    const std::string get_cursor() const { return _cursor; }

    // hashextender.hh:327-353: move the cursor one base (the K-base ring
    // drops its other end) and return the new k-mer's hash; get() is the
    // cursor's hash.
    hash_type shift_right(const char c) {
        auto h = _cursor_shifter().shift_right(_cursor.front(), c);
        _cursor = _cursor.substr(1) + c;
        return h;
    }
    hash_type shift_left(const char c) {
        auto h = _cursor_shifter().shift_left(c, _cursor.back());
        _cursor = c + _cursor.substr(0, K - 1);
        return h;
    }
    hash_type get() { return _cursor_shifter().get(); }

    std::vector<shift_type<DIR_LEFT>> left_extensions() {
        auto& sh = _cursor_shifter();
        std::vector<shift_type<DIR_LEFT>> out;
        const char back = _cursor.back();
        for (const char c : alphabet::SYMBOLS) {   // hashextender.hh:352-364
            out.emplace_back(sh.shift_left(c, back), c);
            sh.shift_right(c, back);
        }
        return out;
    }

    std::vector<shift_type<DIR_RIGHT>> right_extensions() {
        auto& sh = _cursor_shifter();
        std::vector<shift_type<DIR_RIGHT>> out;
        const char front = _cursor.front();
        for (const char c : alphabet::SYMBOLS) {   // hashextender.hh:386-396
            out.emplace_back(sh.shift_right(front, c), c);
            sh.shift_left(front, c);
        }
        return out;
    }

    // unitig_walker.hh:361-371 (present iff query != 0), :407-418
    template <bool Dir>
    size_t count_nodes(const std::vector<shift_type<Dir>>& exts) {
        return filter_nodes(exts).size();
    }

    template <bool Dir>
    std::vector<shift_type<Dir>> filter_nodes(const std::vector<shift_type<Dir>>& exts) {
        std::vector<uint64_t> hs;
        for (const auto& e : exts) hs.push_back(e.value());
        std::vector<count_t> c(hs.size());
        detail::check(gt_query_hashes(S->handle(), hs.data(), hs.size(), c.data()), "filter_nodes");
        std::vector<shift_type<Dir>> out;
        for (size_t i = 0; i < exts.size(); ++i) {
            if (c[i] != 0) out.push_back(exts[i]);
        }
        return out;
    }

    // unitig_walker.hh:386-396, 457-468: present iff query != 0 or the
    // node is among `extras` (nodes about to join the graph).
    template <bool Dir>
    std::vector<shift_type<Dir>> filter_nodes(const std::vector<shift_type<Dir>>& exts,
                                              const std::set<hash_type>& extras) {
        std::vector<uint64_t> hs;
        for (const auto& e : exts) hs.push_back(e.value());
        std::vector<count_t> c(hs.size());
        detail::check(gt_query_hashes(S->handle(), hs.data(), hs.size(), c.data()), "filter_nodes");
        std::vector<shift_type<Dir>> out;
        for (size_t i = 0; i < exts.size(); ++i) {
            if (c[i] != 0 || extras.count(exts[i].hash)) out.push_back(exts[i]);
        }
        return out;
    }
    template <bool Dir>
    size_t count_nodes(const std::vector<shift_type<Dir>>& exts, const std::set<hash_type>& extras) {
        return filter_nodes(exts, extras).size();
    }
    size_t in_degree(const std::set<hash_type>& extras) { return count_nodes(left_extensions(), extras); }
    size_t out_degree(const std::set<hash_type>& extras) { return count_nodes(right_extensions(), extras); }

    size_t in_degree() { return count_nodes(left_extensions()); }
    size_t out_degree() { return count_nodes(right_extensions()); }
    size_t degree() { return in_degree() + out_degree(); }

    std::vector<shift_type<DIR_LEFT>> in_neighbors(const std::string& seed) {
        set_cursor(seed);
        return filter_nodes(left_extensions());
    }
    std::vector<shift_type<DIR_RIGHT>> out_neighbors(const std::string& seed) {
        set_cursor(seed);
        return filter_nodes(right_extensions());
    }
    shift_pair_type neighbors(const std::string& seed) {
        set_cursor(seed);
        return std::make_pair(filter_nodes(left_extensions()), filter_nodes(right_extensions()));
    }

    // unitig_walker.hh:1003-1023
    bool is_decision_kmer(const std::string& node) {
        set_cursor(node);
        return in_degree() > 1 || out_degree() > 1;
    }
    bool is_decision_kmer(const std::string& node, uint8_t& degree) {
        set_cursor(node);
        const size_t l = in_degree(), r = out_degree();
        degree = (uint8_t)(l + r);
        return l > 1 || r > 1;
    }

    // unitig_walker.hh:612-640: neighbour k-mers of a root from its shifts.
    std::vector<kmer_type> build_left_kmers(const std::vector<shift_type<DIR_LEFT>>& nodes,
                                            const std::string& root) {
        std::vector<kmer_type> kmers;
        const std::string pre = root.substr(0, K - 1);
        for (const auto& n : nodes) kmers.push_back(kmer_type(n.hash, n.symbol + pre));
        return kmers;
    }
    std::vector<kmer_type> build_right_kmers(const std::vector<shift_type<DIR_RIGHT>>& nodes,
                                             const std::string& root) {
        std::vector<kmer_type> kmers;
        const std::string suf = root.substr(root.size() - K + 1);
        for (const auto& n : nodes) kmers.push_back(kmer_type(n.hash, suf + n.symbol));
        return kmers;
    }

    // unitig_walker.hh:1085-1117: whether root is a decision k-mer, and then
    // its (left, right) neighbour k-mers in result.
    bool get_decision_neighbors(const std::string& root, neighbor_pair_type& result) {
        set_cursor(root);
        const std::string r = get_cursor();
        auto left = build_left_kmers(filter_nodes(left_extensions()), r);
        auto right = build_right_kmers(filter_nodes(right_extensions()), r);
        if (left.size() > 1 || right.size() > 1) {
            result = std::make_pair(left, right);
            return true;
        }
        return false;
    }

    // unitig_walker.hh:1047-1082: the same with union_nodes counted as
    // present.
    bool get_decision_neighbors(const std::string& root, neighbor_pair_type& result,
                                const std::set<hash_type>& union_nodes) {
        set_cursor(root);
        const std::string r = get_cursor();
        auto left = build_left_kmers(filter_nodes(left_extensions(), union_nodes), r);
        auto right = build_right_kmers(filter_nodes(right_extensions(), union_nodes), r);
        if (left.size() > 1 || right.size() > 1) {
            result = std::make_pair(left, right);
            return true;
        }
        return false;
    }

    // unitig_walker.hh:1025-1045 with the decision neighbours.
    void find_decision_kmers(const std::string& sequence, std::vector<uint32_t>& decision_positions,
                             std::vector<hash_type>& decision_hashes,
                             std::vector<neighbor_pair_type>& decision_neighbors) {
        const size_t first = decision_positions.size();
        find_decision_kmers(sequence, decision_positions, decision_hashes);
        for (size_t i = first; i < decision_positions.size(); ++i) {
            neighbor_pair_type nb;
            get_decision_neighbors(sequence.substr(decision_positions[i], K), nb);
            decision_neighbors.push_back(nb);
        }
    }

    // unitig_walker.hh:1025-1045 (positions and hashes of the decision
    // windows of one sequence; degrees computed on the GPU).
    void find_decision_kmers(const std::string& sequence, std::vector<uint32_t>& decision_positions,
                             std::vector<hash_type>& decision_hashes) {
        check_len(sequence);
        ReadBatch b;
        b.push_back(sequence);
        std::vector<uint8_t> in, out;
        degrees(b, in, out);
        kmer_iter_type it(sequence, K);
        for (uint32_t i = 0; !it.done(); ++i) {
            hash_type h = it.next();
            if (in[i] != 0xFF && (in[i] > 1 || out[i] > 1)) {
                decision_positions.push_back(i);
                decision_hashes.push_back(h);
            }
        }
    }

    std::shared_ptr<StorageType> storage() const { return S; }

  protected:
    std::shared_ptr<StorageType> S;

  private:
    static constexpr int canonical() { return ShifterType::is_canonical ? 1 : 0; }

    void check_len(const std::string& s) const {
        if (s.length() < K) throw SequenceLengthException("Sequence must have length >= K");
    }

    void insert_one(const std::string& s, uint64_t* n_new, std::vector<uint8_t>* flags) {
        const int64_t off[2] = {0, (int64_t)s.size()};
        uint64_t nn = 0, nk = 0;
        if (flags) flags->assign(s.size() - K + 1, 0);
        detail::check(gt_dbg_insert(S->handle(), K, canonical(), _mode, reinterpret_cast<const uint8_t*>(s.data()),
                                    off, 1, &nn, flags ? flags->data() : nullptr, &nk),
                      "insert_sequence");
        if (n_new) *n_new = nn;
    }

    void query_one(const std::string& s, count_t* counts, uint64_t* hashes) {
        const int64_t off[2] = {0, (int64_t)s.size()};
        detail::check(gt_dbg_query(S->handle(), K, canonical(), _mode, reinterpret_cast<const uint8_t*>(s.data()),
                                   off, 1, counts, hashes),
                      "query_sequence");
    }

    ShifterType& _cursor_shifter() {
        if (!_cur) _cur = std::make_unique<ShifterType>(K);
        return *_cur;
    }

    ShifterType _shifter;
    std::unique_ptr<ShifterType> _cur;
    std::string _cursor;
    int _mode = GT_MODE_RAW;
};

}  // namespace goetia

#endif
