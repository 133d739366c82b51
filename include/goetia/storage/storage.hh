/**
 * storage.hh -- storage plugin interface (goetia storage.hh:84-190):
 * StorageTraits, the abstract Storage<uint64_t> with its virtual
 * insert / query / insert_and_query / counters / raw tables / reset, and
 * get_n_primes_near_x.  The concrete storages keep their tables in HBM
 * behind a gt_storage handle of libgoetia_amd.
 */
#ifndef GOETIA_AMD_STORAGE_HH
#define GOETIA_AMD_STORAGE_HH

#include <cmath>
#include <cstdint>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "goetia/goetia_amd.hh"

namespace goetia {

template <class S>
struct StorageTraits {
    static constexpr bool is_counting = false;
    static constexpr bool is_probabilistic = false;
    typedef std::tuple<bool> params_type;
    static constexpr params_type default_params = std::make_tuple(false);
};

// storage.hh:98-143
template <typename ValueType>
class Storage {
  public:
    typedef ValueType value_type;
    virtual ~Storage() {}
    virtual void save(std::string, uint16_t) = 0;
    virtual void load(std::string, uint16_t&) = 0;
    virtual const uint64_t n_occupied() const = 0;
    virtual const uint64_t n_unique_kmers() const = 0;
    virtual const bool insert(value_type khash) = 0;
    virtual const count_t insert_and_query(value_type khash) = 0;
    virtual const count_t query(value_type khash) const = 0;
    virtual byte_t** get_raw_tables() = 0;
    virtual void reset() = 0;
};

// storage.hh:166-190
inline std::vector<uint64_t> get_n_primes_near_x(uint32_t n, uint64_t x) {
    std::vector<uint64_t> out(n ? n : 1);
    uint32_t got = 0;
    detail::check(gt_primes_near_x(n, x, out.data(), &got), "get_n_primes_near_x");
    out.resize(got);
    return out;
}

namespace detail {

// Owner of one device storage handle plus the host mirror that
// get_raw_tables() hands out.
class DeviceStorage {
  public:
    DeviceStorage(int kind, const std::vector<uint64_t>& sizes, uint64_t capacity_hint = 0) : _kind(kind) {
        check(gt_storage_new(kind, sizes.data(), (uint32_t)sizes.size(), capacity_hint, &_h), "gt_storage_new");
    }
    explicit DeviceStorage(gt_storage* h, int kind) : _h(h), _kind(kind) {}
    DeviceStorage(const DeviceStorage&) = delete;
    DeviceStorage& operator=(const DeviceStorage&) = delete;
    virtual ~DeviceStorage() {
        if (_h) gt_storage_free(_h);
    }

    gt_storage* handle() const { return _h; }

    uint32_t n_tables_() const {
        uint32_t n = 0;
        check(gt_storage_n_tables(_h, &n), "n_tables");
        return n;
    }

    std::vector<uint64_t> tablesizes_() const {
        std::vector<uint64_t> v(n_tables_());
        if (!v.empty()) check(gt_storage_tablesizes(_h, v.data()), "get_tablesizes");
        return v;
    }

    std::pair<uint64_t, uint64_t> counters_() const {
        uint64_t u = 0, o = 0;
        check(gt_storage_counters(_h, &u, &o), "counters");
        return {u, o};
    }

    bool insert_(uint64_t h) {
        uint8_t nw = 0;
        check(gt_insert_hashes(_h, &h, 1, &nw), "insert");
        return nw != 0;
    }

    count_t query_(uint64_t h) const {
        count_t c = 0;
        check(gt_query_hashes(_h, &h, 1, &c), "query");
        return c;
    }

    count_t insert_and_query_(uint64_t h) {
        count_t c = 0;
        check(gt_insert_and_query_hashes(_h, &h, 1, &c), "insert_and_query");
        return c;
    }

    byte_t** raw_tables_() {
        const uint32_t n = n_tables_();
        _mirror.resize(n);
        _mirror_ptrs.resize(n);
        for (uint32_t i = 0; i < n; ++i) {
            uint64_t nb = 0;
            check(gt_storage_table_bytes(_h, i, &nb), "table_bytes");
            _mirror[i].resize(nb);
            check(gt_storage_read_table(_h, i, _mirror[i].data()), "get_raw_tables");
            _mirror_ptrs[i] = _mirror[i].data();
        }
        return _mirror_ptrs.data();
    }

    void reset_() { check(gt_storage_reset(_h), "reset"); }

  protected:
    gt_storage* _h = nullptr;
    int _kind;
    std::vector<std::vector<byte_t>> _mirror;
    std::vector<byte_t*> _mirror_ptrs;
};

}  // namespace detail
}  // namespace goetia

#endif
