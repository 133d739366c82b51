/**
 * sparseppstorage.hh -- the exact k-mer set (goetia sparseppstorage.hh:38-108,
 * with the same semantics as PHMapStorage and BTreeStorage): a GPU
 * open-addressing table of 64-bit keys that grows as needed.
 */
#ifndef GOETIA_AMD_SPARSEPPSTORAGE_HH
#define GOETIA_AMD_SPARSEPPSTORAGE_HH

#include <algorithm>

#include "goetia/storage/storage.hh"

namespace goetia {

class SparseppSetStorage;

template <>
struct StorageTraits<SparseppSetStorage> {
    static constexpr bool is_probabilistic = false;
    static constexpr bool is_counting = false;
    typedef std::tuple<bool> params_type;
    static constexpr params_type default_params = std::make_tuple(0);
};

class SparseppSetStorage : public Storage<uint64_t>, public detail::DeviceStorage {
  public:
    using Traits = StorageTraits<SparseppSetStorage>;
    using Storage<uint64_t>::value_type;

    explicit SparseppSetStorage(uint64_t capacity_hint = 0)
        : detail::DeviceStorage(GT_EXACT_STORAGE, {}, capacity_hint) {}

    static std::shared_ptr<SparseppSetStorage> build() { return std::make_shared<SparseppSetStorage>(); }
    static std::shared_ptr<SparseppSetStorage> build(const Traits::params_type&) { return build(); }
    std::shared_ptr<SparseppSetStorage> clone() const { return std::make_shared<SparseppSetStorage>(); }

    const uint64_t n_unique_kmers() const override { return counters_().first; }
    const uint64_t n_buckets() const { return counters_().second; }
    const uint64_t n_occupied() const override { return n_buckets(); }
    const bool insert(value_type h) override { return insert_(h); }
    const count_t insert_and_query(value_type h) override {
        insert_(h);
        return 1;   // sparseppstorage.cc:21-25
    }
    const count_t query(value_type h) const override { return query_(h); }
    byte_t** get_raw_tables() override { return nullptr; }
    void reset() override { reset_(); }
    void save(std::string, uint16_t) override {
        throw GoetiaException("the device exact set is not serialised to sparsepp files");
    }
    void load(std::string, uint16_t&) override {
        throw GoetiaException("the device exact set is not serialised to sparsepp files");
    }

    // Set contents, sorted.
    std::vector<uint64_t> keys() const {
        uint64_t n = 0;
        detail::check(gt_storage_exact_keys(_h, nullptr, 0, &n), "exact_keys");
        std::vector<uint64_t> out(n);
        detail::check(gt_storage_exact_keys(_h, out.data(), n, &n), "exact_keys");
        std::sort(out.begin(), out.end());
        return out;
    }
};

// phmapstorage.hh:43,90-95 and btreestorage.hh:43,90-95: the same set.
// PHMapStorage reports sparsepp's bucket count as n_occupied, not that of
// phmap's parallel set.
class PHMapStorage : public SparseppSetStorage {
  public:
    using SparseppSetStorage::SparseppSetStorage;
    static std::shared_ptr<PHMapStorage> build() { return std::make_shared<PHMapStorage>(); }
};

class BTreeStorage : public SparseppSetStorage {
  public:
    using SparseppSetStorage::SparseppSetStorage;
    static std::shared_ptr<BTreeStorage> build() { return std::make_shared<BTreeStorage>(); }
    // btreestorage.hh:78-84: the btree's max_size() (btree.h:1872), a
    // constant: the largest signed 64-bit value in goetia's build.
    const uint64_t n_buckets() const { return 0x7FFFFFFFFFFFFFFFULL; }
    const uint64_t n_occupied() const override { return n_buckets(); }
};

template <>
struct StorageTraits<PHMapStorage> : StorageTraits<SparseppSetStorage> {};
template <>
struct StorageTraits<BTreeStorage> : StorageTraits<SparseppSetStorage> {};

}  // namespace goetia

#endif
