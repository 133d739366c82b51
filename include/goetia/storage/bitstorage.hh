/**
 * bitstorage.hh -- BitStorage (goetia bitstorage.hh:93-243): a Bloom filter
 * of n_tables bit tables of prime sizes, held in HBM.  get_raw_tables()
 * returns a host copy in the reference layout (p/8+1 bytes per table).
 */
#ifndef GOETIA_AMD_BITSTORAGE_HH
#define GOETIA_AMD_BITSTORAGE_HH

#include "goetia/storage/oxli.hh"
#include "goetia/storage/storage.hh"

namespace goetia {

class BitStorage;

template <>
struct StorageTraits<BitStorage> {
    static constexpr bool is_probabilistic = true;
    static constexpr bool is_counting = false;
    static constexpr int bits_per_slot = 1;
    typedef std::tuple<uint64_t, uint16_t> params_type;
    static constexpr params_type default_params = std::make_tuple(1'000'000, 4);
};

class BitStorage : public Storage<uint64_t>, public detail::DeviceStorage {
  public:
    using Traits = StorageTraits<BitStorage>;
    using Storage<uint64_t>::value_type;

    BitStorage(uint64_t max_table, uint16_t N) : BitStorage(get_n_primes_near_x(N, max_table)) {}
    explicit BitStorage(const std::vector<uint64_t>& tablesizes)
        : detail::DeviceStorage(GT_BIT_STORAGE, tablesizes) {}

    static std::shared_ptr<BitStorage> build(uint64_t max_table, uint16_t N) {
        return std::make_shared<BitStorage>(max_table, N);
    }
    static std::shared_ptr<BitStorage> build(const Traits::params_type& p) {
        return build(std::get<0>(p), std::get<1>(p));
    }
    std::shared_ptr<BitStorage> clone() const { return std::make_shared<BitStorage>(get_tablesizes()); }

    std::vector<uint64_t> get_tablesizes() const { return tablesizes_(); }
    const size_t n_tables() const { return n_tables_(); }
    const uint64_t n_occupied() const override { return counters_().second; }
    const uint64_t n_unique_kmers() const override { return counters_().first; }
    double estimated_fp() {
        const auto s = get_tablesizes();
        return std::pow((double)n_occupied() / (double)s[0], (double)s.size());
    }
    const bool insert(value_type h) override { return insert_(h); }
    const count_t insert_and_query(value_type h) override {
        insert_(h);
        return 1;   // bitstorage.cc:78-84
    }
    const count_t query(value_type h) const override { return query_(h); }
    byte_t** get_raw_tables() override { return raw_tables_(); }
    void reset() override { reset_(); }
    void update_from(const BitStorage& other) { detail::check(gt_storage_update_from(_h, other._h), "update_from"); }
    void save(std::string filename, uint16_t ksize) override { oxli::save(*this, filename, ksize, oxli::kHashbits); }
    void load(std::string filename, uint16_t& ksize) override { ksize = oxli::load(*this, filename, oxli::kHashbits); }
};

}  // namespace goetia

#endif
