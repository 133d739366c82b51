/**
 * bytestorage.hh -- ByteStorage (goetia bytestorage.hh:96-243): a count-min
 * sketch of saturating 8-bit counters (MAX_KCOUNT 255) in HBM, with goetia's
 * optional bigcount map (counts past 255, int16) on the device.
 */
#ifndef GOETIA_AMD_BYTESTORAGE_HH
#define GOETIA_AMD_BYTESTORAGE_HH

#include "goetia/storage/oxli.hh"
#include "goetia/storage/storage.hh"

#define MAX_KCOUNT 255

namespace goetia {

class ByteStorage;

template <>
struct StorageTraits<ByteStorage> {
    static constexpr bool is_probabilistic = true;
    static constexpr bool is_counting = true;
    static constexpr int bits_per_bin = 8;
    typedef std::tuple<uint64_t, uint16_t> params_type;
    static constexpr params_type default_params = std::make_tuple(1'000'000, 4);
};

class ByteStorage : public Storage<uint64_t>, public detail::DeviceStorage {
  public:
    using Traits = StorageTraits<ByteStorage>;
    typedef uint64_t value_type;

    ByteStorage(uint64_t max_table, uint16_t N) : ByteStorage(get_n_primes_near_x(N, max_table)) {}
    explicit ByteStorage(const std::vector<uint64_t>& tablesizes)
        : detail::DeviceStorage(GT_BYTE_STORAGE, tablesizes) {}

    static std::shared_ptr<ByteStorage> build(uint64_t max_table, uint16_t N) {
        return std::make_shared<ByteStorage>(max_table, N);
    }
    static std::shared_ptr<ByteStorage> build(Traits::params_type p) {
        return build(std::get<0>(p), std::get<1>(p));
    }
    std::shared_ptr<ByteStorage> clone() const { return std::make_shared<ByteStorage>(get_tablesizes()); }

    std::vector<uint64_t> get_tablesizes() const { return tablesizes_(); }
    const size_t n_tables() const { return n_tables_(); }
    const uint64_t n_occupied() const override { return counters_().second; }
    const uint64_t n_unique_kmers() const override { return counters_().first; }
    double estimated_fp() {
        const auto s = get_tablesizes();
        return std::pow((double)n_occupied() / (double)s[0], (double)s.size());
    }
    const bool insert(value_type h) override { return insert_(h); }
    const count_t insert_and_query(value_type h) override { return insert_and_query_(h); }
    const count_t query(value_type h) const override { return query_(h); }
    byte_t** get_raw_tables() override { return raw_tables_(); }
    void reset() override { reset_(); }
    void set_use_bigcount(bool b) { detail::check(gt_storage_set_use_bigcount(handle(), b ? 1 : 0), "set_use_bigcount"); }
    bool get_use_bigcount() {
        int v = 0;
        detail::check(gt_storage_get_use_bigcount(handle(), &v), "get_use_bigcount");
        return v != 0;
    }
    void save(std::string filename, uint16_t ksize) override { oxli::save(*this, filename, ksize, oxli::kCounting); }
    void load(std::string filename, uint16_t& ksize) override { ksize = oxli::load(*this, filename, oxli::kCounting); }
};

}  // namespace goetia

#endif
