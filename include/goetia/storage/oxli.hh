/**
 * oxli.hh -- khmer "OXLI" v4 table files (bitstorage.cc:151-307,
 * bytestorage.cc:480-536): signature, version 4, table type, [use_bigcount],
 * ksize u32, n_tables u8, occupied u64, then per table its size u64 and
 * bytes; counting files end with a bigcount count (always 0 here).
 */
#ifndef GOETIA_AMD_OXLI_HH
#define GOETIA_AMD_OXLI_HH

#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "goetia/storage/storage.hh"

namespace goetia {
namespace oxli {

constexpr unsigned char kCounting = 1;   // SAVED_COUNTING_HT, storage.hh:65
constexpr unsigned char kHashbits = 2;   // SAVED_HASHBITS, storage.hh:66

template <class S>
void save(S& store, const std::string& filename, uint16_t ksize, unsigned char type) {
    std::ofstream out(filename, std::ios::binary);
    if (!out) throw GoetiaFileException("cannot open " + filename);
    out.write("OXLI", 4);
    const unsigned char version = 4;
    out.write((const char*)&version, 1);
    out.write((const char*)&type, 1);
    if (type == kCounting) {
        const unsigned char bigcount = 0;
        out.write((const char*)&bigcount, 1);
    }
    const uint32_t k = ksize;
    const auto sizes = store.get_tablesizes();
    const unsigned char nt = (unsigned char)sizes.size();
    const uint64_t occ = store.n_occupied();
    out.write((const char*)&k, 4);
    out.write((const char*)&nt, 1);
    out.write((const char*)&occ, 8);
    byte_t** t = store.get_raw_tables();
    for (size_t i = 0; i < sizes.size(); ++i) {
        const uint64_t sz = sizes[i];
        out.write((const char*)&sz, 8);
        out.write((const char*)t[i], type == kHashbits ? sz / 8 + 1 : sz);
    }
    if (type == kCounting) {
        const uint64_t n_big = 0;
        out.write((const char*)&n_big, 8);
    }
    if (!out) throw GoetiaFileException("write failed: " + filename);
}

template <class S>
uint16_t load(S& store, const std::string& filename, unsigned char type) {
    std::ifstream in(filename, std::ios::binary);
    if (!in) throw GoetiaFileException("cannot open " + filename);
    char sig[4];
    unsigned char version = 0, ftype = 0, bigcount = 0, nt = 0;
    uint32_t k = 0;
    uint64_t occ = 0;
    in.read(sig, 4);
    in.read((char*)&version, 1);
    in.read((char*)&ftype, 1);
    if (!in || std::memcmp(sig, "OXLI", 4) != 0) throw GoetiaFileException(filename + ": not an OXLI file");
    if (version != 4) throw GoetiaFileException(filename + ": unsupported OXLI version");
    if (ftype != type) throw GoetiaFileException(filename + ": wrong OXLI table type");
    if (type == kCounting) in.read((char*)&bigcount, 1);
    in.read((char*)&k, 4);
    in.read((char*)&nt, 1);
    in.read((char*)&occ, 8);
    const auto sizes = store.get_tablesizes();
    if (nt != sizes.size()) throw GoetiaFileException(filename + ": table count differs");
    for (size_t i = 0; i < sizes.size(); ++i) {
        uint64_t sz = 0;
        in.read((char*)&sz, 8);
        if (sz != sizes[i]) throw GoetiaFileException(filename + ": table size differs");
        std::vector<byte_t> buf(type == kHashbits ? sz / 8 + 1 : sz);
        in.read((char*)buf.data(), (std::streamsize)buf.size());
        if (!in) throw GoetiaFileException(filename + ": truncated");
        detail::check(gt_storage_write_table(store.handle(), (uint32_t)i, buf.data()), "load");
    }
    detail::check(gt_storage_set_counters(store.handle(), store.n_unique_kmers(), occ), "load");
    return (uint16_t)k;
}

}  // namespace oxli
}  // namespace goetia

#endif
