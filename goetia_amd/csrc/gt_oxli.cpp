// gt_oxli.cpp -- khmer "OXLI" v4 table files for Bloom and count-min storages,
// byte for byte as goetia writes and reads them:
//   BitStorage::save / load            bitstorage.cc:151-307 (always plain)
//   ByteStorage::save / load           bytestorage.cc:638-667, which pick
//     ByteStorageFile{Writer,Reader}    bytestorage.cc:480-536 / 170-302 (plain)
//     ByteStorageGz{FileWriter,FileReader} bytestorage.cc:538-636 / 304-478
//   by the file name: gzip when the text after its last '.' is "gz"
//   (bytestorage.cc:153-167).
// Layout (little-endian): "OXLI" | version u8 = 4 | type u8 (2 hashbits,
// 1 counting) | [counting: use_bigcount u8] | ksize u32 | n_tables u8 |
// occupied_bins u64 | per table: size u64, then size/8+1 (Bloom) or size
// (count-min) bytes | [counting: n_bigcounts u64, then (k-mer u64, count i16)
// pairs].  Count-min storages write their bigcount flag and map (pairs in
// k-mer order; goetia writes its unordered_map's order) and load both.
//
// Both formats go through one sink / source interface; the gzip one uses
// zlib's gzip file functions with the same open modes goetia uses ("wb",
// "rb"), so a compressed file written here matches goetia's byte for byte.

#include <zlib.h>

#include <climits>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "goetia_amd.h"

namespace gt {
int set_error(int code, const std::string& msg);
// Replaces a Bloom / count-min storage's tables with zeroed tables of the
// given sizes (gt_capi.cpp).
int storage_reshape(gt_storage* st, const uint64_t* sizes, uint32_t n_tables);
constexpr int kMaxTables = 16;
}  // namespace gt

#define GT_TRY(call)             \
    do {                         \
        const int rc_ = (call);  \
        if (rc_ != GT_OK) return rc_; \
    } while (0)

namespace {

constexpr unsigned char kCounting = 1;   // SAVED_COUNTING_HT, storage.hh:65
constexpr unsigned char kHashbits = 2;   // SAVED_HASHBITS, storage.hh:66
constexpr unsigned char kVersion = 4;    // SAVED_FORMAT_VERSION, storage.hh:64

bool gz_name(const std::string& path) {
    const size_t dot = path.find_last_of('.');
    return path.substr(dot == std::string::npos ? 0 : dot + 1) == "gz";
}

struct Sink {
    virtual ~Sink() = default;
    virtual bool put(const void* p, uint64_t n) = 0;
    virtual bool close() = 0;
};

struct FileSink : Sink {
    FILE* f;
    explicit FileSink(FILE* f_) : f(f_) {}
    bool put(const void* p, uint64_t n) override { return n == 0 || fwrite(p, 1, n, f) == n; }
    bool close() override {
        const bool ok = fflush(f) == 0 && !ferror(f);
        return fclose(f) == 0 && ok;
    }
};

struct GzSink : Sink {
    gzFile f;
    explicit GzSink(gzFile f_) : f(f_) {}
    bool put(const void* p, uint64_t n) override {
        const char* c = static_cast<const char*>(p);
        while (n) {   // zlib takes at most INT_MAX bytes per call
            const unsigned chunk = n > (uint64_t)INT_MAX ? (unsigned)INT_MAX : (unsigned)n;
            const int w = gzwrite(f, c, chunk);
            if (w <= 0) return false;
            c += w;
            n -= (uint64_t)w;
        }
        return true;
    }
    bool close() override {
        int err = 0;
        gzerror(f, &err);
        return gzclose(f) == Z_OK && err == Z_OK;
    }
};

struct Source {
    virtual ~Source() = default;
    virtual bool get(void* p, uint64_t n) = 0;
};

struct FileSource : Source {
    FILE* f;
    explicit FileSource(FILE* f_) : f(f_) {}
    ~FileSource() override { fclose(f); }
    bool get(void* p, uint64_t n) override { return n == 0 || fread(p, 1, n, f) == n; }
};

struct GzSource : Source {
    gzFile f;
    explicit GzSource(gzFile f_) : f(f_) {}
    ~GzSource() override { gzclose(f); }
    bool get(void* p, uint64_t n) override {
        char* c = static_cast<char*>(p);
        while (n) {
            const unsigned chunk = n > (uint64_t)INT_MAX ? (unsigned)INT_MAX : (unsigned)n;
            const int r = gzread(f, c, chunk);
            if (r <= 0) return false;
            c += r;
            n -= (uint64_t)r;
        }
        return true;
    }
};

int io_fail(const std::string& msg) { return gt::set_error(GT_ERR_IO, msg); }

// Gzip only for count-min files named *.gz (BitStorage::save ignores the name).
bool use_gz(int kind, const std::string& path) { return kind == GT_BYTE_STORAGE && gz_name(path); }

std::unique_ptr<Source> open_source(int kind, const std::string& path) {
    if (use_gz(kind, path)) {
        gzFile f = gzopen(path.c_str(), "rb");
        return f ? std::unique_ptr<Source>(new GzSource(f)) : nullptr;
    }
    FILE* f = fopen(path.c_str(), "rb");
    return f ? std::unique_ptr<Source>(new FileSource(f)) : nullptr;
}

struct Header {
    uint16_t ksize = 0;
    uint32_t n_tables = 0;
    uint64_t sizes[gt::kMaxTables] = {};
    uint64_t occupied = 0;
    unsigned char use_bigcount = 0;
    std::vector<uint64_t> big_keys;   // the bigcount pairs (read with the tables)
    std::vector<int16_t> big_vals;
};

uint64_t table_bytes(int kind, uint64_t size) { return kind == GT_BIT_STORAGE ? size / 8 + 1 : size; }

// Reads the header and each table's size; leaves `src` at table 0's bytes
// when tables_out is null, else reads the tables too and checks the bigcount
// trailer of counting files.
int read_file(int kind, const std::string& path, Header* h, uint8_t* const* tables_out) {
    if (kind != GT_BIT_STORAGE && kind != GT_BYTE_STORAGE)
        return gt::set_error(GT_ERR_INVALID_ARG, "OXLI files hold Bloom and count-min tables");
    auto src = open_source(kind, path);
    if (!src) return io_fail("Cannot open k-mer " + std::string(kind == GT_BIT_STORAGE ? "graph" : "count") +
                             " file: " + path);
    char sig[4];
    unsigned char version = 0, type = 0;
    if (!src->get(sig, 4) || !src->get(&version, 1) || !src->get(&type, 1))
        return io_fail("Unexpected end of k-mer file: " + path);
    if (std::memcmp(sig, "OXLI", 4) != 0) return io_fail("Does not start with signature for a oxli file: " + path);
    if (version != kVersion)
        return io_fail("Incorrect file format version " + std::to_string(version) + " while reading " + path +
                       "; should be 4");
    const unsigned char want = kind == GT_BIT_STORAGE ? kHashbits : kCounting;
    if (type != want) return io_fail("Incorrect file format type " + std::to_string(type) + " while reading " + path);
    if (type == kCounting && !src->get(&h->use_bigcount, 1)) return io_fail("K-mer count file header read error: " + path);
    uint32_t k = 0;
    unsigned char nt = 0;
    if (!src->get(&k, 4) || !src->get(&nt, 1) || !src->get(&h->occupied, 8))
        return io_fail("K-mer file header read error: " + path);
    h->ksize = (uint16_t)k;
    h->n_tables = nt;
    if (nt == 0 || nt > gt::kMaxTables) return io_fail(path + ": " + std::to_string(nt) + " tables (1..16 supported)");
    std::vector<uint8_t> skip;
    for (uint32_t i = 0; i < nt; ++i) {
        if (!src->get(&h->sizes[i], 8)) return io_fail("K-mer file header read error: " + path);
        if (h->sizes[i] == 0 || h->sizes[i] >= (1ULL << 32)) return io_fail(path + ": table size out of range");
        const uint64_t nb = table_bytes(kind, h->sizes[i]);
        uint8_t* dst = nullptr;
        if (tables_out) {
            dst = tables_out[i];
        } else {
            skip.resize(nb);
            dst = skip.data();
        }
        if (!src->get(dst, nb)) return io_fail("Unexpected end of k-mer file: " + path);
    }
    if (type == kCounting && tables_out) {
        // ByteStorageFileReader (bytestorage.cc:271-285): a pair count, then
        // (k-mer, count) pairs
        uint64_t n_big = 0;
        if (!src->get(&n_big, 8)) return io_fail("K-mer count header read error: " + path);
        h->big_keys.clear();
        h->big_vals.clear();
        for (uint64_t j = 0; j < n_big; ++j) {
            uint64_t key = 0;
            int16_t val = 0;
            if (!src->get(&key, 8) || !src->get(&val, 2)) return io_fail("Unexpected end of k-mer count file: " + path);
            h->big_keys.push_back(key);
            h->big_vals.push_back(val);
        }
    }
    return GT_OK;
}

// Writes one OXLI file; counting files carry use_bigcount and the pairs.
int write_file(const char* path, int kind, uint16_t ksize, uint32_t n_tables, const uint64_t* tablesizes,
               uint64_t occupied, const uint8_t* const* tables, unsigned char use_bigcount, uint64_t n_big,
               const uint64_t* big_keys, const int16_t* big_vals) {
    std::unique_ptr<Sink> out;
    if (use_gz(kind, path)) {
        gzFile f = gzopen(path, "wb");
        if (!f) return io_fail(std::string("cannot open ") + path);
        out.reset(new GzSink(f));
    } else {
        FILE* f = fopen(path, "wb");
        if (!f) return io_fail(std::string("cannot open ") + path + ": " + strerror(errno));
        out.reset(new FileSink(f));
    }
    const unsigned char version = kVersion, type = kind == GT_BIT_STORAGE ? kHashbits : kCounting;
    const uint32_t k = ksize;
    const unsigned char nt = (unsigned char)n_tables;
    bool ok = out->put("OXLI", 4) && out->put(&version, 1) && out->put(&type, 1);
    if (type == kCounting) ok = ok && out->put(&use_bigcount, 1);
    ok = ok && out->put(&k, 4) && out->put(&nt, 1) && out->put(&occupied, 8);
    for (uint32_t i = 0; ok && i < n_tables; ++i)
        ok = out->put(&tablesizes[i], 8) && out->put(tables[i], table_bytes(kind, tablesizes[i]));
    if (type == kCounting) {
        ok = ok && out->put(&n_big, 8);
        for (uint64_t j = 0; ok && j < n_big; ++j) ok = out->put(&big_keys[j], 8) && out->put(&big_vals[j], 2);
    }
    const bool closed = out->close();
    if (!ok || !closed) return io_fail(std::string("write failed: ") + path);
    return GT_OK;
}

}  // namespace

extern "C" {

int gt_oxli_write(const char* path, int kind, uint16_t ksize, uint32_t n_tables, const uint64_t* tablesizes,
                  uint64_t occupied, const uint8_t* const* tables) {
    if (!path || !tablesizes || !tables || n_tables == 0 || n_tables > (uint32_t)gt::kMaxTables)
        return gt::set_error(GT_ERR_INVALID_ARG, "invalid arguments");
    if (kind != GT_BIT_STORAGE && kind != GT_BYTE_STORAGE)
        return gt::set_error(GT_ERR_INVALID_ARG, "OXLI files hold Bloom and count-min tables");
    return write_file(path, kind, ksize, n_tables, tablesizes, occupied, tables, 0, 0, nullptr, nullptr);
}

int gt_oxli_read_header(const char* path, int kind, uint16_t* ksize, uint32_t* n_tables, uint64_t* tablesizes,
                        uint64_t* occupied) {
    if (!path) return gt::set_error(GT_ERR_INVALID_ARG, "null path");
    Header h;
    GT_TRY(read_file(kind, path, &h, nullptr));
    if (ksize) *ksize = h.ksize;
    if (n_tables) *n_tables = h.n_tables;
    if (tablesizes)
        for (uint32_t i = 0; i < h.n_tables; ++i) tablesizes[i] = h.sizes[i];
    if (occupied) *occupied = h.occupied;
    return GT_OK;
}

int gt_oxli_read(const char* path, int kind, uint8_t* const* tables) {
    if (!path || !tables) return gt::set_error(GT_ERR_INVALID_ARG, "invalid arguments");
    Header h;
    return read_file(kind, path, &h, tables);
}

int gt_storage_save(gt_storage* st, const char* path, uint16_t ksize) {
    if (!st || !path) return gt::set_error(GT_ERR_INVALID_ARG, "invalid arguments");
    int kind = 0;
    GT_TRY(gt_storage_kind(st, &kind));
    if (kind == GT_EXACT_STORAGE)
        return gt::set_error(GT_ERR_UNSUPPORTED, "the exact set has no file format (sparseppstorage.cc:52-58)");
    uint32_t nt = 0;
    GT_TRY(gt_storage_n_tables(st, &nt));
    std::vector<uint64_t> sizes(nt);
    GT_TRY(gt_storage_tablesizes(st, sizes.data()));
    uint64_t n_unique = 0, occ = 0;
    GT_TRY(gt_storage_counters(st, &n_unique, &occ));
    std::vector<std::vector<uint8_t>> host(nt);
    std::vector<const uint8_t*> ptrs(nt);
    for (uint32_t i = 0; i < nt; ++i) {
        host[i].resize(table_bytes(kind, sizes[i]));
        GT_TRY(gt_storage_read_table(st, i, host[i].data()));
        ptrs[i] = host[i].data();
    }
    int use_big = 0;
    std::vector<uint64_t> bk;
    std::vector<int16_t> bv;
    if (kind == GT_BYTE_STORAGE) {
        GT_TRY(gt_storage_get_use_bigcount(st, &use_big));
        uint64_t n_big = 0;
        GT_TRY(gt_storage_bigcounts(st, nullptr, nullptr, 0, &n_big));
        bk.resize(n_big);
        bv.resize(n_big);
        GT_TRY(gt_storage_bigcounts(st, bk.data(), bv.data(), n_big, &n_big));
    }
    return write_file(path, kind, ksize, nt, sizes.data(), occ, ptrs.data(), (unsigned char)(use_big ? 1 : 0),
                      bk.size(), bk.data(), bv.data());
}

int gt_storage_load(gt_storage* st, const char* path, uint16_t* ksize) {
    if (!st || !path) return gt::set_error(GT_ERR_INVALID_ARG, "invalid arguments");
    int kind = 0;
    GT_TRY(gt_storage_kind(st, &kind));
    if (kind == GT_EXACT_STORAGE)
        return gt::set_error(GT_ERR_UNSUPPORTED, "the exact set has no file format (sparseppstorage.cc:52-58)");
    Header h;
    GT_TRY(read_file(kind, path, &h, nullptr));
    std::vector<std::vector<uint8_t>> host(h.n_tables);
    std::vector<uint8_t*> ptrs(h.n_tables);
    for (uint32_t i = 0; i < h.n_tables; ++i) {
        host[i].resize(table_bytes(kind, h.sizes[i]));
        ptrs[i] = host[i].data();
    }
    GT_TRY(read_file(kind, path, &h, ptrs.data()));
    // goetia's loaders replace the tables with the file's, sizes included
    GT_TRY(gt::storage_reshape(st, h.sizes, h.n_tables));
    for (uint32_t i = 0; i < h.n_tables; ++i) GT_TRY(gt_storage_write_table(st, i, ptrs[i]));
    uint64_t n_unique = 0;
    GT_TRY(gt_storage_counters(st, &n_unique, nullptr));
    GT_TRY(gt_storage_set_counters(st, n_unique, h.occupied));
    if (kind == GT_BYTE_STORAGE) {
        // the file's flag; its pairs replace the map only when there are any
        GT_TRY(gt_storage_set_use_bigcount(st, h.use_bigcount ? 1 : 0));
        if (!h.big_keys.empty())
            GT_TRY(gt_storage_set_bigcounts(st, h.big_keys.data(), h.big_vals.data(), h.big_keys.size()));
    }
    if (ksize) *ksize = h.ksize;
    return GT_OK;
}

}  // extern "C"
