/**
 * oxli.hh -- khmer "OXLI" v4 table files, as goetia writes and reads them
 * (BitStorage::save / load bitstorage.cc:151-307; ByteStorage::save / load
 * through ByteStorageFile, bytestorage.cc:153-667, gzip when a count-min
 * file's name ends in ".gz").  The library writes and reads the bytes
 * (gt_storage_save / gt_storage_load); a load takes the file's table sizes
 * and restores n_occupied.  Failures throw GoetiaFileException, bigcount
 * files GoetiaException.
 */
#ifndef GOETIA_AMD_OXLI_HH
#define GOETIA_AMD_OXLI_HH

#include <string>

#include "goetia/storage/storage.hh"

namespace goetia {
namespace oxli {

constexpr unsigned char kCounting = 1;   // SAVED_COUNTING_HT, storage.hh:65
constexpr unsigned char kHashbits = 2;   // SAVED_HASHBITS, storage.hh:66

template <class S>
void save(S& store, const std::string& filename, uint16_t ksize, unsigned char) {
    detail::check(gt_storage_save(store.handle(), filename.c_str(), ksize), "save");
}

template <class S>
uint16_t load(S& store, const std::string& filename, unsigned char) {
    uint16_t k = 0;
    detail::check(gt_storage_load(store.handle(), filename.c_str(), &k), "load");
    return k;
}

}  // namespace oxli
}  // namespace goetia

#endif
