#ifndef GOETIA_AMD_STORAGE_TYPES_HH
#define GOETIA_AMD_STORAGE_TYPES_HH

#include "goetia/storage/bitstorage.hh"
#include "goetia/storage/bytestorage.hh"
#include "goetia/storage/sparseppstorage.hh"

#endif
