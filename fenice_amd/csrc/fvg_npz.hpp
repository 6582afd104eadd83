// .npz containers (numpy's format: a zip archive of .npy entries), host only. The run checkpoints of the C ABI
// (fvg_checkpoint_*) are such files, with the keys and dtypes the reference's restart files carry
// (cpp/src/SimulationRestart.cpp:61-85, IO.cpp npz writer/reader), so a checkpoint passes between the reference and
// this library in either direction. Entries are stored (not compressed); sizes and offsets past 32 bits use the
// zip64 extra fields.
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace fvg {
namespace npz {

struct Array {
    std::string descr;                  // numpy dtype: "<f8", "<i8" or "<U<width>"
    std::vector<int64_t> shape;         // C order
    std::vector<uint8_t> store;         // the data when the array owns it
    const uint8_t* data = nullptr;      // the little-endian data (into store, or a caller's buffer)
    uint64_t size = 0;                  // bytes of data

    Array() = default;
    Array(Array&&) = default;
    Array& operator=(Array&&) = default;
    Array(const Array& o) : descr(o.descr), shape(o.shape), store(o.store), data(o.data), size(o.size) {
        if (o.data == o.store.data()) data = store.data();   // an owned copy points at its own store
    }
    Array& operator=(const Array& o) {
        if (this != &o) *this = Array(o);
        return *this;
    }
    uint64_t elements() const;
};

Array scalar_f8(double v);
Array scalar_i8(int64_t v);
Array vector_f8(const std::vector<double>& v);
// a view of n doubles the caller keeps alive until the archive is written
Array view_f8(const double* v, const std::vector<int64_t>& shape);
Array strings(const std::vector<std::string>& v);   // "<U" of the longest string (at least 1), UTF-32 code units
// the value of a one-element "<f8" or "<i8" array
double as_double(const Array& a);

// CRC-32 of the zip format (polynomial 0xEDB88320, reflected), continued from `crc` (0 to start)
uint32_t crc32(const uint8_t* p, uint64_t n, uint32_t crc = 0);

// Writes the arrays as `path` (entries "<name>.npy" in name order), through a temporary file renamed over `path`
// when complete. Entries and offsets from zip64_from bytes on carry zip64 fields. Throws std::runtime_error.
void write(const std::string& path, const std::map<std::string, Array>& arrays, uint64_t zip64_from = 0xFFFFFFFFull);

// Reads every entry of a .npz of stored entries; checks each entry's CRC-32. Throws std::runtime_error.
std::map<std::string, Array> read(const std::string& path);

}   // namespace npz
}   // namespace fvg
