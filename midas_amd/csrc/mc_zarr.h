// Zarr (format 2) store in a zip file, for the caked outputs.
//
// The reference turns <name>.caked.hdf into <name>.caked.hdf.zarr.zip
// (FF_HEDM/workflows/integrator.py:379-426 with Hdf5ToZarr,
// utils/converters/midas2zip.py:31-93): every HDF5 group becomes a Zarr
// group, every dataset a Zarr array of one chunk, every object's attributes
// its .zattrs (size-1 values as scalars).  This writer builds the same
// store directly, from the same calls as the HDF5 writer: arrays carry the
// reference's compressor entry (Blosc, zstd, level 3, bit shuffle) and
// their chunk is a Blosc1 chunk stored without compression (the "memcpyed"
// form every Blosc decoder reads).  Zip entries are stored uncompressed;
// ZIP64 records are added when the archive passes 4 GiB or 65535 entries.
// Not thread-safe: one writer per thread.
#pragma once

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "mc_h5.h"

namespace mc {
namespace zarr {

class Writer {
 public:
  Writer() = default;
  ~Writer();
  Writer(const Writer &) = delete;
  Writer &operator=(const Writer &) = delete;

  // 0 on success; otherwise error() says why.
  int open(const std::string &path);
  int group(const std::string &path);
  // Numeric array (F64, I32 or I64) of the given shape (empty = scalar).
  int dataset(const std::string &path, h5::Kind kind, const std::vector<uint64_t> &dims,
              const void *data);
  int attr_f64(const std::string &path, const std::string &name, const double *v, int n);
  int attr_i32(const std::string &path, const std::string &name, const int32_t *v, int n);
  int attr_str(const std::string &path, const std::string &name, const std::string &value);
  // Writes the groups' and arrays' metadata and the zip directory.
  int close();
  bool is_open() const { return f_ != nullptr; }
  const std::string &error() const { return err_; }

 private:
  struct Entry {
    std::string name;
    uint32_t crc = 0;
    uint64_t size = 0, offset = 0;
  };
  struct Node {
    bool is_group = true;
    std::vector<std::pair<std::string, std::string>> attrs;  // name, JSON value
  };
  int fail(const std::string &msg);
  Node *node(const std::string &path, bool want_group);
  int put(const std::string &name, const void *p, size_t n);
  int put_json(const std::string &name, const std::string &s) {
    return put(name, s.data(), s.size());
  }
  int add_attr(const std::string &path, const std::string &name, std::string json);

  FILE *f_ = nullptr;
  uint64_t pos_ = 0;
  std::string err_;
  std::vector<Entry> entries_;
  std::map<std::string, Node> nodes_;  // key: path without the leading '/', "" = root
};

// CRC-32 (ISO 3309 / zip) of n bytes, continuing from crc.
uint32_t crc32(uint32_t crc, const void *p, size_t n);
// JSON text of a double as Python's json writes it (shortest round trip;
// NaN, Infinity, -Infinity).
std::string json_number(double v);

}  // namespace zarr
}  // namespace mc
