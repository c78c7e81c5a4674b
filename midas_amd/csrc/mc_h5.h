// Minimal HDF5 writer for the caked outputs.
//
// Writes the part of the HDF5 file format (HDF Group, "HDF5 File Format
// Specification Version 3.0") that libhdf5 writes by default and that
// readers of the reference's outputs (h5py, libhdf5) accept:
//   * superblock version 0, 8-byte offsets and lengths,
//   * groups as symbol tables: a version-1 B-tree per group (one node up to
//     65536 members, more levels above that), symbol-table nodes and a
//     local heap,
//   * version-1 object headers,
//   * datasets with contiguous layout: IEEE float64, int32, int64, and
//     variable-length UTF-8 strings (global heap collections),
//   * attributes (version-1 messages): float64, int32, fixed-length ASCII
//     strings.
// Dataset bytes are appended to the file when the dataset is added; the
// metadata (headers, heaps, B-trees) and the superblock are written by
// close().  Objects are named by '/'-separated paths; missing parent groups
// are created.  Not thread-safe: one writer per thread.
#pragma once

#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace mc {
namespace h5 {

enum class Kind { F64, I32, I64, STR, VSTR };

class Writer {
 public:
  Writer();
  ~Writer();
  Writer(const Writer &) = delete;
  Writer &operator=(const Writer &) = delete;

  // 0 on success; otherwise error() says why.
  int open(const std::string &path);
  int group(const std::string &path);
  // Numeric dataset (F64, I32 or I64) of the given shape (empty = scalar).
  int dataset(const std::string &path, Kind kind, const std::vector<uint64_t> &dims,
              const void *data);
  int vstr_dataset(const std::string &path, const std::vector<std::string> &values);
  int attr_f64(const std::string &path, const std::string &name, const double *v, int n);
  int attr_i32(const std::string &path, const std::string &name, const int32_t *v, int n);
  int attr_str(const std::string &path, const std::string &name, const std::string &value);
  int close();
  bool is_open() const { return f_ != nullptr; }
  const std::string &error() const { return err_; }

 private:
  struct Attr {
    std::string name;
    Kind kind;
    int n;                    // elements (strings: 1)
    std::vector<uint8_t> raw; // element bytes
  };
  struct Node {
    bool is_group = true;
    std::map<std::string, std::unique_ptr<Node>> kids;  // byte-wise name order
    std::vector<Attr> attrs;
    Kind kind = Kind::F64;
    std::vector<uint64_t> dims;
    uint64_t data_addr = 0, data_size = 0;
    uint64_t oh_addr = 0, btree_addr = 0, heap_addr = 0;  // after emit
  };

  int fail(const std::string &msg);
  Node *find(const std::string &path, bool create_groups, bool want_group, bool create_leaf);
  int write_at(uint64_t addr, const void *p, size_t n);
  uint64_t append(const void *p, size_t n);
  int emit(Node &n);

  FILE *f_ = nullptr;
  uint64_t eof_ = 0;
  std::string err_;
  std::unique_ptr<Node> root_;
};

}  // namespace h5
}  // namespace mc
