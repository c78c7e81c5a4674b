// mc_zarr.cpp -- Zarr (format 2) store in a zip file; see mc_zarr.h.
#include "mc_zarr.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace mc {
namespace zarr {

namespace {

// Slice-by-8 tables of the reflected CRC-32 polynomial 0xEDB88320.
struct CrcTables {
  uint32_t t[8][256];
  CrcTables() {
    for (uint32_t i = 0; i < 256; i++) {
      uint32_t c = i;
      for (int k = 0; k < 8; k++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
      t[0][i] = c;
    }
    for (uint32_t i = 0; i < 256; i++)
      for (int s = 1; s < 8; s++) t[s][i] = (t[s - 1][i] >> 8) ^ t[0][t[s - 1][i] & 0xff];
  }
};
const CrcTables &crc_tables() {
  static const CrcTables tables;
  return tables;
}

void le16(std::vector<uint8_t> &b, uint32_t v) {
  b.push_back((uint8_t)v);
  b.push_back((uint8_t)(v >> 8));
}
void le32(std::vector<uint8_t> &b, uint32_t v) {
  for (int k = 0; k < 4; k++) b.push_back((uint8_t)(v >> (8 * k)));
}
void le64(std::vector<uint8_t> &b, uint64_t v) {
  for (int k = 0; k < 8; k++) b.push_back((uint8_t)(v >> (8 * k)));
}

std::string json_string(const std::string &s) {
  std::string o = "\"";
  for (unsigned char c : s) {
    switch (c) {
      case '"': o += "\\\""; break;
      case '\\': o += "\\\\"; break;
      case '\n': o += "\\n"; break;
      case '\r': o += "\\r"; break;
      case '\t': o += "\\t"; break;
      case '\b': o += "\\b"; break;
      case '\f': o += "\\f"; break;
      default:
        if (c < 0x20 || c >= 0x7f) {  // ASCII output, as zarr's encoder writes it
          char u[8];
          snprintf(u, sizeof u, "\\u%04x", c);
          o += u;
        } else {
          o += (char)c;
        }
    }
  }
  return o + "\"";
}

// Indented JSON (4 spaces, one list element per line), the layout zarr
// writes its metadata in.
std::string json_list(const std::vector<std::string> &items, int indent) {
  if (items.empty()) return "[]";
  const std::string pad(indent + 4, ' ');
  std::string o = "[\n";
  for (size_t i = 0; i < items.size(); i++)
    o += pad + items[i] + (i + 1 < items.size() ? ",\n" : "\n");
  return o + std::string(indent, ' ') + "]";
}
std::string json_object(const std::vector<std::pair<std::string, std::string>> &kv,
                        int indent) {
  if (kv.empty()) return "{}";
  std::vector<std::pair<std::string, std::string>> s = kv;
  std::stable_sort(s.begin(), s.end(),
                   [](const auto &a, const auto &b) { return a.first < b.first; });
  const std::string pad(indent + 4, ' ');
  std::string o = "{\n";
  for (size_t i = 0; i < s.size(); i++)
    o += pad + json_string(s[i].first) + ": " + s[i].second + (i + 1 < s.size() ? ",\n" : "\n");
  return o + std::string(indent, ' ') + "}";
}

std::string strip(const std::string &path) {
  size_t a = 0, b = path.size();
  while (a < b && path[a] == '/') a++;
  while (b > a && path[b - 1] == '/') b--;
  return path.substr(a, b - a);
}

}  // namespace

uint32_t crc32(uint32_t crc, const void *p, size_t n) {
  const auto &T = crc_tables().t;
  const uint8_t *b = static_cast<const uint8_t *>(p);
  uint32_t c = ~crc;
  while (n >= 8) {
    uint32_t lo, hi;
    memcpy(&lo, b, 4);
    memcpy(&hi, b + 4, 4);
    lo ^= c;
    c = T[7][lo & 0xff] ^ T[6][(lo >> 8) & 0xff] ^ T[5][(lo >> 16) & 0xff] ^ T[4][lo >> 24] ^
        T[3][hi & 0xff] ^ T[2][(hi >> 8) & 0xff] ^ T[1][(hi >> 16) & 0xff] ^ T[0][hi >> 24];
    b += 8;
    n -= 8;
  }
  while (n--) c = T[0][(c ^ *b++) & 0xff] ^ (c >> 8);
  return ~c;
}

std::string json_number(double v) {
  if (std::isnan(v)) return "NaN";
  if (std::isinf(v)) return v > 0 ? "Infinity" : "-Infinity";
  char buf[40];
  for (int prec = 1; prec <= 17; prec++) {
    snprintf(buf, sizeof buf, "%.*g", prec, v);
    if (strtod(buf, nullptr) == v) break;
  }
  std::string s = buf;
  if (s.find_first_of(".e") == std::string::npos) s += ".0";
  return s;
}

Writer::~Writer() {
  if (f_) fclose(f_);
}

int Writer::fail(const std::string &msg) {
  if (err_.empty()) err_ = msg;
  return -1;
}

int Writer::open(const std::string &path) {
  if (f_) return fail("zarr writer already open");
  f_ = fopen(path.c_str(), "wb");
  if (!f_) return fail("cannot create " + path);
  pos_ = 0;
  entries_.clear();
  nodes_.clear();
  nodes_[""] = Node();
  return 0;
}

Writer::Node *Writer::node(const std::string &path, bool want_group) {
  const std::string p = strip(path);
  // parents are groups
  for (size_t k = p.find('/'); k != std::string::npos; k = p.find('/', k + 1)) {
    auto it = nodes_.find(p.substr(0, k));
    if (it == nodes_.end()) {
      nodes_[p.substr(0, k)] = Node();
    } else if (!it->second.is_group) {
      fail("'" + p.substr(0, k) + "' is an array, not a group");
      return nullptr;
    }
  }
  auto it = nodes_.find(p);
  if (it != nodes_.end()) {
    if (it->second.is_group != want_group) {
      fail("'" + p + "' exists with another kind");
      return nullptr;
    }
    return &it->second;
  }
  Node &n = nodes_[p];
  n.is_group = want_group;
  return &n;
}

int Writer::put(const std::string &name, const void *p, size_t n) {
  if (!f_) return fail("zarr writer not open");
  if (n >= 0xFFFFFFFFull) return fail("zip entry of 4 GiB or more: " + name);
  Entry e;
  e.name = name;
  e.size = n;
  e.offset = pos_;
  e.crc = crc32(0, p, n);
  std::vector<uint8_t> h;
  le32(h, 0x04034b50u);
  le16(h, 20);      // version needed: 2.0
  le16(h, 0);       // flags
  le16(h, 0);       // stored
  le16(h, 0);       // time 00:00
  le16(h, 0x21);    // date 1980-01-01
  le32(h, e.crc);
  le32(h, (uint32_t)n);
  le32(h, (uint32_t)n);
  le16(h, (uint32_t)name.size());
  le16(h, 0);
  h.insert(h.end(), name.begin(), name.end());
  if (fwrite(h.data(), 1, h.size(), f_) != h.size() || (n && fwrite(p, 1, n, f_) != n))
    return fail("write failed: " + name);
  pos_ += h.size() + n;
  entries_.push_back(e);
  return 0;
}

int Writer::group(const std::string &path) {
  return node(path, true) ? 0 : -1;
}

int Writer::dataset(const std::string &path, h5::Kind kind, const std::vector<uint64_t> &dims,
                    const void *data) {
  const char *dt = nullptr;
  int es = 0;
  switch (kind) {
    case h5::Kind::F64: dt = "<f8"; es = 8; break;
    case h5::Kind::I32: dt = "<i4"; es = 4; break;
    case h5::Kind::I64: dt = "<i8"; es = 8; break;
    default: return fail("zarr writer: only numeric arrays");
  }
  if (!node(path, false)) return -1;
  const std::string p = strip(path);
  uint64_t count = 1;
  std::vector<std::string> shape, chunks;
  for (uint64_t d : dims) {
    count *= d;
    shape.push_back(std::to_string(d));
    chunks.push_back(std::to_string(d ? d : 1));
  }
  const uint64_t nbytes = count * (uint64_t)es;
  if (nbytes + 16 > 0x7FFFFFFFull) return fail("array too large for one Blosc chunk: " + p);
  const std::string comp = json_object({{"blocksize", "0"},
                                        {"clevel", "3"},
                                        {"cname", json_string("zstd")},
                                        {"id", json_string("blosc")},
                                        {"shuffle", "2"}},
                                       4);
  const std::string meta =
      json_object({{"chunks", json_list(chunks, 4)},
                   {"compressor", comp},
                   {"dtype", json_string(dt)},
                   {"fill_value", kind == h5::Kind::F64 ? "0.0" : "0"},
                   {"filters", "null"},
                   {"order", json_string("C")},
                   {"shape", json_list(shape, 4)},
                   {"zarr_format", "2"}},
                  0);
  if (put_json(p + "/.zarray", meta)) return -1;
  if (count == 0) return 0;  // an empty array has no chunk
  // One Blosc1 chunk, stored as is: version 2, codec version 1, flags
  // "memcpyed", type size, then nbytes, block size, compressed bytes.
  std::vector<uint8_t> chunk;
  chunk.reserve(16 + nbytes);
  chunk.push_back(2);
  chunk.push_back(1);
  chunk.push_back(0x02);
  chunk.push_back((uint8_t)es);
  le32(chunk, (uint32_t)nbytes);
  le32(chunk, (uint32_t)nbytes);
  le32(chunk, (uint32_t)(nbytes + 16));
  const uint8_t *b = static_cast<const uint8_t *>(data);
  chunk.insert(chunk.end(), b, b + nbytes);  // little-endian host
  std::string key;
  for (size_t k = 0; k < std::max<size_t>(dims.size(), 1); k++) key += k ? ".0" : "0";
  return put(p + "/" + key, chunk.data(), chunk.size());
}

int Writer::add_attr(const std::string &path, const std::string &name, std::string json) {
  const std::string p = strip(path);
  auto it = nodes_.find(p);
  if (it == nodes_.end()) return fail("no object '" + p + "' for attribute " + name);
  for (auto &kv : it->second.attrs)
    if (kv.first == name) {
      kv.second = std::move(json);
      return 0;
    }
  it->second.attrs.emplace_back(name, std::move(json));
  return 0;
}

int Writer::attr_f64(const std::string &path, const std::string &name, const double *v, int n) {
  if (n == 1) return add_attr(path, name, json_number(v[0]));
  std::vector<std::string> items;
  for (int i = 0; i < n; i++) items.push_back(json_number(v[i]));
  return add_attr(path, name, json_list(items, 4));
}

int Writer::attr_i32(const std::string &path, const std::string &name, const int32_t *v,
                     int n) {
  if (n == 1) return add_attr(path, name, std::to_string(v[0]));
  std::vector<std::string> items;
  for (int i = 0; i < n; i++) items.push_back(std::to_string(v[i]));
  return add_attr(path, name, json_list(items, 4));
}

int Writer::attr_str(const std::string &path, const std::string &name,
                     const std::string &value) {
  return add_attr(path, name, json_string(value));
}

int Writer::close() {
  if (!f_) return 0;
  int rc = err_.empty() ? 0 : -1;
  // Metadata of every group and array: .zgroup for groups, .zattrs for all
  // (the conversion puts each object's attributes, also when empty).
  const std::string zgroup = json_object({{"zarr_format", "2"}}, 0);
  for (auto &kv : nodes_) {
    if (rc) break;
    const std::string pre = kv.first.empty() ? "" : kv.first + "/";
    if (kv.second.is_group && put_json(pre + ".zgroup", zgroup)) rc = -1;
    if (!rc && put_json(pre + ".zattrs", json_object(kv.second.attrs, 0))) rc = -1;
  }
  // Central directory, with ZIP64 fields where 32 bits do not suffice.
  const uint64_t cd_start = pos_;
  std::vector<uint8_t> cd;
  for (const Entry &e : entries_) {
    const bool big = e.offset >= 0xFFFFFFFFull;
    le32(cd, 0x02014b50u);
    le16(cd, big ? 45 : 20);  // made by
    le16(cd, big ? 45 : 20);  // needed
    le16(cd, 0);
    le16(cd, 0);
    le16(cd, 0);
    le16(cd, 0x21);
    le32(cd, e.crc);
    le32(cd, (uint32_t)e.size);
    le32(cd, (uint32_t)e.size);
    le16(cd, (uint32_t)e.name.size());
    le16(cd, big ? 12 : 0);
    le16(cd, 0);  // comment
    le16(cd, 0);  // disk
    le16(cd, 0);  // internal attributes
    le32(cd, 0);  // external attributes
    le32(cd, big ? 0xFFFFFFFFu : (uint32_t)e.offset);
    cd.insert(cd.end(), e.name.begin(), e.name.end());
    if (big) {
      le16(cd, 0x0001);
      le16(cd, 8);
      le64(cd, e.offset);
    }
  }
  const uint64_t cd_size = cd.size(), n = entries_.size();
  const bool z64 = n >= 0xFFFF || cd_start >= 0xFFFFFFFFull || cd_size >= 0xFFFFFFFFull;
  std::vector<uint8_t> tail;
  if (z64) {
    const uint64_t rec = cd_start + cd_size;
    le32(tail, 0x06064b50u);
    le64(tail, 44);
    le16(tail, 45);
    le16(tail, 45);
    le32(tail, 0);
    le32(tail, 0);
    le64(tail, n);
    le64(tail, n);
    le64(tail, cd_size);
    le64(tail, cd_start);
    le32(tail, 0x07064b50u);
    le32(tail, 0);
    le64(tail, rec);
    le32(tail, 1);
  }
  le32(tail, 0x06054b50u);
  le16(tail, 0);
  le16(tail, 0);
  le16(tail, z64 ? 0xFFFF : (uint32_t)n);
  le16(tail, z64 ? 0xFFFF : (uint32_t)n);
  le32(tail, z64 ? 0xFFFFFFFFu : (uint32_t)cd_size);
  le32(tail, z64 ? 0xFFFFFFFFu : (uint32_t)cd_start);
  le16(tail, 0);
  if (!rc && (fwrite(cd.data(), 1, cd.size(), f_) != cd.size() ||
              fwrite(tail.data(), 1, tail.size(), f_) != tail.size()))
    rc = fail("write failed: zip directory");
  if (fclose(f_) != 0 && !rc) rc = fail("closing the zip file failed");
  f_ = nullptr;
  return rc;
}

}  // namespace zarr
}  // namespace mc
