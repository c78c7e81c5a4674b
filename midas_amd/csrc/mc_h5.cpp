// Minimal HDF5 writer (see mc_h5.h).  Structures follow the HDF5 File Format
// Specification 3.0: superblock version 0 (section II.A), object header
// version 1 (IV.A.1), symbol-table groups with a version-1 B-tree, symbol
// table nodes and a local heap (III.A.1, III.C, III.D), global heap
// collections (III.E), dataspace / datatype / fill value / layout /
// attribute / symbol-table messages (IV.A.2).  Encodings follow what
// libhdf5 1.x writes by default for the same objects.
#include "mc_h5.h"

#include <algorithm>
#include <cstring>

namespace mc {
namespace h5 {
namespace {

constexpr uint64_t kUndef = ~0ull;
constexpr int kLeafK = 64;       // symbol-table node: up to 2K = 128 entries
constexpr int kInternalK = 256;  // B-tree node: up to 2K = 512 children
constexpr uint64_t kSuperblock = 96;
constexpr uint64_t kSnodBytes = 8 + 2 * kLeafK * 40;
constexpr uint64_t kBtreeBytes = 24 + 2 * kInternalK * 8 + (2 * kInternalK + 1) * 8;
constexpr uint64_t kGcolMin = 4096;

struct Buf {
  std::vector<uint8_t> b;
  void u8(uint32_t v) { b.push_back((uint8_t)v); }
  void u16(uint32_t v) { u8(v & 0xff); u8(v >> 8); }
  void u32(uint32_t v) { u16(v & 0xffff); u16(v >> 16); }
  void u64(uint64_t v) { u32((uint32_t)v); u32((uint32_t)(v >> 32)); }
  void raw(const void *p, size_t n) {
    const uint8_t *q = static_cast<const uint8_t *>(p);
    b.insert(b.end(), q, q + n);
  }
  void pad8() { while (b.size() % 8) u8(0); }
  void zeros(size_t n) { b.insert(b.end(), n, 0); }
  size_t size() const { return b.size(); }
};

size_t elem_size(Kind k) {
  switch (k) {
    case Kind::F64: return 8;
    case Kind::I32: return 4;
    case Kind::I64: return 8;
    case Kind::VSTR: return 16;
    default: return 1;
  }
}

// Datatype message body (version 1).  `strlen_` is the size of a fixed
// string type.
Buf datatype(Kind k, uint32_t strlen_ = 0) {
  Buf t;
  switch (k) {
    case Kind::F64:  // IEEE little-endian double
      t.u8(0x11); t.u8(0x20); t.u8(0x3f); t.u8(0x00); t.u32(8);
      t.u16(0); t.u16(64); t.u8(52); t.u8(11); t.u8(0); t.u8(52); t.u32(1023);
      break;
    case Kind::I32:
    case Kind::I64: {  // signed little-endian integer
      const uint32_t n = k == Kind::I32 ? 4 : 8;
      t.u8(0x10); t.u8(0x08); t.u8(0x00); t.u8(0x00); t.u32(n);
      t.u16(0); t.u16(8 * n);
      break;
    }
    case Kind::STR:  // fixed length, null terminated, ASCII
      t.u8(0x13); t.u8(0x00); t.u8(0x00); t.u8(0x00); t.u32(strlen_);
      break;
    case Kind::VSTR:  // variable-length string, null terminated, UTF-8; base u8
      t.u8(0x19); t.u8(0x01); t.u8(0x01); t.u8(0x00); t.u32(16);
      t.u8(0x10); t.u8(0x00); t.u8(0x00); t.u8(0x00); t.u32(1); t.u16(0); t.u16(8);
      break;
  }
  return t;
}

// Dataspace message body (version 1); rank 0 = scalar.
Buf dataspace(const std::vector<uint64_t> &dims) {
  Buf s;
  s.u8(1); s.u8((uint32_t)dims.size()); s.u8(0); s.u8(0); s.u32(0);
  for (uint64_t d : dims) s.u64(d);
  return s;
}

struct Msg {
  uint16_t type;
  uint8_t flags;
  Buf body;
};

Buf object_header(std::vector<Msg> &msgs) {
  Buf h;
  uint32_t size = 0;
  for (auto &m : msgs) {
    m.body.pad8();
    size += 8 + (uint32_t)m.body.size();
  }
  h.u8(1); h.u8(0); h.u16((uint32_t)msgs.size()); h.u32(1); h.u32(size); h.u32(0);
  for (auto &m : msgs) {
    h.u16(m.type); h.u16((uint32_t)m.body.size()); h.u8(m.flags); h.u8(0); h.u16(0);
    h.raw(m.body.b.data(), m.body.size());
  }
  return h;
}

}  // namespace

Writer::Writer() = default;
Writer::~Writer() {
  if (f_) fclose(f_);
}

int Writer::fail(const std::string &msg) {
  err_ = msg;
  return -1;
}

int Writer::open(const std::string &path) {
  if (f_) return fail("writer already open");
  f_ = fopen(path.c_str(), "wb+");
  if (!f_) return fail("cannot create " + path);
  std::vector<uint8_t> z(kSuperblock, 0);  // superblock written by close()
  if (fwrite(z.data(), 1, z.size(), f_) != z.size()) return fail("write failed");
  eof_ = kSuperblock;
  root_ = std::make_unique<Node>();
  return 0;
}

int Writer::write_at(uint64_t addr, const void *p, size_t n) {
  if (fseeko(f_, (off_t)addr, SEEK_SET) != 0 || fwrite(p, 1, n, f_) != n) return -1;
  return 0;
}

uint64_t Writer::append(const void *p, size_t n) {
  const uint64_t a = eof_;
  if (write_at(a, p, n)) return kUndef;
  eof_ += n;
  return a;
}

Writer::Node *Writer::find(const std::string &path, bool create_groups, bool want_group,
                           bool create_leaf) {
  std::vector<std::string> parts;
  size_t i = 0;
  while (i < path.size()) {
    size_t j = path.find('/', i);
    if (j == std::string::npos) j = path.size();
    if (j > i) parts.push_back(path.substr(i, j - i));
    i = j + 1;
  }
  Node *n = root_.get();
  for (size_t k = 0; k < parts.size(); k++) {
    const bool last = k + 1 == parts.size();
    auto it = n->kids.find(parts[k]);
    if (it == n->kids.end()) {
      if (last ? !create_leaf : !create_groups) return nullptr;
      auto c = std::make_unique<Node>();
      c->is_group = last ? want_group : true;
      Node *p = c.get();
      n->kids.emplace(parts[k], std::move(c));
      n = p;
    } else {
      n = it->second.get();
      if (!last && !n->is_group) return nullptr;
      if (last && create_leaf && !want_group) return nullptr;  // name taken
    }
  }
  return n;
}

int Writer::group(const std::string &path) {
  if (!f_) return fail("writer not open");
  Node *n = find(path, true, true, true);
  if (!n || !n->is_group) return fail("cannot create group " + path);
  return 0;
}

int Writer::dataset(const std::string &path, Kind kind, const std::vector<uint64_t> &dims,
                    const void *data) {
  if (!f_) return fail("writer not open");
  if (kind != Kind::F64 && kind != Kind::I32 && kind != Kind::I64)
    return fail("numeric dataset kind expected");
  Node *n = find(path, true, false, true);
  if (!n) return fail("cannot create dataset " + path);
  uint64_t cnt = 1;
  for (uint64_t d : dims) cnt *= d;
  n->is_group = false;
  n->kind = kind;
  n->dims = dims;
  n->data_size = cnt * elem_size(kind);
  n->data_addr = n->data_size ? append(data, n->data_size) : kUndef;
  if (n->data_size && n->data_addr == kUndef) return fail("write failed");
  return 0;
}

int Writer::vstr_dataset(const std::string &path, const std::vector<std::string> &values) {
  if (!f_) return fail("writer not open");
  Node *n = find(path, true, false, true);
  if (!n) return fail("cannot create dataset " + path);
  // Global heap collections hold the strings (objects 1..65535 each), then
  // the dataset's 16-byte references {length, collection, index}.
  Buf refs;
  for (size_t c0 = 0; c0 < values.size() || (c0 == 0 && values.empty()); c0 += 0xffff) {
    const size_t c1 = std::min(values.size(), c0 + 0xffff);
    if (c1 == c0) break;
    Buf g;
    g.raw("GCOL", 4); g.u8(1); g.u8(0); g.u16(0); g.u64(0);  // size patched below
    for (size_t i = c0; i < c1; i++) {
      g.u16((uint32_t)(i - c0 + 1)); g.u16(0); g.u32(0); g.u64(values[i].size());
      g.raw(values[i].data(), values[i].size());
      g.pad8();
    }
    // Free space object (index 0, its 16-byte header included) covers the
    // rest of the collection.
    const uint64_t free_at = g.size();
    const uint64_t total = std::max<uint64_t>(kGcolMin, free_at + 16);
    g.u16(0); g.u16(0); g.u32(0); g.u64(total - free_at);
    g.zeros(total - g.size());
    std::memcpy(&g.b[8], &total, 8);
    const uint64_t coll = append(g.b.data(), g.size());
    if (coll == kUndef) return fail("write failed");
    for (size_t i = c0; i < c1; i++) {
      refs.u32((uint32_t)values[i].size()); refs.u64(coll); refs.u32((uint32_t)(i - c0 + 1));
    }
  }
  n->is_group = false;
  n->kind = Kind::VSTR;
  n->dims = {values.size()};
  n->data_size = refs.size();
  n->data_addr = refs.size() ? append(refs.b.data(), refs.size()) : kUndef;
  return 0;
}

int Writer::attr_f64(const std::string &path, const std::string &name, const double *v, int n) {
  Node *o = f_ ? find(path, false, false, false) : nullptr;
  if (!o) return fail("no object " + path);
  Attr a{name, Kind::F64, n, {}};
  a.raw.assign((const uint8_t *)v, (const uint8_t *)v + 8 * n);
  o->attrs.push_back(std::move(a));
  return 0;
}

int Writer::attr_i32(const std::string &path, const std::string &name, const int32_t *v,
                     int n) {
  Node *o = f_ ? find(path, false, false, false) : nullptr;
  if (!o) return fail("no object " + path);
  Attr a{name, Kind::I32, n, {}};
  a.raw.assign((const uint8_t *)v, (const uint8_t *)v + 4 * n);
  o->attrs.push_back(std::move(a));
  return 0;
}

int Writer::attr_str(const std::string &path, const std::string &name,
                     const std::string &value) {
  Node *o = f_ ? find(path, false, false, false) : nullptr;
  if (!o) return fail("no object " + path);
  Attr a{name, Kind::STR, 1, {}};
  a.raw.assign(value.begin(), value.end());
  a.raw.push_back(0);  // size = strlen + 1, as H5LTset_attribute_string
  o->attrs.push_back(std::move(a));
  return 0;
}

// Writes the metadata of `n` (children first) and sets n.oh_addr.
int Writer::emit(Node &n) {
  std::vector<Msg> msgs;
  if (n.is_group) {
    for (auto &kv : n.kids)
      if (emit(*kv.second)) return -1;
    // Local heap: "" at offset 0, the names, then one free block.
    Buf data;
    data.zeros(8);
    std::vector<uint64_t> name_off;
    for (auto &kv : n.kids) {
      name_off.push_back(data.size());
      data.raw(kv.first.data(), kv.first.size());
      data.u8(0);
      data.pad8();
    }
    const uint64_t free_off = data.size();
    data.u64(1);   // end of the free list
    data.u64(16);  // this free block's size
    const uint64_t data_addr = append(data.b.data(), data.size());
    Buf heap;
    heap.raw("HEAP", 4); heap.u8(0); heap.u8(0); heap.u16(0);
    heap.u64(data.size()); heap.u64(free_off); heap.u64(data_addr);
    n.heap_addr = append(heap.b.data(), heap.size());
    // Symbol-table nodes of up to 2K entries, in name order.
    const size_t per = 2 * kLeafK;
    const size_t nkids = n.kids.size();
    const size_t nsnod = std::max<size_t>(1, (nkids + per - 1) / per);
    std::vector<uint64_t> child, key;  // per child: address, largest name
    auto it = n.kids.begin();
    size_t k = 0;
    for (size_t s = 0; s < nsnod; s++) {
      Buf sn;
      const size_t cnt = std::min(per, nkids - s * per);
      sn.raw("SNOD", 4); sn.u8(1); sn.u8(0); sn.u16((uint32_t)cnt);
      for (size_t e = 0; e < cnt; e++, ++it, ++k) {
        const Node &c = *it->second;
        sn.u64(name_off[k]); sn.u64(c.oh_addr);
        if (c.is_group) {
          sn.u32(1); sn.u32(0); sn.u64(c.btree_addr); sn.u64(c.heap_addr);
        } else {
          sn.u32(0); sn.u32(0); sn.zeros(16);
        }
      }
      sn.zeros(kSnodBytes - sn.size());
      child.push_back(append(sn.b.data(), sn.size()));
      key.push_back(cnt ? name_off[k - 1] : 0);
    }
    // Group B-tree, built level by level: nodes of up to 2K children, key 0
    // the empty name, key i the largest name under child i-1, siblings
    // linked on each level.
    const size_t fan = 2 * (size_t)kInternalK;
    int level = 0;
    for (;;) {
      const size_t nn = (child.size() + fan - 1) / fan;
      const uint64_t base = eof_;
      std::vector<uint64_t> up_child, up_key;
      for (size_t q = 0; q < nn; q++) {
        const size_t c0 = q * fan, c1 = std::min(child.size(), c0 + fan);
        Buf bt;
        bt.raw("TREE", 4); bt.u8(0); bt.u8((uint32_t)level);
        bt.u16((uint32_t)(nkids ? c1 - c0 : 0));
        bt.u64(q > 0 ? base + (q - 1) * kBtreeBytes : kUndef);
        bt.u64(q + 1 < nn ? base + (q + 1) * kBtreeBytes : kUndef);
        bt.u64(0);
        if (nkids)
          for (size_t c = c0; c < c1; c++) {
            bt.u64(child[c]);
            bt.u64(key[c]);
          }
        bt.zeros(kBtreeBytes - bt.size());
        up_child.push_back(append(bt.b.data(), bt.size()));
        up_key.push_back(key[c1 - 1]);
      }
      if (nn == 1) {
        n.btree_addr = up_child[0];
        break;
      }
      if (++level > 255) return fail("group has too many members");
      child.swap(up_child);
      key.swap(up_key);
    }
    Msg m{0x11, 0, {}};
    m.body.u64(n.btree_addr);
    m.body.u64(n.heap_addr);
    msgs.push_back(std::move(m));
  } else {
    Msg sp{0x1, 0, dataspace(n.dims)};
    msgs.push_back(std::move(sp));
    Msg dt{0x3, 1, datatype(n.kind)};
    msgs.push_back(std::move(dt));
    Msg fv{0x5, 1, {}};
    fv.body.u8(2); fv.body.u8(2); fv.body.u8(n.kind == Kind::VSTR ? 0 : 2); fv.body.u8(1);
    fv.body.u32(0);
    msgs.push_back(std::move(fv));
    Msg ly{0x8, 0, {}};
    ly.body.u8(3); ly.body.u8(1); ly.body.u64(n.data_addr); ly.body.u64(n.data_size);
    msgs.push_back(std::move(ly));
  }
  for (const Attr &a : n.attrs) {
    Buf t = a.kind == Kind::STR ? datatype(Kind::STR, (uint32_t)a.raw.size()) : datatype(a.kind);
    Buf s = a.kind == Kind::STR ? dataspace({}) : dataspace({(uint64_t)a.n});
    Msg m{0xC, 0, {}};
    m.body.u8(1); m.body.u8(0);
    m.body.u16((uint32_t)a.name.size() + 1);
    m.body.u16((uint32_t)t.size());
    m.body.u16((uint32_t)s.size());
    m.body.raw(a.name.data(), a.name.size()); m.body.u8(0); m.body.pad8();
    m.body.raw(t.b.data(), t.size()); m.body.pad8();
    m.body.raw(s.b.data(), s.size()); m.body.pad8();
    m.body.raw(a.raw.data(), a.raw.size());
    msgs.push_back(std::move(m));
  }
  Buf oh = object_header(msgs);
  n.oh_addr = append(oh.b.data(), oh.size());
  if (n.oh_addr == kUndef) return fail("write failed");
  return 0;
}

int Writer::close() {
  if (!f_) return fail("writer not open");
  int rc = emit(*root_);
  if (!rc) {
    Buf sb;
    sb.raw("\x89HDF\r\n\x1a\n", 8);
    sb.u8(0); sb.u8(0); sb.u8(0); sb.u8(0); sb.u8(0); sb.u8(8); sb.u8(8); sb.u8(0);
    sb.u16(kLeafK); sb.u16(kInternalK); sb.u32(0);
    sb.u64(0); sb.u64(kUndef); sb.u64(eof_); sb.u64(kUndef);
    // Root group symbol-table entry (cache type 1: B-tree and heap).
    sb.u64(0); sb.u64(root_->oh_addr); sb.u32(1); sb.u32(0);
    sb.u64(root_->btree_addr); sb.u64(root_->heap_addr);
    if (write_at(0, sb.b.data(), sb.size())) rc = fail("write failed");
  }
  if (fclose(f_) != 0 && !rc) rc = fail("close failed");
  f_ = nullptr;
  root_.reset();
  return rc;
}

}  // namespace h5
}  // namespace mc
