// mc_blosc.cpp -- Blosc1 chunk decoder for the Zarr-zip input.
//
// The reference reads its Zarr zips with libzip + blosc1_decompress
// (IntegratorZarrOMP.c:933-1401, 1416-1454, 1646-1669; ZarrReader.c:7-62).
// This decoder implements the Blosc1 container (c-blosc README_HEADER:
// 16-byte header, block starts, per-block streams, byte shuffle and
// bitshuffle) and hands the streams to the codec libraries of the system,
// loaded at run time: libzstd (zstd), liblz4 (lz4, lz4hc) and libz (zlib).
// BloscLZ and Snappy streams are rejected with an error.
#include <dlfcn.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "mc_internal.h"

namespace mc {
namespace {

// ---- codec libraries (dlopen) ----
typedef size_t (*zstd_decompress_t)(void *, size_t, const void *, size_t);
typedef unsigned (*zstd_is_error_t)(size_t);
typedef int (*lz4_decompress_safe_t)(const char *, char *, int, int);
typedef int (*zlib_uncompress_t)(unsigned char *, unsigned long *, const unsigned char *,
                                 unsigned long);

struct Codecs {
  zstd_decompress_t zstd = nullptr;
  zstd_is_error_t zstd_err = nullptr;
  lz4_decompress_safe_t lz4 = nullptr;
  zlib_uncompress_t zlib = nullptr;
};

const Codecs &codecs() {
  static Codecs c;
  static std::once_flag once;
  std::call_once(once, [] {
    if (void *h = dlopen("libzstd.so.1", RTLD_NOW | RTLD_LOCAL)) {
      c.zstd = (zstd_decompress_t)dlsym(h, "ZSTD_decompress");
      c.zstd_err = (zstd_is_error_t)dlsym(h, "ZSTD_isError");
    }
    if (void *h = dlopen("liblz4.so.1", RTLD_NOW | RTLD_LOCAL))
      c.lz4 = (lz4_decompress_safe_t)dlsym(h, "LZ4_decompress_safe");
    if (void *h = dlopen("libz.so.1", RTLD_NOW | RTLD_LOCAL))
      c.zlib = (zlib_uncompress_t)dlsym(h, "uncompress");
  });
  return c;
}

constexpr uint8_t kShuffle = 0x1, kMemcpyed = 0x2, kBitShuffle = 0x4, kDontSplit = 0x10;
enum { kBloscLZ = 0, kLZ4 = 1, kSnappy = 2, kZlib = 3, kZstd = 4 };

uint32_t rd32(const uint8_t *p) {
  uint32_t v;
  memcpy(&v, p, 4);
  return v;  // little-endian host
}

// One stream of one block: n_out bytes from n_in compressed bytes.
int decode_stream(int codec, const uint8_t *in, int64_t n_in, uint8_t *out, int64_t n_out) {
  const Codecs &c = codecs();
  switch (codec) {
    case kZstd: {
      MC_CHECK(c.zstd && c.zstd_err, MC_ERR_FORMAT, "zstd stream but libzstd.so.1 not found");
      const size_t r = c.zstd(out, (size_t)n_out, in, (size_t)n_in);
      MC_CHECK(!c.zstd_err(r) && (int64_t)r == n_out, MC_ERR_FORMAT,
               "zstd stream decoded to %lld of %lld bytes", (long long)(c.zstd_err(r) ? -1 : r),
               (long long)n_out);
      return MC_OK;
    }
    case kLZ4: {
      MC_CHECK(c.lz4, MC_ERR_FORMAT, "lz4 stream but liblz4.so.1 not found");
      const int r = c.lz4((const char *)in, (char *)out, (int)n_in, (int)n_out);
      MC_CHECK(r == n_out, MC_ERR_FORMAT, "lz4 stream decoded to %d of %lld bytes", r,
               (long long)n_out);
      return MC_OK;
    }
    case kZlib: {
      MC_CHECK(c.zlib, MC_ERR_FORMAT, "zlib stream but libz.so.1 not found");
      unsigned long len = (unsigned long)n_out;
      const int r = c.zlib(out, &len, in, (unsigned long)n_in);
      MC_CHECK(r == 0 && (int64_t)len == n_out, MC_ERR_FORMAT,
               "zlib stream: status %d, %lu of %lld bytes", r, len, (long long)n_out);
      return MC_OK;
    }
    default:
      set_error("blosc codec %d (%s) is not supported", codec,
                codec == kBloscLZ ? "blosclz" : codec == kSnappy ? "snappy" : "unknown");
      return MC_ERR_FORMAT;
  }
}

// Byte unshuffle of one block: element e, byte j <- src[j * n_el + e];
// the bytes past the last whole element are copied.
void unshuffle(int ts, int64_t bsize, const uint8_t *src, uint8_t *dst) {
  const int64_t n_el = bsize / ts;
  for (int j = 0; j < ts; j++) {
    const uint8_t *s = src + (int64_t)j * n_el;
    for (int64_t e = 0; e < n_el; e++) dst[e * ts + j] = s[e];
  }
  const int64_t done = n_el * ts;
  memcpy(dst + done, src + done, (size_t)(bsize - done));
}

// 8x8 bit-matrix transpose of the bytes of x (byte r, bit c) -> (byte c,
// bit r) (Hacker's Delight, transpose8).
inline uint64_t transpose8(uint64_t x) {
  uint64_t t;
  t = (x ^ (x >> 7)) & 0x00AA00AA00AA00AAull;
  x = x ^ t ^ (t << 7);
  t = (x ^ (x >> 14)) & 0x0000CCCC0000CCCCull;
  x = x ^ t ^ (t << 14);
  t = (x ^ (x >> 28)) & 0x00000000F0F0F0F0ull;
  x = x ^ t ^ (t << 28);
  return x;
}

// Bit unshuffle of n_el elements of ts bytes (n_el a multiple of 8).  The
// shuffled block holds, for each element byte j and bit k, one bit per
// element (element e at bit e % 8 of byte e / 8) in a row of n_el / 8
// bytes, rows ordered (j, k).
void bitunshuffle(int ts, int64_t n_el, const uint8_t *src, uint8_t *dst) {
  const int64_t rb = n_el / 8;  // bytes per bit row
  for (int j = 0; j < ts; j++) {
    const uint8_t *rows = src + (int64_t)j * 8 * rb;
    for (int64_t g = 0; g < rb; g++) {
      uint64_t x = 0;
      for (int k = 0; k < 8; k++) x |= (uint64_t)rows[k * rb + g] << (8 * k);
      x = transpose8(x);  // byte m: bit k = bit k of element 8g+m's byte j
      for (int m = 0; m < 8; m++) dst[(8 * g + m) * ts + j] = (uint8_t)(x >> (8 * m));
    }
  }
}

int blosc_decode(const uint8_t *src, int64_t n_src, uint8_t *dst, int64_t cap, int64_t *out_n) {
  MC_CHECK(n_src >= 16, MC_ERR_FORMAT, "blosc buffer of %lld bytes has no header",
           (long long)n_src);
  const int version = src[0], flags = src[2], ts = std::max<int>(1, src[3]);
  const int64_t nbytes = rd32(src + 4), blocksize = rd32(src + 8), cbytes = rd32(src + 12);
  MC_CHECK(cbytes <= n_src, MC_ERR_FORMAT, "blosc header says %lld bytes, buffer has %lld",
           (long long)cbytes, (long long)n_src);
  MC_CHECK(nbytes <= cap, MC_ERR_ARG, "blosc chunk of %lld bytes, destination holds %lld",
           (long long)nbytes, (long long)cap);
  *out_n = nbytes;
  if (nbytes == 0) return MC_OK;
  if (flags & kMemcpyed) {
    MC_CHECK(16 + nbytes <= n_src, MC_ERR_FORMAT, "stored blosc chunk is truncated");
    memcpy(dst, src + 16, (size_t)nbytes);
    return MC_OK;
  }
  MC_CHECK(blocksize > 0, MC_ERR_FORMAT, "blosc block size 0");
  const int codec = flags >> 5;
  const bool dont_split = (flags & kDontSplit) != 0;
  const bool shuf = (flags & kShuffle) != 0, bshuf = (flags & kBitShuffle) != 0;
  const int64_t nblocks = (nbytes + blocksize - 1) / blocksize;
  MC_CHECK(16 + 4 * nblocks <= n_src, MC_ERR_FORMAT, "blosc block table is truncated");
  // (a block is never longer than the chunk: a larger stated block size
  // means one short block)
  std::vector<uint8_t> tmp((size_t)std::min(blocksize, nbytes));
  for (int64_t b = 0; b < nblocks; b++) {
    const int64_t bsize = std::min(blocksize, nbytes - b * blocksize);
    const bool leftover = bsize < blocksize;
    int64_t pos = rd32(src + 16 + 4 * b);
    const int nstreams = (!dont_split && !leftover) ? ts : 1;
    const int64_t neblock = bsize / nstreams;
    MC_CHECK(neblock * nstreams == bsize, MC_ERR_FORMAT, "blosc block of %lld bytes in %d streams",
             (long long)bsize, nstreams);
    uint8_t *out = (shuf || bshuf) ? tmp.data() : dst + b * blocksize;
    for (int s = 0; s < nstreams; s++) {
      MC_CHECK(pos + 4 <= n_src, MC_ERR_FORMAT, "blosc stream header past the buffer");
      const int64_t cs = (int32_t)rd32(src + pos);
      pos += 4;
      MC_CHECK(cs >= 0 && pos + cs <= n_src, MC_ERR_FORMAT, "blosc stream past the buffer");
      if (cs == neblock) {
        memcpy(out + s * neblock, src + pos, (size_t)neblock);
      } else {
        int rc = decode_stream(codec, src + pos, cs, out + s * neblock, neblock);
        if (rc) return rc;
      }
      pos += cs;
    }
    uint8_t *d = dst + b * blocksize;
    if (bshuf) {
      int64_t n_el = bsize / ts;
      if (version <= 2) {
        if (n_el % 8) n_el = 0;  // format 2 leaves such blocks unshuffled
      } else {
        n_el -= n_el % 8;
      }
      if (n_el) bitunshuffle(ts, n_el, tmp.data(), d);
      memcpy(d + n_el * ts, tmp.data() + n_el * ts, (size_t)(bsize - n_el * ts));
    } else if (shuf) {
      unshuffle(ts, bsize, tmp.data(), d);
    }
  }
  return MC_OK;
}

}  // namespace
}  // namespace mc

using namespace mc;

extern "C" {

int mc_blosc_decompress(const void *src, int64_t src_size, void *dst, int64_t dst_capacity,
                        int64_t *out_size) {
  MC_CHECK(src && dst && out_size && src_size >= 0 && dst_capacity >= 0, MC_ERR_ARG,
           "bad arguments");
  return blosc_decode((const uint8_t *)src, src_size, (uint8_t *)dst, dst_capacity, out_size);
}

int mc_blosc_decompress_many(const void *const *srcs, const int64_t *src_sizes,
                             void *const *dsts, int64_t dst_capacity, int32_t n,
                             int32_t threads, int64_t *out_sizes) {
  MC_CHECK(srcs && src_sizes && dsts && out_sizes && n >= 0, MC_ERR_ARG, "bad arguments");
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  threads = std::min<int>(threads, std::max<int>(1, n));
  std::atomic<int> next{0};
  std::atomic<int> first_bad{-1};
  std::mutex mu;
  std::string err;
  int err_rc = MC_OK;
  auto work = [&] {
    for (int i = next++; i < n; i = next++) {
      int64_t got = 0;
      int rc = blosc_decode((const uint8_t *)srcs[i], src_sizes[i], (uint8_t *)dsts[i],
                            dst_capacity, &got);
      out_sizes[i] = rc ? -1 : got;
      if (rc) {
        std::lock_guard<std::mutex> lk(mu);
        if (first_bad < 0 || i < first_bad) {
          first_bad = i;
          err = mc_last_error();
          err_rc = rc;
        }
      }
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < threads; t++) pool.emplace_back(work);
  work();
  for (auto &t : pool) t.join();
  if (first_bad >= 0) {
    set_error("chunk %d: %s", (int)first_bad, err.c_str());
    return err_rc;
  }
  return MC_OK;
}

}  // extern "C"
