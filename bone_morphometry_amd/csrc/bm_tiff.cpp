// bm_tiff.cpp -- multi-page 8-bit TIFF volumes into host or device memory,
// and their calibration (bone-morphometry 0.7.0-rc.3, src/tiff_io.rs).
//
// The reference reads pages one after another with the tiff crate (0.11.3,
// Cargo.toml:18).  Here the page directory (the IFD chain) is read first;
// pages are then independent and are decoded on a pool of host threads, each
// page checked and binarised (value > threshold -> 1) as soon as it is
// decoded, while it is still in cache.  The device form decodes into pinned
// staging buffers and streams them to HBM while later pages decode.  Errors
// are those of the first failing page in page order, checked in the
// reference's order within the page, so the message is the one the
// sequential reader gives.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "bm_engine.h"
#include "bm_tiff.h"

namespace bm {

namespace {

std::string fmtf(const char* f, ...) __attribute__((format(printf, 1, 2)));
std::string fmtf(const char* f, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, f);
  vsnprintf(buf, sizeof buf, f, ap);
  va_end(ap);
  return buf;
}

// f64 as Rust's Display writes it: shortest round-trip digits, no exponent.
std::string display_f64(double v) {
  if (std::isnan(v)) return "NaN";
  if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
  char buf[64];
  for (int prec = 0; prec <= 16; ++prec) {
    snprintf(buf, sizeof buf, "%.*e", prec, v);
    if (std::strtod(buf, nullptr) == v) break;
  }
  std::string s = buf;
  const bool neg = s[0] == '-';
  if (neg) s = s.substr(1);
  const size_t e = s.find('e');
  const int exp10 = std::atoi(s.c_str() + e + 1);
  std::string digits;
  for (size_t i = 0; i < e; ++i)
    if (s[i] != '.') digits += s[i];
  while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
  std::string r;
  const int point = exp10 + 1;
  if (point <= 0) r = "0." + std::string(-point, '0') + digits;
  else if ((size_t)point >= digits.size()) r = digits + std::string(point - digits.size(), '0');
  else r = digits.substr(0, point) + "." + digits.substr(point);
  if (r == "0") return neg ? "-0" : "0";
  return neg ? "-" + r : r;
}

// A read-only mapping of the file.
class Mapped {
 public:
  explicit Mapped(const std::string& path) {
    fd_ = ::open(path.c_str(), O_RDONLY);
    if (fd_ < 0) throw Error("opening " + path);
    struct stat st;
    if (fstat(fd_, &st) != 0) throw Error("opening " + path);
    size_ = (uint64_t)st.st_size;
    if (size_) {
      void* p = mmap(nullptr, size_, PROT_READ, MAP_PRIVATE, fd_, 0);
      if (p == MAP_FAILED) throw Error("opening " + path);
      data_ = static_cast<const uint8_t*>(p);
    }
  }
  ~Mapped() {
    if (data_) munmap(const_cast<uint8_t*>(data_), size_);
    if (fd_ >= 0) ::close(fd_);
  }
  Mapped(const Mapped&) = delete;
  Mapped& operator=(const Mapped&) = delete;
  const uint8_t* data() const { return data_; }
  uint64_t size() const { return size_; }

 private:
  int fd_ = -1;
  const uint8_t* data_ = nullptr;
  uint64_t size_ = 0;
};

// Thrown inside the decoder; the caller wraps it in the reference's context.
struct TiffFormatError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

enum : uint16_t {
  kTagWidth = 256, kTagHeight = 257, kTagBits = 258, kTagCompression = 259,
  kTagPhotometric = 262, kTagDescription = 270, kTagStripOffsets = 273, kTagSamples = 277,
  kTagRowsPerStrip = 278, kTagStripCounts = 279, kTagXRes = 282, kTagYRes = 283,
  kTagPlanar = 284, kTagResUnit = 296, kTagPredictor = 317, kTagTileWidth = 322,
  kTagTileLength = 323, kTagTileOffsets = 324, kTagTileCounts = 325, kTagSampleFormat = 339,
};

struct Entry {
  uint16_t type = 0;
  uint64_t count = 0;
  uint64_t value_at = 0;  // file offset of the value bytes
};

struct Ifd {
  std::map<uint16_t, Entry> tags;
};

class TiffFile {
 public:
  explicit TiffFile(const Mapped& m) : m_(m) {
    if (m.size() < 8) throw TiffFormatError("file too short for a TIFF header");
    const uint8_t* d = m.data();
    if (d[0] == 'I' && d[1] == 'I') le_ = true;
    else if (d[0] == 'M' && d[1] == 'M') le_ = false;
    else throw TiffFormatError("not a TIFF byte order mark");
    const uint16_t ver = u16(2);
    if (ver == 42) {
      big_ = false;
      first_ = u32(4);
    } else if (ver == 43) {
      big_ = true;
      if (m.size() < 16 || u16(4) != 8 || u16(6) != 0) throw TiffFormatError("bad BigTIFF header");
      first_ = u64(8);
    } else {
      throw TiffFormatError("not a TIFF version number");
    }
  }

  uint64_t first_ifd() const { return first_; }

  // Reads the IFD at `off`; returns the offset of the next (0: none).
  uint64_t read_ifd(uint64_t off, Ifd& ifd) const {
    const uint64_t head = big_ ? 8 : 2, esz = big_ ? 20 : 12, nsz = big_ ? 8 : 4;
    need(off, head);
    const uint64_t n = big_ ? u64(off) : u16(off);
    need(off + head, n * esz + nsz);
    for (uint64_t i = 0; i < n; ++i) {
      const uint64_t e = off + head + i * esz;
      Entry en;
      const uint16_t tag = u16(e);
      en.type = u16(e + 2);
      en.count = big_ ? u64(e + 4) : u32(e + 4);
      const uint64_t size = type_size(en.type) * en.count;
      const uint64_t inline_cap = big_ ? 8 : 4;
      const uint64_t field = e + (big_ ? 12 : 8);
      en.value_at = size <= inline_cap ? field : (big_ ? u64(field) : u32(field));
      if (type_size(en.type) && size > inline_cap) need(en.value_at, size);
      ifd.tags[tag] = en;
    }
    return big_ ? u64(off + head + n * esz) : u32(off + head + n * esz);
  }

  bool has(const Ifd& ifd, uint16_t tag) const { return ifd.tags.count(tag) != 0; }

  // Unsigned integer values of a tag (BYTE, SHORT, LONG, LONG8, IFD types).
  std::vector<uint64_t> uints(const Ifd& ifd, uint16_t tag) const {
    std::vector<uint64_t> v;
    auto it = ifd.tags.find(tag);
    if (it == ifd.tags.end()) return v;
    const Entry& e = it->second;
    const uint64_t sz = type_size(e.type);
    if (!(e.type == 1 || e.type == 3 || e.type == 4 || e.type == 16 || e.type == 13 ||
          e.type == 18))
      throw TiffFormatError(fmtf("tag %u has a non-integer type %u", tag, e.type));
    v.reserve(e.count);
    for (uint64_t i = 0; i < e.count; ++i) {
      const uint64_t at = e.value_at + i * sz;
      v.push_back(sz == 1 ? m_.data()[at] : sz == 2 ? u16(at) : sz == 4 ? u32(at) : u64(at));
    }
    return v;
  }

  bool uint1(const Ifd& ifd, uint16_t tag, uint64_t& out) const {
    const std::vector<uint64_t> v = uints(ifd, tag);
    if (v.empty()) return false;
    out = v[0];
    return true;
  }

  // A RATIONAL (or other numeric) tag as f64; false when absent or not numeric.
  bool f64(const Ifd& ifd, uint16_t tag, double& out) const {
    auto it = ifd.tags.find(tag);
    if (it == ifd.tags.end() || it->second.count < 1) return false;
    const Entry& e = it->second;
    switch (e.type) {
      case 5: out = (double)u32(e.value_at) / (double)u32(e.value_at + 4); return true;
      case 10: out = (double)(int32_t)u32(e.value_at) / (double)(int32_t)u32(e.value_at + 4); return true;
      case 11: {
        const uint32_t b = u32(e.value_at);
        float f;
        std::memcpy(&f, &b, 4);
        out = f;
        return true;
      }
      case 12: {
        const uint64_t b = u64(e.value_at);
        std::memcpy(&out, &b, 8);
        return true;
      }
      case 1: case 3: case 4: case 16: out = (double)uints(ifd, tag)[0]; return true;
      default: return false;
    }
  }

  bool ascii(const Ifd& ifd, uint16_t tag, std::string& out) const {
    auto it = ifd.tags.find(tag);
    if (it == ifd.tags.end() || it->second.type != 2) return false;
    out.assign(reinterpret_cast<const char*>(m_.data() + it->second.value_at), it->second.count);
    while (!out.empty() && out.back() == '\0') out.pop_back();  // the terminator
    return true;
  }

  const uint8_t* bytes(uint64_t off, uint64_t len) const {
    need(off, len);
    return m_.data() + off;
  }

 private:
  void need(uint64_t off, uint64_t len) const {
    if (off > m_.size() || len > m_.size() - off) throw TiffFormatError("offset past end of file");
  }
  static uint64_t type_size(uint16_t t) {
    switch (t) {
      case 1: case 2: case 6: case 7: return 1;
      case 3: case 8: return 2;
      case 4: case 9: case 11: case 13: return 4;
      case 5: case 10: case 12: case 16: case 17: case 18: return 8;
      default: return 0;
    }
  }
  uint16_t u16(uint64_t o) const {
    need(o, 2);
    const uint8_t* p = m_.data() + o;
    return le_ ? (uint16_t)(p[0] | p[1] << 8) : (uint16_t)(p[0] << 8 | p[1]);
  }
  uint32_t u32(uint64_t o) const {
    need(o, 4);
    const uint8_t* p = m_.data() + o;
    return le_ ? (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24
               : (uint32_t)p[3] | (uint32_t)p[2] << 8 | (uint32_t)p[1] << 16 | (uint32_t)p[0] << 24;
  }
  uint64_t u64(uint64_t o) const {
    const uint64_t a = u32(o), b = u32(o + 4);
    return le_ ? a | b << 32 : b | a << 32;
  }

  const Mapped& m_;
  bool le_ = true, big_ = false;
  uint64_t first_ = 0;
};

// One page of the directory, as the decoder needs it.
struct Page {
  uint64_t width = 0, height = 0;
  bool dims_ok = false, color_ok = false;
  std::string dims_error, color_error, color_name;  // color_name: ColorType's Debug text
  uint16_t compression = 1, predictor = 1, sample_format = 1;
  bool tiled = false;
  uint64_t rows_per_strip = 0, tile_w = 0, tile_h = 0;
  std::vector<uint64_t> offsets, counts;
};

Page describe(const TiffFile& f, const Ifd& ifd) {
  Page p;
  uint64_t w = 0, h = 0;
  try {
    if (!f.uint1(ifd, kTagWidth, w) || !f.uint1(ifd, kTagHeight, h))
      throw TiffFormatError("required tag ImageWidth/ImageLength not found");
    if (w > 0xFFFFFFFFull || h > 0xFFFFFFFFull) throw TiffFormatError("dimensions out of range");
    p.width = w;
    p.height = h;
    p.dims_ok = true;
  } catch (const TiffFormatError& e) {
    p.dims_error = e.what();
    return p;
  }
  try {
    uint64_t photometric = 0, samples = 1;
    if (!f.uint1(ifd, kTagPhotometric, photometric))
      throw TiffFormatError("unknown photometric interpretation");
    f.uint1(ifd, kTagSamples, samples);
    std::vector<uint64_t> bits = f.uints(ifd, kTagBits);
    if (bits.empty()) bits.assign(samples, 1);
    for (uint64_t b : bits)
      if (b != bits[0]) throw TiffFormatError("inconsistent bits per sample");
    const unsigned b = (unsigned)bits[0];
    const char* name = nullptr;
    switch (photometric) {
      case 0: case 1: name = samples == 1 ? "Gray" : samples == 2 ? "GrayA" : nullptr; break;
      case 2: name = samples == 3 ? "RGB" : samples == 4 ? "RGBA" : nullptr; break;
      case 3: name = samples == 1 ? "Palette" : nullptr; break;
      case 5: name = samples == 4 ? "CMYK" : nullptr; break;
      case 6: name = samples == 3 ? "YCbCr" : nullptr; break;
      default: break;
    }
    if (!name) throw TiffFormatError("unsupported color type");
    p.color_name = fmtf("%s(%u)", name, b);
    p.color_ok = true;
  } catch (const TiffFormatError& e) {
    p.color_error = e.what();
    return p;
  }
  uint64_t v = 0;
  if (f.uint1(ifd, kTagCompression, v)) p.compression = (uint16_t)v;
  if (f.uint1(ifd, kTagPredictor, v)) p.predictor = (uint16_t)v;
  if (f.uint1(ifd, kTagSampleFormat, v)) p.sample_format = (uint16_t)v;
  p.tiled = f.has(ifd, kTagTileOffsets);
  if (p.tiled) {
    f.uint1(ifd, kTagTileWidth, p.tile_w);
    f.uint1(ifd, kTagTileLength, p.tile_h);
    p.offsets = f.uints(ifd, kTagTileOffsets);
    p.counts = f.uints(ifd, kTagTileCounts);
  } else {
    p.rows_per_strip = p.height;
    if (f.uint1(ifd, kTagRowsPerStrip, v) && v > 0) p.rows_per_strip = std::min(v, p.height);
    p.offsets = f.uints(ifd, kTagStripOffsets);
    p.counts = f.uints(ifd, kTagStripCounts);
  }
  return p;
}

// --- decompression ------------------------------------------------------------
void unpack_bits(const uint8_t* src, uint64_t n, std::vector<uint8_t>& out, uint64_t want) {
  out.clear();
  uint64_t i = 0;
  while (i < n && out.size() < want) {
    const int8_t c = (int8_t)src[i++];
    if (c >= 0) {
      const uint64_t len = (uint64_t)c + 1;
      if (i + len > n) throw TiffFormatError("truncated PackBits literal");
      out.insert(out.end(), src + i, src + i + len);
      i += len;
    } else if (c != -128) {
      if (i >= n) throw TiffFormatError("truncated PackBits run");
      out.insert(out.end(), (size_t)(1 - c), src[i++]);
    }
  }
}

// TIFF LZW: codes of 9..12 bits, most significant bit first, 256 = clear,
// 257 = end of information, the code width growing one code early.
void unpack_lzw(const uint8_t* src, uint64_t n, std::vector<uint8_t>& out, uint64_t want) {
  out.clear();
  out.reserve(want);
  std::vector<uint16_t> prefix(4096);
  std::vector<uint8_t> suffix(4096), first(4096);
  std::vector<uint16_t> length(4096);
  for (int i = 0; i < 256; ++i) {
    prefix[i] = 0xFFFF;
    suffix[i] = first[i] = (uint8_t)i;
    length[i] = 1;
  }
  uint32_t width = 9, next = 258;
  int prev = -1;
  uint64_t bitpos = 0;
  std::vector<uint8_t> stack;
  const uint64_t nbits = n * 8;
  while (bitpos + width <= nbits) {
    uint32_t code = 0;
    for (uint32_t k = 0; k < width; ++k, ++bitpos)
      code = (code << 1) | ((src[bitpos >> 3] >> (7 - (bitpos & 7))) & 1u);
    if (code == 257) break;
    if (code == 256) {
      width = 9;
      next = 258;
      prev = -1;
      continue;
    }
    uint32_t cur = code;
    uint8_t head;
    if (code < next) {
      head = first[code];
    } else if (code == next && prev >= 0) {
      head = first[prev];
    } else {
      throw TiffFormatError("invalid LZW code");
    }
    if (prev >= 0 && next < 4096) {  // the new entry: prev + head
      prefix[next] = (uint16_t)prev;
      suffix[next] = head;
      first[next] = first[prev];
      length[next] = (uint16_t)(length[prev] + 1);
      ++next;
      if (next + 1 >= (1u << width) && width < 12) ++width;
    }
    stack.clear();
    for (uint32_t c = cur; c != 0xFFFF; c = prefix[c]) stack.push_back(suffix[c]);
    out.insert(out.end(), stack.rbegin(), stack.rend());
    prev = (int)code;
    if (out.size() >= want) break;
  }
}

void unpack_deflate(const uint8_t* src, uint64_t n, std::vector<uint8_t>& out, uint64_t want) {
  out.assign(want, 0);
  z_stream zs;
  std::memset(&zs, 0, sizeof zs);
  if (inflateInit(&zs) != Z_OK) throw TiffFormatError("zlib initialisation failed");
  zs.next_in = const_cast<Bytef*>(src);
  zs.avail_in = (uInt)n;
  zs.next_out = out.data();
  zs.avail_out = (uInt)want;
  const int rc = inflate(&zs, Z_FINISH);
  const uint64_t got = want - zs.avail_out;
  inflateEnd(&zs);
  if (rc != Z_STREAM_END && rc != Z_OK && rc != Z_BUF_ERROR)
    throw TiffFormatError("corrupt Deflate data");
  out.resize(got);
}

void decompress(const Page& p, const uint8_t* src, uint64_t n, std::vector<uint8_t>& out,
                uint64_t want) {
  switch (p.compression) {
    case 1: out.assign(src, src + std::min(n, want)); break;
    case 32773: unpack_bits(src, n, out, want); break;
    case 5: unpack_lzw(src, n, out, want); break;
    case 8: case 32946: unpack_deflate(src, n, out, want); break;
    default: throw TiffFormatError(fmtf("unsupported compression method %u", p.compression));
  }
  if (out.size() < want) throw TiffFormatError("compressed data ended early");
  out.resize(want);
}

// The tiff crate's default decoder limits (Limits::default of tiff 0.11):
// the decoded buffer of one page, and the compressed data of one strip or
// tile.  Larger pages fail to decode, as they do in the reference.
constexpr uint64_t kDecodingBufferLimit = (uint64_t)256 << 20;
constexpr uint64_t kIntermediateBufferLimit = (uint64_t)128 << 20;

uint64_t checked_mul(uint64_t a, uint64_t b) {
  const unsigned __int128 r = (unsigned __int128)a * b;
  if (r >> 64) throw TiffFormatError("size overflow");
  return (uint64_t)r;
}

// Decodes the page's samples (width*height bytes) into dst.  Every size is
// checked before anything is allocated or written: the page and each strip
// or tile within the decoder limits, products without wrap-around.
void decode_page(const TiffFile& f, const Page& p, uint8_t* dst) {
  if (checked_mul(p.width, p.height) > kDecodingBufferLimit)
    throw TiffFormatError("page exceeds the decoding buffer limit");
  if (p.sample_format != 1) throw TiffFormatError("signed/float samples");
  if (p.predictor != 1 && p.predictor != 2) throw TiffFormatError("unsupported predictor");
  std::vector<uint8_t> buf;
  auto unpredict = [&](uint8_t* rows, uint64_t nrows, uint64_t rowlen) {
    if (p.predictor != 2) return;
    for (uint64_t r = 0; r < nrows; ++r) {
      uint8_t* q = rows + r * rowlen;
      for (uint64_t x = 1; x < rowlen; ++x) q[x] = (uint8_t)(q[x] + q[x - 1]);
    }
  };
  if (!p.tiled) {
    const uint64_t rps = p.rows_per_strip ? p.rows_per_strip : p.height;
    const uint64_t strips = p.height ? (p.height + rps - 1) / rps : 0;
    if (p.offsets.size() < strips || p.counts.size() < strips)
      throw TiffFormatError("missing strip offsets or byte counts");
    for (uint64_t s = 0; s < strips; ++s) {
      const uint64_t rows = std::min(rps, p.height - s * rps);
      const uint64_t want = rows * p.width;
      if (p.counts[s] > kIntermediateBufferLimit)
        throw TiffFormatError("strip exceeds the intermediate buffer limit");
      decompress(p, f.bytes(p.offsets[s], p.counts[s]), p.counts[s], buf, want);
      unpredict(buf.data(), rows, p.width);
      std::memcpy(dst + s * rps * p.width, buf.data(), want);
    }
  } else {
    if (!p.tile_w || !p.tile_h) throw TiffFormatError("tile dimensions missing");
    if (p.tile_w > 0xFFFFFFFFull || p.tile_h > 0xFFFFFFFFull)
      throw TiffFormatError("tile dimensions out of range");
    const uint64_t tile_len = checked_mul(p.tile_w, p.tile_h);
    if (tile_len > kDecodingBufferLimit)
      throw TiffFormatError("tile exceeds the decoding buffer limit");
    const uint64_t across = (p.width + p.tile_w - 1) / p.tile_w;
    const uint64_t down = (p.height + p.tile_h - 1) / p.tile_h;
    const uint64_t tiles = checked_mul(across, down);
    if (p.offsets.size() < tiles || p.counts.size() < tiles)
      throw TiffFormatError("missing tile offsets or byte counts");
    for (uint64_t ty = 0; ty < down; ++ty)
      for (uint64_t tx = 0; tx < across; ++tx) {
        const uint64_t t = ty * across + tx;
        if (p.counts[t] > kIntermediateBufferLimit)
          throw TiffFormatError("tile exceeds the intermediate buffer limit");
        decompress(p, f.bytes(p.offsets[t], p.counts[t]), p.counts[t], buf, tile_len);
        unpredict(buf.data(), p.tile_h, p.tile_w);
        const uint64_t x0 = tx * p.tile_w, y0 = ty * p.tile_h;
        const uint64_t cw = std::min(p.tile_w, p.width - x0), ch = std::min(p.tile_h, p.height - y0);
        for (uint64_t r = 0; r < ch; ++r)
          std::memcpy(dst + (y0 + r) * p.width + x0, buf.data() + r * p.tile_w, cw);
      }
  }
}

// The directory: every page, or the error that ends it (at page index `stop`).
struct Directory {
  std::vector<Page> pages;
  std::string chain_error;  // "advancing to the next page of ..." after pages.size() pages
};

Directory read_directory(const TiffFile& f, const std::string& path) {
  Directory dir;
  uint64_t off = f.first_ifd();
  std::vector<uint64_t> seen;
  while (true) {
    Ifd ifd;
    uint64_t next = 0;
    try {
      if (std::find(seen.begin(), seen.end(), off) != seen.end())
        throw TiffFormatError("IFD loop");
      seen.push_back(off);
      next = f.read_ifd(off, ifd);
    } catch (const TiffFormatError&) {
      if (dir.pages.empty()) throw Error("initializing TIFF decoder for " + path);
      dir.chain_error = "advancing to the next page of " + path;
      return dir;
    }
    dir.pages.push_back(describe(f, ifd));
    if (next == 0) return dir;
    off = next;
  }
}

const char* expected_text(TiffSamples a) {
  return a == TiffSamples::ZeroOr255 ? "0 or 255" : "0/1 or 0/255";
}

// Checks one page in the reference's order (src/tiff_io.rs:50-128), decodes,
// validates and binarises it into dst (width*height bytes).  Returns the
// error message, or "" when the page is good.
std::string process_page(const TiffFile& f, const Page& p, uint64_t index, const Page& first,
                         const std::string& path, uint8_t threshold, TiffSamples accepted,
                         uint8_t* dst, std::vector<uint8_t>& scratch) {
  if (!p.dims_ok) return "reading dimensions from " + path;
  if (!p.color_ok) return "reading color type from " + path;
  if (p.color_name != "Gray(8)")
    return path + " must contain 8-bit grayscale pages; found " + p.color_name;
  if (index > 0 && (p.width != first.width || p.height != first.height))
    return fmtf("TIFF pages in %s do not have consistent dimensions: expected %llux%llu, got "
                "%llux%llu",
                path.c_str(), (unsigned long long)first.width, (unsigned long long)first.height,
                (unsigned long long)p.width, (unsigned long long)p.height);
  const uint64_t len = p.width * p.height;
  uint8_t* raw = dst;
  (void)scratch;
  try {
    if (p.sample_format != 1)
      return path + " decoded to an unsupported pixel type; expected unsigned 8-bit samples";
    decode_page(f, p, raw);
  } catch (const std::exception&) {  // format errors, and allocation failures of the decoder
    return fmtf("decoding page %llu of %s", (unsigned long long)(index + 1), path.c_str());
  }
  if (accepted != TiffSamples::Any) {
    for (uint64_t i = 0; i < len; ++i) {
      const uint8_t v = raw[i];
      const bool ok = accepted == TiffSamples::ZeroOr255 ? (v == 0 || v == 255)
                                                          : (v == 0 || v == 1 || v == 255);
      if (!ok)
        return fmtf("%s is not a valid binary image: page %llu, sample %llu has value %u "
                    "(expected %s)",
                    path.c_str(), (unsigned long long)(index + 1), (unsigned long long)i, v,
                    expected_text(accepted));
    }
  }
  for (uint64_t i = 0; i < len; ++i) raw[i] = raw[i] > threshold ? 1 : 0;
  return "";
}

// BinaryVolume::new's checks on the result (src/volume.rs:12-41).
void check_volume(uint64_t w, uint64_t h, uint64_t d) {
  if (w < 2 || h < 2 || d < 2)
    throw Error(fmtf("all dimensions must be at least 2 for 3-D morphometry; got %llux%llux%llu",
                     (unsigned long long)w, (unsigned long long)h, (unsigned long long)d));
}

int pool_size(int threads, uint64_t pages) {
  int n = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
  return (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)n, pages));
}

// Runs process_page over pages [lo, hi) on a pool; returns the first error in
// page order ("" if none).  dst(page) gives each page's destination.
template <class Dst>
std::string process_pages(const TiffFile& f, const Directory& dir, uint64_t lo, uint64_t hi,
                          const std::string& path, uint8_t threshold, TiffSamples accepted,
                          int threads, Dst dst) {
  std::vector<std::string> errors(hi - lo);
  std::atomic<uint64_t> next(lo);
  std::atomic<uint64_t> first_bad(hi);
  auto worker = [&]() {
    std::vector<uint8_t> scratch;
    for (uint64_t k = next++; k < hi; k = next++) {
      if (k > first_bad.load()) continue;  // an earlier page already failed
      std::string e = process_page(f, dir.pages[k], k, dir.pages[0], path, threshold, accepted,
                                   dst(k), scratch);
      if (!e.empty()) {
        errors[k - lo] = e;
        uint64_t cur = first_bad.load();
        while (k < cur && !first_bad.compare_exchange_weak(cur, k)) {
        }
      }
    }
  };
  const int n = pool_size(threads, hi - lo);
  std::vector<std::thread> pool;
  for (int i = 1; i < n; ++i) pool.emplace_back(worker);
  worker();
  for (auto& t : pool) t.join();
  const uint64_t bad = first_bad.load();
  return bad < hi ? errors[bad - lo] : std::string();
}

// The output buffer must hold page_len * depth bytes (compared without
// wrap-around: the product can exceed 2^64 for a crafted directory).
void check_capacity(uint64_t capacity, uint64_t page_len, uint64_t depth) {
  const unsigned __int128 need = (unsigned __int128)page_len * depth;
  if (need > capacity) {
    const std::string voxels = (need >> 64) ? std::string("more than 2^64")
                                            : std::to_string((unsigned long long)need);
    throw Error(fmtf("output buffer holds %llu bytes, the volume has %s voxels",
                     (unsigned long long)capacity, voxels.c_str()));
  }
}

// Width and height of the first page and the number of pages.
TiffVolumeInfo volume_info(const Directory& dir) {
  TiffVolumeInfo info;
  if (!dir.pages.empty() && dir.pages[0].dims_ok) {
    info.width = dir.pages[0].width;
    info.height = dir.pages[0].height;
  }
  info.depth = dir.pages.size();
  return info;
}

}  // namespace

TiffVolumeInfo read_binary_tiff(const std::string& path, uint8_t threshold, TiffSamples accepted,
                                uint8_t* out, uint64_t capacity, int threads) {
  Mapped m(path);
  std::unique_ptr<TiffFile> f;
  try {
    f.reset(new TiffFile(m));
  } catch (const TiffFormatError&) {
    throw Error("initializing TIFF decoder for " + path);
  }
  const Directory dir = read_directory(*f, path);
  TiffVolumeInfo info = volume_info(dir);
  if (!out) return info;  // the size only
  const uint64_t page_len = info.width * info.height;  // fits: both below 2^32
  check_capacity(capacity, page_len, info.depth);
  const std::string err = process_pages(*f, dir, 0, dir.pages.size(), path, threshold, accepted,
                                        threads, [&](uint64_t k) { return out + k * page_len; });
  if (!err.empty()) throw Error(err);
  if (!dir.chain_error.empty()) throw Error(dir.chain_error);
  check_volume(info.width, info.height, info.depth);
  return info;
}

TiffVolumeInfo read_binary_tiff_device(Engine& e, const std::string& path, uint8_t threshold,
                                       TiffSamples accepted, uint8_t* d_out, uint64_t capacity,
                                       int threads, hipStream_t s) {
  Mapped m(path);
  std::unique_ptr<TiffFile> f;
  try {
    f.reset(new TiffFile(m));
  } catch (const TiffFormatError&) {
    throw Error("initializing TIFF decoder for " + path);
  }
  const Directory dir = read_directory(*f, path);
  TiffVolumeInfo info = volume_info(dir);
  const uint64_t page_len = info.width * info.height;  // fits: both below 2^32
  check_capacity(capacity, page_len, info.depth);
  // Two pinned staging buffers of up to 64 MiB (whole pages): the pool decodes
  // a batch into one while the other's copy to HBM runs.
  // (pages above the decoder's buffer limit fail before they are written:
  // their batches stage nothing)
  const bool decodable = page_len <= kDecodingBufferLimit;
  const uint64_t batch_pages =
      decodable ? std::max<uint64_t>(1, std::min<uint64_t>(dir.pages.size(), (64ull << 20) /
                                                            std::max<uint64_t>(page_len, 1)))
                : 1;
  const uint64_t stage_bytes = decodable ? std::max<uint64_t>(1, batch_pages * page_len) : 1;
  uint8_t* pinned[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};
  std::string err;
  try {
    for (int i = 0; i < 2; ++i) {
      BM_HIP(hipHostMalloc((void**)&pinned[i], stage_bytes, hipHostMallocDefault));
      BM_HIP(hipEventCreateWithFlags(&done[i], hipEventDisableTiming));
    }
    int slot = 0;
    bool used[2] = {false, false};
    for (uint64_t lo = 0; lo < dir.pages.size() && err.empty(); lo += batch_pages, slot ^= 1) {
      const uint64_t hi = std::min<uint64_t>(dir.pages.size(), lo + batch_pages);
      if (used[slot]) BM_HIP(hipEventSynchronize(done[slot]));  // its last copy has landed
      uint8_t* stage = pinned[slot];
      err = process_pages(*f, dir, lo, hi, path, threshold, accepted, threads,
                          [&](uint64_t k) { return decodable ? stage + (k - lo) * page_len : stage; });
      if (!err.empty()) break;
      BM_HIP(hipMemcpyAsync(d_out + lo * page_len, stage, (hi - lo) * page_len,
                            hipMemcpyHostToDevice, s));
      BM_HIP(hipEventRecord(done[slot], s));
      used[slot] = true;
    }
    BM_HIP(hipStreamSynchronize(s));
  } catch (...) {
    hipStreamSynchronize(s);
    for (int i = 0; i < 2; ++i) {
      if (pinned[i]) hipHostFree(pinned[i]);
      if (done[i]) hipEventDestroy(done[i]);
    }
    throw;
  }
  for (int i = 0; i < 2; ++i) {
    hipHostFree(pinned[i]);
    hipEventDestroy(done[i]);
  }
  if (!err.empty()) throw Error(err);
  if (!dir.chain_error.empty()) throw Error(dir.chain_error);
  check_volume(info.width, info.height, info.depth);
  return info;
}

// ---------------------------------------------------------------------------
// Calibration, src/tiff_io.rs:149-317.
namespace {

// description_value, src/tiff_io.rs:283-290: the first line "key=value".
bool description_value(const std::string& text, const std::string& key, std::string& out) {
  size_t start = 0;
  while (start <= text.size()) {
    size_t end = text.find('\n', start);
    if (end == std::string::npos) end = text.size();
    std::string line = text.substr(start, end - start);
    if (!line.empty() && line.back() == '\r') line.pop_back();  // str::lines
    // trim_matches('\0') then trim()
    size_t a = 0, b = line.size();
    while (a < b && line[a] == '\0') ++a;
    while (b > a && line[b - 1] == '\0') --b;
    auto ws = [](char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\f' || c == '\v'; };
    while (a < b && ws(line[a])) ++a;
    while (b > a && ws(line[b - 1])) --b;
    const std::string l = line.substr(a, b - a);
    if (l.compare(0, key.size(), key) == 0 && l.size() > key.size() && l[key.size()] == '=') {
      std::string v = l.substr(key.size() + 1);
      size_t c = 0, d = v.size();
      while (c < d && ws(v[c])) ++c;
      while (d > c && ws(v[d - 1])) --d;
      out = v.substr(c, d - c);
      return true;
    }
    if (end == text.size()) break;
    start = end + 1;
  }
  return false;
}

// unit_scale_um, src/tiff_io.rs:292-305.
bool unit_scale_um(const std::string& unit, double& out) {
  std::string u;
  size_t a = 0, b = unit.size();
  while (a < b && isspace((unsigned char)unit[a])) ++a;
  while (b > a && isspace((unsigned char)unit[b - 1])) --b;
  for (size_t i = a; i < b; ++i) {
    const unsigned char c = (unsigned char)unit[i];
    // U+00B5 MICRO SIGN (C2 B5) and U+03BC GREEK SMALL LETTER MU (CE BC) -> 'u'
    if (i + 1 < b && ((c == 0xC2 && (unsigned char)unit[i + 1] == 0xB5) ||
                      (c == 0xCE && (unsigned char)unit[i + 1] == 0xBC))) {
      u += 'u';
      ++i;
    } else {
      u += (char)(c < 0x80 ? tolower(c) : c);
    }
  }
  if (u == "nm" || u == "nanometer" || u == "nanometers") out = 0.001;
  else if (u == "um" || u == "micron" || u == "microns" || u == "micrometer" || u == "micrometers") out = 1.0;
  else if (u == "mm" || u == "millimeter" || u == "millimeters") out = 1000.0;
  else if (u == "cm" || u == "centimeter" || u == "centimeters") out = 10000.0;
  else if (u == "in" || u == "inch" || u == "inches") out = 25400.0;
  else return false;
  return true;
}

// str::parse::<f64>: an optional sign, then digits with an optional point and
// exponent, or inf/infinity/nan (any case); nothing else, no blanks.
bool parse_f64(const std::string& s, double& out) {
  if (s.empty()) return false;
  size_t i = (s[0] == '+' || s[0] == '-') ? 1 : 0;
  std::string rest = s.substr(i);
  std::string low;
  for (char c : rest) low += (char)tolower((unsigned char)c);
  if (low == "inf" || low == "infinity" || low == "nan") {
    out = low == "nan" ? NAN : (s[0] == '-' ? -INFINITY : INFINITY);
    return true;
  }
  bool digits = false, dot = false;
  size_t j = i;
  for (; j < s.size(); ++j) {
    if (isdigit((unsigned char)s[j])) digits = true;
    else if (s[j] == '.' && !dot) dot = true;
    else break;
  }
  if (!digits) return false;
  if (j < s.size() && (s[j] == 'e' || s[j] == 'E')) {
    size_t k = j + 1;
    if (k < s.size() && (s[k] == '+' || s[k] == '-')) ++k;
    size_t e0 = k;
    while (k < s.size() && isdigit((unsigned char)s[k])) ++k;
    if (k == e0) return false;
    j = k;
  }
  if (j != s.size()) return false;
  out = std::strtod(s.c_str(), nullptr);
  return true;
}

}  // namespace

void read_isotropic_spacing_um(const std::string& path, double out[3]) {
  Mapped m(path);
  std::unique_ptr<TiffFile> f;
  Ifd ifd;
  try {
    f.reset(new TiffFile(m));
    f->read_ifd(f->first_ifd(), ifd);
  } catch (const TiffFormatError&) {
    throw Error("initializing TIFF decoder for " + path);
  }
  std::string description;
  bool has_description = f->ascii(ifd, kTagDescription, description);
  double named_unit_um = 0.0;
  std::string unit;
  const bool has_named_unit = has_description && description_value(description, "unit", unit) &&
                              unit_scale_um(unit, named_unit_um);
  uint64_t code = 0;
  bool has_code = false;
  try {
    has_code = f->uint1(ifd, kTagResUnit, code);
  } catch (const TiffFormatError&) {
    throw Error("reading ResolutionUnit from " + path);
  }
  if (has_code && code > 0xFFFF) throw Error("reading ResolutionUnit from " + path);
  bool has_unit = false;
  double unit_um = 0.0;
  if (has_code && code == 2) {
    has_unit = true;
    unit_um = 25400.0;
  } else if (has_code && code == 3) {
    has_unit = true;
    unit_um = 10000.0;
  } else if (!has_code || code == 1) {
    has_unit = has_named_unit;
    unit_um = named_unit_um;
  } else {
    throw Error(fmtf("%s has unsupported TIFF ResolutionUnit code %llu", path.c_str(),
                     (unsigned long long)code));
  }
  double xr = 0, yr = 0;
  const bool hx = f->f64(ifd, kTagXRes, xr), hy = f->f64(ifd, kTagYRes, yr);
  bool has_xy = false;
  double sx = 0, sy = 0;
  if (hx && hy && has_unit) {  // isotropic_spacing_from_resolution, :250-281
    if (!std::isfinite(xr) || !std::isfinite(yr) || xr <= 0.0 || yr <= 0.0)
      throw Error(path + " has invalid X/Y resolution values (" + display_f64(xr) + ", " +
                  display_f64(yr) + ")");
    sx = unit_um / xr;
    sy = unit_um / yr;
    const double mean = 0.5 * (sx + sy);
    if (std::fabs(sx - sy) / mean > 0.01)
      throw Error(path + " is not isotropic in XY: x spacing=" + display_f64(sx) +
                  " \xC2\xB5m, y spacing=" + display_f64(sy) + " \xC2\xB5m");
    has_xy = true;
  }
  bool has_z = false;
  double z = 0;
  std::string value;
  if (has_description && description_value(description, "spacing", value)) {
    double in_units = 0;
    if (!parse_f64(value, in_units)) throw Error("parsing spacing metadata in " + path);
    z = in_units * (has_named_unit ? named_unit_um : 1.0);
    if (!std::isfinite(z) || z <= 0.0)
      throw Error("spacing in " + path + " must be finite and positive; got " + display_f64(z) +
                  " \xC2\xB5m");
    has_z = true;
  }
  // combine_isotropic_spacing, :220-248
  if (!has_xy)
    throw Error("x/y spacing was not found in the TIFF resolution metadata of " + path);
  if (!has_z)
    throw Error("z spacing was not found in the 'spacing=' ImageDescription metadata of " + path +
                "; use --spacing-dir if the binarized TIFF lost calibration");
  const double v[3] = {sx, sy, z};
  double mx = -INFINITY, mn = INFINITY;
  for (double t : v) {
    mx = std::fmax(mx, t);
    mn = std::fmin(mn, t);
  }
  if ((mx - mn) / mx > 1.0e-6)
    throw Error(path + " is not isotropic in 3D: x spacing=" + display_f64(sx) +
                " \xC2\xB5m, y spacing=" + display_f64(sy) + " \xC2\xB5m, z spacing=" +
                display_f64(z) + " \xC2\xB5m");
  out[0] = sx;
  out[1] = sy;
  out[2] = z;
}

}  // namespace bm
