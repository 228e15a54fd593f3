// Byte-string format for H-matrices, instruction programs and factor
// arenas (files, caches, the wire).  The reference has no persistent format.
//
//   string  := header section*
//   header  := magic "PPHL" | u32 version | u32 n_sections | u32 0
//   section := u32 tag | u32 0 | u64 length | payload[length] | pad to 8 |
//              u64 FNV-1a 64 of the payload
//
// All integers little-endian; doubles are IEEE-754 binary64 (float32 where a
// factor arena is kept in FP32).  Tags (four ASCII bytes read as a
// little-endian u32): TREE, ARNA, PROG, FACT.  Readers check every length
// against the bytes that remain and every checksum before they look inside a
// section, and throw wire_error on any mismatch.
#pragma once

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace ppw {

constexpr uint32_t kVersion = 1;

constexpr uint32_t tag(const char (&s)[5]) {
  return (uint32_t)(uint8_t)s[0] | (uint32_t)(uint8_t)s[1] << 8 | (uint32_t)(uint8_t)s[2] << 16 |
         (uint32_t)(uint8_t)s[3] << 24;
}
constexpr uint32_t kMagic = tag("PPHL");
constexpr uint32_t kTree = tag("TREE"), kArena = tag("ARNA"), kProgram = tag("PROG"), kFactor = tag("FACT");

struct wire_error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline uint64_t fnv1a64(const void* p, size_t n, uint64_t h = 1469598103934665603ull) {
  const uint8_t* b = static_cast<const uint8_t*>(p);
  for (size_t i = 0; i < n; ++i) {
    h ^= b[i];
    h *= 1099511628211ull;
  }
  return h;
}

class writer {
public:
  std::vector<uint8_t> bytes;
  template <class T>
  void put(T v) {
    const size_t at = bytes.size();
    bytes.resize(at + sizeof(T));
    std::memcpy(bytes.data() + at, &v, sizeof(T));
  }
  void put_bytes(const void* p, size_t n) {
    const size_t at = bytes.size();
    bytes.resize(at + n);
    if (n) std::memcpy(bytes.data() + at, p, n);
  }
  template <class T>
  void put_vec(const std::vector<T>& v) {
    put_bytes(v.data(), sizeof(T) * v.size());
  }
};

class reader {
  const uint8_t* p;
  size_t n, at = 0;

public:
  reader(const uint8_t* data, size_t len) : p(data), n(len) {}
  size_t left() const { return n - at; }
  const uint8_t* here() const { return p + at; }
  void need(size_t k) const {
    if (k > n - at) throw wire_error("truncated section");
  }
  template <class T>
  T get() {
    need(sizeof(T));
    T v;
    std::memcpy(&v, p + at, sizeof(T));
    at += sizeof(T);
    return v;
  }
  // a count that must fit in what remains at `unit` bytes per item
  int64_t count(int64_t unit) {
    const int64_t c = get<int64_t>();
    if (c < 0 || (unit > 0 && (uint64_t)c > left() / (uint64_t)unit)) throw wire_error("bad count");
    return c;
  }
  void get_bytes(void* dst, size_t k) {
    need(k);
    if (k) std::memcpy(dst, p + at, k);
    at += k;
  }
  void skip(size_t k) {
    need(k);
    at += k;
  }
  template <class T>
  std::vector<T> get_vec(int64_t c) {
    need(sizeof(T) * (size_t)c);
    std::vector<T> v((size_t)c);
    get_bytes(v.data(), sizeof(T) * (size_t)c);
    return v;
  }
};

struct section {
  uint32_t tag;
  const uint8_t* data;
  uint64_t length;
};

// Container: sections are appended, then finish() prepends the header.
class container {
  std::vector<std::pair<uint32_t, std::vector<uint8_t>>> secs;

public:
  void add(uint32_t t, std::vector<uint8_t>&& payload) { secs.emplace_back(t, std::move(payload)); }
  uint64_t size() const {
    uint64_t s = 16;
    for (const auto& x : secs) s += 16 + (x.second.size() + 7) / 8 * 8 + 8;
    return s;
  }
  // writes size() bytes to out
  void write(uint8_t* out) const {
    writer w;
    w.put<uint32_t>(kMagic);
    w.put<uint32_t>(kVersion);
    w.put<uint32_t>((uint32_t)secs.size());
    w.put<uint32_t>(0);
    std::memcpy(out, w.bytes.data(), 16);
    uint64_t at = 16;
    for (const auto& x : secs) {
      writer h;
      h.put<uint32_t>(x.first);
      h.put<uint32_t>(0);
      h.put<uint64_t>(x.second.size());
      std::memcpy(out + at, h.bytes.data(), 16);
      at += 16;
      if (!x.second.empty()) std::memcpy(out + at, x.second.data(), x.second.size());
      const uint64_t padded = (x.second.size() + 7) / 8 * 8;
      std::memset(out + at + x.second.size(), 0, padded - x.second.size());
      at += padded;
      const uint64_t sum = fnv1a64(x.second.data(), x.second.size());
      std::memcpy(out + at, &sum, 8);
      at += 8;
    }
  }
};

// Splits a string into its sections, checking the header, every length and
// every checksum.
inline std::vector<section> parse(const uint8_t* buf, uint64_t len) {
  if (buf == nullptr) throw wire_error("null buffer");
  reader r(buf, len);
  if (len < 16 || r.get<uint32_t>() != kMagic) throw wire_error("not a PPHL string (bad magic)");
  const uint32_t ver = r.get<uint32_t>();
  if (ver != kVersion) throw wire_error("unsupported format version " + std::to_string(ver));
  const uint32_t ns = r.get<uint32_t>();
  (void)r.get<uint32_t>();
  if (ns > r.left() / 24) throw wire_error("bad section count");
  std::vector<section> out;
  for (uint32_t i = 0; i < ns; ++i) {
    section s;
    s.tag = r.get<uint32_t>();
    (void)r.get<uint32_t>();
    s.length = r.get<uint64_t>();
    if (s.length > r.left()) throw wire_error("section longer than the string");
    const uint64_t padded = (s.length + 7) / 8 * 8;
    if (padded + 8 > r.left()) throw wire_error("section longer than the string");
    s.data = r.here();
    r.skip(padded);
    const uint64_t sum = r.get<uint64_t>();
    if (sum != fnv1a64(s.data, s.length)) throw wire_error("section checksum mismatch");
    out.push_back(s);
  }
  if (r.left() != 0) throw wire_error("trailing bytes after the last section");
  return out;
}

inline const section* find(const std::vector<section>& secs, uint32_t t) {
  for (const section& s : secs)
    if (s.tag == t) return &s;
  return nullptr;
}

}  // namespace ppw
