// Basis manifests: the versioned zip archive of ordered basis states that
// QMBED writes and reads as a fixture and exchange format
// (src/archive.rs:333-362 save_basis_zip, :365-416 load_basis_zip).
//
// Entries (deflated): basis_archive_version.npy (a one-element uint8 NPY array
// holding 1), basis_manifest.tsv ("width_bits\t<w>" then one "state\t<decimal>"
// row per basis row, in row order), metadata.tsv ("<key>\t<value>" rows in key
// order). States are arbitrary-width nonnegative integers written in decimal;
// the archive carries no executable content.
//
// The zip container is written here directly (local headers, central
// directory, end record; raw deflate and CRC-32 from zlib). Reading accepts
// stored and deflated entries, and the zip64 extra field of archives written
// by other tools.
#include <zlib.h>

#include <cctype>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <unordered_set>
#include <vector>

#include "archive.hpp"

namespace qmbed {
namespace {

constexpr uint8_t kBasisArchiveVersion = 1;  // BASIS_ARCHIVE_VERSION, src/archive.rs:18

Error archive_error(const std::string& s) { return Error(ErrorKind::Archive, "archive error: " + s); }

std::string os_error(int err) { return std::string(std::strerror(err)) + " (os error " + std::to_string(err) + ")"; }

// ------------------------------------------------------------- decimals --
// A state's decimal text, parsed as BigUint::parse_bytes(text, 10) accepts it
// (num-bigint's from_str_radix: an optional leading '+', then digits with '_'
// separators that may not lead), into its canonical form and its bit length.
bool parse_decimal(const std::string& text, std::string& canonical, size_t& bits) {
  size_t i = 0;
  if (i < text.size() && text[i] == '+') ++i;
  if (i < text.size() && text[i] == '_') return false;
  std::string digits;
  for (; i < text.size(); ++i) {
    const char c = text[i];
    if (c == '_') continue;
    if (c < '0' || c > '9') return false;
    digits.push_back(c);
  }
  if (digits.empty()) return false;
  size_t lead = 0;
  while (lead + 1 < digits.size() && digits[lead] == '0') ++lead;
  canonical = digits.substr(lead);
  std::vector<uint32_t> limbs;  // little-endian base 2^32
  for (char c : canonical) {
    uint64_t carry = uint64_t(c - '0');
    for (auto& l : limbs) {
      const uint64_t v = uint64_t(l) * 10 + carry;
      l = uint32_t(v);
      carry = v >> 32;
    }
    if (carry) limbs.push_back(uint32_t(carry));
  }
  bits = limbs.empty() ? 0 : 32 * (limbs.size() - 1) + size_t(32 - __builtin_clz(limbs.back()));
  return true;
}

// ErasedState::from_decimal (src/basis/mod.rs:5139-5169): the value must fit
// the width, and the width must name a storage class (1 to 16384 bits).
std::string checked_state(size_t width_bits, const std::string& text) {
  std::string canonical;
  size_t bits = 0;
  if (!parse_decimal(text, canonical, bits))
    throw invalid_options("basis state " + debug_string(text) + " is not a nonnegative decimal integer");
  if (bits > width_bits)
    throw unsupported_backend("integer needs " + std::to_string(bits) + " bits but the requested state width is " +
                              std::to_string(width_bits));
  if (width_bits == 0) throw invalid_sector("a basis state needs at least one bit");
  if (width_bits > 16384) throw unsupported_backend("basis state requires more than 16384 bits");
  return canonical;
}

// validate_metadata, src/archive.rs:222-244
void check_metadata(const std::string& key, const std::string& value) {
  bool ok = !key.empty();
  for (unsigned char c : key)
    if (!(std::isalnum(c) || c == '_' || c == '.' || c == '-') || c >= 0x80) ok = false;
  if (!ok) throw archive_error("archive metadata keys may contain only ASCII letters, digits, `_`, `-`, and `.`");
  for (char c : value)
    if (c == '\t' || c == '\r' || c == '\n') throw archive_error("archive metadata values may not contain tabs or newlines");
}

// BasisArchive::new, src/archive.rs:261-280
void check_states(size_t width_bits, const std::vector<std::string>& states) {
  if (width_bits == 0) throw archive_error("basis states must share one positive logical width");
  std::unordered_set<std::string> seen;
  seen.reserve(states.size() * 2);
  for (const auto& s : states)
    if (!seen.insert(s).second) throw archive_error("basis archives require unique physical states");
}

// ------------------------------------------------------------------ NPY --
// Version 1.0 header, padded with spaces and a newline to a multiple of 16
// bytes (src/archive.rs:24-55).
std::string npy_u8(const std::vector<uint8_t>& values) {
  const std::string dict = "{'descr': '|u1', 'fortran_order': False, 'shape': (" + std::to_string(values.size()) + ",), }";
  const size_t pad = (16 - (10 + dict.size() + 1) % 16) % 16;
  const size_t hlen = dict.size() + pad + 1;
  std::string out = "\x93NUMPY";
  out.push_back('\x01');
  out.push_back('\x00');
  out.push_back(char(hlen & 0xFF));
  out.push_back(char(hlen >> 8));
  out += dict;
  out.append(pad, ' ');
  out.push_back('\n');
  out.append(reinterpret_cast<const char*>(values.data()), values.size());
  return out;
}

std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
  while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
  return s.substr(a, b - a);
}

// parse_npy_header + read_version (src/archive.rs:99-180): the descriptor and
// shape read from the header dictionary, then the single uint8 element.
uint8_t read_version(const std::string& data) {
  if (data.size() < 6 || data.compare(0, 6, "\x93NUMPY") != 0) throw archive_error("invalid NPY magic");
  if (data.size() < 8) throw archive_error("failed to fill whole buffer");
  size_t hlen = 0, pos = 0;
  const auto byte = [&](size_t i) { return size_t(static_cast<unsigned char>(data[i])); };
  if (byte(6) == 1) {
    if (data.size() < 10) throw archive_error("failed to fill whole buffer");
    hlen = byte(8) | (byte(9) << 8);
    pos = 10;
  } else if (byte(6) == 2 || byte(6) == 3) {
    if (data.size() < 12) throw archive_error("failed to fill whole buffer");
    hlen = byte(8) | (byte(9) << 8) | (byte(10) << 16) | (byte(11) << 24);
    pos = 12;
  } else {
    throw archive_error("unsupported NPY version");
  }
  if (data.size() < pos + hlen) throw archive_error("failed to fill whole buffer");
  const std::string header = trim(data.substr(pos, hlen));
  pos += hlen;
  if (header.find("'fortran_order': True") != std::string::npos ||
      header.find("\"fortran_order\": True") != std::string::npos)
    throw archive_error("Fortran-ordered NPY matrices are not supported");
  const char* dm = header.find("'descr':") != std::string::npos ? "'descr':" : "\"descr\":";
  size_t at = header.find(dm);
  if (at == std::string::npos) throw archive_error("NPY descriptor is missing");
  std::string tail = header.substr(at + std::strlen(dm));
  size_t k = 0;
  while (k < tail.size() && std::isspace(static_cast<unsigned char>(tail[k]))) ++k;
  if (k >= tail.size() || (tail[k] != '\'' && tail[k] != '"')) throw archive_error("invalid NPY descriptor");
  const size_t close = tail.find(tail[k], k + 1);
  if (close == std::string::npos) throw archive_error("invalid NPY descriptor");
  const std::string descr = tail.substr(k + 1, close - k - 1);
  const char* sm = header.find("'shape':") != std::string::npos ? "'shape':" : "\"shape\":";
  at = header.find(sm);
  if (at == std::string::npos) throw archive_error("NPY shape is missing");
  const std::string stail = header.substr(at + std::strlen(sm));
  const size_t open = stail.find('(');
  const size_t shut = open == std::string::npos ? std::string::npos : stail.find(')', open + 1);
  if (shut == std::string::npos) throw archive_error("invalid NPY shape");
  std::vector<std::string> dims;
  std::string cur;
  for (char c : stail.substr(open + 1, shut - open - 1) + ",") {
    if (c == ',') {
      const std::string d = trim(cur);
      if (!d.empty()) dims.push_back(d);
      cur.clear();
    } else {
      cur.push_back(c);
    }
  }
  std::vector<size_t> shape;
  for (const auto& d : dims) {
    for (char c : d)
      if (c < '0' || c > '9') throw archive_error("invalid digit found in string");
    shape.push_back(size_t(std::stoull(d)));
  }
  if (descr != "|u1" || shape != std::vector<size_t>{1}) throw archive_error("invalid archive-version array");
  if (data.size() < pos + 1) throw archive_error("failed to fill whole buffer");
  return static_cast<uint8_t>(data[pos]);
}

// ------------------------------------------------------------------ zip --
void put16(std::string& o, uint32_t v) {
  o.push_back(char(v & 0xFF));
  o.push_back(char((v >> 8) & 0xFF));
}
void put32(std::string& o, uint32_t v) {
  put16(o, v & 0xFFFF);
  put16(o, v >> 16);
}
uint32_t get16(const std::string& d, size_t p) {
  return uint32_t(static_cast<unsigned char>(d[p])) | (uint32_t(static_cast<unsigned char>(d[p + 1])) << 8);
}
uint32_t get32(const std::string& d, size_t p) { return get16(d, p) | (get16(d, p + 2) << 16); }
uint64_t get64(const std::string& d, size_t p) { return uint64_t(get32(d, p)) | (uint64_t(get32(d, p + 4)) << 32); }

std::string deflate_raw(const std::string& in) {
  z_stream z{};
  if (deflateInit2(&z, Z_DEFAULT_COMPRESSION, Z_DEFLATED, -15, 8, Z_DEFAULT_STRATEGY) != Z_OK)
    throw archive_error("deflate initialisation failed");
  std::string out(deflateBound(&z, uLong(in.size())), '\0');
  z.next_in = reinterpret_cast<Bytef*>(const_cast<char*>(in.data()));
  z.avail_in = uInt(in.size());
  z.next_out = reinterpret_cast<Bytef*>(&out[0]);
  z.avail_out = uInt(out.size());
  const int rc = deflate(&z, Z_FINISH);
  const size_t n = z.total_out;
  deflateEnd(&z);
  if (rc != Z_STREAM_END) throw archive_error("deflate failed");
  out.resize(n);
  return out;
}

std::string inflate_raw(const char* in, size_t n, size_t expect) {
  z_stream z{};
  if (inflateInit2(&z, -15) != Z_OK) throw archive_error("inflate initialisation failed");
  // one spare byte: zlib refuses a null output pointer (empty entries), and
  // a stream longer than its recorded size then shows as a size mismatch
  std::string out(expect + 1, '\0');
  z.next_in = reinterpret_cast<Bytef*>(const_cast<char*>(in));
  z.avail_in = uInt(n);
  z.next_out = reinterpret_cast<Bytef*>(&out[0]);
  z.avail_out = uInt(out.size());
  const int rc = inflate(&z, Z_FINISH);
  const size_t got = z.total_out;
  inflateEnd(&z);
  if (rc != Z_STREAM_END || got != expect) throw archive_error("invalid Zip archive: corrupt deflate stream");
  out.resize(expect);
  return out;
}

struct ZipEntry {
  std::string name, data;
};

// One deflated entry per file; sizes are known before each local header, so
// no data descriptors. Archives beyond the 4 GiB of classic zip fields are
// refused.
void write_zip(const std::string& path, const std::vector<ZipEntry>& entries) {
  std::string out, cd;
  for (const auto& e : entries) {
    const std::string packed = deflate_raw(e.data);
    const uint32_t crc = uint32_t(crc32(0L, reinterpret_cast<const Bytef*>(e.data.data()), uInt(e.data.size())));
    if (e.data.size() >= 0xFFFFFFFFull || packed.size() >= 0xFFFFFFFFull || out.size() >= 0xFFFFFFFFull)
      throw archive_error("large file option has not been set");
    const uint32_t offset = uint32_t(out.size());
    put32(out, 0x04034b50);
    put16(out, 20);      // version needed: deflate
    put16(out, 0x0800);  // UTF-8 names
    put16(out, 8);
    put16(out, 0);       // 00:00:00
    put16(out, 0x0021);  // 1980-01-01
    put32(out, crc);
    put32(out, uint32_t(packed.size()));
    put32(out, uint32_t(e.data.size()));
    put16(out, uint32_t(e.name.size()));
    put16(out, 0);
    out += e.name;
    out += packed;
    put32(cd, 0x02014b50);
    put16(cd, (3 << 8) | 20);  // made by: Unix, 2.0
    put16(cd, 20);
    put16(cd, 0x0800);
    put16(cd, 8);
    put16(cd, 0);
    put16(cd, 0x0021);
    put32(cd, crc);
    put32(cd, uint32_t(packed.size()));
    put32(cd, uint32_t(e.data.size()));
    put16(cd, uint32_t(e.name.size()));
    put16(cd, 0);
    put16(cd, 0);
    put16(cd, 0);
    put16(cd, 0);
    put32(cd, 0100644u << 16);  // regular file, rw-r--r--
    put32(cd, offset);
    cd += e.name;
  }
  if (out.size() + cd.size() >= 0xFFFFFFFFull) throw archive_error("large file option has not been set");
  const uint32_t cd_offset = uint32_t(out.size());
  out += cd;
  put32(out, 0x06054b50);
  put16(out, 0);
  put16(out, 0);
  put16(out, uint32_t(entries.size()));
  put16(out, uint32_t(entries.size()));
  put32(out, uint32_t(cd.size()));
  put32(out, cd_offset);
  put16(out, 0);
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw archive_error(os_error(errno));
  const size_t w = std::fwrite(out.data(), 1, out.size(), f);
  const int err = errno;
  if (std::fclose(f) != 0 || w != out.size()) throw archive_error(os_error(err ? err : EIO));
}

class ZipReader {
 public:
  explicit ZipReader(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw archive_error(os_error(errno));
    char buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) data_.append(buf, n);
    std::fclose(f);
    index();
  }
  bool has(const std::string& name) const { return dir_.count(name) != 0; }
  std::string read(const std::string& name) const {
    const auto it = dir_.find(name);
    if (it == dir_.end()) throw archive_error("specified file not found in archive");
    const Dir& d = it->second;
    if (d.local + 30 > data_.size() || get32(data_, d.local) != 0x04034b50)
      throw archive_error("invalid Zip archive: Invalid local file header");
    const size_t start = d.local + 30 + get16(data_, d.local + 26) + get16(data_, d.local + 28);
    if (start + d.csize > data_.size()) throw archive_error("invalid Zip archive: entry extends past the archive");
    std::string out;
    if (d.method == 0) {
      if (d.csize != d.usize) throw archive_error("invalid Zip archive: stored entry sizes differ");
      out = data_.substr(start, d.csize);
    } else if (d.method == 8) {
      out = inflate_raw(data_.data() + start, d.csize, d.usize);
    } else {
      throw archive_error("unsupported Zip archive: Compression method not supported");
    }
    if (uint32_t(crc32(0L, reinterpret_cast<const Bytef*>(out.data()), uInt(out.size()))) != d.crc)
      throw archive_error("invalid Zip archive: Invalid checksum");
    return out;
  }

 private:
  struct Dir {
    uint32_t method = 0, crc = 0;
    uint64_t csize = 0, usize = 0, local = 0;
  };
  void index() {
    const size_t n = data_.size();
    size_t eocd = std::string::npos;
    for (size_t back = 22; back <= n && back <= 22 + 0xFFFF; ++back)
      if (get32(data_, n - back) == 0x06054b50) {
        eocd = n - back;
        break;
      }
    if (eocd == std::string::npos) throw archive_error("invalid Zip archive: Could not find EOCD");
    uint64_t count = get16(data_, eocd + 10);
    uint64_t cd_off = get32(data_, eocd + 16);
    // zip64 end record, located by the locator just before the classic one
    if ((count == 0xFFFF || cd_off == 0xFFFFFFFFu) && eocd >= 20 && get32(data_, eocd - 20) == 0x07064b50) {
      const uint64_t z = get64(data_, eocd - 20 + 8);
      if (z + 56 > n || get32(data_, size_t(z)) != 0x06064b50) throw archive_error("invalid Zip archive: bad zip64 end");
      count = get64(data_, size_t(z) + 32);
      cd_off = get64(data_, size_t(z) + 48);
    }
    size_t p = size_t(cd_off);
    for (uint64_t i = 0; i < count; ++i) {
      if (p + 46 > n || get32(data_, p) != 0x02014b50) throw archive_error("invalid Zip archive: Invalid Central Directory header");
      Dir d;
      d.method = get16(data_, p + 10);
      d.crc = get32(data_, p + 16);
      d.csize = get32(data_, p + 20);
      d.usize = get32(data_, p + 24);
      const size_t nl = get16(data_, p + 28), xl = get16(data_, p + 30), cl = get16(data_, p + 32);
      d.local = get32(data_, p + 42);
      if (p + 46 + nl + xl + cl > n) throw archive_error("invalid Zip archive: Invalid Central Directory header");
      const std::string name = data_.substr(p + 46, nl);
      // zip64 extended information: the fields whose classic value is all ones, in order
      for (size_t x = p + 46 + nl; x + 4 <= p + 46 + nl + xl;) {
        const uint32_t id = get16(data_, x), len = get16(data_, x + 2);
        if (id == 0x0001) {
          size_t q = x + 4;
          if (d.usize == 0xFFFFFFFFu && q + 8 <= x + 4 + len) d.usize = get64(data_, q), q += 8;
          if (d.csize == 0xFFFFFFFFu && q + 8 <= x + 4 + len) d.csize = get64(data_, q), q += 8;
          if (d.local == 0xFFFFFFFFu && q + 8 <= x + 4 + len) d.local = get64(data_, q), q += 8;
        }
        x += 4 + len;
      }
      dir_[name] = d;
      p += 46 + nl + xl + cl;
    }
  }
  std::string data_;
  std::map<std::string, Dir> dir_;
};

std::vector<std::string> lines_of(const std::string& text) {
  // str::lines: split at '\n', drop one trailing '\r' per line, no final empty line
  std::vector<std::string> out;
  size_t a = 0;
  while (a < text.size()) {
    size_t b = text.find('\n', a);
    if (b == std::string::npos) b = text.size();
    std::string line = text.substr(a, b - a);
    if (!line.empty() && line.back() == '\r') line.pop_back();
    out.push_back(line);
    a = b + 1;
  }
  return out;
}

}  // namespace

void BasisArchive::insert_metadata(const std::string& key, const std::string& value) {
  check_metadata(key, value);
  metadata[key] = value;
}

BasisArchive basis_archive(size_t width_bits, const std::vector<std::string>& states) {
  BasisArchive a;
  a.width_bits = width_bits;
  a.states.reserve(states.size());
  for (const auto& s : states) a.states.push_back(checked_state(width_bits, s));
  check_states(width_bits, a.states);
  return a;
}

BasisArchive basis_archive(size_t width_bits, const uint64_t* states, size_t n) {
  BasisArchive a;
  a.width_bits = width_bits;
  a.states.reserve(n);
  for (size_t i = 0; i < n; ++i) a.states.push_back(checked_state(width_bits, std::to_string(states[i])));
  check_states(width_bits, a.states);
  return a;
}

void save_basis_zip(const std::string& path, const BasisArchive& basis) {
  std::string manifest = "width_bits\t" + std::to_string(basis.width_bits) + "\n";
  for (const auto& s : basis.states) manifest += "state\t" + s + "\n";
  std::string meta;
  for (const auto& [k, v] : basis.metadata) meta += k + "\t" + v + "\n";
  write_zip(path, {{"basis_archive_version.npy", npy_u8({kBasisArchiveVersion})},
                   {"basis_manifest.tsv", manifest},
                   {"metadata.tsv", meta}});
}

BasisArchive load_basis_zip(const std::string& path) {
  const ZipReader zip(path);
  const uint8_t version = read_version(zip.read("basis_archive_version.npy"));
  if (version != kBasisArchiveVersion)
    throw archive_error("unsupported basis archive version " + std::to_string(version));
  const auto lines = lines_of(zip.read("basis_manifest.tsv"));
  const std::string prefix = "width_bits\t";
  if (lines.empty() || lines[0].compare(0, prefix.size(), prefix) != 0) throw archive_error("basis width is missing");
  // usize::from_str: an optional '+', then decimal digits that fit 64 bits
  std::string w = lines[0].substr(prefix.size());
  if (!w.empty() && w[0] == '+') w.erase(0, 1);
  if (w.empty()) throw archive_error("cannot parse integer from empty string");
  for (char c : w)
    if (c < '0' || c > '9') throw archive_error("invalid digit found in string");
  std::string wc;
  size_t wbits = 0;
  parse_decimal(w, wc, wbits);
  if (wbits > 64) throw archive_error("number too large to fit in target type");
  const size_t width_bits = size_t(std::stoull(wc));
  std::vector<std::string> states;
  states.reserve(lines.size() - 1);
  const std::string sp = "state\t";
  for (size_t i = 1; i < lines.size(); ++i) {
    if (lines[i].compare(0, sp.size(), sp) != 0) throw archive_error("invalid basis state row");
    states.push_back(checked_state(width_bits, lines[i].substr(sp.size())));
  }
  BasisArchive basis;
  basis.width_bits = width_bits;
  basis.states = std::move(states);
  check_states(width_bits, basis.states);
  if (zip.has("metadata.tsv")) {
    for (const auto& line : lines_of(zip.read("metadata.tsv"))) {
      const size_t tab = line.find('\t');
      if (tab == std::string::npos) throw archive_error("invalid archive metadata row");
      basis.insert_metadata(line.substr(0, tab), line.substr(tab + 1));
    }
  }
  return basis;
}

bool fits_u64(const std::string& canonical) {
  std::string c;
  size_t bits = 0;
  return parse_decimal(canonical, c, bits) && bits <= 64;
}

}  // namespace qmbed
