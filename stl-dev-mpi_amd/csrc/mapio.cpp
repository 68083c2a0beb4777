// Map I/O (reference crates/stl-mpi/src/io.rs:1-460) and the initial-map
// normal draws (crates/stl-cli/src/lib.rs:523-540).  Formats as the
// reference writes and reads them: `.npy` 1.0 headers padded to 64 bytes,
// FITS headers of 80-byte cards in 2880-byte blocks, big-endian table data.
#include "mapio.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <fstream>
#include <stdexcept>

#include "healpix_device.hpp"

namespace stl {
namespace io {

namespace {

constexpr size_t kBlock = 2880;

[[noreturn]] void fail(const std::string& path, const std::string& msg) {
  throw std::runtime_error(path + ": " + msg);
}

bool is_fits(const std::string& path) {
  auto dot = path.rfind('.');
  if (dot == std::string::npos) return false;
  std::string ext = path.substr(dot + 1);
  return ext == "fits" || ext == "fit";
}

std::vector<char> read_bytes(const std::string& path, uint64_t offset, size_t len) {
  std::ifstream f(path, std::ios::binary);
  if (!f) fail(path, "cannot open");
  f.seekg(std::streamoff(offset));
  std::vector<char> buf(len);
  f.read(buf.data(), std::streamsize(len));
  if (size_t(f.gcount()) != len) fail(path, "truncated file");
  return buf;
}

template <typename U>
U load_bytes(const char* p, bool big) {
  unsigned char b[sizeof(U)];
  std::memcpy(b, p, sizeof(U));
  const uint16_t probe = 1;
  const bool host_little = *reinterpret_cast<const unsigned char*>(&probe) == 1;
  if (big == host_little)
    for (size_t i = 0; i < sizeof(U) / 2; ++i) std::swap(b[i], b[sizeof(U) - 1 - i]);
  U v;
  std::memcpy(&v, b, sizeof(U));
  return v;
}

template <typename U>
void store_bytes(std::vector<char>& out, U v, bool big) {
  unsigned char b[sizeof(U)];
  std::memcpy(b, &v, sizeof(U));
  const uint16_t probe = 1;
  const bool host_little = *reinterpret_cast<const unsigned char*>(&probe) == 1;
  if (big == host_little)
    for (size_t i = 0; i < sizeof(U) / 2; ++i) std::swap(b[i], b[sizeof(U) - 1 - i]);
  out.insert(out.end(), reinterpret_cast<char*>(b), reinterpret_cast<char*>(b) + sizeof(U));
}

// ---------------------------------------------------------------------------
// .npy (io.rs:26-78, 237-300)
// ---------------------------------------------------------------------------

struct NpyLayout {
  uint64_t offset = 0;
  char kind = 'f';  // 'f' float, 'i' integer
  int size = 8;     // bytes per element
  bool big = false;
  std::vector<size_t> shape;
  // Element count; a product that overflows size_t is refused.
  size_t count() const {
    size_t n = 1;
    for (size_t s : shape) {
      if (s != 0 && n > SIZE_MAX / s) throw std::runtime_error(".npy shape overflows the address space");
      n *= s;
    }
    return n;
  }
};

uint64_t file_size(const std::string& path) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) fail(path, "cannot open");
  return uint64_t(f.tellg());
}

// Value of a header dict field: the text after "'key':".
bool dict_field(const std::string& dict, const std::string& key, std::string& out) {
  auto at = dict.find("'" + key + "'");
  if (at == std::string::npos) return false;
  at += key.size() + 2;
  while (at < dict.size() && (dict[at] == ' ' || dict[at] == ':')) ++at;
  out = dict.substr(at);
  return true;
}

NpyLayout npy_layout(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) fail(path, "cannot open");
  unsigned char head[12];
  f.read(reinterpret_cast<char*>(head), 12);
  if (f.gcount() < 10 || std::memcmp(head, "\x93NUMPY", 6) != 0) fail(path, "not a .npy file");
  size_t hlen, start;
  if (head[6] == 1) {
    hlen = size_t(head[8]) | (size_t(head[9]) << 8);
    start = 10;
  } else {
    hlen = size_t(head[8]) | (size_t(head[9]) << 8) | (size_t(head[10]) << 16) | (size_t(head[11]) << 24);
    start = 12;
  }
  auto text = read_bytes(path, start, hlen);
  std::string dict(text.begin(), text.end());
  NpyLayout L;
  L.offset = start + hlen;
  std::string v;
  if (!dict_field(dict, "descr", v)) fail(path, "no descr");
  if (v.empty() || v[0] != '\'') fail(path, "bad descr");
  std::string d = v.substr(1, v.find('\'', 1) - 1);
  if (d.size() < 3) fail(path, "unsupported dtype " + d);
  const char order = d[0];
  const std::string code = d.substr(1);
  if (order == '>') L.big = true;
  else if (order == '<' || order == '|' || order == '=') L.big = false;
  else fail(path, "unsupported dtype " + d);
  if (code == "f8") { L.kind = 'f'; L.size = 8; }
  else if (code == "f4") { L.kind = 'f'; L.size = 4; }
  else if (code == "i8") { L.kind = 'i'; L.size = 8; }
  else if (code == "i4") { L.kind = 'i'; L.size = 4; }
  else fail(path, "unsupported dtype " + d);
  if (dict_field(dict, "fortran_order", v) && v.rfind("True", 0) == 0) fail(path, "Fortran order");
  if (!dict_field(dict, "shape", v)) fail(path, "no shape");
  if (v.empty() || v[0] != '(') fail(path, "bad shape");
  std::string inner = v.substr(1, v.find(')') - 1);
  size_t pos = 0;
  while (pos < inner.size()) {
    size_t comma = inner.find(',', pos);
    std::string tok = inner.substr(pos, comma == std::string::npos ? std::string::npos : comma - pos);
    size_t a = tok.find_first_not_of(' '), b = tok.find_last_not_of(' ');
    if (a != std::string::npos) {
      tok = tok.substr(a, b - a + 1);
      char* end = nullptr;
      unsigned long long n = std::strtoull(tok.c_str(), &end, 10);
      if (end == tok.c_str() || *end) fail(path, "bad shape");
      L.shape.push_back(size_t(n));
    }
    if (comma == std::string::npos) break;
    pos = comma + 1;
  }
  // The data the header announces must be in the file (a malformed shape would
  // otherwise size a read beyond it).
  const size_t n = L.count();
  if (n > SIZE_MAX / size_t(L.size) || L.offset + uint64_t(n) * uint64_t(L.size) > file_size(path))
    fail(path, "shape exceeds the file");
  return L;
}

double npy_value(const NpyLayout& L, const char* p) {
  if (L.kind == 'f') return L.size == 8 ? load_bytes<double>(p, L.big) : double(load_bytes<float>(p, L.big));
  return L.size == 8 ? double(load_bytes<int64_t>(p, L.big)) : double(load_bytes<int32_t>(p, L.big));
}

std::vector<char> npy_header(const std::vector<size_t>& shape) {
  std::string dims;
  for (size_t i = 0; i < shape.size(); ++i) dims += (i ? " " : "") + std::to_string(shape[i]) + ",";
  std::string header = "{'descr': '<f8', 'fortran_order': False, 'shape': (" + dims + "), }";
  // Magic (6) + version (2) + length (2) + header + '\n', aligned on 64 bytes.
  while ((10 + header.size() + 1) % 64 != 0) header.push_back(' ');
  header.push_back('\n');
  std::vector<char> buf = {'\x93', 'N', 'U', 'M', 'P', 'Y', '\x01', '\x00'};
  buf.push_back(char(header.size() & 0xFF));
  buf.push_back(char((header.size() >> 8) & 0xFF));
  buf.insert(buf.end(), header.begin(), header.end());
  return buf;
}

// ---------------------------------------------------------------------------
// FITS (io.rs:80-234, 302-350)
// ---------------------------------------------------------------------------

std::string card(const std::string& key, const std::string& value) {
  std::string c = key;
  c.resize(8, ' ');
  if (!value.empty()) {
    std::string v = value;
    if (v.size() < 20) v = std::string(20 - v.size(), ' ') + v;
    c += "= " + v;
  }
  c.resize(80, ' ');
  return c;
}

std::string string_value(const std::string& s) {
  std::string q = s;
  if (q.size() < 8) q.resize(8, ' ');
  q = "'" + q + "'";
  if (q.size() < 20) q.resize(20, ' ');
  return q;
}

void pad(std::vector<char>& buf, char fill) {
  while (buf.size() % kBlock != 0) buf.push_back(fill);
}

std::vector<char> fits_headers(size_t n, const std::string& column, bool dbl) {
  const int64_t nside = nside_of(n);
  std::vector<char> buf;
  for (const auto& c : {card("SIMPLE", "T"), card("BITPIX", "8"), card("NAXIS", "0"), card("EXTEND", "T"),
                        card("END", "")})
    buf.insert(buf.end(), c.begin(), c.end());
  pad(buf, ' ');
  const int width = dbl ? 8 : 4;
  for (const auto& c : {card("XTENSION", string_value("BINTABLE")), card("BITPIX", "8"), card("NAXIS", "2"),
                        card("NAXIS1", std::to_string(width)), card("NAXIS2", std::to_string(n)),
                        card("PCOUNT", "0"), card("GCOUNT", "1"), card("TFIELDS", "1"),
                        card("TTYPE1", string_value(column)), card("TFORM1", string_value(dbl ? "D" : "E")),
                        card("PIXTYPE", string_value("HEALPIX")), card("ORDERING", string_value("NESTED")),
                        card("COORDSYS", string_value("C")), card("EXTNAME", string_value("xtension")),
                        card("NSIDE", std::to_string(nside)), card("FIRSTPIX", "0"),
                        card("LASTPIX", std::to_string(n - 1)), card("INDXSCHM", string_value("IMPLICIT")),
                        card("OBJECT", string_value("FULLSKY")), card("END", "")})
    buf.insert(buf.end(), c.begin(), c.end());
  pad(buf, ' ');
  return buf;
}

using Cards = std::vector<std::pair<std::string, std::string>>;

// Cards of the HDU header at `at` and the header's length in blocks' bytes;
// false when the bytes run out before END.
bool parse_header(const std::vector<char>& data, size_t at0, Cards& cards, size_t& len) {
  size_t at = at0;
  cards.clear();
  for (;;) {
    if (at + 80 > data.size()) return false;
    std::string text(data.begin() + long(at), data.begin() + long(at + 80));
    at += 80;
    std::string key = text.substr(0, 8);
    key.erase(key.find_last_not_of(' ') + 1);
    if (key == "END") break;
    if (text.compare(8, 2, "= ") == 0) {
      std::string raw = text.substr(10);
      size_t a = raw.find_first_not_of(' ');
      raw = a == std::string::npos ? "" : raw.substr(a);
      std::string value;
      if (!raw.empty() && raw[0] == '\'') {
        value = raw.substr(1, raw.find('\'', 1) == std::string::npos ? std::string::npos : raw.find('\'', 1) - 1);
      } else {
        value = raw.substr(0, raw.find('/'));
      }
      size_t b = value.find_first_not_of(' '), e = value.find_last_not_of(' ');
      value = b == std::string::npos ? "" : value.substr(b, e - b + 1);
      cards.emplace_back(key, value);
    }
  }
  len = (at - at0 + kBlock - 1) / kBlock * kBlock;
  return true;
}

bool get_card(const Cards& cards, const std::string& key, std::string& out) {
  for (const auto& kv : cards)
    if (kv.first == key) {
      out = kv.second;
      return true;
    }
  return false;
}

size_t card_int(const Cards& cards, const std::string& key) {
  std::string v;
  if (!get_card(cards, key, v)) return 0;
  char* end = nullptr;
  unsigned long long n = std::strtoull(v.c_str(), &end, 10);
  return end == v.c_str() ? 0 : size_t(n);
}

struct FitsLayout {
  uint64_t offset = 0;
  size_t row = 0, repeat = 1, width = 4, rows = 0;
  bool nested = true;
};

// Layout of the first binary table (headers read block by block; the data of
// other HDUs skipped).
FitsLayout fits_layout(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) fail(path, "cannot open");
  size_t at = 0;
  for (;;) {
    // The header at `at`, one block at a time until its END card.
    std::vector<char> data;
    Cards cards;
    size_t hlen = 0;
    for (;;) {
      std::vector<char> block(kBlock);
      f.clear();
      f.seekg(std::streamoff(at + data.size()));
      f.read(block.data(), std::streamsize(kBlock));
      if (size_t(f.gcount()) != kBlock) fail(path, "no binary table");
      data.insert(data.end(), block.begin(), block.end());
      if (parse_header(data, 0, cards, hlen)) break;
    }
    const size_t naxis = card_int(cards, "NAXIS");
    size_t size = naxis == 0 ? 0 : 1;
    for (size_t i = 1; i <= naxis; ++i) size *= card_int(cards, "NAXIS" + std::to_string(i));
    std::string xt;
    if (get_card(cards, "XTENSION", xt) && xt == "BINTABLE") {
      std::string form;
      if (!get_card(cards, "TFORM1", form)) fail(path, "no TFORM1");
      size_t split = form.find_first_of("ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz");
      if (split == std::string::npos) fail(path, "bad TFORM1");
      FitsLayout L;
      L.repeat = split == 0 ? 1 : size_t(std::strtoull(form.substr(0, split).c_str(), nullptr, 10));
      const char code = form[split];
      if (code == 'E') L.width = 4;
      else if (code == 'D') L.width = 8;
      else fail(path, std::string("column type ") + code + " is not a float map");
      std::string ordering = "NESTED";
      get_card(cards, "ORDERING", ordering);
      if (ordering != "NESTED" && ordering != "RING") fail(path, "unknown ORDERING " + ordering);
      L.nested = ordering == "NESTED";
      L.offset = at + hlen;
      L.row = card_int(cards, "NAXIS1");
      L.rows = card_int(cards, "NAXIS2");
      if (L.repeat * L.width > L.row) fail(path, "truncated table");
      return L;
    }
    at += hlen + (size + card_int(cards, "PCOUNT") + kBlock - 1) / kBlock * kBlock;
  }
}

double fits_value(const FitsLayout& L, const char* p) {
  return L.width == 4 ? double(load_bytes<float>(p, true)) : load_bytes<double>(p, true);
}

}  // namespace

int64_t nside_of(size_t npix) {
  const int64_t nside = int64_t(std::llround(std::sqrt(double(npix / 12))));
  if (nside <= 0 || 12 * nside * nside != int64_t(npix) || (nside & (nside - 1)) != 0)
    throw std::runtime_error(std::to_string(npix) + " pixels is not a full-sky HEALPix map");
  return nside;
}

int64_t ring2nest(int64_t pix, int order) { return hp::ring2nest(pix, order); }

std::vector<double> read_npy(const std::string& path, std::vector<size_t>* shape) {
  NpyLayout L = npy_layout(path);
  const size_t n = L.count();
  auto bytes = read_bytes(path, L.offset, n * size_t(L.size));
  std::vector<double> out(n);
  for (size_t i = 0; i < n; ++i) out[i] = npy_value(L, bytes.data() + i * size_t(L.size));
  if (shape) *shape = L.shape;
  return out;
}

void write_npy(const std::string& path, const double* values, const std::vector<size_t>& shape) {
  size_t n = 1;
  for (size_t s : shape) n *= s;
  std::vector<char> buf = npy_header(shape);
  buf.reserve(buf.size() + 8 * n);
  for (size_t i = 0; i < n; ++i) store_bytes<double>(buf, values[i], false);
  std::ofstream f(path, std::ios::binary | std::ios::trunc);
  if (!f) fail(path, "cannot create");
  f.write(buf.data(), std::streamsize(buf.size()));
  if (!f) fail(path, "write error");
}

void write_healpix_fits(const std::string& path, const double* values, size_t n, const std::string& column,
                        bool dbl) {
  std::vector<char> buf = fits_headers(n, column, dbl);
  for (size_t i = 0; i < n; ++i) {
    if (dbl) store_bytes<double>(buf, values[i], true);
    else store_bytes<float>(buf, float(values[i]), true);
  }
  pad(buf, 0);
  std::ofstream f(path, std::ios::binary | std::ios::trunc);
  if (!f) fail(path, "cannot create");
  f.write(buf.data(), std::streamsize(buf.size()));
  if (!f) fail(path, "write error");
}

std::vector<double> read_healpix_fits(const std::string& path) {
  FitsLayout L = fits_layout(path);
  const size_t n = L.rows * L.repeat;
  auto body = read_bytes(path, L.offset, L.rows * L.row);
  std::vector<double> values(n);
  for (size_t r = 0; r < L.rows; ++r)
    for (size_t k = 0; k < L.repeat; ++k) values[r * L.repeat + k] = fits_value(L, body.data() + r * L.row + k * L.width);
  const int64_t nside = nside_of(n);
  if (L.nested) return values;
  int order = 0;
  while ((int64_t(1) << order) < nside) ++order;
  std::vector<double> nest(n);
  for (size_t p = 0; p < n; ++p) nest[size_t(hp::ring2nest(int64_t(p), order))] = values[p];
  return nest;
}

std::vector<double> read_map(const std::string& path) {
  return is_fits(path) ? read_healpix_fits(path) : read_npy(path, nullptr);
}

size_t map_len(const std::string& path) {
  if (is_fits(path)) {
    FitsLayout L = fits_layout(path);
    return L.repeat * L.rows;
  }
  return npy_layout(path).count();
}

std::vector<double> read_map_range(const std::string& path, size_t start, size_t end) {
  if (end < start) fail(path, "empty range");
  const size_t n = end - start;
  if (is_fits(path)) {
    FitsLayout L = fits_layout(path);
    if (end > L.rows * L.repeat)
      fail(path, "range " + std::to_string(start) + ".." + std::to_string(end) + " beyond the map");
    if (!L.nested) {
      auto all = read_healpix_fits(path);
      return std::vector<double>(all.begin() + long(start), all.begin() + long(end));
    }
    std::vector<double> out(n);
    if (n == 0) return out;
    const size_t r0 = start / L.repeat, r1 = (end + L.repeat - 1) / L.repeat;
    auto bytes = read_bytes(path, L.offset + r0 * L.row, (r1 - r0) * L.row);
    for (size_t p = start; p < end; ++p)
      out[p - start] = fits_value(L, bytes.data() + (p / L.repeat - r0) * L.row + (p % L.repeat) * L.width);
    return out;
  }
  NpyLayout L = npy_layout(path);
  if (end > L.count())
    fail(path, "range " + std::to_string(start) + ".." + std::to_string(end) + " beyond " +
                   std::to_string(L.count()) + " elements");
  auto bytes = read_bytes(path, L.offset + start * size_t(L.size), n * size_t(L.size));
  std::vector<double> out(n);
  for (size_t i = 0; i < n; ++i) out[i] = npy_value(L, bytes.data() + i * size_t(L.size));
  return out;
}

void sliced_create(const std::string& path, size_t n, const std::string& column, bool dbl) {
  std::vector<char> head;
  size_t data;
  if (is_fits(path)) {
    head = fits_headers(n, column, dbl);
    data = ((dbl ? 8 : 4) * n + kBlock - 1) / kBlock * kBlock;
  } else {
    head = npy_header({n});
    data = 8 * n;
  }
  {
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f) fail(path, "cannot create");
    f.write(head.data(), std::streamsize(head.size()));
    if (data > 0) {
      // Size the file: the last byte of the data area.
      f.seekp(std::streamoff(head.size() + data - 1));
      f.put('\0');
    }
    if (!f) fail(path, "write error");
  }
}

void sliced_write(const std::string& path, bool dbl, size_t start, const double* values, size_t n) {
  const bool fits = is_fits(path);
  // The map length and the data offset come from the file's own header, so a
  // slice beyond the map is refused instead of growing or corrupting the file.
  uint64_t offset;
  size_t len;
  if (fits) {
    const FitsLayout L = fits_layout(path);
    if (L.repeat != 1 || L.width != (dbl ? 8u : 4u)) fail(path, "not a sliced map of this element type");
    offset = L.offset;
    len = L.rows;
  } else {
    const NpyLayout L = npy_layout(path);
    if (L.kind != 'f' || L.size != 8 || L.big || L.shape.size() != 1) fail(path, "not a sliced float64 map");
    offset = L.offset;
    len = L.shape[0];
  }
  if (start > len || n > len - start)
    fail(path, "slice " + std::to_string(start) + ".." + std::to_string(start + n) + " beyond the map of " +
                   std::to_string(len) + " pixels");
  std::vector<char> bytes;
  bytes.reserve(n * 8);
  size_t width = 8;
  for (size_t i = 0; i < n; ++i) {
    if (!fits) store_bytes<double>(bytes, values[i], false);
    else if (dbl) store_bytes<double>(bytes, values[i], true);
    else { store_bytes<float>(bytes, float(values[i]), true); width = 4; }
  }
  std::fstream f(path, std::ios::binary | std::ios::in | std::ios::out);
  if (!f) fail(path, "cannot open for writing");
  f.seekp(std::streamoff(offset + start * width));
  f.write(bytes.data(), std::streamsize(bytes.size()));
  if (!f) fail(path, "write error");
}

namespace {
inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}
}  // namespace

double pixel_normal(uint64_t seed, int64_t id) {
  const uint64_t a = splitmix64(seed ^ splitmix64(uint64_t(id)));
  const uint64_t b = splitmix64(a);
  const double scale = 1.0 / double(uint64_t(1) << 53);
  const double u1 = (double(a >> 11) + 0.5) * scale;
  const double u2 = (double(b >> 11) + 0.5) * scale;
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * 3.14159265358979323846 * u2);
}

}  // namespace io
}  // namespace stl
