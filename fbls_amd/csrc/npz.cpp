// NPZ light-curve inputs (see npz.h).
#include "npz.h"

#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>

namespace fbls {
namespace io {

namespace {

uint16_t rd16(const unsigned char* p) { return (uint16_t)(p[0] | (p[1] << 8)); }
uint32_t rd32(const unsigned char* p) {
  return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
uint64_t rd64(const unsigned char* p) { return (uint64_t)rd32(p) | ((uint64_t)rd32(p + 4) << 32); }

// npz.rs:92-96: the text after `key` up to the next comma, trimmed
bool extract(const std::string& header, const std::string& key, std::string* out) {
  const size_t at = header.find(key);
  if (at == std::string::npos) return false;
  size_t b = at + key.size();
  while (b < header.size() && std::isspace((unsigned char)header[b])) b++;
  const size_t e = header.find(',', b);
  if (e == std::string::npos) return false;
  std::string s = header.substr(b, e - b);
  while (!s.empty() && std::isspace((unsigned char)s.back())) s.pop_back();
  *out = s;
  return true;
}

std::string trim(const std::string& s) {
  size_t b = 0, e = s.size();
  while (b < e && std::isspace((unsigned char)s[b])) b++;
  while (e > b && std::isspace((unsigned char)s[e - 1])) e--;
  return s.substr(b, e - b);
}

template <class T>
T load(const unsigned char* p, bool little) {
  unsigned char b[sizeof(T)];
  for (size_t k = 0; k < sizeof(T); k++) b[k] = little ? p[k] : p[sizeof(T) - 1 - k];
  T v;
  std::memcpy(&v, b, sizeof(T));
  return v;
}

template <class T>
bool read_all(const unsigned char* body, size_t len, size_t count, bool little,
              std::vector<double>* out, std::string* error) {
  if (count > len / sizeof(T)) {
    *error = "truncated NPY body";
    return false;
  }
  out->resize(count);
  for (size_t i = 0; i < count; i++) (*out)[i] = (double)load<T>(body + i * sizeof(T), little);
  return true;
}

bool inflate_raw(const unsigned char* src, size_t n, size_t want, std::vector<unsigned char>* out,
                 std::string* error) {
  if (n > UINT32_MAX || want > UINT32_MAX) {
    *error = "deflated members of 4 GiB or more are not supported";
    return false;
  }
  // deflate expands at most 1032:1; a larger claim is corrupt, not memory to reserve
  if (want / 1032 > n + 1) {
    *error = "corrupt deflate stream";
    return false;
  }
  out->resize(want);
  z_stream z;
  std::memset(&z, 0, sizeof(z));
  if (inflateInit2(&z, -MAX_WBITS) != Z_OK) {
    *error = "inflate initialisation failed";
    return false;
  }
  z.next_in = const_cast<unsigned char*>(src);
  z.avail_in = (uInt)n;
  z.next_out = out->data();
  z.avail_out = (uInt)want;
  const int st = inflate(&z, Z_FINISH);
  inflateEnd(&z);
  if (st != Z_STREAM_END || z.total_out != want) {
    *error = "corrupt deflate stream";
    return false;
  }
  return true;
}

// `x as i64` in the reference: saturating, NaN to zero
int64_t to_i64(double x) {
  if (std::isnan(x)) return 0;
  if (x >= 9223372036854775807.0) return INT64_MAX;
  if (x <= -9223372036854775808.0) return INT64_MIN;
  return (int64_t)x;
}

}  // namespace

bool parse_npy(const unsigned char* bytes, size_t len, Array* out, std::string* error) {
  if (len < 10 || std::memcmp(bytes, "\x93NUMPY", 6) != 0) {
    *error = "not an NPY file";
    return false;
  }
  size_t header_len, offset;
  if (bytes[6] == 1) {
    header_len = rd16(bytes + 8);
    offset = 10;
  } else {
    if (len < 12) {
      *error = "not an NPY file";
      return false;
    }
    header_len = rd32(bytes + 8);
    offset = 12;
  }
  if (offset + header_len > len) {
    *error = "truncated NPY header";
    return false;
  }
  const std::string header((const char*)bytes + offset, header_len);
  std::string descr, fortran;
  if (!extract(header, "'descr':", &descr)) {
    *error = "NPY header without descr";
    return false;
  }
  if (extract(header, "'fortran_order':", &fortran) && fortran.rfind("True", 0) == 0) {
    *error = "fortran_order arrays are not supported";
    return false;
  }
  const size_t sk = header.find("'shape':");
  if (sk == std::string::npos) {
    *error = "NPY header without shape";
    return false;
  }
  const size_t open = header.find('(', sk);
  const size_t close = open == std::string::npos ? open : header.find(')', open);
  if (open == std::string::npos || close == std::string::npos) {
    *error = "invalid shape";
    return false;
  }
  const std::string dims = header.substr(open + 1, close - open - 1);
  out->shape.clear();
  size_t start = 0;
  while (start <= dims.size()) {
    size_t comma = dims.find(',', start);
    if (comma == std::string::npos) comma = dims.size();
    const std::string tok = trim(dims.substr(start, comma - start));
    if (!tok.empty()) {
      char* end = nullptr;
      const unsigned long long v = std::strtoull(tok.c_str(), &end, 10);
      if (*end || tok[0] == '-') {
        *error = "invalid digit found in string";
        return false;
      }
      out->shape.push_back((size_t)v);
    }
    start = comma + 1;
  }
  size_t count = 1;
  for (size_t d : out->shape) {
    if (d != 0 && count > SIZE_MAX / 8 / d) {
      *error = "NPY shape too large";
      return false;
    }
    count *= d;
  }
  const unsigned char* body = bytes + offset + header_len;
  const size_t blen = len - offset - header_len;
  std::string d = descr;
  while (!d.empty() && (d.front() == '\'' || d.front() == '"')) d.erase(d.begin());
  while (!d.empty() && (d.back() == '\'' || d.back() == '"')) d.pop_back();
  const bool little = d.empty() || d[0] != '>';
  const std::string kind = d.size() > 1 ? d.substr(1) : "";
  if (kind == "f8") return read_all<double>(body, blen, count, little, &out->data, error);
  if (kind == "f4") return read_all<float>(body, blen, count, little, &out->data, error);
  if (kind == "i8") return read_all<int64_t>(body, blen, count, little, &out->data, error);
  if (kind == "i4") return read_all<int32_t>(body, blen, count, little, &out->data, error);
  if (kind == "i2") return read_all<int16_t>(body, blen, count, little, &out->data, error);
  if (kind == "u1" || kind == "b1") return read_all<uint8_t>(body, blen, count, true, &out->data, error);
  *error = "unsupported NPY dtype: " + kind;
  return false;
}

bool read_npz(const std::string& path, std::map<std::string, Array>* out, std::string* error) {
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    *error = path + ": No such file or directory (os error 2)";
    return false;
  }
  const std::vector<unsigned char> z((std::istreambuf_iterator<char>(f)),
                                     std::istreambuf_iterator<char>());
  const size_t n = z.size();
  auto bad = [&](const char* what) {
    *error = path + ": " + what;
    return false;
  };
  // end of central directory (and its zip64 form)
  if (n < 22) return bad("invalid Zip archive: could not find central directory end");
  size_t eocd = std::string::npos;
  const size_t lo = n >= 22 + 65535 ? n - 22 - 65535 : 0;
  for (size_t p = n - 22 + 1; p-- > lo;)
    if (rd32(&z[p]) == 0x06054b50u) {
      eocd = p;
      break;
    }
  if (eocd == std::string::npos) return bad("invalid Zip archive: could not find central directory end");
  uint64_t entries = rd16(&z[eocd + 10]);
  uint64_t cd_off = rd32(&z[eocd + 16]);
  if ((entries == 0xffff || cd_off == 0xffffffffu) && eocd >= 20 &&
      rd32(&z[eocd - 20]) == 0x07064b50u) {
    const uint64_t z64 = rd64(&z[eocd - 20 + 8]);
    if (z64 > n || n - z64 < 56 || rd32(&z[z64]) != 0x06064b50u) return bad("invalid Zip64 record");
    entries = rd64(&z[z64 + 32]);
    cd_off = rd64(&z[z64 + 48]);
  }
  size_t p = cd_off;
  out->clear();
  for (uint64_t e = 0; e < entries; e++) {
    if (p > n || n - p < 46 || rd32(&z[p]) != 0x02014b50u) return bad("invalid central directory");
    const uint16_t flags = rd16(&z[p + 8]), method = rd16(&z[p + 10]);
    uint64_t csize = rd32(&z[p + 20]), usize = rd32(&z[p + 24]);
    const uint16_t nlen = rd16(&z[p + 28]), xlen = rd16(&z[p + 30]), clen = rd16(&z[p + 32]);
    uint64_t local = rd32(&z[p + 42]);
    if (n - p - 46 < (size_t)nlen + xlen + clen) return bad("invalid central directory");
    std::string name((const char*)&z[p + 46], nlen);
    // zip64 extended information (header id 1); a record never reads past
    // the extra field
    const size_t xend = p + 46 + nlen + xlen;
    for (size_t x = p + 46 + nlen; x + 4 <= xend;) {
      const uint16_t id = rd16(&z[x]), sz = rd16(&z[x + 2]);
      const size_t rend = std::min(xend, x + 4 + sz);
      if (id == 1) {
        size_t q = x + 4;
        if (usize == 0xffffffffu && q + 8 <= rend) usize = rd64(&z[q]), q += 8;
        if (csize == 0xffffffffu && q + 8 <= rend) csize = rd64(&z[q]), q += 8;
        if (local == 0xffffffffu && q + 8 <= rend) local = rd64(&z[q]), q += 8;
      }
      x += 4 + sz;
    }
    p += 46 + nlen + xlen + clen;
    if (flags & 1) return bad("encrypted archives are not supported");
    if (local > n || n - local < 30 || rd32(&z[local]) != 0x04034b50u)
      return bad("invalid local file header");
    const size_t data = local + 30 + rd16(&z[local + 26]) + rd16(&z[local + 28]);
    if (data > n || csize > n - data) return bad("truncated member");
    std::vector<unsigned char> bytes;
    if (method == 0) {
      bytes.assign(z.begin() + data, z.begin() + data + csize);
    } else if (method == 8) {
      std::string why;
      if (!inflate_raw(&z[data], csize, usize, &bytes, &why)) return bad(why.c_str());
    } else {
      return bad("unsupported compression method");
    }
    while (name.size() >= 4 && name.compare(name.size() - 4, 4, ".npy") == 0)
      name.resize(name.size() - 4);
    Array arr;
    std::string why;
    if (!parse_npy(bytes.data(), bytes.size(), &arr, &why)) {
      *error = path + ":" + name + ": " + why;
      return false;
    }
    (*out)[name] = std::move(arr);
  }
  return true;
}

bool read_inputs(const std::vector<std::string>& paths, Observations* out, std::string* error) {
  *out = Observations();
  bool err_started = false;
  std::map<std::pair<size_t, int64_t>, int64_t> groups;
  for (size_t index = 0; index < paths.size(); index++) {
    const std::string& path = paths[index];
    const bool fits = (path.size() >= 5 && path.compare(path.size() - 5, 5, ".fits") == 0) ||
                      (path.size() >= 8 && path.compare(path.size() - 8, 8, ".fits.gz") == 0);
    if (fits) {
      *error = path + ": FITS input is not supported (convert it to NPZ)";
      return false;
    }
    std::map<std::string, Array> src;
    if (!read_npz(path, &src, error)) return false;
    auto t = src.find("time_bjd");
    auto f = src.find("flux");
    if (t == src.end() || f == src.end()) {
      *error = path + ": required columns are time_bjd and flux";
      return false;
    }
    if (t->second.shape.size() != 1 || f->second.shape != t->second.shape) {
      *error = path + ": time_bjd and flux must be matching vectors";
      return false;
    }
    const size_t n = t->second.data.size();
    auto e = src.find("flux_err");
    if (e != src.end()) {
      if (e->second.shape != t->second.shape) {
        *error = path + ": flux_err must match time_bjd";
        return false;
      }
      if (index == 0) err_started = true;
      if (!err_started) {
        *error = "all input files must either provide flux_err or omit it";
        return false;
      }
      out->err.insert(out->err.end(), e->second.data.begin(), e->second.data.end());
    } else if (err_started) {
      *error = "all input files must either provide flux_err or omit it";
      return false;
    }
    std::vector<int64_t> sector(n, 0);
    auto s = src.find("sector");
    if (s != src.end()) {
      const size_t m = s->second.data.size();
      if (m == 1) {
        sector.assign(n, to_i64(s->second.data[0]));
      } else if (m == n) {
        for (size_t i = 0; i < n; i++) sector[i] = to_i64(s->second.data[i]);
      } else {
        *error = path + ": sector must be a scalar or an observation vector";
        return false;
      }
    }
    out->time.insert(out->time.end(), t->second.data.begin(), t->second.data.end());
    out->flux.insert(out->flux.end(), f->second.data.begin(), f->second.data.end());
    for (int64_t sec : sector) {
      const auto key = std::make_pair(index, sec);
      auto it = groups.find(key);
      if (it == groups.end()) it = groups.emplace(key, (int64_t)groups.size()).first;
      out->group.push_back(it->second);
    }
  }
  out->has_err = err_started;
  if (!err_started) out->err.clear();
  return true;
}

}  // namespace io
}  // namespace fbls
