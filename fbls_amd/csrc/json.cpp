// JSON printing and parsing for the search report (see json.h).
#include "json.h"

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace fbls {
namespace json {

namespace {

bool reads_back(const char* text, double x) { return std::strtod(text, nullptr) == x; }

// Scientific text of |x| with p + 1 significant digits (correctly rounded),
// or its neighbour one unit away in the last digit, that reads back to x.
bool shortest_at(double ax, int p, std::string* digits, int* exp10) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.*e", p, ax);
  const char* e = std::strchr(buf, 'e');
  std::string d(1, buf[0]);
  if (p > 0) d.append((const char*)buf + 2, e);
  const int E = std::atoi(e + 1);
  if (reads_back(buf, ax)) {
    *digits = d;
    *exp10 = E;
    return true;
  }
  // the rounding interval of a power of two is wider above than below, so
  // the nearest candidate can miss while its neighbour reads back
  const uint64_t m = std::strtoull(d.c_str(), nullptr, 10);
  for (int dir = -1; dir <= 1; dir += 2) {
    uint64_t q = m + dir;
    int qe = E;
    std::string s = std::to_string(q);
    if (s.size() > d.size()) {  // 999 + 1
      qe += 1;
      s.pop_back();
    } else if (s.size() < d.size() || q == 0) {  // 100 - 1
      if (q == 0) continue;
      qe -= 1;
      s.push_back('9');
    }
    std::string text = s.substr(0, 1);
    if (s.size() > 1) text += "." + s.substr(1);
    text += "e" + std::to_string(qe);
    if (reads_back(text.c_str(), ax)) {
      *digits = s;
      *exp10 = qe;
      return true;
    }
  }
  return false;
}

void escape(const std::string& s, std::string& out) {
  static const char* hex = "0123456789abcdef";
  out += '"';
  for (unsigned char c : s) {
    switch (c) {
      case '"': out += "\\\""; break;
      case '\\': out += "\\\\"; break;
      case '\b': out += "\\b"; break;
      case '\f': out += "\\f"; break;
      case '\n': out += "\\n"; break;
      case '\r': out += "\\r"; break;
      case '\t': out += "\\t"; break;
      default:
        if (c < 0x20) {
          out += "\\u00";
          out += hex[c >> 4];
          out += hex[c & 15];
        } else {
          out += (char)c;
        }
    }
  }
  out += '"';
}

void emit(const Value& v, size_t indent, std::string& out) {
  switch (v.kind) {
    case Value::Null: out += "null"; return;
    case Value::Bool: out += v.b ? "true" : "false"; return;
    case Value::Int: out += std::to_string(v.i); return;
    case Value::Uint: out += std::to_string(v.u); return;
    case Value::Float: out += format_double(v.f); return;
    case Value::String: escape(v.s, out); return;
    case Value::Array: {
      if (v.a.empty()) {
        out += "[]";
        return;
      }
      out += "[\n";
      for (size_t k = 0; k < v.a.size(); k++) {
        out.append(indent + 2, ' ');
        emit(v.a[k], indent + 2, out);
        out += k + 1 < v.a.size() ? ",\n" : "\n";
      }
      out.append(indent, ' ');
      out += ']';
      return;
    }
    case Value::Object: {
      if (v.o.empty()) {
        out += "{}";
        return;
      }
      out += "{\n";
      size_t k = 0;
      for (const auto& kv : v.o) {
        out.append(indent + 2, ' ');
        escape(kv.first, out);
        out += ": ";
        emit(kv.second, indent + 2, out);
        out += ++k < v.o.size() ? ",\n" : "\n";
      }
      out.append(indent, ' ');
      out += '}';
      return;
    }
  }
}

struct Parser {
  // nesting deeper than serde_json's default recursion limit is refused
  static constexpr int kMaxDepth = 128;
  const std::string& t;
  size_t p = 0;
  int depth = 0;
  std::string err;
  explicit Parser(const std::string& text) : t(text) {}
  void ws() {
    while (p < t.size() && (t[p] == ' ' || t[p] == '\n' || t[p] == '\r' || t[p] == '\t')) p++;
  }
  bool fail(const char* what) {
    if (err.empty()) err = std::string(what) + " at byte " + std::to_string(p);
    return false;
  }
  bool lit(const char* w) {
    const size_t n = std::strlen(w);
    if (t.compare(p, n, w) != 0) return fail("unexpected token");
    p += n;
    return true;
  }
  static void utf8(uint32_t c, std::string& s) {
    if (c < 0x80) {
      s += (char)c;
    } else if (c < 0x800) {
      s += (char)(0xc0 | (c >> 6));
      s += (char)(0x80 | (c & 63));
    } else if (c < 0x10000) {
      s += (char)(0xe0 | (c >> 12));
      s += (char)(0x80 | ((c >> 6) & 63));
      s += (char)(0x80 | (c & 63));
    } else {
      s += (char)(0xf0 | (c >> 18));
      s += (char)(0x80 | ((c >> 12) & 63));
      s += (char)(0x80 | ((c >> 6) & 63));
      s += (char)(0x80 | (c & 63));
    }
  }
  bool hex4(uint32_t* c) {
    if (p + 4 > t.size()) return fail("short escape");
    *c = 0;
    for (int k = 0; k < 4; k++) {
      const char h = t[p++];
      *c <<= 4;
      if (h >= '0' && h <= '9') *c |= h - '0';
      else if (h >= 'a' && h <= 'f') *c |= h - 'a' + 10;
      else if (h >= 'A' && h <= 'F') *c |= h - 'A' + 10;
      else return fail("bad escape");
    }
    return true;
  }
  bool str(std::string* s) {
    if (p >= t.size() || t[p] != '"') return fail("expected string");
    p++;
    while (p < t.size() && t[p] != '"') {
      char c = t[p++];
      if (c != '\\') {
        *s += c;
        continue;
      }
      if (p >= t.size()) return fail("short escape");
      c = t[p++];
      switch (c) {
        case '"': *s += '"'; break;
        case '\\': *s += '\\'; break;
        case '/': *s += '/'; break;
        case 'b': *s += '\b'; break;
        case 'f': *s += '\f'; break;
        case 'n': *s += '\n'; break;
        case 'r': *s += '\r'; break;
        case 't': *s += '\t'; break;
        case 'u': {
          uint32_t u;
          if (!hex4(&u)) return false;
          if (u >= 0xdc00 && u < 0xe000) return fail("lone trailing surrogate in hex escape");
          if (u >= 0xd800 && u < 0xdc00) {
            if (t.compare(p, 2, "\\u") != 0) return fail("lone leading surrogate in hex escape");
            p += 2;
            uint32_t lo;
            if (!hex4(&lo)) return false;
            if (lo < 0xdc00 || lo >= 0xe000) return fail("invalid surrogate pair in hex escape");
            u = 0x10000 + ((u - 0xd800) << 10) + (lo - 0xdc00);
          }
          utf8(u, *s);
          break;
        }
        default: return fail("bad escape");
      }
    }
    if (p >= t.size()) return fail("unterminated string");
    p++;
    return true;
  }
  bool value(Value* v) {
    ws();
    if (p >= t.size()) return fail("unexpected end");
    const char c = t[p];
    if (c == 'n') {
      *v = Value::null();
      return lit("null");
    }
    if (c == 't') {
      *v = Value::boolean(true);
      return lit("true");
    }
    if (c == 'f') {
      *v = Value::boolean(false);
      return lit("false");
    }
    if (c == '"') {
      *v = Value::string("");
      return str(&v->s);
    }
    if ((c == '[' || c == '{') && depth >= kMaxDepth) return fail("recursion limit exceeded");
    if (c == '[') {
      p++;
      depth++;
      *v = Value::array();
      ws();
      if (p < t.size() && t[p] == ']') {
        p++;
        depth--;
        return true;
      }
      for (;;) {
        Value x;
        if (!value(&x)) return false;
        v->a.push_back(std::move(x));
        ws();
        if (p < t.size() && t[p] == ',') {
          p++;
          continue;
        }
        if (p < t.size() && t[p] == ']') {
          p++;
          depth--;
          return true;
        }
        return fail("expected , or ]");
      }
    }
    if (c == '{') {
      p++;
      depth++;
      *v = Value::object();
      ws();
      if (p < t.size() && t[p] == '}') {
        p++;
        depth--;
        return true;
      }
      for (;;) {
        ws();
        std::string key;
        if (!str(&key)) return false;
        ws();
        if (p >= t.size() || t[p] != ':') return fail("expected :");
        p++;
        Value x;
        if (!value(&x)) return false;
        v->o[key] = std::move(x);
        ws();
        if (p < t.size() && t[p] == ',') {
          p++;
          continue;
        }
        if (p < t.size() && t[p] == '}') {
          p++;
          depth--;
          return true;
        }
        return fail("expected , or }");
      }
    }
    if (c == '-' || (c >= '0' && c <= '9')) {
      // RFC 8259 number: -?(0|[1-9][0-9]*)(.[0-9]+)?([eE][+-]?[0-9]+)?
      auto digit = [&](size_t i) { return i < t.size() && t[i] >= '0' && t[i] <= '9'; };
      size_t q = p + (c == '-');
      bool real = false;
      if (!digit(q)) return fail("invalid number");
      if (t[q] == '0') {
        q++;
      } else {
        while (digit(q)) q++;
      }
      if (q < t.size() && t[q] == '.') {
        real = true;
        if (!digit(++q)) return fail("invalid number");
        while (digit(q)) q++;
      }
      if (q < t.size() && (t[q] == 'e' || t[q] == 'E')) {
        real = true;
        q++;
        if (q < t.size() && (t[q] == '+' || t[q] == '-')) q++;
        if (!digit(q)) return fail("invalid number");
        while (digit(q)) q++;
      }
      const std::string tok = t.substr(p, q - p);
      p = q;
      // integers beyond 64 bits, and -0, are read as doubles, as serde_json does
      errno = 0;
      if (!real && c == '-') {
        const long long x = std::strtoll(tok.c_str(), nullptr, 10);
        if (errno != ERANGE && x != 0) {
          *v = Value::integer(x);
          return true;
        }
      } else if (!real) {
        const unsigned long long x = std::strtoull(tok.c_str(), nullptr, 10);
        if (errno != ERANGE) {
          *v = Value::unsigned_(x);
          return true;
        }
      }
      *v = Value::number(std::strtod(tok.c_str(), nullptr));
      return true;
    }
    return fail("unexpected character");
  }
};

}  // namespace

Value numbers(const double* x, size_t n) {
  Value v = Value::array();
  for (size_t i = 0; i < n; i++) v.a.push_back(Value::number(x[i]));
  return v;
}

std::string format_double(double x) {
  if (!std::isfinite(x)) return "null";
  if (x == 0.0) return std::signbit(x) ? "-0.0" : "0.0";
  const double ax = std::fabs(x);
  std::string d;
  int E = 0;
  for (int p = 0; p <= 16; p++)
    if (shortest_at(ax, p, &d, &E)) break;
  while (d.size() > 1 && d.back() == '0') d.pop_back();
  const int n = (int)d.size();
  const int kk = E + 1;  // 10^(kk-1) <= |x| < 10^kk
  const int k = kk - n;  // exponent of the last digit
  std::string out = x < 0 ? "-" : "";
  if (k >= 0 && kk <= 16) {
    out += d;
    out.append(k, '0');
    out += ".0";
  } else if (kk > 0 && kk <= 16) {
    out += d.substr(0, kk) + "." + d.substr(kk);
  } else if (kk > -5 && kk <= 0) {
    out += "0.";
    out.append(-kk, '0');
    out += d;
  } else if (n == 1) {
    out += d + "e" + std::to_string(kk - 1);
  } else {
    out += d.substr(0, 1) + "." + d.substr(1) + "e" + std::to_string(kk - 1);
  }
  return out;
}

std::string pretty(const Value& v) {
  std::string out;
  emit(v, 0, out);
  return out;
}

bool parse(const std::string& text, Value* out, std::string* error) {
  Parser ps(text);
  if (!ps.value(out)) {
    if (error) *error = ps.err;
    return false;
  }
  ps.ws();
  if (ps.p != text.size()) {
    if (error) *error = "trailing characters at byte " + std::to_string(ps.p);
    return false;
  }
  return true;
}

}  // namespace json
}  // namespace fbls
