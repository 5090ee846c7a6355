// json_value.cpp - recursive-descent JSON reader (RFC 8259 grammar, no
// comments, no trailing commas), see json_value.hpp.
#include "json_value.hpp"

#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>

namespace osy {

class JsonParser {
 public:
  explicit JsonParser(const std::string& t) : t_(t) {}

  Json document() {
    skip_ws();
    Json v = value(0);
    skip_ws();
    if (p_ != t_.size()) fail("unexpected trailing characters");
    return v;
  }

 private:
  const std::string& t_;
  size_t p_ = 0;

  [[noreturn]] void fail(const char* what) const {
    throw JsonError(std::string("JSON parse error at byte ") +
                    std::to_string(p_) + ": " + what);
  }
  void skip_ws() {
    while (p_ < t_.size() &&
           (t_[p_] == ' ' || t_[p_] == '\t' || t_[p_] == '\n' || t_[p_] == '\r'))
      p_++;
  }
  bool eat(char c) {
    if (p_ < t_.size() && t_[p_] == c) {
      p_++;
      return true;
    }
    return false;
  }
  void expect_word(const char* w) {
    size_t n = std::strlen(w);
    if (t_.compare(p_, n, w) != 0) fail("invalid literal");
    p_ += n;
  }

  Json value(int depth) {
    if (depth > 256) fail("nesting too deep");
    if (p_ >= t_.size()) fail("unexpected end of input");
    Json v;
    char c = t_[p_];
    if (c == '{') {
      p_++;
      v.kind_ = Json::Kind::Object;
      skip_ws();
      if (eat('}')) return v;
      for (;;) {
        skip_ws();
        if (p_ >= t_.size() || t_[p_] != '"') fail("expected object key");
        std::string key = string_body();
        skip_ws();
        if (!eat(':')) fail("expected ':'");
        skip_ws();
        v.obj_[key] = value(depth + 1);  // last duplicate wins
        skip_ws();
        if (eat(',')) continue;
        if (eat('}')) break;
        fail("expected ',' or '}'");
      }
      return v;
    }
    if (c == '[') {
      p_++;
      v.kind_ = Json::Kind::Array;
      skip_ws();
      if (eat(']')) return v;
      for (;;) {
        skip_ws();
        v.arr_.push_back(value(depth + 1));
        skip_ws();
        if (eat(',')) continue;
        if (eat(']')) break;
        fail("expected ',' or ']'");
      }
      return v;
    }
    if (c == '"') {
      v.kind_ = Json::Kind::String;
      v.s_ = string_body();
      return v;
    }
    if (c == 't') {
      expect_word("true");
      v.kind_ = Json::Kind::Bool;
      v.b_ = true;
      return v;
    }
    if (c == 'f') {
      expect_word("false");
      v.kind_ = Json::Kind::Bool;
      v.b_ = false;
      return v;
    }
    if (c == 'n') {
      expect_word("null");
      return v;
    }
    if (c == '-' || (c >= '0' && c <= '9')) return number();
    fail("unexpected character");
  }

  static void put_utf8(std::string& out, uint32_t cp) {
    if (cp < 0x80) {
      out += char(cp);
    } else if (cp < 0x800) {
      out += char(0xC0 | (cp >> 6));
      out += char(0x80 | (cp & 0x3F));
    } else if (cp < 0x10000) {
      out += char(0xE0 | (cp >> 12));
      out += char(0x80 | ((cp >> 6) & 0x3F));
      out += char(0x80 | (cp & 0x3F));
    } else {
      out += char(0xF0 | (cp >> 18));
      out += char(0x80 | ((cp >> 12) & 0x3F));
      out += char(0x80 | ((cp >> 6) & 0x3F));
      out += char(0x80 | (cp & 0x3F));
    }
  }

  uint32_t hex4() {
    if (p_ + 4 > t_.size()) fail("truncated \\u escape");
    uint32_t v = 0;
    for (int k = 0; k < 4; k++) {
      char c = t_[p_++];
      v <<= 4;
      if (c >= '0' && c <= '9') v |= uint32_t(c - '0');
      else if (c >= 'a' && c <= 'f') v |= uint32_t(c - 'a' + 10);
      else if (c >= 'A' && c <= 'F') v |= uint32_t(c - 'A' + 10);
      else fail("bad hex digit in \\u escape");
    }
    return v;
  }

  std::string string_body() {
    p_++;  // opening quote
    std::string out;
    for (;;) {
      if (p_ >= t_.size()) fail("unterminated string");
      unsigned char c = (unsigned char)t_[p_++];
      if (c == '"') break;
      if (c < 0x20) fail("control character in string");
      if (c != '\\') {
        out += char(c);
        continue;
      }
      if (p_ >= t_.size()) fail("unterminated escape");
      char e = t_[p_++];
      switch (e) {
        case '"': out += '"'; break;
        case '\\': out += '\\'; break;
        case '/': out += '/'; break;
        case 'b': out += '\b'; break;
        case 'f': out += '\f'; break;
        case 'n': out += '\n'; break;
        case 'r': out += '\r'; break;
        case 't': out += '\t'; break;
        case 'u': {
          uint32_t cp = hex4();
          if (cp >= 0xD800 && cp <= 0xDBFF) {
            if (!(eat('\\') && eat('u'))) fail("unpaired surrogate");
            uint32_t lo = hex4();
            if (lo < 0xDC00 || lo > 0xDFFF) fail("bad low surrogate");
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          } else if (cp >= 0xDC00 && cp <= 0xDFFF) {
            fail("unpaired surrogate");
          }
          put_utf8(out, cp);
          break;
        }
        default:
          fail("invalid escape");
      }
    }
    return out;
  }

  Json number() {
    size_t start = p_;
    bool neg = eat('-');
    if (p_ >= t_.size()) fail("truncated number");
    if (t_[p_] == '0') {
      p_++;
    } else if (t_[p_] >= '1' && t_[p_] <= '9') {
      while (p_ < t_.size() && t_[p_] >= '0' && t_[p_] <= '9') p_++;
    } else {
      fail("bad number");
    }
    bool is_float = false;
    if (p_ < t_.size() && t_[p_] == '.') {
      is_float = true;
      p_++;
      if (p_ >= t_.size() || !(t_[p_] >= '0' && t_[p_] <= '9')) fail("bad fraction");
      while (p_ < t_.size() && t_[p_] >= '0' && t_[p_] <= '9') p_++;
    }
    if (p_ < t_.size() && (t_[p_] == 'e' || t_[p_] == 'E')) {
      is_float = true;
      p_++;
      if (p_ < t_.size() && (t_[p_] == '+' || t_[p_] == '-')) p_++;
      if (p_ >= t_.size() || !(t_[p_] >= '0' && t_[p_] <= '9')) fail("bad exponent");
      while (p_ < t_.size() && t_[p_] >= '0' && t_[p_] <= '9') p_++;
    }
    std::string tok = t_.substr(start, p_ - start);
    Json v;
    if (!is_float) {
      errno = 0;
      if (neg) {
        long long x = std::strtoll(tok.c_str(), nullptr, 10);
        if (errno == 0) {
          v.kind_ = Json::Kind::Int;
          v.i_ = x;
          return v;
        }
      } else {
        unsigned long long x = std::strtoull(tok.c_str(), nullptr, 10);
        if (errno == 0) {
          if (x <= (unsigned long long)std::numeric_limits<int64_t>::max()) {
            v.kind_ = Json::Kind::Int;
            v.i_ = (int64_t)x;
          } else {
            v.kind_ = Json::Kind::UInt;
            v.u_ = x;
          }
          return v;
        }
      }
      // integer overflow: nlohmann falls back to a floating-point number
    }
    v.kind_ = Json::Kind::Float;
    v.f_ = std::strtod(tok.c_str(), nullptr);
    return v;
  }
};

Json Json::parse(const std::string& text) { return JsonParser(text).document(); }

const char* Json::type_name() const {
  switch (kind_) {
    case Kind::Null: return "null";
    case Kind::Bool: return "boolean";
    case Kind::Int:
    case Kind::UInt:
    case Kind::Float: return "number";
    case Kind::String: return "string";
    case Kind::Array: return "array";
    case Kind::Object: return "object";
  }
  return "?";
}

bool Json::contains(const std::string& key) const {
  return kind_ == Kind::Object && obj_.count(key) != 0;
}

const Json& Json::at(const std::string& key) const {
  if (kind_ != Kind::Object)
    throw JsonError(std::string("cannot use at() with ") + type_name());
  auto it = obj_.find(key);
  if (it == obj_.end()) throw JsonError("key '" + key + "' not found");
  return it->second;
}

int64_t Json::to_int64() const {
  switch (kind_) {
    case Kind::Int: return i_;
    case Kind::UInt: return (int64_t)u_;
    case Kind::Float: {
      if (!(f_ > -9.2e18 && f_ < 9.2e18))
        throw JsonError("number out of integer range");
      return (int64_t)f_;  // truncation, as static_cast
    }
    case Kind::Bool: return b_ ? 1 : 0;
    default:
      throw JsonError(std::string("type must be number, but is ") + type_name());
  }
}

double Json::to_double() const {
  switch (kind_) {
    case Kind::Int: return (double)i_;
    case Kind::UInt: return (double)u_;
    case Kind::Float: return f_;
    case Kind::Bool: return b_ ? 1.0 : 0.0;
    default:
      throw JsonError(std::string("type must be number, but is ") + type_name());
  }
}

bool Json::to_bool() const {
  if (kind_ != Kind::Bool)
    throw JsonError(std::string("type must be boolean, but is ") + type_name());
  return b_;
}

const std::string& Json::str() const {
  if (kind_ != Kind::String)
    throw JsonError(std::string("type must be string, but is ") + type_name());
  return s_;
}

}  // namespace osy
