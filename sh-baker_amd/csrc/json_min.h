// A small JSON document model and parser (RFC 8259) for reading glTF:
// null, booleans, numbers (double), strings (UTF-8, \u escapes incl.
// surrogate pairs), arrays and objects (insertion order kept).
#ifndef SHB_JSON_MIN_H_
#define SHB_JSON_MIN_H_

#include <cstdint>
#include <cstdlib>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace shb {
namespace json {

struct Value {
  enum Type { Null, Bool, Number, String, Array, Object } type = Null;
  bool b = false;
  double num = 0.0;
  std::string str;
  std::vector<Value> arr;
  std::vector<std::pair<std::string, Value>> obj;

  bool is_null() const { return type == Null; }
  bool is_number() const { return type == Number; }
  bool is_string() const { return type == String; }
  bool is_array() const { return type == Array; }
  bool is_object() const { return type == Object; }
  // member lookup (nullptr when absent or not an object)
  const Value* get(const std::string& key) const {
    if (type != Object) return nullptr;
    for (const auto& kv : obj)
      if (kv.first == key) return &kv.second;
    return nullptr;
  }
  bool has(const std::string& key) const { return get(key) != nullptr; }
  size_t size() const { return type == Array ? arr.size() : type == Object ? obj.size() : 0; }
  const Value& operator[](size_t i) const { return arr[i]; }
  double number_or(double d) const { return type == Number ? num : d; }
};

// number member or default
inline double num(const Value& v, const char* key, double d) {
  const Value* m = v.get(key);
  return m && m->is_number() ? m->num : d;
}
// An int field; absent, not a number, NaN or outside the int range: d.
inline int integer(const Value& v, const char* key, int d) {
  const Value* m = v.get(key);
  if (!m || !m->is_number()) return d;
  const double x = m->num;
  return x >= -2147483648.0 && x <= 2147483647.0 ? (int)x : d;
}
inline std::string string(const Value& v, const char* key, const std::string& d = "") {
  const Value* m = v.get(key);
  return m && m->is_string() ? m->str : d;
}

class Parser {
 public:
  Parser(const char* p, size_t n) : p_(p), end_(p + n) {}
  bool parse(Value* out, std::string* err) {
    ws();
    if (!value(out, 0)) {
      *err = "JSON: " + err_ + " at byte " + std::to_string(pos());
      return false;
    }
    ws();
    if (p_ != end_) {
      *err = "JSON: trailing characters at byte " + std::to_string(pos());
      return false;
    }
    return true;
  }

 private:
  const char* p_;
  const char* end_;
  const char* begin_ = p_;
  std::string err_;
  size_t pos() const { return (size_t)(p_ - begin_); }
  bool fail(const char* m) {
    err_ = m;
    return false;
  }
  void ws() {
    while (p_ < end_ && (*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r')) ++p_;
  }
  bool lit(const char* s) {
    const char* q = p_;
    for (; *s; ++s, ++q)
      if (q >= end_ || *q != *s) return false;
    p_ = q;
    return true;
  }
  bool value(Value* v, int depth) {
    if (depth > 256) return fail("nesting too deep");
    if (p_ >= end_) return fail("unexpected end");
    switch (*p_) {
      case '{': return object(v, depth);
      case '[': return array(v, depth);
      case '"': v->type = Value::String; return string(&v->str);
      case 't': v->type = Value::Bool; v->b = true; return lit("true") || fail("bad literal");
      case 'f': v->type = Value::Bool; v->b = false; return lit("false") || fail("bad literal");
      case 'n': v->type = Value::Null; return lit("null") || fail("bad literal");
      default: return number(v);
    }
  }
  bool number(Value* v) {
    const char* s = p_;
    if (p_ < end_ && *p_ == '-') ++p_;
    if (p_ >= end_ || !(*p_ >= '0' && *p_ <= '9')) return fail("bad number");
    while (p_ < end_ && ((*p_ >= '0' && *p_ <= '9') || *p_ == '.' || *p_ == 'e' || *p_ == 'E' ||
                         *p_ == '+' || *p_ == '-'))
      ++p_;
    std::string t(s, p_);
    char* e = nullptr;
    v->type = Value::Number;
    v->num = std::strtod(t.c_str(), &e);
    if (!e || *e) return fail("bad number");
    return true;
  }
  static void utf8(uint32_t c, std::string* o) {
    if (c < 0x80) {
      o->push_back((char)c);
    } else if (c < 0x800) {
      o->push_back((char)(0xC0 | (c >> 6)));
      o->push_back((char)(0x80 | (c & 0x3F)));
    } else if (c < 0x10000) {
      o->push_back((char)(0xE0 | (c >> 12)));
      o->push_back((char)(0x80 | ((c >> 6) & 0x3F)));
      o->push_back((char)(0x80 | (c & 0x3F)));
    } else {
      o->push_back((char)(0xF0 | (c >> 18)));
      o->push_back((char)(0x80 | ((c >> 12) & 0x3F)));
      o->push_back((char)(0x80 | ((c >> 6) & 0x3F)));
      o->push_back((char)(0x80 | (c & 0x3F)));
    }
  }
  bool hex4(uint32_t* c) {
    if (end_ - p_ < 4) return fail("bad \\u escape");
    uint32_t v = 0;
    for (int i = 0; i < 4; ++i) {
      char h = *p_++;
      v <<= 4;
      if (h >= '0' && h <= '9') v |= (uint32_t)(h - '0');
      else if (h >= 'a' && h <= 'f') v |= (uint32_t)(h - 'a' + 10);
      else if (h >= 'A' && h <= 'F') v |= (uint32_t)(h - 'A' + 10);
      else return fail("bad \\u escape");
    }
    *c = v;
    return true;
  }
  bool string(std::string* o) {
    ++p_;  // opening quote
    o->clear();
    while (p_ < end_ && *p_ != '"') {
      char ch = *p_++;
      if ((unsigned char)ch < 0x20) return fail("control character in string");
      if (ch != '\\') {
        o->push_back(ch);
        continue;
      }
      if (p_ >= end_) return fail("unexpected end in string");
      char e = *p_++;
      switch (e) {
        case '"': o->push_back('"'); break;
        case '\\': o->push_back('\\'); break;
        case '/': o->push_back('/'); break;
        case 'b': o->push_back('\b'); break;
        case 'f': o->push_back('\f'); break;
        case 'n': o->push_back('\n'); break;
        case 'r': o->push_back('\r'); break;
        case 't': o->push_back('\t'); break;
        case 'u': {
          uint32_t c;
          if (!hex4(&c)) return false;
          if (c >= 0xD800 && c < 0xDC00 && end_ - p_ >= 6 && p_[0] == '\\' && p_[1] == 'u') {
            p_ += 2;
            uint32_t lo;
            if (!hex4(&lo)) return false;
            if (lo >= 0xDC00 && lo < 0xE000) c = 0x10000 + ((c - 0xD800) << 10) + (lo - 0xDC00);
          }
          utf8(c, o);
          break;
        }
        default: return fail("bad escape");
      }
    }
    if (p_ >= end_) return fail("unterminated string");
    ++p_;
    return true;
  }
  bool array(Value* v, int depth) {
    v->type = Value::Array;
    ++p_;
    ws();
    if (p_ < end_ && *p_ == ']') {
      ++p_;
      return true;
    }
    while (true) {
      v->arr.emplace_back();
      ws();
      if (!value(&v->arr.back(), depth + 1)) return false;
      ws();
      if (p_ >= end_) return fail("unterminated array");
      if (*p_ == ',') {
        ++p_;
        continue;
      }
      if (*p_ == ']') {
        ++p_;
        return true;
      }
      return fail("expected , or ]");
    }
  }
  bool object(Value* v, int depth) {
    v->type = Value::Object;
    ++p_;
    ws();
    if (p_ < end_ && *p_ == '}') {
      ++p_;
      return true;
    }
    while (true) {
      ws();
      if (p_ >= end_ || *p_ != '"') return fail("expected member name");
      std::string key;
      if (!string(&key)) return false;
      ws();
      if (p_ >= end_ || *p_ != ':') return fail("expected :");
      ++p_;
      ws();
      v->obj.emplace_back(std::move(key), Value());
      if (!value(&v->obj.back().second, depth + 1)) return false;
      ws();
      if (p_ >= end_) return fail("unterminated object");
      if (*p_ == ',') {
        ++p_;
        continue;
      }
      if (*p_ == '}') {
        ++p_;
        return true;
      }
      return fail("expected , or }");
    }
  }
};

inline bool parse(const std::string& text, Value* out, std::string* err) {
  Parser p(text.data(), text.size());
  return p.parse(out, err);
}

}  // namespace json
}  // namespace shb

#endif  // SHB_JSON_MIN_H_
