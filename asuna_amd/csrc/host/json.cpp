#include "json.h"

#include <cmath>
#include <cstdlib>
#include <sstream>

namespace ash {

class JsonParser {
 public:
  explicit JsonParser(const std::string& s) : s_(s) {}
  Json parseDoc() {
    Json v = value();
    ws();
    if (p_ != s_.size()) fail("trailing characters");
    return v;
  }

 private:
  const std::string& s_;
  size_t p_ = 0;

  [[noreturn]] void fail(const char* what) {
    size_t line = 1;
    for (size_t i = 0; i < p_ && i < s_.size(); ++i) line += s_[i] == '\n';
    std::ostringstream os;
    os << "json parse error at line " << line << ": " << what;
    throw JsonError(os.str());
  }
  void ws() {
    while (p_ < s_.size()) {
      char c = s_[p_];
      if (c == ' ' || c == '\t' || c == '\n' || c == '\r') {
        ++p_;
      } else if (c == '/' && p_ + 1 < s_.size() && s_[p_ + 1] == '/') {  // tolerate // comments
        while (p_ < s_.size() && s_[p_] != '\n') ++p_;
      } else {
        break;
      }
    }
  }
  bool match(const char* lit) {
    size_t n = 0;
    while (lit[n]) ++n;
    if (s_.compare(p_, n, lit) == 0) {
      p_ += n;
      return true;
    }
    return false;
  }
  Json value() {
    ws();
    if (p_ >= s_.size()) fail("unexpected end");
    char c = s_[p_];
    Json v;
    if (c == '{') {
      v.type_ = Json::Type::Object;
      ++p_;
      ws();
      if (p_ < s_.size() && s_[p_] == '}') {
        ++p_;
        return v;
      }
      for (;;) {
        ws();
        if (p_ >= s_.size() || s_[p_] != '"') fail("expected key");
        std::string k = string();
        ws();
        if (p_ >= s_.size() || s_[p_] != ':') fail("expected ':'");
        ++p_;
        v.obj_[k] = value();
        ws();
        if (p_ < s_.size() && s_[p_] == ',') {
          ++p_;
          continue;
        }
        if (p_ < s_.size() && s_[p_] == '}') {
          ++p_;
          break;
        }
        fail("expected ',' or '}'");
      }
    } else if (c == '[') {
      v.type_ = Json::Type::Array;
      ++p_;
      ws();
      if (p_ < s_.size() && s_[p_] == ']') {
        ++p_;
        return v;
      }
      for (;;) {
        v.arr_.push_back(value());
        ws();
        if (p_ < s_.size() && s_[p_] == ',') {
          ++p_;
          continue;
        }
        if (p_ < s_.size() && s_[p_] == ']') {
          ++p_;
          break;
        }
        fail("expected ',' or ']'");
      }
    } else if (c == '"') {
      v.type_ = Json::Type::String;
      v.str_ = string();
    } else if (match("true")) {
      v.type_ = Json::Type::Bool;
      v.bool_ = true;
    } else if (match("false")) {
      v.type_ = Json::Type::Bool;
      v.bool_ = false;
    } else if (match("null")) {
      v.type_ = Json::Type::Null;
    } else if (c == '-' || (c >= '0' && c <= '9')) {
      const char* b = s_.c_str() + p_;
      char* e = nullptr;
      double d = std::strtod(b, &e);
      if (e == b) fail("bad number");
      p_ += size_t(e - b);
      v.type_ = Json::Type::Number;
      v.num_ = d;
    } else {
      fail("unexpected character");
    }
    return v;
  }
  static void utf8(std::string& o, unsigned cp) {
    if (cp < 0x80) {
      o += char(cp);
    } else if (cp < 0x800) {
      o += char(0xC0 | (cp >> 6));
      o += char(0x80 | (cp & 0x3F));
    } else if (cp < 0x10000) {
      o += char(0xE0 | (cp >> 12));
      o += char(0x80 | ((cp >> 6) & 0x3F));
      o += char(0x80 | (cp & 0x3F));
    } else {
      o += char(0xF0 | (cp >> 18));
      o += char(0x80 | ((cp >> 12) & 0x3F));
      o += char(0x80 | ((cp >> 6) & 0x3F));
      o += char(0x80 | (cp & 0x3F));
    }
  }
  unsigned hex4() {
    if (p_ + 4 > s_.size()) fail("bad \\u escape");
    unsigned v = 0;
    for (int i = 0; i < 4; ++i) {
      char c = s_[p_++];
      v <<= 4;
      if (c >= '0' && c <= '9') v |= unsigned(c - '0');
      else if (c >= 'a' && c <= 'f') v |= unsigned(c - 'a' + 10);
      else if (c >= 'A' && c <= 'F') v |= unsigned(c - 'A' + 10);
      else fail("bad hex digit");
    }
    return v;
  }
  std::string string() {
    ++p_;  // opening quote
    std::string o;
    while (p_ < s_.size() && s_[p_] != '"') {
      char c = s_[p_++];
      if (c != '\\') {
        o += c;
        continue;
      }
      if (p_ >= s_.size()) fail("bad escape");
      char e = s_[p_++];
      switch (e) {
        case '"': o += '"'; break;
        case '\\': o += '\\'; break;
        case '/': o += '/'; break;
        case 'b': o += '\b'; break;
        case 'f': o += '\f'; break;
        case 'n': o += '\n'; break;
        case 'r': o += '\r'; break;
        case 't': o += '\t'; break;
        case 'u': {
          unsigned cp = hex4();
          if (cp >= 0xD800 && cp < 0xDC00 && p_ + 1 < s_.size() && s_[p_] == '\\' && s_[p_ + 1] == 'u') {
            p_ += 2;
            unsigned lo = hex4();
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          }
          utf8(o, cp);
          break;
        }
        default: fail("bad escape");
      }
    }
    if (p_ >= s_.size()) fail("unterminated string");
    ++p_;
    return o;
  }
};

Json Json::parse(const std::string& text) { return JsonParser(text).parseDoc(); }

bool Json::contains(const std::string& key) const { return type_ == Type::Object && obj_.count(key) != 0; }

const Json& Json::operator[](const std::string& key) const {
  if (type_ != Type::Object) throw JsonError("json: not an object when reading key '" + key + "'");
  auto it = obj_.find(key);
  if (it == obj_.end()) throw JsonError("json: missing key '" + key + "'");
  return it->second;
}
const Json& Json::operator[](size_t i) const {
  if (type_ != Type::Array) throw JsonError("json: not an array");
  if (i >= arr_.size()) throw JsonError("json: array index out of range");
  return arr_[i];
}
size_t Json::size() const {
  if (type_ == Type::Array) return arr_.size();
  if (type_ == Type::Object) return obj_.size();
  return 0;
}
double Json::num() const {
  if (type_ == Type::Number) return num_;
  if (type_ == Type::Bool) return bool_ ? 1.0 : 0.0;
  throw JsonError("json: expected a number, got " + dump());
}
bool Json::b() const {
  if (type_ == Type::Bool) return bool_;
  if (type_ == Type::Number) return num_ != 0;
  throw JsonError("json: expected a bool, got " + dump());
}
const std::string& Json::str() const {
  if (type_ != Type::String) throw JsonError("json: expected a string, got " + dump());
  return str_;
}
const std::vector<Json>& Json::arr() const {
  if (type_ != Type::Array) throw JsonError("json: expected an array");
  return arr_;
}
const std::map<std::string, Json>& Json::obj() const {
  if (type_ != Type::Object) throw JsonError("json: expected an object");
  return obj_;
}
std::string Json::dump() const {
  std::ostringstream os;
  switch (type_) {
    case Type::Null: os << "null"; break;
    case Type::Bool: os << (bool_ ? "true" : "false"); break;
    case Type::Number: os << num_; break;
    case Type::String: os << '"' << str_ << '"'; break;
    case Type::Array: {
      os << '[';
      for (size_t i = 0; i < arr_.size(); ++i) os << (i ? "," : "") << arr_[i].dump();
      os << ']';
      break;
    }
    case Type::Object: {
      os << '{';
      bool first = true;
      for (auto& kv : obj_) {
        os << (first ? "" : ",") << '"' << kv.first << "\":" << kv.second.dump();
        first = false;
      }
      os << '}';
      break;
    }
  }
  return os.str();
}

}  // namespace ash
