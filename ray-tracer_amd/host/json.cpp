// json.cpp — parser and writer for the scene document model.
#include "json.hpp"

#include <charconv>
#include <cmath>
#include <cstdio>
#include <string_view>

namespace rtm {

Json Json::boolean(bool b) {
  Json j;
  j.type_ = Type::Bool;
  j.b_ = b;
  return j;
}
Json Json::number(double d) {
  Json j;
  j.type_ = Type::Number;
  j.n_ = d;
  return j;
}
Json Json::string(std::string s) {
  Json j;
  j.type_ = Type::String;
  j.s_ = std::move(s);
  return j;
}
Json Json::array() {
  Json j;
  j.type_ = Type::Array;
  return j;
}
Json Json::object() {
  Json j;
  j.type_ = Type::Object;
  return j;
}

static const char* type_name(Json::Type t) {
  switch (t) {
    case Json::Type::Null: return "null";
    case Json::Type::Bool: return "bool";
    case Json::Type::Number: return "number";
    case Json::Type::String: return "string";
    case Json::Type::Array: return "array";
    default: return "object";
  }
}

static void need(Json::Type have, Json::Type want) {
  if (have != want) throw JsonError(std::string("expected ") + type_name(want) + ", got " + type_name(have));
}

bool Json::as_bool() const { need(type_, Type::Bool); return b_; }
double Json::as_number() const { need(type_, Type::Number); return n_; }
const std::string& Json::as_string() const { need(type_, Type::String); return s_; }
const std::vector<Json>& Json::items() const { need(type_, Type::Array); return a_; }
const std::vector<std::pair<std::string, Json>>& Json::members() const { need(type_, Type::Object); return o_; }

const Json* Json::get(const std::string& key) const {
  if (type_ != Type::Object) return nullptr;
  for (const auto& kv : o_)
    if (kv.first == key) return &kv.second;
  return nullptr;
}

Json& Json::push(Json v) {
  need(type_, Type::Array);
  a_.push_back(std::move(v));
  return a_.back();
}

Json& Json::set(const std::string& key, Json v) {
  need(type_, Type::Object);
  for (auto& kv : o_)
    if (kv.first == key) {
      kv.second = std::move(v);
      return kv.second;
    }
  o_.emplace_back(key, std::move(v));
  return o_.back().second;
}

bool Json::operator==(const Json& o) const {
  if (type_ != o.type_) return false;
  switch (type_) {
    case Type::Null: return true;
    case Type::Bool: return b_ == o.b_;
    case Type::Number: return n_ == o.n_;
    case Type::String: return s_ == o.s_;
    case Type::Array: return a_ == o.a_;
    default: return o_ == o.o_;
  }
}

// ---------------------------------------------------------------- writer --
namespace {

void write_string(std::string& out, const std::string& s) {
  out += '"';
  for (unsigned char c : s) {
    switch (c) {
      case '"': out += "\\\""; break;
      case '\\': out += "\\\\"; break;
      case '\n': out += "\\n"; break;
      case '\r': out += "\\r"; break;
      case '\t': out += "\\t"; break;
      case '\b': out += "\\b"; break;
      case '\f': out += "\\f"; break;
      default:
        if (c < 0x20) {
          char buf[8];
          std::snprintf(buf, sizeof buf, "\\u%04x", c);
          out += buf;
        } else {
          out += (char)c;
        }
    }
  }
  out += '"';
}

void write(std::string& out, const Json& j, int indent, int level) {
  auto nl = [&](int lvl) {
    if (indent < 0) return;
    out += '\n';
    out.append((size_t)(indent * lvl), ' ');
  };
  switch (j.type()) {
    case Json::Type::Null: out += "null"; break;
    case Json::Type::Bool: out += j.as_bool() ? "true" : "false"; break;
    case Json::Type::Number: {
      const double d = j.as_number();
      if (!std::isfinite(d)) {
        out += "null";
        break;
      }
      char buf[32];
      const auto r = std::to_chars(buf, buf + sizeof buf, d);
      out.append(buf, r.ptr);
      break;
    }
    case Json::Type::String: write_string(out, j.as_string()); break;
    case Json::Type::Array: {
      const auto& a = j.items();
      bool scalars = a.size() <= 4;  // vectors and colours stay on one line
      for (const auto& e : a) scalars = scalars && (e.is_number() || e.is_bool() || e.is_null());
      out += '[';
      for (size_t i = 0; i < a.size(); ++i) {
        if (i) out += (scalars && indent >= 0) ? ", " : ",";
        if (!scalars) nl(level + 1);
        write(out, a[i], indent, level + 1);
      }
      if (!scalars && !a.empty()) nl(level);
      out += ']';
      break;
    }
    case Json::Type::Object: {
      const auto& o = j.members();
      out += '{';
      for (size_t i = 0; i < o.size(); ++i) {
        if (i) out += ',';
        nl(level + 1);
        write_string(out, o[i].first);
        out += indent >= 0 ? ": " : ":";
        write(out, o[i].second, indent, level + 1);
      }
      if (!o.empty()) nl(level);
      out += '}';
      break;
    }
  }
}

// ---------------------------------------------------------------- parser --
class Parser {
 public:
  Parser(const std::string& s, int max_depth) : s_(s), max_depth_(max_depth) {}

  Json document() {
    Json v = value(0);
    ws();
    if (p_ != s_.size()) error("trailing characters after JSON value");
    return v;
  }

 private:
  [[noreturn]] void error(const std::string& what) const {
    size_t line = 1, col = 1;
    for (size_t i = 0; i < p_ && i < s_.size(); ++i) {
      if (s_[i] == '\n') {
        ++line;
        col = 1;
      } else {
        ++col;
      }
    }
    throw JsonError("JSON parse error at line " + std::to_string(line) + ", column " + std::to_string(col) + ": " +
                    what);
  }
  void ws() {
    while (p_ < s_.size() && (s_[p_] == ' ' || s_[p_] == '\t' || s_[p_] == '\n' || s_[p_] == '\r')) ++p_;
  }
  char cur() const { return p_ < s_.size() ? s_[p_] : '\0'; }
  bool is_digit() const { return cur() >= '0' && cur() <= '9'; }
  void eat(char c) {
    if (cur() != c) error(std::string("expected '") + c + "'");
    ++p_;
  }
  bool word(std::string_view w) {
    if (std::string_view(s_).substr(p_, w.size()) == w) {
      p_ += w.size();
      return true;
    }
    return false;
  }

  Json value(int depth) {
    if (depth > max_depth_) error("nesting too deep");
    ws();
    switch (cur()) {
      case '{': return object(depth);
      case '[': return array(depth);
      case '"': return Json::string(str());
      case '\0':
        if (p_ >= s_.size()) error("unexpected end of input");
        break;
      default: break;
    }
    if (cur() == '-' || is_digit()) return Json::number(num());
    if (word("true")) return Json::boolean(true);
    if (word("false")) return Json::boolean(false);
    if (word("null")) return Json::null();
    error(std::string("unexpected character '") + cur() + "'");
  }

  Json object(int depth) {
    eat('{');
    Json o = Json::object();
    ws();
    if (cur() == '}') {
      ++p_;
      return o;
    }
    for (;;) {
      ws();
      if (cur() != '"') error("expected string key");
      std::string key = str();
      if (o.get(key)) error("duplicate key '" + key + "'");
      ws();
      eat(':');
      o.set(key, value(depth + 1));
      ws();
      if (cur() == ',') {
        ++p_;
        continue;
      }
      eat('}');
      return o;
    }
  }

  Json array(int depth) {
    eat('[');
    Json a = Json::array();
    ws();
    if (cur() == ']') {
      ++p_;
      return a;
    }
    for (;;) {
      a.push(value(depth + 1));
      ws();
      if (cur() == ',') {
        ++p_;
        continue;
      }
      eat(']');
      return a;
    }
  }

  uint32_t hex4() {
    if (p_ + 4 > s_.size()) error("truncated \\u escape");
    uint32_t v = 0;
    for (int i = 0; i < 4; ++i) {
      const char h = s_[p_++];
      int d;
      if (h >= '0' && h <= '9') d = h - '0';
      else if (h >= 'a' && h <= 'f') d = h - 'a' + 10;
      else if (h >= 'A' && h <= 'F') d = h - 'A' + 10;
      else error("bad hex digit in \\u escape");
      v = (v << 4) | (uint32_t)d;
    }
    return v;
  }

  static void utf8(std::string& out, uint32_t cp) {
    if (cp < 0x80) {
      out += (char)cp;
    } else if (cp < 0x800) {
      out += (char)(0xC0 | (cp >> 6));
      out += (char)(0x80 | (cp & 0x3F));
    } else if (cp < 0x10000) {
      out += (char)(0xE0 | (cp >> 12));
      out += (char)(0x80 | ((cp >> 6) & 0x3F));
      out += (char)(0x80 | (cp & 0x3F));
    } else {
      out += (char)(0xF0 | (cp >> 18));
      out += (char)(0x80 | ((cp >> 12) & 0x3F));
      out += (char)(0x80 | ((cp >> 6) & 0x3F));
      out += (char)(0x80 | (cp & 0x3F));
    }
  }

  std::string str() {
    eat('"');
    std::string out;
    for (;;) {
      if (p_ >= s_.size()) error("unterminated string");
      const char c = s_[p_++];
      if (c == '"') return out;
      if ((unsigned char)c < 0x20) error("control character in string");
      if (c != '\\') {
        out += c;
        continue;
      }
      if (p_ >= s_.size()) error("unterminated escape");
      const char e = s_[p_++];
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
            if (!word("\\u")) error("unpaired surrogate");
            const uint32_t lo = hex4();
            if (lo < 0xDC00 || lo > 0xDFFF) error("invalid low surrogate");
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          } else if (cp >= 0xDC00 && cp <= 0xDFFF) {
            error("unpaired surrogate");
          }
          utf8(out, cp);
          break;
        }
        default: error("invalid escape");
      }
    }
  }

  double num() {
    const size_t start = p_;
    if (cur() == '-') ++p_;
    if (cur() == '0') {
      ++p_;
    } else if (cur() >= '1' && cur() <= '9') {
      while (is_digit()) ++p_;
    } else {
      error("invalid number");
    }
    if (cur() == '.') {
      ++p_;
      if (!is_digit()) error("digit expected after '.'");
      while (is_digit()) ++p_;
    }
    if (cur() == 'e' || cur() == 'E') {
      ++p_;
      if (cur() == '+' || cur() == '-') ++p_;
      if (!is_digit()) error("digit expected in exponent");
      while (is_digit()) ++p_;
    }
    double d = 0;
    const auto r = std::from_chars(s_.data() + start, s_.data() + p_, d);
    if (r.ec != std::errc() || r.ptr != s_.data() + p_ || !std::isfinite(d)) error("number out of range");
    return d;
  }

  const std::string& s_;
  size_t p_ = 0;
  int max_depth_;
};

}  // namespace

std::string Json::dump(int indent) const {
  std::string out;
  write(out, *this, indent, 0);
  return out;
}

Json Json::parse(const std::string& text, int max_depth) { return Parser(text, max_depth).document(); }

}  // namespace rtm
