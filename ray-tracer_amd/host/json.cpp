// json.cpp — the scene document model: accessors, a printer and a parser.
//
// The parser is a pushdown loop over an explicit stack of open containers
// (no recursion, so nesting depth costs heap, not call stack). What it
// accepts, the wording of each error and the position an error names are
// the scene format's contract (the reference reader, src/io/json.cpp), and
// so is the printed layout: compact, or indented with arrays of at most four
// scalars kept on one line.
#include "json.hpp"

#include <charconv>
#include <cmath>
#include <cstdio>
#include <string_view>

namespace rtm {

// ------------------------------------------------------------ the model --
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

namespace {
constexpr const char* kTypeNames[] = {"null", "bool", "number", "string", "array", "object"};
void require_type(Json::Type have, Json::Type want) {
  if (have == want) return;
  throw JsonError(std::string("expected ") + kTypeNames[(int)want] + ", got " + kTypeNames[(int)have]);
}
}  // namespace

bool Json::as_bool() const {
  require_type(type_, Type::Bool);
  return b_;
}
double Json::as_number() const {
  require_type(type_, Type::Number);
  return n_;
}
const std::string& Json::as_string() const {
  require_type(type_, Type::String);
  return s_;
}
const std::vector<Json>& Json::items() const {
  require_type(type_, Type::Array);
  return a_;
}
const std::vector<std::pair<std::string, Json>>& Json::members() const {
  require_type(type_, Type::Object);
  return o_;
}

const Json* Json::get(const std::string& key) const {
  if (type_ != Type::Object) return nullptr;
  for (const auto& m : o_)
    if (m.first == key) return &m.second;
  return nullptr;
}

Json& Json::push(Json v) {
  require_type(type_, Type::Array);
  a_.push_back(std::move(v));
  return a_.back();
}

Json& Json::set(const std::string& key, Json v) {
  require_type(type_, Type::Object);
  for (auto& m : o_)
    if (m.first == key) return m.second = std::move(v);
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
    case Type::Object: return o_ == o.o_;
  }
  return false;
}

// ---------------------------------------------------------- the printer --
namespace {

class Printer {
 public:
  explicit Printer(int indent) : indent_(indent) {}
  std::string take() { return std::move(out_); }

  void value(const Json& j, int level) {
    switch (j.type()) {
      case Json::Type::Null: out_ += "null"; return;
      case Json::Type::Bool: out_ += j.as_bool() ? "true" : "false"; return;
      case Json::Type::Number: number(j.as_number()); return;
      case Json::Type::String: quoted(j.as_string()); return;
      case Json::Type::Array: list(j.items(), level); return;
      case Json::Type::Object: record(j.members(), level); return;
    }
  }

 private:
  bool pretty() const { return indent_ >= 0; }
  void line_break(int level) {
    if (!pretty()) return;
    out_.push_back('\n');
    out_.append((size_t)indent_ * (size_t)level, ' ');
  }
  // JSON has no inf or nan; finite values print in their shortest
  // round-trip form.
  void number(double d) {
    if (!std::isfinite(d)) {
      out_ += "null";
      return;
    }
    char buf[32];
    out_.append(buf, std::to_chars(buf, buf + sizeof buf, d).ptr);
  }
  void quoted(const std::string& s) {
    static const char* const kShort[32] = {
        nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, "\\b", "\\t", "\\n",
        nullptr, "\\f",   "\\r",   nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
        nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    out_.push_back('"');
    for (const char ch : s) {
      const auto c = (unsigned char)ch;
      if (c == '"' || c == '\\') {
        out_.push_back('\\');
        out_.push_back(ch);
      } else if (c >= 0x20) {
        out_.push_back(ch);
      } else if (kShort[c]) {
        out_ += kShort[c];
      } else {
        char buf[8];
        std::snprintf(buf, sizeof buf, "\\u%04x", c);
        out_ += buf;
      }
    }
    out_.push_back('"');
  }
  // Up to four scalars (a vector, a colour) share one line.
  static bool one_line(const std::vector<Json>& a) {
    if (a.size() > 4) return false;
    for (const Json& e : a)
      if (e.is_array() || e.is_object() || e.is_string()) return false;
    return true;
  }
  void list(const std::vector<Json>& a, int level) {
    const bool inline_items = one_line(a);
    out_.push_back('[');
    for (size_t i = 0; i < a.size(); ++i) {
      if (i) out_ += inline_items && pretty() ? ", " : ",";
      if (!inline_items) line_break(level + 1);
      value(a[i], level + 1);
    }
    if (!inline_items && !a.empty()) line_break(level);
    out_.push_back(']');
  }
  void record(const std::vector<std::pair<std::string, Json>>& o, int level) {
    out_.push_back('{');
    for (size_t i = 0; i < o.size(); ++i) {
      if (i) out_.push_back(',');
      line_break(level + 1);
      quoted(o[i].first);
      out_ += pretty() ? ": " : ":";
      value(o[i].second, level + 1);
    }
    if (!o.empty()) line_break(level);
    out_.push_back('}');
  }

  int indent_;
  std::string out_;
};

// ----------------------------------------------------------- the parser --
// Reads tokens from the text; every error names the line and column of the
// position the scanner has reached.
class Scanner {
 public:
  explicit Scanner(std::string_view text) : t_(text) {}

  size_t pos() const { return i_; }
  bool at_end() const { return i_ >= t_.size(); }
  char peek() const { return at_end() ? '\0' : t_[i_]; }
  void skip() { ++i_; }
  void skip_space() {
    while (!at_end() && (t_[i_] == ' ' || t_[i_] == '\n' || t_[i_] == '\t' || t_[i_] == '\r')) ++i_;
  }
  void expect(char c) {
    if (peek() != c) fail(std::string("expected '") + c + "'");
    ++i_;
  }

  [[noreturn]] void fail(const std::string& what) const {
    const size_t end = std::min(i_, t_.size());
    size_t line = 1, line_start = 0;
    for (size_t k = 0; k < end; ++k)
      if (t_[k] == '\n') {
        ++line;
        line_start = k + 1;
      }
    throw JsonError("JSON parse error at line " + std::to_string(line) + ", column " +
                    std::to_string(end - line_start + 1) + ": " + what);
  }

  // A scalar value starting at the current position (after white space).
  Json scalar() {
    const char c = peek();
    if (c == '"') return Json::string(quoted());
    if (c == '-' || (c >= '0' && c <= '9')) return Json::number(number());
    if (keyword("true")) return Json::boolean(true);
    if (keyword("false")) return Json::boolean(false);
    if (keyword("null")) return Json::null();
    if (at_end()) fail("unexpected end of input");
    fail(std::string("unexpected character '") + c + "'");
  }

  std::string quoted() {
    expect('"');
    std::string s;
    while (true) {
      if (at_end()) fail("unterminated string");
      const char c = t_[i_++];
      if (c == '"') return s;
      if ((unsigned char)c < 0x20) fail("control character in string");
      if (c == '\\') escape(s);
      else s.push_back(c);
    }
  }

 private:
  bool keyword(std::string_view w) {
    if (t_.compare(i_, w.size(), w) != 0) return false;
    i_ += w.size();
    return true;
  }

  void escape(std::string& s) {
    if (at_end()) fail("unterminated escape");
    const char e = t_[i_++];
    static constexpr std::string_view kFrom = "\"\\/bfnrt", kTo = "\"\\/\b\f\n\r\t";
    if (const size_t k = kFrom.find(e); k != std::string_view::npos) {
      s.push_back(kTo[k]);
      return;
    }
    if (e != 'u') fail("invalid escape");
    uint32_t cp = code_unit();
    const bool high = cp >= 0xD800 && cp < 0xDC00, low = cp >= 0xDC00 && cp < 0xE000;
    if (low) fail("unpaired surrogate");
    if (high) {
      if (!keyword("\\u")) fail("unpaired surrogate");
      const uint32_t second = code_unit();
      if (second < 0xDC00 || second >= 0xE000) fail("invalid low surrogate");
      cp = 0x10000 + (((cp & 0x3FF) << 10) | (second & 0x3FF));
    }
    put_utf8(s, cp);
  }

  // Four hex digits of a \u escape.
  uint32_t code_unit() {
    if (i_ + 4 > t_.size()) fail("truncated \\u escape");
    uint32_t v = 0;
    for (int n = 0; n < 4; ++n) {
      const char h = t_[i_++];
      const int d = h >= '0' && h <= '9' ? h - '0' : h >= 'a' && h <= 'f' ? h - 'a' + 10 : h >= 'A' && h <= 'F' ? h - 'A' + 10 : -1;
      if (d < 0) fail("bad hex digit in \\u escape");
      v = v * 16 + (uint32_t)d;
    }
    return v;
  }

  static void put_utf8(std::string& s, uint32_t cp) {
    if (cp < 0x80) {
      s.push_back((char)cp);
      return;
    }
    const int n = cp < 0x800 ? 2 : cp < 0x10000 ? 3 : 4;
    static constexpr unsigned char kLead[5] = {0, 0, 0xC0, 0xE0, 0xF0};
    char bytes[4];
    for (int k = n - 1; k > 0; --k, cp >>= 6) bytes[k] = (char)(0x80 | (cp & 0x3F));
    bytes[0] = (char)(kLead[n] | cp);
    s.append(bytes, (size_t)n);
  }

  bool digit() const { return peek() >= '0' && peek() <= '9'; }
  void digits() {
    while (digit()) ++i_;
  }

  // RFC 8259 number grammar; the value must be a finite double.
  double number() {
    const size_t first = i_;
    if (peek() == '-') ++i_;
    if (peek() == '0') ++i_;
    else if (digit()) digits();
    else fail("invalid number");
    if (peek() == '.') {
      ++i_;
      if (!digit()) fail("digit expected after '.'");
      digits();
    }
    if (peek() == 'e' || peek() == 'E') {
      ++i_;
      if (peek() == '+' || peek() == '-') ++i_;
      if (!digit()) fail("digit expected in exponent");
      digits();
    }
    double d = 0;
    const char* b = t_.data() + first;
    const char* e = t_.data() + i_;
    const auto r = std::from_chars(b, e, d);
    if (r.ec != std::errc() || r.ptr != e || !std::isfinite(d)) fail("number out of range");
    return d;
  }

  std::string_view t_;
  size_t i_ = 0;
};

// An open array or object and, for an object, the key of the member whose
// value is being read.
struct Open {
  Json node;
  std::string key;
};

// Reads a member key and its colon; the value follows.
void member_key(Scanner& sc, Open& o) {
  sc.skip_space();
  if (sc.peek() != '"') sc.fail("expected string key");
  o.key = sc.quoted();
  if (o.node.get(o.key)) sc.fail("duplicate key '" + o.key + "'");
  sc.skip_space();
  sc.expect(':');
}

Json parse_document(std::string_view text, int max_depth) {
  Scanner sc(text);
  std::vector<Open> open;
  while (true) {
    // A value is due at depth open.size(); the limit applies before any
    // white space in front of it is skipped.
    if ((int)open.size() > max_depth) sc.fail("nesting too deep");
    sc.skip_space();
    Json done;
    const char c = sc.peek();
    if (c == '[' || c == '{') {
      sc.skip();
      const char close = c == '[' ? ']' : '}';
      open.push_back(Open{c == '[' ? Json::array() : Json::object(), {}});
      sc.skip_space();
      if (sc.peek() != close) {
        if (c == '{') member_key(sc, open.back());
        continue;  // the first element or member value
      }
      sc.skip();
      done = std::move(open.back().node);
      open.pop_back();
    } else {
      done = sc.scalar();
    }
    // Attach the finished value to its container; close every container
    // that ends here.
    while (true) {
      if (open.empty()) {
        sc.skip_space();
        if (!sc.at_end()) sc.fail("trailing characters after JSON value");
        return done;
      }
      Open& top = open.back();
      if (top.node.is_array()) top.node.push(std::move(done));
      else top.node.set(top.key, std::move(done));
      sc.skip_space();
      if (sc.peek() == ',') {
        sc.skip();
        if (top.node.is_object()) member_key(sc, top);
        break;  // the next element or member value
      }
      sc.expect(top.node.is_array() ? ']' : '}');
      done = std::move(top.node);
      open.pop_back();
    }
  }
}

}  // namespace

std::string Json::dump(int indent) const {
  Printer p(indent);
  p.value(*this, 0);
  return p.take();
}

Json Json::parse(const std::string& text, int max_depth) { return parse_document(text, max_depth); }

}  // namespace rtm
