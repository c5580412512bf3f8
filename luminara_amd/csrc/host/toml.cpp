// TOML subset reader (see toml.h).
#include "toml.h"

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <limits>

namespace lmn {

const TomlValue* TomlValue::get(const std::string& key) const {
  for (auto& kv : table)
    if (kv.first == key) return &kv.second;
  return nullptr;
}
TomlValue* TomlValue::get_mut(const std::string& key) {
  for (auto& kv : table)
    if (kv.first == key) return &kv.second;
  return nullptr;
}
TomlValue& TomlValue::put(const std::string& key, TomlValue v) {
  table.emplace_back(key, std::move(v));
  return table.back().second;
}

namespace {

struct Parser {
  const std::string& s;
  size_t i = 0;
  int line = 1;
  std::string err;

  explicit Parser(const std::string& text) : s(text) {}

  bool fail(const std::string& m) {
    if (err.empty()) err = "line " + std::to_string(line) + ": " + m;
    return false;
  }
  bool eof() const { return i >= s.size(); }
  char peek() const { return eof() ? '\0' : s[i]; }
  void skip_ws() {  // spaces and tabs only
    while (!eof() && (s[i] == ' ' || s[i] == '\t')) ++i;
  }
  void skip_comment() {
    if (peek() == '#')
      while (!eof() && s[i] != '\n') ++i;
  }
  // whitespace, newlines and comments (inside arrays)
  void skip_all() {
    for (;;) {
      skip_ws();
      if (peek() == '#') {
        skip_comment();
        continue;
      }
      if (peek() == '\r') {
        ++i;
        continue;
      }
      if (peek() == '\n') {
        ++i;
        ++line;
        continue;
      }
      break;
    }
  }
  bool end_of_line() {
    skip_ws();
    skip_comment();
    if (peek() == '\r') ++i;
    if (eof()) return true;
    if (peek() == '\n') {
      ++i;
      ++line;
      return true;
    }
    return fail("expected end of line");
  }

  static bool bare_char(char c) { return std::isalnum((unsigned char)c) || c == '_' || c == '-'; }

  bool parse_key_part(std::string& out) {
    skip_ws();
    if (peek() == '"' || peek() == '\'') {
      TomlValue v;
      if (!parse_string(v)) return false;
      out = v.str;
      return true;
    }
    size_t b = i;
    while (!eof() && bare_char(s[i])) ++i;
    if (i == b) return fail("expected a key");
    out = s.substr(b, i - b);
    return true;
  }
  bool parse_key(std::vector<std::string>& parts) {
    parts.clear();
    for (;;) {
      std::string k;
      if (!parse_key_part(k)) return false;
      parts.push_back(k);
      skip_ws();
      if (peek() == '.') {
        ++i;
        continue;
      }
      return true;
    }
  }

  bool parse_string(TomlValue& v) {
    v.kind = TomlValue::STRING;
    char q = s[i];
    bool multi = s.compare(i, 3, std::string(3, q)) == 0;
    if (multi) {
      i += 3;
      if (peek() == '\n') {
        ++i;
        ++line;
      }
    } else {
      ++i;
    }
    std::string out;
    for (;;) {
      if (eof()) return fail("unterminated string");
      char c = s[i];
      if (multi && s.compare(i, 3, std::string(3, q)) == 0) {
        i += 3;
        break;
      }
      if (!multi && c == q) {
        ++i;
        break;
      }
      if (!multi && c == '\n') return fail("newline in string");
      if (c == '\n') ++line;
      if (q == '"' && c == '\\') {
        ++i;
        char e = peek();
        ++i;
        switch (e) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case 'r': out += '\r'; break;
          case '\\': out += '\\'; break;
          case '"': out += '"'; break;
          case 'b': out += '\b'; break;
          case 'f': out += '\f'; break;
          case 'u':
          case 'U': {
            int n = e == 'u' ? 4 : 8;
            if (i + n > s.size()) return fail("bad unicode escape");
            unsigned long cp = std::strtoul(s.substr(i, n).c_str(), nullptr, 16);
            i += n;
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
            break;
          }
          case '\n':  // line-ending backslash in multi-line strings
            ++line;
            while (!eof() && std::isspace((unsigned char)s[i])) {
              if (s[i] == '\n') ++line;
              ++i;
            }
            break;
          default: return fail("bad escape in string");
        }
        continue;
      }
      out += c;
      ++i;
    }
    v.str = out;
    return true;
  }

  bool parse_number_or_bool(TomlValue& v) {
    size_t b = i;
    while (!eof() && (std::isalnum((unsigned char)s[i]) || s[i] == '+' || s[i] == '-' || s[i] == '.' ||
                      s[i] == '_'))
      ++i;
    std::string tok = s.substr(b, i - b);
    if (tok == "true" || tok == "false") {
      v.kind = TomlValue::BOOLEAN;
      v.boolean = tok == "true";
      return true;
    }
    std::string t;
    for (char c : tok)
      if (c != '_') t += c;
    if (t.empty()) return fail("expected a value");
    std::string body = (t[0] == '+' || t[0] == '-') ? t.substr(1) : t;
    bool neg = t[0] == '-';
    if (body == "inf" || body == "nan") {
      v.kind = TomlValue::FLOAT;
      v.flt = body == "inf" ? std::numeric_limits<double>::infinity() : std::numeric_limits<double>::quiet_NaN();
      if (neg) v.flt = -v.flt;
      return true;
    }
    if (body.size() > 2 && body[0] == '0' && (body[1] == 'x' || body[1] == 'o' || body[1] == 'b')) {
      int base = body[1] == 'x' ? 16 : body[1] == 'o' ? 8 : 2;
      char* end = nullptr;
      long long x = std::strtoll(body.c_str() + 2, &end, base);
      if (*end) return fail("bad integer '" + tok + "'");
      v.kind = TomlValue::INTEGER;
      v.integer = neg ? -x : x;
      return true;
    }
    bool is_float = t.find_first_of(".eE") != std::string::npos;
    char* end = nullptr;
    if (is_float) {
      v.kind = TomlValue::FLOAT;
      v.flt = std::strtod(t.c_str(), &end);
    } else {
      v.kind = TomlValue::INTEGER;
      v.integer = std::strtoll(t.c_str(), &end, 10);
    }
    if (end == nullptr || *end != '\0') return fail("bad value '" + tok + "'");
    return true;
  }

  bool parse_inline_table(TomlValue& v) {
    v.kind = TomlValue::TABLE;
    ++i;  // '{'
    skip_ws();
    if (peek() == '}') {
      ++i;
      return true;
    }
    for (;;) {
      std::vector<std::string> key;
      if (!parse_key(key)) return false;
      skip_ws();
      if (peek() != '=') return fail("expected '=' in inline table");
      ++i;
      skip_ws();
      TomlValue val;
      if (!parse_value(val)) return false;
      if (!assign(v, key, std::move(val))) return false;
      skip_ws();
      if (peek() == ',') {
        ++i;
        skip_ws();
        continue;
      }
      if (peek() == '}') {
        ++i;
        return true;
      }
      return fail("expected ',' or '}' in inline table");
    }
  }

  bool parse_array(TomlValue& v) {
    v.kind = TomlValue::ARRAY;
    ++i;  // '['
    for (;;) {
      skip_all();
      if (peek() == ']') {
        ++i;
        return true;
      }
      TomlValue e;
      if (!parse_value(e)) return false;
      v.array.push_back(std::move(e));
      skip_all();
      if (peek() == ',') {
        ++i;
        continue;
      }
      if (peek() == ']') {
        ++i;
        return true;
      }
      return fail("expected ',' or ']' in array");
    }
  }

  bool parse_value(TomlValue& v) {
    skip_ws();
    char c = peek();
    if (c == '"' || c == '\'') return parse_string(v);
    if (c == '[') return parse_array(v);
    if (c == '{') return parse_inline_table(v);
    return parse_number_or_bool(v);
  }

  // Assigns `val` at dotted key path below table t (creating implicit tables).
  bool assign(TomlValue& t, const std::vector<std::string>& key, TomlValue val) {
    TomlValue* cur = &t;
    for (size_t k = 0; k + 1 < key.size(); ++k) {
      TomlValue* nxt = cur->get_mut(key[k]);
      if (!nxt) {
        TomlValue tab;
        tab.kind = TomlValue::TABLE;
        nxt = &cur->put(key[k], std::move(tab));
      }
      if (nxt->kind != TomlValue::TABLE) return fail("key '" + key[k] + "' is not a table");
      cur = nxt;
    }
    if (cur->get(key.back())) return fail("duplicate key '" + key.back() + "'");
    cur->put(key.back(), std::move(val));
    return true;
  }

  // Resolves a [header] / [[header]] path; returns the table to fill.
  TomlValue* open_header(TomlValue& root, const std::vector<std::string>& key, bool array_of_tables) {
    TomlValue* cur = &root;
    for (size_t k = 0; k < key.size(); ++k) {
      bool last = k + 1 == key.size();
      TomlValue* nxt = cur->get_mut(key[k]);
      if (last && array_of_tables) {
        if (!nxt) {
          TomlValue arr;
          arr.kind = TomlValue::ARRAY;
          arr.is_array_of_tables = true;
          nxt = &cur->put(key[k], std::move(arr));
        }
        if (nxt->kind != TomlValue::ARRAY || !nxt->is_array_of_tables) {
          fail("'" + key[k] + "' is not an array of tables");
          return nullptr;
        }
        TomlValue tab;
        tab.kind = TomlValue::TABLE;
        nxt->array.push_back(std::move(tab));
        return &nxt->array.back();
      }
      if (!nxt) {
        TomlValue tab;
        tab.kind = TomlValue::TABLE;
        nxt = &cur->put(key[k], std::move(tab));
      }
      if (nxt->kind == TomlValue::ARRAY && nxt->is_array_of_tables) {
        nxt = &nxt->array.back();
      } else if (nxt->kind != TomlValue::TABLE) {
        fail("'" + key[k] + "' is not a table");
        return nullptr;
      }
      if (last) {
        if (nxt->defined_by_header) {
          fail("table '" + key[k] + "' defined twice");
          return nullptr;
        }
        nxt->defined_by_header = true;
      }
      cur = nxt;
    }
    return cur;
  }

  bool run(TomlValue& root) {
    root = TomlValue();
    root.kind = TomlValue::TABLE;
    TomlValue* cur = &root;
    for (;;) {
      skip_all();
      if (eof()) return true;
      if (peek() == '[') {
        bool aot = s.compare(i, 2, "[[") == 0;
        i += aot ? 2 : 1;
        std::vector<std::string> key;
        if (!parse_key(key)) return false;
        skip_ws();
        if (aot ? s.compare(i, 2, "]]") != 0 : peek() != ']') return fail("unterminated table header");
        i += aot ? 2 : 1;
        cur = open_header(root, key, aot);
        if (!cur) return false;
        if (!end_of_line()) return false;
        continue;
      }
      std::vector<std::string> key;
      if (!parse_key(key)) return false;
      skip_ws();
      if (peek() != '=') return fail("expected '='");
      ++i;
      TomlValue val;
      if (!parse_value(val)) return false;
      if (!assign(*cur, key, std::move(val))) return false;
      if (!end_of_line()) return false;
    }
  }
};

}  // namespace

bool toml_parse(const std::string& text, TomlValue& root, std::string& err) {
  Parser p(text);
  bool ok = p.run(root);
  if (!ok) err = p.err;
  return ok;
}

}  // namespace lmn
