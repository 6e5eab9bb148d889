// A small JSON value and parser (RFC 8259 subset: no surrogate pairs), used
// by the ONTC container reader and the texture manifest front end.
#pragma once

#include <cctype>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

namespace openntc
{
namespace json
{
struct Json
{
  enum Kind
  {
    Null,
    Bool,
    Number,
    String,
    Array,
    Object
  } kind = Null;
  bool b = false;
  double num = 0.0;
  bool is_int = false;
  long long inum = 0;
  std::string str;
  std::vector<Json> arr;
  std::vector<std::pair<std::string, Json>> obj;

  const Json* find(const char* k) const
  {
    if (kind != Object) return nullptr;
    for (const auto& kv : obj)
      if (kv.first == k) return &kv.second;
    return nullptr;
  }
};

class Parser
{
public:
  Parser(const char* b, const char* e) : p_(b), e_(e) {}
  bool Parse(Json& out)
  {
    if (!Value(out, 0)) return false;
    Ws();
    return p_ == e_;
  }

private:
  const char* p_;
  const char* e_;
  void Ws()
  {
    while (p_ < e_ && (*p_ == ' ' || *p_ == '\n' || *p_ == '\r' || *p_ == '\t')) ++p_;
  }
  bool Lit(const char* s)
  {
    const size_t n = strlen(s);
    if ((size_t)(e_ - p_) < n || strncmp(p_, s, n) != 0) return false;
    p_ += n;
    return true;
  }
  bool Str(std::string& out)
  {
    if (p_ >= e_ || *p_ != '"') return false;
    ++p_;
    while (p_ < e_ && *p_ != '"')
    {
      char c = *p_++;
      if (c == '\\')
      {
        if (p_ >= e_) return false;
        char esc = *p_++;
        switch (esc)
        {
        case '"': out += '"'; break;
        case '\\': out += '\\'; break;
        case '/': out += '/'; break;
        case 'b': out += '\b'; break;
        case 'f': out += '\f'; break;
        case 'n': out += '\n'; break;
        case 'r': out += '\r'; break;
        case 't': out += '\t'; break;
        case 'u':
        {
          if (e_ - p_ < 4) return false;
          unsigned v = 0;
          for (int i = 0; i < 4; ++i)
          {
            const char h = p_[i];
            v <<= 4;
            if (h >= '0' && h <= '9') v |= (unsigned)(h - '0');
            else if (h >= 'a' && h <= 'f') v |= (unsigned)(h - 'a' + 10);
            else if (h >= 'A' && h <= 'F') v |= (unsigned)(h - 'A' + 10);
            else return false;
          }
          p_ += 4;
          if (v < 0x80) out += (char)v;
          else if (v < 0x800)
          {
            out += (char)(0xC0 | (v >> 6));
            out += (char)(0x80 | (v & 0x3F));
          }
          else
          {
            out += (char)(0xE0 | (v >> 12));
            out += (char)(0x80 | ((v >> 6) & 0x3F));
            out += (char)(0x80 | (v & 0x3F));
          }
          break;
        }
        default: return false;
        }
      }
      else
        out += c;
    }
    if (p_ >= e_) return false;
    ++p_;
    return true;
  }
  bool Value(Json& v, int depth)
  {
    if (depth > 64) return false;
    Ws();
    if (p_ >= e_) return false;
    const char c = *p_;
    if (c == '{')
    {
      ++p_;
      v.kind = Json::Object;
      Ws();
      if (p_ < e_ && *p_ == '}')
      {
        ++p_;
        return true;
      }
      for (;;)
      {
        Ws();
        std::string k;
        if (!Str(k)) return false;
        Ws();
        if (p_ >= e_ || *p_ != ':') return false;
        ++p_;
        Json child;
        if (!Value(child, depth + 1)) return false;
        v.obj.emplace_back(std::move(k), std::move(child));
        Ws();
        if (p_ < e_ && *p_ == ',')
        {
          ++p_;
          continue;
        }
        if (p_ < e_ && *p_ == '}')
        {
          ++p_;
          return true;
        }
        return false;
      }
    }
    if (c == '[')
    {
      ++p_;
      v.kind = Json::Array;
      Ws();
      if (p_ < e_ && *p_ == ']')
      {
        ++p_;
        return true;
      }
      for (;;)
      {
        Json child;
        if (!Value(child, depth + 1)) return false;
        v.arr.push_back(std::move(child));
        Ws();
        if (p_ < e_ && *p_ == ',')
        {
          ++p_;
          continue;
        }
        if (p_ < e_ && *p_ == ']')
        {
          ++p_;
          return true;
        }
        return false;
      }
    }
    if (c == '"')
    {
      v.kind = Json::String;
      return Str(v.str);
    }
    if (Lit("true"))
    {
      v.kind = Json::Bool;
      v.b = true;
      return true;
    }
    if (Lit("false"))
    {
      v.kind = Json::Bool;
      return true;
    }
    if (Lit("null")) return true;
    // number
    const char* s = p_;
    if (p_ < e_ && (*p_ == '-' || *p_ == '+')) ++p_;
    bool frac = false;
    while (p_ < e_ && (isdigit((unsigned char)*p_) || *p_ == '.' || *p_ == 'e' || *p_ == 'E' ||
                       ((*p_ == '-' || *p_ == '+') && (p_[-1] == 'e' || p_[-1] == 'E'))))
    {
      if (*p_ == '.' || *p_ == 'e' || *p_ == 'E') frac = true;
      ++p_;
    }
    if (p_ == s) return false;
    const std::string tok(s, p_);
    char* end = nullptr;
    v.kind = Json::Number;
    v.num = strtod(tok.c_str(), &end);
    if (end != tok.c_str() + tok.size()) return false;
    v.is_int = !frac;
    if (v.is_int) v.inum = strtoll(tok.c_str(), nullptr, 10);
    return true;
  }
};

} // namespace json
} // namespace openntc
