// json.cpp -- recursive-descent JSON reader (RFC 8259 subset used by the
// condition files: no surrogate-pair decoding beyond \uXXXX -> UTF-8).
#include "json.h"

#include <cstdlib>
#include <fstream>
#include <sstream>

namespace gample {

namespace {

struct Parser {
    const std::string &s;
    size_t i = 0;
    explicit Parser(const std::string &t) : s(t) {}

    [[noreturn]] void fail(const char *what)
    {
        throw std::runtime_error(std::string("JSON parse error: ") + what + " at offset " + std::to_string(i));
    }
    void ws()
    {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) i++;
    }
    bool lit(const char *w)
    {
        size_t n = 0;
        while (w[n]) n++;
        if (s.compare(i, n, w) == 0) { i += n; return true; }
        return false;
    }
    std::string string_()
    {
        if (s[i] != '"') fail("expected string");
        i++;
        std::string out;
        while (i < s.size() && s[i] != '"') {
            char c = s[i++];
            if (c == '\\') {
                if (i >= s.size()) fail("bad escape");
                char e = s[i++];
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
                    if (i + 4 > s.size()) fail("bad \\u escape");
                    unsigned cp = (unsigned)std::strtoul(s.substr(i, 4).c_str(), nullptr, 16);
                    i += 4;
                    if (cp < 0x80) out += (char)cp;
                    else if (cp < 0x800) { out += (char)(0xC0 | (cp >> 6)); out += (char)(0x80 | (cp & 0x3F)); }
                    else { out += (char)(0xE0 | (cp >> 12)); out += (char)(0x80 | ((cp >> 6) & 0x3F));
                           out += (char)(0x80 | (cp & 0x3F)); }
                    break;
                }
                default: fail("bad escape");
                }
            } else {
                out += c;
            }
        }
        if (i >= s.size()) fail("unterminated string");
        i++;
        return out;
    }
    Json value()
    {
        ws();
        if (i >= s.size()) fail("unexpected end");
        Json j;
        char c = s[i];
        if (c == '{') {
            i++;
            j.type = Json::Object;
            ws();
            if (s[i] == '}') { i++; return j; }
            for (;;) {
                ws();
                std::string k = string_();
                ws();
                if (s[i] != ':') fail("expected ':'");
                i++;
                Json v = value();
                j.set(k, v);
                ws();
                if (s[i] == ',') { i++; continue; }
                if (s[i] == '}') { i++; break; }
                fail("expected ',' or '}'");
            }
        } else if (c == '[') {
            i++;
            j.type = Json::Array;
            ws();
            if (s[i] == ']') { i++; return j; }
            for (;;) {
                j.arr.push_back(value());
                ws();
                if (s[i] == ',') { i++; continue; }
                if (s[i] == ']') { i++; break; }
                fail("expected ',' or ']'");
            }
        } else if (c == '"') {
            j.type = Json::String;
            j.str = string_();
        } else if (lit("true")) {
            j.type = Json::Bool; j.b = true;
        } else if (lit("false")) {
            j.type = Json::Bool; j.b = false;
        } else if (lit("null")) {
            j.type = Json::Null;
        } else {
            const char *b = s.c_str() + i;
            char *e = nullptr;
            double v = std::strtod(b, &e);
            if (e == b) fail("unexpected character");
            i += (size_t)(e - b);
            j.type = Json::Number;
            j.num = v;
        }
        return j;
    }
};

}  // namespace

Json json_parse(const std::string &text)
{
    Parser p(text);
    Json j = p.value();
    p.ws();
    if (p.i != text.size()) p.fail("trailing characters");
    return j;
}

Json json_load(const std::string &path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return json_parse(ss.str());
}

}  // namespace gample
