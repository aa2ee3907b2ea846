/*
 * Minimal JSON reader for Fumeng scene files (RFC 8259 subset: objects,
 * arrays, strings with escapes, numbers, true/false/null).  Object keys keep
 * the last duplicate, like the reference's nlohmann::json parse.
 */
#ifndef FM_HOST_JSON_H
#define FM_HOST_JSON_H

#include <cstdint>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace fmh {

struct Json {
    enum Type { NUL, BOOL, INT, REAL, STRING, ARRAY, OBJECT } type = NUL;
    bool b = false;
    int64_t i = 0;
    double d = 0.0;
    std::string s;
    std::vector<Json> arr;
    std::map<std::string, Json> obj;

    bool is_number() const { return type == INT || type == REAL; }
    double num() const { return type == INT ? (double)i : d; }
    const Json* find(const std::string& k) const {
        if (type != OBJECT) return nullptr;  /* nlohmann find() on non-objects: end() */
        auto it = obj.find(k);
        return it == obj.end() ? nullptr : &it->second;
    }
};

class JsonParser {
public:
    explicit JsonParser(const std::string& text) : t_(text) {}
    Json parse() {
        Json v = value();
        ws();
        if (p_ != t_.size()) fail("trailing characters");
        return v;
    }

private:
    const std::string& t_;
    size_t p_ = 0;

    [[noreturn]] void fail(const char* what) {
        throw std::runtime_error(std::string("JSON parse error at offset ") + std::to_string(p_) + ": " + what);
    }
    void ws() {
        while (p_ < t_.size() && (t_[p_] == ' ' || t_[p_] == '\t' || t_[p_] == '\n' || t_[p_] == '\r')) p_++;
    }
    bool lit(const char* w) {
        size_t n = 0;
        while (w[n]) n++;
        if (t_.compare(p_, n, w) == 0) { p_ += n; return true; }
        return false;
    }
    Json value() {
        ws();
        if (p_ >= t_.size()) fail("unexpected end");
        char c = t_[p_];
        Json v;
        if (c == '{') {
            v.type = Json::OBJECT;
            p_++;
            ws();
            if (p_ < t_.size() && t_[p_] == '}') { p_++; return v; }
            for (;;) {
                ws();
                if (p_ >= t_.size() || t_[p_] != '"') fail("expected key");
                std::string k = str();
                ws();
                if (p_ >= t_.size() || t_[p_] != ':') fail("expected ':'");
                p_++;
                v.obj[k] = value();
                ws();
                if (p_ < t_.size() && t_[p_] == ',') { p_++; continue; }
                if (p_ < t_.size() && t_[p_] == '}') { p_++; return v; }
                fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.type = Json::ARRAY;
            p_++;
            ws();
            if (p_ < t_.size() && t_[p_] == ']') { p_++; return v; }
            for (;;) {
                v.arr.push_back(value());
                ws();
                if (p_ < t_.size() && t_[p_] == ',') { p_++; continue; }
                if (p_ < t_.size() && t_[p_] == ']') { p_++; return v; }
                fail("expected ',' or ']'");
            }
        }
        if (c == '"') { v.type = Json::STRING; v.s = str(); return v; }
        if (lit("true")) { v.type = Json::BOOL; v.b = true; return v; }
        if (lit("false")) { v.type = Json::BOOL; v.b = false; return v; }
        if (lit("null")) { v.type = Json::NUL; return v; }
        return number();
    }
    static void put_utf8(std::string& o, uint32_t cp) {
        if (cp < 0x80) o += (char)cp;
        else if (cp < 0x800) { o += (char)(0xC0 | (cp >> 6)); o += (char)(0x80 | (cp & 0x3F)); }
        else if (cp < 0x10000) { o += (char)(0xE0 | (cp >> 12)); o += (char)(0x80 | ((cp >> 6) & 0x3F)); o += (char)(0x80 | (cp & 0x3F)); }
        else { o += (char)(0xF0 | (cp >> 18)); o += (char)(0x80 | ((cp >> 12) & 0x3F)); o += (char)(0x80 | ((cp >> 6) & 0x3F)); o += (char)(0x80 | (cp & 0x3F)); }
    }
    uint32_t hex4() {
        if (p_ + 4 > t_.size()) fail("bad \\u escape");
        uint32_t v = 0;
        for (int k = 0; k < 4; k++) {
            char h = t_[p_++];
            v <<= 4;
            if (h >= '0' && h <= '9') v |= (uint32_t)(h - '0');
            else if (h >= 'a' && h <= 'f') v |= (uint32_t)(h - 'a' + 10);
            else if (h >= 'A' && h <= 'F') v |= (uint32_t)(h - 'A' + 10);
            else fail("bad hex digit");
        }
        return v;
    }
    std::string str() {
        std::string o;
        p_++;  /* opening quote */
        for (;;) {
            if (p_ >= t_.size()) fail("unterminated string");
            char c = t_[p_++];
            if (c == '"') return o;
            if ((unsigned char)c < 0x20) fail("control character in string");
            if (c != '\\') { o += c; continue; }
            if (p_ >= t_.size()) fail("bad escape");
            char e = t_[p_++];
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
                    uint32_t cp = hex4();
                    if (cp >= 0xD800 && cp < 0xDC00 && p_ + 1 < t_.size() && t_[p_] == '\\' && t_[p_ + 1] == 'u') {
                        p_ += 2;
                        uint32_t lo = hex4();
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    put_utf8(o, cp);
                    break;
                }
                default: fail("bad escape");
            }
        }
    }
    Json number() {
        size_t s = p_;
        if (p_ < t_.size() && t_[p_] == '-') p_++;
        bool is_real = false;
        while (p_ < t_.size()) {
            char c = t_[p_];
            if (c >= '0' && c <= '9') { p_++; continue; }
            if (c == '.' || c == 'e' || c == 'E' || c == '+' || c == '-') { is_real = true; p_++; continue; }
            break;
        }
        if (p_ == s || (p_ == s + 1 && t_[s] == '-')) fail("unexpected character");
        std::string tok = t_.substr(s, p_ - s);
        Json v;
        char* end = nullptr;
        if (!is_real) {
            errno = 0;
            long long x = std::strtoll(tok.c_str(), &end, 10);
            if (errno == 0 && end && *end == 0) { v.type = Json::INT; v.i = x; return v; }
        }
        v.type = Json::REAL;
        v.d = std::strtod(tok.c_str(), &end);
        if (!end || *end != 0) fail("bad number");
        return v;
    }
};

}  // namespace fmh
#endif
