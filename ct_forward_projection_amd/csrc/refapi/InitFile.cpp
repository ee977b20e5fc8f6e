// InitFile.cpp -- clean-room INI reader; see InitFile.h for the grammar.
#include "InitFile.h"

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <strings.h>

namespace {

std::string trim(const std::string &s)
{
    size_t b = 0, e = s.size();
    while (b < e && (s[b] == ' ' || s[b] == '\t' || s[b] == '\r' || s[b] == '\n')) b++;
    while (e > b && (s[e - 1] == ' ' || s[e - 1] == '\t' || s[e - 1] == '\r' || s[e - 1] == '\n')) e--;
    return s.substr(b, e - b);
}

bool iequal(const std::string &a, const char *b) { return strcasecmp(a.c_str(), b) == 0; }

bool parse_ll(const std::string &v, long long &out)
{
    if (v.empty()) return false;
    char *end = nullptr;
    errno = 0;
    long long x = std::strtoll(v.c_str(), &end, 10);
    if (errno || end == v.c_str()) return false;
    // accept "60000", "6.0e4" style integers too
    if (*end != '\0') {
        double d = std::strtod(v.c_str(), &end);
        if (*end != '\0' || !std::isfinite(d) || d != std::floor(d)) return false;
        if (d < -9.2e18 || d > 9.2e18) return false;
        x = (long long)d;
    }
    out = x;
    return true;
}

bool parse_d(const std::string &v, double &out)
{
    if (v.empty()) return false;
    char *end = nullptr;
    errno = 0;
    double d = std::strtod(v.c_str(), &end);
    if (errno || end == v.c_str() || *end != '\0') return false;
    out = d;
    return true;
}

} // namespace

CInitFile::CInitFile() {}
CInitFile::~CInitFile() {}

void CInitFile::parse(const std::string &text)
{
    m_sections.clear();
    std::istringstream in(text);
    std::string line;
    Section *cur = nullptr;
    bool first = true;
    while (std::getline(in, line)) {
        if (first) { // UTF-8 byte-order mark
            if (line.size() >= 3 && (unsigned char)line[0] == 0xEF && (unsigned char)line[1] == 0xBB &&
                (unsigned char)line[2] == 0xBF)
                line = line.substr(3);
            first = false;
        }
        std::string t = trim(line);
        if (t.empty() || t[0] == ';' || t[0] == '#') continue;
        if (t[0] == '[') {
            size_t close = t.find(']');
            std::string name = trim(t.substr(1, close == std::string::npos ? std::string::npos : close - 1));
            m_sections.push_back(Section{name, {}});
            cur = &m_sections.back();
            continue;
        }
        size_t eq = t.find('=');
        if (eq == std::string::npos || !cur) continue;
        Entry e;
        e.key = trim(t.substr(0, eq));
        std::string rest = t.substr(eq + 1);
        size_t semi = rest.find(';');
        if (semi != std::string::npos) {
            e.comment = trim(rest.substr(semi + 1));
            rest = rest.substr(0, semi);
        }
        e.value = trim(rest);
        cur->entries.push_back(e);
    }
}

void CInitFile::GetFileName(const char *szFileName)
{
    m_szModName = szFileName ? szFileName : "";
    m_loaded = false;
    m_sections.clear();
    std::ifstream f(m_szModName, std::ios::binary);
    if (!f.is_open()) return;
    std::stringstream ss;
    ss << f.rdbuf();
    parse(ss.str());
    m_loaded = true;
}

const CInitFile::Entry *CInitFile::find(const char *section, const char *key) const
{
    if (!section || !key) return nullptr;
    for (int pass = 0; pass < 2; pass++) {
        for (const Section &s : m_sections) {
            bool sm = pass == 0 ? s.name == section : iequal(s.name, section);
            if (!sm) continue;
            for (const Entry &e : s.entries)
                if (pass == 0 ? e.key == key : iequal(e.key, key)) return &e;
        }
    }
    return nullptr;
}

CInitFile::Entry *CInitFile::findOrAdd(const char *section, const char *key)
{
    Section *sec = nullptr;
    for (Section &s : m_sections)
        if (s.name == section) sec = &s;
    if (!sec) {
        m_sections.push_back(Section{section, {}});
        sec = &m_sections.back();
    }
    for (Entry &e : sec->entries)
        if (e.key == key) return &e;
    sec->entries.push_back(Entry{key, "", ""});
    return &sec->entries.back();
}

BOOL CInitFile::GetEntryValue(const char *s, const char *k, const char *def, std::string &out)
{
    const Entry *e = find(s, k);
    if (!e) {
        out = def ? def : "";
        return 0;
    }
    out = e->value;
    return 1;
}

#define CTFP_INI_INTEGER(T)                                                              \
    BOOL CInitFile::GetEntryValue(const char *s, const char *k, T##_default_t def, T &out) \
    {                                                                                    \
        const Entry *e = find(s, k);                                                     \
        long long v;                                                                     \
        if (!e || !parse_ll(e->value, v) || v < (long long)std::numeric_limits<T>::min() ||  \
            (v > 0 && (unsigned long long)v > (unsigned long long)std::numeric_limits<T>::max())) { \
            out = (T)def;                                                                \
            return 0;                                                                    \
        }                                                                                \
        out = (T)v;                                                                      \
        return 1;                                                                        \
    }

typedef int char_default_t;
typedef int short_default_t;
typedef int int_default_t;
typedef long long_default_t;
typedef long long llong_default_t;
typedef long long llong;
CTFP_INI_INTEGER(char)
CTFP_INI_INTEGER(short)
CTFP_INI_INTEGER(int)
CTFP_INI_INTEGER(long)

BOOL CInitFile::GetEntryValue(const char *s, const char *k, long long def, long long &out)
{
    const Entry *e = find(s, k);
    long long v;
    if (!e || !parse_ll(e->value, v)) {
        out = def;
        return 0;
    }
    out = v;
    return 1;
}

BOOL CInitFile::GetEntryValue(const char *s, const char *k, size_t def, size_t &out)
{
    // A negative or malformed value stores the default (the reference's
    // sentinel -100 becomes a huge size_t, which callers must reject).
    const Entry *e = find(s, k);
    long long v;
    if (!e || !parse_ll(e->value, v) || v < 0) {
        out = def;
        return 0;
    }
    out = (size_t)v;
    return 1;
}

BOOL CInitFile::GetEntryValue(const char *s, const char *k, double def, double &out)
{
    const Entry *e = find(s, k);
    double v;
    if (!e || !parse_d(e->value, v)) {
        out = def;
        return 0;
    }
    out = v;
    return 1;
}

BOOL CInitFile::GetEntryValue(const char *s, const char *k, float def, float &out)
{
    double v;
    BOOL ok = GetEntryValue(s, k, (double)def, v);
    out = ok ? (float)v : def;
    return ok;
}

BOOL CInitFile::SetEntryValue(const char *s, const char *k, std::string v)
{
    if (!s || !k) return 0;
    findOrAdd(s, k)->value = v;
    return 1;
}
BOOL CInitFile::SetEntryValue(const char *s, const char *k, char v) { return SetEntryValue(s, k, std::to_string((int)v)); }
BOOL CInitFile::SetEntryValue(const char *s, const char *k, short v) { return SetEntryValue(s, k, std::to_string(v)); }
BOOL CInitFile::SetEntryValue(const char *s, const char *k, int v) { return SetEntryValue(s, k, std::to_string(v)); }
BOOL CInitFile::SetEntryValue(const char *s, const char *k, unsigned long v) { return SetEntryValue(s, k, std::to_string(v)); }
BOOL CInitFile::SetEntryValue(const char *s, const char *k, long long v) { return SetEntryValue(s, k, std::to_string(v)); }
BOOL CInitFile::SetEntryValue(const char *s, const char *k, double v)
{
    char buf[64];
    snprintf(buf, sizeof buf, "%.17g", v);
    return SetEntryValue(s, k, std::string(buf));
}
BOOL CInitFile::SetEntryValue(const char *s, const char *k, float v)
{
    char buf[64];
    snprintf(buf, sizeof buf, "%.9g", (double)v);
    return SetEntryValue(s, k, std::string(buf));
}

void CInitFile::SaveFile(const char *szFileName)
{
    std::string path = szFileName ? szFileName : m_szModName;
    std::ofstream f(path);
    for (const Section &s : m_sections) {
        f << "[" << s.name << "]\n";
        for (const Entry &e : s.entries) {
            f << e.key << "=" << e.value;
            if (!e.comment.empty()) f << "\t;" << e.comment;
            f << "\n";
        }
        f << "\n";
    }
}

void CInitFile::creatFile(const char *szFileName)
{
    m_szModName = szFileName ? szFileName : "";
    m_sections.clear();
    m_loaded = true;
    SaveFile();
}
