// C++ host layer: scene ingestion, packing and the Integrator classes over libzgl.so.
// FP32 arithmetic is done one IEEE operation at a time in a fixed order (built with
// -ffp-contract=off); sin, cos and tan are the FP64 functions rounded to FP32.  The Python
// host (zillumgl_amd/loader.py, scene.py) follows the same operation order, and the two
// produce the same buffers bit for bit.
#include "zgl_host.hpp"

#include <sys/stat.h>

#include <algorithm>
#include <array>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>

namespace zgl {
namespace host {

namespace {

[[noreturn]] void bad(const std::string& msg) { throw std::invalid_argument(msg); }

void check(zgl_status st, zgl_context* ctx, const char* what)
{
    if (st != 0) throw std::runtime_error(std::string(what) + ": " + (ctx ? zgl_last_error(ctx) : "error"));
}

bool fileExists(const std::string& p)
{
    struct stat sb;
    return stat(p.c_str(), &sb) == 0 && S_ISREG(sb.st_mode);
}

std::string dirOf(const std::string& path)
{
    const size_t k = path.find_last_of('/');
    return k == std::string::npos ? std::string(".") : (k == 0 ? std::string("/") : path.substr(0, k));
}

std::string joinPath(const std::string& base, const std::string& p)
{
    if (!p.empty() && p[0] == '/') return p;
    return base.empty() ? p : base + "/" + p;
}

std::string readFile(const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

std::vector<std::string> splitWs(const std::string& s)
{
    std::vector<std::string> out;
    size_t i = 0;
    while (i < s.size()) {
        while (i < s.size() && std::isspace((unsigned char)s[i])) i++;
        size_t j = i;
        while (j < s.size() && !std::isspace((unsigned char)s[j])) j++;
        if (j > i) out.push_back(s.substr(i, j - i));
        i = j;
    }
    return out;
}

std::string joinFrom(const std::vector<std::string>& t, size_t k)
{
    std::string s;
    for (size_t i = k; i < t.size(); i++) s += (i > k ? " " : "") + t[i];
    return s;
}

// A decimal floating-point token as Python's float() reads it (no hexadecimal forms).
bool parseDouble(const std::string& tok, double& out)
{
    if (tok.empty() || tok.find_first_of("xX") != std::string::npos) return false;
    errno = 0;
    char* end = nullptr;
    const double v = std::strtod(tok.c_str(), &end);
    if (end != tok.c_str() + tok.size()) return false;
    out = v;
    return true;
}

double toDouble(const std::string& tok, const std::string& where)
{
    double v;
    if (!parseDouble(tok, v)) bad(where + ": malformed number '" + tok + "'");
    return v;
}

long toInt(const std::string& tok, const std::string& where)
{
    std::string t = tok;
    while (!t.empty() && std::isspace((unsigned char)t.back())) t.pop_back();
    size_t b = 0;
    while (b < t.size() && std::isspace((unsigned char)t[b])) b++;
    t = t.substr(b);
    char* end = nullptr;
    errno = 0;
    const long v = std::strtol(t.c_str(), &end, 10);
    if (t.empty() || end != t.c_str() + t.size() || errno) bad(where + ": malformed integer '" + tok + "'");
    return v;
}

// stringstream >> float, n times: missing or malformed values stay 0 (the rest too).
std::vector<double> floatsOf(const std::string* s, int n)
{
    std::vector<double> out(n, 0.0);
    if (!s) return out;
    const auto t = splitWs(*s);
    for (int i = 0; i < n && i < (int)t.size(); i++) {
        double v;
        if (!parseDouble(t[i], v)) break;
        out[i] = v;
    }
    return out;
}

// ----------------------------------------------------------------------------- XML
struct XmlNode {
    std::string tag;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::vector<XmlNode> children;
    const std::string* get(const char* name) const
    {
        for (const auto& a : attrs)
            if (a.first == name) return &a.second;
        return nullptr;
    }
    const XmlNode* find(const std::string& path) const
    {
        // ElementTree's find: the first match, in document order, of a child path
        std::vector<const XmlNode*> cur = {this};
        size_t b = 0;
        while (b <= path.size()) {
            size_t e = path.find('/', b);
            if (e == std::string::npos) e = path.size();
            const std::string step = path.substr(b, e - b);
            std::vector<const XmlNode*> next;
            for (const XmlNode* n : cur)
                for (const XmlNode& c : n->children)
                    if (c.tag == step) next.push_back(&c);
            cur.swap(next);
            b = e + 1;
        }
        return cur.empty() ? nullptr : cur[0];
    }
};

class XmlParser {
public:
    explicit XmlParser(const std::string& text, const std::string& name) : s(text), file(name) {}
    XmlNode parseDocument()
    {
        if (s.compare(0, 3, "\xEF\xBB\xBF") == 0) i = 3;
        skipMisc();
        if (i >= s.size() || s[i] != '<') fail("no root element");
        XmlNode root = element();
        skipMisc();
        if (i != s.size()) fail("content after the root element");
        return root;
    }

private:
    const std::string& s;
    std::string file;
    size_t i = 0;
    [[noreturn]] void fail(const std::string& m) { bad(file + ": malformed XML (" + m + ")"); }
    void ws()
    {
        while (i < s.size() && std::isspace((unsigned char)s[i])) i++;
    }
    bool startsWith(const char* p) const { return s.compare(i, std::strlen(p), p) == 0; }
    void skipUntil(const char* end)
    {
        const size_t k = s.find(end, i);
        if (k == std::string::npos) fail(std::string("missing ") + end);
        i = k + std::strlen(end);
    }
    void skipMisc()
    {
        for (;;) {
            ws();
            if (startsWith("<?")) skipUntil("?>");
            else if (startsWith("<!--")) skipUntil("-->");
            else if (startsWith("<!")) skipUntil(">");
            else return;
        }
    }
    std::string name()
    {
        const size_t b = i;
        while (i < s.size() && !std::isspace((unsigned char)s[i]) && s[i] != '>' && s[i] != '/' && s[i] != '=')
            i++;
        if (i == b) fail("expected a name");
        return s.substr(b, i - b);
    }
    std::string decode(const std::string& v)
    {
        std::string o;
        for (size_t k = 0; k < v.size(); k++) {
            if (v[k] != '&') {
                o += v[k];
                continue;
            }
            const size_t e = v.find(';', k);
            if (e == std::string::npos) fail("bad entity");
            const std::string ent = v.substr(k + 1, e - k - 1);
            if (ent == "amp") o += '&';
            else if (ent == "lt") o += '<';
            else if (ent == "gt") o += '>';
            else if (ent == "quot") o += '"';
            else if (ent == "apos") o += '\'';
            else if (!ent.empty() && ent[0] == '#') {
                const long c = (ent.size() > 1 && (ent[1] == 'x' || ent[1] == 'X')) ? std::strtol(ent.c_str() + 2, nullptr, 16)
                                                                                   : std::strtol(ent.c_str() + 1, nullptr, 10);
                if (c <= 0 || c > 0x10FFFF) fail("bad character reference");
                if (c < 0x80) o += (char)c;
                else if (c < 0x800) { o += (char)(0xC0 | (c >> 6)); o += (char)(0x80 | (c & 0x3F)); }
                else if (c < 0x10000) { o += (char)(0xE0 | (c >> 12)); o += (char)(0x80 | ((c >> 6) & 0x3F)); o += (char)(0x80 | (c & 0x3F)); }
                else { o += (char)(0xF0 | (c >> 18)); o += (char)(0x80 | ((c >> 12) & 0x3F)); o += (char)(0x80 | ((c >> 6) & 0x3F)); o += (char)(0x80 | (c & 0x3F)); }
            } else fail("unknown entity &" + ent + ";");
            k = e;
        }
        return o;
    }
    XmlNode element()
    {
        XmlNode n;
        i++;  // '<'
        n.tag = name();
        for (;;) {
            ws();
            if (i >= s.size()) fail("unterminated tag");
            if (s[i] == '/') {
                if (s.compare(i, 2, "/>") != 0) fail("bad tag end");
                i += 2;
                return n;
            }
            if (s[i] == '>') {
                i++;
                break;
            }
            const std::string an = name();
            ws();
            if (i >= s.size() || s[i] != '=') fail("expected '='");
            i++;
            ws();
            if (i >= s.size() || (s[i] != '"' && s[i] != '\'')) fail("expected a quoted value");
            const char q = s[i++];
            const size_t e = s.find(q, i);
            if (e == std::string::npos) fail("unterminated value");
            n.attrs.emplace_back(an, decode(s.substr(i, e - i)));
            i = e + 1;
        }
        for (;;) {
            const size_t lt = s.find('<', i);
            if (lt == std::string::npos) fail("unterminated element " + n.tag);
            i = lt;
            if (startsWith("</")) {
                i += 2;
                if (name() != n.tag) fail("mismatched end tag");
                ws();
                if (i >= s.size() || s[i] != '>') fail("bad end tag");
                i++;
                return n;
            }
            if (startsWith("<!--")) skipUntil("-->");
            else if (startsWith("<![CDATA[")) skipUntil("]]>");
            else if (startsWith("<?")) skipUntil("?>");
            else n.children.push_back(element());
        }
    }
};

// ------------------------------------------------------------------------ glm, FP32
using V4 = std::array<float, 4>;
struct M4 {
    V4 c[4];  // columns
};

M4 identity()
{
    M4 m;
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) m.c[c][r] = (r == c) ? 1.0f : 0.0f;
    return m;
}

const float kDegToRad = (float)0.01745329251994329576923690768489;
float radians(double deg) { return (float)deg * kDegToRad; }
float trigF(double (*fn)(double), float x) { return (float)fn((double)x); }

std::array<float, 3> normalize3(std::array<float, 3> v)
{
    float d = v[0] * v[0] + v[1] * v[1];
    d = d + v[2] * v[2];
    const float k = 1.0f / std::sqrt(d);
    return {v[0] * k, v[1] * k, v[2] * k};
}

std::array<float, 3> cross3(const std::array<float, 3>& a, const std::array<float, 3>& b)
{
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}

M4 translate(const M4& m, const double v[3])
{
    M4 o = m;
    const float x = (float)v[0], y = (float)v[1], z = (float)v[2];
    for (int r = 0; r < 4; r++) o.c[3][r] = ((m.c[0][r] * x + m.c[1][r] * y) + m.c[2][r] * z) + m.c[3][r];
    return o;
}

M4 scale(const M4& m, const double v[3])
{
    M4 o = m;
    for (int k = 0; k < 3; k++)
        for (int r = 0; r < 4; r++) o.c[k][r] = m.c[k][r] * (float)v[k];
    return o;
}

// glm::rotate(m, angle, axis) = m * R(angle, normalize(axis))
M4 rotate(const M4& m, float angle, std::array<float, 3> axis)
{
    const float c = trigF(std::cos, angle), s = trigF(std::sin, angle);
    const auto a = normalize3(axis);
    const float omc = 1.0f - c;
    const float t[3] = {omc * a[0], omc * a[1], omc * a[2]};
    const float R[3][3] = {{c + t[0] * a[0], t[0] * a[1] + s * a[2], t[0] * a[2] - s * a[1]},
                           {t[1] * a[0] - s * a[2], c + t[1] * a[1], t[1] * a[2] + s * a[0]},
                           {t[2] * a[0] + s * a[1], t[2] * a[1] - s * a[0], c + t[2] * a[2]}};
    M4 o;
    for (int k = 0; k < 3; k++)
        for (int r = 0; r < 4; r++) o.c[k][r] = (m.c[0][r] * R[k][0] + m.c[1][r] * R[k][1]) + m.c[2][r] * R[k][2];
    o.c[3] = m.c[3];
    return o;
}

M4 matmul(const M4& a, const M4& b)
{
    M4 o;
    for (int k = 0; k < 4; k++)
        for (int r = 0; r < 4; r++)
            o.c[k][r] = ((a.c[0][r] * b.c[k][0] + a.c[1][r] * b.c[k][1]) + a.c[2][r] * b.c[k][2]) + a.c[3][r] * b.c[k][3];
    return o;
}

// glm::inverse of a mat4 (cofactors, then one division by the determinant)
M4 inverse(const M4& m)
{
    auto g = [&](int c, int r) { return m.c[c][r]; };
    const float c00 = g(2, 2) * g(3, 3) - g(3, 2) * g(2, 3), c02 = g(1, 2) * g(3, 3) - g(3, 2) * g(1, 3);
    const float c03 = g(1, 2) * g(2, 3) - g(2, 2) * g(1, 3), c04 = g(2, 1) * g(3, 3) - g(3, 1) * g(2, 3);
    const float c06 = g(1, 1) * g(3, 3) - g(3, 1) * g(1, 3), c07 = g(1, 1) * g(2, 3) - g(2, 1) * g(1, 3);
    const float c08 = g(2, 1) * g(3, 2) - g(3, 1) * g(2, 2), c10 = g(1, 1) * g(3, 2) - g(3, 1) * g(1, 2);
    const float c11 = g(1, 1) * g(2, 2) - g(2, 1) * g(1, 2), c12 = g(2, 0) * g(3, 3) - g(3, 0) * g(2, 3);
    const float c14 = g(1, 0) * g(3, 3) - g(3, 0) * g(1, 3), c15 = g(1, 0) * g(2, 3) - g(2, 0) * g(1, 3);
    const float c16 = g(2, 0) * g(3, 2) - g(3, 0) * g(2, 2), c18 = g(1, 0) * g(3, 2) - g(3, 0) * g(1, 2);
    const float c19 = g(1, 0) * g(2, 2) - g(2, 0) * g(1, 2), c20 = g(2, 0) * g(3, 1) - g(3, 0) * g(2, 1);
    const float c22 = g(1, 0) * g(3, 1) - g(3, 0) * g(1, 1), c23 = g(1, 0) * g(2, 1) - g(2, 0) * g(1, 1);
    const V4 fac0 = {c00, c00, c02, c03}, fac1 = {c04, c04, c06, c07}, fac2 = {c08, c08, c10, c11};
    const V4 fac3 = {c12, c12, c14, c15}, fac4 = {c16, c16, c18, c19}, fac5 = {c20, c20, c22, c23};
    const V4 vec0 = {g(1, 0), g(0, 0), g(0, 0), g(0, 0)}, vec1 = {g(1, 1), g(0, 1), g(0, 1), g(0, 1)};
    const V4 vec2 = {g(1, 2), g(0, 2), g(0, 2), g(0, 2)}, vec3 = {g(1, 3), g(0, 3), g(0, 3), g(0, 3)};
    auto comb = [](const V4& a, const V4& fa, const V4& b, const V4& fb, const V4& c, const V4& fc) {
        V4 o;
        for (int k = 0; k < 4; k++) o[k] = (a[k] * fa[k] - b[k] * fb[k]) + c[k] * fc[k];
        return o;
    };
    const V4 inv[4] = {comb(vec1, fac0, vec2, fac1, vec3, fac2), comb(vec0, fac0, vec2, fac3, vec3, fac4),
                       comb(vec0, fac1, vec1, fac3, vec3, fac5), comb(vec0, fac2, vec1, fac4, vec2, fac5)};
    const float sa[4] = {1, -1, 1, -1}, sb[4] = {-1, 1, -1, 1};
    M4 cols;
    for (int c = 0; c < 4; c++)
        for (int k = 0; k < 4; k++) cols.c[c][k] = inv[c][k] * ((c % 2 == 0) ? sa[k] : sb[k]);
    const float d0 = m.c[0][0] * cols.c[0][0], d1 = m.c[0][1] * cols.c[1][0];
    const float d2 = m.c[0][2] * cols.c[2][0], d3 = m.c[0][3] * cols.c[3][0];
    const float det = (d0 + d1) + (d2 + d3);
    const float oneOver = 1.0f / det;
    for (int c = 0; c < 4; c++)
        for (int k = 0; k < 4; k++) cols.c[c][k] = cols.c[c][k] * oneOver;
    return cols;
}

M4 constRot()
{
    static const M4 r = rotate(identity(), radians(90.0), {1.0f, 0.0f, 0.0f});
    return r;
}

// ------------------------------------------------------------------------ OBJ / MTL
void readMtl(const std::string& path, std::map<std::string, std::vector<float>>& mats)
{
    std::istringstream in(readFile(path));
    std::string line, cur;
    bool has = false;
    while (std::getline(in, line)) {
        const auto t = splitWs(line);
        if (t.empty() || t[0][0] == '#') continue;
        if (t[0] == "newmtl") {
            cur = joinFrom(t, 1);
            has = true;
            mats[cur] = {0.0f, 0.0f, 0.0f};
        } else if (t[0] == "Kd" && has && t.size() >= 4) {
            mats[cur] = {(float)toDouble(t[1], path), (float)toDouble(t[2], path), (float)toDouble(t[3], path)};
        }
    }
}

struct Corner {
    long p, t, n;
    bool operator<(const Corner& o) const
    {
        return p != o.p ? p < o.p : (t != o.t ? t < o.t : n < o.n);
    }
};

long objIndex(const std::string& tok, size_t n, const std::string& where)
{
    const long i = toInt(tok, where);
    return i > 0 ? i - 1 : (long)n + i;
}

// Per-corner normal = normalised sum of the unit (Newell) normals of the polygons of the
// mesh touching its position, each polygon counted once; FP64, sums left to right.
void smoothNormals(const std::vector<float>& pos, const std::vector<std::vector<uint32_t>>& polys,
                   std::vector<float>& nrm, const std::vector<char>& missing)
{
    const size_t n = pos.size() / 3;
    std::map<std::array<float, 3>, int> keys;  // -0 and +0 compare equal, as Python's float keys do
    std::vector<int> slot(n);
    for (size_t i = 0; i < n; i++) {
        const std::array<float, 3> k = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
        auto it = keys.find(k);
        if (it == keys.end()) it = keys.emplace(k, (int)keys.size()).first;
        slot[i] = it->second;
    }
    std::vector<std::array<double, 3>> acc(keys.size(), {0.0, 0.0, 0.0});
    for (const auto& ids : polys) {
        double nx = 0.0, ny = 0.0, nz = 0.0;
        const size_t k = ids.size();
        for (size_t i = 0; i < k; i++) {
            const float* a = &pos[3 * ids[i]];
            const float* b = &pos[3 * ids[(i + 1) % k]];
            const double a0 = a[0], a1 = a[1], a2 = a[2], b0 = b[0], b1 = b[1], b2 = b[2];
            nx += (a1 - b1) * (a2 + b2);
            ny += (a2 - b2) * (a0 + b0);
            nz += (a0 - b0) * (a1 + b1);
        }
        const double ln = std::sqrt((nx * nx + ny * ny) + nz * nz);
        if (ln > 0) {
            std::set<int> touched;
            for (uint32_t v : ids) touched.insert(slot[v]);
            for (int j : touched) {
                acc[j][0] += nx / ln;
                acc[j][1] += ny / ln;
                acc[j][2] += nz / ln;
            }
        }
    }
    for (size_t i = 0; i < n; i++) {
        if (!missing[i]) continue;
        const auto& a = acc[slot[i]];
        const double ln = std::sqrt((a[0] * a[0] + a[1] * a[1]) + a[2] * a[2]);
        for (int c = 0; c < 3; c++) nrm[3 * i + c] = ln > 0 ? (float)(a[c] / ln) : 0.0f;
    }
}

// ------------------------------------------------------------------------ images
std::vector<float> readPfm(const std::string& path, int& w, int& h)
{
    const std::string d = readFile(path);
    size_t pos = 0;
    auto line = [&]() {
        const size_t e = d.find('\n', pos);
        if (e == std::string::npos) bad(path + ": truncated PFM header");
        std::string l = d.substr(pos, e - pos);
        pos = e + 1;
        return l;
    };
    const auto magic = splitWs(line());
    if (magic.size() != 1 || magic[0] != "PF") bad(path + ": not an RGB PFM file");
    const auto wh = splitWs(line());
    if (wh.size() != 2) bad(path + ": bad PFM size");
    w = (int)toInt(wh[0], path);
    h = (int)toInt(wh[1], path);
    const double sc = toDouble(splitWs(line()).at(0), path);
    if (w < 1 || h < 1 || d.size() - pos < (size_t)w * h * 12) bad(path + ": truncated PFM data");
    std::vector<float> out((size_t)w * h * 3);
    for (int y = 0; y < h; y++) {
        for (size_t k = 0; k < (size_t)w * 3; k++) {
            unsigned char b[4];
            std::memcpy(b, d.data() + pos + ((size_t)y * w * 3 + k) * 4, 4);
            if (sc >= 0) std::swap(b[0], b[3]), std::swap(b[1], b[2]);  // big-endian file
            float v;
            std::memcpy(&v, b, 4);
            out[((size_t)(h - 1 - y) * w * 3) + k] = v;  // rows bottom-up in the file
        }
    }
    return out;
}

std::vector<float> readHdr(const std::string& path, int& w, int& h)
{
    const std::string d = readFile(path);
    size_t pos = 0;
    for (;;) {
        const size_t e = d.find('\n', pos);
        if (e == std::string::npos) bad(path + ": truncated HDR header");
        std::string l = d.substr(pos, e - pos);
        pos = e + 1;
        size_t b0 = 0, e0 = l.size();
        while (b0 < e0 && std::isspace((unsigned char)l[b0])) b0++;
        while (e0 > b0 && std::isspace((unsigned char)l[e0 - 1])) e0--;
        const std::string stripped = l.substr(b0, e0 - b0);
        if (stripped.empty()) break;
        if (stripped.compare(0, 7, "FORMAT=") == 0 && stripped != "FORMAT=32-bit_rle_rgbe")
            bad(path + ": unsupported HDR format");
    }
    const size_t e = d.find('\n', pos);
    if (e == std::string::npos) bad(path + ": missing HDR size line");
    const auto t = splitWs(d.substr(pos, e - pos));
    if (t.size() != 4 || t[0] != "-Y" || t[2] != "+X") bad(path + ": unsupported HDR orientation");
    h = (int)toInt(t[1], path);
    w = (int)toInt(t[3], path);
    pos = e + 1;
    if (w < 1 || h < 1) bad(path + ": bad HDR size");
    std::vector<unsigned char> rgbe((size_t)w * h * 4);
    auto byte = [&](size_t k) -> unsigned char {
        if (k >= d.size()) bad(path + ": truncated HDR data");
        return (unsigned char)d[k];
    };
    for (int y = 0; y < h; y++) {
        unsigned char* row = rgbe.data() + (size_t)y * w * 4;
        if (w >= 8 && w < 32768 && byte(pos) == 2 && byte(pos + 1) == 2 && ((byte(pos + 2) << 8) | byte(pos + 3)) == w) {
            pos += 4;
            for (int c = 0; c < 4; c++) {
                int x = 0;
                while (x < w) {
                    int n = byte(pos++);
                    if (n > 128) {
                        n -= 128;
                        if (x + n > w) bad(path + ": bad HDR run");
                        const unsigned char v = byte(pos++);
                        for (int k = 0; k < n; k++) row[(x + k) * 4 + c] = v;
                    } else {
                        if (n == 0 || x + n > w) bad(path + ": bad HDR run");
                        for (int k = 0; k < n; k++) row[(x + k) * 4 + c] = byte(pos++);
                    }
                    x += n;
                }
            }
        } else {
            for (size_t k = 0; k < (size_t)w * 4; k++) row[k] = byte(pos++);
        }
    }
    std::vector<float> out((size_t)w * h * 3);
    for (size_t i = 0; i < (size_t)w * h; i++) {
        const int ex = rgbe[4 * i + 3];
        const double f = ex > 0 ? std::ldexp(1.0, ex - 136) : 0.0;
        for (int c = 0; c < 3; c++) out[3 * i + c] = (float)((double)rgbe[4 * i + c] * f);
    }
    return out;
}

std::string resolve(const std::string& p, const std::string& base)
{
    const std::string a = joinPath(base, p);
    if (fileExists(a)) return a;
    if (fileExists(p)) return p;
    throw std::runtime_error("file not found: " + p);
}

Material loadMaterial(const XmlNode* node, bool& isDefault)
{
    const std::string* tp = node ? node->get("type") : nullptr;
    const std::string typ = tp ? *tp : "";
    isDefault = (typ == "default");
    Material m;
    if (isDefault) return m;
    auto f = [&](const char* name, float& dst) {
        const XmlNode* c = node->find(name);
        if (c) dst = (float)floatsOf(c->get("value"), 1)[0];
    };
    auto v3 = [&](const char* name, float* dst) {
        const XmlNode* c = node->find(name);
        if (c) {
            const auto v = floatsOf(c->get("value"), 3);
            for (int k = 0; k < 3; k++) dst[k] = (float)v[k];
        }
    };
    if (typ == "principled") {
        v3("baseColor", m.baseColor);
        f("subsurface", m.subsurface); f("metallic", m.metallic); f("roughness", m.roughness);
        f("specular", m.specular); f("specularTint", m.specularTint); f("sheen", m.sheen);
        f("sheenTint", m.sheenTint); f("clearcoat", m.clearcoat); f("clearcoatGloss", m.clearcoatGloss);
        m.type = Principled;
    } else if (typ == "metalWorkflow") {
        v3("baseColor", m.baseColor);
        f("metallic", m.metallic);
        f("roughness", m.roughness);
        m.type = MetalWorkflow;
    } else if (typ == "dielectric") {
        v3("baseColor", m.baseColor);
        f("ior", m.ior);
        f("roughness", m.roughness);
        m.type = Dielectric;
    } else if (typ == "thinDielectric") {
        v3("baseColor", m.baseColor);
        f("ior", m.ior);
        m.type = ThinDielectric;
    } else if (typ == "lambertian") {
        v3("baseColor", m.baseColor);
        m.type = Lambertian;
    }
    return m;
}

float lum601(const float c[3]) { return (c[0] * 0.299f + c[1] * 0.587f) + c[2] * 0.114f; }

}  // namespace

// ============================================================================ public
void Material::record(float out[16]) const
{
    const float r[16] = {baseColor[0], baseColor[1], baseColor[2], roughness, subsurface, metallic, specular,
                         specularTint, sheen, sheenTint, clearcoat, clearcoatGloss, ior, 0.0f, 0.0f, 0.0f};
    std::memcpy(out, r, sizeof(r));
    std::memcpy(out + 13, &type, 4);
}

Mat4 modelMatrix(const double pos[3], const double scl[3], const double rotDeg[3])
{
    M4 m = identity();
    m = translate(m, pos);
    m = scale(m, scl);
    m = matmul(m, constRot());
    m = rotate(m, radians(rotDeg[0]), {0.0f, 0.0f, 1.0f});
    m = rotate(m, radians(rotDeg[1]), {1.0f, 0.0f, 0.0f});
    m = rotate(m, radians(rotDeg[2]), {0.0f, 1.0f, 0.0f});
    Mat4 o;
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) o.m[c * 4 + r] = m.c[c][r];
    return o;
}

ObjModel readObj(const std::string& path)
{
    const std::string base = dirOf(path);
    std::vector<double> P, T, N;
    std::vector<std::string> order;  // material groups in order of first use ("\1" = none)
    std::map<std::string, std::vector<std::vector<Corner>>> groups;
    std::map<std::string, std::vector<float>> mtl;
    std::string cur = "\1";
    std::string text = readFile(path);
    for (char& ch : text)
        if (ch == '\r') ch = '\n';
    std::istringstream in(text);
    std::string line;
    while (std::getline(in, line)) {
        const auto t = splitWs(line);
        if (t.empty() || t[0][0] == '#') continue;
        const std::string& k = t[0];
        auto need = [&](size_t n) {
            if (t.size() < n) bad(path + ": short line '" + line + "'");
        };
        if (k == "v") {
            need(4);
            for (int c = 1; c <= 3; c++) P.push_back(toDouble(t[c], path));
        } else if (k == "vt") {
            need(2);
            T.push_back(toDouble(t[1], path));
            T.push_back(t.size() > 2 ? toDouble(t[2], path) : 0.0);
        } else if (k == "vn") {
            need(4);
            for (int c = 1; c <= 3; c++) N.push_back(toDouble(t[c], path));
        } else if (k == "mtllib") {
            const std::string p = joinPath(base, joinFrom(t, 1));
            if (fileExists(p)) readMtl(p, mtl);
        } else if (k == "usemtl") {
            cur = joinFrom(t, 1);
        } else if (k == "f") {
            std::vector<Corner> corners;
            const size_t nP = P.size() / 3, nT = T.size() / 2, nN = N.size() / 3;
            for (size_t i = 1; i < t.size(); i++) {
                std::vector<std::string> parts;
                size_t b = 0;
                for (;;) {
                    const size_t e = t[i].find('/', b);
                    parts.push_back(t[i].substr(b, e == std::string::npos ? std::string::npos : e - b));
                    if (e == std::string::npos) break;
                    b = e + 1;
                }
                const long pi = objIndex(parts[0], nP, path);
                const long ti = (parts.size() > 1 && !parts[1].empty()) ? objIndex(parts[1], nT, path) : -1;
                const long ni = (parts.size() > 2 && !parts[2].empty()) ? objIndex(parts[2], nN, path) : -1;
                if (!(pi >= 0 && pi < (long)nP) || ti >= (long)nT || ni >= (long)nN || ti < -1 || ni < -1)
                    bad(path + ": face index out of range in '" + line + "'");
                corners.push_back({pi, ti, ni});
            }
            if (corners.size() < 3) continue;
            if (!groups.count(cur)) order.push_back(cur);
            groups[cur].push_back(corners);
        }
    }
    ObjModel model;
    for (const std::string& name : order) {
        auto it = mtl.find(name);
        model.materials.emplace_back(name == "\1" ? std::string() : name,
                                     (name != "\1" && it != mtl.end()) ? it->second : std::vector<float>{1.0f, 1.0f, 1.0f});
        std::map<Corner, uint32_t> keyIndex;
        std::vector<Corner> out;
        std::vector<std::vector<uint32_t>> polys;
        Mesh mesh;
        for (const auto& face : groups[name]) {
            std::vector<uint32_t> ids;
            for (const Corner& c : face) {
                auto f = keyIndex.find(c);
                if (f == keyIndex.end()) {
                    f = keyIndex.emplace(c, (uint32_t)out.size()).first;
                    out.push_back(c);
                }
                ids.push_back(f->second);
            }
            for (size_t i = 1; i + 1 < ids.size(); i++) {
                mesh.indices.push_back(ids[0]);
                mesh.indices.push_back(ids[i]);
                mesh.indices.push_back(ids[i + 1]);
            }
            polys.push_back(ids);
        }
        std::vector<char> missing(out.size(), 0);
        bool anyMissing = false;
        for (size_t i = 0; i < out.size(); i++) {
            const Corner& c = out[i];
            for (int k = 0; k < 3; k++) mesh.positions.push_back((float)P[3 * c.p + k]);
            const float u = c.t >= 0 ? (float)T[2 * c.t] : 0.0f;
            const float v = c.t >= 0 ? (float)T[2 * c.t + 1] : 0.0f;
            mesh.texcoords.push_back(u);
            mesh.texcoords.push_back(1.0f - v);  // aiProcess_FlipUVs
            for (int k = 0; k < 3; k++) mesh.normals.push_back(c.n >= 0 ? (float)N[3 * c.n + k] : 0.0f);
            if (c.n < 0) missing[i] = 1, anyMissing = true;
        }
        if (anyMissing) smoothNormals(mesh.positions, polys, mesh.normals, missing);
        model.meshes.push_back(std::move(mesh));
    }
    return model;
}

SceneDesc loadScene(const std::string& path)
{
    const std::string base = dirOf(path);
    const std::string text = readFile(path);
    const XmlNode root = XmlParser(text, path).parseDocument();
    if (root.tag != "scene") bad(path + ": root element is not <scene>");
    SceneDesc d;
    const XmlNode* size = root.find("integrator/size");
    auto intAttr = [&](const char* a) {
        const std::string* v = size ? size->get(a) : nullptr;
        return v ? (int)toInt(*v, path) : 0;
    };
    d.width = intAttr("width");
    d.height = intAttr("height");
    if (d.width < 1 || d.height < 1) bad(path + ": integrator/size must give a positive width and height");
    const XmlNode* smp = root.find("sampler");
    const std::string* st = smp ? smp->get("type") : nullptr;
    d.sampler = (st && *st == "sobol") ? 1 : 0;
    const XmlNode* cam = root.find("camera");
    auto cv = [&](const char* name, int n) {
        const XmlNode* c = cam ? cam->find(name) : nullptr;
        return floatsOf(c ? c->get("value") : nullptr, n);
    };
    const auto cp = cv("position", 3), ca = cv("angle", 3);
    for (int k = 0; k < 3; k++) d.camera.pos[k] = cp[k], d.camera.angle[k] = ca[k];
    d.camera.fov = cv("fov", 1)[0];
    d.camera.lensRadius = cv("lensRadius", 1)[0];
    d.camera.focalDist = cv("focalDistance", 1)[0];

    std::map<std::string, ObjModel> cache;
    const XmlNode* insts = root.find("modelInstances");
    if (insts) {
        for (const XmlNode& inst : insts->children) {
            const std::string* mp = inst.get("path");
            const std::string mpath = resolve(mp ? *mp : std::string(), base);
            if (!cache.count(mpath)) cache[mpath] = readObj(mpath);
            const ObjModel& model = cache[mpath];
            const XmlNode* tr = inst.find("transform");
            const auto pos = floatsOf(tr ? tr->get("translate") : nullptr, 3);
            const auto scl = floatsOf(tr ? tr->get("scale") : nullptr, 3);
            const auto rot = floatsOf(tr ? tr->get("rotate") : nullptr, 3);
            const Mat4 Mm = modelMatrix(pos.data(), scl.data(), rot.data());
            M4 M;
            for (int c = 0; c < 4; c++)
                for (int r = 0; r < 4; r++) M.c[c][r] = Mm.m[c * 4 + r];
            const M4 inv = inverse(M);
            // mat3(transpose(inverse(M))): column c = row c of the inverse's upper 3 x 3
            float n3[3][3];
            for (int c = 0; c < 3; c++)
                for (int r = 0; r < 3; r++) n3[c][r] = inv.c[r][c];
            std::vector<Mesh> world;
            for (const Mesh& mesh : model.meshes) {
                Mesh w;
                const size_t n = mesh.vertexCount();
                w.positions.resize(3 * n);
                w.normals.resize(3 * n);
                for (size_t i = 0; i < n; i++) {
                    const float x = mesh.positions[3 * i], y = mesh.positions[3 * i + 1], z = mesh.positions[3 * i + 2];
                    for (int r = 0; r < 3; r++)
                        w.positions[3 * i + r] = (M.c[0][r] * x + M.c[1][r] * y) + (M.c[2][r] * z + M.c[3][r] * 1.0f);
                    const float nx = mesh.normals[3 * i], ny = mesh.normals[3 * i + 1], nz = mesh.normals[3 * i + 2];
                    float t[3];
                    for (int r = 0; r < 3; r++) t[r] = (n3[0][r] * nx + n3[1][r] * ny) + n3[2][r] * nz;
                    const float dd = (t[0] * t[0] + t[1] * t[1]) + t[2] * t[2];
                    const float k = 1.0f / std::sqrt(dd);
                    for (int r = 0; r < 3; r++) w.normals[3 * i + r] = t[r] * k;
                }
                w.indices = mesh.indices;
                w.texcoords = mesh.texcoords;
                world.push_back(std::move(w));
            }
            const std::string* type = inst.get("type");
            if (type && *type == "light") {
                const XmlNode* rn = inst.find("radiance");
                const auto rad = floatsOf(rn ? rn->get("value") : nullptr, 3);
                // every mesh of a light model gets the whole radiance as its power, spread
                // over its triangles by area (Scene.cpp:213-238)
                for (Mesh& w : world) d.lights.emplace_back(std::move(w), std::vector<float>{(float)rad[0], (float)rad[1], (float)rad[2]});
                continue;
            }
            bool isDefault = false;
            const Material over = loadMaterial(inst.find("material"), isDefault);
            for (size_t k = 0; k < world.size(); k++) {
                Material mat = over;
                if (isDefault) {
                    mat = Material();
                    const auto& kd = model.materials[k].second;
                    for (int c = 0; c < 3; c++) mat.baseColor[c] = kd[c];
                }
                d.objects.emplace_back(std::move(world[k]), mat);
            }
        }
    }
    const XmlNode* env = root.find("envMap");
    const std::string* ep = env ? env->get("path") : nullptr;
    if (ep && !ep->empty()) {
        const std::string p = resolve(*ep, base);
        const size_t dot = p.find_last_of('.');
        std::string ext = dot == std::string::npos ? "" : p.substr(dot);
        for (char& ch : ext) ch = (char)std::tolower((unsigned char)ch);
        if (ext == ".hdr") d.env = readHdr(p, d.envWidth, d.envHeight);
        else if (ext == ".pfm") d.env = readPfm(p, d.envWidth, d.envHeight);
        else bad("unsupported environment map format (the C++ host reads .hdr and .pfm): " + p);
    }
    return d;
}

zgl_camera cameraUniforms(const CameraDesc& cam, int width, int height)
{
    using V3 = std::array<float, 3>;
    const float ax = radians(cam.angle[0]), ay = radians(cam.angle[1]);
    const float x = trigF(std::sin, ax) * trigF(std::cos, ay);
    const float y = trigF(std::cos, ax) * trigF(std::cos, ay);
    const float z = trigF(std::sin, ay);
    const V3 front = normalize3({x, y, z});
    V3 right = normalize3(cross3(front, {0.0f, 0.0f, 1.0f}));
    const float roll = (float)cam.angle[2];
    if (roll != 0.0f) {
        // glm::rotate(mat4(1), roll, front) applied to right (Camera.cpp:159-160)
        const float c = trigF(std::cos, roll), s = trigF(std::sin, roll);
        const V3 a = normalize3(front);
        const float omc = 1.0f - c;
        const float t[3] = {omc * a[0], omc * a[1], omc * a[2]};
        const float R[3][3] = {{c + t[0] * a[0], t[0] * a[1] + s * a[2], t[0] * a[2] - s * a[1]},
                               {t[1] * a[0] - s * a[2], c + t[1] * a[1], t[1] * a[2] + s * a[0]},
                               {t[2] * a[0] + s * a[1], t[2] * a[1] - s * a[0], c + t[2] * a[2]}};
        V3 r2;
        for (int k = 0; k < 3; k++) r2[k] = (R[0][k] * right[0] + R[1][k] * right[1]) + R[2][k] * right[2];
        right = normalize3(r2);
    }
    const V3 up = normalize3(cross3(right, front));
    // glm::inverse(mat3(right, up, front)) (NaivePath.cpp:53-54)
    const V3 m[3] = {right, up, front};
    auto g = [&](int c, int r) { return m[c][r]; };
    const float det = (g(0, 0) * (g(1, 1) * g(2, 2) - g(2, 1) * g(1, 2)) - g(1, 0) * (g(0, 1) * g(2, 2) - g(2, 1) * g(0, 2))) +
                      g(2, 0) * (g(0, 1) * g(1, 2) - g(1, 1) * g(0, 2));
    const float o = 1.0f / det;
    const float inv[9] = {(g(1, 1) * g(2, 2) - g(2, 1) * g(1, 2)) * o, -(g(0, 1) * g(2, 2) - g(2, 1) * g(0, 2)) * o,
                          (g(0, 1) * g(1, 2) - g(1, 1) * g(0, 2)) * o, -(g(1, 0) * g(2, 2) - g(2, 0) * g(1, 2)) * o,
                          (g(0, 0) * g(2, 2) - g(2, 0) * g(0, 2)) * o, -(g(0, 0) * g(1, 2) - g(1, 0) * g(0, 2)) * o,
                          (g(1, 0) * g(2, 1) - g(2, 0) * g(1, 1)) * o, -(g(0, 0) * g(2, 1) - g(2, 0) * g(0, 1)) * o,
                          (g(0, 0) * g(1, 1) - g(1, 0) * g(0, 1)) * o};
    const double fov = std::min(std::max(cam.fov, 0.1), 90.0);
    zgl_camera c{};
    for (int k = 0; k < 3; k++) {
        c.front[k] = front[k];
        c.right[k] = right[k];
        c.up[k] = up[k];
        c.pos[k] = (float)cam.pos[k];
    }
    std::memcpy(c.mat_inv, inv, sizeof(inv));
    c.tan_fov = trigF(std::tan, radians((double)((float)fov * 0.5f)));
    c.aspect = (float)width / (float)height;
    c.lens_radius = (float)cam.lensRadius;
    c.focal_dist = (float)cam.focalDist;
    c.film_width = width;
    c.film_height = height;
    return c;
}

std::vector<float> noiseTexture(int width, int height, uint64_t seed)
{
    // splitmix64 of seed ^ "noise", top 24 bits as a float in [0, 1)
    const uint64_t s = seed ^ 0x6E6F697365ull;
    const size_t n = (size_t)width * height * 2;
    std::vector<float> out(n);
    for (size_t i = 0; i < n; i++) {
        uint64_t z = s + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        out[i] = (float)((double)(z >> 40) * (1.0 / 16777216.0));
    }
    return out;
}

PackedScene pack(const SceneDesc& desc, bool deviceBvh)
{
    PackedScene ps;
    int64_t offV = 0;
    int32_t offM = 0, objPrims = 0;
    for (const auto& om : desc.objects) {
        const Mesh& mesh = om.first;
        float rec[16];
        om.second.record(rec);
        ps.materials.insert(ps.materials.end(), rec, rec + 16);
        const size_t n = mesh.vertexCount();
        if (mesh.normals.size() != 3 * n || mesh.indices.size() % 3) bad("mesh buffers of inconsistent sizes");
        ps.vertices.insert(ps.vertices.end(), mesh.positions.begin(), mesh.positions.end());
        ps.normals.insert(ps.normals.end(), mesh.normals.begin(), mesh.normals.end());
        if (mesh.texcoords.size() == 2 * n) ps.texcoords.insert(ps.texcoords.end(), mesh.texcoords.begin(), mesh.texcoords.end());
        else ps.texcoords.insert(ps.texcoords.end(), 2 * n, 0.0f);
        for (uint32_t i : mesh.indices) ps.indices.push_back((int32_t)(i + offV));
        const size_t m = mesh.indices.size() / 3;
        ps.matTexIndices.insert(ps.matTexIndices.end(), m, offM - 65536);  // texIndex -1 << 16 | matIndex 0
        offV += n;
        objPrims += (int32_t)m;
        offM += 1;
    }
    std::vector<float> pdf;
    float lightSum = 0.0f;
    for (const auto& lm : desc.lights) {
        const Mesh& mesh = lm.first;
        const size_t n = mesh.vertexCount();
        ps.vertices.insert(ps.vertices.end(), mesh.positions.begin(), mesh.positions.end());
        ps.normals.insert(ps.normals.end(), mesh.normals.begin(), mesh.normals.end());
        for (uint32_t i : mesh.indices) ps.indices.push_back((int32_t)(i + offV));
        const size_t m = mesh.indices.size() / 3;
        std::vector<float> areas(m);
        float sumArea = 0.0f;
        for (size_t t = 0; t < m; t++) {
            const float* va = &mesh.positions[3 * mesh.indices[3 * t]];
            const float* vb = &mesh.positions[3 * mesh.indices[3 * t + 1]];
            const float* vc = &mesh.positions[3 * mesh.indices[3 * t + 2]];
            const float e1[3] = {vc[0] - va[0], vc[1] - va[1], vc[2] - va[2]};
            const float e2[3] = {vb[0] - va[0], vb[1] - va[1], vb[2] - va[2]};
            const float cx = e1[1] * e2[2] - e1[2] * e2[1];
            const float cy = e1[2] * e2[0] - e1[0] * e2[2];
            const float cz = e1[0] * e2[1] - e1[1] * e2[0];
            areas[t] = std::sqrt((cx * cx + cy * cy) + cz * cz);
            sumArea = sumArea + areas[t];
        }
        for (size_t t = 0; t < m; t++) {
            float tp[3];
            for (int c = 0; c < 3; c++) tp[c] = (lm.second[c] * areas[t]) / sumArea;
            ps.lightPower.insert(ps.lightPower.end(), tp, tp + 3);
            const float l = lum601(tp);
            pdf.push_back(l);
            lightSum = lightSum + l;
        }
        offV += n;
    }
    if (ps.materials.empty()) {
        float rec[16];
        Material().record(rec);
        ps.materials.assign(rec, rec + 16);
    }
    const int64_t nTri = (int64_t)ps.indices.size() / 3;
    if (nTri < 1) bad("the scene has no triangles");
    ps.objPrimCount = objPrims;
    ps.nLightTriangles = (int32_t)(nTri - objPrims);
    if (!deviceBvh) {
        ps.bounds.resize((size_t)(2 * nTri - 1) * 6);
        ps.hitTable.resize((size_t)6 * (2 * nTri - 1) * 3);
        std::vector<uint32_t> idx(ps.indices.begin(), ps.indices.end());
        check(zgl_build_mtbvh(ps.vertices.data(), (int64_t)ps.vertices.size() / 3, idx.data(), nTri, ps.bounds.data(),
                              ps.hitTable.data()),
              nullptr, "MTBVH build");
    }
    if (!pdf.empty()) {
        ps.lightAlias.resize(pdf.size());
        ps.lightProb.resize(pdf.size());
        check(zgl_alias_build(pdf.data(), (int32_t)pdf.size(), ps.lightAlias.data(), ps.lightProb.data()), nullptr,
              "light alias table");
    }
    ps.lightSum = lightSum;
    std::vector<float> env = desc.env;
    int ew = desc.envWidth, eh = desc.envHeight;
    if (env.empty()) {
        env.assign(3, 0.0f);
        ew = eh = 1;
    }
    if (env.size() != (size_t)ew * eh * 3) bad("environment map of inconsistent size");
    ps.envWidth = ew;
    ps.envHeight = eh;
    ps.envAlias.resize((size_t)(ew + 1) * eh);
    ps.envProb.resize((size_t)(ew + 1) * eh);
    float envSum = 0.0f;
    check(zgl_env_alias_build(env.data(), ew, eh, ps.envAlias.data(), ps.envProb.data(), &envSum), nullptr,
          "environment alias table");
    ps.envSum = std::isfinite(envSum) ? (float)(int64_t)envSum : 0.0f;  // EnvironmentMap::sumPdf() -> int
    ps.envHalf.resize(env.size());
    zgl_float_to_half(env.data(), ps.envHalf.data(), (int64_t)env.size());
    ps.envRotation = desc.envRotation;
    ps.sampler = desc.sampler;
    ps.camera = cameraUniforms(desc.camera, desc.width, desc.height);
    ps.width = desc.width;
    ps.height = desc.height;
    ps.noise = noiseTexture(desc.width, desc.height, desc.noiseSeed);
    return ps;
}

zgl_scene_buffers PackedScene::buffers() const
{
    zgl_scene_buffers sb{};
    sb.vertices = vertices.data();
    sb.normals = normals.data();
    sb.n_vertices = (int64_t)vertices.size() / 3;
    sb.texcoords = texcoords.data();
    sb.n_texcoords = (int64_t)texcoords.size() / 2;
    sb.indices = indices.data();
    sb.n_triangles = (int64_t)indices.size() / 3;
    sb.bounds = bounds.empty() ? nullptr : bounds.data();
    sb.hit_table = hitTable.empty() ? nullptr : hitTable.data();
    sb.bvh_size = (int32_t)(2 * sb.n_triangles - 1);
    sb.mat_tex_indices = matTexIndices.data();
    sb.obj_prim_count = objPrimCount;
    sb.materials = materials.data();
    sb.n_materials = (int32_t)(materials.size() / 16);
    sb.light_power = lightPower.data();
    sb.light_alias = lightAlias.data();
    sb.light_prob = lightProb.data();
    sb.n_light_triangles = nLightTriangles;
    sb.light_sum = lightSum;
    sb.env_rgb16f = envHalf.data();
    sb.env_width = envWidth;
    sb.env_height = envHeight;
    sb.env_alias = envAlias.data();
    sb.env_prob = envProb.data();
    sb.env_sum = envSum;
    sb.env_rotation = envRotation;
    sb.sampler = sampler;
    return sb;
}

void upload(zgl_context* ctx, const PackedScene& ps)
{
    const zgl_scene_buffers sb = ps.buffers();
    check(zgl_upload_scene(ctx, &sb), ctx, "zgl_upload_scene");
    check(zgl_set_camera(ctx, &ps.camera), ctx, "zgl_set_camera");
    check(zgl_set_noise(ctx, ps.noise.data(), ps.width, ps.height), ctx, "zgl_set_noise");
}

// ============================================================================ integrators
void Integrator::init(zgl_context* ctx, int width, int height)
{
    if (!ctx || width < 1 || height < 1) bad("Integrator::init: context and film size required");
    mCtx = ctx;
    mWidth = width;
    mHeight = height;
    reset();
}

void Integrator::reset()
{
    check(zgl_clear_frame(mCtx), mCtx, "zgl_clear_frame");
    mCurSample = 0;
}

std::vector<float> Integrator::getFrame() const
{
    std::vector<float> out((size_t)mWidth * mHeight * 4);
    check(zgl_read_frame(mCtx, out.data(), (int64_t)out.size()), mCtx, "zgl_read_frame");
    return out;
}

void Integrator::render(int32_t integrator, int n, int32_t blocksOnePass, int32_t loopsPerPass)
{
    if (!mCtx) throw std::runtime_error("Integrator: init() first");
    if (n < 1) bad("Integrator::renderOnePass: n must be positive");
    mFreeCounter += 1;
    zgl_pass_params p{};
    p.integrator = integrator;
    p.first_pass = mCurSample;
    p.n_passes = n;
    p.free_counter = mFreeCounter;
    p.max_depth = mMaxDepth;
    p.russian_roulette = mRussianRoulette;
    p.sample_light = mSampleLight;
    p.light_env_uniform = mLightEnvUniform;
    p.light_portion = mLightPortion;
    p.blocks_one_pass = blocksOnePass;
    p.loops_per_pass = loopsPerPass;
    check(zgl_render(mCtx, &p), mCtx, "zgl_render");
    mFreeCounter += n - 1;
    mCurSample += n;
}

NaivePathIntegrator::NaivePathIntegrator(int maxDepth, bool russianRoulette, bool sampleLight, bool lightEnvUniform,
                                         float lightPortion)
{
    mMaxDepth = maxDepth;
    mRussianRoulette = russianRoulette;
    mSampleLight = sampleLight;
    mLightEnvUniform = lightEnvUniform;
    mLightPortion = lightPortion;
}

void NaivePathIntegrator::renderOnePass(int n) { render(ZGL_INTEGRATOR_NAIVE, n); }

GlobalQueuePathIntegrator::GlobalQueuePathIntegrator(int maxDepth, bool russianRoulette, bool sampleLight,
                                                     bool lightEnvUniform, float lightPortion)
{
    mMaxDepth = maxDepth;
    mRussianRoulette = russianRoulette;
    mSampleLight = sampleLight;
    mLightEnvUniform = lightEnvUniform;
    mLightPortion = lightPortion;
}

void GlobalQueuePathIntegrator::reset()
{
    Integrator::reset();
    mSamplePerPixel = 1.0;
}

void GlobalQueuePathIntegrator::renderOnePass(int n)
{
    render(ZGL_INTEGRATOR_GLOBAL_QUEUE, n);
    mSamplePerPixel += n;
}

LightPathIntegrator::LightPathIntegrator(int maxDepth, bool russianRoulette, int threadBlocksOnePass)
    : mBlocks(threadBlocksOnePass)
{
    mMaxDepth = maxDepth;
    mRussianRoulette = russianRoulette;
}

void LightPathIntegrator::reset()
{
    Integrator::reset();
    mSamplePerPixel = (double)mBlocks * WorkgroupSize / ((double)mWidth * mHeight);
}

void LightPathIntegrator::renderOnePass(int n)
{
    render(ZGL_INTEGRATOR_LIGHT, n, mBlocks, 1);
    mSamplePerPixel += (double)n * mBlocks * WorkgroupSize / ((double)mWidth * mHeight);
}

TriplePathIntegrator::TriplePathIntegrator(int maxDepth, bool russianRoulette, int lptBlocksOnePass, int lptLoopsPerPass)
    : mBlocks(lptBlocksOnePass), mLoops(lptLoopsPerPass)
{
    mMaxDepth = maxDepth;
    mRussianRoulette = russianRoulette;
}

void TriplePathIntegrator::reset()
{
    Integrator::reset();
    mSamplePerPixel = 1.0;
}

void TriplePathIntegrator::renderOnePass(int n)
{
    render(ZGL_INTEGRATOR_TRIPLE, n, mBlocks, mLoops);
    mSamplePerPixel += n;
}

}  // namespace host
}  // namespace zgl
