// io.cpp -- the files on either side of the time-step path (host code only).
//
//  * Mesh readers for unstructured grids of lines, triangles, quadrangles,
//    tetrahedra or hexahedra with their cell data: VTK XML (.vtu; ascii, inline base64 with or without zlib
//    compression, raw or base64 appended data) and legacy VTK (.vtk; ASCII or
//    BINARY, the classic CELLS layout and the OFFSETS/CONNECTIVITY layout of
//    file version 5).  They take the place of TNL::Meshes::loadMesh
//    (control.h:20-48) and of getMeshReader(...)->readCellData for the
//    "InitialCondition[Z<i>]" arrays (Solver.hpp:296-306).  TNL's readers are
//    not vendored with the reference; the formats are the published VTK file
//    formats.
//  * The VTU writer of snapshots (HeatEquationModel.h:164-231, which uses TNL's
//    VTUWriter with FileFormat::zlib_compressed): FieldData CYCLE and TIME,
//    Points, Cells and cell data.
//  * TNL's binary object format for BoundaryConditionsStorage
//    (BoundaryConditionsStorage.h:16-72) and for the dumped vectors: the magic
//    "TNLMN", the serialization type as an int32 length and its characters,
//    and for an array its element count as a 64-bit integer followed by the
//    raw elements.
//  * The linear-system dump of a failed solve (Solver::saveLinearSystem,
//    Solver.hpp:966-977): the matrix in Matrix Market coordinate form, the
//    last iterate and the right-hand side as TNL arrays.
//
// Nothing here launches GPU work; mhfem_save_linear_system reads the system
// back through the public export calls.  Errors leave their message for
// mhfem_last_error(NULL).
#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/mhfem.h"

namespace mhfem {
void set_global_error(const std::string &msg);  // capi.cpp: message of mhfem_last_error(NULL)
}

struct mhfem_mesh_file {
    int dim = 0, nvc = 0;         // dimension, vertices per cell
    std::vector<double> points;   // dim per vertex
    std::vector<int32_t> cells;   // nvc per cell
    std::map<std::string, std::vector<double>> cell_data;
    std::vector<std::string> cell_data_order;
};

namespace {

struct IoError {
    int code;
    std::string msg;
};

[[noreturn]] void fail(const std::string &msg, int code = MHFEM_E_FILE) { throw IoError{code, msg}; }

template <typename F>
int io_guarded(F &&f)
{
    try {
        f();
        return MHFEM_OK;
    } catch (const IoError &e) {
        mhfem::set_global_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc &) {
        mhfem::set_global_error("out of host memory");
        return MHFEM_E_FILE;
    } catch (const std::exception &e) {
        mhfem::set_global_error(e.what());
        return MHFEM_E_FILE;
    }
}

std::string read_file(const std::string &path)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) fail("cannot open file '" + path + "': " + std::strerror(errno));
    std::string s;
    char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, got);
    const bool bad = std::ferror(f);
    std::fclose(f);
    if (bad) fail("error reading file '" + path + "'");
    return s;
}

struct OutFile {
    FILE *f = nullptr;
    std::string path;
    explicit OutFile(const std::string &p) : path(p)
    {
        f = std::fopen(p.c_str(), "wb");
        if (!f) fail("cannot open file '" + p + "' for writing: " + std::strerror(errno));
    }
    ~OutFile()
    {
        if (f) std::fclose(f);
    }
    void write(const void *data, size_t bytes)
    {
        if (bytes && std::fwrite(data, 1, bytes, f) != bytes) fail("error writing file '" + path + "'");
    }
    void put(const std::string &s) { write(s.data(), s.size()); }
    void close()
    {
        const int rc = std::fclose(f);
        f = nullptr;
        if (rc) fail("error writing file '" + path + "'");
    }
};

// ---------------------------------------------------------------------------
// Element types of both VTK formats

enum class VT { I8, U8, I16, U16, I32, U32, I64, U64, F32, F64 };

size_t vt_size(VT t)
{
    switch (t) {
    case VT::I8: case VT::U8: return 1;
    case VT::I16: case VT::U16: return 2;
    case VT::I32: case VT::U32: case VT::F32: return 4;
    default: return 8;
    }
}

VT xml_type(const std::string &s)
{
    static const std::map<std::string, VT> m = {
        {"Int8", VT::I8},   {"UInt8", VT::U8},   {"Int16", VT::I16}, {"UInt16", VT::U16},
        {"Int32", VT::I32}, {"UInt32", VT::U32}, {"Int64", VT::I64}, {"UInt64", VT::U64},
        {"Float32", VT::F32}, {"Float64", VT::F64}};
    auto it = m.find(s);
    if (it == m.end()) fail("unsupported VTK data type '" + s + "'");
    return it->second;
}

VT legacy_type(std::string s)
{
    for (auto &c : s) c = (char)std::tolower((unsigned char)c);
    static const std::map<std::string, VT> m = {
        {"char", VT::I8},          {"unsigned_char", VT::U8},  {"short", VT::I16},
        {"unsigned_short", VT::U16}, {"int", VT::I32},         {"unsigned_int", VT::U32},
        {"long", VT::I64},         {"unsigned_long", VT::U64}, {"vtktypeint64", VT::I64},
        {"vtktypeuint64", VT::U64}, {"vtkidtype", VT::I64},    {"vtktypeint32", VT::I32},
        {"vtktypeuint32", VT::U32}, {"float", VT::F32},        {"double", VT::F64}};
    auto it = m.find(s);
    if (it == m.end()) fail("unsupported legacy VTK data type '" + s + "'");
    return it->second;
}

// little-endian (native) element k of a raw buffer, as double or as int64
double elem_f(const uint8_t *p, VT t)
{
    switch (t) {
    case VT::I8: return (double)*(const int8_t *)p;
    case VT::U8: return (double)*p;
    case VT::I16: { int16_t v; std::memcpy(&v, p, 2); return v; }
    case VT::U16: { uint16_t v; std::memcpy(&v, p, 2); return v; }
    case VT::I32: { int32_t v; std::memcpy(&v, p, 4); return v; }
    case VT::U32: { uint32_t v; std::memcpy(&v, p, 4); return v; }
    case VT::I64: { int64_t v; std::memcpy(&v, p, 8); return (double)v; }
    case VT::U64: { uint64_t v; std::memcpy(&v, p, 8); return (double)v; }
    case VT::F32: { float v; std::memcpy(&v, p, 4); return v; }
    default: { double v; std::memcpy(&v, p, 8); return v; }
    }
}

int64_t elem_i(const uint8_t *p, VT t)
{
    switch (t) {
    case VT::F32: case VT::F64: {
        const double v = elem_f(p, t);
        if (v != std::floor(v) || std::fabs(v) > 9.0e15) fail("non-integer value in an index array");
        return (int64_t)v;
    }
    case VT::U64: { uint64_t v; std::memcpy(&v, p, 8); if (v > (uint64_t)INT64_MAX) fail("index out of range"); return (int64_t)v; }
    case VT::I64: { int64_t v; std::memcpy(&v, p, 8); return v; }
    default: return (int64_t)elem_f(p, t);
    }
}

void swap_bytes(std::vector<uint8_t> &b, size_t width)
{
    if (width <= 1) return;
    for (size_t k = 0; k + width <= b.size(); k += width) std::reverse(b.begin() + k, b.begin() + k + width);
}

// one decoded data array: native-endian raw elements
struct Raw {
    VT type = VT::F64;
    long ncomp = 1;
    std::vector<uint8_t> bytes;
    size_t count() const { return bytes.size() / vt_size(type); }
    std::vector<double> as_double() const
    {
        const size_t w = vt_size(type), n = count();
        std::vector<double> out(n);
        for (size_t k = 0; k < n; k++) out[k] = elem_f(bytes.data() + k * w, type);
        return out;
    }
    std::vector<int64_t> as_int() const
    {
        const size_t w = vt_size(type), n = count();
        std::vector<int64_t> out(n);
        for (size_t k = 0; k < n; k++) out[k] = elem_i(bytes.data() + k * w, type);
        return out;
    }
};

// text numbers -> raw elements of type t
void append_text_value(std::vector<uint8_t> &out, const char *tok, VT t)
{
    char *end = nullptr;
    uint8_t buf[8];
    errno = 0;
    if (t == VT::F32 || t == VT::F64) {
        const double v = std::strtod(tok, &end);
        if (end == tok) fail(std::string("invalid number '") + tok + "'");
        if (t == VT::F32) { const float f = (float)v; std::memcpy(buf, &f, 4); }
        else std::memcpy(buf, &v, 8);
    } else {
        const bool is_unsigned = t == VT::U8 || t == VT::U16 || t == VT::U32 || t == VT::U64;
        int64_t v = 0;
        uint64_t u = 0;
        if (is_unsigned) u = std::strtoull(tok, &end, 10);
        else v = std::strtoll(tok, &end, 10);
        if (end == tok || errno == ERANGE) fail(std::string("invalid integer '") + tok + "'");
        switch (t) {
        case VT::I8: { int8_t x = (int8_t)v; std::memcpy(buf, &x, 1); break; }
        case VT::U8: { uint8_t x = (uint8_t)u; std::memcpy(buf, &x, 1); break; }
        case VT::I16: { int16_t x = (int16_t)v; std::memcpy(buf, &x, 2); break; }
        case VT::U16: { uint16_t x = (uint16_t)u; std::memcpy(buf, &x, 2); break; }
        case VT::I32: { int32_t x = (int32_t)v; std::memcpy(buf, &x, 4); break; }
        case VT::U32: { uint32_t x = (uint32_t)u; std::memcpy(buf, &x, 4); break; }
        case VT::I64: std::memcpy(buf, &v, 8); break;
        default: std::memcpy(buf, &u, 8); break;
        }
    }
    out.insert(out.end(), buf, buf + vt_size(t));
}

// ---------------------------------------------------------------------------
// Base64 (RFC 4648).  The decoder works in groups of four characters and
// accepts padding at the end of any group, so that a header and a data block
// encoded one after the other (as VTK writes them) decode to the
// concatenation of both.

int b64_value(unsigned char c)
{
    if (c >= 'A' && c <= 'Z') return c - 'A';
    if (c >= 'a' && c <= 'z') return c - 'a' + 26;
    if (c >= '0' && c <= '9') return c - '0' + 52;
    if (c == '+') return 62;
    if (c == '/') return 63;
    return -1;
}

std::vector<uint8_t> b64_decode(const char *s, size_t len)
{
    std::vector<uint8_t> out;
    out.reserve(len / 4 * 3 + 3);
    uint32_t q[4];
    int k = 0, pad = 0;
    for (size_t i = 0; i < len; i++) {
        const unsigned char c = (unsigned char)s[i];
        if (std::isspace(c)) continue;
        if (c == '=') {
            if (k < 2) fail("invalid base64 data");
            q[k++] = 0;
            pad++;
        } else {
            const int v = b64_value(c);
            if (v < 0 || pad) fail("invalid base64 data");
            q[k++] = (uint32_t)v;
        }
        if (k == 4) {
            const uint32_t w = q[0] << 18 | q[1] << 12 | q[2] << 6 | q[3];
            out.push_back((uint8_t)(w >> 16));
            if (pad < 2) out.push_back((uint8_t)(w >> 8));
            if (pad < 1) out.push_back((uint8_t)w);
            k = 0;
            pad = 0;
        }
    }
    if (k) fail("truncated base64 data");
    return out;
}

std::string b64_encode(const uint8_t *p, size_t n)
{
    static const char tab[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
    std::string out;
    out.reserve((n + 2) / 3 * 4);
    size_t i = 0;
    for (; i + 3 <= n; i += 3) {
        const uint32_t w = (uint32_t)p[i] << 16 | (uint32_t)p[i + 1] << 8 | p[i + 2];
        out += tab[w >> 18];
        out += tab[(w >> 12) & 63];
        out += tab[(w >> 6) & 63];
        out += tab[w & 63];
    }
    if (i < n) {
        const uint32_t w = (uint32_t)p[i] << 16 | (i + 1 < n ? (uint32_t)p[i + 1] << 8 : 0u);
        out += tab[w >> 18];
        out += tab[(w >> 12) & 63];
        out += i + 1 < n ? tab[(w >> 6) & 63] : '=';
        out += '=';
    }
    return out;
}

// ---------------------------------------------------------------------------
// A small XML reader: elements, attributes and the extent of each element's
// text.  Enough for VTK XML files; the appended-data section is located but
// not parsed (raw binary follows its '_' marker).

struct XNode {
    std::string name;
    std::map<std::string, std::string> attr;
    size_t text_begin = 0, text_end = 0;
    int parent = -1;
    std::vector<int> children;
    std::string get(const std::string &k, const std::string &def = "") const
    {
        auto it = attr.find(k);
        return it == attr.end() ? def : it->second;
    }
};

std::string xml_unescape(const std::string &s)
{
    std::string out;
    for (size_t i = 0; i < s.size(); i++) {
        if (s[i] != '&') { out += s[i]; continue; }
        const size_t semi = s.find(';', i);
        if (semi == std::string::npos) fail("invalid XML entity");
        const std::string e = s.substr(i + 1, semi - i - 1);
        if (e == "amp") out += '&';
        else if (e == "lt") out += '<';
        else if (e == "gt") out += '>';
        else if (e == "quot") out += '"';
        else if (e == "apos") out += '\'';
        else if (!e.empty() && e[0] == '#') {
            const long v = e.size() > 1 && (e[1] == 'x' || e[1] == 'X') ? std::strtol(e.c_str() + 2, nullptr, 16)
                                                                      : std::strtol(e.c_str() + 1, nullptr, 10);
            if (v <= 0 || v > 127) fail("unsupported XML character reference");
            out += (char)v;
        } else fail("unknown XML entity &" + e + ";");
        i = semi;
    }
    return out;
}

std::string xml_escape(const std::string &s)
{
    std::string out;
    for (char c : s) {
        if (c == '&') out += "&amp;";
        else if (c == '<') out += "&lt;";
        else if (c == '>') out += "&gt;";
        else if (c == '"') out += "&quot;";
        else out += c;
    }
    return out;
}

struct XDoc {
    const std::string &src;
    std::vector<XNode> nodes;
    size_t appended_begin = std::string::npos, appended_end = std::string::npos;
    int appended = -1;

    explicit XDoc(const std::string &s) : src(s) { parse(); }

    void parse()
    {
        std::vector<int> stack;
        size_t i = 0;
        const size_t n = src.size();
        while (i < n) {
            const size_t lt = src.find('<', i);
            if (lt == std::string::npos) break;
            if (src.compare(lt, 4, "<!--") == 0) {
                const size_t e = src.find("-->", lt + 4);
                if (e == std::string::npos) fail("unterminated XML comment");
                i = e + 3;
                continue;
            }
            if (src.compare(lt, 2, "<?") == 0 || src.compare(lt, 2, "<!") == 0) {
                const size_t e = src.find('>', lt);
                if (e == std::string::npos) fail("malformed XML");
                i = e + 1;
                continue;
            }
            if (src.compare(lt, 2, "</") == 0) {
                const size_t e = src.find('>', lt);
                if (e == std::string::npos) fail("malformed XML");
                std::string name = src.substr(lt + 2, e - lt - 2);
                while (!name.empty() && std::isspace((unsigned char)name.back())) name.pop_back();
                if (stack.empty() || nodes[stack.back()].name != name) fail("mismatched XML end tag </" + name + ">");
                nodes[stack.back()].text_end = lt;
                stack.pop_back();
                i = e + 1;
                continue;
            }
            // start tag
            size_t p = lt + 1;
            XNode nd;
            while (p < n && !std::isspace((unsigned char)src[p]) && src[p] != '>' && src[p] != '/') nd.name += src[p++];
            bool self_close = false;
            for (;;) {
                while (p < n && std::isspace((unsigned char)src[p])) p++;
                if (p >= n) fail("malformed XML");
                if (src[p] == '>') { p++; break; }
                if (src[p] == '/') {
                    if (p + 1 >= n || src[p + 1] != '>') fail("malformed XML");
                    self_close = true;
                    p += 2;
                    break;
                }
                std::string key;
                while (p < n && src[p] != '=' && !std::isspace((unsigned char)src[p])) key += src[p++];
                while (p < n && std::isspace((unsigned char)src[p])) p++;
                if (p >= n || src[p] != '=') fail("malformed XML attribute '" + key + "'");
                p++;
                while (p < n && std::isspace((unsigned char)src[p])) p++;
                if (p >= n || (src[p] != '"' && src[p] != '\'')) fail("malformed XML attribute '" + key + "'");
                const char q = src[p++];
                const size_t e = src.find(q, p);
                if (e == std::string::npos) fail("malformed XML attribute '" + key + "'");
                nd.attr[key] = xml_unescape(src.substr(p, e - p));
                p = e + 1;
            }
            nd.parent = stack.empty() ? -1 : stack.back();
            nd.text_begin = nd.text_end = p;
            const int id = (int)nodes.size();
            if (nd.parent >= 0) nodes[nd.parent].children.push_back(id);
            nodes.push_back(std::move(nd));
            if (nodes[id].name == "AppendedData") {
                if (self_close) fail("empty AppendedData section");
                const size_t us = src.find('_', p);
                if (us == std::string::npos) fail("AppendedData without the '_' marker");
                appended = id;
                appended_begin = us + 1;
                const size_t close = src.rfind("</AppendedData>");
                appended_end = close == std::string::npos || close < appended_begin ? n : close;
                return;  // binary data follows: nothing after it is parsed
            }
            if (!self_close) stack.push_back(id);
            i = p;
        }
        if (!stack.empty()) fail("unterminated XML element <" + nodes[stack.back()].name + ">");
    }

    int child(int parent, const std::string &name) const
    {
        for (int c : nodes[parent].children)
            if (nodes[c].name == name) return c;
        return -1;
    }
    std::vector<int> children(int parent, const std::string &name) const
    {
        std::vector<int> out;
        for (int c : nodes[parent].children)
            if (nodes[c].name == name) out.push_back(c);
        return out;
    }
};

// ---------------------------------------------------------------------------
// VTK XML (.vtu)

struct VtuContext {
    explicit VtuContext(const XDoc &d) : doc(d) {}
    const XDoc &doc;
    bool big_endian = false;
    size_t header_size = 4;  // UInt32 unless header_type says UInt64
    bool zlib = false;
    bool appended_raw = true;
    std::vector<size_t> appended_offsets;  // sorted, for the extent of base64 blocks

    uint64_t header_value(const uint8_t *p) const
    {
        uint8_t b[8] = {};
        std::memcpy(b, p, header_size);
        if (big_endian) std::reverse(b, b + header_size);
        if (header_size == 4) {
            uint32_t v;
            std::memcpy(&v, b, 4);
            return v;
        }
        uint64_t v;
        std::memcpy(&v, b, 8);
        return v;
    }

    // the data of one array from [header][payload] bytes; `avail` bounds the input
    std::vector<uint8_t> unpack(const uint8_t *p, size_t avail) const
    {
        if (!zlib) {
            if (avail < header_size) fail("truncated VTK data block");
            const uint64_t nbytes = header_value(p);
            if (nbytes > avail - header_size) fail("truncated VTK data block");
            return std::vector<uint8_t>(p + header_size, p + header_size + nbytes);
        }
        if (avail < 3 * header_size) fail("truncated compressed VTK data block");
        const uint64_t nb = header_value(p), bs = header_value(p + header_size),
                       last = header_value(p + 2 * header_size);
        if (nb > (avail / header_size)) fail("corrupt compressed VTK header");
        const size_t hdr = (3 + nb) * header_size;
        if (hdr > avail) fail("truncated compressed VTK header");
        std::vector<uint8_t> out;
        size_t at = hdr;
        for (uint64_t b = 0; b < nb; b++) {
            const uint64_t cs = header_value(p + (3 + b) * header_size);
            const uint64_t us = (b + 1 == nb && last) ? last : bs;
            if (cs > avail - at) fail("truncated compressed VTK data block");
            const size_t o = out.size();
            out.resize(o + us);
            uLongf dest = (uLongf)us;
            const int rc = uncompress(out.data() + o, &dest, p + at, (uLong)cs);
            if (rc != Z_OK || dest != us) fail("zlib decompression failed in a VTK data block");
            at += cs;
        }
        return out;
    }

    Raw read(int id) const
    {
        const XNode &nd = doc.nodes[id];
        Raw r;
        r.type = xml_type(nd.get("type"));
        r.ncomp = std::atol(nd.get("NumberOfComponents", "1").c_str());
        if (r.ncomp < 1) fail("invalid NumberOfComponents");
        const std::string format = nd.get("format", "ascii");
        const size_t w = vt_size(r.type);
        if (format == "ascii") {
            const std::string text = doc.src.substr(nd.text_begin, nd.text_end - nd.text_begin);
            const char *s = text.c_str();
            while (*s) {
                while (*s && std::isspace((unsigned char)*s)) s++;
                if (!*s) break;
                const char *e = s;
                while (*e && !std::isspace((unsigned char)*e)) e++;
                append_text_value(r.bytes, std::string(s, e).c_str(), r.type);
                s = e;
            }
            return r;  // parsed into native order
        }
        if (format == "binary") {
            const std::vector<uint8_t> b = b64_decode(doc.src.data() + nd.text_begin, nd.text_end - nd.text_begin);
            r.bytes = unpack(b.data(), b.size());
        } else if (format == "appended") {
            if (doc.appended < 0) fail("appended data array without an AppendedData section");
            const size_t off = (size_t)std::atoll(nd.get("offset", "0").c_str());
            const size_t begin = doc.appended_begin + off;
            if (begin > doc.appended_end) fail("appended data offset out of range");
            if (appended_raw) {
                r.bytes = unpack((const uint8_t *)doc.src.data() + begin, doc.appended_end - begin);
            } else {
                auto it = std::upper_bound(appended_offsets.begin(), appended_offsets.end(), off);
                const size_t end = it == appended_offsets.end() ? doc.appended_end : doc.appended_begin + *it;
                const std::vector<uint8_t> b = b64_decode(doc.src.data() + begin, end - begin);
                r.bytes = unpack(b.data(), b.size());
            }
        } else {
            fail("unsupported DataArray format '" + format + "'");
        }
        if (r.bytes.size() % w) fail("VTK data block is not a whole number of elements");
        if (big_endian) swap_bytes(r.bytes, w);
        return r;
    }
};

// VTK cell types of the supported shapes: dimension and vertices per cell
struct CellShape {
    int dim, nvc;
};
CellShape cell_shape(int64_t vtk_type)
{
    switch (vtk_type) {
    case 3: return {1, 2};    // VTK_LINE
    case 5: return {2, 3};    // VTK_TRIANGLE
    case 9: return {2, 4};    // VTK_QUAD
    case 10: return {3, 4};   // VTK_TETRA
    case 12: return {3, 8};   // VTK_HEXAHEDRON
    default: break;
    }
    fail("unsupported cell type " + std::to_string(vtk_type) +
         " (lines, triangles, quadrangles, tetrahedra and hexahedra only)");
}

// common checks once points, connectivity and types are known
void finish_mesh(mhfem_mesh_file &m, const std::vector<double> &pts3, long np, const std::vector<int64_t> &conn,
                 const std::vector<int64_t> &offsets_end, const std::vector<int64_t> &types)
{
    const long nc = (long)types.size();
    if (nc == 0) fail("the mesh has no cells");
    if ((long)pts3.size() != 3 * np) fail("the Points array does not hold 3 coordinates per point");
    const CellShape sh = cell_shape(types[0]);
    const int dim = sh.dim, nvc = sh.nvc;
    for (long k = 0; k < nc; k++)
        if (types[k] != types[0]) fail("mixed cell types are not supported");
    if ((long)offsets_end.size() != nc) fail("offsets and types differ in length");
    int64_t prev = 0;
    for (long k = 0; k < nc; k++) {
        if (offsets_end[k] - prev != nvc) fail("cell " + std::to_string(k) + " does not have " + std::to_string(nvc) + " vertices");
        prev = offsets_end[k];
    }
    if ((long)conn.size() != nc * nvc) fail("connectivity length does not match the cells");
    if (np > INT32_MAX || nc > INT32_MAX / 8) fail("mesh too large for 32-bit indices");
    m.dim = dim;
    m.nvc = nvc;
    m.cells.resize(conn.size());
    for (size_t q = 0; q < conn.size(); q++) {
        if (conn[q] < 0 || conn[q] >= np) fail("connectivity index out of range");
        m.cells[q] = (int32_t)conn[q];
    }
    m.points.resize((size_t)np * dim);
    for (long v = 0; v < np; v++) {
        for (int c = 0; c < dim; c++) m.points[(size_t)v * dim + c] = pts3[(size_t)v * 3 + c];
        for (int c = dim; c < 3; c++)
            if (pts3[(size_t)v * 3 + c] != 0.0)
                fail(dim == 1 ? "a mesh of lines must lie on the x axis" : "a 2-D mesh must lie in the z = 0 plane");
    }
}

std::unique_ptr<mhfem_mesh_file> read_vtu(const std::string &path)
{
    const std::string src = read_file(path);
    XDoc doc(src);
    if (doc.nodes.empty() || doc.nodes[0].name != "VTKFile") fail("'" + path + "' is not a VTK XML file");
    const XNode &root = doc.nodes[0];
    if (root.get("type") != "UnstructuredGrid")
        fail("unsupported VTK XML file type '" + root.get("type") + "' (UnstructuredGrid only)");
    VtuContext cx(doc);
    const std::string bo = root.get("byte_order", "LittleEndian");
    if (bo == "BigEndian") cx.big_endian = true;
    else if (bo != "LittleEndian") fail("unknown byte_order '" + bo + "'");
    const std::string ht = root.get("header_type", "UInt32");
    if (ht == "UInt64") cx.header_size = 8;
    else if (ht != "UInt32") fail("unsupported header_type '" + ht + "'");
    const std::string comp = root.get("compressor");
    if (comp == "vtkZLibDataCompressor") cx.zlib = true;
    else if (!comp.empty()) fail("unsupported compressor '" + comp + "'");
    if (doc.appended >= 0) {
        const std::string enc = doc.nodes[doc.appended].get("encoding", "raw");
        if (enc == "base64") cx.appended_raw = false;
        else if (enc != "raw") fail("unsupported AppendedData encoding '" + enc + "'");
        for (const XNode &nd : doc.nodes)
            if (nd.name == "DataArray" && nd.get("format") == "appended")
                cx.appended_offsets.push_back((size_t)std::atoll(nd.get("offset", "0").c_str()));
        std::sort(cx.appended_offsets.begin(), cx.appended_offsets.end());
    }
    const int grid = doc.child(0, "UnstructuredGrid");
    if (grid < 0) fail("missing <UnstructuredGrid>");
    const std::vector<int> pieces = doc.children(grid, "Piece");
    if (pieces.size() != 1) fail("expected exactly one <Piece>, found " + std::to_string(pieces.size()));
    const int piece = pieces[0];
    const long np = std::atol(doc.nodes[piece].get("NumberOfPoints", "-1").c_str());
    const long nc = std::atol(doc.nodes[piece].get("NumberOfCells", "-1").c_str());
    if (np <= 0 || nc <= 0) fail("invalid NumberOfPoints / NumberOfCells");

    const int pts = doc.child(piece, "Points");
    if (pts < 0) fail("missing <Points>");
    const std::vector<int> pdas = doc.children(pts, "DataArray");
    if (pdas.empty()) fail("missing points DataArray");
    const Raw P = cx.read(pdas[0]);
    if (P.ncomp != 3) fail("points must have 3 components");
    const int cells = doc.child(piece, "Cells");
    if (cells < 0) fail("missing <Cells>");
    Raw conn, offs, types;
    bool have[3] = {};
    for (int c : doc.children(cells, "DataArray")) {
        const std::string name = doc.nodes[c].get("Name");
        if (name == "connectivity") { conn = cx.read(c); have[0] = true; }
        else if (name == "offsets") { offs = cx.read(c); have[1] = true; }
        else if (name == "types") { types = cx.read(c); have[2] = true; }
    }
    if (!have[0] || !have[1] || !have[2]) fail("<Cells> needs connectivity, offsets and types");
    if ((long)types.count() != nc) fail("types length differs from NumberOfCells");
    auto m = std::make_unique<mhfem_mesh_file>();
    finish_mesh(*m, P.as_double(), np, conn.as_int(), offs.as_int(), types.as_int());

    const int cd = doc.child(piece, "CellData");
    if (cd >= 0)
        for (int c : doc.children(cd, "DataArray")) {
            const std::string name = doc.nodes[c].get("Name");
            const Raw r = cx.read(c);
            if ((long)r.count() != nc * r.ncomp) fail("cell data '" + name + "' does not have NumberOfCells tuples");
            if (!m->cell_data.count(name)) m->cell_data_order.push_back(name);
            m->cell_data[name] = r.as_double();
        }
    return m;
}

// ---------------------------------------------------------------------------
// Legacy VTK (.vtk)

struct Legacy {
    const std::string &s;
    size_t p = 0;
    bool binary = false;
    explicit Legacy(const std::string &src) : s(src) {}

    std::string line()
    {
        const size_t e = s.find('\n', p);
        std::string out = s.substr(p, (e == std::string::npos ? s.size() : e) - p);
        p = e == std::string::npos ? s.size() : e + 1;
        if (!out.empty() && out.back() == '\r') out.pop_back();
        return out;
    }
    bool eof()
    {
        while (p < s.size() && std::isspace((unsigned char)s[p])) p++;
        return p >= s.size();
    }
    std::string token()
    {
        if (eof()) fail("unexpected end of legacy VTK file");
        const size_t b = p;
        while (p < s.size() && !std::isspace((unsigned char)s[p])) p++;
        return s.substr(b, p - b);
    }
    long number()
    {
        const std::string t = token();
        char *e = nullptr;
        const long v = std::strtol(t.c_str(), &e, 10);
        if (*e) fail("expected an integer in legacy VTK file, found '" + t + "'");
        return v;
    }
    bool peek_number()
    {
        if (eof()) return false;
        return std::isdigit((unsigned char)s[p]) != 0;
    }
    // `count` values of type t: text tokens, or big-endian binary after the end of the keyword line
    Raw values(VT t, size_t count)
    {
        Raw r;
        r.type = t;
        if (!binary) {
            for (size_t k = 0; k < count; k++) append_text_value(r.bytes, token().c_str(), t);
            return r;
        }
        const size_t nl = s.find('\n', p);
        if (nl == std::string::npos) fail("unexpected end of legacy VTK file");
        p = nl + 1;
        const size_t bytes = count * vt_size(t);
        if (bytes > s.size() - p) fail("truncated binary data in legacy VTK file");
        r.bytes.assign(s.begin() + p, s.begin() + p + bytes);
        p += bytes;
        swap_bytes(r.bytes, vt_size(t));  // legacy binary files are big-endian
        return r;
    }
    void skip_metadata()
    {
        // METADATA ... ends at the first empty line
        line();
        for (;;) {
            if (p >= s.size()) return;
            std::string l = line();
            bool blank = true;
            for (char c : l) if (!std::isspace((unsigned char)c)) blank = false;
            if (blank) return;
        }
    }
};

std::unique_ptr<mhfem_mesh_file> read_legacy_vtk(const std::string &path)
{
    const std::string src = read_file(path);
    Legacy L(src);
    const std::string head = L.line();
    if (head.rfind("# vtk DataFile Version", 0) != 0) fail("'" + path + "' is not a legacy VTK file");
    const double version = std::atof(head.c_str() + 22);
    L.line();  // title
    std::string fmt = L.line();
    for (auto &c : fmt) c = (char)std::toupper((unsigned char)c);
    while (!fmt.empty() && std::isspace((unsigned char)fmt.back())) fmt.pop_back();
    if (fmt == "BINARY") L.binary = true;
    else if (fmt != "ASCII") fail("legacy VTK format must be ASCII or BINARY, found '" + fmt + "'");
    if (L.token() != "DATASET") fail("legacy VTK: expected DATASET");
    const std::string ds = L.token();
    if (ds != "UNSTRUCTURED_GRID") fail("unsupported legacy VTK dataset '" + ds + "' (UNSTRUCTURED_GRID only)");

    std::vector<double> pts3;
    long np = -1, nc = -1;
    std::vector<int64_t> conn, offs_end, types;
    auto m = std::make_unique<mhfem_mesh_file>();
    long data_cells = -1;    // > 0 inside CELL_DATA, the tuple count
    long data_points = -1;   // inside POINT_DATA (read and dropped)
    auto store = [&](const std::string &name, const Raw &r, long ncomp) {
        if (data_cells >= 0) {
            if ((long)r.count() != data_cells * ncomp) fail("cell data '" + name + "' has the wrong length");
            if (!m->cell_data.count(name)) m->cell_data_order.push_back(name);
            m->cell_data[name] = r.as_double();
        }
    };
    auto tuples = [&]() -> long {
        if (data_cells >= 0) return data_cells;
        if (data_points >= 0) return data_points;
        fail("legacy VTK: attribute data outside CELL_DATA / POINT_DATA");
    };
    while (!L.eof()) {
        std::string kw = L.token();
        for (auto &c : kw) c = (char)std::toupper((unsigned char)c);
        if (kw == "POINTS") {
            np = L.number();
            const VT t = legacy_type(L.token());
            pts3 = L.values(t, (size_t)np * 3).as_double();
        } else if (kw == "METADATA") {
            L.skip_metadata();
        } else if (kw == "CELLS") {
            const long a = L.number(), b = L.number();
            if (version >= 5.0) {
                // CELLS <offsets count> <connectivity size>, then OFFSETS and CONNECTIVITY
                if (L.token() != "OFFSETS") fail("legacy VTK 5: expected OFFSETS");
                const std::vector<int64_t> o = L.values(legacy_type(L.token()), (size_t)a).as_int();
                if (L.token() != "CONNECTIVITY") fail("legacy VTK 5: expected CONNECTIVITY");
                conn = L.values(legacy_type(L.token()), (size_t)b).as_int();
                if (o.empty() || o[0] != 0) fail("legacy VTK 5: offsets must start at 0");
                offs_end.assign(o.begin() + 1, o.end());
                nc = a - 1;
            } else {
                nc = a;
                const std::vector<int64_t> raw = L.values(VT::I32, (size_t)b).as_int();
                size_t at = 0;
                conn.clear();
                offs_end.clear();
                for (long k = 0; k < nc; k++) {
                    if (at >= raw.size()) fail("legacy VTK: CELLS list too short");
                    const int64_t cnt = raw[at++];
                    if (cnt < 0 || at + cnt > raw.size()) fail("legacy VTK: CELLS list too short");
                    conn.insert(conn.end(), raw.begin() + at, raw.begin() + at + cnt);
                    offs_end.push_back((int64_t)conn.size());
                    at += cnt;
                }
                if (at != raw.size()) fail("legacy VTK: CELLS size does not match its list");
            }
        } else if (kw == "CELL_TYPES") {
            const long k = L.number();
            types = L.values(VT::I32, (size_t)k).as_int();
        } else if (kw == "CELL_DATA") {
            data_cells = L.number();
            data_points = -1;
            if (nc >= 0 && data_cells != nc) fail("CELL_DATA count differs from the number of cells");
        } else if (kw == "POINT_DATA") {
            data_points = L.number();
            data_cells = -1;
        } else if (kw == "SCALARS") {
            const std::string name = L.token();
            const VT t = legacy_type(L.token());
            long ncomp = 1;
            if (L.peek_number()) ncomp = L.number();
            std::string lt = L.token();
            if (lt != "LOOKUP_TABLE") fail("legacy VTK: expected LOOKUP_TABLE after SCALARS " + name);
            L.token();  // table name
            store(name, L.values(t, (size_t)(tuples() * ncomp)), ncomp);
        } else if (kw == "VECTORS" || kw == "NORMALS") {
            const std::string name = L.token();
            const VT t = legacy_type(L.token());
            store(name, L.values(t, (size_t)(tuples() * 3)), 3);
        } else if (kw == "TENSORS") {
            const std::string name = L.token();
            const VT t = legacy_type(L.token());
            store(name, L.values(t, (size_t)(tuples() * 9)), 9);
        } else if (kw == "FIELD") {
            L.token();  // field name
            const long arrays = L.number();
            for (long a = 0; a < arrays; a++) {
                const std::string name = L.token();
                if (name == "NULL_ARRAY") continue;
                const long ncomp = L.number(), nt = L.number();
                const VT t = legacy_type(L.token());
                const Raw r = L.values(t, (size_t)(ncomp * nt));
                if (data_cells >= 0 && nt == data_cells) store(name, r, ncomp);
            }
        } else if (kw == "LOOKUP_TABLE") {
            L.token();
            const long size = L.number();
            if (L.binary) L.values(VT::U8, (size_t)size * 4);
            else L.values(VT::F32, (size_t)size * 4);
        } else {
            fail("unsupported legacy VTK section '" + kw + "'");
        }
    }
    if (np < 0) fail("legacy VTK: no POINTS section");
    if (nc < 0) fail("legacy VTK: no CELLS section");
    if ((long)types.size() != nc) fail("legacy VTK: CELL_TYPES count differs from CELLS");
    finish_mesh(*m, pts3, np, conn, offs_end, types);
    return m;
}

std::string extension(const std::string &path)
{
    const size_t slash = path.find_last_of('/');
    const size_t dot = path.find_last_of('.');
    if (dot == std::string::npos || (slash != std::string::npos && dot < slash)) return "";
    std::string e = path.substr(dot + 1);
    for (auto &c : e) c = (char)std::tolower((unsigned char)c);
    return e;
}

// ---------------------------------------------------------------------------
// VTU writer

struct VtuWriter {
    OutFile &out;
    int format;  // MHFEM_VTK_*

    template <typename T>
    void array(const char *vtk_type, const std::string &name, const T *data, size_t count, int ncomp,
               const char *extra = "")
    {
        static const char *fmt_name[] = {"ascii", "binary", "binary"};
        out.put(std::string("<DataArray type=\"") + vtk_type + "\" Name=\"" + xml_escape(name) + "\"" +
                (ncomp > 1 ? " NumberOfComponents=\"" + std::to_string(ncomp) + "\"" : std::string()) + extra +
                " format=\"" + fmt_name[format] + "\">\n");
        if (format == MHFEM_VTK_ASCII) {
            std::string line;
            char buf[40];
            for (size_t k = 0; k < count; k++) {
                if constexpr (std::is_floating_point<T>::value) std::snprintf(buf, sizeof buf, "%.17g", (double)data[k]);
                else std::snprintf(buf, sizeof buf, "%lld", (long long)data[k]);
                line += buf;
                line += (k + 1) % 6 == 0 || k + 1 == count ? '\n' : ' ';
                if (line.size() > (1 << 16)) { out.put(line); line.clear(); }
            }
            out.put(line);
        } else if (format == MHFEM_VTK_BINARY) {
            // header (UInt64 byte count) and data, each encoded separately
            const uint64_t nbytes = count * sizeof(T);
            out.put(b64_encode((const uint8_t *)&nbytes, 8));
            out.put(b64_encode((const uint8_t *)data, nbytes));
            out.put("\n");
        } else {
            // one zlib block: header {1, size, size, compressed size}
            const uint64_t nbytes = count * sizeof(T);
            uLongf csize = compressBound((uLong)nbytes);
            std::vector<uint8_t> z(csize);
            if (nbytes && compress2(z.data(), &csize, (const Bytef *)data, (uLong)nbytes, Z_DEFAULT_COMPRESSION) != Z_OK)
                fail("zlib compression failed");
            if (!nbytes) csize = 0;
            const uint64_t hdr[4] = {nbytes ? 1u : 0u, nbytes, nbytes, (uint64_t)csize};
            out.put(b64_encode((const uint8_t *)hdr, nbytes ? 32 : 24));
            out.put(b64_encode(z.data(), csize));
            out.put("\n");
        }
        out.put("</DataArray>\n");
    }
};

// TNL binary objects
void tnl_write_type(OutFile &f, const std::string &type)
{
    f.write("TNLMN", 5);
    const int32_t len = (int32_t)type.size();
    f.write(&len, 4);
    f.write(type.data(), type.size());
}

struct TnlReader {
    std::string s;
    std::string path;
    size_t p = 0;
    void need(size_t n)
    {
        if (n > s.size() - p) fail("unexpected end of file '" + path + "'");
    }
    std::string object_type()
    {
        need(5);
        if (s.compare(p, 5, "TNLMN") != 0)
            fail("wrong magic number - file '" + path + "' is not in a TNL-compatible format");
        p += 5;
        need(4);
        int32_t len;
        std::memcpy(&len, s.data() + p, 4);
        p += 4;
        if (len < 0) fail("corrupt object type in '" + path + "'");
        need((size_t)len);
        std::string t = s.substr(p, len);
        p += len;
        return t;
    }
    template <typename T>
    T value()
    {
        need(sizeof(T));
        T v;
        std::memcpy(&v, s.data() + p, sizeof(T));
        p += sizeof(T);
        return v;
    }
    // an Array< elem >: the serialization type may carry more template arguments
    // (device, index) after the element type, as older TNL versions wrote it
    template <typename T>
    std::vector<T> array(const std::string &elem)
    {
        const std::string t = object_type();
        const std::string want = "TNL::Containers::Array< " + elem;
        if (t.compare(0, want.size(), want) != 0 ||
            !(t.compare(want.size(), 2, " >") == 0 || t.compare(want.size(), 2, ", ") == 0))
            fail("object type does not match (expected TNL::Containers::Array< " + elem + " >, found " + t + ").");
        const uint64_t n = value<uint64_t>();
        if (n > (s.size() - p) / sizeof(T)) fail("unexpected end of file '" + path + "'");
        std::vector<T> out(n);
        if (n) std::memcpy(out.data(), s.data() + p, n * sizeof(T));
        p += n * sizeof(T);
        return out;
    }
};

const char *BC_TYPE = "TNL::MHFEM::BoundaryConditionsStorage< double >";
const char *BC_TAG_TYPE = "TNL::MHFEM::BoundaryConditionsType";

void write_tnl_array(OutFile &f, const char *elem, const void *data, uint64_t n, size_t width)
{
    tnl_write_type(f, std::string("TNL::Containers::Array< ") + elem + " >");
    f.write(&n, 8);
    f.write(data, n * width);
}

}  // namespace

extern "C" {

mhfem_mesh_file *mhfem_mesh_file_open(const char *path, const char *format)
{
    std::unique_ptr<mhfem_mesh_file> m;
    const int rc = io_guarded([&] {
        if (!path) fail("null path", MHFEM_E_ARGUMENT);
        std::string f = format && *format ? format : "auto";
        if (f == "auto") f = extension(path);
        if (f == "vtu") m = read_vtu(path);
        else if (f == "vtk") m = read_legacy_vtk(path);
        else if (f == "pvtu") fail("distributed meshes (.pvtu) are not supported: one GPU reads one .vtu or .vtk mesh");
        else fail("unsupported mesh format '" + f + "' (vtu or vtk)");
    });
    return rc == MHFEM_OK ? m.release() : nullptr;
}

void mhfem_mesh_file_close(mhfem_mesh_file *f) { delete f; }

int mhfem_mesh_file_info(const mhfem_mesh_file *f, int *dim, int *vertices_per_cell, long *points, long *cells)
{
    if (!f) return MHFEM_E_ARGUMENT;
    if (dim) *dim = f->dim;
    if (vertices_per_cell) *vertices_per_cell = f->nvc;
    if (points) *points = (long)(f->points.size() / f->dim);
    if (cells) *cells = (long)(f->cells.size() / f->nvc);
    return MHFEM_OK;
}

int mhfem_mesh_file_read(const mhfem_mesh_file *f, double *points, int32_t *cell_vertices)
{
    if (!f) return MHFEM_E_ARGUMENT;
    if (points) std::memcpy(points, f->points.data(), f->points.size() * sizeof(double));
    if (cell_vertices) std::memcpy(cell_vertices, f->cells.data(), f->cells.size() * sizeof(int32_t));
    return MHFEM_OK;
}

long mhfem_mesh_file_cell_data(const mhfem_mesh_file *f, const char *name, double *out, long capacity)
{
    if (!f || !name) return MHFEM_E_ARGUMENT;
    auto it = f->cell_data.find(name);
    if (it == f->cell_data.end()) {
        mhfem::set_global_error(std::string("cell data array '") + name + "' was not found in the mesh file");
        return MHFEM_E_FILE;
    }
    const long n = (long)it->second.size();
    if (out) {
        if (capacity < n) {
            mhfem::set_global_error("output buffer too small for cell data '" + std::string(name) + "'");
            return MHFEM_E_ARGUMENT;
        }
        std::memcpy(out, it->second.data(), n * sizeof(double));
    }
    return n;
}

int mhfem_mesh_file_cell_data_count(const mhfem_mesh_file *f)
{
    return f ? (int)f->cell_data_order.size() : MHFEM_E_ARGUMENT;
}

const char *mhfem_mesh_file_cell_data_name(const mhfem_mesh_file *f, int index)
{
    if (!f || index < 0 || index >= (int)f->cell_data_order.size()) return nullptr;
    return f->cell_data_order[index].c_str();
}

int mhfem_write_vtu(const char *path, const mhfem_mesh_desc *mesh, int vertices_per_cell, long cycle, double time,
                    int fields, const char *const *names, const double *const *values, int format)
{
    return io_guarded([&] {
        if (!path || !mesh || (fields > 0 && (!names || !values))) fail("null argument", MHFEM_E_ARGUMENT);
        const int dim = mesh->dim, nv = vertices_per_cell;
        int vtk_type = 0;
        if (dim == 1 && nv == 2) vtk_type = 3;
        else if (dim == 2 && nv == 3) vtk_type = 5;
        else if (dim == 2 && nv == 4) vtk_type = 9;
        else if (dim == 3 && nv == 4) vtk_type = 10;
        else if (dim == 3 && nv == 8) vtk_type = 12;
        else fail("unsupported cells: " + std::to_string(nv) + " vertices in " + std::to_string(dim) + "-D",
                  MHFEM_E_ARGUMENT);
        if (format < MHFEM_VTK_ASCII || format > MHFEM_VTK_ZLIB) fail("unknown VTU format", MHFEM_E_ARGUMENT);
        const long np = mesh->num_points, nc = mesh->num_cells;
        if (np <= 0 || nc <= 0 || !mesh->points || !mesh->cell_vertices) fail("empty mesh", MHFEM_E_ARGUMENT);
        OutFile out(path);
        VtuWriter w{out, format};
        out.put("<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" "
                "byte_order=\"LittleEndian\" header_type=\"UInt64\"");
        out.put(format == MHFEM_VTK_ZLIB ? " compressor=\"vtkZLibDataCompressor\">\n" : ">\n");
        out.put("<UnstructuredGrid>\n");
        if (cycle >= 0 || time >= 0) {
            out.put("<FieldData>\n");
            if (cycle >= 0) {
                const int32_t c = (int32_t)cycle;
                w.array("Int32", "CYCLE", &c, 1, 1, " NumberOfTuples=\"1\"");
            }
            if (time >= 0) w.array("Float64", "TIME", &time, 1, 1, " NumberOfTuples=\"1\"");
            out.put("</FieldData>\n");
        }
        out.put("<Piece NumberOfPoints=\"" + std::to_string(np) + "\" NumberOfCells=\"" + std::to_string(nc) + "\">\n");
        std::vector<double> p3((size_t)np * 3, 0.0);
        for (long v = 0; v < np; v++)
            for (int c = 0; c < mesh->dim; c++) p3[(size_t)v * 3 + c] = mesh->points[(size_t)v * mesh->dim + c];
        out.put("<Points>\n");
        w.array("Float64", "Points", p3.data(), p3.size(), 3);
        out.put("</Points>\n<Cells>\n");
        for (long q = 0; q < nc * nv; q++)
            if (mesh->cell_vertices[q] < 0 || mesh->cell_vertices[q] >= np) fail("cell vertex out of range", MHFEM_E_ARGUMENT);
        w.array("Int32", "connectivity", mesh->cell_vertices, (size_t)nc * nv, 1);
        std::vector<int64_t> offs(nc);
        for (long k = 0; k < nc; k++) offs[k] = (int64_t)(k + 1) * nv;
        w.array("Int64", "offsets", offs.data(), offs.size(), 1);
        std::vector<uint8_t> types(nc, (uint8_t)vtk_type);
        w.array("UInt8", "types", types.data(), types.size(), 1);
        out.put("</Cells>\n");
        if (fields > 0) {
            out.put("<CellData>\n");
            for (int k = 0; k < fields; k++) {
                if (!names[k] || !values[k]) fail("null cell data", MHFEM_E_ARGUMENT);
                w.array("Float64", names[k], values[k], (size_t)nc, 1);
            }
            out.put("</CellData>\n");
        }
        out.put("</Piece>\n</UnstructuredGrid>\n</VTKFile>\n");
        out.close();
    });
}

int mhfem_bc_save(const char *path, int64_t dof_size, const uint8_t *tags, const double *values,
                  const double *dirichlet)
{
    return io_guarded([&] {
        if (!path || dof_size < 0 || (dof_size > 0 && (!tags || !values || !dirichlet)))
            fail("null argument", MHFEM_E_ARGUMENT);
        for (int64_t k = 0; k < dof_size; k++)
            if (tags[k] > 3) fail("invalid boundary condition type " + std::to_string(tags[k]), MHFEM_E_ARGUMENT);
        OutFile f(path);
        tnl_write_type(f, BC_TYPE);
        f.write(&dof_size, 8);
        write_tnl_array(f, BC_TAG_TYPE, tags, (uint64_t)dof_size, 1);
        write_tnl_array(f, "double", values, (uint64_t)dof_size, 8);
        write_tnl_array(f, "double", dirichlet, (uint64_t)dof_size, 8);
        f.close();
    });
}

int mhfem_bc_load(const char *path, int64_t *dof_size, uint8_t *tags, double *values, double *dirichlet,
                  int64_t capacity)
{
    return io_guarded([&] {
        if (!path || !dof_size) fail("null argument", MHFEM_E_ARGUMENT);
        TnlReader r{read_file(path), path};
        const std::string t = r.object_type();
        if (t != BC_TYPE)
            fail(std::string("object type does not match (expected ") + BC_TYPE + ", found " + t + ").");
        const int64_t n = r.value<int64_t>();
        const std::vector<uint8_t> tg = r.array<uint8_t>(BC_TAG_TYPE);
        const std::vector<double> v = r.array<double>("double");
        const std::vector<double> d = r.array<double>("double");
        if ((int64_t)tg.size() != n || (int64_t)v.size() != n || (int64_t)d.size() != n)
            fail("Invalid dofSize in BoundaryConditionsStorage: dofSize = " + std::to_string(n) +
                 ", tags.getSize() = " + std::to_string(tg.size()) + ", values.getSize() = " +
                 std::to_string(v.size()) + ", dirichletValues.getSize() = " + std::to_string(d.size()) + ".");
        for (int64_t k = 0; k < n; k++)
            if (tg[k] > 3) fail("invalid boundary condition type " + std::to_string(tg[k]) + " in '" + std::string(path) + "'");
        *dof_size = n;
        if (tags || values || dirichlet) {
            if (capacity < n) fail("output buffers too small for dofSize " + std::to_string(n), MHFEM_E_ARGUMENT);
            if (tags) std::memcpy(tags, tg.data(), n);
            if (values) std::memcpy(values, v.data(), n * 8);
            if (dirichlet) std::memcpy(dirichlet, d.data(), n * 8);
        }
    });
}

int mhfem_save_vector(const char *path, const double *data, int64_t count)
{
    return io_guarded([&] {
        if (!path || count < 0 || (count && !data)) fail("null argument", MHFEM_E_ARGUMENT);
        OutFile f(path);
        write_tnl_array(f, "double", data, (uint64_t)count, 8);
        f.close();
    });
}

long mhfem_load_vector(const char *path, double *data, int64_t capacity)
{
    long n = -1;
    const int rc = io_guarded([&] {
        if (!path) fail("null argument", MHFEM_E_ARGUMENT);
        TnlReader r{read_file(path), path};
        const std::vector<double> v = r.array<double>("double");
        if (data) {
            if (capacity < (int64_t)v.size()) fail("output buffer too small", MHFEM_E_ARGUMENT);
            std::memcpy(data, v.data(), v.size() * 8);
        }
        n = (long)v.size();
    });
    return rc == MHFEM_OK ? n : rc;
}

int mhfem_save_linear_system(mhfem_ctx *ctx, const char *directory)
{
    return io_guarded([&] {
        int rc = MHFEM_OK;
        if (!ctx || !directory) fail("null argument", MHFEM_E_ARGUMENT);
        int rows = 0;
        long nnz = 0;
        if ((rc = mhfem_system_size(ctx, &rows, &nnz)) != MHFEM_OK) fail("linear system not set up", rc);
        std::vector<int32_t> rowptr(rows + 1), col(nnz);
        std::vector<double> val(nnz), rhs(rows), dof(rows);
        if ((rc = mhfem_export_system(ctx, rowptr.data(), col.data(), val.data(), rhs.data())) != MHFEM_OK ||
            (rc = mhfem_get_array(ctx, MHFEM_Z_iF, dof.data(), rows)) != MHFEM_OK)
            fail(mhfem_last_error(ctx), rc);
        const std::string dir(directory);
        {
            OutFile f(dir + "/matrix.mtx");
            f.put("%%MatrixMarket matrix coordinate real general\n");
            f.put(std::to_string(rows) + " " + std::to_string(rows) + " " + std::to_string(nnz) + "\n");
            std::string buf;
            char line[80];
            for (int r = 0; r < rows; r++)
                for (int32_t q = rowptr[r]; q < rowptr[r + 1]; q++) {
                    std::snprintf(line, sizeof line, "%d %d %.17g\n", r + 1, col[q] + 1, val[q]);
                    buf += line;
                    if (buf.size() > (1 << 20)) { f.put(buf); buf.clear(); }
                }
            f.put(buf);
            f.close();
        }
        {
            OutFile f(dir + "/dof.vec.tnl");
            write_tnl_array(f, "double", dof.data(), (uint64_t)rows, 8);
            f.close();
        }
        {
            OutFile f(dir + "/rhs.vec.tnl");
            write_tnl_array(f, "double", rhs.data(), (uint64_t)rows, 8);
            f.close();
        }
    });
}

}  // extern "C"
