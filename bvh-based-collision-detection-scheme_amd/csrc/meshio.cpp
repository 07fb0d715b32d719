// Mesh-sequence ingest and the per-frame log text (host C++).
//
// The reference reads one mesh file per frame through assimp and concatenates every mesh of the
// file into one triangle soup, offsetting each mesh's vertex ids by the vertices before it
// (world/Scene.cpp:14-42); a benchmark names its frames with a printf pattern, a first frame
// and a count (setting/CDBenchmarkSettings.cpp:17-60, e.g. "funnel.plys\%03d.ply" from 1,
// "flowing_cloth\%04d_00.obj" from 0). This file reads the two formats those sequences come in
// without assimp:
//  * Wavefront OBJ: "v x y z" vertices, "f" faces with 1-based or negative (relative) indices in
//    the forms i, i/t, i//n, i/t/n; polygons are split into a fan (0, k, k+1). "o" and "g"
//    groups share the file's vertex list, so ids are already global.
//  * PLY (ascii, binary_little_endian, binary_big_endian): the "vertex" element's x, y, z (any
//    scalar type) and the "face" element's vertex_indices / vertex_index list (fan split);
//    other properties and elements are skipped.
// Coordinates are converted to float (the reference's ExtentType) once.
//
// The frame log restates the reference's Logger output (system/Log/Logger.hpp:68-94 with
// utility/CudaTimer.hpp:43-75): the operation messages and the inspectResults line
// "front (int, ext) cpNum bp actualCpNum np" (collision/bvtt_front/BvttFrontLooseIntra.cu:50-61),
// then one "@section: ms" line per timed stage and "#cd_total_time: ms", "###".
#include "../../include/bvhcd.h"

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <new>
#include <sstream>
#include <string>
#include <vector>

struct bvhcd_mesh {
    std::vector<float> v;     // 3 per vertex
    std::vector<int32_t> f;   // 3 per triangle
    std::string err;
};

namespace {

bool fan(std::vector<int32_t>& out, const std::vector<int64_t>& poly, int64_t nv) {
    for (int64_t id : poly)
        if (id < 0 || id >= nv) return false;
    for (size_t k = 1; k + 1 < poly.size(); ++k) {
        out.push_back((int32_t)poly[0]);
        out.push_back((int32_t)poly[k]);
        out.push_back((int32_t)poly[k + 1]);
    }
    return true;
}

int load_obj(std::istream& in, bvhcd_mesh& m) {
    std::string line;
    std::vector<std::vector<int64_t>> faces;  // resolved after all vertices are known
    int64_t nv = 0;
    while (std::getline(in, line)) {
        size_t p = line.find_first_not_of(" \t\r");
        if (p == std::string::npos || line[p] == '#') continue;
        const char* s = line.c_str() + p;
        if (s[0] == 'v' && (s[1] == ' ' || s[1] == '\t')) {
            char* end = nullptr;
            const char* q = s + 2;
            double x[3];
            for (int k = 0; k < 3; ++k) {
                errno = 0;
                x[k] = std::strtod(q, &end);
                if (end == q) return BVHCD_E_INVALID;
                q = end;
            }
            m.v.insert(m.v.end(), {(float)x[0], (float)x[1], (float)x[2]});
            ++nv;
        } else if (s[0] == 'f' && (s[1] == ' ' || s[1] == '\t')) {
            std::vector<int64_t> poly;
            std::istringstream ss(s + 2);
            std::string tok;
            while (ss >> tok) {
                char* end = nullptr;
                const long long id = std::strtoll(tok.c_str(), &end, 10);
                if (end == tok.c_str() || id == 0) return BVHCD_E_INVALID;
                poly.push_back(id > 0 ? id - 1 : nv + id);  // negative: relative to the vertices so far
            }
            if (poly.size() < 3) return BVHCD_E_INVALID;
            faces.push_back(std::move(poly));
        }
    }
    for (const auto& poly : faces)
        if (!fan(m.f, poly, nv)) return BVHCD_E_INVALID;
    return BVHCD_OK;
}

enum PlyType { P_NONE, P_I8, P_U8, P_I16, P_U16, P_I32, P_U32, P_F32, P_F64 };

PlyType ply_type(const std::string& t) {
    if (t == "char" || t == "int8") return P_I8;
    if (t == "uchar" || t == "uint8") return P_U8;
    if (t == "short" || t == "int16") return P_I16;
    if (t == "ushort" || t == "uint16") return P_U16;
    if (t == "int" || t == "int32") return P_I32;
    if (t == "uint" || t == "uint32") return P_U32;
    if (t == "float" || t == "float32") return P_F32;
    if (t == "double" || t == "float64") return P_F64;
    return P_NONE;
}
int ply_size(PlyType t) {
    switch (t) {
        case P_I8: case P_U8: return 1;
        case P_I16: case P_U16: return 2;
        case P_I32: case P_U32: case P_F32: return 4;
        case P_F64: return 8;
        default: return 0;
    }
}

struct PlyProp {
    std::string name;
    PlyType type = P_NONE, count_type = P_NONE;  // count_type != P_NONE: a list
};
struct PlyElem {
    std::string name;
    int64_t count = 0;
    std::vector<PlyProp> props;
};

// Reads one scalar of type t: binary (swap = big endian on this little-endian host) or ascii.
bool ply_read(std::istream& in, int format, PlyType t, double& out) {
    if (format == 0) {
        std::string tok;
        if (!(in >> tok)) return false;
        char* end = nullptr;
        out = std::strtod(tok.c_str(), &end);
        return end != tok.c_str();
    }
    unsigned char b[8];
    const int n = ply_size(t);
    if (!in.read(reinterpret_cast<char*>(b), n)) return false;
    if (format == 2) std::reverse(b, b + n);
    switch (t) {
        case P_I8: out = (int8_t)b[0]; break;
        case P_U8: out = b[0]; break;
        case P_I16: { int16_t x; std::memcpy(&x, b, 2); out = x; break; }
        case P_U16: { uint16_t x; std::memcpy(&x, b, 2); out = x; break; }
        case P_I32: { int32_t x; std::memcpy(&x, b, 4); out = x; break; }
        case P_U32: { uint32_t x; std::memcpy(&x, b, 4); out = x; break; }
        case P_F32: { float x; std::memcpy(&x, b, 4); out = x; break; }
        case P_F64: { double x; std::memcpy(&x, b, 8); out = x; break; }
        default: return false;
    }
    return true;
}

int load_ply(std::istream& in, bvhcd_mesh& m) {
    std::string line;
    if (!std::getline(in, line) || line.compare(0, 3, "ply") != 0) return BVHCD_E_INVALID;
    int format = -1;  // 0 ascii, 1 binary little endian, 2 binary big endian
    std::vector<PlyElem> elems;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        std::istringstream ss(line);
        std::string kw;
        ss >> kw;
        if (kw == "format") {
            std::string f;
            ss >> f;
            format = f == "ascii" ? 0 : f == "binary_little_endian" ? 1 : f == "binary_big_endian" ? 2 : -1;
        } else if (kw == "element") {
            PlyElem e;
            ss >> e.name >> e.count;
            if (e.count < 0) return BVHCD_E_INVALID;
            elems.push_back(e);
        } else if (kw == "property") {
            if (elems.empty()) return BVHCD_E_INVALID;
            PlyProp p;
            std::string t;
            ss >> t;
            if (t == "list") {
                std::string ct, it;
                ss >> ct >> it >> p.name;
                p.count_type = ply_type(ct);
                p.type = ply_type(it);
                if (p.count_type == P_NONE || p.type == P_NONE) return BVHCD_E_INVALID;
            } else {
                p.type = ply_type(t);
                ss >> p.name;
                if (p.type == P_NONE) return BVHCD_E_INVALID;
            }
            elems.back().props.push_back(p);
        } else if (kw == "end_header") {
            break;
        }
    }
    if (format < 0) return BVHCD_E_INVALID;
    int64_t nv = -1;
    for (const auto& e : elems) {
        const bool is_v = e.name == "vertex", is_f = e.name == "face";
        int xyz[3] = {-1, -1, -1};
        int fl = -1;
        for (size_t k = 0; k < e.props.size(); ++k) {
            const auto& p = e.props[k];
            if (is_v && p.count_type == P_NONE && (p.name == "x" || p.name == "y" || p.name == "z"))
                xyz[p.name[0] - 'x'] = (int)k;
            if (is_f && p.count_type != P_NONE && (p.name == "vertex_indices" || p.name == "vertex_index")) fl = (int)k;
        }
        if (is_v && (xyz[0] < 0 || xyz[1] < 0 || xyz[2] < 0)) return BVHCD_E_INVALID;
        if (is_v) m.v.reserve(3 * (size_t)e.count);
        std::vector<int64_t> poly;
        for (int64_t i = 0; i < e.count; ++i) {
            double pos[3] = {0, 0, 0};
            for (size_t k = 0; k < e.props.size(); ++k) {
                const auto& p = e.props[k];
                if (p.count_type != P_NONE) {
                    double cnt;
                    if (!ply_read(in, format, p.count_type, cnt) || cnt < 0 || cnt > 1e6) return BVHCD_E_INVALID;
                    if ((int)k == fl) poly.clear();
                    for (int64_t j = 0; j < (int64_t)cnt; ++j) {
                        double x;
                        if (!ply_read(in, format, p.type, x)) return BVHCD_E_INVALID;
                        if ((int)k == fl) poly.push_back((int64_t)x);
                    }
                } else {
                    double x;
                    if (!ply_read(in, format, p.type, x)) return BVHCD_E_INVALID;
                    for (int c = 0; c < 3; ++c)
                        if ((int)k == xyz[c]) pos[c] = x;
                }
            }
            if (is_v) m.v.insert(m.v.end(), {(float)pos[0], (float)pos[1], (float)pos[2]});
            if (is_f && fl >= 0) {
                if (nv < 0 || poly.size() < 3 || !fan(m.f, poly, nv)) return BVHCD_E_INVALID;
            }
        }
        if (is_v) nv = e.count;
    }
    return nv >= 0 ? BVHCD_OK : BVHCD_E_INVALID;
}

bool ends_with(const std::string& s, const char* suf) {
    const size_t n = std::strlen(suf);
    if (s.size() < n) return false;
    for (size_t k = 0; k < n; ++k)
        if (std::tolower((unsigned char)s[s.size() - n + k]) != suf[k]) return false;
    return true;
}

}  // namespace

extern "C" {

int bvhcd_mesh_load(const char* path, bvhcd_mesh** out) {
    if (!path || !out) return BVHCD_E_INVALID;
    *out = nullptr;
    const std::string p(path);
    std::ifstream in(p, std::ios::binary);
    if (!in) return BVHCD_E_INVALID;
    bvhcd_mesh* m = new (std::nothrow) bvhcd_mesh();
    if (!m) return BVHCD_E_OOM;
    int st;
    if (ends_with(p, ".obj")) st = load_obj(in, *m);
    else if (ends_with(p, ".ply")) st = load_ply(in, *m);
    else st = BVHCD_E_INVALID;
    if (st != BVHCD_OK || m->v.size() / 3 > (size_t)INT32_MAX / 2 || m->f.size() / 3 > (size_t)INT32_MAX / 2) {
        delete m;
        return st != BVHCD_OK ? st : BVHCD_E_INVALID;
    }
    *out = m;
    return BVHCD_OK;
}

void bvhcd_mesh_destroy(bvhcd_mesh* m) { delete m; }
int64_t bvhcd_mesh_num_verts(const bvhcd_mesh* m) { return m ? (int64_t)m->v.size() / 3 : -1; }
int64_t bvhcd_mesh_num_tris(const bvhcd_mesh* m) { return m ? (int64_t)m->f.size() / 3 : -1; }

int bvhcd_mesh_faces(const bvhcd_mesh* m, int32_t* faces) {
    if (!m || !faces) return BVHCD_E_INVALID;
    std::memcpy(faces, m->f.data(), sizeof(int32_t) * m->f.size());
    return BVHCD_OK;
}

int bvhcd_mesh_verts(const bvhcd_mesh* m, float* xyz) {
    if (!m || !xyz) return BVHCD_E_INVALID;
    std::memcpy(xyz, m->v.data(), sizeof(float) * m->v.size());
    return BVHCD_OK;
}

int bvhcd_frame_path(const char* pattern, int32_t index, char* out, int64_t cap) {
    if (!pattern || !out || cap <= 0) return BVHCD_E_INVALID;
    // exactly one integer conversion (%d, %03d, %4d, ...) and no other conversion but %%
    int convs = 0;
    for (const char* s = pattern; *s; ++s) {
        if (*s != '%') continue;
        if (s[1] == '%') { ++s; continue; }
        const char* q = s + 1;
        while (*q == '0' || *q == '-' || *q == ' ' || *q == '+') ++q;
        while (std::isdigit((unsigned char)*q)) ++q;
        if (*q != 'd' && *q != 'i') return BVHCD_E_INVALID;
        ++convs;
        s = q;
    }
    if (convs != 1) return BVHCD_E_INVALID;
    const int n = std::snprintf(out, (size_t)cap, pattern, index);
    return (n < 0 || n >= cap) ? BVHCD_E_INVALID : BVHCD_OK;
}

}  // extern "C"
