// OBJ ingestion (io/obj.rs:16-124 OBJ::read, then TriangleMesh::new,
// shapes/triangle.rs:48-90): the reference's own scene loader, so that a
// bonsai scene built from an OBJ file reaches the GPU path unchanged.
//
// Parsing follows the reference line by line: after trimming, a line starting
// with 'f' contributes every whitespace-separated token after the first as one
// face vertex ("v", "v/t", "v/t/n" or "v//n", 1-based indices); a line starting
// with 'v' is "vt" (2 floats), "vn" (3 floats) or a position (3 floats).
// Triangles are consecutive triples of the face-vertex list (faces with more
// than three vertices are not fanned, as in the reference). Positions go
// through Point3::transform (point3.rs:51-69) and normals through
// Normal::transform (normal.rs:21-30) unless the transform is the identity.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "bg_internal.h"

namespace {

struct ObjData {
    std::vector<float> verts, normals, uvs;  // 3, 3, 2 floats per entry
    std::vector<uint64_t> vi, ni, ti;         // face-vertex indices, 1-based
};

bool parse_index(const std::string& s, uint64_t* out) {
    if (s.empty()) return false;
    size_t k = 0;
    if (s[0] == '+') k = 1;  // Rust's usize parser accepts a leading '+'
    if (k == s.size()) return false;
    uint64_t v = 0;
    for (; k < s.size(); ++k) {
        if (s[k] < '0' || s[k] > '9') return false;
        v = v * 10 + (uint64_t)(s[k] - '0');
    }
    *out = v;
    return true;
}

bool parse_float(const std::string& s, float* out) {
    if (s.empty()) return false;
    if (s.find('x') != std::string::npos || s.find('X') != std::string::npos) return false;  // no hex floats
    char* end = nullptr;
    *out = std::strtof(s.c_str(), &end);
    return end && *end == '\0';
}

std::vector<std::string> split_ws(const std::string& line) {
    std::vector<std::string> out;
    std::istringstream is(line);
    std::string tok;
    while (is >> tok) out.push_back(tok);
    return out;
}

std::vector<std::string> split_on(const std::string& s, const std::string& sep) {
    std::vector<std::string> out;
    size_t start = 0;
    for (;;) {
        const size_t p = s.find(sep, start);
        if (p == std::string::npos) {
            out.push_back(s.substr(start));
            return out;
        }
        out.push_back(s.substr(start, p - start));
        start = p + sep.size();
    }
}

std::string trim(const std::string& s) {
    const char* ws = " \t\r\n\v\f";
    const size_t a = s.find_first_not_of(ws);
    if (a == std::string::npos) return "";
    const size_t b = s.find_last_not_of(ws);
    return s.substr(a, b - a + 1);
}

int parse(const char* path, ObjData* d) {
    std::ifstream f(path);
    if (!f) return BG_ERR_INVALID;
    std::string raw;
    while (std::getline(f, raw)) {
        const std::string line = trim(raw);
        if (!line.empty() && line[0] == 'f') {
            const std::vector<std::string> tok = split_ws(line);
            for (size_t k = 1; k < tok.size(); ++k) {
                const std::string& v = tok[k];
                uint64_t a, b, c;
                if (v.find("//") != std::string::npos) {
                    const auto p = split_on(v, "//");
                    if (p.size() != 2 || !parse_index(p[0], &a) || !parse_index(p[1], &b)) return BG_ERR_INVALID;
                    d->vi.push_back(a);
                    d->ni.push_back(b);
                } else if (v.find('/') != std::string::npos) {
                    const auto p = split_on(v, "/");
                    if ((p.size() != 2 && p.size() != 3) || !parse_index(p[0], &a) || !parse_index(p[1], &b))
                        return BG_ERR_INVALID;
                    d->vi.push_back(a);
                    d->ti.push_back(b);
                    if (p.size() == 3) {
                        if (!parse_index(p[2], &c)) return BG_ERR_INVALID;
                        d->ni.push_back(c);
                    }
                } else {
                    if (!parse_index(v, &a)) return BG_ERR_INVALID;
                    d->vi.push_back(a);
                }
            }
        }
        if (!line.empty() && line[0] == 'v') {
            const std::vector<std::string> tok = split_ws(line);
            std::vector<float> p;
            for (size_t k = 1; k < tok.size(); ++k) {
                float x;
                if (!parse_float(tok[k], &x)) return BG_ERR_INVALID;
                p.push_back(x);
            }
            if (line.rfind("vt", 0) == 0) {
                if (p.size() < 2) return BG_ERR_INVALID;
                d->uvs.insert(d->uvs.end(), {p[0], p[1]});
            } else if (line.rfind("vn", 0) == 0) {
                if (p.size() < 3) return BG_ERR_INVALID;
                d->normals.insert(d->normals.end(), {p[0], p[1], p[2]});
            } else {
                if (p.size() < 3) return BG_ERR_INVALID;
                d->verts.insert(d->verts.end(), {p[0], p[1], p[2]});
            }
        }
    }
    return BG_OK;
}

bool is_identity(const float* m) {
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c)
            if (m[4 * r + c] != (r == c ? 1.0f : 0.0f)) return false;
    return true;
}

// Point3::transform (point3.rs:51-69)
void xform_point(const float* m, const float* p, float* out) {
    const float xp = m[0] * p[0] + m[1] * p[1] + m[2] * p[2] + m[3];
    const float yp = m[4] * p[0] + m[5] * p[1] + m[6] * p[2] + m[7];
    const float zp = m[8] * p[0] + m[9] * p[1] + m[10] * p[2] + m[11];
    const float wp = m[12] * p[0] + m[13] * p[1] + m[14] * p[2] + m[15];
    if (wp == 1.0f) {
        out[0] = xp; out[1] = yp; out[2] = zp;
        return;
    }
    const float iw = 1.0f / wp;
    out[0] = xp * iw; out[1] = yp * iw; out[2] = zp * iw;
}

// Normal::transform (normal.rs:21-30): the transpose of m_inverse
void xform_normal(const float* mi, const float* n, float* out) {
    out[0] = mi[0] * n[0] + mi[4] * n[1] + mi[8] * n[2];
    out[1] = mi[1] * n[0] + mi[5] * n[1] + mi[9] * n[2];
    out[2] = mi[2] * n[0] + mi[6] * n[1] + mi[10] * n[2];
}

}  // namespace

extern "C" {

int bg_obj_read(const char* path, const float object_to_world[16], const float world_to_object[16], uint64_t* n_tris,
                uint32_t* has, float* positions, float* normals, float* uvs) {
    if (!path || !object_to_world || !world_to_object || !n_tris || !has) return BG_ERR_INVALID;
    ObjData d;
    const int rc = parse(path, &d);
    if (rc) return rc;
    // every index in range (the reference indexes its vectors and panics otherwise)
    for (uint64_t v : d.vi)
        if (v == 0 || v > d.verts.size() / 3) return BG_ERR_INVALID;
    for (uint64_t v : d.ni)
        if (v == 0 || v > d.normals.size() / 3) return BG_ERR_INVALID;
    for (uint64_t v : d.ti)
        if (v == 0 || v > d.uvs.size() / 2) return BG_ERR_INVALID;
    const uint64_t nv = d.vi.size();
    // per-vertex normals / uvs exist when the reference's mesh would hold one per position
    if ((!d.ni.empty() && d.ni.size() != nv) || (!d.ti.empty() && d.ti.size() != nv)) return BG_ERR_UNSUPPORTED;
    *n_tris = nv / 3;
    *has = (d.ni.empty() ? 0u : BG_OBJ_NORMALS) | (d.ti.empty() ? 0u : BG_OBJ_UVS);
    if (!positions) return BG_OK;
    const bool ident = is_identity(object_to_world);
    for (uint64_t k = 0; k < 3 * *n_tris; ++k) {
        const float* p = &d.verts[3 * (d.vi[k] - 1)];
        if (ident) std::memcpy(positions + 3 * k, p, 3 * sizeof(float));
        else xform_point(object_to_world, p, positions + 3 * k);
        if (normals && !d.ni.empty()) {
            const float* n = &d.normals[3 * (d.ni[k] - 1)];
            if (ident) std::memcpy(normals + 3 * k, n, 3 * sizeof(float));
            else xform_normal(world_to_object, n, normals + 3 * k);
        }
        if (uvs && !d.ti.empty()) std::memcpy(uvs + 2 * k, &d.uvs[2 * (d.ti[k] - 1)], 2 * sizeof(float));
    }
    return BG_OK;
}

}  // extern "C"
