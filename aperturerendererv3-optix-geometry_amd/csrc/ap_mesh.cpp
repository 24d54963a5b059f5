// Mesh ingestion (include/aperture/ap_mesh.h: ap_mesh_load_obj): an OBJ
// file to triangle SceneObjects as LoadTexturedMesh builds them
// (core/helper/scene_helper.h:453-564), over the reading tinyobjloader 2.0
// gives the file (core/helper/tiny_obj_loader.h, vendored in the reference:
// LoadObj with triangulate = true and no mapbox earcut).
//
// Parsed records: v x y z [...], vt u [v] [...], vn x y z, f with
// i | i/j | i//k | i/j/k vertices (1-based, or negative counting back from
// the records read so far). Every other record (o, g, s, usemtl, mtllib, l,
// p, comments) leaves the faces as they are: tinyobjloader splits shapes
// there, but LoadTexturedMesh walks every shape in file order, so the
// triangles come out in face order either way.
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "aperture/ap_mesh.h"
#include "ap_internal.h"

namespace {

using ap::f3;

struct Idx {
    int v, vt, vn;  // 0-based, -1 when absent
};

struct Obj {
    std::vector<float> v, vt, vn;  // 3, 2 and 3 floats per record
    std::vector<Idx> tri;          // three per triangle, in face order
};

bool is_space(char c) { return c == ' ' || c == '\t'; }

// A real the way tinyobjloader reads one: a double, then real_t (float).
float parse_real(const char*& p, float dflt, bool* got = nullptr) {
    while (is_space(*p)) ++p;
    char* end = nullptr;
    const double d = std::strtod(p, &end);
    if (end == p) {
        if (got) *got = false;
        return dflt;
    }
    p = end;
    if (got) *got = true;
    return static_cast<float>(d);
}

// fixIndex: 1-based index or a negative one relative to the count so far.
bool fix_index(long idx, size_t n, int* out) {
    if (idx > 0) {
        *out = static_cast<int>(idx - 1);
        return true;
    }
    if (idx < 0) {
        *out = static_cast<int>(static_cast<long>(n) + idx);
        return true;
    }
    return false;  // 0 is not an OBJ index
}

bool parse_triple(const char*& p, const Obj& o, Idx* out) {
    char* end = nullptr;
    out->v = out->vt = out->vn = -1;
    const long a = std::strtol(p, &end, 10);
    if (end == p || !fix_index(a, o.v.size() / 3, &out->v)) return false;
    p = end;
    if (*p != '/') return true;
    ++p;
    if (*p == '/') {  // i//k
        ++p;
        const long c = std::strtol(p, &end, 10);
        if (end == p || !fix_index(c, o.vn.size() / 3, &out->vn)) return false;
        p = end;
        return true;
    }
    const long b = std::strtol(p, &end, 10);  // i/j or i/j/k
    if (end == p || !fix_index(b, o.vt.size() / 2, &out->vt)) return false;
    p = end;
    if (*p != '/') return true;
    ++p;
    const long c = std::strtol(p, &end, 10);
    if (end == p || !fix_index(c, o.vn.size() / 3, &out->vn)) return false;
    p = end;
    return true;
}

bool valid(const Obj& o, const Idx& i) {
    return i.v >= 0 && static_cast<size_t>(i.v) * 3 + 2 < o.v.size() &&
           (i.vt < 0 || static_cast<size_t>(i.vt) * 2 + 1 < o.vt.size()) &&
           (i.vn < 0 || static_cast<size_t>(i.vn) * 3 + 2 < o.vn.size());
}

// Even-odd point-in-polygon test of (tx, ty) (pnpoly, tiny_obj_loader.h:7113-7123).
bool point_in_triangle(const float vx[3], const float vy[3], float tx, float ty) {
    bool c = false;
    for (int i = 0, j = 2; i < 3; j = i++)
        if (((vy[i] > ty) != (vy[j] > ty)) && (tx < (vx[j] - vx[i]) * (ty - vy[i]) / (vy[j] - vy[i]) + vx[i])) c = !c;
    return c;
}

// tinyobjloader's built-in ear clipping of a face of five or more vertices
// (tiny_obj_loader.h:7431-7610): project on the two axes the first corner's
// normal does not dominate, then walk the polygon from a guessed vertex,
// clipping a corner whose turn agrees with its sign test and whose triangle
// holds no other vertex, until three are left or a full pass clips nothing.
void ear_clip(Obj& o, const std::vector<Idx>& face) {
    const float* V = o.v.data();
    auto P = [&](const Idx& i, int a) { return V[static_cast<size_t>(i.v) * 3 + a]; };
    const size_t n0 = face.size();
    int axes[2] = {1, 2};
    for (size_t k = 0; k < n0; ++k) {
        const Idx& a = face[k % n0];
        const Idx& b = face[(k + 1) % n0];
        const Idx& c = face[(k + 2) % n0];
        const float e0x = P(b, 0) - P(a, 0), e0y = P(b, 1) - P(a, 1), e0z = P(b, 2) - P(a, 2);
        const float e1x = P(c, 0) - P(b, 0), e1y = P(c, 1) - P(b, 1), e1z = P(c, 2) - P(b, 2);
        const float cx = std::fabs(e0y * e1z - e0z * e1y);
        const float cy = std::fabs(e0z * e1x - e0x * e1z);
        const float cz = std::fabs(e0x * e1y - e0y * e1x);
        if (cx > FLT_EPSILON || cy > FLT_EPSILON || cz > FLT_EPSILON) {
            if (!(cx > cy && cx > cz)) {
                axes[0] = 0;
                if (cz > cx && cz > cy) axes[1] = 1;
            }
            break;
        }
    }
    std::vector<Idx> rem = face;
    size_t guess = 0;
    size_t iterations = rem.size();
    size_t previous = rem.size();
    while (rem.size() > 3 && iterations > 0) {
        const size_t n = rem.size();
        if (guess >= n) guess -= n;
        if (previous != n) {
            previous = n;
            iterations = n;
        } else {
            --iterations;
        }
        Idx ind[3];
        float vx[3], vy[3];
        for (int k = 0; k < 3; ++k) {
            ind[k] = rem[(guess + k) % n];
            vx[k] = P(ind[k], axes[0]);
            vy[k] = P(ind[k], axes[1]);
        }
        const float e0x = vx[1] - vx[0], e0y = vy[1] - vy[0];
        const float e1x = vx[2] - vx[1], e1y = vy[2] - vy[1];
        const float cross = e0x * e1y - e0y * e1x;
        const float area = (vx[0] * vy[1] - vy[0] * vx[1]) * 0.5f;
        if (cross * area < 0.0f) {
            guess += 1;
            continue;
        }
        bool overlap = false;
        for (size_t other = 3; other < n; ++other) {
            const Idx& t = rem[(guess + other) % n];
            if (point_in_triangle(vx, vy, P(t, axes[0]), P(t, axes[1]))) {
                overlap = true;
                break;
            }
        }
        if (overlap) {
            guess += 1;
            continue;
        }
        o.tri.insert(o.tri.end(), ind, ind + 3);
        rem.erase(rem.begin() + static_cast<std::ptrdiff_t>((guess + 1) % n));
    }
    if (rem.size() == 3) o.tri.insert(o.tri.end(), rem.begin(), rem.end());
}

// exportGroupsToShape's triangulation of one face (tiny_obj_loader.h:7157-7290):
// triangles as they are, quads split along the shorter diagonal, larger
// faces ear-clipped.
int add_face(Obj& o, const std::vector<Idx>& f, std::string* err) {
    const size_t n = f.size();
    if (n < 3) return AP_OK;  // degenerate face: skipped
    for (const Idx& i : f)
        if (!valid(o, i)) {
            *err = "face with an index out of range";
            return AP_ERR_INVALID_ARG;
        }
    if (n == 3) {
        o.tri.insert(o.tri.end(), f.begin(), f.end());
        return AP_OK;
    }
    if (n > 4) {
        ear_clip(o, f);
        return AP_OK;
    }
    const float* V = o.v.data();
    auto at = [&](int k, int c) { return V[static_cast<size_t>(f[k].v) * 3 + c]; };
    const float e02x = at(2, 0) - at(0, 0), e02y = at(2, 1) - at(0, 1), e02z = at(2, 2) - at(0, 2);
    const float e13x = at(3, 0) - at(1, 0), e13y = at(3, 1) - at(1, 1), e13z = at(3, 2) - at(1, 2);
    const float sqr02 = e02x * e02x + e02y * e02y + e02z * e02z;
    const float sqr13 = e13x * e13x + e13y * e13y + e13z * e13z;
    const int order[2][6] = {{0, 1, 2, 0, 2, 3}, {0, 1, 3, 1, 2, 3}};
    const int* ord = order[sqr02 < sqr13 ? 0 : 1];
    for (int k = 0; k < 6; ++k) o.tri.push_back(f[ord[k]]);
    return AP_OK;
}

int read_obj(const char* path, Obj& o, std::string* err) {
    FILE* fp = std::fopen(path, "rb");
    if (!fp) {
        *err = std::string("cannot open ") + path;
        return AP_ERR_INVALID_ARG;
    }
    std::string line;
    std::vector<Idx> face;
    int rc = AP_OK;
    int ch = 0;
    while (rc == AP_OK && ch != EOF) {
        line.clear();
        while ((ch = std::fgetc(fp)) != EOF && ch != '\n' && ch != '\r') line.push_back(static_cast<char>(ch));
        const char* p = line.c_str();
        while (is_space(*p)) ++p;
        if (p[0] == 'v' && is_space(p[1])) {
            p += 2;
            for (int c = 0; c < 3; ++c) o.v.push_back(parse_real(p, 0.0f));
        } else if (p[0] == 'v' && p[1] == 'n' && is_space(p[2])) {
            p += 3;
            for (int c = 0; c < 3; ++c) o.vn.push_back(parse_real(p, 0.0f));
        } else if (p[0] == 'v' && p[1] == 't' && is_space(p[2])) {
            p += 3;
            o.vt.push_back(parse_real(p, 0.0f));
            o.vt.push_back(parse_real(p, 0.0f));
        } else if (p[0] == 'f' && is_space(p[1])) {
            p += 2;
            face.clear();
            for (;;) {
                while (is_space(*p)) ++p;
                if (*p == '\0') break;
                Idx i;
                if (!parse_triple(p, o, &i)) {
                    *err = "failed to parse an f record: " + line;
                    rc = AP_ERR_INVALID_ARG;
                    break;
                }
                face.push_back(i);
                while (*p && !is_space(*p)) ++p;
            }
            if (rc == AP_OK) rc = add_face(o, face, err);
        }
    }
    std::fclose(fp);
    return rc;
}

}  // namespace

using namespace ap;

extern "C" int ap_mesh_load_obj(const char* path, const float position[3], float scale, const uint64_t textures[4],
                                int32_t flags, ap_scene_object* out, size_t capacity, size_t* count) {
    if (!path || !position || !count) return fail(AP_ERR_INVALID_ARG, "ap_mesh_load_obj: NULL argument");
    *count = 0;
    Obj o;
    std::string err;
    const int rc = read_obj(path, o, &err);
    if (rc != AP_OK) return fail(rc, "ap_mesh_load_obj: " + err);
    const size_t n_tri = o.tri.size() / 3;
    *count = n_tri;
    if (!out) return AP_OK;
    const f3 c{position[0], position[1], position[2]};
    const m3 I{{1.0f, 0.0f, 0.0f}, {0.0f, 1.0f, 0.0f}, {0.0f, 0.0f, 1.0f}};
    for (size_t t = 0; t < n_tri && t < capacity; ++t) {
        ap_scene_object& s = out[t];
        std::memset(&s, 0, sizeof(s));
        // SceneObject(position, (scale, scale, scale), CreateGGXPureColor(White, 0.4, 0, 0), 0, OBJ_TRIANGLE).
        s.center = {position[0], position[1], position[2]};
        s.extent = {std::fabs(scale), std::fabs(scale), std::fabs(scale)};
        s.material.shading_model = AP_MAT_GGX;
        ap_gbuffer& g = s.material.texture.base;
        g.albedo = {1.0f, 1.0f, 1.0f};
        g.roughness = 0.4f;
        g.ior = 1.5f;
        g.sigma_a = 100.0f;
        g.normal = {0.0f, 0.0f, 1.0f};
        g.abbe_number = 20.0f;
        if (textures) {
            g.albedo_tex = textures[0];
            g.normal_tex = textures[1];
            g.metallic_tex = textures[2];
            g.roughness_tex = textures[3];
        }
        const float* Mi = &I.x.x;
        std::memcpy(&s.world_to_object, Mi, sizeof(float) * 9);
        std::memcpy(&s.object_to_world, Mi, sizeof(float) * 9);
        s.type = AP_OBJ_TRIANGLE;
        s.additional.sdf_info.cached_volume_index = -1;
        ap_vertex* vs[3] = {&s.additional.v0, &s.additional.v1, &s.additional.v2};
        f3 p[3];
        for (int k = 0; k < 3; ++k) {
            const Idx& i = o.tri[3 * t + k];
            const float* V = &o.v[static_cast<size_t>(i.v) * 3];
            p[k] = f3{V[0] * scale, V[1] * scale, V[2] * scale};
            vs[k]->position = {p[k].x, p[k].y, p[k].z};
            if (i.vn >= 0) {
                const float* N = &o.vn[static_cast<size_t>(i.vn) * 3];
                vs[k]->normal = {N[0], N[1], N[2]};
            } else {
                vs[k]->normal = {0.0f, 1.0f, 0.0f};
            }
            if (i.vt >= 0) {
                const float* T = &o.vt[static_cast<size_t>(i.vt) * 2];
                vs[k]->uv = {T[0], 1.0f - T[1]};
            } else {
                vs[k]->uv = {0.0f, 0.0f};
            }
        }
        if (o.tri[3 * t].vn < 0) {  // the face's first vertex has no normal: the face normal for all three
            const f3 n = normalize(cross(p[1] - p[0], p[2] - p[0]));
            for (int k = 0; k < 3; ++k) vs[k]->normal = {n.x, n.y, n.z};
        }
        // UpdateProxy (object.cuh:309-342) for a triangle.
        f3 w[3];
        for (int k = 0; k < 3; ++k) w[k] = mul(I, p[k]) + c;
        // The reference's centre (the mesh position), or the centroid the builder can bin.
        if (flags & AP_MESH_REFERENCE_PROXY) {
            s.proxy.center = s.center;
        } else {
            const f3 m = ((w[0] + w[1]) + w[2]) / 3.0f;
            s.proxy.center = {m.x, m.y, m.z};
        }
        s.proxy.bound_min = {std::fmin(std::fmin(w[0].x, w[1].x), w[2].x), std::fmin(std::fmin(w[0].y, w[1].y), w[2].y),
                             std::fmin(std::fmin(w[0].z, w[1].z), w[2].z)};
        s.proxy.bound_max = {std::fmax(std::fmax(w[0].x, w[1].x), w[2].x), std::fmax(std::fmax(w[0].y, w[1].y), w[2].y),
                             std::fmax(std::fmax(w[0].z, w[1].z), w[2].z)};
        s.object_index = -1;
        s.light_index = -1;
        s.visible = 1;
    }
    return AP_OK;
}
