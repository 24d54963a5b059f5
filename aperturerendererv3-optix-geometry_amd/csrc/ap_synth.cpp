// Synthetic benchmark scenes and ray sets (host C++).
#include <cmath>
#include <cstring>
#include <limits>

#include <map>
#include <utility>

#include "aperture/ap_synth.h"
#include "ap_internal.h"
#include "ap_pcg.h"
#include "ap_sample.h"

namespace ap {

// Default-constructed SceneObject with an identity transform
// (SceneObject(center, extent, material, rotation = 0, type),
// core/render/object.cuh:344-364; Material / GBuffer defaults,
// object.cuh:60-91, 124-127).
void init_object(ap_scene_object& o, f3 center, f3 extent, int type) {
    std::memset(&o, 0, sizeof(o));
    o.center = {center.x, center.y, center.z};
    o.extent = {fabsf(extent.x), fabsf(extent.y), fabsf(extent.z)};
    ap_gbuffer& g = o.material.texture.base;
    o.material.shading_model = AP_MAT_DIFFUSE;
    g.albedo = {1.0f, 1.0f, 1.0f};
    g.emissive = {0.0f, 0.0f, 0.0f};
    g.ior = 1.5f;
    g.sigma_a = 100.0f;
    g.normal = {0.0f, 0.0f, 1.0f};
    g.abbe_number = 20.0f;
    o.world_to_object = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    o.object_to_world = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    o.type = type;
    o.additional.sdf_info.cached_volume_index = -1;
    o.object_index = -1;
    o.light_index = -1;
    o.visible = 1;
}

// SceneObject::UpdateProxy (core/render/object.cuh:309-342).
void update_proxy(ap_scene_object& o) {
    const f3 c{o.center.x, o.center.y, o.center.z};
    o.proxy.center = o.center;
    const m3 m{{o.object_to_world.x.x, o.object_to_world.x.y, o.object_to_world.x.z},
               {o.object_to_world.y.x, o.object_to_world.y.y, o.object_to_world.y.z},
               {o.object_to_world.z.x, o.object_to_world.z.y, o.object_to_world.z.z}};
    if (o.type == AP_OBJ_CUBE || o.type == AP_OBJ_SDF) {
        const f3 ax = mul(m, f3{1, 0, 0}), ay = mul(m, f3{0, 1, 0}), az = mul(m, f3{0, 0, 1});
        const f3 e{o.extent.x, o.extent.y, o.extent.z};
        f3 h;
        h.x = fabsf(ax.x) * e.x + fabsf(ay.x) * e.y + fabsf(az.x) * e.z;
        h.y = fabsf(ax.y) * e.x + fabsf(ay.y) * e.y + fabsf(az.y) * e.z;
        h.z = fabsf(ax.z) * e.x + fabsf(ay.z) * e.y + fabsf(az.z) * e.z;
        o.proxy.bound_min = {c.x - h.x, c.y - h.y, c.z - h.z};
        o.proxy.bound_max = {c.x + h.x, c.y + h.y, c.z + h.z};
    } else if (o.type == AP_OBJ_SPHERE) {
        const float r = fmaxf(o.extent.x, fmaxf(o.extent.y, o.extent.z));
        o.proxy.bound_min = {c.x - r, c.y - r, c.z - r};
        o.proxy.bound_max = {c.x + r, c.y + r, c.z + r};
    } else if (o.type == AP_OBJ_TRIANGLE) {
        auto P = [&](const ap_vertex& v) { return mul(m, f3{v.position.x, v.position.y, v.position.z}) + c; };
        const f3 p0 = P(o.additional.v0), p1 = P(o.additional.v1), p2 = P(o.additional.v2);
        o.proxy.bound_min = {fminf(fminf(p0.x, p1.x), p2.x), fminf(fminf(p0.y, p1.y), p2.y),
                             fminf(fminf(p0.z, p1.z), p2.z)};
        o.proxy.bound_max = {fmaxf(fmaxf(p0.x, p1.x), p2.x), fmaxf(fmaxf(p0.y, p1.y), p2.y),
                             fmaxf(fmaxf(p0.z, p1.z), p2.z)};
    }
}

// Rotates v by the unit quaternion (x, y, z, w).
static f3 qrot(const float q[4], f3 v) {
    const f3 u{q[0], q[1], q[2]};
    const float w = q[3];
    const f3 t = 2.0f * cross(u, v);
    return v + w * t + cross(u, t);
}

// Unit icosphere: 12 vertices, 20 faces, each subdivision splits a face in
// four with normalised edge midpoints.
static void icosphere(int subdiv, std::vector<f3>& verts, std::vector<int>& tris) {
    const float t = (1.0f + std::sqrt(5.0f)) / 2.0f;
    verts = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
             {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (f3& v : verts) v = normalize(v);
    tris = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
            3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    for (int s = 0; s < subdiv; ++s) {
        std::vector<int> next;
        std::map<std::pair<int, int>, int> mid;
        auto midpoint = [&](int a, int b) {
            const std::pair<int, int> key{std::min(a, b), std::max(a, b)};
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            verts.push_back(normalize((verts[a] + verts[b]) * 0.5f));
            const int id = static_cast<int>(verts.size()) - 1;
            mid.emplace(key, id);
            return id;
        };
        for (size_t k = 0; k < tris.size(); k += 3) {
            const int a = tris[k], b = tris[k + 1], c = tris[k + 2];
            const int ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            const int nt[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            next.insert(next.end(), nt, nt + 12);
        }
        tris.swap(next);
    }
}

// One triangle object: centre = centroid, local vertices relative to it.
static void triangle_object(ap_scene_object& o, f3 p0, f3 p1, f3 p2, f3 n0, f3 n1, f3 n2, f3 albedo) {
    const f3 c = (p0 + p1 + p2) / 3.0f;
    init_object(o, c, f3{1.0f, 1.0f, 1.0f}, AP_OBJ_TRIANGLE);
    o.material.texture.base.albedo = {albedo.x, albedo.y, albedo.z};
    ap_vertex* vs[3] = {&o.additional.v0, &o.additional.v1, &o.additional.v2};
    const f3 ps[3] = {p0 - c, p1 - c, p2 - c};
    const f3 ns[3] = {n0, n1, n2};
    for (int k = 0; k < 3; ++k) {
        vs[k]->position = {ps[k].x, ps[k].y, ps[k].z};
        vs[k]->normal = {ns[k].x, ns[k].y, ns[k].z};
        vs[k]->uv = {0.0f, 0.0f};
    }
    update_proxy(o);
}

// Value noise on the integer lattice: a SplitMix64 hash of (ix, iz, octave,
// seed) mapped to [-1, 1], blended with the quintic fade 6t^5 - 15t^4 + 10t^3.
static float lattice(int64_t ix, int64_t iz, int octave, uint64_t seed) {
    uint64_t h = splitmix64(seed ^ splitmix64(static_cast<uint64_t>(ix) * 0x9E3779B97F4A7C15ULL ^
                                              (static_cast<uint64_t>(iz) << 21) ^ (static_cast<uint64_t>(octave) << 58)));
    return static_cast<float>(h >> 40) * 0x1p-23f - 1.0f;
}

static float fade(float t) { return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f); }

static float value_noise(float x, float z, int octave, uint64_t seed) {
    const float fx = floorf(x), fz = floorf(z);
    const int64_t ix = static_cast<int64_t>(fx), iz = static_cast<int64_t>(fz);
    const float tx = fade(x - fx), tz = fade(z - fz);
    const float a = lattice(ix, iz, octave, seed), b = lattice(ix + 1, iz, octave, seed);
    const float c = lattice(ix, iz + 1, octave, seed), d = lattice(ix + 1, iz + 1, octave, seed);
    const float ab = a + (b - a) * tx, cd = c + (d - c) * tx;
    return ab + (cd - ab) * tz;
}

// Fractal Brownian motion: octaves of value noise at base frequency 2 over
// [-1, 1]^2, doubling frequency and halving amplitude, scaled so the octave
// amplitudes sum to `amplitude`.
static float fbm(float x, float z, int octaves, float amplitude, uint64_t seed) {
    float sum = 0.0f, amp = 1.0f, norm = 0.0f, freq = 2.0f;
    for (int o = 0; o < octaves; ++o) {
        sum += amp * value_noise(x * freq, z * freq, o, seed);
        norm += amp;
        amp *= 0.5f;
        freq *= 2.0f;
    }
    return amplitude * sum / norm;
}

}  // namespace ap

using namespace ap;

extern "C" {

int ap_synth_triangle_soup(size_t n, uint64_t seed, float edge_min, float edge_max, ap_scene_object* out) {
    if (n > 0 && !out) return fail(AP_ERR_INVALID_ARG, "ap_synth_triangle_soup: NULL out");
    if (!(edge_min > 0.0f) || !(edge_max >= edge_min))
        return fail(AP_ERR_INVALID_ARG, "ap_synth_triangle_soup: bad edge range");
    Pcg32 rng = pcg32_seed(seed, 0);
    const double lmin = std::log(static_cast<double>(edge_min));
    const double lmax = std::log(static_cast<double>(edge_max));
    for (size_t i = 0; i < n; ++i) {
        const f3 c{pcg32_float(rng), pcg32_float(rng), pcg32_float(rng)};
        const float s = static_cast<float>(std::exp(lmin + (lmax - lmin) * pcg32_float(rng)));
        // Uniform rotation (Shoemake 1992).
        const double u1 = pcg32_float(rng), u2 = pcg32_float(rng), u3 = pcg32_float(rng);
        const double a = std::sqrt(1.0 - u1), b = std::sqrt(u1);
        const double tau = 6.283185307179586;
        const float q[4] = {static_cast<float>(a * std::sin(tau * u2)), static_cast<float>(a * std::cos(tau * u2)),
                            static_cast<float>(b * std::sin(tau * u3)), static_cast<float>(b * std::cos(tau * u3))};
        const float R = s / std::sqrt(3.0f);
        const f3 v0 = qrot(q, f3{0.0f, R, 0.0f});
        const f3 v1 = qrot(q, f3{-0.5f * s, -0.5f * R, 0.0f});
        const f3 v2 = qrot(q, f3{0.5f * s, -0.5f * R, 0.0f});
        ap_scene_object& o = out[i];
        init_object(o, c, f3{s, s, s}, AP_OBJ_TRIANGLE);
        const f3 nrm = normalize(cross(v1 - v0, v2 - v0));
        ap_vertex* vs[3] = {&o.additional.v0, &o.additional.v1, &o.additional.v2};
        const f3 ps[3] = {v0, v1, v2};
        for (int k = 0; k < 3; ++k) {
            vs[k]->position = {ps[k].x, ps[k].y, ps[k].z};
            vs[k]->normal = {nrm.x, nrm.y, nrm.z};
            vs[k]->uv = {0.0f, 0.0f};
        }
        update_proxy(o);
    }
    return AP_OK;
}

int ap_synth_box_scene(uint64_t seed, int sphere_subdiv, float light, ap_scene_object* out, size_t capacity,
                       size_t* n_out) {
    if (!n_out || sphere_subdiv < 0 || sphere_subdiv > 7) return fail(AP_ERR_INVALID_ARG, "ap_synth_box_scene: bad arguments");
    std::vector<f3> sv;
    std::vector<int> st;
    icosphere(sphere_subdiv, sv, st);
    const size_t n = 10 + 1 + 2 * (st.size() / 3);
    *n_out = n;
    if (!out) return AP_OK;
    if (capacity < n) return fail(AP_ERR_INVALID_ARG, "ap_synth_box_scene: capacity too small");
    Pcg32 rng = pcg32_seed(seed, 0);
    auto albedo = [&]() {
        const float r = 0.1f + 0.8f * pcg32_float(rng);
        const float g = 0.1f + 0.8f * pcg32_float(rng);
        const float b = 0.1f + 0.8f * pcg32_float(rng);
        return f3{r, g, b};
    };
    size_t k = 0;
    // Quads as (corner, edge u, edge v) with the inward normal u x v.
    struct Quad { f3 c, u, v; };
    const Quad quads[5] = {
        {{-1, -1, -1}, {0, 0, 2}, {2, 0, 0}},   // floor y = -1, normal +y
        {{-1, 1, -1}, {2, 0, 0}, {0, 0, 2}},    // ceiling y = +1, normal -y
        {{-1, -1, -1}, {2, 0, 0}, {0, 2, 0}},   // back z = -1, normal +z
        {{-1, -1, -1}, {0, 2, 0}, {0, 0, 2}},   // left x = -1, normal +x
        {{1, -1, -1}, {0, 0, 2}, {0, 2, 0}},    // right x = +1, normal -x
    };
    for (const Quad& q : quads) {
        const f3 a = albedo();
        const f3 nrm = normalize(cross(q.u, q.v));
        const f3 p00 = q.c, p10 = q.c + q.u, p11 = q.c + q.u + q.v, p01 = q.c + q.v;
        triangle_object(out[k++], p00, p10, p11, nrm, nrm, nrm, a);
        triangle_object(out[k++], p00, p11, p01, nrm, nrm, nrm, a);
    }
    {
        ap_scene_object& L = out[k++];
        init_object(L, f3{0.0f, 1.0f - 0.01f, 0.0f}, f3{0.25f, 0.01f, 0.25f}, AP_OBJ_CUBE);
        L.material.shading_model = AP_MAT_LIGHT;
        L.material.texture.base.albedo = {0.0f, 0.0f, 0.0f};
        L.material.texture.base.emissive = {light, light, light};
        update_proxy(L);
    }
    const f3 centres[2] = {{-0.45f, -0.7f, -0.3f}, {0.45f, -0.7f, 0.35f}};
    for (const f3& c : centres) {
        const f3 a = albedo();
        for (size_t t = 0; t < st.size(); t += 3) {
            const f3 n0 = sv[st[t]], n1 = sv[st[t + 1]], n2 = sv[st[t + 2]];
            triangle_object(out[k++], c + n0 * 0.3f, c + n1 * 0.3f, c + n2 * 0.3f, n0, n1, n2, a);
        }
    }
    return AP_OK;
}

int ap_synth_heightfield_scene(int grid, int n_spheres, uint64_t seed, float light, ap_scene_object* out,
                               size_t capacity, size_t* n_out) {
    if (!n_out || grid < 1 || grid > 8192 || n_spheres < 0 || n_spheres > (1 << 20))
        return fail(AP_ERR_INVALID_ARG, "ap_synth_heightfield_scene: bad arguments");
    std::vector<f3> sv;
    std::vector<int> st;
    icosphere(2, sv, st);
    const size_t per_sphere = st.size() / 3;
    const size_t n_field = 2 * static_cast<size_t>(grid) * grid;
    const size_t n = n_field + static_cast<size_t>(n_spheres) * per_sphere + 4;
    *n_out = n;
    if (!out) return AP_OK;
    if (capacity < n) return fail(AP_ERR_INVALID_ARG, "ap_synth_heightfield_scene: capacity too small");
    constexpr int kOctaves = 6;
    constexpr float kAmplitude = 0.2f;
    const uint64_t noise_seed = seed;
    const int nv = grid + 1;
    const float cell = 2.0f / static_cast<float>(grid);
    std::vector<f3> pos(static_cast<size_t>(nv) * nv), nrm(pos.size());
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nv; ++j)
        for (int i = 0; i < nv; ++i) {
            const float x = -1.0f + cell * static_cast<float>(i), z = -1.0f + cell * static_cast<float>(j);
            pos[static_cast<size_t>(j) * nv + i] = f3{x, fbm(x, z, kOctaves, kAmplitude, noise_seed), z};
        }
    // Vertex normals from central differences of the height samples.
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nv; ++j)
        for (int i = 0; i < nv; ++i) {
            const int i0 = i > 0 ? i - 1 : i, i1 = i < grid ? i + 1 : i;
            const int j0 = j > 0 ? j - 1 : j, j1 = j < grid ? j + 1 : j;
            const float dx = pos[static_cast<size_t>(j) * nv + i1].y - pos[static_cast<size_t>(j) * nv + i0].y;
            const float dz = pos[static_cast<size_t>(j1) * nv + i].y - pos[static_cast<size_t>(j0) * nv + i].y;
            nrm[static_cast<size_t>(j) * nv + i] =
                normalize(f3{-dx / (cell * static_cast<float>(i1 - i0)), 1.0f, -dz / (cell * static_cast<float>(j1 - j0))});
        }
    // Materials: per mesh, Lambert or GGX with equal odds, roughness U[0.05, 1],
    // albedo U[0.05, 0.95] per channel, from PCG32(seed + 1).
    Pcg32 mrng = pcg32_seed(seed + 1, 0);
    struct Mat { int model; float rough; f3 albedo; };
    auto draw_mat = [&]() {
        Mat m;
        m.model = pcg32_float(mrng) < 0.5f ? AP_MAT_DIFFUSE : AP_MAT_GGX;
        m.rough = 0.05f + 0.95f * pcg32_float(mrng);
        const float r = 0.05f + 0.9f * pcg32_float(mrng);
        const float g = 0.05f + 0.9f * pcg32_float(mrng);
        const float b = 0.05f + 0.9f * pcg32_float(mrng);
        m.albedo = f3{r, g, b};
        return m;
    };
    auto apply = [](ap_scene_object& o, const Mat& m) {
        o.material.shading_model = m.model;
        o.material.texture.base.roughness = m.rough;
        o.material.texture.base.specular = 0.5f;
        o.material.texture.base.metallic = 0.0f;
    };
    const Mat field = draw_mat();
#pragma omp parallel for schedule(static)
    for (int j = 0; j < grid; ++j)
        for (int i = 0; i < grid; ++i) {
            const size_t a = static_cast<size_t>(j) * nv + i, b = a + 1, c = a + nv, d = c + 1;
            const size_t k = 2 * (static_cast<size_t>(j) * grid + i);
            triangle_object(out[k], pos[a], pos[c], pos[d], nrm[a], nrm[c], nrm[d], field.albedo);
            triangle_object(out[k + 1], pos[a], pos[d], pos[b], nrm[a], nrm[d], nrm[b], field.albedo);
            apply(out[k], field);
            apply(out[k + 1], field);
        }
    // Spheres resting on the surface: centre (x, h(x, z) + r, z), radius
    // log-uniform in [0.01, 0.1], from PCG32(seed + 2).
    Pcg32 srng = pcg32_seed(seed + 2, 0);
    size_t k = n_field;
    for (int sidx = 0; sidx < n_spheres; ++sidx) {
        const float r = 0.01f * expf(logf(10.0f) * pcg32_float(srng));
        const float x = -0.9f + 1.8f * pcg32_float(srng), z = -0.9f + 1.8f * pcg32_float(srng);
        const f3 c{x, fbm(x, z, kOctaves, kAmplitude, noise_seed) + r, z};
        const Mat m = draw_mat();
        for (size_t t = 0; t < st.size(); t += 3) {
            const f3 n0 = sv[st[t]], n1 = sv[st[t + 1]], n2 = sv[st[t + 2]];
            triangle_object(out[k], c + n0 * r, c + n1 * r, c + n2 * r, n0, n1, n2, m.albedo);
            apply(out[k], m);
            ++k;
        }
    }
    // Four thin-cube area lights (0.3 x 0.3) over the field.
    const f3 lights[4] = {{-0.5f, 1.2f, -0.5f}, {0.5f, 1.3f, -0.4f}, {-0.4f, 1.25f, 0.5f}, {0.45f, 1.15f, 0.45f}};
    for (const f3& c : lights) {
        ap_scene_object& L = out[k++];
        init_object(L, c, f3{0.15f, 0.005f, 0.15f}, AP_OBJ_CUBE);
        L.material.shading_model = AP_MAT_LIGHT;
        L.material.texture.base.albedo = {0.0f, 0.0f, 0.0f};
        L.material.texture.base.emissive = {light, light, light};
        update_proxy(L);
    }
    return AP_OK;
}

int ap_synth_camera_rays(int width, int height, const float origin[3], const float direction[3], float fov_deg,
                         ap_ray* out) {
    if (width <= 0 || height <= 0 || !origin || !direction || !out)
        return fail(AP_ERR_INVALID_ARG, "ap_synth_camera_rays: bad arguments");
    const f3 o{origin[0], origin[1], origin[2]};
    const f3 dir{direction[0], direction[1], direction[2]};
    const float fov = fov_deg * 3.1415926f / 180.0f;
    const float aspect = static_cast<float>(width) / fmaxf(1.0f, static_cast<float>(height));
    const float tan_y = tanf(0.5f * fov);
    const float tan_x = tan_y * aspect;
    const f3 right = normalize(cross(dir, f3{0.0f, 1.0f, 0.0f}));
    const f3 up = normalize(cross(dir, right));
    const float inf = std::numeric_limits<float>::infinity();
#pragma omp parallel for schedule(static)
    for (int j = 0; j < height; ++j) {
        for (int i = 0; i < width; ++i) {
            const float u = (static_cast<float>(i) + 0.5f) / static_cast<float>(width);
            const float v = (static_cast<float>(j) + 0.5f) / static_cast<float>(height);
            const float nx = u * 2.0f - 1.0f;
            const float ny = static_cast<float>((1.0 - static_cast<double>(v)) * 2.0f - 1.0f);
            const f3 d = normalize(dir + right * (nx * tan_x) + up * (ny * tan_y));
            ap_ray& r = out[static_cast<size_t>(j) * width + i];
            r = ap_ray{o.x, o.y, o.z, 0.0f, d.x, d.y, d.z, inf};
        }
    }
    return AP_OK;
}

int ap_synth_secondary_rays(const ap_scene* scene, const ap_ray* prim, const ap_hit* hits, size_t n,
                            uint64_t seed, float offset, ap_ray* out, size_t* n_out) {
    if (!scene || (n > 0 && (!prim || !hits || !out)) || !n_out)
        return fail(AP_ERR_INVALID_ARG, "ap_synth_secondary_rays: bad arguments");
    const float inf = std::numeric_limits<float>::infinity();
    // Prefix positions so the output stays in primary order.
    std::vector<size_t> pos(n + 1, 0);
    for (size_t i = 0; i < n; ++i) pos[i + 1] = pos[i] + (hits[i].id >= 0 ? 1 : 0);
    const int64_t nn = static_cast<int64_t>(n);
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < nn; ++ii) {
        const size_t i = static_cast<size_t>(ii);
        const int32_t id = hits[i].id;
        if (id < 0) continue;
        const ap_scene_object& ob = scene->objects[id];
        const f3 o{prim[i].ox, prim[i].oy, prim[i].oz};
        const f3 d{prim[i].dx, prim[i].dy, prim[i].dz};
        f3 nrm{0.0f, 1.0f, 0.0f};
        if (ob.type == AP_OBJ_TRIANGLE) {
            const m3 m{{ob.object_to_world.x.x, ob.object_to_world.x.y, ob.object_to_world.x.z},
                       {ob.object_to_world.y.x, ob.object_to_world.y.y, ob.object_to_world.y.z},
                       {ob.object_to_world.z.x, ob.object_to_world.z.y, ob.object_to_world.z.z}};
            const f3 c{ob.center.x, ob.center.y, ob.center.z};
            auto P = [&](const ap_vertex& v) { return mul(m, f3{v.position.x, v.position.y, v.position.z}) + c; };
            const f3 p0 = P(ob.additional.v0), p1 = P(ob.additional.v1), p2 = P(ob.additional.v2);
            nrm = normalize(cross(p1 - p0, p2 - p0));
        }
        if (dot(nrm, d) > 0.0f) nrm = -nrm;
        const f3 hp = o + d * hits[i].t;
        const f3 org = hp + nrm * offset;
        Pcg32 rng = pcg32_seed(seed, static_cast<uint64_t>(i));
        const float e0 = pcg32_float(rng), e1 = pcg32_float(rng), e2 = pcg32_float(rng);
        const DirSample s = cosine_hemisphere(f3{e0, e1, e2}, nrm);
        out[pos[i]] = ap_ray{org.x, org.y, org.z, 0.0f, s.dir.x, s.dir.y, s.dir.z, inf};
    }
    *n_out = pos[n];
    return AP_OK;
}

}  // extern "C"
