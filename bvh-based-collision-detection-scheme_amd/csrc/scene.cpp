// Seeded synthetic deforming meshes (host C++).
//
// The reference ingests one mesh file per frame through assimp and concatenates the meshes
// of a frame into one triangle soup with vertex-id offsets (world/Scene.cpp:14-49,
// world/Scene.h:16-20). This build has no mesh data, so scenes are generated: every vertex
// position is computed in double and rounded once to float, from a splitmix64-seeded
// xoshiro256** stream, so a (kind, params, frame) triple always gives the same bits.
//
// Frozen interpretation choices:
//  * cloth: vertex (i, j) of an n x n grid has material coords x0 = i/(n-1), y = j/(n-1);
//    z = sum_k A(s) sin(kx_k x0 + ky_k y + phi_k) with kx, ky ~ U[2pi, 16pi], phi ~ U[0, 2pi),
//    A(s) = amplitude * s, s = t/(F-1); then x = x0 (1 - compress s). Quad (i, j) gives the
//    triangles (v00, v10, v11) and (v00, v11, v01).
//  * fold (cloth kinds with fold > 0): the part of the sheet on one side of a seeded line is
//    bent about that line over a cylinder of radius fold_radius, by the angle
//    theta(s) = fold * (1 - (1 - s)^4). The line passes through (cx (1 - compress s), cy) with
//    cx, cy ~ U[0.35, 0.65] and direction angle alpha ~ U[0, pi), drawn after the sinusoids.
//    The height z is an offset along the bent sheet's normal. Past theta = pi the flap
//    descends through the unbent part, so distant parts of the sheet interpenetrate.
//  * two spheres: icosahedron + `subdiv` midpoint subdivisions projected to `radius`; centres
//    at -/+ (spacing/2) on x, each moving `speed/2` per frame towards the other; per frame,
//    per vertex, per axis jitter N(0, sigma) added to the rest shape (not accumulated).
//  * layers: `layers` cloths stacked at z offset k*spacing, each with its own sinusoids.
//  * sheet + sphere: a cloth plus an icosphere of `radius` moving on a seeded straight line
//    from z = +0.5 to z = -0.5 through the sheet.
#include "../../include/bvhcd.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <new>
#include <thread>
#include <utility>
#include <vector>

namespace {

struct Rng {  // xoshiro256**, seeded by splitmix64
    uint64_t s[4];
    static uint64_t splitmix(uint64_t& x) {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (auto& v : s) v = splitmix(x);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    // Box-Muller pair in double.
    void normal2(double& z0, double& z1) {
        double u1 = 1.0 - uniform();  // (0, 1]
        double u2 = uniform();
        double r = std::sqrt(-2.0 * std::log(u1));
        z0 = r * std::cos(2.0 * M_PI * u2);
        z1 = r * std::sin(2.0 * M_PI * u2);
    }
};

uint64_t mix_seed(uint64_t a, uint64_t b) {
    uint64_t x = a * 0x9E3779B97F4A7C15ull + b + 0x632BE59BD9B4E019ull;
    return Rng::splitmix(x);
}

struct Mesh {
    std::vector<double> v;    // rest positions, 3 per vertex
    std::vector<int32_t> f;   // 3 per triangle
};

Mesh icosphere(int subdiv, double radius) {
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    Mesh m;
    const double base[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0},
                                {0, -1, t}, {0, 1, t}, {0, -1, -t}, {0, 1, -t},
                                {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (auto& p : base) {
        double l = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        m.v.insert(m.v.end(), {p[0] / l, p[1] / l, p[2] / l});
    }
    m.f = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6,
           7, 1, 8, 3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7,
           9, 8, 1};
    for (int s = 0; s < subdiv; ++s) {
        std::map<std::pair<int, int>, int> mid;
        auto midpoint = [&](int a, int b) {
            auto key = std::make_pair(std::min(a, b), std::max(a, b));
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            double x = m.v[3 * a] + m.v[3 * b], y = m.v[3 * a + 1] + m.v[3 * b + 1],
                   z = m.v[3 * a + 2] + m.v[3 * b + 2];
            double l = std::sqrt(x * x + y * y + z * z);
            int id = (int)(m.v.size() / 3);
            m.v.insert(m.v.end(), {x / l, y / l, z / l});
            mid.emplace(key, id);
            return id;
        };
        std::vector<int32_t> nf;
        nf.reserve(m.f.size() * 4);
        for (size_t i = 0; i < m.f.size(); i += 3) {
            int a = m.f[i], b = m.f[i + 1], c = m.f[i + 2];
            int ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            nf.insert(nf.end(), {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca});
        }
        m.f.swap(nf);
    }
    for (auto& x : m.v) x *= radius;
    return m;
}

struct Cloth {
    int n = 0;
    int frames = 1;
    double amplitude = 0, compress = 0, zoff = 0;
    double kx[4], ky[4], phi[4];
    double fold = 0, fold_r = 0, fcx = 0, fcy = 0, fdx = 1, fdy = 0;
    void init(uint64_t seed, int n_, int frames_, double amp, double comp, double z0,
              double fold_ = 0.0, double fold_radius = 0.0) {
        n = n_; frames = frames_; amplitude = amp; compress = comp; zoff = z0;
        Rng r(seed);
        for (int k = 0; k < 4; ++k) {
            kx[k] = r.uniform(2.0 * M_PI, 16.0 * M_PI);
            ky[k] = r.uniform(2.0 * M_PI, 16.0 * M_PI);
            phi[k] = r.uniform(0.0, 2.0 * M_PI);
        }
        fold = fold_ > 0.0 ? fold_ : 0.0;
        fold_r = fold_radius > 0.0 ? fold_radius : 0.05;
        fcx = r.uniform(0.35, 0.65);
        fcy = r.uniform(0.35, 0.65);
        const double alpha = r.uniform(0.0, M_PI);
        fdx = std::cos(alpha);
        fdy = std::sin(alpha);
    }
    int64_t nv() const { return (int64_t)n * n; }
    int64_t nt() const { return 2 * (int64_t)(n - 1) * (n - 1); }
    void faces(int32_t* f, int32_t off) const {
        int64_t t = 0;
        for (int j = 0; j + 1 < n; ++j)
            for (int i = 0; i + 1 < n; ++i) {
                int32_t v00 = off + j * n + i, v10 = v00 + 1, v01 = v00 + n, v11 = v01 + 1;
                int32_t* p = f + 6 * t;
                p[0] = v00; p[1] = v10; p[2] = v11;
                p[3] = v00; p[4] = v11; p[5] = v01;
                ++t;
            }
    }
    double s_of(int frame) const { return frames > 1 ? (double)frame / (double)(frames - 1) : 0.0; }
    double theta_of(double s) const {
        const double u = 1.0 - s;
        return fold * (1.0 - u * u * u * u);
    }
    // Bends the point (px, py, h) (h = offset along the unbent normal +z) about the fold line.
    void bend(double theta, double cx, double& px, double& py, double& h) const {
        const double ox = px - cx, oy = py - fcy;
        const double u = ox * fdx + oy * fdy;     // along the line
        const double sg = -ox * fdy + oy * fdx;   // signed distance (normal (-dy, dx))
        if (sg <= 0.0 || theta <= 0.0) return;
        const double R = fold_r, arc = R * theta;
        double a, b;
        if (sg <= arc) {
            const double ph = sg / R, sp = std::sin(ph), cp = std::cos(ph);
            a = (R - h) * sp;
            b = R * (1.0 - cp) + h * cp;
        } else {
            const double st = std::sin(theta), ct = std::cos(theta), d = sg - arc;
            a = R * st + d * ct - h * st;
            b = R * (1.0 - ct) + d * st + h * ct;
        }
        px = cx + u * fdx - a * fdy;
        py = fcy + u * fdy + a * fdx;
        h = b;
    }
    // rows [j0, j1)
    void frame_rows(int frame, float* xyz, int j0, int j1) const {
        const double s = s_of(frame), A = amplitude * s, cx = 1.0 - compress * s;
        const double theta = theta_of(s), lx = fcx * cx;
        const double inv = n > 1 ? 1.0 / (double)(n - 1) : 0.0;
        for (int j = j0; j < j1; ++j) {
            const double y = (double)j * inv;
            for (int i = 0; i < n; ++i) {
                const double x0 = (double)i * inv;
                double z = 0.0;
                for (int k = 0; k < 4; ++k) z += A * std::sin(kx[k] * x0 + ky[k] * y + phi[k]);
                double px = x0 * cx, py = y;
                if (fold > 0.0) bend(theta, lx, px, py, z);
                float* p = xyz + 3 * ((int64_t)j * n + i);
                p[0] = (float)px;
                p[1] = (float)py;
                p[2] = (float)(z + zoff);
            }
        }
    }
};

template <class F>
void parallel_rows(int rows, int threads, F&& fn) {
    if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
    if (threads < 1) threads = 1;
    if (threads > rows) threads = rows > 0 ? rows : 1;
    if (threads == 1) { fn(0, rows); return; }
    std::vector<std::thread> ts;
    for (int t = 0; t < threads; ++t) {
        int a = (int)((int64_t)rows * t / threads), b = (int)((int64_t)rows * (t + 1) / threads);
        ts.emplace_back([=, &fn] { fn(a, b); });
    }
    for (auto& th : ts) th.join();
}

}  // namespace

struct bvhcd_scene {
    int kind;
    bvhcd_scene_params p;
    std::vector<Cloth> cloths;
    Mesh sphere;
    double path0[3] = {0, 0, 0}, path1[3] = {0, 0, 0};
    int64_t nv = 0, nt = 0;
};

extern "C" {

int bvhcd_scene_create(int kind, const bvhcd_scene_params* p, bvhcd_scene** out) {
    if (!p || !out) return BVHCD_E_INVALID;
    *out = nullptr;
    bvhcd_scene* s = new (std::nothrow) bvhcd_scene();
    if (!s) return BVHCD_E_OOM;
    s->kind = kind;
    s->p = *p;
    if (s->p.frames < 1) s->p.frames = 1;
    switch (kind) {
        case BVHCD_SCENE_TWO_SPHERES: {
            if (p->subdiv < 0 || p->subdiv > 8) { delete s; return BVHCD_E_INVALID; }
            s->sphere = icosphere(p->subdiv, p->radius);
            int64_t v = (int64_t)s->sphere.v.size() / 3, t = (int64_t)s->sphere.f.size() / 3;
            s->nv = 2 * v; s->nt = 2 * t;
            break;
        }
        case BVHCD_SCENE_CLOTH:
        case BVHCD_SCENE_LAYERS: {
            int layers = kind == BVHCD_SCENE_CLOTH ? 1 : p->layers;
            if (p->n < 2 || layers < 1) { delete s; return BVHCD_E_INVALID; }
            s->cloths.resize(layers);
            for (int k = 0; k < layers; ++k) {
                uint64_t seed = kind == BVHCD_SCENE_CLOTH ? p->seed : mix_seed(p->seed, (uint64_t)k);
                s->cloths[k].init(seed, p->n, s->p.frames, p->amplitude, p->compress,
                                  kind == BVHCD_SCENE_CLOTH ? 0.0 : (double)k * p->spacing,
                                  p->fold, p->fold_radius);
                s->nv += s->cloths[k].nv();
                s->nt += s->cloths[k].nt();
            }
            break;
        }
        case BVHCD_SCENE_SHEET_SPHERE: {
            if (p->n < 2 || p->subdiv < 0 || p->subdiv > 8) { delete s; return BVHCD_E_INVALID; }
            s->cloths.resize(1);
            s->cloths[0].init(p->seed, p->n, s->p.frames, p->amplitude, p->compress, 0.0, p->fold,
                              p->fold_radius);
            s->sphere = icosphere(p->subdiv, p->radius);
            Rng r(mix_seed(p->seed, 0x5A5A5A5Aull));
            s->path0[0] = r.uniform(0.2, 0.8); s->path0[1] = r.uniform(0.2, 0.8); s->path0[2] = 0.5;
            s->path1[0] = r.uniform(0.2, 0.8); s->path1[1] = r.uniform(0.2, 0.8); s->path1[2] = -0.5;
            s->nv = s->cloths[0].nv() + (int64_t)s->sphere.v.size() / 3;
            s->nt = s->cloths[0].nt() + (int64_t)s->sphere.f.size() / 3;
            break;
        }
        default:
            delete s;
            return BVHCD_E_INVALID;
    }
    if (s->nt > INT32_MAX / 2 || s->nv > INT32_MAX / 2) { delete s; return BVHCD_E_INVALID; }
    *out = s;
    return BVHCD_OK;
}

void bvhcd_scene_destroy(bvhcd_scene* s) { delete s; }
int64_t bvhcd_scene_num_tris(const bvhcd_scene* s) { return s ? s->nt : -1; }
int64_t bvhcd_scene_num_verts(const bvhcd_scene* s) { return s ? s->nv : -1; }

int bvhcd_scene_faces(const bvhcd_scene* s, int32_t* faces) {
    if (!s || !faces) return BVHCD_E_INVALID;
    int64_t toff = 0;
    int32_t voff = 0;
    for (const auto& c : s->cloths) {
        c.faces(faces + 3 * toff, voff);
        toff += c.nt();
        voff += (int32_t)c.nv();
    }
    if (s->kind == BVHCD_SCENE_TWO_SPHERES || s->kind == BVHCD_SCENE_SHEET_SPHERE) {
        const int copies = s->kind == BVHCD_SCENE_TWO_SPHERES ? 2 : 1;
        const int32_t sv = (int32_t)(s->sphere.v.size() / 3);
        for (int c = 0; c < copies; ++c) {
            for (size_t i = 0; i < s->sphere.f.size(); ++i) faces[3 * toff + i] = s->sphere.f[i] + voff;
            toff += (int64_t)s->sphere.f.size() / 3;
            voff += sv;
        }
    }
    return BVHCD_OK;
}

int bvhcd_scene_frame(const bvhcd_scene* s, int32_t frame, float* xyz, int32_t threads) {
    if (!s || !xyz || frame < 0 || frame >= s->p.frames) return BVHCD_E_INVALID;
    int64_t voff = 0;
    for (const auto& c : s->cloths) {
        float* base = xyz + 3 * voff;
        parallel_rows(c.n, threads, [&](int a, int b) { c.frame_rows(frame, base, a, b); });
        voff += c.nv();
    }
    const int64_t sv = (int64_t)s->sphere.v.size() / 3;
    if (s->kind == BVHCD_SCENE_TWO_SPHERES) {
        Rng r(mix_seed(s->p.seed, (uint64_t)frame + 1));
        const double half = 0.5 * s->p.spacing, step = 0.5 * s->p.speed * frame;
        const double cx[2] = {-half + step, half - step};
        for (int c = 0; c < 2; ++c)
            for (int64_t v = 0; v < sv; ++v) {
                double j[3], spare;
                r.normal2(j[0], j[1]);
                r.normal2(j[2], spare);
                float* p = xyz + 3 * (voff + c * sv + v);
                p[0] = (float)(cx[c] + s->sphere.v[3 * v] + s->p.sigma * j[0]);
                p[1] = (float)(s->sphere.v[3 * v + 1] + s->p.sigma * j[1]);
                p[2] = (float)(s->sphere.v[3 * v + 2] + s->p.sigma * j[2]);
            }
    } else if (s->kind == BVHCD_SCENE_SHEET_SPHERE) {
        const double u = s->p.frames > 1 ? (double)frame / (double)(s->p.frames - 1) : 0.0;
        double c[3];
        for (int k = 0; k < 3; ++k) c[k] = s->path0[k] + (s->path1[k] - s->path0[k]) * u;
        for (int64_t v = 0; v < sv; ++v) {
            float* p = xyz + 3 * (voff + v);
            for (int k = 0; k < 3; ++k) p[k] = (float)(c[k] + s->sphere.v[3 * v + k]);
        }
    }
    return BVHCD_OK;
}

}  // extern "C"
