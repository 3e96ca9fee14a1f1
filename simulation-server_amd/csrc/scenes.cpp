// scenes.cpp -- reproducible scene construction: cell models, vein meshes and
// initial particle states (host only).
//
// The reference hard-codes one scene at compile time (presets in
// src/config/blood_cell_presets.hpp, a generated Y-shaped vein in
// src/config/vein_definition.hpp) and places cells with time-seeded curand
// (src/simulation/simulation_controller.cu:156-243: x = centroid + model
// vertex, v = cell velocity, F = 0).  Here every piece is generated from
// parameters and a 64-bit seed so that the GPU path and the CPU oracle start
// from the same bytes.
#include "bcs_internal.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <memory>
#include <unordered_map>
#include <vector>

using bcs::fail;

namespace {

struct D3 {
    double x, y, z;
};
inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 operator*(double s, D3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 crossd(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm(D3 a) { return std::sqrt(dot(a, a)); }

// SplitMix64 stream (Steele, Lea, Flood 2014).
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * 0x1p-53; }  // [0, 1)
    double normal()                                                 // Box-Muller
    {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

// ------------------------------------------------------------ cell models --

struct Model {
    int ppc = 0;
    std::vector<float> v;  // 3*ppc
    std::vector<int32_t> sa, sb;
    std::vector<float> sl;
    float contact_radius = 0.f;
    float bound = 0.f;     // bounding radius of the model
};

float spring_rest(const std::vector<float>& v, int a, int b)
{
    float dx = v[3 * a] - v[3 * b], dy = v[3 * a + 1] - v[3 * b + 1], dz = v[3 * a + 2] - v[3 * b + 2];
    return std::sqrt(dx * dx + dy * dy + dz * dz);
}

std::vector<D3> icosahedron_unit()
{
    const double p = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<D3> v = {{-1, p, 0}, {1, p, 0}, {-1, -p, 0}, {1, -p, 0}, {0, -1, p}, {0, 1, p},
                         {0, -1, -p}, {0, 1, -p}, {p, 0, -1}, {p, 0, 1}, {-p, 0, -1}, {-p, 0, 1}};
    for (auto& a : v) a = (1.0 / norm(a)) * a;
    return v;
}

// Edges = vertex pairs at the minimum distance; faces = 3-cliques of edges.
void mesh_edges_faces(const std::vector<D3>& v, std::vector<std::array<int, 2>>& edges,
                      std::vector<std::array<int, 3>>& faces)
{
    const int n = (int)v.size();
    double dmin = 1e300;
    for (int a = 0; a < n; ++a)
        for (int b = a + 1; b < n; ++b) dmin = std::min(dmin, norm(v[a] - v[b]));
    std::vector<std::vector<char>> adj(n, std::vector<char>(n, 0));
    for (int a = 0; a < n; ++a)
        for (int b = a + 1; b < n; ++b)
            if (norm(v[a] - v[b]) < dmin * 1.0001) {
                edges.push_back({a, b});
                adj[a][b] = adj[b][a] = 1;
            }
    for (int a = 0; a < n; ++a)
        for (int b = a + 1; b < n; ++b)
            if (adj[a][b])
                for (int c = b + 1; c < n; ++c)
                    if (adj[a][c] && adj[b][c]) faces.push_back({a, b, c});
}

Model make_icosahedron(float radius, float contact)
{
    auto u = icosahedron_unit();
    std::vector<std::array<int, 2>> edges;
    std::vector<std::array<int, 3>> faces;
    mesh_edges_faces(u, edges, faces);
    Model m;
    m.ppc = 12;
    for (auto& a : u) {
        m.v.push_back((float)(radius * a.x));
        m.v.push_back((float)(radius * a.y));
        m.v.push_back((float)(radius * a.z));
    }
    for (auto& e : edges) {
        m.sa.push_back(e[0]);
        m.sb.push_back(e[1]);
        m.sl.push_back(spring_rest(m.v, e[0], e[1]));
    }
    m.contact_radius = contact;
    m.bound = radius;
    return m;
}

// Icosphere of one subdivision (42 vertices, 80 faces, 120 edges) flattened
// along y, plus 20 cross springs between antipodal vertices (the 20 pairs
// with the largest vertical separation).
Model make_disc42(float radius, float thickness, float contact)
{
    auto u = icosahedron_unit();
    std::vector<std::array<int, 2>> e0;
    std::vector<std::array<int, 3>> f0;
    mesh_edges_faces(u, e0, f0);
    std::vector<D3> v = u;
    std::unordered_map<uint64_t, int> mid;
    auto midpoint = [&](int a, int b) {
        uint64_t key = (uint64_t)std::min(a, b) * 1024u + (uint64_t)std::max(a, b);
        auto it = mid.find(key);
        if (it != mid.end()) return it->second;
        D3 m = 0.5 * (v[a] + v[b]);
        m = (1.0 / norm(m)) * m;
        v.push_back(m);
        mid[key] = (int)v.size() - 1;
        return (int)v.size() - 1;
    };
    std::vector<std::array<int, 3>> faces;
    for (auto& f : f0) {
        int ab = midpoint(f[0], f[1]), bc = midpoint(f[1], f[2]), ca = midpoint(f[2], f[0]);
        faces.push_back({f[0], ab, ca});
        faces.push_back({f[1], bc, ab});
        faces.push_back({f[2], ca, bc});
        faces.push_back({ab, bc, ca});
    }
    std::vector<std::array<int, 2>> edges;
    {
        std::vector<uint64_t> keys;
        for (auto& f : faces)
            for (int k = 0; k < 3; ++k) {
                int a = f[k], b = f[(k + 1) % 3];
                keys.push_back((uint64_t)std::min(a, b) * 1024u + (uint64_t)std::max(a, b));
            }
        std::sort(keys.begin(), keys.end());
        keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
        for (auto k : keys) edges.push_back({(int)(k / 1024u), (int)(k % 1024u)});
    }
    Model m;
    m.ppc = (int)v.size();  // 42
    for (auto& a : v) {
        m.v.push_back((float)(radius * a.x));
        m.v.push_back((float)(radius * thickness * a.y));
        m.v.push_back((float)(radius * a.z));
    }
    for (auto& e : edges) {
        m.sa.push_back(e[0]);
        m.sb.push_back(e[1]);
        m.sl.push_back(spring_rest(m.v, e[0], e[1]));
    }
    // antipodal pairs
    std::vector<std::array<int, 2>> anti;
    for (int a = 0; a < m.ppc; ++a)
        for (int b = a + 1; b < m.ppc; ++b)
            if (norm(v[a] + v[b]) < 1e-9) anti.push_back({a, b});
    std::stable_sort(anti.begin(), anti.end(), [&](const std::array<int, 2>& p, const std::array<int, 2>& q) {
        return std::fabs(v[p[0]].y - v[p[1]].y) > std::fabs(v[q[0]].y - v[q[1]].y) + 1e-12;
    });
    for (int k = 0; k < 20 && k < (int)anti.size(); ++k) {
        m.sa.push_back(anti[k][0]);
        m.sb.push_back(anti[k][1]);
        m.sl.push_back(spring_rest(m.v, anti[k][0], anti[k][1]));
    }
    m.contact_radius = contact;
    m.bound = radius;
    return m;
}

// ------------------------------------------------------------ vein meshes --

struct Mesh {
    std::vector<float> pos;      // 3*nv
    std::vector<uint32_t> idx;   // 3*nt
    std::vector<float> endc, endr;
};

// Straight tube along -y: rings k = 0..rings at y = -k*L/rings, segments m at
// angle 2*pi*m/S; two triangles per quad with the reference's winding
// (first triangles of src/config/vein_definition.hpp:15966 are
// {ring1 m, ring0 m, ring0 m+1}, {ring1 m, ring0 m+1, ring1 m+1}).
Mesh make_cylinder(float R, float L, int S, int rings)
{
    Mesh me;
    for (int k = 0; k <= rings; ++k) {
        double y = -(double)k * L / rings;
        for (int m = 0; m < S; ++m) {
            double th = 2.0 * M_PI * m / S;
            me.pos.push_back((float)(R * std::cos(th)));
            me.pos.push_back((float)y);
            me.pos.push_back((float)(R * std::sin(th)));
        }
    }
    for (int k = 0; k < rings; ++k)
        for (int m = 0; m < S; ++m) {
            uint32_t a = k * S + m, b = k * S + (m + 1) % S;
            uint32_t c = (k + 1) * S + m, d = (k + 1) * S + (m + 1) % S;
            me.idx.insert(me.idx.end(), {c, a, b, c, b, d});
        }
    me.endc = {0.f, -L, 0.f};
    me.endr = {R};
    return me;
}

struct YShape {
    double Rt, Lt, Rb, Lb, ang, k;
    D3 junction, up, d1, d2;
    double ray_dist(D3 p, D3 o, D3 d) const
    {
        double t = std::max(0.0, dot(p - o, d));
        return norm(p - o - t * d);
    }
    static double smin(double a, double b, double k)
    {
        double h = std::max(k - std::fabs(a - b), 0.0) / k;
        return std::min(a, b) - h * h * k * 0.25;
    }
    // < 0 inside the lumen
    double sdf(D3 p) const
    {
        double a = ray_dist(p, junction, up) - Rt;
        double b = ray_dist(p, junction, d1) - Rb;
        double c = ray_dist(p, junction, d2) - Rb;
        return smin(smin(a, b, k), c, k);
    }
};

YShape make_yshape(double Rt, double Lt, double Rb, double Lb, double ang_deg)
{
    YShape y;
    y.Rt = Rt; y.Lt = Lt; y.Rb = Rb; y.Lb = Lb;
    y.ang = ang_deg * M_PI / 180.0;
    y.k = 0.25 * Rb;
    y.junction = {0, -Lt, 0};
    y.up = {0, 1, 0};
    y.d1 = {std::sin(y.ang), -std::cos(y.ang), 0};
    y.d2 = {-std::sin(y.ang), -std::cos(y.ang), 0};
    return y;
}

// Marching tetrahedra (Kuhn decomposition of every voxel into 6 tetrahedra
// sharing the main diagonal) over the lumen SDF.  Tubes are open where the
// sampling box cuts them: the trunk at y = 0, the branches at the bottom
// plane.  Triangles are wound like the reference mesh: cross(v1-v0, v2-v0)
// points out of the lumen.
Mesh make_yvein(const YShape& Y, double h)
{
    const double ybot = -(Y.Lt + Y.Lb * std::cos(Y.ang));
    const double X = Y.Lb * std::sin(Y.ang) + Y.Rb / std::cos(Y.ang) + 2 * h;
    const double Z = std::max(Y.Rt, Y.Rb) + 2 * h;
    const int nx = (int)std::ceil(2 * X / h), ny = (int)std::ceil(-ybot / h), nz = (int)std::ceil(2 * Z / h);
    const double x0 = -X, y0 = ybot, z0 = -Z;
    const int64_t px = nx + 1, pz = nz + 1;
    auto point = [&](int i, int j, int l) { return D3{x0 + i * h, y0 + j * h, z0 + l * h}; };
    auto pid = [&](int i, int j, int l) { return ((int64_t)j * pz + l) * px + i; };
    std::vector<float> lay0((size_t)px * pz), lay1((size_t)px * pz);
    auto fill = [&](std::vector<float>& lay, int j) {
        for (int l = 0; l <= nz; ++l)
            for (int i = 0; i <= nx; ++i) {
                double f = Y.sdf(point(i, j, l));
                if (f == 0.0) f = 1e-7;
                lay[(size_t)l * px + i] = (float)f;
            }
    };
    Mesh me;
    std::unordered_map<uint64_t, uint32_t> vmap;
    vmap.reserve(1 << 20);
    const int64_t npts = px * (int64_t)(ny + 1) * pz;
    static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    fill(lay0, 0);
    for (int j = 0; j < ny; ++j) {
        fill(lay1, j + 1);
        for (int l = 0; l < nz; ++l)
            for (int i = 0; i < nx; ++i) {
                // quick reject: all 8 corners on one side
                float cv[2][2][2];
                bool anyin = false, anyout = false;
                for (int dj = 0; dj < 2; ++dj)
                    for (int dl = 0; dl < 2; ++dl)
                        for (int di = 0; di < 2; ++di) {
                            float f = (dj ? lay1 : lay0)[(size_t)(l + dl) * px + i + di];
                            cv[di][dj][dl] = f;
                            if (f < 0) anyin = true; else anyout = true;
                        }
                if (!anyin || !anyout) continue;
                for (auto& pm : perms) {
                    int c[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {1, 1, 1}};
                    c[1][pm[0]] = 1;
                    c[2][pm[0]] = 1;
                    c[2][pm[1]] = 1;
                    int gi[4], gj[4], gl[4];
                    float fv[4];
                    for (int q = 0; q < 4; ++q) {
                        gi[q] = i + c[q][0]; gj[q] = j + c[q][1]; gl[q] = l + c[q][2];
                        fv[q] = cv[c[q][0]][c[q][1]][c[q][2]];
                    }
                    int in[4], out[4], nin = 0, nout = 0;
                    for (int q = 0; q < 4; ++q) {
                        if (fv[q] < 0) in[nin++] = q; else out[nout++] = q;
                    }
                    if (nin == 0 || nout == 0) continue;
                    auto evert = [&](int a, int b) -> uint32_t {
                        int64_t pa = pid(gi[a], gj[a], gl[a]), pb = pid(gi[b], gj[b], gl[b]);
                        if (pa > pb) { std::swap(pa, pb); std::swap(a, b); }
                        uint64_t key = (uint64_t)pa * (uint64_t)npts + (uint64_t)pb;
                        auto it = vmap.find(key);
                        if (it != vmap.end()) return it->second;
                        double t = fv[a] / (double)(fv[a] - fv[b]);
                        t = std::min(0.9, std::max(0.1, t));
                        D3 A = point(gi[a], gj[a], gl[a]), B = point(gi[b], gj[b], gl[b]);
                        D3 P = A + t * (B - A);
                        uint32_t id = (uint32_t)(me.pos.size() / 3);
                        me.pos.push_back((float)P.x);
                        me.pos.push_back((float)P.y);
                        me.pos.push_back((float)P.z);
                        vmap[key] = id;
                        return id;
                    };
                    D3 cin{0, 0, 0}, cout{0, 0, 0};
                    for (int q = 0; q < nin; ++q) cin = cin + point(gi[in[q]], gj[in[q]], gl[in[q]]);
                    for (int q = 0; q < nout; ++q) cout = cout + point(gi[out[q]], gj[out[q]], gl[out[q]]);
                    D3 outdir = (1.0 / nout) * cout - (1.0 / nin) * cin;
                    auto emit = [&](uint32_t a, uint32_t b, uint32_t c2) {
                        D3 A{me.pos[3 * a], me.pos[3 * a + 1], me.pos[3 * a + 2]};
                        D3 B{me.pos[3 * b], me.pos[3 * b + 1], me.pos[3 * b + 2]};
                        D3 C{me.pos[3 * c2], me.pos[3 * c2 + 1], me.pos[3 * c2 + 2]};
                        if (dot(crossd(B - A, C - A), outdir) < 0) std::swap(b, c2);
                        me.idx.insert(me.idx.end(), {a, b, c2});
                    };
                    if (nin == 1 || nin == 3) {
                        int lone = nin == 1 ? in[0] : out[0];
                        int o[3], k = 0;
                        for (int q = 0; q < 4; ++q)
                            if (q != lone) o[k++] = q;
                        emit(evert(lone, o[0]), evert(lone, o[1]), evert(lone, o[2]));
                    } else {
                        uint32_t ac = evert(in[0], out[0]), ad = evert(in[0], out[1]);
                        uint32_t bd = evert(in[1], out[1]), bc = evert(in[1], out[0]);
                        emit(ac, ad, bd);
                        emit(ac, bd, bc);
                    }
                }
            }
        std::swap(lay0, lay1);
    }
    D3 e1 = Y.junction + Y.Lb * Y.d1, e2 = Y.junction + Y.Lb * Y.d2;
    me.endc = {(float)e1.x, (float)e1.y, (float)e1.z, (float)e2.x, (float)e2.y, (float)e2.z};
    me.endr = {(float)Y.Rb, (float)Y.Rb};
    return me;
}

// ------------------------------------------------------------------ scene --

struct SceneImpl {
    bcs_scene_params params;
    bcs_physics physics;
    std::vector<Model> models;
    std::vector<bcs_cell_type> types;
    Mesh mesh;
    int vein_kind = 0;
    float cyl_R = 0, cyl_L = 0;
    YShape yshape{};
    int np = 0, nc = 0;
    std::vector<float> pos, vel, frc, centers;

    double sdf(D3 p) const
    {
        if (vein_kind == BCS_VEIN_Y) {
            // the sampled surface is open at y = 0 (inlet) and at the bottom plane (outlets)
            const double ybot = -(yshape.Lt + yshape.Lb * std::cos(yshape.ang));
            return std::max(yshape.sdf(p), std::max(p.y, ybot - p.y));
        }
        double r = std::sqrt(p.x * p.x + p.z * p.z);
        double d = r - cyl_R;
        // open ends: inside between y = -L and y = 0
        double ya = p.y - 0.0, yb = -cyl_L - p.y;
        return std::max(d, std::max(ya, yb));
    }
};

}  // namespace

struct bcs_scene {
    SceneImpl s;
};

extern "C" int bcs_scene_default_params(int32_t preset, bcs_scene_params* out)
{
    if (!out) return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_scene_default_params: null output");
    std::memset(out, 0, sizeof(*out));
    bcs_default_physics(&out->physics);
    bcs_physics& ph = out->physics;
    ph.particle_k_sniff = 50.0f;
    ph.particle_d_fact = 0.5f;
    for (int d = 0; d < 3; ++d) {
        ph.particle_cell[d] = 1.0f;
        ph.triangle_cell[d] = 10.0f;
    }
    out->scale = 1.0f;
    out->init_speed = 0.5f;
    out->init_noise_sigma = 0.05f;
    out->seed = 0;
    out->packing_fraction = 0.35f;
    switch (preset) {
    case 0:
        out->vein_kind = BCS_VEIN_CYLINDER;
        out->radius = 40.0f;
        out->length = 400.0f;
        out->segments = 64;
        out->rings = 100;
        out->n_cells = 1;
        out->cells[0] = {BCS_CELL_ICOSAHEDRON, 500, 2.0f, 1.0f, 0.5f};
        out->placement = BCS_PLACE_RANDOM;
        ph.dt = 0.01f;
        break;
    case 1:
    case 2:
    case 4:
        out->vein_kind = BCS_VEIN_Y;
        out->radius = 60.0f;
        out->length = 600.0f;
        out->branch_radius = 42.0f;
        out->branch_length = 400.0f;
        out->branch_angle_deg = 30.0f;
        out->mesh_spacing = 4.3f;
        ph.dt = 0.005f;
        if (preset == 1) {
            out->n_cells = 1;
            out->cells[0] = {BCS_CELL_ICOSAHEDRON, 20000, 2.0f, 1.0f, 0.5f};
            out->placement = BCS_PLACE_RANDOM;
        } else if (preset == 2) {
            out->scale = 3.0f;
            out->n_cells = 2;
            out->cells[0] = {BCS_CELL_DISC42, 140000, 2.0f, 0.35f, 0.5f};
            out->cells[1] = {BCS_CELL_ICOSAHEDRON, 60000, 2.0f, 1.0f, 0.5f};
            out->placement = BCS_PLACE_LATTICE;
        } else {
            out->scale = 3.0f;
            out->n_cells = 1;
            out->cells[0] = {BCS_CELL_DISC42, 1000000, 2.0f, 0.35f, 0.5f};
            out->placement = BCS_PLACE_LATTICE;
        }
        break;
    default:
        return fail(BCS_ERR_INVALID_ARGUMENT, "unknown scene preset");
    }
    return BCS_OK;
}

// Vein `index` of an ensemble of n_veins straight veins: radius uniform in
// [20, 80], length 10 radius, cell count log-uniform in [500, 50000] (70%
// 42-particle discs, 30% 12-particle icosahedra).  Radii and counts are
// drawn independently and paired by rank (the k-th smallest count goes to
// the k-th smallest vein), so that every vein holds its cells at the 35%
// lattice packing; the pairs are then dealt to the veins in a shuffled order.
extern "C" int bcs_scene_ensemble_params(int32_t n_veins, int32_t index, uint64_t seed, double cell_fraction,
                                         bcs_scene_params* out)
{
    if (!out || n_veins < 1 || index < 0 || index >= n_veins || !(cell_fraction > 0.0))
        return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_scene_ensemble_params: bad arguments");
    Rng rng(seed ^ 0xE5E5E5E5ull);
    std::vector<double> r(n_veins), n(n_veins);
    for (int k = 0; k < n_veins; ++k) {
        r[k] = 20.0 + 60.0 * rng.uniform();
        n[k] = std::exp(std::log(500.0) + rng.uniform() * (std::log(50000.0) - std::log(500.0)));
    }
    std::sort(r.begin(), r.end());
    std::sort(n.begin(), n.end());
    std::vector<int> deal(n_veins);
    for (int k = 0; k < n_veins; ++k) deal[k] = k;
    for (int k = n_veins - 1; k > 0; --k) std::swap(deal[k], deal[(int)(rng.uniform() * (k + 1)) % (k + 1)]);
    const int q = deal[index];
    const double R = r[q];
    const int cells = std::max(2, (int)std::lround(n[q] * cell_fraction));
    const int rc = bcs_scene_default_params(0, out);
    if (rc != BCS_OK) return rc;
    out->vein_kind = BCS_VEIN_CYLINDER;
    out->radius = (float)R;
    out->length = (float)(10.0 * R);
    out->segments = std::max(16, (int)std::lround(2.0 * M_PI * R / 4.0));
    out->rings = std::max(10, (int)std::lround(10.0 * R / 4.0));
    out->n_cells = 2;
    const int discs = (int)std::lround(0.7 * cells);
    out->cells[0] = {BCS_CELL_DISC42, discs, 2.0f, 0.35f, 0.5f};
    out->cells[1] = {BCS_CELL_ICOSAHEDRON, cells - discs, 2.0f, 1.0f, 0.5f};
    out->placement = BCS_PLACE_LATTICE;
    out->packing_fraction = 0.35f;
    out->physics.dt = 0.005f;
    out->seed = seed * 1000003ull + (uint64_t)index;
    out->physics.seed = out->seed;
    return BCS_OK;
}

extern "C" int bcs_scene_build(const bcs_scene_params* p, bcs_scene** out)
{
    if (!p || !out) return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_scene_build: null argument");
    if (p->n_cells < 1 || p->n_cells > 4) return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_scene_build: 1..4 cell kinds");
    std::unique_ptr<bcs_scene> sc(new bcs_scene());
    SceneImpl& s = sc->s;
    s.params = *p;
    s.physics = p->physics;
    const double scale = p->scale > 0 ? p->scale : 1.0;
    // vein
    s.vein_kind = p->vein_kind;
    if (p->vein_kind == BCS_VEIN_CYLINDER) {
        if (p->segments < 3 || p->rings < 1) return fail(BCS_ERR_INVALID_ARGUMENT, "cylinder resolution");
        s.cyl_R = (float)(p->radius * scale);
        s.cyl_L = (float)(p->length * scale);
        s.mesh = make_cylinder(s.cyl_R, s.cyl_L, p->segments, p->rings);
    } else if (p->vein_kind == BCS_VEIN_Y) {
        if (!(p->mesh_spacing > 0)) return fail(BCS_ERR_INVALID_ARGUMENT, "Y vein needs mesh_spacing > 0");
        s.yshape = make_yshape(p->radius * scale, p->length * scale, p->branch_radius * scale,
                               p->branch_length * scale, p->branch_angle_deg);
        s.mesh = make_yvein(s.yshape, p->mesh_spacing);
    } else {
        return fail(BCS_ERR_INVALID_ARGUMENT, "unknown vein kind");
    }
    s.physics.cylinder_radius = (float)(p->radius * scale);
    // cell models (type order = params order)
    for (int k = 0; k < p->n_cells; ++k) {
        const bcs_scene_cell& c = p->cells[k];
        if (c.count < 0) return fail(BCS_ERR_INVALID_ARGUMENT, "negative cell count");
        if (c.kind == BCS_CELL_ICOSAHEDRON) s.models.push_back(make_icosahedron(c.radius, c.contact_radius));
        else if (c.kind == BCS_CELL_DISC42)
            s.models.push_back(make_disc42(c.radius, c.thickness > 0 ? c.thickness : 0.35f, c.contact_radius));
        else return fail(BCS_ERR_INVALID_ARGUMENT, "unknown cell kind");
    }
    for (int k = 0; k < p->n_cells; ++k) {
        const Model& m = s.models[k];
        bcs_cell_type t;
        std::memset(&t, 0, sizeof(t));
        t.count = p->cells[k].count;
        t.particles_in_cell = m.ppc;
        t.spring_count = (int32_t)m.sa.size();
        t.spring_start = m.sa.data();
        t.spring_end = m.sb.data();
        t.spring_length = m.sl.data();
        t.vertices = m.v.data();
        t.contact_radius = m.contact_radius;
        s.types.push_back(t);
        s.np += t.count * m.ppc;
        s.nc += t.count;
    }
    // placement: one centroid + velocity per cell (type-major storage)
    Rng rng(p->seed * 0x2545F4914F6CDD1Dull + 0x1234567ull);
    double bmax = 0;
    for (auto& m : s.models) bmax = std::max(bmax, (double)m.bound);
    const double gap = 0.3;
    const double b = bmax + gap;  // bounding radius used for clearance
    std::vector<D3> cent;
    std::vector<int> kind_of;  // kind index per placed cell (placement order)
    cent.reserve(s.nc);
    // lumen sampling box
    D3 lo, hi;
    {
        float mn[3] = {1e30f, 1e30f, 1e30f}, mx[3] = {-1e30f, -1e30f, -1e30f};
        for (size_t i = 0; i < s.mesh.pos.size(); i += 3)
            for (int d = 0; d < 3; ++d) {
                mn[d] = std::min(mn[d], s.mesh.pos[i + d]);
                mx[d] = std::max(mx[d], s.mesh.pos[i + d]);
            }
        lo = {mn[0], mn[1], mn[2]};
        hi = {mx[0], mx[1], mx[2]};
    }
    std::vector<int> labels;
    for (int k = 0; k < p->n_cells; ++k) labels.insert(labels.end(), p->cells[k].count, k);
    for (size_t i = labels.size(); i > 1; --i) std::swap(labels[i - 1], labels[rng.next() % i]);
    if (p->placement == BCS_PLACE_RANDOM) {
        const double cs = 2 * b;
        std::unordered_map<uint64_t, std::vector<int>> hash;
        auto hkey = [&](D3 c) {
            int64_t ix = (int64_t)std::floor((c.x - lo.x) / cs), iy = (int64_t)std::floor((c.y - lo.y) / cs),
                    iz = (int64_t)std::floor((c.z - lo.z) / cs);
            return ((uint64_t)(ix & 0x1FFFFF) << 42) | ((uint64_t)(iy & 0x1FFFFF) << 21) | (uint64_t)(iz & 0x1FFFFF);
        };
        int64_t attempts = 0, max_attempts = (int64_t)s.nc * 200 + 1000;
        while ((int)cent.size() < s.nc) {
            if (++attempts > max_attempts)
                return fail(BCS_ERR_INVALID_ARGUMENT, "random placement: lumen too small for the cell count");
            D3 c{lo.x + rng.uniform() * (hi.x - lo.x), lo.y + rng.uniform() * (hi.y - lo.y),
                 lo.z + rng.uniform() * (hi.z - lo.z)};
            if (s.sdf(c) > -(b + 1.0)) continue;
            bool ok = true;
            int64_t ix = (int64_t)std::floor((c.x - lo.x) / cs), iy = (int64_t)std::floor((c.y - lo.y) / cs),
                    iz = (int64_t)std::floor((c.z - lo.z) / cs);
            for (int dx = -1; dx <= 1 && ok; ++dx)
                for (int dy = -1; dy <= 1 && ok; ++dy)
                    for (int dz = -1; dz <= 1 && ok; ++dz) {
                        uint64_t k = ((uint64_t)((ix + dx) & 0x1FFFFF) << 42) | ((uint64_t)((iy + dy) & 0x1FFFFF) << 21) |
                                     (uint64_t)((iz + dz) & 0x1FFFFF);
                        auto it = hash.find(k);
                        if (it == hash.end()) continue;
                        for (int q : it->second)
                            if (norm(cent[q] - c) < 2 * b) { ok = false; break; }
                    }
            if (!ok) continue;
            hash[hkey(c)].push_back((int)cent.size());
            cent.push_back(c);
        }
    } else {
        // cubic lattice of spacing 2b: bounding-sphere volume fraction 0.5236;
        // sites are kept with probability packing/0.5236, filled from the
        // inlet (largest y) downwards, jittered inside their cubes.
        const double a = 2 * b;
        const double keep = std::min(1.0, (double)p->packing_fraction / (M_PI / 6.0));
        const double jit = gap * 0.5;
        std::vector<D3> sites;
        int64_t nx = (int64_t)std::floor((hi.x - lo.x) / a), ny = (int64_t)std::floor((hi.y - lo.y) / a),
                nz = (int64_t)std::floor((hi.z - lo.z) / a);
        for (int64_t j = ny - 1; j >= 0 && (int64_t)sites.size() < (int64_t)s.nc; --j)
            for (int64_t l = 0; l < nz; ++l)
                for (int64_t i = 0; i < nx; ++i) {
                    D3 c{lo.x + (i + 0.5) * a, lo.y + (j + 0.5) * a, lo.z + (l + 0.5) * a};
                    if (s.sdf(c) > -(b + 1.0)) continue;
                    if (rng.uniform() >= keep) continue;
                    c = c + D3{(rng.uniform() * 2 - 1) * jit, (rng.uniform() * 2 - 1) * jit, (rng.uniform() * 2 - 1) * jit};
                    sites.push_back(c);
                }
        if ((int64_t)sites.size() < (int64_t)s.nc)
            return fail(BCS_ERR_INVALID_ARGUMENT, "lattice placement: lumen too small for the cell count");
        sites.resize(s.nc);
        cent = sites;
    }
    // assign placements to types: the i-th cell of kind k takes the i-th site labelled k
    std::vector<std::vector<int>> sites_of(p->n_cells);
    for (int i = 0; i < s.nc; ++i) sites_of[labels[i]].push_back(i);
    s.pos.assign((size_t)3 * s.np, 0.f);
    s.vel.assign((size_t)3 * s.np, 0.f);
    s.frc.assign((size_t)3 * s.np, 0.f);
    s.centers.assign((size_t)3 * s.nc, 0.f);
    const size_t N = (size_t)s.np, NC = (size_t)s.nc;
    int pi = 0, ci = 0;
    for (int k = 0; k < p->n_cells; ++k) {
        const Model& m = s.models[k];
        for (int c = 0; c < p->cells[k].count; ++c, ++ci) {
            D3 C = cent[sites_of[k][c]];
            float cx = (float)C.x, cy = (float)C.y, cz = (float)C.z;
            float vx = (float)(p->init_noise_sigma * rng.normal());
            float vy = (float)(-p->init_speed + p->init_noise_sigma * rng.normal());
            float vz = (float)(p->init_noise_sigma * rng.normal());
            s.centers[ci] = cx;
            s.centers[NC + ci] = cy;
            s.centers[2 * NC + ci] = cz;
            for (int a = 0; a < m.ppc; ++a, ++pi) {
                // setBloodCellsPositionsFromRandom: centroid + model vertex
                s.pos[pi] = cx + m.v[3 * a];
                s.pos[N + pi] = cy + m.v[3 * a + 1];
                s.pos[2 * N + pi] = cz + m.v[3 * a + 2];
                s.vel[pi] = vx;
                s.vel[N + pi] = vy;
                s.vel[2 * N + pi] = vz;
            }
        }
    }
    *out = sc.release();
    return BCS_OK;
}

extern "C" int bcs_scene_preset(int32_t preset, double cell_fraction, uint64_t seed, bcs_scene** out)
{
    bcs_scene_params p;
    int rc = bcs_scene_default_params(preset, &p);
    if (rc) return rc;
    if (cell_fraction > 0 && cell_fraction != 1.0)
        for (int k = 0; k < p.n_cells; ++k) {
            int c = (int)std::llround(p.cells[k].count * cell_fraction);
            p.cells[k].count = std::max(c, 1);
        }
    p.seed = seed;
    return bcs_scene_build(&p, out);
}

extern "C" int bcs_scene_get_view(bcs_scene* scene, bcs_scene_view* out)
{
    if (!scene || !out) return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_scene_get_view: null argument");
    SceneImpl& s = scene->s;
    std::memset(out, 0, sizeof(*out));
    out->physics = s.physics;
    out->n_types = (int32_t)s.types.size();
    out->types = s.types.data();
    out->mesh.vertex_count = (int32_t)(s.mesh.pos.size() / 3);
    out->mesh.positions = s.mesh.pos.data();
    out->mesh.triangle_count = (int32_t)(s.mesh.idx.size() / 3);
    out->mesh.indices = s.mesh.idx.data();
    out->mesh.ending_count = (int32_t)s.mesh.endr.size();
    out->mesh.ending_centers = s.mesh.endc.data();
    out->mesh.ending_radii = s.mesh.endr.data();
    out->particle_count = s.np;
    out->positions = s.pos.data();
    out->velocities = s.vel.data();
    out->forces = s.frc.data();
    out->cell_centers = s.centers.data();
    out->cell_count = s.nc;
    return BCS_OK;
}

extern "C" int bcs_scene_destroy(bcs_scene* scene)
{
    delete scene;
    return BCS_OK;
}

extern "C" int bcs_scene_lumen_sdf(bcs_scene* scene, const float* points, int32_t n, float* out)
{
    if (!scene || (!points && n > 0) || (!out && n > 0)) return fail(BCS_ERR_INVALID_ARGUMENT, "bcs_scene_lumen_sdf");
    for (int32_t i = 0; i < n; ++i)
        out[i] = (float)scene->s.sdf(D3{points[3 * i], points[3 * i + 1], points[3 * i + 2]});
    return BCS_OK;
}
