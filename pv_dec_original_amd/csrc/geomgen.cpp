// geomgen.cpp — synthetic vessel meshes (host C++).
//
// Parametric tube, and marching-tetrahedra contours of implicit vessel fields: Y-bifurcation,
// seeded tortuous vessels and a seeded aneurysm.  The contour is watertight: every cube is cut
// into the same six tetrahedra around its main diagonal, so neighbouring cubes share face
// diagonals, and each edge vertex is interpolated in a fixed (low -> high grid id) direction so
// both tetrahedra that share an edge produce the same bits.  Triangles are oriented outward.

#include "../../include/pvgen.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

struct pvgen_mesh {
    std::vector<double> tri;
};

namespace {

constexpr double kPi = 3.14159265358979323846;

void push_tri(std::vector<double>& out, const double a[3], const double b[3], const double c[3]) {
    double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double n[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
    double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (len > 0) {
        n[0] /= len;
        n[1] /= len;
        n[2] /= len;
    }
    const double* p[3] = {a, b, c};
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) out.push_back(p[i][k]);
    for (int k = 0; k < 3; k++) out.push_back(n[k]);
}

// splitmix64
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uni(double a, double b) { return a + (b - a) * ((next() >> 11) * (1.0 / 9007199254740992.0)); }
    int irange(int a, int b) { return a + (int)(next() % (uint64_t)(b - a + 1)); }
};

struct Seg {
    double a[3], b[3], ra, rb;
    // optional flat end: the segment keeps p[end_axis] >= end_val (end_keep > 0) or <= (< 0)
    int end_axis = -1;
    double end_val = 0.0;
    int end_keep = 1;
};

struct Grid {
    double lo[3], sp;
    int64_t g[3];
    std::vector<double> phi;
    double pos(int axis, int64_t i) const { return lo[axis] + (double)i * sp; }
    int64_t id(int64_t i, int64_t j, int64_t k) const { return i + g[0] * (j + g[1] * k); }
};

Grid make_grid(const double lo[3], const double hi[3], double sp) {
    Grid G;
    G.sp = sp;
    for (int a = 0; a < 3; a++) {
        G.lo[a] = lo[a];
        G.g[a] = (int64_t)std::floor((hi[a] - lo[a]) / sp) + 1;
    }
    G.phi.assign((size_t)(G.g[0] * G.g[1] * G.g[2]), 1e30);
    return G;
}

double seg_field(const double p[3], const Seg& s) {
    double ab[3] = {s.b[0] - s.a[0], s.b[1] - s.a[1], s.b[2] - s.a[2]};
    double ap[3] = {p[0] - s.a[0], p[1] - s.a[1], p[2] - s.a[2]};
    double ll = ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2];
    double t = ll > 0 ? (ap[0] * ab[0] + ap[1] * ab[1] + ap[2] * ab[2]) / ll : 0.0;
    t = std::min(1.0, std::max(0.0, t));
    double d[3] = {ap[0] - t * ab[0], ap[1] - t * ab[1], ap[2] - t * ab[2]};
    double v = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) - (s.ra + t * (s.rb - s.ra));
    if (s.end_axis >= 0) {
        const double cut = s.end_keep > 0 ? s.end_val - p[s.end_axis] : p[s.end_axis] - s.end_val;
        v = std::max(v, cut);
    }
    return v;
}

// phi = min(phi, field of each segment), evaluated only inside each segment's padded bbox
void splat_segments(Grid& G, const std::vector<Seg>& segs) {
    const double pad = 2.0 * G.sp + 1.0;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t k = 0; k < G.g[2]; k++) {
        for (const Seg& s : segs) {
            double r = std::max(s.ra, s.rb) + pad;
            double z0 = std::min(s.a[2], s.b[2]) - r, z1 = std::max(s.a[2], s.b[2]) + r;
            double z = G.pos(2, k);
            if (z < z0 || z > z1) continue;
            int64_t i0 = std::max<int64_t>(0, (int64_t)std::floor((std::min(s.a[0], s.b[0]) - r - G.lo[0]) / G.sp));
            int64_t i1 = std::min<int64_t>(G.g[0] - 1, (int64_t)std::ceil((std::max(s.a[0], s.b[0]) + r - G.lo[0]) / G.sp));
            int64_t j0 = std::max<int64_t>(0, (int64_t)std::floor((std::min(s.a[1], s.b[1]) - r - G.lo[1]) / G.sp));
            int64_t j1 = std::min<int64_t>(G.g[1] - 1, (int64_t)std::ceil((std::max(s.a[1], s.b[1]) + r - G.lo[1]) / G.sp));
            for (int64_t j = j0; j <= j1; j++)
                for (int64_t i = i0; i <= i1; i++) {
                    double p[3] = {G.pos(0, i), G.pos(1, j), z};
                    double v = seg_field(p, s);
                    double& dst = G.phi[G.id(i, j, k)];
                    if (v < dst) dst = v;
                }
        }
    }
}

void cut_slab(Grid& G, double x0, double x1) {
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < G.g[2]; k++)
        for (int64_t j = 0; j < G.g[1]; j++)
            for (int64_t i = 0; i < G.g[0]; i++) {
                double x = G.pos(0, i);
                double& v = G.phi[G.id(i, j, k)];
                v = std::max(v, std::max(x0 - x, x - x1));
            }
}

// keep the solid inside the box [blo, bhi] on every axis
void cut_box(Grid& G, const double blo[3], const double bhi[3]) {
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < G.g[2]; k++)
        for (int64_t j = 0; j < G.g[1]; j++)
            for (int64_t i = 0; i < G.g[0]; i++) {
                const double p[3] = {G.pos(0, i), G.pos(1, j), G.pos(2, k)};
                double& v = G.phi[G.id(i, j, k)];
                for (int a = 0; a < 3; a++) v = std::max(v, std::max(blo[a] - p[a], p[a] - bhi[a]));
            }
}

void add_sphere(Grid& G, const double c[3], double r) {
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < G.g[2]; k++)
        for (int64_t j = 0; j < G.g[1]; j++)
            for (int64_t i = 0; i < G.g[0]; i++) {
                double p[3] = {G.pos(0, i) - c[0], G.pos(1, j) - c[1], G.pos(2, k) - c[2]};
                double v = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) - r;
                double& dst = G.phi[G.id(i, j, k)];
                if (v < dst) dst = v;
            }
}

// Marching tetrahedra over the grid.
void contour(const Grid& G, std::vector<double>& out) {
    // Kuhn split: tet = (0, e_a, e_a + e_b, 7) for each axis order (a, b, c); corner bits x=1,y=2,z=4
    static const int tets[6][4] = {{0, 1, 3, 7}, {0, 1, 5, 7}, {0, 2, 3, 7},
                                   {0, 2, 6, 7}, {0, 4, 5, 7}, {0, 4, 6, 7}};
    const int64_t gx = G.g[0], gy = G.g[1], gz = G.g[2];
    int nth = 1;
#ifdef _OPENMP
    nth = omp_get_max_threads();
#endif
    std::vector<std::vector<double>> parts((size_t)std::max<int64_t>(1, gz - 1));
#pragma omp parallel for schedule(dynamic, 1) num_threads(nth)
    for (int64_t k = 0; k < gz - 1; k++) {
        std::vector<double>& o = parts[k];
        for (int64_t j = 0; j < gy - 1; j++)
            for (int64_t i = 0; i < gx - 1; i++) {
                int64_t gid[8];
                double ph[8], P[8][3];
                int nin = 0;
                for (int c = 0; c < 8; c++) {
                    int64_t a = i + (c & 1), b = j + ((c >> 1) & 1), d = k + ((c >> 2) & 1);
                    gid[c] = G.id(a, b, d);
                    double v = G.phi[gid[c]];
                    if (v == 0.0) v = 1e-12;
                    ph[c] = v;
                    P[c][0] = G.pos(0, a);
                    P[c][1] = G.pos(1, b);
                    P[c][2] = G.pos(2, d);
                    nin += v < 0;
                }
                if (nin == 0 || nin == 8) continue;
                for (int t = 0; t < 6; t++) {
                    const int* T = tets[t];
                    int ins[4], outs[4], ni = 0, no = 0;
                    for (int v = 0; v < 4; v++) {
                        if (ph[T[v]] < 0) ins[ni++] = T[v];
                        else outs[no++] = T[v];
                    }
                    if (ni == 0 || ni == 4) continue;
                    auto edge = [&](int a, int b, double p[3]) {
                        if (gid[a] > gid[b]) std::swap(a, b);
                        double s = ph[a] / (ph[a] - ph[b]);
                        for (int m = 0; m < 3; m++) p[m] = P[a][m] + s * (P[b][m] - P[a][m]);
                    };
                    // outward = from the inside vertices' mean towards the outside vertices' mean
                    double cin[3] = {0, 0, 0}, cout_[3] = {0, 0, 0};
                    for (int v = 0; v < ni; v++)
                        for (int m = 0; m < 3; m++) cin[m] += P[ins[v]][m] / ni;
                    for (int v = 0; v < no; v++)
                        for (int m = 0; m < 3; m++) cout_[m] += P[outs[v]][m] / no;
                    double dir[3] = {cout_[0] - cin[0], cout_[1] - cin[1], cout_[2] - cin[2]};
                    auto emit = [&](const double* a, const double* b, const double* c) {
                        double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
                        double v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
                        double n[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2],
                                       u[0] * v[1] - u[1] * v[0]};
                        if (n[0] * dir[0] + n[1] * dir[1] + n[2] * dir[2] < 0) push_tri(o, a, c, b);
                        else push_tri(o, a, b, c);
                    };
                    if (ni == 1 || ni == 3) {
                        int lone = (ni == 1) ? ins[0] : outs[0];
                        const int* others = (ni == 1) ? outs : ins;
                        double p0[3], p1[3], p2[3];
                        edge(lone, others[0], p0);
                        edge(lone, others[1], p1);
                        edge(lone, others[2], p2);
                        emit(p0, p1, p2);
                    } else {
                        double ac[3], ad[3], bd[3], bc[3];
                        edge(ins[0], outs[0], ac);
                        edge(ins[0], outs[1], ad);
                        edge(ins[1], outs[1], bd);
                        edge(ins[1], outs[0], bc);
                        emit(ac, ad, bd);
                        emit(ac, bd, bc);
                    }
                }
            }
    }
    for (auto& p : parts) out.insert(out.end(), p.begin(), p.end());
}

// Catmull-Rom spline through control points (uniform parameter), sampled into segments
void spline_segments(const std::vector<std::array<double, 4>>& cp, int samples_per_span,
                     std::vector<Seg>& segs) {
    auto at = [&](int i) { return cp[(size_t)std::min<int>((int)cp.size() - 1, std::max(0, i))]; };
    std::vector<std::array<double, 4>> pts;
    for (int s = 0; s + 1 < (int)cp.size(); s++) {
        auto p0 = at(s - 1), p1 = at(s), p2 = at(s + 1), p3 = at(s + 2);
        for (int m = 0; m < samples_per_span; m++) {
            double t = (double)m / samples_per_span, t2 = t * t, t3 = t2 * t;
            std::array<double, 4> q;
            for (int c = 0; c < 4; c++)
                q[c] = 0.5 * ((2 * p1[c]) + (-p0[c] + p2[c]) * t + (2 * p0[c] - 5 * p1[c] + 4 * p2[c] - p3[c]) * t2 +
                              (-p0[c] + 3 * p1[c] - 3 * p2[c] + p3[c]) * t3);
            pts.push_back(q);
        }
    }
    pts.push_back(cp.back());
    for (size_t i = 0; i + 1 < pts.size(); i++) {
        Seg s;
        for (int c = 0; c < 3; c++) {
            s.a[c] = pts[i][c];
            s.b[c] = pts[i + 1][c];
        }
        s.ra = pts[i][3];
        s.rb = pts[i + 1][3];
        segs.push_back(s);
    }
}

void ybif_segments(double rp, double rd, double angle_deg, double xb, double y0, double z0,
                   double xcap0, double xcap1, std::vector<Seg>& segs) {
    Seg p = {{xcap0 - 20.0, y0, z0}, {xb, y0, z0}, rp, rp};
    segs.push_back(p);
    double th = angle_deg * kPi / 180.0;
    double L = (xcap1 + 20.0 - xb) / std::cos(th);
    for (int sgn = -1; sgn <= 1; sgn += 2) {
        Seg d = {{xb, y0, z0}, {xb + L * std::cos(th), y0 + sgn * L * std::sin(th), z0}, rd, rd};
        segs.push_back(d);
    }
}

int finish(Grid& G, double xcap0, double xcap1, pvgen_mesh** out) {
    cut_slab(G, xcap0, xcap1);
    pvgen_mesh* m = new pvgen_mesh();
    contour(G, m->tri);
    *out = m;
    return 0;
}

}  // namespace

extern "C" int pvgen_tube(double R, double L, int nseg, pvgen_mesh** out) {
    if (!out || nseg < 3 || R <= 0 || L <= 0) return 1;
    pvgen_mesh* m = new pvgen_mesh();
    std::vector<std::array<double, 2>> ring((size_t)nseg);
    for (int k = 0; k < nseg; k++) {
        double th = (k + 0.5) * 2.0 * kPi / nseg;
        ring[k] = {R * std::cos(th), R * std::sin(th)};
    }
    for (int k = 0; k < nseg; k++) {
        int k1 = (k + 1) % nseg;
        double a[3] = {0.0, ring[k][0], ring[k][1]}, b[3] = {L, ring[k][0], ring[k][1]};
        double c[3] = {L, ring[k1][0], ring[k1][1]}, d[3] = {0.0, ring[k1][0], ring[k1][1]};
        // side quad a-b-c-d; (a, c, b) and (a, d, c) wind with the normal pointing away from the axis
        push_tri(m->tri, a, c, b);
        push_tri(m->tri, a, d, c);
    }
    double c0[3] = {0.0, 0.0, 0.0}, c1[3] = {L, 0.0, 0.0};
    for (int k = 0; k < nseg; k++) {
        int k1 = (k + 1) % nseg;
        double a0[3] = {0.0, ring[k][0], ring[k][1]}, b0[3] = {0.0, ring[k1][0], ring[k1][1]};
        double a1[3] = {L, ring[k][0], ring[k][1]}, b1[3] = {L, ring[k1][0], ring[k1][1]};
        push_tri(m->tri, b0, a0, c0);  // x = 0 cap, normal -x, centre as vertex 3
        push_tri(m->tri, a1, b1, c1);  // x = L cap, normal +x, centre as vertex 3
    }
    *out = m;
    return 0;
}

extern "C" int pvgen_ybifurcation(double rp, double rd, double angle_deg, double xb, double y0,
                                  double z0, double xcap0, double xcap1, const double lo[3],
                                  const double hi[3], double spacing, pvgen_mesh** out) {
    if (!out || spacing <= 0) return 1;
    Grid G = make_grid(lo, hi, spacing);
    std::vector<Seg> segs;
    ybif_segments(rp, rd, angle_deg, xb, y0, z0, xcap0, xcap1, segs);
    splat_segments(G, segs);
    return finish(G, xcap0, xcap1, out);
}

extern "C" int pvgen_vessel(uint64_t seed, double n, double rmin, double rmax, int nbif,
                            double xcap0, double xcap1, const double lo[3], const double hi[3],
                            double spacing, pvgen_mesh** out) {
    if (!out || spacing <= 0 || rmax < rmin) return 1;
    Rng rng(seed * 0x2545F4914F6CDD1Dull + 12345);
    const double m = rmax + 3.0;  // keep every tube inside the box in y and z
    // Inlet and outlets are straight legs along x (flow extensions) crossing the planes x = 0 and
    // x = n: the Zou-He closures of the x faces take the wall to run along x at the boundary
    // plane, and an oblique vessel end leaves rim nodes whose closure is unstable.
    const double x_in = 0.12 * n, x_out = 0.88 * n, leg0 = -0.2 * n, leg1 = 1.2 * n;
    std::vector<Seg> segs;
    int ncp = rng.irange(4, 6);
    std::vector<std::array<double, 4>> main;
    for (int i = 0; i < ncp; i++) {
        double x = x_in + (x_out - x_in) * i / (ncp - 1);
        main.push_back({x, rng.uni(m, n - m), rng.uni(m, n - m), rng.uni(rmin, rmax)});
    }
    auto straight = [&](const std::array<double, 4>& p, double x) {
        Seg s = {{p[0], p[1], p[2]}, {x, p[1], p[2]}, p[3], p[3]};
        segs.push_back(s);
    };
    straight(main.front(), leg0);
    straight(main.back(), leg1);
    size_t first = segs.size();
    spline_segments(main, 24, segs);
    size_t nmain = segs.size() - first;
    std::vector<std::array<double, 3>> exits = {{main.back()[1], main.back()[2], main.back()[3]}};
    for (int b = 0; b < std::max(1, std::min(2, nbif)); b++) {
        // branch from a point of the main line to an exit of its own at x = n
        size_t at = first + (size_t)(nmain * rng.uni(0.2, 0.55));
        const Seg& s0 = segs[std::min(at, segs.size() - 1)];
        double r0 = std::max(rmin, 0.75 * s0.ra);
        std::vector<std::array<double, 4>> br;
        br.push_back({s0.a[0], s0.a[1], s0.a[2], r0});
        int nb = rng.irange(2, 3);
        for (int i = 1; i <= nb; i++) {
            double x = s0.a[0] + (x_out - s0.a[0]) * i / nb;
            br.push_back({x, rng.uni(m, n - m), rng.uni(m, n - m), rng.uni(rmin, std::max(rmin, r0))});
        }
        // keep the exit apart from the others where the box allows it (a bounded number of draws)
        for (int tries = 0; tries < 32; tries++) {
            bool apart = true;
            for (const auto& e : exits) {
                double dy = br.back()[1] - e[0], dz = br.back()[2] - e[1];
                apart = apart && std::sqrt(dy * dy + dz * dz) > br.back()[3] + e[2] + 4.0;
            }
            if (apart) break;
            br.back()[1] = rng.uni(m, n - m);
            br.back()[2] = rng.uni(m, n - m);
        }
        exits.push_back({br.back()[1], br.back()[2], br.back()[3]});
        straight(br.back(), leg1);
        spline_segments(br, 24, segs);
    }
    Grid G = make_grid(lo, hi, spacing);
    splat_segments(G, segs);
    return finish(G, xcap0, xcap1, out);
}

extern "C" int pvgen_aneurysm(uint64_t seed, double nx, double ny, double nz, double rt,
                              double rs, double xcap0, double xcap1, const double lo[3],
                              const double hi[3], double spacing, pvgen_mesh** out) {
    if (!out || spacing <= 0) return 1;
    Rng rng(seed * 0x9E3779B97F4A7C15ull + 777);
    double amp = rng.uni(0.4, 0.8) * (0.5 * ny - rt - 8.0) * 0.5;
    double phase = rng.uni(0.0, 2.0 * kPi);
    double cy = 0.5 * ny, cz = 0.5 * nz - 0.5 * rs;
    std::vector<std::array<double, 4>> cp;
    for (int i = 0; i <= 8; i++) {
        double x = xcap0 - 20.0 + (xcap1 - xcap0 + 40.0) * i / 8.0;
        cp.push_back({x, cy + amp * std::sin(kPi * x / nx + phase), cz, rt});
    }
    std::vector<Seg> segs;
    spline_segments(cp, 16, segs);
    Grid G = make_grid(lo, hi, spacing);
    splat_segments(G, segs);
    // saccular bulge: sphere centred beside the tube near mid-length, offset in +z
    double xs = nx * rng.uni(0.42, 0.58);
    double ys = cy + amp * std::sin(kPi * xs / nx + phase);
    double c[3] = {xs, ys, cz + rt + 0.45 * rs};
    add_sphere(G, c, rs);
    return finish(G, xcap0, xcap1, out);
}

extern "C" int pvgen_tree(int nseg, const double* seg, const double box_lo[3], const double box_hi[3],
                          const double lo[3], const double hi[3], double spacing, pvgen_mesh** out) {
    if (!out || spacing <= 0 || nseg <= 0 || !seg) return 1;
    std::vector<Seg> segs((size_t)nseg);
    for (int i = 0; i < nseg; i++) {
        const double* v = seg + 11 * i;
        Seg& S = segs[(size_t)i];
        for (int a = 0; a < 3; a++) {
            S.a[a] = v[a];
            S.b[a] = v[3 + a];
        }
        S.ra = v[6];
        S.rb = v[7];
        S.end_axis = (int)v[8];
        S.end_val = v[9];
        S.end_keep = v[10] < 0 ? -1 : 1;
        if (S.end_axis > 2 || S.ra <= 0 || S.rb <= 0) return 1;
    }
    Grid G = make_grid(lo, hi, spacing);
    splat_segments(G, segs);
    cut_box(G, box_lo, box_hi);
    pvgen_mesh* m = new pvgen_mesh();
    contour(G, m->tri);
    *out = m;
    return 0;
}

extern "C" int64_t pvgen_mesh_size(const pvgen_mesh* m) { return m ? (int64_t)(m->tri.size() / 12) : 0; }

extern "C" void pvgen_mesh_copy(const pvgen_mesh* m, double* tri) {
    if (m && tri) std::memcpy(tri, m->tri.data(), m->tri.size() * sizeof(double));
}

extern "C" void pvgen_mesh_free(pvgen_mesh* m) { delete m; }

extern "C" int pvgen_write_stl_ascii(const char* path, const double* tri, int64_t ntri,
                                     const char* name) {
    FILE* fp = std::fopen(path, "w");
    if (!fp) return 1;
    const char* nm = name ? name : "pvgen";
    std::fprintf(fp, "solid %s\n", nm);
    for (int64_t t = 0; t < ntri; t++) {
        const double* T = tri + 12 * t;
        std::fprintf(fp, "facet normal %.17g %.17g %.17g\n", T[9], T[10], T[11]);
        std::fprintf(fp, "outer loop\n");
        for (int v = 0; v < 3; v++)
            std::fprintf(fp, "vertex %.17g %.17g %.17g\n", T[3 * v], T[3 * v + 1], T[3 * v + 2]);
        std::fprintf(fp, "endloop\nendfacet\n");
    }
    std::fprintf(fp, "endsolid %s\n", nm);
    return std::fclose(fp) == 0 ? 0 : 1;
}

extern "C" int pvgen_write_stl_binary(const char* path, const double* tri, int64_t ntri, const char* name) {
    if (ntri < 0 || ntri > (int64_t)UINT32_MAX) return 1;
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return 1;
    char header[80] = {0};
    std::snprintf(header, sizeof header, "%s", name ? name : "pvgen");
    const uint32_t n = (uint32_t)ntri;
    bool ok = std::fwrite(header, 1, 80, fp) == 80 && std::fwrite(&n, 4, 1, fp) == 1;
    for (int64_t t = 0; ok && t < ntri; t++) {
        const double* T = tri + 12 * t;
        float f[12] = {(float)T[9], (float)T[10], (float)T[11]};
        for (int v = 0; v < 9; v++) f[3 + v] = (float)T[v];
        const uint16_t attr = 0;
        ok = std::fwrite(f, 4, 12, fp) == 12 && std::fwrite(&attr, 2, 1, fp) == 1;
    }
    return (std::fclose(fp) == 0 && ok) ? 0 : 1;
}

extern "C" int pvgen_field(int shape, const double* params, int64_t npts, const double* xyz,
                           double* phi) {
    if (shape != 1) return 1;
    std::vector<Seg> segs;
    ybif_segments(params[0], params[1], params[2], params[3], params[4], params[5], params[6],
                  params[7], segs);
    for (int64_t p = 0; p < npts; p++) {
        double v = 1e30;
        for (const Seg& s : segs) v = std::min(v, seg_field(xyz + 3 * p, s));
        double x = xyz[3 * p];
        v = std::max(v, std::max(params[6] - x, x - params[7]));
        phi[p] = v;
    }
    return 0;
}
