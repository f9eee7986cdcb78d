// pvoracle.cpp — TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of the reference solver (Resacd/PV_Dec_Original) that the tests, the smoke
// check and the CPU-baseline leg of bench.py use as the checker.  The product path (the HIP
// library) never links, loads or calls this file.  Each function cites the reference lines it
// follows and keeps their floating-point operation order; the Makefile compiles it with
// -O2 -ffp-contract=off so that the rounding is strict IEEE in that order.
//
// Layout choices are the oracle's own (flat arrays, node ids, neighbour table), not the
// reference's heap object graph; they do not change any computed value.

#include "pvoracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------------------------
// Lattice tables.  D3Q19: LatticeModel.cpp:53-92 (weights, velocities, rest = 18), opposites
// :108-151.  D3Q27: LatticeModel.cpp:210-266 (rest = 0), opposites :274-334.  The index order
// is the reference's and every closure formula below depends on it.
// ---------------------------------------------------------------------------------------------
const int C19[19][3] = {{1, 0, 0},  {-1, 0, 0}, {0, 1, 0},   {0, -1, 0}, {0, 0, 1},
                        {0, 0, -1}, {1, 1, 0},  {1, -1, 0},  {1, 0, 1},  {1, 0, -1},
                        {-1, 1, 0}, {-1, -1, 0}, {-1, 0, 1}, {-1, 0, -1}, {0, 1, 1},
                        {0, 1, -1}, {0, -1, 1}, {0, -1, -1}, {0, 0, 0}};
const double W19[19] = {2.0 / 36.0, 2.0 / 36.0, 2.0 / 36.0, 2.0 / 36.0, 2.0 / 36.0,
                        2.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                        1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                        1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 12.0 / 36.0};
const int OPP19[19] = {1, 0, 3, 2, 5, 4, 11, 10, 13, 12, 7, 6, 9, 8, 17, 16, 15, 14, 18};

const int C27[27][3] = {{0, 0, 0},   {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},
                        {0, 0, 1},   {0, 0, -1},  {1, 1, 0},   {1, -1, 0},  {-1, 1, 0},
                        {-1, -1, 0}, {1, 0, 1},   {-1, 0, 1},  {1, 0, -1},  {-1, 0, -1},
                        {0, 1, 1},   {0, 1, -1},  {0, -1, 1},  {0, -1, -1}, {1, 1, 1},
                        {-1, -1, -1}, {1, 1, -1}, {-1, -1, 1}, {1, -1, 1},  {-1, 1, -1},
                        {1, -1, -1}, {-1, 1, 1}};
const double W27[27] = {8.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,
                        2.0 / 27.0,  2.0 / 27.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,
                        1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,
                        1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 216.0,
                        1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0,
                        1.0 / 216.0, 1.0 / 216.0};
const int OPP27[27] = {0,  2,  1,  4,  3,  6,  5,  10, 9,  8,  7,  14, 13, 12,
                       11, 18, 17, 16, 15, 20, 19, 22, 21, 24, 23, 26, 25};

struct Lat {
    int q;
    const int (*c)[3];
    const double* w;
    const int* opp;
};

Lat lattice(int q) {
    if (q == 27) return Lat{27, C27, W27, OPP27};
    return Lat{19, C19, W19, OPP19};
}

// ---------------------------------------------------------------------------------------------
// Macroscopic moments, Simulation::Collision (Simulation.cpp:52-58 density, :69-105 D3Q19
// velocity, :112-175 D3Q27 velocity).
// ---------------------------------------------------------------------------------------------
inline void moments(int q, const double* f, double& rho, double u[3]) {
    rho = 0.0;
    for (int d = 0; d < q; d++) rho += f[d];
    if (q == 19) {
        u[0] = ((f[0] + f[6] + f[7] + f[8] + f[9]) - (f[1] + f[10] + f[11] + f[12] + f[13])) / rho;
        u[1] = ((f[2] + f[6] + f[10] + f[14] + f[15]) - (f[3] + f[7] + f[11] + f[16] + f[17])) / rho;
        u[2] = ((f[4] + f[8] + f[12] + f[14] + f[16]) - (f[5] + f[9] + f[13] + f[15] + f[17])) / rho;
    } else {
        u[0] = ((f[1] + f[7] + f[8] + f[11] + f[13] + f[19] + f[21] + f[23] + f[25]) -
                (f[2] + f[9] + f[10] + f[12] + f[14] + f[20] + f[22] + f[24] + f[26])) /
               rho;
        u[1] = ((f[3] + f[7] + f[9] + f[15] + f[16] + f[19] + f[21] + f[24] + f[26]) -
                (f[4] + f[8] + f[10] + f[17] + f[18] + f[20] + f[22] + f[23] + f[25])) /
               rho;
        u[2] = ((f[5] + f[11] + f[12] + f[15] + f[17] + f[19] + f[22] + f[23] + f[26]) -
                (f[6] + f[13] + f[14] + f[16] + f[18] + f[20] + f[21] + f[24] + f[25])) /
               rho;
    }
}

// Equilibrium, Simulation.cpp:184-197.
inline void equilibrium(const Lat& L, double rho, const double u[3], double* feq) {
    for (int d = 0; d < L.q; d++) {
        double du = u[0] * (double)L.c[d][0] + u[1] * (double)L.c[d][1] + u[2] * (double)L.c[d][2];
        double u_sqr = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        feq[d] = L.w[d] * rho * (1.0 + 3.0 * du + 4.5 * du * du - 1.5 * u_sqr);
    }
}

// BGK relaxation, Simulation.cpp:201-205.
inline void relax(int q, double tau, const double* f, const double* feq, double* fnew) {
    for (int d = 0; d < q; d++) fnew[d] = f[d] + (1.0 / tau) * (feq[d] - f[d]);
}

// ---------------------------------------------------------------------------------------------
// Zou-He closures.
// ---------------------------------------------------------------------------------------------
// D3Q19 velocity inlet on the x-min face, u = (ux, 0, 0): Boundary::ApplyBc(double&)
// (Boundary.cpp:129-157).
void inlet19(double* m, double ux) {
    double uin0 = ux, uin1 = 0.0, uin2 = 0.0;
    double rhoN = (m[2] + m[3] + m[4] + m[5] + m[14] + m[15] + m[16] + m[17] + m[18] +
                   2 * (m[1] + m[10] + m[11] + m[12] + m[13])) /
                  (1 - uin0);
    double Ny = (1.0 / 2) * (m[2] + m[14] + m[15] - m[3] - m[16] - m[17]) - (1.0 / 3) * (rhoN * uin1);
    double Nz = (1.0 / 2) * (m[4] + m[16] + m[14] - m[5] - m[15] - m[17]) - (1.0 / 3) * (rhoN * uin2);
    m[0] = m[1] + (1 / 3.0) * rhoN * uin0;
    m[7] = m[10] + (1 / 6.0) * rhoN * (uin0 - uin1) + Ny;
    m[6] = m[11] + (1 / 6.0) * rhoN * (uin0 + uin1) - Ny;
    m[8] = m[13] + (1 / 6.0) * rhoN * (uin0 + uin2) - Nz;
    m[9] = m[12] + (1 / 6.0) * rhoN * (uin0 - uin2) + Nz;
}

// D3Q27 velocity inlet on the x-min face: the algebra of the reference's D3Q27 Boundary::ApplyBc
// (Boundary.cpp:87-120) with u = (ux, 0, 0).  (The reference's D3Q27 build does not compile and
// its trailing MacroscopicValue call only refreshes output fields, so it is not restated.)
void inlet27(double* m, double ux) {
    double u0 = ux, u1 = 0.0, u2 = 0.0;
    double rho = ((m[0] + m[3] + m[4] + m[5] + m[6] + m[15] + m[16] + m[17] + m[18]) +
                  2 * (m[2] + m[9] + m[10] + m[12] + m[14] + m[20] + m[22] + m[24] + m[26])) /
                 (1 - u0);
    double Ny = (1.0 / 6) * (m[3] + m[15] + m[16] - m[4] - m[17] - m[18]) - (1.0 / 9) * (rho * u1);
    double Nz = (1.0 / 6) * (m[5] + m[15] + m[17] - m[6] - m[16] - m[18]) - (1.0 / 9) * (rho * u2);
    m[1] = m[2] + (4 / 9.0) * rho * u0;
    m[7] = m[10] + (1 / 9.0) * rho * (u0 + u1) - Ny;
    m[8] = m[9] + (1 / 9.0) * rho * (u0 - u1) + Ny;
    m[11] = m[14] + (1 / 9.0) * rho * (u0 + u2) - Nz;
    m[13] = m[12] + (1 / 9.0) * rho * (u0 - u2) + Nz;
    m[19] = m[20] + (1 / 36.0) * rho * (u0 + u1 + u2) - Ny - Nz;
    m[21] = m[22] + (1 / 36.0) * rho * (u0 + u1 - u2) - Ny + Nz;
    m[23] = m[24] + (1 / 36.0) * rho * (u0 - u1 + u2) + Ny - Nz;
    m[25] = m[26] + (1 / 36.0) * rho * (u0 - u1 - u2) + Ny + Nz;
}

// D3Q19 pressure outlet on the y-max face, Simulation::ApplyOutletBc list 1
// (Simulation.cpp:642-695); rho is the prescribed density (1.0 in the reference, :630).
void outlet19_ymax(double* m, double rho) {
    double vx = 0.0, vy = 0.0, vz = 0.0;
    vy = -1 + (1.0 / rho) * (m[0] + m[1] + m[4] + m[5] + m[8] + m[9] + m[12] + m[13] + m[18] +
                             2.0 * (m[2] + m[6] + m[10] + m[14] + m[15]));
    double Nyx = 0.5 * (m[0] + m[8] + m[9] - (m[1] + m[12] + m[13])) - (1.0 / 3.0) * rho * vx;
    double Nyz = 0.5 * (m[4] + m[8] + m[12] - (m[5] + m[9] + m[13])) - (1.0 / 3.0) * rho * vz;
    m[3] = m[2] - (1.0 / 3.0) * rho * vy;
    m[11] = m[6] + (1.0 / 6.0) * rho * (-vx - vy) + Nyx;
    m[7] = m[10] + (1.0 / 6.0) * rho * (vx - vy) - Nyx;
    m[17] = m[14] + (1.0 / 6.0) * rho * (-vy - vz) + Nyz;
    m[16] = m[15] + (1.0 / 6.0) * rho * (-vy + vz) - Nyz;
}

// D3Q19 pressure outlet on the z-min face, lists 2 and 3 (Simulation.cpp:710-766, :780-836).
void outlet19_zmin(double* m, double rho) {
    double vx = 0.0, vy = 0.0, vz = 0.0;
    vz = 1 - (1.0 / rho) * (m[0] + m[1] + m[2] + m[3] + m[6] + m[7] + m[10] + m[11] + m[18] +
                            2.0 * (m[5] + m[9] + m[13] + m[15] + m[17]));
    double Nzx = 0.5 * (m[0] + m[6] + m[7] - (m[1] + m[10] + m[11])) - (1.0 / 3.0) * rho * vx;
    double Nzy = 0.5 * (m[2] + m[6] + m[10] - (m[3] + m[7] + m[11])) - (1.0 / 3.0) * rho * vy;
    m[4] = m[5] + (1.0 / 3.0) * rho * vz;
    m[8] = m[13] + (1.0 / 6.0) * rho * (vz + vx) - Nzx;
    m[12] = m[9] + (1.0 / 6.0) * rho * (vz - vx) + Nzx;
    m[14] = m[17] + (1.0 / 6.0) * rho * (vz + vy) - Nzy;
    m[16] = m[15] + (1.0 / 6.0) * rho * (vz - vy) + Nzy;
}

// D3Q19 pressure outlet on the x-max face.  No reference formula exists; this is
// outlet19_ymax with every population index mapped by the x<->y permutation
// sigma = (0 2)(1 3)(7 10)(8 14)(9 15)(12 16)(13 17) and vx<->vy.
void outlet19_xmax(double* m, double rho) {
    double vx = 0.0, vy = 0.0, vz = 0.0;
    vx = -1 + (1.0 / rho) * (m[2] + m[3] + m[4] + m[5] + m[14] + m[15] + m[16] + m[17] + m[18] +
                             2.0 * (m[0] + m[6] + m[7] + m[8] + m[9]));
    double Nxy = 0.5 * (m[2] + m[14] + m[15] - (m[3] + m[16] + m[17])) - (1.0 / 3.0) * rho * vy;
    double Nxz = 0.5 * (m[4] + m[14] + m[16] - (m[5] + m[15] + m[17])) - (1.0 / 3.0) * rho * vz;
    m[1] = m[0] - (1.0 / 3.0) * rho * vx;
    m[11] = m[6] + (1.0 / 6.0) * rho * (-vy - vx) + Nxy;
    m[10] = m[7] + (1.0 / 6.0) * rho * (vy - vx) - Nxy;
    m[13] = m[8] + (1.0 / 6.0) * rho * (-vx - vz) + Nxz;
    m[12] = m[9] + (1.0 / 6.0) * rho * (-vx + vz) - Nxz;
}

// D3Q27 pressure outlet on the x-max face.  No reference formula exists; this is inlet27 under
// the mirror x -> -x (index map 1<->2, 7<->9, 8<->10, 11<->12, 13<->14, 19<->26, 20<->25,
// 21<->24, 22<->23) with the density prescribed and u_x solved from the known populations.
void outlet27_xmax(double* m, double rho) {
    double u0 = 0.0, u1 = 0.0, u2 = 0.0;
    u0 = -1 + ((m[0] + m[3] + m[4] + m[5] + m[6] + m[15] + m[16] + m[17] + m[18]) +
               2 * (m[1] + m[7] + m[8] + m[11] + m[13] + m[19] + m[21] + m[23] + m[25])) /
                  rho;
    double Ny = (1.0 / 6) * (m[3] + m[15] + m[16] - m[4] - m[17] - m[18]) - (1.0 / 9) * (rho * u1);
    double Nz = (1.0 / 6) * (m[5] + m[15] + m[17] - m[6] - m[16] - m[18]) - (1.0 / 9) * (rho * u2);
    m[2] = m[1] + (4 / 9.0) * rho * (-u0);
    m[9] = m[8] + (1 / 9.0) * rho * (-u0 + u1) - Ny;
    m[10] = m[7] + (1 / 9.0) * rho * (-u0 - u1) + Ny;
    m[12] = m[13] + (1 / 9.0) * rho * (-u0 + u2) - Nz;
    m[14] = m[11] + (1 / 9.0) * rho * (-u0 - u2) + Nz;
    m[26] = m[25] + (1 / 36.0) * rho * (-u0 + u1 + u2) - Ny - Nz;
    m[24] = m[23] + (1 / 36.0) * rho * (-u0 + u1 - u2) - Ny + Nz;
    m[22] = m[21] + (1 / 36.0) * rho * (-u0 - u1 + u2) + Ny - Nz;
    m[20] = m[19] + (1 / 36.0) * rho * (-u0 - u1 - u2) + Ny + Nz;
}

void zouhe(int q, int closure, double* m, double ux, double rho) {
    if (q == 19) {
        switch (closure) {
            case PVLBM_INLET_XMIN: inlet19(m, ux); break;
            case PVLBM_OUTLET_YMAX: outlet19_ymax(m, rho); break;
            case PVLBM_OUTLET_ZMIN: outlet19_zmin(m, rho); break;
            case PVLBM_OUTLET_XMAX: outlet19_xmax(m, rho); break;
        }
    } else {
        switch (closure) {
            case PVLBM_INLET_XMIN: inlet27(m, ux); break;
            case PVLBM_OUTLET_XMAX: outlet27_xmax(m, rho); break;
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Ray casting (ComplexGeometry.cpp).
// ---------------------------------------------------------------------------------------------
// Geometry::isPointInsideTriangle (ComplexGeometry.cpp:620-674).
bool point_in_triangle(double x, double y, double z, const double* T) {
    double x1 = T[0], y1 = T[1], z1 = T[2];
    double x2 = T[3], y2 = T[4], z2 = T[5];
    double x3 = T[6], y3 = T[7], z3 = T[8];
    double dirX = 0.0, dirY = 0.0, dirZ = 1.0;
    double uX = x2 - x1, uY = y2 - y1, uZ = z2 - z1;
    double vX = x3 - x1, vY = y3 - y1, vZ = z3 - z1;
    double nX = uY * vZ - uZ * vY;
    double nY = uZ * vX - uX * vZ;
    double nZ = uX * vY - uY * vX;
    double nLengthSq = nX * nX + nY * nY + nZ * nZ;
    const double tol = 1e-16;
    if (nLengthSq < tol * tol) return false;
    double w0X = x - x1, w0Y = y - y1, w0Z = z - z1;
    double a = -(nX * w0X + nY * w0Y + nZ * w0Z);
    double b = nX * dirX + nY * dirY + nZ * dirZ;
    if (std::fabs(b) < tol) return false;
    double pos = a / b;
    double ix = x + pos * dirX;
    double iy = y + pos * dirY;
    double iz = z + pos * dirZ;
    double uu = uX * uX + uY * uY + uZ * uZ;
    double uv = uX * vX + uY * vY + uZ * vZ;
    double vv = vX * vX + vY * vY + vZ * vZ;
    double wX = ix - x1, wY = iy - y1, wZ = iz - z1;
    double wu = wX * uX + wY * uY + wZ * uZ;
    double wv = wX * vX + wY * vY + wZ * vZ;
    double D = uv * uv - uu * vv;
    double s = (uv * wv - vv * wu) / D;
    if (s < 0.0 || s > 1.0) return false;
    double t = (uv * wu - uu * wv) / D;
    if (t < 0.0 || (s + t) > 1.0) return false;
    return true;
}

// One triangle of Geometry::countIntersectinsTriangle (ComplexGeometry.cpp:807-845).
inline bool crosses(const double* T, const double s[3], const double e[3]) {
    const double threshold = 1e-16;
    double px = s[0], py = s[1], pz = s[2];
    double dx = e[0] - s[0], dy = e[1] - s[1], dz = e[2] - s[2];
    double nx = T[9], ny = T[10], nz = T[11];
    double qx = T[0], qy = T[1], qz = T[2];
    double numerator = (qx - px) * nx + (qy - py) * ny + (qz - pz) * nz;
    double denominator = dx * nx + dy * ny + dz * nz;
    if (std::fabs(denominator) < threshold) return false;
    double t = numerator / denominator;
    if (t >= 0.0 && t <= 1.0) {
        double ix = px + t * dx;
        double iy = py + t * dy;
        double iz = pz + t * dz;
        if (point_in_triangle(ix, iy, iz, T)) return true;
    }
    return false;
}

// The +z ray of one point: labelPoints (ComplexGeometry.cpp:989-1003) then
// adjustLineCoordinates case 0 (:855-884); rayOffset = 0.02 (:979).
inline void make_ray(double x, double y, double z, double h, const double bb[6], double s[3],
                     double e[3]) {
    const double rayOffset = 0.02;
    s[0] = x;
    s[1] = y;
    s[2] = z + h / 2;
    double rayLength = bb[5] - bb[4];
    e[0] = x;
    e[1] = y;
    e[2] = z + rayLength + 0.1;
    if (s[0] == bb[0]) { s[0] += rayOffset; e[0] += rayOffset; }
    if (s[0] == bb[1]) { s[0] -= rayOffset; e[0] -= rayOffset; }
    if (s[1] == bb[2]) { s[1] += rayOffset; e[1] += rayOffset; }
    if (s[1] == bb[3]) { s[1] -= rayOffset; e[1] -= rayOffset; }
    if (s[2] == bb[4]) { s[2] += rayOffset; }
    if (e[2] == bb[5]) { e[2] -= rayOffset; }
}

int set_threads(int threads) {
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
    return threads;
#else
    (void)threads;
    return 1;
#endif
}

}  // namespace

// =============================================================================================
// Geometry
// =============================================================================================

extern "C" int64_t orc_read_stl_binary(const char* path, double* tri, int64_t capacity) {
    // Reader::readSTLBinary (GeometryReader.cpp:119-176), stopping at the end of the file
    std::ifstream file(path, std::ios::in | std::ios::binary);
    if (!file) return -1;
    char header[80];
    file.read(header, 80);
    uint32_t num_triangles = 0;
    file.read(reinterpret_cast<char*>(&num_triangles), sizeof(num_triangles));
    int64_t count = 0;
    for (uint32_t i = 0; i < num_triangles; i++) {
        float nx, ny, nz;
        file.read(reinterpret_cast<char*>(&nx), sizeof(float));
        file.read(reinterpret_cast<char*>(&ny), sizeof(float));
        file.read(reinterpret_cast<char*>(&nz), sizeof(float));
        float norm = std::sqrt(nx * nx + ny * ny + nz * nz);
        nx /= norm;
        ny /= norm;
        nz /= norm;
        float v[9];
        file.read(reinterpret_cast<char*>(v), sizeof v);
        uint16_t attr_byte_count;
        file.read(reinterpret_cast<char*>(&attr_byte_count), sizeof(attr_byte_count));
        if (!file) break;
        if (tri && count < capacity) {
            double* T = tri + 12 * count;
            for (int k = 0; k < 9; k++) T[k] = v[k];
            T[9] = nx;
            T[10] = ny;
            T[11] = nz;
        }
        count++;
    }
    return count;
}

extern "C" int64_t orc_read_stl_ascii(const char* path, double* tri, int64_t capacity) {
    // Reader::readSTLASCII (GeometryReader.cpp:33-117): facet normal / outer loop / 3 vertex /
    // endloop / endfacet; a facet whose lines do not parse is skipped.
    std::ifstream file(path);
    if (!file.is_open()) return -1;
    std::string line, facet, normal, outer, loop, vertex, endloop, endfacet;
    std::stringstream ss;
    double x, y, z;
    double t[12] = {0};
    int64_t count = 0;
    while (std::getline(file, line)) {
        ss.clear();
        ss.str(line);
        ss >> facet >> normal >> t[9] >> t[10] >> t[11];
        if (ss.fail() || facet != "facet" || normal != "normal") continue;
        std::getline(file, line);
        ss.clear();
        ss.str(line);
        ss >> outer >> loop;
        if (ss.fail() || outer != "outer" || loop != "loop") continue;
        for (int i = 0; i < 3; ++i) {
            std::getline(file, line);
            ss.clear();
            ss.str(line);
            ss >> vertex >> x >> y >> z;
            if (ss.fail() || vertex != "vertex") break;
            t[3 * i + 0] = x;
            t[3 * i + 1] = y;
            t[3 * i + 2] = z;
        }
        std::getline(file, line);
        ss.clear();
        ss.str(line);
        ss >> endloop;
        if (ss.fail() || endloop != "endloop") continue;
        std::getline(file, line);
        ss.clear();
        ss.str(line);
        ss >> endfacet;
        if (ss.fail() || endfacet != "endfacet") continue;
        if (tri && count < capacity) std::memcpy(tri + 12 * count, t, sizeof t);
        count++;
    }
    return count;
}

extern "C" void orc_bbox(const double* tri, int64_t ntri, double bb[6]) {
    // ComplexGeometry.cpp:142-154: x over v1..v3, y and z over v1, v2 only.
    double inf = std::numeric_limits<double>::infinity();
    double mnx = inf, mny = inf, mnz = inf, mxx = -inf, mxy = -inf, mxz = -inf;
    for (int64_t i = 0; i < ntri; i++) {
        const double* T = tri + 12 * i;
        mnx = std::min({mnx, T[0], T[3], T[6]});
        mny = std::min({mny, T[1], T[4]});
        mnz = std::min({mnz, T[2], T[5]});
        mxx = std::max({mxx, T[0], T[3], T[6]});
        mxy = std::max({mxy, T[1], T[4]});
        mxz = std::max({mxz, T[2], T[5]});
    }
    bb[0] = mnx; bb[1] = mxx; bb[2] = mny; bb[3] = mxy; bb[4] = mnz; bb[5] = mxz;
}

extern "C" void orc_grid_dims(const double bb[6], double h, int32_t dims[3]) {
    // ComplexGeometry.cpp:163-165 (nz has no +1).
    dims[0] = (int32_t)(std::ceil((bb[1] - bb[0]) / h) + 1);
    dims[1] = (int32_t)(std::ceil((bb[3] - bb[2]) / h) + 1);
    dims[2] = (int32_t)std::ceil((bb[5] - bb[4]) / h);
}

extern "C" int orc_point_in_triangle(double x, double y, double z, const double* tri12) {
    return point_in_triangle(x, y, z, tri12) ? 1 : 0;
}

extern "C" int orc_count_crossings(const double* tri, int64_t ntri, const double start[3],
                                   const double end[3]) {
    int count = 0;
    for (int64_t t = 0; t < ntri; t++)
        if (crosses(tri + 12 * t, start, end)) count++;
    return count;
}

extern "C" void orc_label_points(const double* tri, int64_t ntri, const double bb[6],
                                 const int32_t dims[3], const double origin[3], double h,
                                 uint8_t* fluid, int threads) {
    const int64_t nx = dims[0], ny = dims[1], nz = dims[2];
    threads = set_threads(threads);
#pragma omp parallel for schedule(dynamic, 1) collapse(2) num_threads(threads)
    for (int64_t k = 0; k < nz; k++)
        for (int64_t j = 0; j < ny; j++)
            for (int64_t i = 0; i < nx; i++) {
                // point = min + idx*spacing (ComplexGeometry.cpp:337-340)
                double x = origin[0] + (double)i * h;
                double y = origin[1] + (double)j * h;
                double z = origin[2] + (double)k * h;
                double s[3], e[3];
                make_ray(x, y, z, h, bb, s, e);
                int c = 0;
                for (int64_t t = 0; t < ntri; t++)
                    if (crosses(tri + 12 * t, s, e)) c++;
                fluid[i + nx * (j + ny * k)] = (c % 2 == 1) ? 1 : 0;
            }
}

extern "C" void orc_label_points_binned(const double* tri, int64_t ntri, const double bb[6],
                                        const int32_t dims[3], const double origin[3], double h,
                                        uint8_t* fluid, int threads) {
    const int64_t nx = dims[0], ny = dims[1], nz = dims[2];
    // ray x (per i) and y (per j) after the bbox nudges, which depend on x and y only
    std::vector<double> xr(nx), yr(ny);
    for (int64_t i = 0; i < nx; i++) {
        double s[3], e[3];
        make_ray(origin[0] + (double)i * h, origin[1], origin[2], h, bb, s, e);
        xr[i] = s[0];
    }
    for (int64_t j = 0; j < ny; j++) {
        double s[3], e[3];
        make_ray(origin[0], origin[1] + (double)j * h, origin[2], h, bb, s, e);
        yr[j] = s[1];
    }
    double scale = 1.0;
    for (int a = 0; a < 6; a++) scale = std::max(scale, std::fabs(bb[a]));
    const double margin = 1e-6 * scale;
    std::vector<std::vector<int64_t>> cols((size_t)(nx * ny));
    const int64_t slack = 1 + (int64_t)std::ceil(0.02 / h);
    for (int64_t t = 0; t < ntri; t++) {
        const double* T = tri + 12 * t;
        double x0 = std::min({T[0], T[3], T[6]}) - margin, x1 = std::max({T[0], T[3], T[6]}) + margin;
        double y0 = std::min({T[1], T[4], T[7]}) - margin, y1 = std::max({T[1], T[4], T[7]}) + margin;
        int64_t i0 = std::max<int64_t>(0, (int64_t)std::floor((x0 - origin[0]) / h) - slack);
        int64_t i1 = std::min<int64_t>(nx - 1, (int64_t)std::ceil((x1 - origin[0]) / h) + slack);
        int64_t j0 = std::max<int64_t>(0, (int64_t)std::floor((y0 - origin[1]) / h) - slack);
        int64_t j1 = std::min<int64_t>(ny - 1, (int64_t)std::ceil((y1 - origin[1]) / h) + slack);
        for (int64_t j = j0; j <= j1; j++) {
            if (yr[j] < y0 || yr[j] > y1) continue;
            for (int64_t i = i0; i <= i1; i++)
                if (xr[i] >= x0 && xr[i] <= x1) cols[i + nx * j].push_back(t);
        }
    }
    threads = set_threads(threads);
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
    for (int64_t col = 0; col < nx * ny; col++) {
        int64_t i = col % nx, j = col / nx;
        const std::vector<int64_t>& cand = cols[col];
        for (int64_t k = 0; k < nz; k++) {
            double x = origin[0] + (double)i * h;
            double y = origin[1] + (double)j * h;
            double z = origin[2] + (double)k * h;
            double s[3], e[3];
            make_ray(x, y, z, h, bb, s, e);
            int c = 0;
            for (int64_t t : cand)
                if (crosses(tri + 12 * t, s, e)) c++;
            fluid[i + nx * (j + ny * k)] = (c % 2 == 1) ? 1 : 0;
        }
    }
}

// =============================================================================================
// Simulation
// =============================================================================================

struct orc_sim {
    Lat L;
    int q;
    int64_t nx, ny, nz, nbox;
    double ox, oy, oz, h, tau, dt;
    int threads = 1;
    std::vector<pvlbm_face> faces;
    // nodes in the reference's traversal order, i outer, j, k inner (Subdomain.cpp:583-596)
    std::vector<int64_t> node_box;
    std::vector<int32_t> box_node;
    std::vector<uint8_t> kind;  // enum pvlbm_kind
    std::vector<int8_t> face;   // matched face index for inlet/outlet nodes
    std::vector<int32_t> nbr;   // [n][q], -1 = null (Subdomain.cpp:173-197)
    std::vector<double> f, fnew, feq, rho, u, prof;
    std::vector<int32_t> lists[7];  // LIST_FLUID..LIST_NTW as node ids
    std::vector<double> normal;     // [nw][3]
    std::vector<double> wss, tau3, acc;  // [nw], [nw][3], [nw][4]
    // pull state: post-collision lattice A, stale populations P, previous macros M
    bool pull_ready = false;
    std::vector<double> A, A2, P, M;
    int64_t gate = 0;
    int64_t iter = 0;
};

namespace {

bool in_box(const orc_sim* s, int64_t i, int64_t j, int64_t k) {
    return i >= 0 && i < s->nx && j >= 0 && j < s->ny && k >= 0 && k < s->nz;
}

// Face match of ComplexGeometry.cpp:1370-1408 in long double.
bool face_match(const pvlbm_face& F, long double c[3]) {
    long double tolerance = F.tol;
    int a = (F.axis == 0) ? 1 : 0;
    int b = (F.axis == 2) ? 1 : 2;
    long double plane = (long double)F.plane + (long double)F.plane_lo;
    return (std::abs(c[F.axis] - plane) <= tolerance) && (F.lo[0] < c[a] && c[a] < F.hi[0]) &&
           (F.lo[1] < c[b] && c[b] < F.hi[1]);
}

}  // namespace

// classifyBoundaryPoints (ComplexGeometry.cpp:1369-1409) and the precedence of CreateNodes
// (Subdomain.cpp:124-155) for one boundary point: the index of the face it is assigned to, or -1
// (a wall point).  The reference pushes a point onto every list whose plane and window it matches;
// CreateNodes then tries the inlet list first, then outlets 1, 2 and 3, so the first role that
// matches wins (and here, within a role, the first face in array order).
extern "C" int orc_classify_point(const pvlbm_face* faces, int nfaces, const double c[3]) {
    long double c3[3] = {(long double)c[0], (long double)c[1], (long double)c[2]};
    for (int role = PVLBM_INLET; role <= PVLBM_OUTLET3; role++)
        for (int fi = 0; fi < nfaces; fi++)
            if (faces[fi].role == role && face_match(faces[fi], c3)) return fi;
    return -1;
}

namespace {

// Simulation::ComputeWallShearStress for one node (Simulation.cpp:880-898).
void wss_node(const Lat& L, const double* f, const double* feq, const double* n, double C,
              double tau_i[3]) {
    int dim = 3;
    for (int i = 0; i < dim; i++) {
        double sum_alpha = 0.0;
        for (int alpha = 0; alpha < L.q; alpha++) {
            double sum_k = 0.0;
            for (int k = 0; k < dim; k++) sum_k += (double)L.c[alpha][k] * n[k] * n[i];
            double sum_j = 0.0;
            for (int j = 0; j < dim; j++)
                sum_j += (double)L.c[alpha][j] * n[j] * ((double)L.c[alpha][i] - sum_k);
            double f_ne_alpha = f[alpha] - feq[alpha];
            sum_alpha += f_ne_alpha * sum_j;
        }
        tau_i[i] = C * sum_alpha;
    }
}

void wss_node(const orc_sim* s, const double* f, const double* feq, const double* n,
              double C, double tau_i[3]) {
    wss_node(s->L, f, feq, n, C, tau_i);
}

double wss_constant(const orc_sim* s) {
    // Simulation.cpp:880-881
    double phys_den = 1060;
    double dtdt = s->dt * s->dt;
    double dxdx = s->h * s->h * 1e-6;
    return phys_den * dtdt / dxdx;
}

void accumulate_wss(orc_sim* s, int64_t w, const double tau_i[3]) {
    double WSS = std::sqrt(tau_i[0] * tau_i[0] + tau_i[1] * tau_i[1] + tau_i[2] * tau_i[2]);
    s->wss[w] = WSS;
    for (int i = 0; i < 3; i++) {
        s->tau3[3 * w + i] = tau_i[i];
        s->acc[4 * w + i] += tau_i[i];
    }
    s->acc[4 * w + 3] += WSS;
}

// calculateBoundaryNormals for one point (Subdomain.cpp:301-369): sum over every triangle, in
// file order, of the unit face normal times area / distance^2 to the centroid, then normalised.
void normal_at(const double* tri, int64_t ntri, double x, double y, double z, double out[3]) {
    double n0 = 0.0, n1 = 0.0, n2 = 0.0;
    for (int64_t t = 0; t < ntri; t++) {
        const double* T = tri + 12 * t;
        double ux = T[3] - T[0], uy = T[4] - T[1], uz = T[5] - T[2];
        double vx = T[6] - T[0], vy = T[7] - T[1], vz = T[8] - T[2];
        double nx = uy * vz - uz * vy;
        double ny = uz * vx - ux * vz;
        double nz = ux * vy - uy * vx;
        double norm = std::sqrt(nx * nx + ny * ny + nz * nz);
        if (norm > 0) {
            nx /= norm;
            ny /= norm;
            nz /= norm;
        }
        double area = 0.5 * norm;
        double cX = (T[0] + T[3] + T[6]) / 3.0;
        double cY = (T[1] + T[4] + T[7]) / 3.0;
        double cZ = (T[2] + T[5] + T[8]) / 3.0;
        double distance = std::sqrt((x - cX) * (x - cX) + (y - cY) * (y - cY) + (z - cZ) * (z - cZ));
        double weight = area / (distance * distance);
        n0 += nx * weight;
        n1 += ny * weight;
        n2 += nz * weight;
    }
    double norm = std::sqrt(n0 * n0 + n1 * n1 + n2 * n2);
    if (norm > 0) {
        n0 /= norm;
        n1 /= norm;
        n2 /= norm;
    }
    out[0] = n0;
    out[1] = n1;
    out[2] = n2;
}

}  // namespace

extern "C" orc_sim* orc_sim_create(const uint8_t* fluid, const int32_t dims[3],
                                   const double origin[3], double h, int q,
                                   const pvlbm_face* faces, int nfaces, const double* tri,
                                   int64_t ntri, double tau, double dt, int normals) {
    if (q != 19 && q != 27) return nullptr;
    orc_sim* s = new orc_sim();
    s->L = lattice(q);
    s->q = q;
    s->nx = dims[0]; s->ny = dims[1]; s->nz = dims[2];
    s->nbox = s->nx * s->ny * s->nz;
    s->ox = origin[0]; s->oy = origin[1]; s->oz = origin[2];
    s->h = h; s->tau = tau; s->dt = dt;
    s->faces.assign(faces, faces + nfaces);
    s->box_node.assign((size_t)s->nbox, -1);
    auto bidx = [&](int64_t i, int64_t j, int64_t k) { return i + s->nx * (j + s->ny * k); };
    // nodes = fluid points (Subdomain.cpp:113-157), traversal i, j, k
    for (int64_t i = 0; i < s->nx; i++)
        for (int64_t j = 0; j < s->ny; j++)
            for (int64_t k = 0; k < s->nz; k++)
                if (fluid[bidx(i, j, k)]) {
                    s->box_node[bidx(i, j, k)] = (int32_t)s->node_box.size();
                    s->node_box.push_back(bidx(i, j, k));
                }
    const int64_t nn = (int64_t)s->node_box.size();
    s->kind.assign(nn, PVLBM_FLUID);
    s->face.assign(nn, -1);
    s->nbr.assign((size_t)(nn * q), -1);
    for (int64_t n = 0; n < nn; n++) {
        int64_t b = s->node_box[n];
        int64_t i = b % s->nx, j = (b / s->nx) % s->ny, k = b / (s->nx * s->ny);
        // findBoundaryPoints (ComplexGeometry.cpp:1269-1286): any Q neighbour outside or not fluid
        bool boundary = false;
        for (int d = 0; d < q; d++) {
            int64_t a = i + s->L.c[d][0], bj = j + s->L.c[d][1], c = k + s->L.c[d][2];
            if (in_box(s, a, bj, c)) {
                int32_t m = s->box_node[bidx(a, bj, c)];
                s->nbr[n * q + d] = m;
                if (m < 0) boundary = true;
            } else {
                boundary = true;
            }
        }
        if (!boundary) continue;
        const double c3[3] = {s->ox + (double)i * h, s->oy + (double)j * h, s->oz + (double)k * h};
        const int matched = orc_classify_point(faces, nfaces, c3);
        if (matched >= 0) {
            s->kind[n] = faces[matched].role == PVLBM_INLET ? PVLBM_KINLET : PVLBM_KOUTLET;
            s->face[n] = (int8_t)matched;
        } else {
            s->kind[n] = PVLBM_WALL;
        }
    }
    // IdentifyNextToWallLattices (Subdomain.cpp:201-218)
    for (int64_t n = 0; n < nn; n++) {
        if (s->kind[n] != PVLBM_WALL) continue;
        for (int d = 0; d < q; d++) {
            int32_t m = s->nbr[n * q + d];
            if (m >= 0 && (s->kind[m] == PVLBM_FLUID || s->kind[m] == PVLBM_NTW))
                s->kind[m] = PVLBM_NTW;
        }
    }
    for (int64_t n = 0; n < nn; n++) {
        s->lists[PVLBM_LIST_FLUID].push_back((int32_t)n);
        switch (s->kind[n]) {
            case PVLBM_KINLET: s->lists[PVLBM_LIST_INLET].push_back((int32_t)n); break;
            case PVLBM_KOUTLET:
                s->lists[PVLBM_LIST_OUTLET1 + faces[s->face[n]].role - 1].push_back((int32_t)n);
                break;
            case PVLBM_WALL: s->lists[PVLBM_LIST_WALL].push_back((int32_t)n); break;
            case PVLBM_NTW: s->lists[PVLBM_LIST_NTW].push_back((int32_t)n); break;
        }
    }
    // next-to-wall list in box-index order (the reference's unordered_set order is unspecified)
    std::sort(s->lists[PVLBM_LIST_NTW].begin(), s->lists[PVLBM_LIST_NTW].end(),
              [&](int32_t a, int32_t b) { return s->node_box[a] < s->node_box[b]; });
    const int64_t nw = (int64_t)s->lists[PVLBM_LIST_NTW].size();
    s->normal.assign((size_t)(3 * nw), 0.0);
    if (normals) {
        // calculateBoundaryNormals for the next-to-wall nodes (Subdomain.cpp:301-369)
        int th = set_threads(s->threads);
        (void)th;
#pragma omp parallel for schedule(dynamic, 8)
        for (int64_t w = 0; w < nw; w++) {
            int64_t b = s->node_box[s->lists[PVLBM_LIST_NTW][w]];
            int64_t i = b % s->nx, j = (b / s->nx) % s->ny, k = b / (s->nx * s->ny);
            normal_at(tri, ntri, s->ox + (double)i * h, s->oy + (double)j * h, s->oz + (double)k * h,
                      &s->normal[3 * w]);
        }
    }
    // Initialize: f = w, f_new = 0 (Node.cpp:30-37), f_eq = 0 (Node.cpp:7), rho/u = 0
    s->f.resize((size_t)(nn * q));
    for (int64_t n = 0; n < nn; n++)
        for (int d = 0; d < q; d++) s->f[n * q + d] = s->L.w[d];
    s->fnew.assign((size_t)(nn * q), 0.0);
    s->feq.assign((size_t)(nn * q), 0.0);
    s->rho.assign(nn, 0.0);
    s->u.assign((size_t)(3 * nn), 0.0);
    s->prof.assign(nn, 1.0);
    s->wss.assign(nw, 0.0);
    s->tau3.assign((size_t)(3 * nw), 0.0);
    s->acc.assign((size_t)(4 * nw), 0.0);  // InitializeShearStressAccumulators (:32-38)
    return s;
}

extern "C" void orc_sim_destroy(orc_sim* s) { delete s; }

extern "C" void orc_sim_set_threads(orc_sim* s, int threads) { s->threads = set_threads(threads); }

extern "C" void orc_sim_counts(const orc_sim* s, int64_t out[7]) {
    for (int l = 0; l < 7; l++) out[l] = (int64_t)s->lists[l].size();
}

extern "C" void orc_sim_kinds(const orc_sim* s, uint8_t* kinds) {
    std::memset(kinds, 0, (size_t)s->nbox);
    for (size_t n = 0; n < s->node_box.size(); n++) kinds[s->node_box[n]] = s->kind[n];
}

extern "C" int64_t orc_sim_list(const orc_sim* s, int list, int64_t* idx) {
    if (list < 0 || list > 6) return -1;
    const std::vector<int32_t>& L = s->lists[list];
    if (idx)
        for (size_t i = 0; i < L.size(); i++) idx[i] = s->node_box[L[i]];
    return (int64_t)L.size();
}

extern "C" void orc_sim_set_inlet_profile(orc_sim* s, const double* prof_box) {
    for (size_t n = 0; n < s->node_box.size(); n++) s->prof[n] = prof_box[s->node_box[n]];
}

namespace {

// Zou-He / wall closures of one node's post-stream populations, in the reference's call
// order ApplyZouHeBc, ApplyOutletBc, ApplyWallBc (Simulation.cpp:582-584); each node is in at
// most one list, so the order between nodes does not matter.
void apply_bc(const orc_sim* s, int64_t n, double* m, double scale) {
    const int q = s->q;
    switch (s->kind[n]) {
        case PVLBM_KINLET: {
            const pvlbm_face& F = s->faces[s->face[n]];
            double vel = scale * s->prof[n];
            zouhe(q, F.closure, m, vel, F.rho);
            break;
        }
        case PVLBM_KOUTLET: {
            const pvlbm_face& F = s->faces[s->face[n]];
            zouhe(q, F.closure, m, 0.0, F.rho);
            break;
        }
        case PVLBM_WALL:
            // ApplyWallBc (Simulation.cpp:846-862): in direction order, f[opp] = f[dir] when the
            // neighbour in dir is null; later directions see earlier assignments.
            for (int d = 0; d < q; d++)
                if (s->nbr[n * q + d] < 0) m[s->L.opp[d]] = m[d];
            break;
        default: break;
    }
}

}  // namespace

extern "C" void orc_sim_step_push(orc_sim* s, int nsteps, const double* inlet_scale,
                                  int wss_every) {
    const int q = s->q;
    const int64_t nn = (int64_t)s->node_box.size();
    const double C = wss_constant(s);
    if (wss_every < 1) wss_every = 1;
    s->pull_ready = false;
    for (int step = 0; step < nsteps; step++) {
        double scale = inlet_scale ? inlet_scale[step] : 1.0;
        // Collision (Simulation.cpp:43-208)
#pragma omp parallel for schedule(static) num_threads(s->threads)
        for (int64_t n = 0; n < nn; n++) {
            double* f = &s->f[n * q];
            moments(q, f, s->rho[n], &s->u[3 * n]);
            equilibrium(s->L, s->rho[n], &s->u[3 * n], &s->feq[n * q]);
            relax(q, s->tau, f, &s->feq[n * q], &s->fnew[n * q]);
        }
        // Streaming (Simulation.cpp:211-226): push where the neighbour exists and f_new != 0
#pragma omp parallel for schedule(static) num_threads(s->threads)
        for (int64_t n = 0; n < nn; n++)
            for (int d = 0; d < q; d++) {
                int32_t m = s->nbr[n * q + d];
                if (m >= 0 && s->fnew[n * q + d] != 0.0) s->f[(int64_t)m * q + d] = s->fnew[n * q + d];
            }
        // Boundary closures (Simulation.cpp:580-584)
        for (int l = PVLBM_LIST_INLET; l <= PVLBM_LIST_WALL; l++) {
            const std::vector<int32_t>& Ls = s->lists[l];
#pragma omp parallel for schedule(static) num_threads(s->threads)
            for (int64_t i = 0; i < (int64_t)Ls.size(); i++)
                apply_bc(s, Ls[i], &s->f[(int64_t)Ls[i] * q], scale);
        }
        // ComputeWallShearStress (Simulation.cpp:880-898)
        if (s->iter % wss_every == 0) {
            const std::vector<int32_t>& W = s->lists[PVLBM_LIST_NTW];
#pragma omp parallel for schedule(static) num_threads(s->threads)
            for (int64_t w = 0; w < (int64_t)W.size(); w++) {
                int64_t n = W[w];
                double t3[3];
                wss_node(s, &s->f[n * q], &s->feq[n * q], &s->normal[3 * w], C, t3);
                accumulate_wss(s, w, t3);
            }
        }
        s->iter++;
    }
}

extern "C" void orc_sim_step_pull(orc_sim* s, int nsteps, const double* inlet_scale,
                                  int wss_every) {
    const int q = s->q;
    const int64_t nn = (int64_t)s->node_box.size();
    const double C = wss_constant(s);
    if (wss_every < 1) wss_every = 1;
    if (!s->pull_ready) {
        // A = collide(f^n), P = f^n, M = (rho^n, u^n)
        s->A.resize((size_t)(nn * q));
        s->A2.resize((size_t)(nn * q));
        s->P = s->f;
        s->M.resize((size_t)(4 * nn));
        for (int64_t n = 0; n < nn; n++) {
            double rho, u[3], feq[27];
            moments(q, &s->f[n * q], rho, u);
            equilibrium(s->L, rho, u, feq);
            relax(q, s->tau, &s->f[n * q], feq, &s->A[n * q]);
            s->M[4 * n] = rho;
            s->M[4 * n + 1] = u[0]; s->M[4 * n + 2] = u[1]; s->M[4 * n + 3] = u[2];
        }
        s->pull_ready = true;
    }
    std::vector<int32_t> ntw_slot(nn, -1);
    for (size_t w = 0; w < s->lists[PVLBM_LIST_NTW].size(); w++)
        ntw_slot[s->lists[PVLBM_LIST_NTW][w]] = (int32_t)w;
    for (int step = 0; step < nsteps; step++) {
        double scale = inlet_scale ? inlet_scale[step] : 1.0;
        bool do_wss = (s->iter % wss_every == 0);
        int64_t gate = 0;
#pragma omp parallel for schedule(static) num_threads(s->threads) reduction(+ : gate)
        for (int64_t n = 0; n < nn; n++) {
            double g[27];
            bool bnd = s->kind[n] >= PVLBM_KINLET;
            // gather f_d(x) = A_d(x - c_d); x - c_d is the neighbour in direction opp(d)
            for (int d = 0; d < q; d++) {
                int32_t up = s->nbr[n * q + s->L.opp[d]];
                if (up >= 0) {
                    double v = s->A[(int64_t)up * q + d];
                    if (v != 0.0) {
                        g[d] = v;
                    } else {
                        g[d] = s->P[n * q + d];
                        gate++;
                    }
                } else {
                    g[d] = s->P[n * q + d];
                }
            }
            apply_bc(s, n, g, scale);
            // the oracle keeps every node's previous populations so that a gated pull can be
            // resolved exactly; the GPU keeps them for boundary nodes only and counts the rest
            (void)bnd;
            std::memcpy(&s->P[n * q], g, sizeof(double) * q);
            int32_t w = ntw_slot[n];
            if (w >= 0 && do_wss) {
                double feq[27], t3[3];
                equilibrium(s->L, s->M[4 * n], &s->M[4 * n + 1], feq);
                wss_node(s, g, feq, &s->normal[3 * w], C, t3);
                accumulate_wss(s, w, t3);
            }
            double rho, u[3], feq[27];
            moments(q, g, rho, u);
            s->M[4 * n] = rho;
            s->M[4 * n + 1] = u[0]; s->M[4 * n + 2] = u[1]; s->M[4 * n + 3] = u[2];
            equilibrium(s->L, rho, u, feq);
            relax(q, s->tau, g, feq, &s->A2[n * q]);
        }
        s->gate += gate;
        std::swap(s->A, s->A2);
        s->iter++;
    }
    s->f = s->P;  // the post-closure populations of the last iteration
}

extern "C" void orc_sim_macros(const orc_sim* s, double* rho, double* u) {
    const int q = s->q;
    std::memset(rho, 0, sizeof(double) * (size_t)s->nbox);
    std::memset(u, 0, sizeof(double) * (size_t)(3 * s->nbox));
    for (size_t n = 0; n < s->node_box.size(); n++) {
        double r, v[3];
        moments(q, &s->f[n * q], r, v);
        int64_t b = s->node_box[n];
        rho[b] = r;
        u[b] = v[0];
        u[s->nbox + b] = v[1];
        u[2 * s->nbox + b] = v[2];
    }
}

extern "C" void orc_sim_populations(const orc_sim* s, double* f) {
    // nodes sorted by box index
    std::vector<int32_t> order(s->node_box.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = (int32_t)i;
    std::sort(order.begin(), order.end(),
              [&](int32_t a, int32_t b) { return s->node_box[a] < s->node_box[b]; });
    for (size_t i = 0; i < order.size(); i++)
        std::memcpy(f + i * s->q, &s->f[(size_t)order[i] * s->q], sizeof(double) * s->q);
}

extern "C" void orc_sim_set_populations(orc_sim* s, const double* f) {
    std::vector<int32_t> order(s->node_box.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = (int32_t)i;
    std::sort(order.begin(), order.end(),
              [&](int32_t a, int32_t b) { return s->node_box[a] < s->node_box[b]; });
    for (size_t i = 0; i < order.size(); i++)
        std::memcpy(&s->f[(size_t)order[i] * s->q], f + i * s->q, sizeof(double) * s->q);
    s->pull_ready = false;
}

extern "C" void orc_sim_wss(const orc_sim* s, double* out6, double* tau3, double* acc4) {
    const std::vector<int32_t>& W = s->lists[PVLBM_LIST_NTW];
    const int64_t nw = (int64_t)W.size();
    for (int64_t w = 0; w < nw; w++) {
        int64_t b = s->node_box[W[w]];
        int64_t i = b % s->nx, j = (b / s->nx) % s->ny, k = b / (s->nx * s->ny);
        const double* a = &s->acc[4 * w];
        // ComputeOSIAndTAWSS (Simulation.cpp:899-905), period 0.9 (pulsatile.h:12)
        double tawss = a[3] / 0.9;
        double norm_acc = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        double osi = 0.5 * (1.0 - (norm_acc / a[3]));
        if (out6) {
            out6[6 * w + 0] = s->ox + (double)i * s->h;
            out6[6 * w + 1] = s->oy + (double)j * s->h;
            out6[6 * w + 2] = s->oz + (double)k * s->h;
            out6[6 * w + 3] = s->wss[w];
            out6[6 * w + 4] = osi;
            out6[6 * w + 5] = tawss;
        }
        if (tau3)
            for (int c = 0; c < 3; c++) tau3[c * nw + w] = s->tau3[3 * w + c];
        if (acc4)
            for (int c = 0; c < 4; c++) acc4[c * nw + w] = a[c];
    }
}

extern "C" void orc_sim_normals(const orc_sim* s, double* n3) {
    const int64_t nw = (int64_t)s->lists[PVLBM_LIST_NTW].size();
    for (int64_t w = 0; w < nw; w++)
        for (int c = 0; c < 3; c++) n3[c * nw + w] = s->normal[3 * w + c];
}

extern "C" void orc_sim_set_normals(orc_sim* s, const double* n3) {
    const int64_t nw = (int64_t)s->lists[PVLBM_LIST_NTW].size();
    for (int64_t w = 0; w < nw; w++)
        for (int c = 0; c < 3; c++) s->normal[3 * w + c] = n3[c * nw + w];
}

extern "C" void orc_normal_at(const double* tri, int64_t ntri, const double* xyz, int64_t n, double* n3) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t p = 0; p < n; p++) {
        double o[3];
        normal_at(tri, ntri, xyz[3 * p], xyz[3 * p + 1], xyz[3 * p + 2], o);
        for (int c = 0; c < 3; c++) n3[c * n + p] = o[c];
    }
}

extern "C" int64_t orc_sim_gate_events(const orc_sim* s) { return s->gate; }

namespace {

// Geometry::FindLineTriangleIntersection (ComplexGeometry.cpp:689-730).
bool find_line_triangle(const double start[3], const double end[3], const double* TRI, double ip[3]) {
    const double threshold = 1e-7;
    double px = start[0], py = start[1], pz = start[2];
    double dx = end[0] - start[0], dy = end[1] - start[1], dz = end[2] - start[2];
    double nx = TRI[9], ny = TRI[10], nz = TRI[11];
    double qx = TRI[0], qy = TRI[1], qz = TRI[2];
    double denom = (dx * nx + dy * ny + dz * nz);
    if (std::abs(denom) < threshold) return false;
    double t = ((qx - px) * nx + (qy - py) * ny + (qz - pz) * nz) / denom;
    if (t < 0.0 || t > 1.0) return false;
    double ix = px + t * dx;
    double iy = py + t * dy;
    double iz = pz + t * dz;
    if (point_in_triangle(ix, iy, iz, TRI)) {
        ip[0] = ix;
        ip[1] = iy;
        ip[2] = iz;
        return true;
    }
    return false;
}

}  // namespace

// Geometry::CalculateWallDistances (ComplexGeometry.cpp:550-608) over the wall nodes in the
// oracle's list order: for dir < ndirs (the reference hard-codes 19), a neighbour in the box that
// is not fluid gives the segment point -> point + c * h, and the first triangle in file order that
// it intersects gives q = distance / h.  box_out[w] = box index, q_out[w][q] (-1 without a link).
extern "C" void orc_sim_link_fractions(const orc_sim* s, const double* tri, int64_t ntri, int ndirs,
                                       int64_t* box_out, double* q_out) {
    const std::vector<int32_t>& W = s->lists[PVLBM_LIST_WALL];
    for (size_t w = 0; w < W.size(); w++) {
        const int64_t b = s->node_box[W[w]];
        const int64_t i = b % s->nx, j = (b / s->nx) % s->ny, k = b / (s->nx * s->ny);
        box_out[w] = b;
        for (int d = 0; d < s->q; d++) q_out[w * s->q + d] = -1.0;
        for (int dir = 0; dir < ndirs; ++dir) {
            const int64_t xn = i + s->L.c[dir][0], yn = j + s->L.c[dir][1], zn = k + s->L.c[dir][2];
            // solid = a box point that is not a node
            if (!in_box(s, xn, yn, zn) || s->box_node[(size_t)(xn + s->nx * (yn + s->ny * zn))] >= 0) continue;
            double start[3] = {s->ox + (double)i * s->h, s->oy + (double)j * s->h, s->oz + (double)k * s->h};
            double link[3] = {s->L.c[dir][0] * s->h, s->L.c[dir][1] * s->h, s->L.c[dir][2] * s->h};
            double end[3] = {start[0] + link[0], start[1] + link[1], start[2] + link[2]};
            for (int64_t t = 0; t < ntri; t++) {
                double ip[3];
                if (find_line_triangle(start, end, tri + 12 * t, ip)) {
                    double distanceToIntersection =
                        std::sqrt((ip[0] - start[0]) * (ip[0] - start[0]) + (ip[1] - start[1]) * (ip[1] - start[1]) +
                                  (ip[2] - start[2]) * (ip[2] - start[2]));
                    q_out[w * s->q + dir] = distanceToIntersection / s->h;
                    break;
                }
            }
        }
    }
}

extern "C" void orc_collide(int q, double tau, const double* f, double* feq, double* fnew,
                            double* out4) {
    Lat L = lattice(q);
    double rho, u[3];
    moments(q, f, rho, u);
    equilibrium(L, rho, u, feq);
    relax(q, tau, f, feq, fnew);
    out4[0] = rho;
    out4[1] = u[0];
    out4[2] = u[1];
    out4[3] = u[2];
}

extern "C" void orc_wss_node(int q, const double* f, const double* feq, const double n[3], double C,
                             double tau[3]) {
    wss_node(lattice(q), f, feq, n, C, tau);
}

extern "C" void orc_zouhe(int q, int closure, double* f, double ux, double rho_out) {
    zouhe(q, closure, f, ux, rho_out);
}

extern "C" double orc_pulsatile_velocity(double x, double period) {
    // Simulation::getInletVelocity (Simulation.cpp:548-563) with Pulsatile (pulsatile.h:16-28)
    double vel = 0.0;
    double t = std::fmod(x, period);
    if (t >= 0 && t < (0.1 * period)) {
        vel = (-8.579e+08 * std::pow(t, 8) + 4.059e+08 * std::pow(t, 7) - 7.817e+07 * std::pow(t, 6) +
               7.994e+06 * std::pow(t, 5) - 4.835e+05 * std::pow(t, 4) + 1.782e+04 * std::pow(t, 3) -
               321.6 * std::pow(t, 2) + 3.215 * t + 0.1423);
    } else if (t >= (0.1 * period) && t < (0.3 * period)) {
        vel = (-4.32e+06 * std::pow(t, 8) + 7.94e+06 * std::pow(t, 7) - 6.298e+06 * std::pow(t, 6) +
               2.82e+06 * std::pow(t, 5) - 7.809e+05 * std::pow(t, 4) + 1.37e+05 * std::pow(t, 3) -
               1.484e+04 * std::pow(t, 2) + 901.9 * t - 22.87);
    } else if (t >= (0.3 * period) && t < (0.9 * period)) {
        vel = (-5.123 * std::pow(t, 4) + 14.15 * std::pow(t, 3) - 14.21 * std::pow(t, 2) + 5.865 * t - 0.5696);
    }
    return vel;
}

extern "C" void orc_lattice(int q, int32_t* c, double* w, int32_t* opp) {
    Lat L = lattice(q);
    for (int d = 0; d < L.q; d++) {
        for (int a = 0; a < 3; a++) c[3 * d + a] = L.c[d][a];
        w[d] = L.w[d];
        opp[d] = L.opp[d];
    }
}
