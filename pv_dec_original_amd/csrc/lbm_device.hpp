// lbm_device.hpp — lattice tables and per-node physics shared by the gfx950 kernels.
//
// Every function keeps the reference's floating-point operation order (the whole library is
// built with -ffp-contract=off), so that the FP64 instantiation rounds exactly as the reference's
// source order does.  R is the arithmetic type: double for the plain path, float for the
// shifted-storage path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pvlbm {

// Lattice tables: the reference's index order, which the closures depend on.
// D3Q19: LatticeModel.cpp:53-92 (rest = 18), opposites :108-151.
// D3Q27: LatticeModel.cpp:210-266 (rest = 0), opposites :274-334.
template <int Q>
struct Lat;

template <>
struct Lat<19> {
    static constexpr int cx[19] = {1, -1, 0, 0, 0, 0, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0, 0};
    static constexpr int cy[19] = {0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0};
    static constexpr int cz[19] = {0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, 0};
    static constexpr int opp[19] = {1, 0, 3, 2, 5, 4, 11, 10, 13, 12, 7, 6, 9, 8, 17, 16, 15, 14, 18};
    static constexpr double w[19] = {2.0 / 36.0, 2.0 / 36.0, 2.0 / 36.0, 2.0 / 36.0, 2.0 / 36.0,
                                     2.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                                     1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                                     1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 12.0 / 36.0};
};

template <>
struct Lat<27> {
    static constexpr int cx[27] = {0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, -1, 1,
                                   -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1};
    static constexpr int cy[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0,
                                   0, 1, 1, -1, -1, 1, -1, 1, -1, -1, 1, -1, 1};
    static constexpr int cz[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1,
                                   -1, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1, -1, 1};
    static constexpr int opp[27] = {0,  2,  1,  4,  3,  6,  5,  10, 9,  8,  7,  14, 13, 12,
                                    11, 18, 17, 16, 15, 20, 19, 22, 21, 24, 23, 26, 25};
    static constexpr double w[27] = {8.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,
                                     2.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,  1.0 / 54.0,
                                     1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,
                                     1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,
                                     1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 216.0,
                                     1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0,
                                     1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0};
};

// Closure ids carried per boundary-table slot (wall, then pvlbm_closure + 1).
enum BClose : uint8_t { BC_WALL = 0, BC_INLET_XMIN = 1, BC_OUT_YMAX = 2, BC_OUT_ZMIN = 3, BC_OUT_XMAX = 4 };

// Node kinds in the device kind array (values of enum pvlbm_kind).
enum Kind : uint8_t { K_SOLID = 0, K_FLUID = 1, K_NTW = 2, K_INLET = 3, K_OUTLET = 4, K_WALL = 5 };

// ---- moments: Simulation::Collision (Simulation.cpp:52-58, :69-105 D3Q19, :112-175 D3Q27) ----
template <class R>
__device__ __forceinline__ void moments(const R (&f)[19], R& rho, R (&u)[3]) {
    rho = R(0.0);
#pragma unroll
    for (int d = 0; d < 19; d++) rho += f[d];
    u[0] = ((f[0] + f[6] + f[7] + f[8] + f[9]) - (f[1] + f[10] + f[11] + f[12] + f[13])) / rho;
    u[1] = ((f[2] + f[6] + f[10] + f[14] + f[15]) - (f[3] + f[7] + f[11] + f[16] + f[17])) / rho;
    u[2] = ((f[4] + f[8] + f[12] + f[14] + f[16]) - (f[5] + f[9] + f[13] + f[15] + f[17])) / rho;
}

template <class R>
__device__ __forceinline__ void moments(const R (&f)[27], R& rho, R (&u)[3]) {
    rho = R(0.0);
#pragma unroll
    for (int d = 0; d < 27; d++) rho += f[d];
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

// ---- equilibrium of one direction, Simulation.cpp:186-196 ----
template <int Q, int D, class R>
__device__ __forceinline__ R feq_dir(R rho, const R (&u)[3]) {
    R du = u[0] * R(Lat<Q>::cx[D]) + u[1] * R(Lat<Q>::cy[D]) + u[2] * R(Lat<Q>::cz[D]);
    R u_sqr = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    return R(Lat<Q>::w[D]) * rho * (R(1.0) + R(3.0) * du + R(4.5) * du * du - R(1.5) * u_sqr);
}

template <int Q, class R, int D = 0>
__device__ __forceinline__ void equilibrium(R rho, const R (&u)[3], R (&feq)[Q]) {
    if constexpr (D < Q) {
        feq[D] = feq_dir<Q, D>(rho, u);
        equilibrium<Q, R, D + 1>(rho, u, feq);
    }
}

// ---- Zou-He closures (see oracle/pvoracle.cpp for the same algebra on the CPU) ----
// D3Q19 velocity inlet, x-min face: Boundary::ApplyBc(double&) (Boundary.cpp:129-157).
template <class R>
__device__ __forceinline__ void inlet_xmin(R (&m)[19], R ux) {
    R uin0 = ux, uin1 = R(0.0), uin2 = R(0.0);
    R rhoN = (m[2] + m[3] + m[4] + m[5] + m[14] + m[15] + m[16] + m[17] + m[18] +
              R(2) * (m[1] + m[10] + m[11] + m[12] + m[13])) /
             (R(1) - uin0);
    R Ny = R(1.0 / 2) * (m[2] + m[14] + m[15] - m[3] - m[16] - m[17]) - R(1.0 / 3) * (rhoN * uin1);
    R Nz = R(1.0 / 2) * (m[4] + m[16] + m[14] - m[5] - m[15] - m[17]) - R(1.0 / 3) * (rhoN * uin2);
    m[0] = m[1] + R(1 / 3.0) * rhoN * uin0;
    m[7] = m[10] + R(1 / 6.0) * rhoN * (uin0 - uin1) + Ny;
    m[6] = m[11] + R(1 / 6.0) * rhoN * (uin0 + uin1) - Ny;
    m[8] = m[13] + R(1 / 6.0) * rhoN * (uin0 + uin2) - Nz;
    m[9] = m[12] + R(1 / 6.0) * rhoN * (uin0 - uin2) + Nz;
}

// D3Q27 velocity inlet, x-min face: algebra of the reference's D3Q27 Boundary::ApplyBc
// (Boundary.cpp:87-120), u = (ux, 0, 0).
template <class R>
__device__ __forceinline__ void inlet_xmin(R (&m)[27], R ux) {
    R u0 = ux, u1 = R(0.0), u2 = R(0.0);
    R rho = ((m[0] + m[3] + m[4] + m[5] + m[6] + m[15] + m[16] + m[17] + m[18]) +
             R(2) * (m[2] + m[9] + m[10] + m[12] + m[14] + m[20] + m[22] + m[24] + m[26])) /
            (R(1) - u0);
    R Ny = R(1.0 / 6) * (m[3] + m[15] + m[16] - m[4] - m[17] - m[18]) - R(1.0 / 9) * (rho * u1);
    R Nz = R(1.0 / 6) * (m[5] + m[15] + m[17] - m[6] - m[16] - m[18]) - R(1.0 / 9) * (rho * u2);
    m[1] = m[2] + R(4 / 9.0) * rho * u0;
    m[7] = m[10] + R(1 / 9.0) * rho * (u0 + u1) - Ny;
    m[8] = m[9] + R(1 / 9.0) * rho * (u0 - u1) + Ny;
    m[11] = m[14] + R(1 / 9.0) * rho * (u0 + u2) - Nz;
    m[13] = m[12] + R(1 / 9.0) * rho * (u0 - u2) + Nz;
    m[19] = m[20] + R(1 / 36.0) * rho * (u0 + u1 + u2) - Ny - Nz;
    m[21] = m[22] + R(1 / 36.0) * rho * (u0 + u1 - u2) - Ny + Nz;
    m[23] = m[24] + R(1 / 36.0) * rho * (u0 - u1 + u2) + Ny - Nz;
    m[25] = m[26] + R(1 / 36.0) * rho * (u0 - u1 - u2) + Ny + Nz;
}

// D3Q19 pressure outlet, y-max face: ApplyOutletBc list 1 (Simulation.cpp:642-695).
template <class R>
__device__ __forceinline__ void outlet_ymax(R (&m)[19], R rho) {
    R vx = R(0.0), vy = R(0.0), vz = R(0.0);
    vy = R(-1) + (R(1.0) / rho) * (m[0] + m[1] + m[4] + m[5] + m[8] + m[9] + m[12] + m[13] + m[18] +
                                   R(2.0) * (m[2] + m[6] + m[10] + m[14] + m[15]));
    R Nyx = R(0.5) * (m[0] + m[8] + m[9] - (m[1] + m[12] + m[13])) - R(1.0 / 3.0) * rho * vx;
    R Nyz = R(0.5) * (m[4] + m[8] + m[12] - (m[5] + m[9] + m[13])) - R(1.0 / 3.0) * rho * vz;
    m[3] = m[2] - R(1.0 / 3.0) * rho * vy;
    m[11] = m[6] + R(1.0 / 6.0) * rho * (-vx - vy) + Nyx;
    m[7] = m[10] + R(1.0 / 6.0) * rho * (vx - vy) - Nyx;
    m[17] = m[14] + R(1.0 / 6.0) * rho * (-vy - vz) + Nyz;
    m[16] = m[15] + R(1.0 / 6.0) * rho * (-vy + vz) - Nyz;
}

// D3Q19 pressure outlet, z-min face: ApplyOutletBc lists 2 and 3 (Simulation.cpp:710-836).
template <class R>
__device__ __forceinline__ void outlet_zmin(R (&m)[19], R rho) {
    R vx = R(0.0), vy = R(0.0), vz = R(0.0);
    vz = R(1) - (R(1.0) / rho) * (m[0] + m[1] + m[2] + m[3] + m[6] + m[7] + m[10] + m[11] + m[18] +
                                  R(2.0) * (m[5] + m[9] + m[13] + m[15] + m[17]));
    R Nzx = R(0.5) * (m[0] + m[6] + m[7] - (m[1] + m[10] + m[11])) - R(1.0 / 3.0) * rho * vx;
    R Nzy = R(0.5) * (m[2] + m[6] + m[10] - (m[3] + m[7] + m[11])) - R(1.0 / 3.0) * rho * vy;
    m[4] = m[5] + R(1.0 / 3.0) * rho * vz;
    m[8] = m[13] + R(1.0 / 6.0) * rho * (vz + vx) - Nzx;
    m[12] = m[9] + R(1.0 / 6.0) * rho * (vz - vx) + Nzx;
    m[14] = m[17] + R(1.0 / 6.0) * rho * (vz + vy) - Nzy;
    m[16] = m[15] + R(1.0 / 6.0) * rho * (vz - vy) + Nzy;
}

// D3Q19 pressure outlet, x-max face: the y-max formula under the x<->y index permutation
// (0 2)(1 3)(7 10)(8 14)(9 15)(12 16)(13 17) with vx<->vy.  No reference formula exists.
template <class R>
__device__ __forceinline__ void outlet_xmax(R (&m)[19], R rho) {
    R vx = R(0.0), vy = R(0.0), vz = R(0.0);
    vx = R(-1) + (R(1.0) / rho) * (m[2] + m[3] + m[4] + m[5] + m[14] + m[15] + m[16] + m[17] + m[18] +
                                   R(2.0) * (m[0] + m[6] + m[7] + m[8] + m[9]));
    R Nxy = R(0.5) * (m[2] + m[14] + m[15] - (m[3] + m[16] + m[17])) - R(1.0 / 3.0) * rho * vy;
    R Nxz = R(0.5) * (m[4] + m[14] + m[16] - (m[5] + m[15] + m[17])) - R(1.0 / 3.0) * rho * vz;
    m[1] = m[0] - R(1.0 / 3.0) * rho * vx;
    m[11] = m[6] + R(1.0 / 6.0) * rho * (-vy - vx) + Nxy;
    m[10] = m[7] + R(1.0 / 6.0) * rho * (vy - vx) - Nxy;
    m[13] = m[8] + R(1.0 / 6.0) * rho * (-vx - vz) + Nxz;
    m[12] = m[9] + R(1.0 / 6.0) * rho * (-vx + vz) - Nxz;
}

// D3Q27 pressure outlet, x-max face: the D3Q27 inlet mirrored x -> -x with the density
// prescribed.  No reference formula exists.
template <class R>
__device__ __forceinline__ void outlet_xmax(R (&m)[27], R rho) {
    R u0 = R(0.0), u1 = R(0.0), u2 = R(0.0);
    u0 = R(-1) + ((m[0] + m[3] + m[4] + m[5] + m[6] + m[15] + m[16] + m[17] + m[18]) +
                  R(2) * (m[1] + m[7] + m[8] + m[11] + m[13] + m[19] + m[21] + m[23] + m[25])) /
                     rho;
    R Ny = R(1.0 / 6) * (m[3] + m[15] + m[16] - m[4] - m[17] - m[18]) - R(1.0 / 9) * (rho * u1);
    R Nz = R(1.0 / 6) * (m[5] + m[15] + m[17] - m[6] - m[16] - m[18]) - R(1.0 / 9) * (rho * u2);
    m[2] = m[1] + R(4 / 9.0) * rho * (-u0);
    m[9] = m[8] + R(1 / 9.0) * rho * (-u0 + u1) - Ny;
    m[10] = m[7] + R(1 / 9.0) * rho * (-u0 - u1) + Ny;
    m[12] = m[13] + R(1 / 9.0) * rho * (-u0 + u2) - Nz;
    m[14] = m[11] + R(1 / 9.0) * rho * (-u0 - u2) + Nz;
    m[26] = m[25] + R(1 / 36.0) * rho * (-u0 + u1 + u2) - Ny - Nz;
    m[24] = m[23] + R(1 / 36.0) * rho * (-u0 + u1 - u2) - Ny + Nz;
    m[22] = m[21] + R(1 / 36.0) * rho * (-u0 - u1 + u2) + Ny - Nz;
    m[20] = m[19] + R(1 / 36.0) * rho * (-u0 - u1 - u2) + Ny + Nz;
}

// D3Q27 has no y-max / z-min closure; the host rejects those faces for lattice 27.
template <class R>
__device__ __forceinline__ void outlet_ymax(R (&)[27], R) {}
template <class R>
__device__ __forceinline__ void outlet_zmin(R (&)[27], R) {}

// Wall rule, ApplyWallBc (Simulation.cpp:846-862): in direction order, f[opp(d)] = f[d] when the
// neighbour in direction d is missing; later directions see earlier assignments.
template <int Q, class R, int D = 0>
__device__ __forceinline__ void wall_rule(R (&m)[Q], uint32_t nmiss) {
    if constexpr (D < Q) {
        if ((nmiss >> D) & 1u) m[Lat<Q>::opp[D]] = m[D];
        wall_rule<Q, R, D + 1>(m, nmiss);
    }
}

// Wall shear stress of one node, Simulation::ComputeWallShearStress (Simulation.cpp:882-892):
// the weight sum_j(A, I) = sum_k c_A,k n_k (c_A,I - sum_l c_A,l n_l n_I) that multiplies the
// non-equilibrium population of direction A in component I, with the loop's operation order.
// Terms with a zero velocity component are left out: for a finite normal each is a signed zero,
// and adding a signed zero to these sums, which start at +0 and so never hold -0, leaves them
// unchanged, so the weight rounds exactly as the full loop's.
template <int Q, class R, int I, int A>
__device__ __forceinline__ R wss_weight(const R (&n)[3]) {
    constexpr int c0 = Lat<Q>::cx[A], c1 = Lat<Q>::cy[A], c2 = Lat<Q>::cz[A];
    constexpr int ci = (I == 0) ? c0 : (I == 1) ? c1 : c2;
    R sum_k = R(0.0);
    if constexpr (c0 != 0) sum_k += R(c0) * n[0] * n[I];
    if constexpr (c1 != 0) sum_k += R(c1) * n[1] * n[I];
    if constexpr (c2 != 0) sum_k += R(c2) * n[2] * n[I];
    R sum_j = R(0.0);
    if constexpr (c0 != 0) sum_j += R(c0) * n[0] * (R(ci) - sum_k);
    if constexpr (c1 != 0) sum_j += R(c1) * n[1] * (R(ci) - sum_k);
    if constexpr (c2 != 0) sum_j += R(c2) * n[2] * (R(ci) - sum_k);
    return sum_j;
}

}  // namespace pvlbm
