// k_setup.hip — set-up kernels for gfx950: ray-cast voxelisation, boundary detection,
// next-to-wall marking and boundary normals.

#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "lbm_device.hpp"

namespace pvlbm {

namespace {

// Geometry::isPointInsideTriangle (ComplexGeometry.cpp:620-674), same operation order.
__device__ __forceinline__ bool point_in_triangle(double x, double y, double z, const double* T) {
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
    if (fabs(b) < tol) return false;
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
__device__ __forceinline__ bool crosses(const double* T, const double s[3], const double e[3]) {
    const double threshold = 1e-16;
    double px = s[0], py = s[1], pz = s[2];
    double dx = e[0] - s[0], dy = e[1] - s[1], dz = e[2] - s[2];
    double nx = T[9], ny = T[10], nz = T[11];
    double qx = T[0], qy = T[1], qz = T[2];
    double numerator = (qx - px) * nx + (qy - py) * ny + (qz - pz) * nz;
    double denominator = dx * nx + dy * ny + dz * nz;
    if (fabs(denominator) < threshold) return false;
    double t = numerator / denominator;
    if (t >= 0.0 && t <= 1.0) {
        double ix = px + t * dx;
        double iy = py + t * dy;
        double iz = pz + t * dz;
        if (point_in_triangle(ix, iy, iz, T)) return true;
    }
    return false;
}

// Geometry::FindLineTriangleIntersection (ComplexGeometry.cpp:689-730), same operation order:
// plane of the file normal through vertex 1, |denominator| >= 1e-7, 0 <= t <= 1, then the
// point-in-triangle test of the voxeliser.
__device__ __forceinline__ bool line_triangle(const double* T, const double* s, const double* e, double ip[3]) {
    const double threshold = 1e-7;
    double px = s[0], py = s[1], pz = s[2];
    double dx = e[0] - s[0], dy = e[1] - s[1], dz = e[2] - s[2];
    double nx = T[9], ny = T[10], nz = T[11];
    double qx = T[0], qy = T[1], qz = T[2];
    double denom = (dx * nx + dy * ny + dz * nz);
    if (fabs(denom) < threshold) return false;
    double t = ((qx - px) * nx + (qy - py) * ny + (qz - pz) * nz) / denom;
    if (t < 0.0 || t > 1.0) return false;
    double ix = px + t * dx;
    double iy = py + t * dy;
    double iz = pz + t * dz;
    if (point_in_triangle(ix, iy, iz, T)) {
        ip[0] = ix;
        ip[1] = iy;
        ip[2] = iz;
        return true;
    }
    return false;
}

// One wave per (i, j) column; its lanes walk the column's points in z.  The candidate triangles
// of a column are the same for every lane, so their loads are wave-uniform.
__global__ __launch_bounds__(64) void k_voxelize(const VoxParams p) {
    const int64_t col = blockIdx.x;
    const int64_t i = col % p.nx, j = col / p.nx;
    const int64_t c0 = p.col_off[col], c1 = p.col_off[col + 1];
    const double rayOffset = 0.02;
    for (int64_t k = threadIdx.x; k < p.nz; k += 64) {
        // point = min + idx*spacing (ComplexGeometry.cpp:337-340); ray of labelPoints (:989-1003)
        // and adjustLineCoordinates case 0 (:855-884)
        double x = p.ox + (double)i * p.h;
        double y = p.oy + (double)j * p.h;
        double z = p.oz + (double)k * p.h;
        double s[3] = {x, y, z + p.h / 2};
        double rayLength = p.bb[5] - p.bb[4];
        double e[3] = {x, y, z + rayLength + 0.1};
        if (s[0] == p.bb[0]) { s[0] += rayOffset; e[0] += rayOffset; }
        if (s[0] == p.bb[1]) { s[0] -= rayOffset; e[0] -= rayOffset; }
        if (s[1] == p.bb[2]) { s[1] += rayOffset; e[1] += rayOffset; }
        if (s[1] == p.bb[3]) { s[1] -= rayOffset; e[1] -= rayOffset; }
        if (s[2] == p.bb[4]) { s[2] += rayOffset; }
        if (e[2] == p.bb[5]) { e[2] -= rayOffset; }
        int count = 0;
        for (int64_t c = c0; c < c1; c++) {
            const double* T = p.tri + 12 * (int64_t)p.col_tri[c];
            if (crosses(T, s, e)) count++;
        }
        p.fluid[i + (int64_t)p.nx * (j + (int64_t)p.ny * k)] = (count % 2 == 1) ? 1 : 0;
    }
}

// findBoundaryPoints (ComplexGeometry.cpp:1259-1300) over the Q directions: a fluid node is a
// boundary node when a Q neighbour is outside the box or not fluid.
template <int Q>
__global__ __launch_bounds__(256) void k_boundary(const uint8_t* fluid, int64_t nxp, int64_t nyp,
                                                 int64_t nx, int64_t ny, int64_t nz, uint8_t* bflag) {
    const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t total = nxp * nyp * nz;
    if (n >= total) return;
    const int64_t i = n % nxp, j = (n / nxp) % nyp, k = n / (nxp * nyp);
    if (!fluid[n]) {
        bflag[n] = 0;
        return;
    }
    uint32_t m = 0;
#pragma unroll
    for (int d = 0; d < Q; d++) {
        const int64_t a = i + Lat<Q>::cx[d], b = j + Lat<Q>::cy[d], c = k + Lat<Q>::cz[d];
        const bool inside = a >= 0 && a < nx && b >= 0 && b < ny && c >= 0 && c < nz;
        if (!inside || !fluid[a + nxp * (b + nyp * c)]) m |= 1u << d;
    }
    bflag[n] = m != 0;
}

// IdentifyNextToWallLattices (Subdomain.cpp:201-218): interior fluid nodes linked to a wall node.
// Interior nodes have all Q neighbours inside the box.
template <int Q>
__global__ __launch_bounds__(256) void k_ntw(uint8_t* kind, int64_t nxp, int64_t nyp, int64_t nz) {
    const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (n >= nxp * nyp * nz) return;
    if (kind[n] != K_FLUID) return;
    bool near_wall = false;
#pragma unroll
    for (int d = 0; d < Q; d++) {
        const int64_t o = Lat<Q>::cx[d] + nxp * (Lat<Q>::cy[d] + nyp * Lat<Q>::cz[d]);
        near_wall |= kind[n + o] == K_WALL;
    }
    if (near_wall) kind[n] = K_NTW;
}

// Per-triangle quantities of calculateBoundaryNormals (Subdomain.cpp:318-341): unit
// cross-product normal, area, centroid.  prep planes: nx, ny, nz, area, cx, cy, cz.
__global__ __launch_bounds__(256) void k_tri_prep(const double* tri, int64_t ntri, double* prep) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= ntri) return;
    const double* T = tri + 12 * t;
    double ux = T[3] - T[0], uy = T[4] - T[1], uz = T[5] - T[2];
    double vx = T[6] - T[0], vy = T[7] - T[1], vz = T[8] - T[2];
    double nx = uy * vz - uz * vy;
    double ny = uz * vx - ux * vz;
    double nz = ux * vy - uy * vx;
    double norm = sqrt(nx * nx + ny * ny + nz * nz);
    if (norm > 0) {
        nx /= norm;
        ny /= norm;
        nz /= norm;
    }
    prep[t] = nx;
    prep[ntri + t] = ny;
    prep[2 * ntri + t] = nz;
    prep[3 * ntri + t] = 0.5 * norm;
    prep[4 * ntri + t] = (T[0] + T[3] + T[6]) / 3.0;
    prep[5 * ntri + t] = (T[1] + T[4] + T[7]) / 3.0;
    prep[6 * ntri + t] = (T[2] + T[5] + T[8]) / 3.0;
}

// Normal of each listed node: sum over all triangles of unit normal * area / distance^2, then
// normalised (Subdomain.cpp:343-368).  One thread per node, summing in triangle file order as the
// reference does; the block's threads stage the per-triangle quantities in LDS, 256 triangles at a
// time, so each is read from memory once per block instead of once per node.
__global__ __launch_bounds__(256) void k_normals(const double* prep, int64_t ntri, const double* xyz,
                                                int64_t n, double* nrm) {
    constexpr int CH = 256;
    __shared__ double P[7][CH];
    const int64_t w = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool act = w < n;
    const double x = act ? xyz[w] : 0.0, y = act ? xyz[n + w] : 0.0, z = act ? xyz[2 * n + w] : 0.0;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int64_t t0 = 0; t0 < ntri; t0 += CH) {
        const int64_t t = t0 + threadIdx.x;
        __syncthreads();  // the previous chunk is consumed
        if (t < ntri)
            for (int c = 0; c < 7; c++) P[c][threadIdx.x] = prep[c * ntri + t];
        __syncthreads();
        const int m = (int)(ntri - t0 < CH ? ntri - t0 : CH);
#pragma unroll 4
        for (int k = 0; k < m; k++) {
            const double cX = P[4][k], cY = P[5][k], cZ = P[6][k];
            const double distance = sqrt((x - cX) * (x - cX) + (y - cY) * (y - cY) + (z - cZ) * (z - cZ));
            const double weight = P[3][k] / (distance * distance);
            a0 += P[0][k] * weight;
            a1 += P[1][k] * weight;
            a2 += P[2][k] * weight;
        }
    }
    if (!act) return;
    double n0 = a0, n1 = a1, n2 = a2;
    const double norm = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
    if (norm > 0) {
        n0 /= norm;
        n1 /= norm;
        n2 /= norm;
    }
    nrm[w] = n0;
    nrm[n + w] = n1;
    nrm[2 * n + w] = n2;
}

}  // namespace

hipError_t launch_voxelize(const VoxParams& p, hipStream_t s) {
    const int64_t cols = (int64_t)p.nx * p.ny;
    if (cols <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_voxelize, dim3((unsigned)cols), dim3(64), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_boundary(int q, const uint8_t* fluid_pad, int64_t nxp, int64_t nyp, int64_t nx,
                           int64_t ny, int64_t nz, uint8_t* bflag, hipStream_t s) {
    const int64_t total = nxp * nyp * nz;
    const unsigned g = (unsigned)((total + 255) / 256);
    if (q == 19) hipLaunchKernelGGL(k_boundary<19>, dim3(g), dim3(256), 0, s, fluid_pad, nxp, nyp, nx, ny, nz, bflag);
    else hipLaunchKernelGGL(k_boundary<27>, dim3(g), dim3(256), 0, s, fluid_pad, nxp, nyp, nx, ny, nz, bflag);
    return hipGetLastError();
}

hipError_t launch_ntw(int q, uint8_t* kind, int64_t nxp, int64_t nyp, int64_t nz, hipStream_t s) {
    const int64_t total = nxp * nyp * nz;
    const unsigned g = (unsigned)((total + 255) / 256);
    if (q == 19) hipLaunchKernelGGL(k_ntw<19>, dim3(g), dim3(256), 0, s, kind, nxp, nyp, nz);
    else hipLaunchKernelGGL(k_ntw<27>, dim3(g), dim3(256), 0, s, kind, nxp, nyp, nz);
    return hipGetLastError();
}

hipError_t launch_tri_prep(const double* tri, int64_t ntri, double* prep, hipStream_t s) {
    if (ntri <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tri_prep, dim3((unsigned)((ntri + 255) / 256)), dim3(256), 0, s, tri, ntri, prep);
    return hipGetLastError();
}

hipError_t launch_normals(const double* prep, int64_t ntri, const double* xyz, int64_t n, double* nrm,
                          hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_normals, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, prep, ntri, xyz, n, nrm);
    return hipGetLastError();
}

}  // namespace pvlbm

namespace pvlbm {

namespace {

// Streaming copy at the attainable HBM rate, used to calibrate HBM traffic counters against a
// kernel whose bytes are known exactly and as the copy ceiling the benchmark reports: each block
// moves one 16 KiB chunk, each thread four 16-byte vectors 4 KiB apart (four loads in flight,
// then four nontemporal stores); a partial last chunk is copied vector by vector.
typedef double dvec2 __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(256) void k_copy(const dvec2* __restrict__ src, dvec2* __restrict__ dst, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 1024 + threadIdx.x;
    if (i + 768 < n) {
        const dvec2 v0 = src[i], v1 = src[i + 256], v2 = src[i + 512], v3 = src[i + 768];
        __builtin_nontemporal_store(v0, dst + i);
        __builtin_nontemporal_store(v1, dst + i + 256);
        __builtin_nontemporal_store(v2, dst + i + 512);
        __builtin_nontemporal_store(v3, dst + i + 768);
    } else {
        for (int k = 0; k < 4; k++)
            if (i + 256 * k < n) dst[i + 256 * k] = src[i + 256 * k];
    }
}

}  // namespace

// link fractions: one thread per segment, candidates from the cells its bounding box touches
__global__ void k_link_q(const double* seg, int64_t npairs, const double* tri, const int32_t* cell_start,
                         const int32_t* cell_tri, double ogx, double ogy, double ogz, double cs, int ncx, int ncy,
                         int ncz, double h, int nearest, double* q) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    const double* s = seg + 6 * p;
    const double* e = s + 3;
    const double og[3] = {ogx, ogy, ogz};
    const int nc[3] = {ncx, ncy, ncz};
    int c0[3], c1[3];
    for (int a = 0; a < 3; a++) {
        const double lo = fmin(s[a], e[a]) - 1e-9 * cs, hi = fmax(s[a], e[a]) + 1e-9 * cs;
        c0[a] = min(max((int)floor((lo - og[a]) / cs), 0), nc[a] - 1);
        c1[a] = min(max((int)floor((hi - og[a]) / cs), 0), nc[a] - 1);
    }
    // nearest == 0: the triangle of smallest index (the reference's first in file order);
    // nearest == 1: the intersection closest to the start (the wall the link meets first)
    int32_t best = INT32_MAX;
    double bestd = 1e300;
    double ip[3] = {0.0, 0.0, 0.0};
    for (int cz = c0[2]; cz <= c1[2]; cz++)
        for (int cy = c0[1]; cy <= c1[1]; cy++)
            for (int cx = c0[0]; cx <= c1[0]; cx++) {
                const int64_t c = ((int64_t)cz * nc[1] + cy) * nc[0] + cx;
                for (int32_t k = cell_start[c]; k < cell_start[c + 1]; k++) {
                    const int32_t t = cell_tri[k];
                    if (!nearest && t >= best) break;  // ascending within a cell
                    double hit[3];
                    if (line_triangle(tri + 12 * (int64_t)t, s, e, hit)) {
                        const double dd = (hit[0] - s[0]) * (hit[0] - s[0]) + (hit[1] - s[1]) * (hit[1] - s[1]) +
                                          (hit[2] - s[2]) * (hit[2] - s[2]);
                        if (nearest && (dd > bestd || (dd == bestd && t > best))) continue;
                        best = t;
                        bestd = dd;
                        ip[0] = hit[0];
                        ip[1] = hit[1];
                        ip[2] = hit[2];
                        if (!nearest) break;
                    }
                }
            }
    if (best == INT32_MAX) {
        q[p] = -1.0;
        return;
    }
    const double d = sqrt((ip[0] - s[0]) * (ip[0] - s[0]) + (ip[1] - s[1]) * (ip[1] - s[1]) +
                          (ip[2] - s[2]) * (ip[2] - s[2]));
    q[p] = d / h;
}

hipError_t launch_link_q(const double* seg, int64_t npairs, const double* tri, const int32_t* cell_start,
                         const int32_t* cell_tri, const double og[3], double cs, const int32_t nc[3], double h,
                         int nearest, double* q, hipStream_t s) {
    if (npairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_link_q, dim3((unsigned)((npairs + 255) / 256)), dim3(256), 0, s, seg, npairs, tri,
                       cell_start, cell_tri, og[0], og[1], og[2], cs, nc[0], nc[1], nc[2], h, nearest, q);
    return hipGetLastError();
}

// coarse-field expansion for starts from a network prediction
__device__ __forceinline__ float coarse_coord(int64_t i, int64_t n, int c, int& i0, int& i1) {
    float x = ((float)i + 0.5f) * (float)c / (float)n - 0.5f;
    x = fminf(fmaxf(x, 0.f), (float)(c - 1));
    i0 = (int)floorf(x);
    i1 = min(i0 + 1, c - 1);
    return x - (float)i0;
}

__global__ void k_expand_coarse(const float* rho_c, const float* u_c, int cx, int cy, int cz, int64_t nx,
                                int64_t ny, int64_t nz, int64_t nxp, int64_t nyp, int64_t ns, double rlo,
                                double rhi, double* out, unsigned long long* nonfin) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= ns) return;
    const int64_t x = n % nxp, pr = n / nxp, y = pr % nyp, z = pr / nyp;
    double r = 1.0, u0 = 0.0, u1 = 0.0, u2 = 0.0;
    if (x < nx && y < ny && z < nz) {
        int x0, x1, y0, y1, z0, z1;
        const float fx = coarse_coord(x, nx, cx, x0, x1);
        const float fy = coarse_coord(y, ny, cy, y0, y1);
        const float fz = coarse_coord(z, nz, cz, z0, z1);
        const int64_t plane = (int64_t)cx * cy * cz;
        float v[4];
        for (int c = 0; c < 4; c++) {
            const float* f = c == 0 ? rho_c : u_c + (c - 1) * plane;
            auto at = [&](int zz, int yy, int xx) { return f[((int64_t)zz * cy + yy) * cx + xx]; };
            const float a = at(z0, y0, x0) * (1.f - fx) + at(z0, y0, x1) * fx;
            const float b = at(z0, y1, x0) * (1.f - fx) + at(z0, y1, x1) * fx;
            const float cc = at(z1, y0, x0) * (1.f - fx) + at(z1, y0, x1) * fx;
            const float d = at(z1, y1, x0) * (1.f - fx) + at(z1, y1, x1) * fx;
            v[c] = (a * (1.f - fy) + b * fy) * (1.f - fz) + (cc * (1.f - fy) + d * fy) * fz;
        }
        if (isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]) && isfinite(v[3])) {
            r = fmin(fmax((double)v[0], rlo), rhi);
            u0 = v[1];
            u1 = v[2];
            u2 = v[3];
        } else {
            atomicAdd(nonfin, 1ull);
        }
    }
    out[n] = r;
    out[ns + n] = u0;
    out[2 * ns + n] = u1;
    out[3 * ns + n] = u2;
}

// flux of u_x through each plane x over the fluid nodes: one block per plane, fixed-order sums
__global__ __launch_bounds__(256) void k_slab_flux(const double* ux, const uint8_t* kind, int64_t ny, int64_t nz,
                                                   int64_t nxp, int64_t nyp, double* flux) {
    __shared__ double sh[256];
    const int64_t x = blockIdx.x;
    double sum = 0.0;
    for (int64_t r = threadIdx.x; r < ny * nz; r += 256) {
        const int64_t n = x + nxp * ((r % ny) + nyp * (r / ny));
        if (kind[n] != 0) sum += ux[n];
    }
    sh[threadIdx.x] = sum;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) flux[x] = sh[0];
}

// velocity of plane x scaled by target / flux[x] (limited to [1/4, 4]; planes without flux keep
// theirs), then speeds clipped to umax
__global__ void k_scale_clip(double* out, int64_t ns, int64_t nx, int64_t nxp, const double* flux, double target,
                             const double* targets, const double* drho, double rlo, double rhi, double umax) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= ns) return;
    const int64_t x = n % nxp;
    if (x < nx && targets) target = targets[x];
    double f = 1.0;
    if (flux && target > 0.0 && x < nx && flux[x] > 1e-12 * target) f = fmin(fmax(target / flux[x], 0.25), 4.0);
    if (drho && x < nx) out[n] = fmin(fmax(out[n] + drho[x], rlo), rhi);
    double u0 = out[ns + n] * f, u1 = out[2 * ns + n] * f, u2 = out[3 * ns + n] * f;
    const double sp = sqrt(u0 * u0 + u1 * u1 + u2 * u2);
    if (sp > umax) {
        const double c = umax / sp;
        u0 *= c;
        u1 *= c;
        u2 *= c;
    }
    out[ns + n] = u0;
    out[2 * ns + n] = u1;
    out[3 * ns + n] = u2;
}

hipError_t launch_expand_coarse(const float* rho_c, const float* u_c, int cx, int cy, int cz, int64_t nx,
                                int64_t ny, int64_t nz, int64_t nxp, int64_t nyp, int64_t ns, double umax,
                                double rlo, double rhi, const uint8_t* kind, double flux_x, const double* targets,
                                const double* drho, double* flux, double* out, unsigned long long* nonfin,
                                hipStream_t s) {
    const int64_t blocks = (ns + 255) / 256;
    hipLaunchKernelGGL(k_expand_coarse, dim3((unsigned)blocks), dim3(256), 0, s, rho_c, u_c, cx, cy, cz, nx, ny,
                       nz, nxp, nyp, ns, rlo, rhi, out, nonfin);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const bool scale = flux_x > 0.0 || targets;
    if (scale) {
        hipLaunchKernelGGL(k_slab_flux, dim3((unsigned)nx), dim3(256), 0, s, out + ns, kind, ny, nz, nxp, nyp, flux);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_scale_clip, dim3((unsigned)blocks), dim3(256), 0, s, out, ns, nx, nxp,
                       scale ? flux : nullptr, flux_x, targets, drho, rlo, rhi, umax);
    return hipGetLastError();
}

// one thread per box point; reads the padded slot of the same point
__global__ void k_macros_f32(const double* macro, const uint8_t* kind, int64_t nx, int64_t ny, int64_t nbox,
                             int64_t nxp, int64_t nyp, int64_t ns, float* rho, float* u) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    const int64_t i = b % nx, r = b / nx;
    const int64_t n = (r % ny + nyp * (r / ny)) * nxp + i;
    const bool node = kind[n] != 0;
    rho[b] = node ? (float)macro[n] : 1.0f;
    for (int c = 0; c < 3; c++) u[c * nbox + b] = node ? (float)macro[(c + 1) * ns + n] : 0.0f;
}

hipError_t launch_macros_f32(const double* macro, const uint8_t* kind, int64_t nx, int64_t ny, int64_t nz,
                             int64_t nxp, int64_t nyp, int64_t ns, float* rho, float* u, hipStream_t s) {
    const int64_t nbox = nx * ny * nz;
    if (nbox <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_macros_f32, dim3((unsigned)((nbox + 255) / 256)), dim3(256), 0, s, macro, kind, nx, ny,
                       nbox, nxp, nyp, ns, rho, u);
    return hipGetLastError();
}

// one thread per output point of a separable index grid
__global__ void k_sample_kinds(const uint8_t* kind, int64_t nxp, int64_t nyp, const int32_t* ix, int nx,
                               const int32_t* iy, int ny, const int32_t* iz, int nz, uint8_t* out) {
    const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= (int64_t)nx * ny * nz) return;
    const int i = (int)(o % nx);
    const int64_t r = o / nx;
    const int j = (int)(r % ny), k = (int)(r / ny);
    out[o] = kind[ix[i] + nxp * (iy[j] + nyp * (int64_t)iz[k])];
}

hipError_t launch_sample_kinds(const uint8_t* kind, int64_t nxp, int64_t nyp, const int32_t* ix, int nx,
                               const int32_t* iy, int ny, const int32_t* iz, int nz, uint8_t* out, hipStream_t s) {
    const int64_t n = (int64_t)nx * ny * nz;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sample_kinds, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, kind, nxp, nyp, ix, nx,
                       iy, ny, iz, nz, out);
    return hipGetLastError();
}

hipError_t launch_copy(const void* src, void* dst, int64_t bytes, hipStream_t s) {
    const int64_t n = bytes / 16;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy, dim3((unsigned)((n + 1023) / 1024)), dim3(256), 0, s, static_cast<const dvec2*>(src),
                       static_cast<dvec2*>(dst), n);
    return hipGetLastError();
}

}  // namespace pvlbm
