// k_step.hip — the fused lattice Boltzmann step for gfx950.
//
// One launch = the streaming and boundary closures of reference iteration n-1 followed by the
// collision of iteration n (Simulation::Iterate body, Simulation.cpp:577-586, regrouped): every
// lattice node pulls its populations from the post-collision lattice A, boundary nodes take the
// populations that have no upstream node from their stale table P (the reference keeps the old
// value there, Simulation.cpp:218-224), apply their closure, next-to-wall nodes evaluate WSS
// against the equilibrium of their previous macros M (Simulation.cpp:880-898), and every node
// collides (Simulation.cpp:43-208) into B.
//
// Layout: populations are SoA planes over a dense padded box (x padded to a multiple of 64), so
// a wave of 64 lanes handles one 64-node x-segment ("tile") and each direction's gather is one
// contiguous 512-byte (FP64) or 128-byte (16-bit) run shifted by the stencil offset.  Only
// tiles holding at least one lattice node are launched, and the block -> tile map keeps runs of
// consecutive tiles on one XCD so the partial cache lines at tile ends stay in one L2.

#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "lbm_device.hpp"

namespace pvlbm {

namespace {

// Storage traits: plain FP64, or the shifted population f - w in FP16 / BF16 with FP32 math.
struct StFP64 {
    using S = double;
    using R = double;
    static constexpr bool kGate = true;
    __device__ static R load(S v, double) { return v; }
    __device__ static S store(R v, double) { return v; }
};

struct StFP16 {
    using S = _Float16;
    using R = float;
    static constexpr bool kGate = false;
    __device__ static R load(S v, double w) { return (float)v + (float)w; }
    __device__ static S store(R v, double w) { return (_Float16)(v - (float)w); }
};

struct StBF16 {
    using S = uint16_t;
    using R = float;
    static constexpr bool kGate = false;
    __device__ static R load(S v, double w) { return __uint_as_float((uint32_t)v << 16) + (float)w; }
    __device__ static S store(R v, double w) {
        float x = v - (float)w;
        uint32_t b = __float_as_uint(x);
        if ((b & 0x7f800000u) == 0x7f800000u && (b & 0x007fffffu)) return (uint16_t)((b >> 16) | 0x40u);
        b += 0x7fffu + ((b >> 16) & 1u);  // round to nearest even
        return (uint16_t)(b >> 16);
    }
};

__device__ __forceinline__ int xcd_remap(int b, int nb) {
    // Blocks are dispatched round-robin over the 8 XCDs; give each XCD one contiguous run of
    // logical blocks (consecutive tiles) instead of every eighth one.
    const int X = 8;
    const int q = nb / X, r = nb % X;
    const int x = b % X, i = b / X;
    return (x < r) ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Byte offset of node n inside a plane: the guard keeps n + guard - offset >= 0 for every stencil
// offset, so the per-lane part of every address is an unsigned 32-bit byte offset and the plane
// base stays scalar (global_load ... saddr).
template <class S>
__device__ __forceinline__ uint32_t node_bytes(int64_t n, int64_t guard) {
    return (uint32_t)((n + guard) * (int64_t)sizeof(S));
}

// All Q pulls issued back to back, unconditionally: for a boundary node the upstream of a missing
// link is outside the fluid (or the box) and reads a guard or solid slot, whose value is replaced
// from the stale table afterwards.
template <int Q, class ST, int D = 0>
__device__ __forceinline__ void load_all(const StepParams& p, uint32_t nb, typename ST::S (&v)[Q]) {
    if constexpr (D < Q) {
        using S = typename ST::S;
        constexpr int64_t cx = Lat<Q>::cx[D], cy = Lat<Q>::cy[D], cz = Lat<Q>::cz[D];
        const int64_t offb = (cx + cy * p.sy + cz * p.sz) * (int64_t)sizeof(S);
        const char* plane = static_cast<const char*>(p.A) + D * p.ns * (int64_t)sizeof(S);
        v[D] = *reinterpret_cast<const S*>(plane + (uint32_t)(nb - (uint32_t)offb));
        load_all<Q, ST, D + 1>(p, nb, v);
    }
}

template <int Q, class ST, int D = 0>
__device__ __forceinline__ void save_stale(const StepParams& p, uint32_t nmiss, int32_t slot,
                                           const typename ST::R (&f)[Q]) {
    if constexpr (D < Q) {
        using R = typename ST::R;
        if ((nmiss >> Lat<Q>::opp[D]) & 1u) static_cast<R*>(p.P)[D * p.pst + slot] = f[D];
        save_stale<Q, ST, D + 1>(p, nmiss, slot, f);
    }
}

template <int Q, class ST, int D = 0>
__device__ __forceinline__ void collide_store(const StepParams& p, uint32_t nb, typename ST::R rho,
                                              const typename ST::R (&u)[3],
                                              const typename ST::R (&f)[Q], typename ST::R omega) {
    if constexpr (D < Q) {
        using R = typename ST::R;
        using S = typename ST::S;
        const R feq = feq_dir<Q, D>(rho, u);
        const R fo = f[D] + omega * (feq - f[D]);
        char* plane = static_cast<char*>(p.B) + D * p.ns * (int64_t)sizeof(S);
        *reinterpret_cast<S*>(plane + nb) = ST::store(fo, Lat<Q>::w[D]);
        collide_store<Q, ST, D + 1>(p, nb, rho, u, f, omega);
    }
}

template <int Q, class ST, bool WSS>
__device__ __forceinline__ bool node_update(const StepParams& p, double (&red)[4]) {
    using R = typename ST::R;
    using S = typename ST::S;
    const int t = xcd_remap(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
    if (t >= p.ntiles) return false;
    const int64_t n = (int64_t)p.tiles[t] * 64 + (threadIdx.x & 63);
    const uint8_t kd = p.kind[n];
    if (kd == K_SOLID) return false;
    const bool bnd = kd >= K_INLET;
    const bool init = (p.flags & F_INIT) != 0;
    const uint32_t nb = node_bytes<S>(n, p.guard);
    S v[Q];
    // issue the pulls first; the side-table lookups below overlap with them
    if (!init) load_all<Q, ST>(p, nb, v);
    const int32_t slot = (kd >= K_NTW) ? p.aux[n] : -1;
    const uint32_t nmiss = bnd ? p.bmask[slot] : 0u;

    R f[Q];
    if (!init) {
#pragma unroll
        for (int d = 0; d < Q; d++) f[d] = ST::load(v[d], Lat<Q>::w[d]);
        if constexpr (ST::kGate) {
            // the reference streams only non-zero values (Simulation.cpp:222); a zero pull on an
            // existing link would keep a value this scheme does not store: count it
            bool gate = false;
#pragma unroll
            for (int d = 0; d < Q; d++) gate |= (v[d] == S(0)) && !((nmiss >> Lat<Q>::opp[d]) & 1u);
            if (gate) atomicAdd(p.gate, 1ull);
        }
        if (bnd) {
            const R* P = static_cast<const R*>(p.P);
            R pv[Q];
#pragma unroll
            for (int d = 0; d < Q; d++) pv[d] = P[d * p.pst + slot];
#pragma unroll
            for (int d = 0; d < Q; d++)
                if ((nmiss >> Lat<Q>::opp[d]) & 1u) f[d] = pv[d];
            const uint8_t cl = p.bclose[slot];
            if (cl == BC_WALL) {
                wall_rule<Q>(f, nmiss);
            } else if (cl == BC_INLET_XMIN) {
                inlet_xmin(f, R(p.scale * p.bprof[slot]));
            } else if (cl == BC_OUT_XMAX) {
                outlet_xmax(f, R(p.brho[slot]));
            } else if (cl == BC_OUT_YMAX) {
                outlet_ymax(f, R(p.brho[slot]));
            } else if (cl == BC_OUT_ZMIN) {
                outlet_zmin(f, R(p.brho[slot]));
            }
        }
    } else {
        if (p.init_rho) {
            R r0 = R(p.init_rho[n]);
            R u0[3] = {R(p.init_u[n]), R(p.init_u[p.ns + n]), R(p.init_u[2 * p.ns + n])};
            equilibrium<Q>(r0, u0, f);
        } else {
#pragma unroll
            for (int d = 0; d < Q; d++) f[d] = R(Lat<Q>::w[d]);
        }
    }
    if (bnd) save_stale<Q, ST>(p, nmiss, slot, f);

    R rho, u[3];
    moments(f, rho, u);

    if (kd == K_NTW) {
        R* M = static_cast<R*>(p.M);
        const int64_t ms = p.mst;
        if constexpr (WSS) {
            if (!init) {
                R mu[3] = {M[ms + slot], M[2 * ms + slot], M[3 * ms + slot]};
                R feq_old[Q];
                equilibrium<Q>(M[slot], mu, feq_old);
                R fne[Q];
#pragma unroll
                for (int d = 0; d < Q; d++) fne[d] = f[d] - feq_old[d];
                const R nv[3] = {R(p.nrm[slot]), R(p.nrm[ms + slot]), R(p.nrm[2 * ms + slot])};
                R s0 = R(0.0), s1 = R(0.0), s2 = R(0.0);
                wss_sum<Q, R, 0>(fne, nv, s0);
                wss_sum<Q, R, 1>(fne, nv, s1);
                wss_sum<Q, R, 2>(fne, nv, s2);
                const R C = R(p.C);
                const R t0 = C * s0, t1 = C * s1, t2 = C * s2;
                const R W = sqrt(t0 * t0 + t1 * t1 + t2 * t2);
                p.wss[slot] = (double)W;
                p.tau3[slot] = (double)t0;
                p.tau3[ms + slot] = (double)t1;
                p.tau3[2 * ms + slot] = (double)t2;
                p.acc[slot] += (double)t0;
                p.acc[ms + slot] += (double)t1;
                p.acc[2 * ms + slot] += (double)t2;
                p.acc[3 * ms + slot] += (double)W;
            }
        }
        M[slot] = rho;
        M[ms + slot] = u[0];
        M[2 * ms + slot] = u[1];
        M[3 * ms + slot] = u[2];
    }

    collide_store<Q, ST>(p, nb, rho, u, f, R(p.omega));

    if (p.flags & F_MACRO) {
        p.mrho[n] = (double)rho;
        p.mrho[p.ns + n] = (double)u[0];
        p.mrho[2 * p.ns + n] = (double)u[1];
        p.mrho[3 * p.ns + n] = (double)u[2];
    }
    if (p.flags & F_REDUCE) {
        const double r = (double)rho, jx = (double)rho * (double)u[0];
        red[0] = r;
        red[1] = (kd == K_INLET) ? jx : 0.0;
        red[2] = (kd == K_OUTLET) ? jx : 0.0;
        const bool bad = !isfinite(r) || !isfinite((double)u[0]) || !isfinite((double)u[1]) ||
                         !isfinite((double)u[2]);
        red[3] = bad ? 1.0 : 0.0;
    }
    return true;
}

template <int Q, class ST, bool WSS>
__global__ __launch_bounds__(256) void k_step(const StepParams p) {
    double red[4] = {0.0, 0.0, 0.0, 0.0};
    node_update<Q, ST, WSS>(p, red);
    if (p.flags & F_REDUCE) {
        __shared__ double part[4][4];
        const int wv = threadIdx.x >> 6;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const double s = wave_sum(red[k]);
            if ((threadIdx.x & 63) == 0) part[wv][k] = s;
        }
        __syncthreads();
        if (threadIdx.x < 4) {
            const int k = threadIdx.x;
            p.partials[4 * (int64_t)blockIdx.x + k] = ((part[0][k] + part[1][k]) + part[2][k]) + part[3][k];
        }
    }
}

// Deterministic final sum of the block partials: fixed strided order, fixed tree.
__global__ __launch_bounds__(256) void k_reduce_final(const double* partials, int nparts, double* out4) {
    __shared__ double s[4][256];
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < nparts; i += 256)
        for (int k = 0; k < 4; k++) a[k] += partials[4 * (int64_t)i + k];
    for (int k = 0; k < 4; k++) s[k][threadIdx.x] = a[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 4; k++) s[k][threadIdx.x] += s[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 4) out4[threadIdx.x] = s[threadIdx.x][0];
}

}  // namespace

template <int Q, class ST>
void launch_q(const StepParams& p, dim3 g, hipStream_t s) {
    if (p.flags & F_WSS) hipLaunchKernelGGL((k_step<Q, ST, true>), g, dim3(256), 0, s, p);
    else hipLaunchKernelGGL((k_step<Q, ST, false>), g, dim3(256), 0, s, p);
}

hipError_t launch_step(int q, int storage, const StepParams& p, int nblocks, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    dim3 g(nblocks);
    if (q == 19) {
        if (storage == 0) launch_q<19, StFP64>(p, g, s);
        else if (storage == 1) launch_q<19, StFP16>(p, g, s);
        else launch_q<19, StBF16>(p, g, s);
    } else {
        if (storage == 0) launch_q<27, StFP64>(p, g, s);
        else if (storage == 1) launch_q<27, StFP16>(p, g, s);
        else launch_q<27, StBF16>(p, g, s);
    }
    return hipGetLastError();
}

hipError_t launch_reduce_final(const double* partials, int nparts, double* out4, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, s, partials, nparts, out4);
    return hipGetLastError();
}

}  // namespace pvlbm
