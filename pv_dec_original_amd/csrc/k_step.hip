// k_step.hip — the fused lattice Boltzmann step for gfx950.
//
// One launch = the streaming and boundary closures of reference iteration n-1 followed by the
// collision of iteration n (Simulation::Iterate body, Simulation.cpp:577-586, regrouped).  Two
// schemes (enum Scheme): the pull from the post-collision lattice A into B, or the in-place AA
// pattern on one lattice, whose even launches read and write each node's own slots (directions
// reversed) and whose odd launches read from and write to the neighbours' slots; each node then
// owns the slots it touches, so no launch needs a second lattice, and the working set per step
// is half the pull's.  A link with no upstream node keeps its previous post-closure value (the
// reference does not stream into it, Simulation.cpp:218-224); that value is parked in the slot
// the next launch reads for that link, which in both schemes is a slot no other node touches,
// so every lane does the same branch-free gather.
// Boundary nodes then apply their closure (Zou-He or the wall rule); next-to-wall nodes evaluate
// WSS against the equilibrium of their previous macros M (Simulation.cpp:880-898).
//
// Work unit: one wave64 per 64-node x-segment ("tile") of the box, which is padded to at least
// one extra column in x (to a multiple of 64) and one extra row in y, so that no out-of-box
// neighbour position aliases a node.  A tile descriptor carries three 64-bit lane masks (lattice
// node, next-to-wall, boundary) and the first side-table slot of each kind, so a lane finds its
// slot with mbcnt and no per-lane lookup.  Each gather and each store is one contiguous 512-byte
// (FP64) / 128-byte (16-bit) run of a direction plane, shifted by the stencil offset or not; every
// plane has a guard before and after so that all addresses are in bounds.  An address is a 64-bit wave-uniform base
// (plane, stencil offset and tile) plus the lane's byte offset within the tile, so one VGPR serves
// every direction and a plane may exceed 4 GiB.

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

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// number of set bits of m below this lane
__device__ __forceinline__ int32_t lanes_below(uint64_t m) {
    return (int32_t)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Byte offset from a lattice base of direction D's slot for the node at byte index 0, by kind of
// launch: the slot it reads, the slot its relaxed population goes to, and the slot a link of D
// without an upstream node is parked in (the slot the next launch reads for D).
template <int Q, int SCH, int D>
__device__ __forceinline__ int64_t load_off(const StepParams& p) {
    if constexpr (SCH == SCH_PULL) return p.poff_pull[D];       // plane D at x - c_D
    else if constexpr (SCH == SCH_EVEN) return D * p.poff_plane;  // plane D at x
    else return p.poff_push[Lat<Q>::opp[D]];                     // plane opp D at x - c_D
}

template <int Q, int SCH, int D>
__device__ __forceinline__ int64_t store_off(const StepParams& p) {
    if constexpr (SCH == SCH_PULL) return D * p.poff_plane;                    // plane D at x
    else if constexpr (SCH == SCH_EVEN) return Lat<Q>::opp[D] * p.poff_plane;  // plane opp D at x
    else return p.poff_push[D];                                                // plane D at x + c_D
}

template <int Q, int SCH, int D>
__device__ __forceinline__ int64_t park_off(const StepParams& p) {
    if constexpr (SCH == SCH_PULL) return p.poff_pull[D];
    else if constexpr (SCH == SCH_EVEN) return load_off<Q, SCH_ODD, D>(p);
    else return load_off<Q, SCH_EVEN, D>(p);
}

// Addresses: `tb` is the byte index of the tile's first lane within a plane (wave-uniform, 64-bit)
// and `lo` the lane's byte offset within the tile (lane * sizeof(S)), so that each access is a
// 64-bit scalar base per direction plus the one 32-bit lane offset all directions share.
template <class S, class P>
__device__ __forceinline__ S* at(P* base, int64_t sbytes, uint32_t lo) {
    return reinterpret_cast<S*>(reinterpret_cast<typename std::conditional<std::is_const<P>::value, const char, char>::type*>(base) + sbytes + lo);
}

// All Q gathers issued back to back, before any store of the lane (in the in-place scheme a lane
// writes the slots it read).  NL: nontemporal loads (in place, for a lattice near or beyond the
// size of the last-level cache).
template <int Q, class ST, int SCH, bool NL, int D = 0>
__device__ __forceinline__ void load_all(const StepParams& p, int64_t tb, uint32_t lo, typename ST::S (&v)[Q]) {
    if constexpr (D < Q) {
        const typename ST::S* src = at<const typename ST::S>(p.A, load_off<Q, SCH, D>(p) + tb, lo);
        if constexpr (NL) v[D] = __builtin_nontemporal_load(src);
        else v[D] = *src;
        load_all<Q, ST, SCH, NL, D + 1>(p, tb, lo, v);
    }
}

// Stale populations live in the lattice itself.  In the pull scheme a link of node x with no
// upstream node reads slot x - c_D of plane D, a non-node slot (solid, or padding / guard outside
// the box) that only x ever addresses in that plane; in the in-place scheme an odd launch reads it
// from slot x - c_D of plane opp D (likewise only x's), an even launch from x's own slot of
// plane D, which no neighbour writes since x - c_D is not a node.  x parks its post-closure value
// there for the next launch.
template <int Q, class ST, int SCH, int D = 0>
__device__ __forceinline__ void save_stale(const StepParams& p, int64_t tb, uint32_t lo, uint32_t nmiss,
                                           const typename ST::R (&f)[Q]) {
    if constexpr (D < Q) {
        using S = typename ST::S;
        if ((nmiss >> Lat<Q>::opp[D]) & 1u)
            *at<S>(p.B, park_off<Q, SCH, D>(p) + tb, lo) = ST::store(f[D], Lat<Q>::w[D]);
        save_stale<Q, ST, SCH, D + 1>(p, tb, lo, nmiss, f);
    }
}

// Interpolated bounce-back on the wall links of a wall node (Bouzidi, Firdaouss and Lallemand
// 2001, linear): for each population D with no upstream node, I = opp D points at the wall and
// q = bq[slot][I] is where the link meets it.  f*(x) is this node's post-collision population of
// the previous launch, read from its own slot of A (pull scheme only); f*_I(x - c_I) is the
// population it pulled in direction I (valid when x - c_I, i.e. x + c_D, is a node).
template <int Q, class ST, int D = 0>
__device__ __forceinline__ void wall_interpolated(const StepParams& p, int64_t tb, uint32_t lo, uint32_t nmiss,
                                                  const float* bq, typename ST::R (&f)[Q]) {
    if constexpr (D < Q) {
        using R = typename ST::R;
        using S = typename ST::S;
        constexpr int I = Lat<Q>::opp[D];
        if ((nmiss >> I) & 1u) {
            const R aI = ST::load(*at<const S>(p.A, I * p.poff_plane + tb, lo), Lat<Q>::w[I]);
            const R q = R(bq[I]);
            R v = aI;  // no crossing found, or a thin wall: halfway bounce-back
            if (q >= R(0.5)) {
                const R aD = ST::load(*at<const S>(p.A, D * p.poff_plane + tb, lo), Lat<Q>::w[D]);
                v = aI / (R(2) * q) + (R(2) * q - R(1)) / (R(2) * q) * aD;
            } else if (q >= R(0) && !((nmiss >> D) & 1u)) {
                v = R(2) * q * aI + (R(1) - R(2) * q) * f[I];
            }
            f[D] = v;
        }
        wall_interpolated<Q, ST, D + 1>(p, tb, lo, nmiss, bq, f);
    }
}

template <int Q, class ST, bool NT, int SCH, int D = 0>
__device__ __forceinline__ void collide_store(const StepParams& p, int64_t tb, uint32_t lo, typename ST::R rho,
                                              const typename ST::R (&u)[3],
                                              const typename ST::R (&f)[Q], typename ST::R omega) {
    if constexpr (D < Q) {
        using R = typename ST::R;
        using S = typename ST::S;
        const R feq = feq_dir<Q, D>(rho, u);
        const R fo = f[D] + omega * (feq - f[D]);
        S* dst = at<S>(p.B, store_off<Q, SCH, D>(p) + tb, lo);
        if constexpr (NT) __builtin_nontemporal_store(ST::store(fo, Lat<Q>::w[D]), dst);
        else *dst = ST::store(fo, Lat<Q>::w[D]);
        collide_store<Q, ST, NT, SCH, D + 1>(p, tb, lo, rho, u, f, omega);
    }
}

template <int Q, class R>
__device__ __forceinline__ void init_populations(const StepParams& p, int64_t n, R (&f)[Q]) {
    if (p.init_rho) {
        R r0 = R(p.init_rho[n]);
        R u0[3] = {R(p.init_u[n]), R(p.init_u[p.ns + n]), R(p.init_u[2 * p.ns + n])};
        equilibrium<Q>(r0, u0, f);
    } else {
#pragma unroll
        for (int d = 0; d < Q; d++) f[d] = R(Lat<Q>::w[d]);
    }
}

// Collision and store fused with the WSS sums of Simulation::ComputeWallShearStress
// (Simulation.cpp:882-892): for each component I the reference sums, over directions A in index
// order, fne_A * sum_j(A, I) with fne_A = f_A - f_eq,A(M).  The three sums are carried through the
// store loop together (each still in its own A order, so every sum rounds exactly as the
// reference's loop nest does), and f_A dies once its relaxed value is stored, so no Q-long array
// stays live next to f.
template <int Q, class ST, bool NT, int SCH, int D = 0>
__device__ __forceinline__ void collide_store_wss(const StepParams& p, int64_t tb, uint32_t lo, typename ST::R rho,
                                                  const typename ST::R (&u)[3], const typename ST::R (&f)[Q],
                                                  typename ST::R omega, typename ST::R mrho,
                                                  const typename ST::R (&mu)[3], const typename ST::R (&n)[3],
                                                  typename ST::R& s0, typename ST::R& s1, typename ST::R& s2) {
    if constexpr (D < Q) {
        using R = typename ST::R;
        using S = typename ST::S;
        const R feq = feq_dir<Q, D>(rho, u);
        const R fo = f[D] + omega * (feq - f[D]);
        S* dst = at<S>(p.B, store_off<Q, SCH, D>(p) + tb, lo);
        if constexpr (NT) __builtin_nontemporal_store(ST::store(fo, Lat<Q>::w[D]), dst);
        else *dst = ST::store(fo, Lat<Q>::w[D]);
        // the rest direction has weight zero: its product is a signed zero, which leaves the
        // sums unchanged (see wss_weight)
        if constexpr (Lat<Q>::cx[D] != 0 || Lat<Q>::cy[D] != 0 || Lat<Q>::cz[D] != 0) {
            const R fne = f[D] - feq_dir<Q, D>(mrho, mu);
            s0 += fne * wss_weight<Q, R, 0, D>(n);
            s1 += fne * wss_weight<Q, R, 1, D>(n);
            s2 += fne * wss_weight<Q, R, 2, D>(n);
        }
        collide_store_wss<Q, ST, NT, SCH, D + 1>(p, tb, lo, rho, u, f, omega, mrho, mu, n, s0, s1, s2);
    }
}

// WI: the wall nodes take the interpolated closure (a separate instance, so that the default
// reference closure does not carry its registers); it reads the node's previous populations from
// its own slots, which exist only in the pull scheme.  SCH: the kind of launch (enum Scheme); an
// F_INIT launch of the in-place scheme is an even one without the gathers.
template <int Q, class ST, bool WSS, bool NT, bool WI, int SCH, bool NL>
__device__ __forceinline__ bool node_update(const StepParams& p, int blk, double (&red)[4]) {
    static_assert(!WI || SCH == SCH_PULL, "the interpolated closure needs the pull scheme");
    using R = typename ST::R;
    using S = typename ST::S;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int t = blk * 4 + wave;
    if (t >= p.ntiles) return false;
    const TileDesc T = p.tdesc[t];  // wave-uniform: scalar loads
    const int lane = threadIdx.x & 63;
    const uint64_t bit = 1ull << lane;
    if (!(T.node & bit)) return false;
    const int64_t n = (int64_t)T.tile * 64 + lane;
    const int64_t tb = ((int64_t)T.tile * 64 + p.guard) * (int64_t)sizeof(S);
    const uint32_t lo = (uint32_t)lane * (uint32_t)sizeof(S);
    const bool init = (p.flags & F_INIT) != 0;
    const bool bnd = (T.bnd & bit) != 0;
    const int32_t bslot = T.bslot + lanes_below(T.bnd);
    // A boundary node's word holds its missing links (bits 0..Q-1) and its closure (bits 28..31),
    // so a wall node needs no further load after the pulls.  In WSS launches the inlet profile,
    // outlet density and a next-to-wall node's previous macros, normal and accumulators are loaded
    // here too, before the pulls, so that the wave waits for memory once rather than again after
    // them (3-4% on the D3Q27 FP64 and 16-bit WSS launches, +0.9% on D3Q19 FP64).  In the other
    // launches profile and density are read after the pulls by the few inlet and outlet nodes.
    static_assert(Q <= 28, "the closure shares the boundary word with the link bits");
    const uint32_t bword = bnd ? p.bmask[bslot] : 0u;
    const uint32_t nmiss = bword & ((1u << Q) - 1u);
    const uint8_t bcl = (uint8_t)(bword >> 28);
    const double bprof = (WSS && bnd) ? p.bprof[bslot] : 0.0, brho = (WSS && bnd) ? p.brho[bslot] : 0.0;
    auto closure = [&]() -> uint8_t { return bcl; };
    auto profile = [&]() -> double { return WSS ? bprof : p.bprof[bslot]; };
    auto density = [&]() -> double { return WSS ? brho : p.brho[bslot]; };
    const bool ntw = (T.ntw & bit) != 0;
    const int32_t ws = T.wslot + lanes_below(T.ntw);
    R* M = static_cast<R*>(p.M);
    const int64_t ms = p.mst;
    R mr = R(1), mu[3] = {R(0), R(0), R(0)}, nv[3] = {R(0), R(0), R(0)};
    double pacc[4] = {0.0, 0.0, 0.0, 0.0};  // the FP64 accumulators, updated after the sums
    if (WSS && !init && ntw) {
        pacc[0] = p.acc[ws];
        pacc[1] = p.acc[ms + ws];
        pacc[2] = p.acc[2 * ms + ws];
        pacc[3] = p.acc[3 * ms + ws];
        mr = M[ws];
        mu[0] = M[ms + ws];
        mu[1] = M[2 * ms + ws];
        mu[2] = M[3 * ms + ws];
        nv[0] = R(p.nrm[ws]);
        nv[1] = R(p.nrm[ms + ws]);
        nv[2] = R(p.nrm[2 * ms + ws]);
    }
    uint8_t cl = 0;

    R f[Q];
    if (!init) {
        S v[Q];
        load_all<Q, ST, SCH, NL>(p, tb, lo, v);
#pragma unroll
        for (int d = 0; d < Q; d++) f[d] = ST::load(v[d], Lat<Q>::w[d]);
        if constexpr (ST::kGate) {
            // the reference streams only non-zero values (Simulation.cpp:222); a zero pulled over
            // an existing link would keep an old value this scheme does not store: count it
            bool gate = false;
#pragma unroll
            for (int d = 0; d < Q; d++) gate |= (v[d] == S(0)) && !((nmiss >> Lat<Q>::opp[d]) & 1u);
            if (gate) atomicAdd(p.gate, 1ull);
        }
        if (bnd) {
            // closures in the reference's call order (Simulation.cpp:582-584); one per node
            cl = closure();
            if (cl == BC_WALL) {
                if constexpr (WI)
                    wall_interpolated<Q, ST>(p, tb, lo, nmiss, p.bq + (int64_t)bslot * Q, f);
                else
                    wall_rule<Q>(f, nmiss);
            } else if (cl == BC_INLET_XMIN) {
                inlet_xmin(f, R(p.scale * profile()));
            } else if (cl == BC_OUT_XMAX) {
                outlet_xmax(f, R(density()));
            } else if (cl == BC_OUT_YMAX) {
                outlet_ymax(f, R(density()));
            } else if (cl == BC_OUT_ZMIN) {
                outlet_zmin(f, R(density()));
            }
        }
    } else {
        init_populations<Q>(p, n, f);
        if (bnd) cl = closure();
    }
    if (bnd) {
        if (init && p.init_rho) {
            // a start from given fields keeps the rest value w in the populations of links
            // without an upstream node: the reference starts from rest and never writes them
            // again, so any other value would move the steady state
            R fw[Q];
#pragma unroll
            for (int d = 0; d < Q; d++) fw[d] = R(Lat<Q>::w[d]);
            save_stale<Q, ST, SCH>(p, tb, lo, nmiss, fw);
        } else {
            save_stale<Q, ST, SCH>(p, tb, lo, nmiss, f);
        }
    }

    R rho, u[3];
    moments(f, rho, u);

    // In WSS launches a tile with a next-to-wall lane runs every lane through the fused pass;
    // lanes that are not next-to-wall take M = rest, n = 0 (set above) and discard the sums.
    // Tiles without one (a wave-uniform test) take the plain store.
    if (WSS && !init && T.ntw != 0) {
        R s0 = R(0.0), s1 = R(0.0), s2 = R(0.0);
        collide_store_wss<Q, ST, NT, SCH>(p, tb, lo, rho, u, f, R(p.omega), mr, mu, nv, s0, s1, s2);
        if (ntw) {
            const R C = R(p.C);
            const R t0 = C * s0, t1 = C * s1, t2 = C * s2;
            const R W = sqrt(t0 * t0 + t1 * t1 + t2 * t2);
            p.wss[ws] = (double)W;
            p.tau3[ws] = (double)t0;
            p.tau3[ms + ws] = (double)t1;
            p.tau3[2 * ms + ws] = (double)t2;
            p.acc[ws] = pacc[0] + (double)t0;
            p.acc[ms + ws] = pacc[1] + (double)t1;
            p.acc[2 * ms + ws] = pacc[2] + (double)t2;
            p.acc[3 * ms + ws] = pacc[3] + (double)W;
        }
    } else {
        collide_store<Q, ST, NT, SCH>(p, tb, lo, rho, u, f, R(p.omega));
    }
    if (ntw) {
        M[ws] = rho;
        M[ms + ws] = u[0];
        M[2 * ms + ws] = u[1];
        M[3 * ms + ws] = u[2];
    }

    if (p.flags & F_MACRO) {
        p.mrho[n] = (double)rho;
        p.mrho[p.ns + n] = (double)u[0];
        p.mrho[2 * p.ns + n] = (double)u[1];
        p.mrho[3 * p.ns + n] = (double)u[2];
    }
    if (p.flags & F_REDUCE) {
        const double r = (double)rho, jx = (double)rho * (double)u[0];
        red[0] = r;
        red[1] = (bnd && cl == BC_INLET_XMIN) ? jx : 0.0;
        red[2] = (bnd && cl >= BC_OUT_YMAX) ? jx : 0.0;
        const bool bad = !isfinite(r) || !isfinite((double)u[0]) || !isfinite((double)u[1]) ||
                         !isfinite((double)u[2]);
        red[3] = bad ? 1.0 : 0.0;
    }
    return true;
}

// Block b takes tiles 4b .. 4b+3, or with p.reverse block b takes the tiles of block
// blocks-1-b: workgroups start in index order, so the sweep runs from the other end of the
// lattice, while each group of four tiles (and so each block partial sum) stays the same.
template <int Q, class ST, bool WSS, bool NT, bool WI, int SCH, bool NL>
__global__ __launch_bounds__(256) void k_step(const StepParams p) {
    double red[4] = {0.0, 0.0, 0.0, 0.0};
    const int blk = p.reverse ? p.blocks - 1 - (int)blockIdx.x : (int)blockIdx.x;
    node_update<Q, ST, WSS, NT, WI, SCH, NL>(p, blk, red);
    if (p.flags & F_REDUCE) {
        // FP64 block sums of (mass, inlet flux, outlet flux, non-finite count)
        __shared__ double s[4][4];
        const int wv = threadIdx.x >> 6;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const double v = wave_sum(red[k]);
            if ((threadIdx.x & 63) == 0) s[wv][k] = v;
        }
        __syncthreads();
        if (threadIdx.x < 4) {
            const int k = threadIdx.x;
            p.partials[4 * (int64_t)blk + k] = ((s[0][k] + s[1][k]) + s[2][k]) + s[3][k];
        }
    }
}

// State moved between layouts, from p.A into p.B: node x's relaxed population of direction D from
// where a launch of kind KF stored it to where a launch of kind KT stores it, and each parked link
// likewise.  Only values move, so the fields are unchanged bit for bit.
template <int Q, class S, int KF, int KT, int D = 0>
__device__ __forceinline__ void relayout_dirs(const StepParams& p, int64_t tb, uint32_t lo, uint32_t nmiss) {
    if constexpr (D < Q) {
        *at<S>(p.B, store_off<Q, KT, D>(p) + tb, lo) = *at<const S>(p.A, store_off<Q, KF, D>(p) + tb, lo);
        if ((nmiss >> Lat<Q>::opp[D]) & 1u)
            *at<S>(p.B, park_off<Q, KT, D>(p) + tb, lo) = *at<const S>(p.A, park_off<Q, KF, D>(p) + tb, lo);
        relayout_dirs<Q, S, KF, KT, D + 1>(p, tb, lo, nmiss);
    }
}

template <int Q, class S, int KF, int KT>
__global__ __launch_bounds__(256) void k_relayout(const StepParams p) {
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int t = blockIdx.x * 4 + wave;
    if (t >= p.ntiles) return;
    const TileDesc T = p.tdesc[t];
    const int lane = threadIdx.x & 63;
    const uint64_t bit = 1ull << lane;
    if (!(T.node & bit)) return;
    const int64_t tb = ((int64_t)T.tile * 64 + p.guard) * (int64_t)sizeof(S);
    const uint32_t lo = (uint32_t)lane * (uint32_t)sizeof(S);
    const uint32_t nmiss = (T.bnd & bit) ? (p.bmask[T.bslot + lanes_below(T.bnd)] & ((1u << Q) - 1u)) : 0u;
    relayout_dirs<Q, S, KF, KT>(p, tb, lo, nmiss);
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

template <int Q, class ST, bool WI, int SCH>
void launch_qws(const StepParams& p, hipStream_t s) {
    const dim3 g(p.blocks), b(256);
    // nontemporal stores keep a lattice larger than the caches from evicting the gathers; a
    // per-step working set that fits the last-level cache is served from it with plain stores
    if (p.stream_stores) {
        if (p.flags & F_WSS) hipLaunchKernelGGL((k_step<Q, ST, true, true, WI, SCH, false>), g, b, 0, s, p);
        else hipLaunchKernelGGL((k_step<Q, ST, false, true, WI, SCH, false>), g, b, 0, s, p);
    } else if constexpr (SCH != SCH_PULL) {
        // in place, nontemporal loads when the host asks for them (by the lattice's size)
        if (p.stream_loads) {
            if (p.flags & F_WSS) hipLaunchKernelGGL((k_step<Q, ST, true, false, WI, SCH, true>), g, b, 0, s, p);
            else hipLaunchKernelGGL((k_step<Q, ST, false, false, WI, SCH, true>), g, b, 0, s, p);
        } else {
            if (p.flags & F_WSS) hipLaunchKernelGGL((k_step<Q, ST, true, false, WI, SCH, false>), g, b, 0, s, p);
            else hipLaunchKernelGGL((k_step<Q, ST, false, false, WI, SCH, false>), g, b, 0, s, p);
        }
    } else {
        if (p.flags & F_WSS) hipLaunchKernelGGL((k_step<Q, ST, true, false, WI, SCH, false>), g, b, 0, s, p);
        else hipLaunchKernelGGL((k_step<Q, ST, false, false, WI, SCH, false>), g, b, 0, s, p);
    }
}

template <int Q, class ST>
hipError_t launch_q(const StepParams& p, hipStream_t s) {
    if (p.bq) {
        if (p.scheme != SCH_PULL) return hipErrorInvalidValue;
        launch_qws<Q, ST, true, SCH_PULL>(p, s);
    } else if (p.scheme == SCH_PULL) {
        launch_qws<Q, ST, false, SCH_PULL>(p, s);
    } else if (p.scheme == SCH_EVEN) {
        launch_qws<Q, ST, false, SCH_EVEN>(p, s);
    } else if (p.scheme == SCH_ODD) {
        launch_qws<Q, ST, false, SCH_ODD>(p, s);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int Q, class S>
hipError_t relayout_q(const StepParams& p, int from, int to, hipStream_t s) {
    const dim3 g(p.blocks), b(256);
    if (from == SCH_PULL && to == SCH_EVEN) hipLaunchKernelGGL((k_relayout<Q, S, SCH_PULL, SCH_EVEN>), g, b, 0, s, p);
    else if (from == SCH_EVEN && to == SCH_PULL) hipLaunchKernelGGL((k_relayout<Q, S, SCH_EVEN, SCH_PULL>), g, b, 0, s, p);
    else if (from == SCH_ODD && to == SCH_PULL) hipLaunchKernelGGL((k_relayout<Q, S, SCH_ODD, SCH_PULL>), g, b, 0, s, p);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace

hipError_t launch_step(int q, int storage, const StepParams& p, hipStream_t s) {
    if (p.blocks <= 0) return hipSuccess;
    if (p.scheme != SCH_PULL && p.A != p.B) return hipErrorInvalidValue;  // in place: one lattice
    if (q == 19) {
        if (storage == 0) return launch_q<19, StFP64>(p, s);
        if (storage == 1) return launch_q<19, StFP16>(p, s);
        return launch_q<19, StBF16>(p, s);
    }
    if (storage == 0) return launch_q<27, StFP64>(p, s);
    if (storage == 1) return launch_q<27, StFP16>(p, s);
    return launch_q<27, StBF16>(p, s);
}

hipError_t launch_relayout(int q, int storage, const StepParams& p, int from, int to, hipStream_t s) {
    if (p.blocks <= 0) return hipSuccess;
    if (p.A == p.B) return hipErrorInvalidValue;
    if (q == 19) return storage == 0 ? relayout_q<19, double>(p, from, to, s) : relayout_q<19, uint16_t>(p, from, to, s);
    return storage == 0 ? relayout_q<27, double>(p, from, to, s) : relayout_q<27, uint16_t>(p, from, to, s);
}

hipError_t launch_reduce_final(const double* partials, int nparts, double* out4, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, s, partials, nparts, out4);
    return hipGetLastError();
}

}  // namespace pvlbm
