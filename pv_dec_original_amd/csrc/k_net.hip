// k_net.hip — kernels of the warm-start network and the C ABI of include/pvnet.h (gfx950).
//
// The convolutions are implicit GEMMs on the CDNA4 matrix cores (v_mfma_f32_16x16x32_bf16):
//   forward / input gradient   C[m][n] = sum_k A[m][k] B[k][n],  m = voxel, n = output channel,
//                              k = tap * cin + ci, A gathered from the channels-last activations
//                              with zero padding, B the packed bf16 weights [n][k].
//   weight gradient            C[n][k] = sum_m dy[m][n] x[m + off(tap(k))][ci(k)], split over the
//                              voxels into FP32 partials that a second kernel sums in fixed order.
// Tiles are staged through LDS in rows of 32 k-values (one MFMA depth); the next tile's global
// loads are issued before the current tile's MFMAs so that their latency overlaps the math.
//
// MFMA 16x16x32 bf16 lane layout (64 lanes, g = lane / 16, c = lane % 16):
//   A operand: A[c][8g + j], j = 0..7      B operand: B[8g + j][c]
//   D:         D[4g + e][c], e = 0..3

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>

#include "../../include/pvnet.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define NET_HIP(x)                                                                      \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if (e_ != hipSuccess) return fail(2, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define NET_LAUNCH()                                                                    \
    do {                                                                                \
        hipError_t e_ = hipGetLastError();                                              \
        if (e_ != hipSuccess) return fail(2, std::string("launch: ") + hipGetErrorString(e_)); \
    } while (0)

constexpr int BK = 32;  // k-values per MFMA step
constexpr int LP = 40;  // LDS row pitch in bf16 (32 + 8 padding: 80-byte rows)

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// round to nearest even; NaN stays NaN
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// bf16 bit pattern strictly greater than zero (NaN is not)
__device__ __forceinline__ bool bf_pos(uint16_t b) { return (b & 0x8000u) == 0 && b != 0 && b <= 0x7f80u; }

// Workgroups are dispatched to the 8 XCDs round-robin (block b to XCD b % 8), each XCD with its own
// L2.  Neighbouring output tiles share most of their input (the 3x3x3 halo), so block b works on
// tile xcd_tile(b, n): XCD x takes a contiguous run of tiles, and neighbours meet in one L2.
constexpr int NXCD = 8;

__device__ __forceinline__ int xcd_tile(int b, int n) {
    const int q = n / NXCD, r = n % NXCD, x = b % NXCD, i = b / NXCD;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}

__device__ __forceinline__ void tap_offset(int ntaps, int tap, int& dz, int& dy, int& dx) {
    if (ntaps == 27) {
        dz = tap / 9 - 1;
        dy = (tap / 3) % 3 - 1;
        dx = tap % 3 - 1;
    } else {
        dz = dy = dx = 0;
    }
}

// ---------------------------------------------------------------------------------------------
// forward / input-gradient convolution: block tile 4*WM voxels x BN channels, 4 waves of WM x BN

template <int BN, int WM>
__global__ __launch_bounds__(256) void k_conv_fwd(pvnet_conv a, int M, int K) {
    constexpr int BM = 4 * WM;
    constexpr int NA = BM / 64;  // A rows loaded per thread and step
    __shared__ __attribute__((aligned(16))) uint16_t As[BM * LP];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[BN * LP];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int g = lane >> 4, c16 = lane & 15;
    const int m0 = xcd_tile(blockIdx.x, gridDim.x) * BM, n0 = blockIdx.y * BN;
    const int d = a.s.d, h = a.s.h, w = a.s.w;

    // load roles: A rows r + 64 i, 8-wide k chunk q; B row bn (if < BN), same chunk.  Per row, a
    // mask of the taps whose neighbour lies in the box and the row's base pointer are set once;
    // the chunk's (tap, channel) advance by 32 k-values per step.
    const int q = t & 3, r = t >> 2, bn = t >> 2;
    uint32_t tmask[NA];
    const uint16_t* rowp[NA];
    #pragma unroll
    for (int i = 0; i < NA; i++) {
        const int m = m0 + r + 64 * i;
        const int mm = m < M ? m : 0;
        const int xx = mm % w, tq = mm / w, yy = tq % h, zz = (tq / h) % d;
        uint32_t mk = 0;
        if (m < M) {
            for (int tp = 0; tp < a.ntaps; tp++) {
                int dz, dy, dx;
                tap_offset(a.ntaps, tp, dz, dy, dx);
                if ((unsigned)(zz + dz) < (unsigned)d && (unsigned)(yy + dy) < (unsigned)h && (unsigned)(xx + dx) < (unsigned)w)
                    mk |= 1u << tp;
            }
        }
        tmask[i] = mk;
        rowp[i] = a.x + (int64_t)mm * a.ldx;
    }
    const bool brow = bn < BN && n0 + bn < a.cout;
    const uint16_t* wp = a.w + (int64_t)(n0 + (brow ? bn : 0)) * a.kpad + q * 8;
    const int nsteps = a.kpad / BK;
    uint4 ra[NA], rb;
    int tap = (q * 8) / a.cin, ci = q * 8 - tap * a.cin;  // chunk of step 0

    auto load = [&](int s) {
        const bool kin = tap < a.ntaps;  // k = s * 32 + 8 q < K
        int dz, dy, dx;
        tap_offset(a.ntaps, kin ? tap : 0, dz, dy, dx);
        const int64_t off = (((int64_t)dz * h + dy) * w + dx) * a.ldx + ci;
        #pragma unroll
        for (int i = 0; i < NA; i++) {
            ra[i] = make_uint4(0, 0, 0, 0);
            if (kin && ((tmask[i] >> tap) & 1u)) ra[i] = *(const uint4*)(rowp[i] + off);
        }
        rb = make_uint4(0, 0, 0, 0);
        if (brow) rb = *(const uint4*)(wp + s * BK);
        ci += BK;  // the chunk of the next step
        while (ci >= a.cin) {
            ci -= a.cin;
            tap++;
        }
    };

    f32x4 acc[WM / 16][BN / 16];
    #pragma unroll
    for (int i = 0; i < WM / 16; i++)
        for (int j = 0; j < BN / 16; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    load(0);
    for (int s = 0; s < nsteps; s++) {
        #pragma unroll
        for (int i = 0; i < NA; i++) *(uint4*)(As + (r + 64 * i) * LP + q * 8) = ra[i];
        if (bn < BN) *(uint4*)(Bs + bn * LP + q * 8) = rb;
        __syncthreads();
        if (s + 1 < nsteps) load(s + 1);
        bf16x8 af[WM / 16], bfr[BN / 16];
        #pragma unroll
        for (int i = 0; i < WM / 16; i++) af[i] = *(const bf16x8*)(As + (wv * WM + i * 16 + c16) * LP + g * 8);
        #pragma unroll
        for (int j = 0; j < BN / 16; j++) bfr[j] = *(const bf16x8*)(Bs + (j * 16 + c16) * LP + g * 8);
        #pragma unroll
        for (int i = 0; i < WM / 16; i++)
            for (int j = 0; j < BN / 16; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }

    const bool relu = a.flags & PVNET_RELU, f32out = a.flags & PVNET_F32OUT;
    #pragma unroll
    for (int i = 0; i < WM / 16; i++)
        for (int j = 0; j < BN / 16; j++) {
            const int n = n0 + j * 16 + c16;
            if (n >= a.cout) continue;
            const float b = a.bias ? a.bias[n] : 0.f;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int m = m0 + wv * WM + i * 16 + g * 4 + e;
                if (m >= M) continue;
                float v = acc[i][j][e] + b;
                if (relu) v = v < 0.f ? 0.f : v;
                if (a.relu_of && !bf_pos(a.relu_of[(int64_t)m * a.ldr + n])) v = 0.f;
                if (f32out)
                    ((float*)a.y)[(int64_t)m * a.ldy + n] = v;
                else
                    ((uint16_t*)a.y)[(int64_t)m * a.ldy + n] = f2bf(v);
            }
        }
}

// ---------------------------------------------------------------------------------------------
// 1x1x1 convolutions with K = cin <= 32, one MFMA k-step: every wave takes PWT tiles of 16
// voxels and feeds the matrix core straight from global memory, with no LDS and no barrier.  Lane
// (c16, g) holds the input channels 8g .. 8g+7 of voxel c16 of a tile (zero beyond cin) and the
// weight row 16 j + c16 at the same k-values: the operands k_conv_fwd stages for the same
// convolution, so the outputs are the same bit for bit.
constexpr int PWT = 4;

template <int NJ>
__global__ __launch_bounds__(256) void k_conv_pw(pvnet_conv a, int M) {
    const int lane = threadIdx.x & 63, g = lane >> 4, c16 = lane & 15;
    const int64_t tile0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * PWT;
    bf16x8 bfr[NJ];
    float bias[NJ];
#pragma unroll
    for (int j = 0; j < NJ; j++) {
        const int n = 16 * j + c16;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (n < a.cout) v = *(const uint4*)(a.w + (int64_t)n * a.kpad + g * 8);
        bfr[j] = *reinterpret_cast<const bf16x8*>(&v);
        bias[j] = (n < a.cout && a.bias) ? a.bias[n] : 0.f;
    }
    uint4 ra[PWT];
#pragma unroll
    for (int i = 0; i < PWT; i++) {
        const int64_t m = (tile0 + i) * 16 + c16;
        ra[i] = make_uint4(0, 0, 0, 0);
        if (m < M && g * 8 < a.cin) ra[i] = *(const uint4*)(a.x + m * a.ldx + g * 8);
    }
    const bool relu = a.flags & PVNET_RELU, f32out = a.flags & PVNET_F32OUT;
#pragma unroll
    for (int i = 0; i < PWT; i++)
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            const f32x4 acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(*reinterpret_cast<const bf16x8*>(&ra[i]), bfr[j],
                                                                      f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
            const int n = 16 * j + c16;
            if (n >= a.cout) continue;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int64_t m = (tile0 + i) * 16 + g * 4 + e;  // MFMA row 4 g + e
                if (m >= M) continue;
                float v = acc[e] + bias[j];
                if (relu) v = v < 0.f ? 0.f : v;
                if (a.relu_of && !bf_pos(a.relu_of[m * a.ldr + n])) v = 0.f;
                if (f32out)
                    ((float*)a.y)[m * a.ldy + n] = v;
                else
                    ((uint16_t*)a.y)[m * a.ldy + n] = f2bf(v);
            }
        }
}

// ---------------------------------------------------------------------------------------------
// 3x3x3 convolution with 16 NT output channels per block (NT = 1..4 column tiles; grid.y takes
// further groups of 16 NT channels) on grids of whole 16 x 4 x 4
// bricks: each block stages one brick of one box with its one-voxel halo (18 x 6 x 6 voxels, all
// cin channels) in LDS once and runs the k-steps of k_conv_fwd over it, where k_conv_fwd gathers
// every tap's rows from L2 again.  The k-steps, their operands and their order are those of
// k_conv_fwd, so each output is the same bit for bit.  Wave wv computes brick plane z = wv as 4
// MFMA row tiles (y = tile) of 16 voxels (x = MFMA row).  The halo is stored chunk-major
// (halo[chunk][voxel], a chunk = 8 channels, 16 bytes), so the 16 lanes that read one A fragment
// row group address 16 consecutive 16-byte rows: one pass over the 64 banks.

// bricks of BRX x BY x BRZ voxels (BY = 4 or 8, one wave per 4 rows of a plane), halo HLX x (BY + 2)
// x HLZ
constexpr int BRX = 16, BRZ = 4;
constexpr int HLX = BRX + 2, HLZ = BRZ + 2;
template <int BY>
constexpr int hl_vox() { return HLX * (BY + 2) * HLZ; }

// step table: entry 4 s + g = the halo offset of k-values s * 32 + 8 g .. + 7 relative to the
// output voxel (chunk * HLV + tap shift), plus BR_TAB_BIAS; 0 where they lie beyond K
constexpr int BR_TAB_BIAS = 1 << 20;

// WL: the block's weight rows are staged in LDS after the step table (row pitch kpad + 8 bf16, an
// odd number of 16-byte units, so the 16 rows one fragment reads fall on distinct banks), instead
// of being read from global memory one step ahead.
template <int NT, int BY, bool WL>
__global__ __launch_bounds__(64 * BY) void k_conv_fwd_brick(pvnet_conv a, int K) {
    constexpr int HLY = BY + 2, HLV = hl_vox<BY>(), NTH = 64 * BY;
    extern __shared__ uint4 halo[];  // [cin / 8 chunks][HLV voxels], the step table, WL: weights
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int zz = wv % BRZ, yq = wv / BRZ;  // this wave: plane zz, rows 4 yq .. 4 yq + 3
    const int g = lane >> 4, c16 = lane & 15;
    const int d = a.s.d, h = a.s.h, w = a.s.w;
    const int nbx = w / BRX, nby = h / BY, nbz = d / BRZ;
    int bid = xcd_tile(blockIdx.x, gridDim.x);
    const int bx = bid % nbx;
    bid /= nbx;
    const int by = bid % nby;
    bid /= nby;
    const int bz = bid % nbz;
    const int box = bid / nbz;
    const int x0 = bx * BRX, y0 = by * BY, z0 = bz * BRZ;
    const int n0 = blockIdx.y * 16 * NT;
    const int cch = a.cin >> 3;
    const int nsteps = a.kpad / BK;
    int* tab = reinterpret_cast<int*>(halo + HLV * cch);
    for (int i = t; i < 4 * nsteps; i += NTH) {
        const int k = (i >> 2) * BK + (i & 3) * 8;
        int v = 0;
        if (k < K) {
            const int tap = k / a.cin, ci = k - tap * a.cin;
            int dz, dy, dx;
            tap_offset(27, tap, dz, dy, dx);
            v = (ci >> 3) * HLV + (dz * HLY + dy) * HLX + dx + BR_TAB_BIAS;
        }
        tab[i] = v;
    }
    uint4* wl = reinterpret_cast<uint4*>(tab + 4 * nsteps);
    const int WP = 4 * nsteps + 1;  // 16-byte units per weight row in LDS
    if constexpr (WL) {
        for (int i = t; i < 16 * NT * 4 * nsteps; i += NTH) {
            const int row = i / (4 * nsteps), col = i - row * (4 * nsteps);
            wl[row * WP + col] = *(const uint4*)(a.w + (int64_t)(n0 + row) * a.kpad + col * 8);
        }
    }
    for (int i = t; i < HLV * cch; i += NTH) {
        const int ch = i / HLV, v = i - ch * HLV;
        const int x = x0 - 1 + v % HLX, y = y0 - 1 + (v / HLX) % HLY, z = z0 - 1 + v / (HLX * HLY);
        uint4 val = make_uint4(0, 0, 0, 0);
        if ((unsigned)x < (unsigned)w && (unsigned)y < (unsigned)h && (unsigned)z < (unsigned)d) {
            const int64_t m = (((int64_t)box * d + z) * h + y) * w + x;
            val = *(const uint4*)(a.x + m * a.ldx + ch * 8);
        }
        halo[i] = val;
    }
    __syncthreads();

    f32x4 acc[4][NT];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < NT; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    // halo voxel of (c16, tile 0, wv); tile i adds i rows.  Only the last step can reach beyond K
    // (kpad - K < 32), so the others run without a check, each prefetching the next step's table
    // entry and weight fragments.
    const int vox = ((zz + 1) * HLY + 4 * yq + 1) * HLX + c16 + 1 - BR_TAB_BIAS;
    const uint16_t* wrow[NT];
#pragma unroll
    for (int j = 0; j < NT; j++) wrow[j] = a.w + (int64_t)(n0 + 16 * j + c16) * a.kpad + g * 8;
    // weight fragment of column tile j at step s
    auto wfrag = [&](int j, int s) -> bf16x8 {
        if constexpr (WL)
            return *reinterpret_cast<const bf16x8*>(&wl[(16 * j + c16) * WP + 4 * s + g]);
        else
            return *(const bf16x8*)(wrow[j] + s * BK);
    };
    int tv = tab[g];
    bf16x8 bfr[NT];
#pragma unroll
    for (int j = 0; j < NT; j++) bfr[j] = wfrag(j, 0);
    for (int s = 0; s + 1 < nsteps; s++) {
        const int tvn = tab[4 * (s + 1) + g];
        bf16x8 bn[NT];
#pragma unroll
        for (int j = 0; j < NT; j++) bn[j] = wfrag(j, s + 1);
        const int base = vox + tv;
        uint4 r[4];
#pragma unroll
        for (int i = 0; i < 4; i++) r[i] = halo[base + i * HLX];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < NT; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(*reinterpret_cast<const bf16x8*>(&r[i]), bfr[j],
                                                                    acc[i][j], 0, 0, 0);
        tv = tvn;
#pragma unroll
        for (int j = 0; j < NT; j++) bfr[j] = bn[j];
    }
    {
        const int base = vox + tv;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            uint4 r = make_uint4(0, 0, 0, 0);
            if (tv) r = halo[base + i * HLX];
#pragma unroll
            for (int j = 0; j < NT; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(*reinterpret_cast<const bf16x8*>(&r), bfr[j],
                                                                    acc[i][j], 0, 0, 0);
        }
    }

    const bool relu = a.flags & PVNET_RELU, f32out = a.flags & PVNET_F32OUT;
#pragma unroll
    for (int j = 0; j < NT; j++) {
        const int n = n0 + 16 * j + c16;
        const float bias = a.bias ? a.bias[n] : 0.f;
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                // MFMA row 4 g + e: voxel x0 + 4 g + e of row y0 + 4 yq + i in plane z0 + zz
                const int64_t m = (((int64_t)box * d + z0 + zz) * h + y0 + 4 * yq + i) * w + x0 + 4 * g + e;
                float v = acc[i][j][e] + bias;
                if (relu) v = v < 0.f ? 0.f : v;
                if (a.relu_of && !bf_pos(a.relu_of[m * a.ldr + n])) v = 0.f;
                if (f32out)
                    ((float*)a.y)[m * a.ldy + n] = v;
                else
                    ((uint16_t*)a.y)[m * a.ldy + n] = f2bf(v);
            }
    }
}

// ---------------------------------------------------------------------------------------------
// weight gradient: block tile BN output channels x 64 k-values, reduction over its voxel chunk in
// LDS fills of 128 voxels.  Threads 0..127 each load an 8-voxel x 8-channel block of the gathered
// input (eight 16-byte loads), transpose it in registers and write eight 16-byte LDS rows of 8
// voxels; threads 128.. do the same for dy.  The MFMA reduction index is the voxel.
//
// The input's LDS rows are swizzled: the 16-byte chunk q of row r sits at chunk q ^ ((r >> 3) & 7).
// Eight consecutive x-role lanes write rows 8 kc + c (kc = 0..7) at one chunk, and any 16-byte row
// pitch puts rows eight apart on the same two bank groups (a 4-way conflict); the swizzle spreads
// them over eight, and the MFMA fragment reads (16 rows, one chunk) stay conflict-free.

// 8x8 transpose of bf16: r[i] = row i (8 channels, channel c in 16-bit half c); o[c] = column c
// (rows 0..7 in its halves)
__device__ __forceinline__ void transpose8(const uint4 (&r)[8], uint4 (&o)[8]) {
    const uint32_t* rw = reinterpret_cast<const uint32_t*>(r);
    uint32_t* ow = reinterpret_cast<uint32_t*>(o);
#pragma unroll
    for (int c = 0; c < 8; c++)
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t lo = rw[(2 * k) * 4 + (c >> 1)], hi = rw[(2 * k + 1) * 4 + (c >> 1)];
            ow[c * 4 + k] = (c & 1) ? ((lo >> 16) | (hi & 0xffff0000u)) : ((lo & 0xffffu) | (hi << 16));
        }
}

template <int BN>
__global__ __launch_bounds__(256) void k_conv_wgrad(pvnet_wgrad a, int M, int K, int mchunk) {
    constexpr int BMW = 128;  // voxels per LDS fill
    constexpr int PW = BMW + 8;  // LDS row pitch in bf16 (272 bytes)
    __shared__ __attribute__((aligned(16))) uint16_t As[BN * PW];  // dy^T  [n][m]
    __shared__ __attribute__((aligned(16))) uint16_t Bs[64 * PW];  // x^T   [k][m]
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int g = lane >> 4, c16 = lane & 15;
    const int k0 = blockIdx.x * 64, n0 = blockIdx.y * BN, split = blockIdx.z;
    const int mb = split * mchunk, me = min(M, mb + mchunk);
    const int d = a.s.d, h = a.s.h, w = a.s.w;

    // roles: x blocks (t < 128): voxel group mg, k chunk kc; dy blocks (128 <= t < 128 + 2 BN)
    const bool xrole = t < 128;
    const int u = t - 128;
    const bool drole = !xrole && u < 2 * BN;
    const int mg = xrole ? (t >> 3) : drole ? u / (BN / 8) : 0;
    const int kc = t & 7, nc = drole ? u % (BN / 8) : 0;
    const int k = k0 + kc * 8;
    const bool kin = xrole && k < K;
    const int tap = kin ? k / a.cin : 0;
    const int ci = k - tap * a.cin;
    int dz, dy, dx;
    tap_offset(a.ntaps, tap, dz, dy, dx);
    const int64_t off = ((int64_t)dz * h + dy) * w + dx;
    const bool dcol = drole && n0 + nc * 8 < a.cout;

    f32x4 acc[BN / 16];
    for (int j = 0; j < BN / 16; j++) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};

    uint4 r[8];
    auto load = [&](int mbase) {
        const int m0 = mbase + mg * 8;
        if (xrole) {
            int x = m0 % w, tq = m0 / w, y = tq % h, z = (tq / h) % d;
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int m = m0 + i;
                const int z2 = z + dz, y2 = y + dy, x2 = x + dx;
                const bool in = kin && m < me && (unsigned)z2 < (unsigned)d && (unsigned)y2 < (unsigned)h &&
                                (unsigned)x2 < (unsigned)w;
                r[i] = make_uint4(0, 0, 0, 0);
                if (in) r[i] = *(const uint4*)(a.x + ((int64_t)m + off) * a.ldx + ci);
                if (++x == w) {  // next voxel
                    x = 0;
                    if (++y == h) {
                        y = 0;
                        if (++z == d) z = 0;
                    }
                }
            }
        } else {
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int m = m0 + i;
                r[i] = make_uint4(0, 0, 0, 0);
                if (dcol && m < me) r[i] = *(const uint4*)(a.dy + (int64_t)m * a.lddy + n0 + nc * 8);
            }
        }
    };

    if (mb < me) load(mb);
    for (int mbase = mb; mbase < me; mbase += BMW) {
        if (xrole || drole) {
            uint4 o[8];
            transpose8(r, o);
            uint16_t* dst = xrole ? Bs + (kc * 8) * PW + (mg ^ kc) * 8 : As + (nc * 8) * PW + mg * 8;
#pragma unroll
            for (int c = 0; c < 8; c++) *(uint4*)(dst + c * PW) = o[c];
        }
        __syncthreads();
        if (mbase + BMW < me) load(mbase + BMW);
#pragma unroll
        for (int kk = 0; kk < BMW / 32; kk++) {
            const int brow = wv * 16 + c16;
            const bf16x8 bx = *(const bf16x8*)(Bs + brow * PW + ((kk * 4 + g) ^ ((brow >> 3) & 7)) * 8);
#pragma unroll
            for (int j = 0; j < BN / 16; j++) {
                const bf16x8 ad = *(const bf16x8*)(As + (j * 16 + c16) * PW + kk * 32 + g * 8);
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ad, bx, acc[j], 0, 0, 0);
            }
        }
        __syncthreads();
    }

    const int kk = k0 + wv * 16 + c16;
    if (kk >= K) return;
    for (int j = 0; j < BN / 16; j++)
        for (int e = 0; e < 4; e++) {
            const int n = n0 + j * 16 + g * 4 + e;
            if (n < a.cout) a.ws[((int64_t)split * a.cout + n) * K + kk] = acc[j][e];
        }
}

// ---------------------------------------------------------------------------------------------
// weight gradient of a 16- or 32-output 3x3x3 convolution on grids of whole 8 x 8 x 4 bricks.
// Block `split` reduces a contiguous run of bricks; per brick it stages dy^T once (16 or 32 rows of
// 256 voxels) and holds its MFMA A fragments (8 per 16 rows, 32 voxels each) in registers for
// every tap, so that each B fragment feeds one MFMA per 16 output channels, then stages
// the input one 8-channel group at a time as three channel-major copies, one per x shift dx, of
// the brick's rows with their z/y halo: X[dx][ci][zh][yh] = the 8 voxels x0 + dx + (0..7) of row
// (z0 - 1 + zh, y0 - 1 + yh), one 16-byte LDS row.  A B fragment (32 voxels x 16 k-columns) is
// then one aligned 16-byte read per lane.  A 16-column C tile pairs the taps 2p and 2p + 1 of the
// group's 8 channels (14 tiles, the last half empty); wave w owns tiles w, w + 4, ...
//
// LDS rows are 16 bytes; the 64 banks span 16 rows.  The channel stride of 65 rows (= 1 mod 16)
// puts the 8 channels a half of a 16-lane group reads on 8 consecutive bank groups, and the copy
// stride of 520 rows (= 8 mod 16) puts the other half, the next tap's copy, on the other 8; dy^T
// rows are 33 apart (= 1 mod 16) for the same reason.

constexpr int WBRX = 8, WBRY = 8, WBRZ = 4;  // weight-gradient bricks
constexpr int WBX_CH = 65;                 // rows per channel (6 x 10 + 5)
constexpr int WBX_CPY = 8 * WBX_CH;        // rows per dx copy
constexpr int WBX_ROWS = 3 * WBX_CPY;      // input rows
constexpr int WBX_ZERO = 40;               // zero rows after them (row offsets within a tile <= 37)
constexpr int WBD_P = 33;                  // dy^T row pitch in 16-byte rows (256 voxels + 1)
constexpr int WB_SPLITS = 1024;            // blocks (partial rows) at most along the voxels
constexpr int WB_NGB = 2;                  // 8-channel groups per block

// column c of the 8x8 bf16 block r (rows = voxels, 16-bit halves = channels): 8 voxels of channel c
template <int c>
__device__ __forceinline__ uint4 tcol(const uint4 (&r)[8]) {
    const uint32_t* rw = reinterpret_cast<const uint32_t*>(r);
    uint32_t o[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t lo = rw[(2 * k) * 4 + (c >> 1)], hi = rw[(2 * k + 1) * 4 + (c >> 1)];
        o[k] = (c & 1) ? ((lo >> 16) | (hi & 0xffff0000u)) : ((lo & 0xffffu) | (hi << 16));
    }
    return make_uint4(o[0], o[1], o[2], o[3]);
}

template <int c = 0>
__device__ __forceinline__ void store_cols(const uint4 (&r)[8], uint4* dst, int stride) {
    if constexpr (c < 8) {
        dst[c * stride] = tcol<c>(r);
        store_cols<c + 1>(r, dst, stride);
    }
}

// block (split, gy): bricks split * per .. ; channel groups gy * NGB + (0 .. NGB - 1); NR row tiles
// of 16 output channels (cout = 16 NR).  The grid is one-dimensional, ngy blocks per split: the
// blocks of one split take consecutive slots of one XCD, so that they walk the same bricks at the
// same time through one L2 (each reads its own channels of the same input rows).
//
// dboff >= 0 (16-output layers): the blocks of the first channel group also sum dy over their
// bricks for the bias gradient, as one more MFMA per fragment with B = ones in their first wave,
// into the row ws[dboff + split * cout ..] (the split sums reduce those rows like the weights').
template <int NGB, int NR>
__global__ __launch_bounds__(256) void k_conv_wgrad_brick(pvnet_wgrad a, int K, int nbricks, int per, int ngy,
                                                          int64_t dboff) {
    // X ends in WBX_ZERO zero rows: the lanes of the empty half of the last tap pair read there
    __shared__ uint4 X[WBX_ROWS + WBX_ZERO];
    __shared__ uint4 DY[16 * NR * WBD_P];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int g = lane >> 4, c16 = lane & 15;
    const int d = a.s.d, h = a.s.h, w = a.s.w;
    const int nbx = w / WBRX, nby = h / WBRY, nbz = d / WBRZ;
    const int slot = xcd_tile(blockIdx.x, gridDim.x);
    const int split = slot / ngy, cg0 = (slot % ngy) * NGB;
    const int ngrp = min(NGB, a.cin / 8 - cg0);
    const int b0 = split * per, b1 = min(nbricks, b0 + per);

    f32x4 acc[NGB][4][NR];
#pragma unroll
    for (int j = 0; j < NGB; j++)
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int rt = 0; rt < NR; rt++) acc[j][i][rt] = f32x4{0.f, 0.f, 0.f, 0.f};
    const bool dob = NR == 1 && dboff >= 0 && cg0 == 0 && wv == 0;  // wave-uniform
    f32x4 accb = f32x4{0.f, 0.f, 0.f, 0.f};
    bf16x8 ones;
#pragma unroll
    for (int e = 0; e < 8; e++) ones[e] = (__bf16)1.0f;

    const int cl = c16 & 7;  // this lane's channel within the group; its tap is 2 p + (c16 >> 3)
    if (t < WBX_ZERO) X[WBX_ROWS + t] = make_uint4(0, 0, 0, 0);  // ordered by the first barrier

    for (int b = b0; b < b1; b++) {
        int bid = b;
        const int bx = bid % nbx;
        bid /= nbx;
        const int by = bid % nby;
        bid /= nby;
        const int bz = bid % nbz;
        const int64_t box = bid / nbz;
        const int x0 = bx * WBRX, y0 = by * WBRY, z0 = bz * WBRZ;

        // dy task (row rr, channel chunk nc): dy^T rows n = 8 nc .. 8 nc + 7 at voxels 8 rr ..
        for (int u = t; u < 64 * NR; u += 256) {
            const int rr = u / (2 * NR), nc = u % (2 * NR);
            const int64_t row = ((box * d + z0 + (rr >> 3)) * h + y0 + (rr & 7)) * (int64_t)w + x0;
            uint4 q[8];
#pragma unroll
            for (int i = 0; i < 8; i++) q[i] = *(const uint4*)(a.dy + (row + i) * a.lddy + nc * 8);
            store_cols(q, DY + nc * 8 * WBD_P + rr, WBD_P);
        }
#pragma unroll
        for (int j = 0; j < NGB; j++) {
            if (j >= ngrp) break;  // block-uniform
            const int cg = cg0 + j;
            if (j > 0) __syncthreads();  // the previous group's reads of X are done
            if (t < 180) {
                // input task (dx, zh, yh): 8 voxels x 8 channels, written as 8 channel rows
                const int dxi = t / 60, zh = (t / 10) % 6, yh = t % 10;
                const int z = z0 - 1 + zh, y = y0 - 1 + yh;
                const bool zy = (unsigned)z < (unsigned)d && (unsigned)y < (unsigned)h;
                const int64_t row = ((box * d + z) * h + y) * (int64_t)w;
                uint4 r[8];
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    const int x = x0 + i + dxi - 1;
                    r[i] = make_uint4(0, 0, 0, 0);
                    if (zy && (unsigned)x < (unsigned)w) r[i] = *(const uint4*)(a.x + (row + x) * a.ldx + cg * 8);
                }
                store_cols(r, X + dxi * WBX_CPY + zh * 10 + yh, WBX_CH);
            }
            __syncthreads();
            // A fragments: dy^T[n = 16 rt + c16][voxels 32 s + 8 g .. + 7] of brick row 4 s + g
            bf16x8 af[NR][8];
#pragma unroll
            for (int rt = 0; rt < NR; rt++)
#pragma unroll
                for (int s2 = 0; s2 < 8; s2++)
                    af[rt][s2] = *reinterpret_cast<const bf16x8*>(&DY[(16 * rt + c16) * WBD_P + 4 * s2 + g]);
            if (NR == 1 && dob && j == 0) {
                // every column of dy^T x ones is the sum of dy over the brick's voxels
#pragma unroll
                for (int s2 = 0; s2 < 8; s2++)
                    accb = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[0][s2], ones, accb, 0, 0, 0);
            }
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int p = wv + 4 * i;
                if (p >= 14) break;
                const int tap = 2 * p + (c16 >> 3);
                int dz, dy, dx;
                tap_offset(27, tap < 27 ? tap : 0, dz, dy, dx);
                const int base = tap < 27 ? (dx + 1) * WBX_CPY + cl * WBX_CH + (1 + dz) * 10 + (1 + dy) : WBX_ROWS;
#pragma unroll
                for (int s2 = 0; s2 < 8; s2++) {
                    const int r = 4 * s2 + g;  // brick row: z = r / 8, y = r % 8
                    const uint4 v = X[base + (r >> 3) * 10 + (r & 7)];
#pragma unroll
                    for (int rt = 0; rt < NR; rt++)
                        acc[j][i][rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            af[rt][s2], *reinterpret_cast<const bf16x8*>(&v), acc[j][i][rt], 0, 0, 0);
                }
            }
        }
        __syncthreads();  // X and DY are restaged for the next brick
    }

    // D[4 g + e][c16] of row tile rt: n = 16 rt + 4 g + e; column c16 of tile p -> k = tap * cin +
    // 8 cg + (c16 & 7)
#pragma unroll
    for (int j = 0; j < NGB; j++) {
        if (j >= ngrp) break;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int p = wv + 4 * i;
            if (p >= 14) break;
            const int tap = 2 * p + (c16 >> 3);
            if (tap >= 27) continue;
            const int k = tap * a.cin + (cg0 + j) * 8 + cl;
#pragma unroll
            for (int rt = 0; rt < NR; rt++)
#pragma unroll
                for (int e = 0; e < 4; e++)
                    a.ws[((int64_t)split * 16 * NR + 16 * rt + 4 * g + e) * K + k] = acc[j][i][rt][e];
        }
    }
    // D[4 g + e][c16] = sum over this block's voxels of dy[.][4 g + e], the same in every column
    if (dob && c16 == 0) {
#pragma unroll
        for (int e = 0; e < 4; e++) a.ws[dboff + (int64_t)split * 16 + 4 * g + e] = accb[e];
    }
}

// column sums of dy over a chunk of voxels per block: part[blk][n].  Thread t sums the 8-channel
// group t % g over the rows r = t / g, t / g + R, ... (g = cout / 8, R = 256 / g); the R partial
// rows are then added in ascending order.
__global__ __launch_bounds__(256) void k_colsum(const uint16_t* dy, int ld, int cout, int M, int mchunk,
                                                float* part) {
    __shared__ float sp[256 * 8];
    const int g = cout / 8, R = 256 / g;
    const int t = threadIdx.x, c8 = t % g, rg = t / g;
    const int mb = blockIdx.x * mchunk, me = min(M, mb + mchunk);
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (rg < R)
        for (int m = mb + rg; m < me; m += R) {
            const uint4 v = *(const uint4*)(dy + (int64_t)m * ld + c8 * 8);
            const uint16_t* pv = (const uint16_t*)&v;
            for (int j = 0; j < 8; j++) acc[j] += bf2f(pv[j]);
        }
    for (int j = 0; j < 8; j++) sp[t * 8 + j] = acc[j];
    __syncthreads();
    if (t < cout) {
        const int cg = t / 8, j = t % 8;
        float s = 0.f;
        for (int r = 0; r < R; r++) s += sp[(r * g + cg) * 8 + j];
        part[(int64_t)blockIdx.x * cout + t] = s;
    }
}

// dst[i] = sum_s src[s * len + i] in a fixed order: a block takes 16 columns, its 16 row groups
// sum the rows rg, rg + 16, ... in ascending order, and the 16 partials are added in ascending rg
constexpr int SUM_COLS = 16;

// One reduction of split partials: dst[(i / rowlen) ldd + i % rowlen] = sum_s src[s len + i].
struct SumJob {
    const float* src;
    float* dst;
    int64_t len, ldd, rowlen;
    int nsplit;
    int blocks;  // blocks_for(len, SUM_COLS)
};

// Two reductions in one launch (the weight and bias gradients of a layer): blocks below
// a.blocks take job a, the others job b, each summing exactly as alone.
__global__ __launch_bounds__(256) void k_sum_splits(const SumJob a, const SumJob b) {
    __shared__ float sp[16][SUM_COLS + 1];
    const bool second = (int)blockIdx.x >= a.blocks;
    const float* src = second ? b.src : a.src;
    float* dst = second ? b.dst : a.dst;
    const int64_t len = second ? b.len : a.len, ldd = second ? b.ldd : a.ldd, rowlen = second ? b.rowlen : a.rowlen;
    const int nsplit = second ? b.nsplit : a.nsplit;
    const int blk = second ? (int)blockIdx.x - a.blocks : (int)blockIdx.x;
    const int t = threadIdx.x, col = t % SUM_COLS, rg = t / SUM_COLS;
    const int64_t i = (int64_t)blk * SUM_COLS + col;
    float s = 0.f;
    if (i < len)
        for (int p = rg; p < nsplit; p += 16) s += src[(int64_t)p * len + i];
    sp[rg][col] = s;
    __syncthreads();
    if (t < SUM_COLS && i < len) {
        float acc = 0.f;
        for (int r = 0; r < 16; r++) acc += sp[r][t];
        dst[(i / rowlen) * ldd + i % rowlen] = acc;
    }
}

// ---------------------------------------------------------------------------------------------
// pooling, upsampling, ReLU backward, features

struct Coarse {
    int64_t b;
    int z, y, x;
};

__device__ __forceinline__ int64_t fine_index(const pvnet_shape& s, int64_t b, int z, int y, int x) {
    return ((b * s.d + z) * s.h + y) * (int64_t)s.w + x;
}

// coarse voxel and 8-channel chunk of thread i (32-bit arithmetic: the host keeps the coarse
// voxel count times the chunk count below 2^31)
__device__ __forceinline__ bool coarse_of(const pvnet_shape& s, int c8n, int64_t i, Coarse& o, int& ch) {
    const uint32_t d2 = s.d / 2, h2 = s.h / 2, w2 = s.w / 2;
    const int64_t n = (int64_t)s.nb * d2 * h2 * w2 * c8n;
    if (i >= n) return false;
    uint32_t v = (uint32_t)i;
    ch = (int)(v % (uint32_t)c8n) * 8;
    v /= (uint32_t)c8n;
    o.x = (int)(v % w2);
    v /= w2;
    o.y = (int)(v % h2);
    v /= h2;
    o.z = (int)(v % d2);
    o.b = v / d2;
    return true;
}

__device__ __forceinline__ int64_t coarse_index(const pvnet_shape& s, const Coarse& o) {
    return ((o.b * (s.d / 2) + o.z) * (s.h / 2) + o.y) * (int64_t)(s.w / 2) + o.x;
}

// fine voxel f and 8-channel chunk of thread i (consecutive threads: the chunks of one voxel, then
// the next voxel of the row), and the index of its parent coarse voxel; 32-bit arithmetic as above
__device__ __forceinline__ bool fine_of(const pvnet_shape& s, int c8n, int64_t i, int64_t& f, int64_t& cidx, int& ch) {
    const uint32_t w = s.w, h = s.h, d = s.d;
    const int64_t n = (int64_t)s.nb * d * h * w * c8n;
    if (i >= n) return false;
    uint32_t v = (uint32_t)i;
    ch = (int)(v % (uint32_t)c8n) * 8;
    v /= (uint32_t)c8n;
    f = v;
    const uint32_t x = v % w;
    v /= w;
    const uint32_t y = v % h;
    v /= h;
    const uint32_t z = v % d, b = v / d;
    cidx = (((int64_t)b * (d / 2) + z / 2) * (h / 2) + y / 2) * (int64_t)(w / 2) + x / 2;
    return true;
}

__global__ void k_pool2_fwd(pvnet_shape s, int c, const uint16_t* x, int ldx, uint16_t* y, int ldy) {
    Coarse o;
    int ch;
    if (!coarse_of(s, c / 8, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, o, ch)) return;
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < 8; k++) {
        const int64_t f = fine_index(s, o.b, 2 * o.z + (k >> 2), 2 * o.y + ((k >> 1) & 1), 2 * o.x + (k & 1));
        const uint4 v = *(const uint4*)(x + f * ldx + ch);
        const uint16_t* p = (const uint16_t*)&v;
        for (int j = 0; j < 8; j++) acc[j] += bf2f(p[j]);
    }
    uint4 out;
    uint16_t* po = (uint16_t*)&out;
    for (int j = 0; j < 8; j++) po[j] = f2bf(acc[j] * 0.125f);
    *(uint4*)(y + coarse_index(s, o) * ldy + ch) = out;
}

// one thread per fine voxel chunk: dx = dx (accumulate) + dy(parent) / 8.  relu_of (optional): the
// result is multiplied by (relu_of > 0), the backward of a ReLU whose output is relu_of, in the same
// pass
__global__ void k_pool2_bwd(pvnet_shape s, int c, const uint16_t* dy, int lddy, uint16_t* dx, int lddx, int accum,
                            const uint16_t* relu_of, int ldr) {
    int64_t f, ci;
    int ch;
    if (!fine_of(s, c / 8, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, f, ci, ch)) return;
    const uint4 v = *(const uint4*)(dy + ci * lddy + ch);
    const uint16_t* p = (const uint16_t*)&v;
    uint4* dst = (uint4*)(dx + f * lddx + ch);
    uint4 out;
    uint16_t* po = (uint16_t*)&out;
    if (accum) {
        const uint4 old = *dst;
        const uint16_t* pold = (const uint16_t*)&old;
        for (int j = 0; j < 8; j++) po[j] = f2bf(bf2f(pold[j]) + bf2f(p[j]) * 0.125f);
    } else {
        for (int j = 0; j < 8; j++) po[j] = f2bf(bf2f(p[j]) * 0.125f);
    }
    if (relu_of) {
        const uint4 av = *(const uint4*)(relu_of + f * ldr + ch);
        const uint16_t* pa = (const uint16_t*)&av;
        for (int j = 0; j < 8; j++)
            if (!bf_pos(pa[j])) po[j] = 0;
    }
    *dst = out;
}

// one thread per fine voxel chunk, copying its parent's
__global__ void k_up2_fwd(pvnet_shape s, int c, const uint16_t* x, int ldx, uint16_t* y, int ldy) {
    int64_t f, ci;
    int ch;
    if (!fine_of(s, c / 8, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, f, ci, ch)) return;
    *(uint4*)(y + f * ldy + ch) = *(const uint4*)(x + ci * ldx + ch);
}

__global__ void k_up2_bwd(pvnet_shape s, int c, const uint16_t* dy, int lddy, uint16_t* dx, int lddx) {
    Coarse o;
    int ch;
    if (!coarse_of(s, c / 8, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, o, ch)) return;
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < 8; k++) {
        const int64_t f = fine_index(s, o.b, 2 * o.z + (k >> 2), 2 * o.y + ((k >> 1) & 1), 2 * o.x + (k & 1));
        const uint4 v = *(const uint4*)(dy + f * lddy + ch);
        const uint16_t* p = (const uint16_t*)&v;
        for (int j = 0; j < 8; j++) acc[j] += bf2f(p[j]);
    }
    uint4 out;
    uint16_t* po = (uint16_t*)&out;
    for (int j = 0; j < 8; j++) po[j] = f2bf(acc[j]);
    *(uint4*)(dx + coarse_index(s, o) * lddx + ch) = out;
}

// grid.y: the 8-channel chunk
__global__ void k_relu_bwd(int64_t m, const uint16_t* a, int lda, uint16_t* dx, int lddx) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= m) return;
    const int ch = blockIdx.y * 8;
    const uint4 av = *(const uint4*)(a + v * lda + ch);
    uint4* dp = (uint4*)(dx + v * lddx + ch);
    uint4 dv = *dp;
    const uint16_t* pa = (const uint16_t*)&av;
    uint16_t* pd = (uint16_t*)&dv;
    for (int j = 0; j < 8; j++)
        if (!bf_pos(pa[j])) pd[j] = 0;
    *dp = dv;
}

// squared-distance transform along one axis: g(i) = min_j (g_in(j) + (i - j)^2); the first pass
// starts from g_in = 0 on solid voxels and "infinite" elsewhere.  One thread per line.
constexpr float EDT_INF = 1e30f;

// One thread per line; the block's lines are staged in LDS (row pitch n + 1: the threads' reads of
// one column fall in distinct banks).  The first pass (binary input: 0 on solid voxels) is the
// squared distance to the nearest solid voxel of the line, from one sweep each way.  The later
// passes give each voxel min_j (f(j) + (i - j)^2) by scanning outward from i and stopping at the
// first r with r^2 >= the best value so far: f >= 0, so no farther j can lower it.  Every term is
// an exact integer in FP32 (or EDT_INF), so the result is that of the full scan bit for bit.
__global__ __launch_bounds__(64) void k_edt_pass(pvnet_shape s, int axis, const uint8_t* code, float* g) {
    extern __shared__ float sh[];
    const int64_t L = (int64_t)s.nb * s.d * s.h * s.w;
    const int n = axis == 0 ? s.w : axis == 1 ? s.h : s.d;
    const int64_t nlines = L / n;
    const int64_t line = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float* in = sh + (int64_t)threadIdx.x * (n + 1);
    if (line >= nlines) return;
    int64_t base, stride;
    if (axis == 0) {
        base = line * s.w;
        stride = 1;
    } else if (axis == 1) {
        const int64_t x = line % s.w, bz = line / s.w;
        base = bz * (int64_t)s.h * s.w + x;
        stride = s.w;
    } else {
        const int64_t xy = line % ((int64_t)s.w * s.h), b = line / ((int64_t)s.w * s.h);
        base = b * (int64_t)s.d * s.h * s.w + xy;
        stride = (int64_t)s.w * s.h;
    }
    if (axis == 0) {
        // in[j] = distance to the nearest solid voxel at or left of j (n + n if none), then the
        // sweep back takes the right side
        int last = -2 * n;
        for (int j = 0; j < n; j++) {
            if (code[base + j] == 0) last = j;
            in[j] = (float)(j - last);
        }
        int next = 3 * n;
        for (int j = n - 1; j >= 0; j--) {
            if (code[base + j] == 0) next = j;
            const int r = min((int)in[j], next - j);
            g[base + j] = r >= n ? EDT_INF : (float)(r * r);
        }
        return;
    }
    for (int j = 0; j < n; j++) in[j] = g[base + j * stride];
    for (int i = 0; i < n; i++) {
        float best = in[i];
        for (int r = 1; r < n; r++) {
            const float rr = (float)(r * r);
            if (rr >= best) break;
            if (i - r >= 0) best = fminf(best, in[i - r] + rr);
            if (i + r < n) best = fminf(best, in[i + r] + rr);
        }
        g[base + i * stride] = axis == 2 ? fminf(sqrtf(best), 255.f) : best;
    }
}

__global__ void k_features(pvnet_shape s, const uint8_t* code, const float* prm, const float* dist, uint16_t* xo) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)s.nb * s.d * s.h * s.w;
    if (i >= n) return;
    const int x = (int)(i % s.w);
    const int64_t b = i / ((int64_t)s.w * s.h * s.d);
    const int cd = code[i];
    const float fl = cd > 0 ? 1.f : 0.f;
    float f[8];
    f[0] = fl;
    f[1] = cd == 2 ? 1.f : 0.f;
    f[2] = cd == 3 ? 1.f : 0.f;
    f[3] = fl * (prm[3 * b + 1] - 0.5f) / 0.4f;
    f[4] = fl * prm[3 * b] / 0.05f;
    f[5] = fl * ((float)x / (float)max(1, s.w - 1) - 0.5f);
    f[6] = fl * dist[i] / 16.f;
    f[7] = fl * prm[3 * b + 2] / 804.247719f;  // pi * 16^2
    uint4 out;
    uint16_t* po = (uint16_t*)&out;
    for (int j = 0; j < 8; j++) po[j] = f2bf(f[j]);
    *(uint4*)(xo + i * 8) = out;
}

// ---------------------------------------------------------------------------------------------
// masked MSE loss: per-block FP64 partials, one final block, then the gradient

constexpr int LOSS_BLOCK = 256;

__global__ __launch_bounds__(LOSS_BLOCK) void k_loss_partial(int64_t M, int nc, const float* y, int ldy,
                                                             const uint16_t* t, const uint8_t* code,
                                                             const float* wc, double* part) {
    __shared__ double ss[LOSS_BLOCK], sc[LOSS_BLOCK];
    const int64_t i = (int64_t)blockIdx.x * LOSS_BLOCK + threadIdx.x;
    double e = 0.0, c = 0.0;
    if (i < M && code[i] > 0) {
        c = 1.0;
        for (int k = 0; k < nc; k++) {
            const float r = y[i * ldy + k] - bf2f(t[i * nc + k]);
            e += (double)(wc[k] * r * r);
        }
    }
    ss[threadIdx.x] = e;
    sc[threadIdx.x] = c;
    __syncthreads();
    for (int st = LOSS_BLOCK / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            ss[threadIdx.x] += ss[threadIdx.x + st];
            sc[threadIdx.x] += sc[threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = ss[0];
        part[2 * blockIdx.x + 1] = sc[0];
    }
}

__global__ __launch_bounds__(LOSS_BLOCK) void k_loss_final(int nblk, int nc, const double* part, double* loss) {
    __shared__ double ss[LOSS_BLOCK], sc[LOSS_BLOCK];
    double e = 0.0, c = 0.0;
    for (int b = threadIdx.x; b < nblk; b += LOSS_BLOCK) {
        e += part[2 * b];
        c += part[2 * b + 1];
    }
    ss[threadIdx.x] = e;
    sc[threadIdx.x] = c;
    __syncthreads();
    for (int st = LOSS_BLOCK / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            ss[threadIdx.x] += ss[threadIdx.x + st];
            sc[threadIdx.x] += sc[threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        loss[0] = sc[0] > 0 ? ss[0] / (sc[0] * nc) : 0.0;
        loss[1] = sc[0];
    }
}

__global__ void k_loss_grad(int64_t M, int nc, const float* y, int ldy, const uint16_t* t, const uint8_t* code,
                            const float* wc, const double* loss, uint16_t* dy, int lddy) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const double cnt = loss[1];
    const float sc = cnt > 0 ? (float)(2.0 / (cnt * nc)) : 0.f;
    const bool fl = code[i] > 0;
    for (int k = 0; k < lddy; k++) {
        float gk = 0.f;
        if (fl && k < nc) gk = sc * wc[k] * (y[i * ldy + k] - bf2f(t[i * nc + k]));
        dy[i * lddy + k] = f2bf(gk);
    }
}

// ---------------------------------------------------------------------------------------------
// optimizer and weight packing

__global__ void k_adam(int64_t n, float* p, const float* g, float* m, float* v, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_pack(const float* w32, int ntaps, int cin, int cout, int kpad, int kpad_d, uint16_t* wt,
                       uint16_t* wd) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nt = (int64_t)cout * kpad, nd = (int64_t)cin * kpad_d;
    const int K = ntaps * cin, Kd = ntaps * cout;
    if (i < nt) {
        const int n = (int)(i / kpad), k = (int)(i % kpad);
        wt[i] = k < K ? f2bf(w32[(int64_t)n * K + k]) : 0;
    } else if (i < nt + nd) {
        const int64_t j = i - nt;
        const int ci = (int)(j / kpad_d), k = (int)(j % kpad_d);
        uint16_t v = 0;
        if (k < Kd) {
            const int tapd = k / cout, co = k % cout;
            v = f2bf(w32[(int64_t)co * K + (ntaps - 1 - tapd) * cin + ci]);
        }
        wd[j] = v;
    }
}

// ---------------------------------------------------------------------------------------------
// host helpers

int64_t voxels(const pvnet_shape& s) { return (int64_t)s.nb * s.d * s.h * s.w; }

bool shape_ok(const pvnet_shape* s) { return s && s->nb > 0 && s->d > 0 && s->h > 0 && s->w > 0; }

bool al16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

int kpad_of(int ntaps, int c) { return (ntaps * c + BK - 1) / BK * BK; }

struct WgradPlan {
    int ktiles, ntiles, bn, nsplit, mchunk, cblk, cchunk;
    int64_t ws_main, ws_total;
    bool brick;        // k_conv_wgrad_brick (nsplit blocks of `per` bricks each)
    int nbricks, per;
};

// the brick kernel takes 16-output 3x3x3 layers with up to 48 input channels on grids of whole
// bricks; PVNET_GATHER keeps them on k_conv_wgrad
bool wgrad_brick_ok(const pvnet_shape& s, int ntaps, int cin, int cout) {
    return ntaps == 27 && (cout == 16 || cout == 32) && cin % 8 == 0 && s.d % WBRZ == 0 && s.h % WBRY == 0 &&
           s.w % WBRX == 0;
}

int wgrad_bn(int cout) { return cout <= 16 ? 16 : (cout <= 32 || cout % 64 != 0) ? 32 : 64; }

WgradPlan plan_wgrad(const pvnet_shape& s, int ntaps, int cin, int cout, int flags) {
    WgradPlan p{};
    const int64_t M = voxels(s);
    const int K = ntaps * cin;
    p.bn = wgrad_bn(cout);
    p.ktiles = (K + 63) / 64;
    p.ntiles = (cout + p.bn - 1) / p.bn;
    const int64_t want = 2048 / (p.ktiles * p.ntiles) + 1;
    int64_t steps = (M + 127) / 128;  // LDS fills of 128 voxels
    int64_t ns = std::min<int64_t>(std::max<int64_t>(want, 1), std::max<int64_t>(steps, 1));
    int64_t per = (steps + ns - 1) / ns;
    p.mchunk = (int)(per * 128);
    p.nsplit = (int)((M + p.mchunk - 1) / p.mchunk);
    p.ws_main = (int64_t)p.nsplit * cout * K;
    p.brick = !(flags & PVNET_GATHER) && wgrad_brick_ok(s, ntaps, cin, cout);
    if (p.brick) {
        p.nbricks = (int)(M / (WBRX * WBRY * WBRZ));
        const int ns2 = std::min(p.nbricks, WB_SPLITS);
        p.per = (p.nbricks + ns2 - 1) / ns2;
        p.nsplit = (p.nbricks + p.per - 1) / p.per;
        p.ws_main = (int64_t)p.nsplit * cout * K;
    }
    p.cchunk = (int)std::max<int64_t>(256, ((M + 1023) / 1024 + 255) & ~(int64_t)255);  // ~1024 blocks
    p.cblk = (int)((M + p.cchunk - 1) / p.cchunk);
    // the bias rows: one per column-sum block, or one per split when the brick kernel sums them
    p.ws_total = p.ws_main + (int64_t)std::max(p.cblk, p.brick ? p.nsplit : 0) * cout;
    return p;
}

int blocks_for(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

int pool_args(const pvnet_shape* s, int c, const void* a, int lda, const void* b, int ldb, const char* who) {
    if (!shape_ok(s) || s->d % 2 || s->h % 2 || s->w % 2) return fail(1, std::string(who) + ": sizes must be even");
    if (c <= 0 || c % 8 || lda < c || ldb < c || lda % 8 || ldb % 8) return fail(1, std::string(who) + ": bad channels");
    if (!a || !b || !al16(a) || !al16(b)) return fail(1, std::string(who) + ": null or misaligned pointer");
    if (voxels(*s) * (c / 8) >= ((int64_t)1 << 31)) return fail(1, std::string(who) + ": too many voxels");
    return 0;
}

}  // namespace

extern "C" {

const char* pvnet_last_error(void) { return g_err.c_str(); }

int pvnet_conv_fwd(const pvnet_conv* c, void* stream) {
    if (!c || !shape_ok(&c->s)) return fail(1, "pvnet_conv_fwd: bad shape");
    if (c->ntaps != 1 && c->ntaps != 27) return fail(1, "pvnet_conv_fwd: ntaps must be 1 or 27");
    if (c->cin <= 0 || c->cin % 8 || c->cout <= 0) return fail(1, "pvnet_conv_fwd: cin must be a positive multiple of 8");
    if (c->kpad != kpad_of(c->ntaps, c->cin)) return fail(1, "pvnet_conv_fwd: kpad must be roundup(ntaps*cin, 32)");
    if (c->ldx < c->cin || c->ldx % 8 || c->ldy < c->cout) return fail(1, "pvnet_conv_fwd: bad row lengths");
    if (!c->x || !c->w || !c->y || !al16(c->x) || !al16(c->w)) return fail(1, "pvnet_conv_fwd: null or misaligned pointer");
    if (c->relu_of && c->ldr < c->cout) return fail(1, "pvnet_conv_fwd: bad relu_of row length");
    const int64_t M = voxels(c->s);
    if (M >= (int64_t)1 << 31) return fail(1, "pvnet_conv_fwd: too many voxels");
    hipStream_t st = (hipStream_t)stream;
    const int K = c->ntaps * c->cin;
    // tiles: 256-voxel blocks for narrow outputs on large grids; 64-channel tiles only when the
    // grid still has >= 2048 blocks (the K loop is long, so the device needs many blocks in flight)
    const int mt = blocks_for(M, 128);
    // brick kernel: 16 NT channels per block, NT = cout / 16 up to 4 (48 and 96: NT = 3); bricks of
    // 16 x 8 x 4 voxels where their halo fits LIM of LDS, else 16 x 4 x 4
    const size_t tab_bytes = (size_t)(c->kpad / BK) * 4 * sizeof(int);
    const size_t halo8 = (size_t)hl_vox<8>() * c->cin * 2 + tab_bytes, halo4 = (size_t)hl_vox<4>() * c->cin * 2 + tab_bytes;
    const int bnt = c->cout % 16 ? 0 : c->cout <= 64 ? c->cout / 16 : c->cout % 48 == 0 ? 3 : c->cout % 64 == 0 ? 4 : 0;
    const bool bok = bnt && c->ntaps == 27 && !(c->flags & PVNET_GATHER) && c->s.d % BRZ == 0 && c->s.w % BRX == 0 &&
                     M >= (int64_t)256 * 2048;
    // the first of (16 x 8 x 4, weights in LDS), (16 x 8 x 4), (16 x 4 x 4, weights in LDS),
    // (16 x 4 x 4) whose LDS fits LIM: the taller brick first (more waves per CU for the same
    // LDS), then the weights in LDS (no global load in the k-loop).  A workgroup may take all
    // 160 KB of a CU's LDS; LIM = 112 KB measured best on the warm-start network's layers (taller
    // bricks pay even at one block per CU, weights in LDS only while about two blocks fit)
    const size_t wbytes = (size_t)16 * bnt * (2 * (size_t)c->kpad + 16);
    const size_t LIM = 112 * 1024;
    bool by8 = false, wl = false;
    size_t lds = 0;
    if (bok && c->s.h % 8 == 0 && halo8 + wbytes <= LIM) {
        by8 = wl = true;
        lds = halo8 + wbytes;
    } else if (bok && c->s.h % 8 == 0 && halo8 <= LIM) {
        by8 = true;
        lds = halo8;
    } else if (bok && c->s.h % 4 == 0 && halo4 + wbytes <= LIM) {
        wl = true;
        lds = halo4 + wbytes;
    } else if (bok && c->s.h % 4 == 0 && halo4 <= LIM) {
        lds = halo4;
    }
    if (lds) {
        const unsigned nbr = (unsigned)(M / (BRX * (by8 ? 8 : 4) * BRZ));
        const dim3 gb(nbr, c->cout / (16 * bnt));
#define PVNET_BRICK(NTV, BYV)                                                                         \
    do {                                                                                              \
        if (wl)                                                                                       \
            hipLaunchKernelGGL((k_conv_fwd_brick<NTV, BYV, true>), gb, dim3(64 * BYV), lds, st, *c, K);  \
        else                                                                                          \
            hipLaunchKernelGGL((k_conv_fwd_brick<NTV, BYV, false>), gb, dim3(64 * BYV), lds, st, *c, K); \
    } while (0)
        if (by8) {
            switch (bnt) {
                case 1: PVNET_BRICK(1, 8); break;
                case 2: PVNET_BRICK(2, 8); break;
                case 3: PVNET_BRICK(3, 8); break;
                default: PVNET_BRICK(4, 8); break;
            }
        } else {
            switch (bnt) {
                case 1: PVNET_BRICK(1, 4); break;
                case 2: PVNET_BRICK(2, 4); break;
                case 3: PVNET_BRICK(3, 4); break;
                default: PVNET_BRICK(4, 4); break;
            }
        }
#undef PVNET_BRICK
    } else if (c->ntaps == 1 && K <= BK && c->cout <= 64 && !(c->flags & PVNET_GATHER)) {
        const dim3 gp((unsigned)((M + 64 * PWT - 1) / (64 * PWT)));
        switch ((c->cout + 15) / 16) {
            case 1: hipLaunchKernelGGL((k_conv_pw<1>), gp, dim3(256), 0, st, *c, (int)M); break;
            case 2: hipLaunchKernelGGL((k_conv_pw<2>), gp, dim3(256), 0, st, *c, (int)M); break;
            case 3: hipLaunchKernelGGL((k_conv_pw<3>), gp, dim3(256), 0, st, *c, (int)M); break;
            default: hipLaunchKernelGGL((k_conv_pw<4>), gp, dim3(256), 0, st, *c, (int)M); break;
        }
    } else if (c->cout <= 16 && M >= (int64_t)256 * 2048) {
        hipLaunchKernelGGL((k_conv_fwd<16, 64>), dim3(blocks_for(M, 256), 1), dim3(256), 0, st, *c, (int)M, K);
    } else if (c->cout <= 16) {
        hipLaunchKernelGGL((k_conv_fwd<16, 32>), dim3(mt, 1), dim3(256), 0, st, *c, (int)M, K);
    } else if (c->cout % 64 == 0 && (int64_t)mt * (c->cout / 64) >= 2048) {
        hipLaunchKernelGGL((k_conv_fwd<64, 32>), dim3(mt, c->cout / 64), dim3(256), 0, st, *c, (int)M, K);
    } else {
        hipLaunchKernelGGL((k_conv_fwd<32, 32>), dim3(mt, (c->cout + 31) / 32), dim3(256), 0, st, *c, (int)M, K);
    }
    NET_LAUNCH();
    return 0;
}

int64_t pvnet_wgrad_ws_floats(const pvnet_shape* s, int32_t ntaps, int32_t cin, int32_t cout) {
    if (!shape_ok(s) || cin <= 0 || cout <= 0) return -1;
    // either kernel's partials fit (the call's flags pick one)
    return std::max(plan_wgrad(*s, ntaps, cin, cout, 0).ws_total, plan_wgrad(*s, ntaps, cin, cout, PVNET_GATHER).ws_total);
}

int pvnet_conv_wgrad(const pvnet_wgrad* a, void* stream) {
    if (!a || !shape_ok(&a->s)) return fail(1, "pvnet_conv_wgrad: bad shape");
    if (a->ntaps != 1 && a->ntaps != 27) return fail(1, "pvnet_conv_wgrad: ntaps must be 1 or 27");
    if (a->cin <= 0 || a->cin % 8 || a->cout <= 0 || a->cout % 8)
        return fail(1, "pvnet_conv_wgrad: cin and cout must be positive multiples of 8");
    if (a->ldx < a->cin || a->ldx % 8 || a->lddy < a->cout || a->lddy % 8) return fail(1, "pvnet_conv_wgrad: bad row lengths");
    if (!a->x || !a->dy || !a->ws || !a->dw || !al16(a->x) || !al16(a->dy))
        return fail(1, "pvnet_conv_wgrad: null or misaligned pointer");
    const int64_t M = voxels(a->s);
    if (M % 32) return fail(1, "pvnet_conv_wgrad: voxel count must be a multiple of 32");
    if (M >= (int64_t)1 << 31) return fail(1, "pvnet_conv_wgrad: too many voxels");
    hipStream_t st = (hipStream_t)stream;
    const WgradPlan p = plan_wgrad(a->s, a->ntaps, a->cin, a->cout, a->flags);
    const int K = a->ntaps * a->cin;
    dim3 grid(p.ktiles, p.ntiles, p.nsplit);
    int nsplit = p.nsplit;
    bool fused_db = false;
    if (p.brick) {
        // one round of resident blocks along the voxels (at most the plan's WB_SPLITS partial rows)
        const int ng = a->cin / 8;
        void (*kern)(pvnet_wgrad, int, int, int, int, int64_t) =
            a->cout == 32 ? (ng == 1 ? k_conv_wgrad_brick<1, 2> : k_conv_wgrad_brick<WB_NGB, 2>)
                          : (ng == 1 ? k_conv_wgrad_brick<1, 1> : k_conv_wgrad_brick<WB_NGB, 1>);
        int dev = 0, cus = 0, occ = 0;
        NET_HIP(hipGetDevice(&dev));
        NET_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        NET_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, 256, 0));
        const int gy = ng == 1 ? 1 : (ng + WB_NGB - 1) / WB_NGB;
        // one round of resident blocks (at most the plan's WB_SPLITS partial rows)
        const int want = std::max(1, std::min({p.nbricks, WB_SPLITS, std::max(1, cus * occ / gy)}));
        const int per = (p.nbricks + want - 1) / want;
        nsplit = (p.nbricks + per - 1) / per;
        // 16-output layers sum the bias gradient in the same pass (rows after the weight partials)
        fused_db = a->db && a->cout == 16;
        hipLaunchKernelGGL(kern, dim3(nsplit * gy), dim3(256), 0, st, *a, K, p.nbricks, per, gy,
                           fused_db ? p.ws_main : (int64_t)-1);
    } else if (p.bn == 16)
        hipLaunchKernelGGL(k_conv_wgrad<16>, grid, dim3(256), 0, st, *a, (int)M, K, p.mchunk);
    else if (p.bn == 32)
        hipLaunchKernelGGL(k_conv_wgrad<32>, grid, dim3(256), 0, st, *a, (int)M, K, p.mchunk);
    else
        hipLaunchKernelGGL(k_conv_wgrad<64>, grid, dim3(256), 0, st, *a, (int)M, K, p.mchunk);
    NET_LAUNCH();
    const int64_t len = (int64_t)a->cout * K;
    const SumJob jw{a->ws, a->dw, len, (int64_t)K, (int64_t)K, nsplit, (int)blocks_for(len, SUM_COLS)};
    SumJob jb{nullptr, nullptr, 0, 1, 1, 0, 0};
    if (fused_db) {
        jb = SumJob{a->ws + p.ws_main, a->db, (int64_t)a->cout, (int64_t)a->cout, (int64_t)a->cout, nsplit,
                    (int)blocks_for(a->cout, SUM_COLS)};
    } else if (a->db) {
        // the bias column sums read only dy: they run before the split sums, whose launch then
        // reduces the weight and the bias partials together
        if (a->cout > 256 || a->cout % 8) return fail(1, "pvnet_conv_wgrad: bias gradient needs cout <= 256, a multiple of 8");
        float* part = a->ws + p.ws_main;
        hipLaunchKernelGGL(k_colsum, dim3(p.cblk), dim3(256), 0, st, a->dy, a->lddy, a->cout, (int)M, p.cchunk, part);
        NET_LAUNCH();
        jb = SumJob{part, a->db, (int64_t)a->cout, (int64_t)a->cout, (int64_t)a->cout, p.cblk,
                    (int)blocks_for(a->cout, SUM_COLS)};
    }
    hipLaunchKernelGGL(k_sum_splits, dim3(jw.blocks + jb.blocks), dim3(256), 0, st, jw, jb);
    NET_LAUNCH();
    return 0;
}

int pvnet_pack_weights(const float* w32, int32_t ntaps, int32_t cin, int32_t cout, uint16_t* wt, uint16_t* wd,
                       void* stream) {
    if (!w32 || !wt || !wd || (ntaps != 1 && ntaps != 27) || cin <= 0 || cout <= 0)
        return fail(1, "pvnet_pack_weights: bad argument");
    const int kpad = kpad_of(ntaps, cin), kpad_d = kpad_of(ntaps, cout);
    const int64_t n = (int64_t)cout * kpad + (int64_t)cin * kpad_d;
    hipLaunchKernelGGL(k_pack, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, w32, ntaps, cin, cout,
                       kpad, kpad_d, wt, wd);
    NET_LAUNCH();
    return 0;
}

int pvnet_pool2_fwd(const pvnet_shape* s, int32_t c, const uint16_t* x, int32_t ldx, uint16_t* y, int32_t ldy,
                    void* stream) {
    if (int rc = pool_args(s, c, x, ldx, y, ldy, "pvnet_pool2_fwd")) return rc;
    const int64_t n = voxels(*s) / 8 * (c / 8);
    hipLaunchKernelGGL(k_pool2_fwd, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, *s, c, x, ldx, y, ldy);
    NET_LAUNCH();
    return 0;
}

int pvnet_pool2_bwd(const pvnet_shape* s, int32_t c, const uint16_t* dy, int32_t lddy, uint16_t* dx, int32_t lddx,
                    int32_t accumulate, void* stream) {
    if (int rc = pool_args(s, c, dy, lddy, dx, lddx, "pvnet_pool2_bwd")) return rc;
    const int64_t n = voxels(*s) * (c / 8);  // fine voxel chunks
    hipLaunchKernelGGL(k_pool2_bwd, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, *s, c, dy, lddy, dx,
                       lddx, accumulate ? 1 : 0, (const uint16_t*)nullptr, 0);
    NET_LAUNCH();
    return 0;
}

int pvnet_pool2_bwd_relu(const pvnet_shape* s, int32_t c, const uint16_t* dy, int32_t lddy, uint16_t* dx,
                         int32_t lddx, const uint16_t* relu_of, int32_t ldr, void* stream) {
    if (int rc = pool_args(s, c, dy, lddy, dx, lddx, "pvnet_pool2_bwd_relu")) return rc;
    if (!relu_of || !al16(relu_of) || ldr < c || ldr % 8) return fail(1, "pvnet_pool2_bwd_relu: bad relu_of");
    const int64_t n = voxels(*s) * (c / 8);  // fine voxel chunks
    hipLaunchKernelGGL(k_pool2_bwd, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, *s, c, dy, lddy, dx,
                       lddx, 1, relu_of, ldr);
    NET_LAUNCH();
    return 0;
}

int pvnet_up2_fwd(const pvnet_shape* s, int32_t c, const uint16_t* x, int32_t ldx, uint16_t* y, int32_t ldy,
                  void* stream) {
    if (int rc = pool_args(s, c, x, ldx, y, ldy, "pvnet_up2_fwd")) return rc;
    const int64_t n = voxels(*s) * (c / 8);  // fine voxel chunks
    hipLaunchKernelGGL(k_up2_fwd, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, *s, c, x, ldx, y, ldy);
    NET_LAUNCH();
    return 0;
}

int pvnet_up2_bwd(const pvnet_shape* s, int32_t c, const uint16_t* dy, int32_t lddy, uint16_t* dx, int32_t lddx,
                  void* stream) {
    if (int rc = pool_args(s, c, dy, lddy, dx, lddx, "pvnet_up2_bwd")) return rc;
    const int64_t n = voxels(*s) / 8 * (c / 8);
    hipLaunchKernelGGL(k_up2_bwd, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, *s, c, dy, lddy, dx, lddx);
    NET_LAUNCH();
    return 0;
}

int pvnet_relu_bwd(int64_t m, int32_t c, const uint16_t* a, int32_t lda, uint16_t* dx, int32_t lddx, void* stream) {
    if (m <= 0 || c <= 0 || c % 8 || lda < c || lddx < c || lda % 8 || lddx % 8 || !a || !dx || !al16(a) || !al16(dx))
        return fail(1, "pvnet_relu_bwd: bad argument");
    hipLaunchKernelGGL(k_relu_bwd, dim3(blocks_for(m, 256), c / 8), dim3(256), 0, (hipStream_t)stream, m, a, lda, dx, lddx);
    NET_LAUNCH();
    return 0;
}

int pvnet_wall_distance(const pvnet_shape* s, const uint8_t* code, float* dist, void* stream) {
    if (!shape_ok(s) || !code || !dist) return fail(1, "pvnet_wall_distance: bad argument");
    if (s->d > PVNET_EDT_MAX || s->h > PVNET_EDT_MAX || s->w > PVNET_EDT_MAX)
        return fail(1, "pvnet_wall_distance: box sizes are limited to PVNET_EDT_MAX");
    const int64_t n = voxels(*s);
    const int len[3] = {s->w, s->h, s->d};
    for (int axis = 0; axis < 3; axis++) {
        const int64_t lines = n / len[axis];
        const int threads = len[axis] <= 128 ? 64 : len[axis] <= 256 ? 32 : 16;  // LDS <= 33 KB
        const size_t lds = (size_t)threads * (len[axis] + 1) * sizeof(float);
        hipLaunchKernelGGL(k_edt_pass, dim3(blocks_for(lines, threads)), dim3(threads), lds, (hipStream_t)stream, *s,
                           axis, code, dist);
        NET_LAUNCH();
    }
    return 0;
}

int pvnet_features(const pvnet_shape* s, const uint8_t* code, const float* prm, const float* dist, uint16_t* x,
                   void* stream) {
    if (!shape_ok(s) || !code || !prm || !dist || !x || !al16(x)) return fail(1, "pvnet_features: bad argument");
    const int64_t n = voxels(*s);
    hipLaunchKernelGGL(k_features, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, *s, code, prm, dist,
                       x);
    NET_LAUNCH();
    return 0;
}

int64_t pvnet_loss_ws_doubles(int64_t m) { return m > 0 ? 2 * ((m + LOSS_BLOCK - 1) / LOSS_BLOCK) : -1; }

int pvnet_mse_loss(int64_t m, int32_t nc, const float* y, int32_t ldy, const uint16_t* t, const uint8_t* code,
                   const float* wc, uint16_t* dy, int32_t lddy, double* part, double* loss, void* stream) {
    if (m <= 0 || nc <= 0 || ldy < nc || lddy < nc || !y || !t || !code || !wc || !dy || !part || !loss)
        return fail(1, "pvnet_mse_loss: bad argument");
    hipStream_t st = (hipStream_t)stream;
    const int nblk = blocks_for(m, LOSS_BLOCK);
    hipLaunchKernelGGL(k_loss_partial, dim3(nblk), dim3(LOSS_BLOCK), 0, st, m, nc, y, ldy, t, code, wc, part);
    NET_LAUNCH();
    hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(LOSS_BLOCK), 0, st, nblk, nc, part, loss);
    NET_LAUNCH();
    hipLaunchKernelGGL(k_loss_grad, dim3(blocks_for(m, 256)), dim3(256), 0, st, m, nc, y, ldy, t, code, wc, loss, dy, lddy);
    NET_LAUNCH();
    return 0;
}

int pvnet_adam(int64_t n, float* p, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
               int32_t t, void* stream) {
    if (n <= 0 || !p || !g || !m || !v || t < 1) return fail(1, "pvnet_adam: bad argument");
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, n, p, g, m, v, lr, b1, b2,
                       eps, c1, c2);
    NET_LAUNCH();
    return 0;
}

}  // extern "C"
