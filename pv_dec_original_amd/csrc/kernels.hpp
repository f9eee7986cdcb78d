// kernels.hpp — launch parameter blocks and launcher declarations shared by the host C-ABI code
// (capi.cpp) and the gfx950 kernels (k_*.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pvlbm {

// Step-kernel flags
enum StepFlags : int {
    F_WSS = 1,     // evaluate WSS and its accumulators on next-to-wall nodes
    F_MACRO = 2,   // write rho/u of this launch's collision to the macro planes
    F_REDUCE = 4,  // write FP64 block partial sums (mass, inlet/outlet flux, non-finite count)
    F_INIT = 8,    // no gather/closure: start from f = w or f = f_eq(init_rho, init_u)
};

// One launch of the fused step kernel: the stream of iteration n-1, its closures and WSS, then the
// collision of iteration n.  Kinds of launch (StepParams::scheme):
//   SCH_PULL: two lattices; node x reads direction d from A[d][x - c_d] and writes B[d][x].
//   SCH_EVEN / SCH_ODD: one lattice (B == A), alternating (the AA pattern): an even launch reads
//   A[d][x] and writes A[opp d][x]; an odd launch reads A[opp d][x - c_d] and writes
//   A[d][x + c_d].  In either kind a node's relaxed populations go back to the slots it read,
//   and no two nodes address the same slot, so the update is in place.
// A link with no upstream node parks its post-closure value where the next launch reads it.
enum Scheme : int { SCH_PULL = 0, SCH_EVEN = 1, SCH_ODD = 2 };

// One active 64-node x-segment of the padded box.  Lane l of the tile is node 64*tile + l; the
// masks mark which lanes are lattice nodes, next-to-wall nodes and boundary (inlet/outlet/wall)
// nodes.  Side-table slots are assigned in node order, so the slot of a lane is the first slot of
// the tile plus the number of masked lanes below it.
struct alignas(16) TileDesc {
    int32_t tile;
    int32_t bslot;   // first boundary-table slot of the tile
    int32_t wslot;   // first next-to-wall-table slot of the tile
    int32_t pad;
    uint64_t node;
    uint64_t ntw;
    uint64_t bnd;
};

struct StepParams {
    const void* A;      // post-collision populations of the previous launch (allocation base)
    void* B;            // post-collision populations written by this launch (allocation base)
    int64_t ns;         // elements per direction plane: guard + padded node count + guard
    int64_t guard;      // guard elements before node 0 of each plane (> largest stencil offset)
    int64_t sy, sz;     // index strides of y and z (x stride 1)
    // Byte offsets from a lattice base: plane d starts at poff_plane * d; the pull of direction d
    // (and the stale slot it reads) sits poff_pull[d] from the base plus the node's byte index, so
    // a wave addresses every plane from one 64-bit scalar base per direction and one 32-bit lane
    // offset, with no 32-bit limit on the size of a plane.
    int64_t poff_plane;
    int64_t poff_pull[27];
    int64_t poff_push[27];  // plane d, shifted by +c_d
    const TileDesc* tdesc;  // active tiles
    int32_t ntiles;
    int32_t blocks;         // launched blocks (4 tiles each)
    int32_t flags;
    int32_t stream_stores;  // 1: nontemporal population stores (per-step traffic above the cache)
    int32_t stream_loads;   // 1: nontemporal population loads (in place with plain stores only)
    int32_t scheme;         // enum Scheme
    int32_t reverse;        // 1: blocks take their tiles from the end (a sweep in the other direction)
    // boundary table (inlet, outlet and wall nodes)
    const uint32_t* bmask;  // bit d set: neighbour x + c_d is not a node
    const uint8_t* bclose;  // enum BClose
    const double* bprof;    // inlet velocity profile
    const double* brho;     // prescribed outlet density
    const float* bq;        // [slot][Q] wall-link fractions for the interpolated closure, or NULL
    // next-to-wall table
    void* M;                // previous (rho, ux, uy, uz) [4][mst] (arithmetic type)
    int64_t mst;
    const double* nrm;      // normals [3][mst]
    double* wss;            // [mst]
    double* tau3;           // [3][mst]
    double* acc;            // [4][mst]
    // scalars
    double omega;           // 1/tau
    double scale;           // inlet velocity scale of the iteration being closed
    double C;               // WSS constant 1060 dt^2 / (h^2 1e-6)
    // outputs
    double* mrho;           // macro planes [4][ns] (rho, ux, uy, uz), F_MACRO
    double* partials;       // [blocks][4], F_REDUCE
    unsigned long long* gate;  // stream-gate events
    const double* init_rho; // F_INIT warm start (padded layout), NULL = rest
    const double* init_u;   // [3][ns]
};

// Voxeliser launch: one wave per (i, j) column, candidate triangles in CSR form.
struct VoxParams {
    const double* tri;       // [T][12]
    const int64_t* col_off;  // [nx*ny + 1]
    const int32_t* col_tri;  // candidate triangle ids
    double ox, oy, oz, h;
    double bb[6];
    int32_t nx, ny, nz;
    uint8_t* fluid;          // nx*ny*nz, x fastest
};

// storage codes match enum pvlbm_storage
// one step over all active tiles
hipError_t launch_step(int q, int storage, const StepParams& p, hipStream_t s);
// Moves the state left by a launch of kind `from` to the layout a launch of kind `to` leaves, from
// p.A into p.B (distinct arrays); from/to: SCH_PULL -> SCH_EVEN, or SCH_EVEN / SCH_ODD -> SCH_PULL.
hipError_t launch_relayout(int q, int storage, const StepParams& p, int from, int to, hipStream_t s);
hipError_t launch_reduce_final(const double* partials, int nparts, double* out4, hipStream_t s);
hipError_t launch_voxelize(const VoxParams& p, hipStream_t s);
// boundary flags: per padded node, 1 if fluid and a Q neighbour is missing (padded box nxp x nyp
// x nz holding the nx x ny x nz grid)
hipError_t launch_boundary(int q, const uint8_t* fluid_pad, int64_t nxp, int64_t nyp, int64_t nx,
                           int64_t ny, int64_t nz, uint8_t* bflag, hipStream_t s);
// next-to-wall marking: FLUID nodes with a WALL Q-neighbour become NTW
hipError_t launch_ntw(int q, uint8_t* kind, int64_t nxp, int64_t nyp, int64_t nz, hipStream_t s);
// per-triangle unit normal, area, centroid (Subdomain.cpp:244-267)
hipError_t launch_tri_prep(const double* tri, int64_t ntri, double* prep /*[8][ntri]*/, hipStream_t s);
// area/d^2-weighted normal of each listed node (Subdomain.cpp:301-369)
hipError_t launch_normals(const double* prep, int64_t ntri, const double* xyz /*[3][n]*/, int64_t n,
                          double* nrm /*[3][n]*/, hipStream_t s);

// Link fractions of wall links (Geometry::CalculateWallDistances, ComplexGeometry.cpp:550-608):
// for each segment seg[6*p .. 6*p+5] = (start, end), the triangle of smallest file index that
// FindLineTriangleIntersection (:689-730) accepts (the reference takes the first in file order)
// gives q = |intersection - start| / h; q = -1 without one (nearest = 1: the intersection
// closest to the start instead).  Candidates come from a uniform grid of cells (size cs, origin
// og, nc cells per axis) in CSR form: cell c holds cell_tri[cell_start[c] .. cell_start[c+1]),
// ascending.
hipError_t launch_link_q(const double* seg, int64_t npairs, const double* tri, const int32_t* cell_start,
                         const int32_t* cell_tri, const double og[3], double cs, const int32_t nc[3], double h,
                         int nearest, double* q, hipStream_t s);

// Device address of `device`'s scratch array of PVLBM_NOALLOC_SCRATCH_BYTES (noalloc.hip): looked
// up once per device with that device current, then cached.
hipError_t noalloc_scratch(int device, void** ptr);

// streaming copy of `bytes` (multiple of 16) for counter calibration
hipError_t launch_copy(const void* src, void* dst, int64_t bytes, hipStream_t s);

// Dense initial fields out[4][ns] (rho, ux, uy, uz per padded slot) sampled trilinearly from a
// coarse FP32 field (rho_c[cz][cy][cx], u_c[3][cz][cy][cx], voxel centres at fine coordinates
// (i + 0.5) n / c - 0.5), densities clipped into [rlo, rhi]; padding slots and non-finite samples
// (counted in *nonfin) get rest.  With flux_x > 0 the velocity of every plane x is scaled so that
// the flux of u_x over its fluid nodes (kind != 0) is flux_x, or targets[x] when targets (nx
// device doubles) is given (flux: nx doubles of scratch); speeds are then clipped to umax.  With
// drho (nx device doubles) every slot of plane x gets drho[x] added to its density, clipped again.
hipError_t launch_expand_coarse(const float* rho_c, const float* u_c, int cx, int cy, int cz, int64_t nx,
                                int64_t ny, int64_t nz, int64_t nxp, int64_t nyp, int64_t ns, double umax,
                                double rlo, double rhi, const uint8_t* kind, double flux_x, const double* targets,
                                const double* drho, double* flux, double* out, unsigned long long* nonfin,
                                hipStream_t s);

// Compact FP32 box fields from the macro planes (macro[4][ns], padded slots): rho[b] and
// u[c][b] for every box point b = i + nx (j + ny k); points that are not nodes (kind 0) get
// rho = 1, u = 0.
hipError_t launch_macros_f32(const double* macro, const uint8_t* kind, int64_t nx, int64_t ny, int64_t nz,
                             int64_t nxp, int64_t nyp, int64_t ns, float* rho, float* u, hipStream_t s);

// out[(k ny + j) nx + i] = kind of box point (ix[i], iy[j], iz[k]) (padded layout nxp x nyp);
// the indices must lie in the box (device arrays).
hipError_t launch_sample_kinds(const uint8_t* kind, int64_t nxp, int64_t nyp, const int32_t* ix, int nx,
                               const int32_t* iy, int ny, const int32_t* iz, int nz, uint8_t* out, hipStream_t s);

}  // namespace pvlbm
