/*
 * pvlbm.h — C ABI of the MI355X-native lattice Boltzmann engine for vessel blood flow.
 *
 * The reference solver (Resacd/PV_Dec_Original) exposes no FFI: its interface is four C++
 * constructors called from main() (LBM.cpp:69-72).  Each entry point below replaces one step of
 * that object pipeline; the comment on each names the reference code it stands in for.
 *
 * Conventions
 *   - Every call returns an int status: PVLBM_OK (0) or a positive error code.  The text of the
 *     last error is returned by pvlbm_last_error(handle) (or pvlbm_last_error(NULL) for calls that
 *     take no handle).  No C++ exception crosses this boundary.
 *   - Host pointers are owned by the caller and are copied synchronously; every call has finished
 *     its device work when it returns.  A handle owns all of its device memory and its own HIP
 *     stream; use one handle per host thread.
 *   - Box arrays are dense over the grid with x fastest: index = i + nx*(j + ny*k), the node at
 *     (origin_x + i*h, origin_y + j*h, origin_z + k*h) (ComplexGeometry.cpp:337-340).
 *   - Triangles are 12 doubles each: v1, v2, v3, then the facet normal exactly as the STL file
 *     gives it (Struct.h:117-125 order; the normal is not recomputed, GeometryReader.cpp:249).
 */
#ifndef PVLBM_H
#define PVLBM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pvlbm_sim* pvlbm_handle;

enum pvlbm_status_code {
    PVLBM_OK = 0,
    PVLBM_EINVAL = 1, /* bad argument */
    PVLBM_EHIP = 2,   /* HIP runtime error (no GPU, launch failure, ...) */
    PVLBM_ENOMEM = 3, /* device or host allocation failed */
    PVLBM_EIO = 4,    /* file could not be read */
    PVLBM_ESTATE = 5  /* call not valid in the handle's current state */
};

/* Population storage of a handle.  FP64 is the plain path (parity with the reference).  FP16S and
 * BF16S store the shifted populations f_i - w_i in 16 bits and collide in FP32; global sums stay
 * in FP64.  The shifted forms have no reference counterpart. */
enum pvlbm_storage { PVLBM_FP64 = 0, PVLBM_FP16S = 1, PVLBM_BF16S = 2 };

/* Roles of an open boundary (ComplexGeometry.h:46-49: inlet, outlet 1..3). */
enum pvlbm_role { PVLBM_INLET = 0, PVLBM_OUTLET1 = 1, PVLBM_OUTLET2 = 2, PVLBM_OUTLET3 = 3 };

/* Closure applied on a face.  INLET_XMIN and OUTLET_YMAX / OUTLET_ZMIN are the reference's
 * formulas (Boundary.cpp:129-157 and :87-123; Simulation.cpp:642-695 and :710-836).  OUTLET_XMAX
 * has no reference formula: for D3Q19 it is OUTLET_YMAX under the x<->y index permutation, for
 * D3Q27 the mirror image of the D3Q27 inlet with the density prescribed. */
enum pvlbm_closure {
    PVLBM_INLET_XMIN = 0,
    PVLBM_OUTLET_YMAX = 1,
    PVLBM_OUTLET_ZMIN = 2,
    PVLBM_OUTLET_XMAX = 3
};

/* Node kinds reported by pvlbm_get_kinds. */
enum pvlbm_kind {
    PVLBM_SOLID = 0,  /* not a lattice node (NodeType::NotFluid) */
    PVLBM_FLUID = 1,  /* interior fluid node */
    PVLBM_NTW = 2,    /* interior fluid node next to a wall node (Subdomain.cpp:201-218) */
    PVLBM_KINLET = 3, /* inlet boundary node */
    PVLBM_KOUTLET = 4,/* outlet boundary node (any of outlets 1..3) */
    PVLBM_WALL = 5    /* wall boundary node */
};

/* Lists that pvlbm_count / pvlbm_list report. */
enum pvlbm_list_id {
    PVLBM_LIST_FLUID = 0,   /* every lattice node */
    PVLBM_LIST_INLET = 1,
    PVLBM_LIST_OUTLET1 = 2,
    PVLBM_LIST_OUTLET2 = 3,
    PVLBM_LIST_OUTLET3 = 4,
    PVLBM_LIST_WALL = 5,
    PVLBM_LIST_NTW = 6,
    PVLBM_LIST_TILES = 7    /* active 64-node tiles of the step kernel */
};

/* A face specification: the reference hard-codes four planes of PV.stl
 * (ComplexGeometry.cpp:1378-1408); this generalises each to {axis, plane, tolerance, window}.
 * A boundary node at (c0, c1, c2) matches when |c[axis] - P| <= tol and, for the two other
 * axes a < b in increasing order, lo[0] < c[a] < hi[0] and lo[1] < c[b] < hi[1] (open intervals),
 * compared in long double as the reference does, with P = (long double)plane + (long double)plane_lo:
 * the reference's planes are long double literals (-16.9538L, -170.016L), which a double pair
 * carries exactly and a single double does not (plane_lo = 0 for a plane that is a double).
 * Matching is tried in the order inlet, outlet1, outlet2, outlet3, and among the faces of one role
 * in array order (Subdomain.cpp:124-155); a boundary node that matches nothing is a wall node. */
typedef struct pvlbm_face {
    int32_t role;    /* enum pvlbm_role */
    int32_t closure; /* enum pvlbm_closure */
    int32_t axis;    /* 0 = x, 1 = y, 2 = z */
    int32_t reserved;
    double plane;
    double tol;
    double lo[2];
    double hi[2];
    double rho;      /* prescribed density of a pressure outlet (the reference uses 1.0) */
    double plane_lo; /* low part of the plane coordinate (see above); 0 for a double plane */
} pvlbm_face;

/* A lattice grid: node (i,j,k) sits at origin + (i,j,k)*h. */
typedef struct pvlbm_grid {
    double origin[3];
    int32_t dims[3];
    int32_t reserved;
} pvlbm_grid;

/* Text of the last error.  h == NULL gives the calling thread's last handle-less error. */
const char* pvlbm_last_error(pvlbm_handle h);

/* Number of HIP devices visible (0 when none).  No reference counterpart. */
int pvlbm_device_count(int* count);

/* Parse an ASCII STL file.  tri == NULL returns only the count.  Malformed facets are skipped
 * silently, as the reference does.  Replaces Reader::readSTLASCII (GeometryReader.cpp:33-117). */
int pvlbm_read_stl_ascii(const char* path, double* tri, int64_t capacity, int64_t* ntri);

/* Parse a binary STL file (80-byte header, uint32 count, 50-byte facets); facet normals are
 * normalised in single precision as the reference does (a zero normal becomes NaN).  A file
 * shorter than its facet count is an error (PVLBM_EIO), where the reference would read past the
 * end.  Replaces Reader::readSTLBinary (GeometryReader.cpp:119-176). */
int pvlbm_read_stl_binary(const char* path, double* tri, int64_t capacity, int64_t* ntri);

/* Bounding box {min_x,max_x,min_y,max_y,min_z,max_z} by the reference rule: x over all three
 * vertices, y and z over vertices 1 and 2 only.  Replaces the bbox part of
 * Geometry::GetSubBoundingBox (ComplexGeometry.cpp:142-154). */
int pvlbm_bbox(const double* tri, int64_t ntri, double bbox[6]);

/* The reference grid of a bbox: origin = min corner, nx = ceil(dx/h)+1, ny = ceil(dy/h)+1,
 * nz = ceil(dz/h) (ComplexGeometry.cpp:163-165). */
int pvlbm_grid_from_bbox(const double bbox[6], double h, pvlbm_grid* grid);

/* Ray-cast voxelisation on the GPU: a node is fluid when the +z ray from (x, y, z + h/2) crosses
 * the mesh an odd number of times, with every floating-point step of the reference
 * (Geometry::labelPoints and helpers, ComplexGeometry.cpp:620-674, 803-1053).  grid == NULL uses
 * the reference grid of the mesh bbox; otherwise the given grid is used and the ray's bbox is the
 * union of the mesh bbox and the grid.  grid_out receives the grid used.  fluid_out (may be NULL
 * to query the grid) receives nx*ny*nz bytes, 1 = fluid, 0 = not fluid.  Replaces
 * GetSubBoundingBox + generatePointsInSubBox + labelPoints of the Geometry constructor
 * (ComplexGeometry.cpp:8-22). */
int pvlbm_voxelize(const double* tri, int64_t ntri, double h, const pvlbm_grid* grid,
                   pvlbm_grid* grid_out, uint8_t* fluid_out);

/* Work of this thread's last pvlbm_voxelize call that labelled points: the number of ray-triangle
 * tests (every point of a column against the column's candidate triangles) and the labelling
 * kernel's time in milliseconds (HIP events).  No counterpart in the reference (whose
 * labelPoints tests every point against all triangles, ComplexGeometry.cpp:135-136, 1039). */
int pvlbm_voxelize_stats(int64_t* tests, double* kernel_ms);

/* Build a device-resident simulation from voxel flags: boundary detection over the Q
 * neighbours (ComplexGeometry.cpp:1259-1300), face classification (:1369-1429), node kinds
 * (Subdomain.cpp:91-169), neighbour links (:173-197), next-to-wall set (:201-218), area/d^2
 * weighted normals over all triangles (:224-370), f = w initialisation (Subdomain.cpp:549-564,
 * Node.cpp:30-37) and the first collision.  lattice is 19 or 27.  tau and dt play the roles of
 * Simulation(SubDomain&, tau, pressure, del_t) (Simulation.h:26); h enters the WSS constant as
 * the reference's spacing (Simulation.cpp:880).  Replaces SubDomain::SubDomain and the
 * set-up half of Simulation::Simulation.  The populations take one lattice of Q planes over the
 * padded box (two when the streaming scheme takes the pull, pvlbm_set_streaming). */
int pvlbm_create(const uint8_t* fluid, const pvlbm_grid* grid, double h, int lattice, int storage,
                 const pvlbm_face* faces, int nfaces, const double* tri, int64_t ntri, double tau,
                 double dt, pvlbm_handle* out);

/* Change the relaxation time (Simulation::SetRelaxation, Simulation.cpp:867-869). */
int pvlbm_set_tau(pvlbm_handle h, double tau);

/* Change the time step del_t that enters the WSS constant 1060*dt^2/(h^2*1e-6)
 * (Simulation::m_dt, Simulation.cpp:10 and :880). */
int pvlbm_set_dt(pvlbm_handle h, double dt);

/* Size and members (box indices) of a node list. */
int pvlbm_count(pvlbm_handle h, int list, int64_t* n);
int pvlbm_list(pvlbm_handle h, int list, int64_t* idx);

/* Node kind of every box point (enum pvlbm_kind), nx*ny*nz bytes. */
int pvlbm_get_kinds(pvlbm_handle h, uint8_t* kinds);

/* Node kinds on a separable grid of box points: out[(k ny + j) nx + i] = kind of point
 * (ix[i], iy[j], iz[k]); every index must lie in the box.  No counterpart in the reference: a
 * coarse view of the kinds (the warm-start network's input) without copying the whole box. */
int pvlbm_sample_kinds(pvlbm_handle h, const int32_t* ix, int32_t nx, const int32_t* iy, int32_t ny,
                       const int32_t* iz, int32_t nz, uint8_t* out);

/* As pvlbm_sample_kinds, allocating nothing.  No counterpart in the reference: it serves callers
 * that cannot allocate device memory (serving loops that sample many handles, harnesses that
 * forbid hipMalloc).  The sample indices and the sampled kinds are staged in a device array of
 * PVLBM_NOALLOC_SCRATCH_BYTES (8 MiB) that every GPU the library runs on holds from the first
 * pvlbm_create on it until the process ends; a request needing more than that, 4 * (nx + ny + nz)
 * bytes rounded up to 256 plus nx * ny * nz bytes, fails with PVLBM_EINVAL.  All calls on one GPU
 * share that array (each writes every byte it reads back), so calls that overlap in time on one
 * GPU race. */
#define PVLBM_NOALLOC_SCRATCH_BYTES (8 << 20)
int pvlbm_sample_kinds_noalloc(pvlbm_handle h, const int32_t* ix, int32_t nx, const int32_t* iy, int32_t ny,
                               const int32_t* iz, int32_t nz, uint8_t* out);

/* Inlet velocity of each inlet-list node; the x-velocity applied at step s is
 * inlet_scale[s] * u[node].  Default 1.0 everywhere, which makes inlet_scale the reference's
 * uniform InletVelocity_ (Simulation.cpp:580, Boundary.cpp:131-133). */
int pvlbm_set_inlet_profile(pvlbm_handle h, const double* u);

/* Restart from f = f_eq(rho, u) (warm start; no reference counterpart).  rho and u are dense box
 * arrays (u as three planes x, y, z); rho == NULL means rest (f = w, the reference start,
 * Node.cpp:30-37).  Resets the step counter, the WSS accumulators and the stale populations;
 * those (populations of links without an upstream node, which the reference never writes after
 * its start from rest) take the rest value w_i in both cases, so a start from given fields
 * converges to the same steady state as a start from rest. */
int pvlbm_init_fields(pvlbm_handle h, const double* rho, const double* u);

/* As pvlbm_init_fields from a coarse FP32 field in device memory (on this handle's device; the call
 * first waits for all work on the device, so fields still being produced on any stream of the
 * caller are complete before they are read): rho_c[cz][cy][cx] and
 * u_c[3][cz][cy][cx] (x fastest; cdims = {cx, cy, cz}) cover the grid, coarse voxel (i, j, k)
 * centred at fine coordinates ((i + 0.5) nx / cx - 0.5, ...); each node takes the trilinear
 * sample (clamped at the edges) with its density clipped into [rho_lo, rho_hi].  With
 * flux_x > 0 the velocity of every plane x is then scaled (by a factor within [1/4, 4]) so that
 * the flux of u_x over the plane's fluid nodes equals flux_x: for a grid whose inlet is the x-min
 * face and whose outlets are all on the x-max face, mass conservation makes the inlet flux the
 * flux of every plane.  plane_flux (host, nx values, may be NULL) gives a target per plane
 * instead, for a start at an instant where the planes' fluxes differ (the compressible part of a
 * pulsatile flow); plane_drho (host, nx values, may be NULL) is added to the density of every
 * node of plane x, which is then clipped again.  Speeds are finally clipped to umax.  Nodes whose
 * sample is not finite start at rest; their number is returned in *nonfinite.  No counterpart in
 * the reference (it always starts from rest, Simulation.cpp:13-15): the start from the learned
 * warm-start network, without a host copy of the full-size field. */
int pvlbm_init_fields_coarse(pvlbm_handle h, const float* rho_c, const float* u_c,
                             const int32_t cdims[3], double umax, double rho_lo, double rho_hi,
                             double flux_x, const double* plane_flux, const double* plane_drho,
                             int64_t* nonfinite);

/* Link fractions of the wall nodes (list PVLBM_LIST_WALL, list order): for each direction
 * d < ndirs whose neighbour is in the box and solid, the segment from the node's point to point +
 * c_d * h is tested against the triangles, and the first one in file order that it intersects
 * gives q = distance / h.  q_out[nwall][Q]; -1 where there is no such link.  ndirs = 19 is the
 * reference's loop (it hard-codes 19 directions, also for D3Q27); ndirs = Q covers every
 * direction.  Replaces Geometry::CalculateWallDistances (ComplexGeometry.cpp:550-608) with
 * FindLineTriangleIntersection (:689-730); the reference computes these and never uses them. */
int pvlbm_link_fractions(pvlbm_handle h, int ndirs, double* q_out);

/* Closure of the wall nodes.  PVLBM_WALL_REFERENCE (default): the reference's rule,
 * f[opp d] = f[d] on the post-stream populations in direction order (Simulation.cpp:843-863).
 * PVLBM_WALL_INTERPOLATED: linear interpolated bounce-back (Bouzidi, Firdaouss and Lallemand
 * 2001) on each link from a wall node to a solid neighbour, with q the fraction of the link at
 * which the segment from the node meets the surface first; the segment starts at the node
 * shifted by h/2 in z, where the voxelisation tests the node (ComplexGeometry.cpp:989-1003).
 * For q < 1/2: f_k = 2q f*_i(x) + (1 - 2q) f*_i(x - c_i); for q >= 1/2:
 * f_k = f*_i(x)/(2q) + (2q - 1)/(2q) f*_k(x)  (c_i towards the wall, k = opp i, f* the
 * post-collision populations); a link without a surface crossing, or whose other neighbour is
 * solid too, takes halfway bounce-back f_k = f*_i(x).  No counterpart in the reference: it
 * computes link fractions (ComplexGeometry.cpp:550-608) but never uses them. */
enum { PVLBM_WALL_REFERENCE = 0, PVLBM_WALL_INTERPOLATED = 1 };
int pvlbm_set_wall_closure(pvlbm_handle h, int mode);

/* Streaming scheme of the step (no counterpart in the reference, whose Streaming copies each
 * population into the neighbour's array slot, Simulation.cpp:211-237; results are identical under
 * every scheme).  In place, on one lattice (the AA pattern: alternate launches read and write each
 * node's own slots, then its neighbours' slots), a step moves the same bytes as the pull from one
 * lattice into another but touches half the memory, so a lattice near the size of the MI355X's
 * 256 MB last-level cache is read from that cache; far beyond it the pull is faster.
 * PVLBM_STREAMING_AUTO (default): in place while the lattice, Q * bytes per population * nodes,
 * is at most PVLBM_STREAMING_AUTO_INPLACE_MAX_BYTES (512 MiB), the pull beyond.
 * PVLBM_STREAMING_TWO_LATTICES / _IN_PLACE force one (in place halves the population memory, for
 * domains too large for two lattices).  The interpolated wall closure (pvlbm_set_wall_closure)
 * reads the node's previous populations from its own slots and always takes the pull; it cannot
 * be combined with PVLBM_STREAMING_IN_PLACE (PVLBM_EINVAL).  A change of scheme moves the current
 * state into the new layout at once (the second lattice is allocated when first needed);
 * pvlbm_get_streaming reports the scheme in use (1 = in place). */
enum { PVLBM_STREAMING_AUTO = 0, PVLBM_STREAMING_TWO_LATTICES = 1, PVLBM_STREAMING_IN_PLACE = 2 };
#define PVLBM_STREAMING_AUTO_INPLACE_MAX_BYTES (512ll << 20)
int pvlbm_set_streaming(pvlbm_handle h, int mode);
int pvlbm_get_streaming(pvlbm_handle h, int* in_place);

/* How the step writes and reads the populations (no counterpart in the reference; results are
 * identical under every policy).  PVLBM_STORES_AUTO (default), with two lattices: plain stores
 * while one step's traffic, 2 * Q * bytes per population * nodes, stays within
 * PVLBM_STORES_AUTO_BYTES (128 MiB), so that the next launch reads the lattice from the MI355X's
 * 256 MB last-level cache, and nontemporal stores beyond, which keep the written populations
 * from evicting the lines the gathers are about to read.  In place: plain stores, every other
 * launch sweeping the lattice in reverse so that it starts on the lines the previous launch
 * touched last, and nontemporal loads by the size of the lattice's lines that hold nodes (lines
 * per plane x 128 B x Q): none up to PVLBM_LOADS_AUTO_INPLACE_BYTES (180 MiB), in every launch up
 * to PVLBM_LOADS_AUTO_EVEN_BYTES (240 MiB), in the even launches (those that read and write each
 * node's own slots) beyond.  PVLBM_STORES_REGULAR / _STREAMING force plain or nontemporal stores,
 * with plain loads.  pvlbm_get_store_policy reports the stores in use (1 = nontemporal),
 * pvlbm_get_load_policy the loads (0 = plain, 1 = nontemporal in the even launches, 2 = in every
 * launch). */
enum { PVLBM_STORES_AUTO = 0, PVLBM_STORES_REGULAR = 1, PVLBM_STORES_STREAMING = 2 };
#define PVLBM_STORES_AUTO_BYTES (128ll << 20)
#define PVLBM_LOADS_AUTO_INPLACE_BYTES (180ll << 20)
#define PVLBM_LOADS_AUTO_EVEN_BYTES (240ll << 20)
int pvlbm_set_store_policy(pvlbm_handle h, int policy);
int pvlbm_get_store_policy(pvlbm_handle h, int* streaming);
int pvlbm_get_load_policy(pvlbm_handle h, int* streaming);

/* Every how many iterations WSS and its accumulators are evaluated (reference: every iteration,
 * Simulation.cpp:585-586).  Default 1, the reference's quantities.  With every > 1 the accumulators
 * sum only the evaluated iterations (index multiple of every), so the OSI and TAWSS that
 * pvlbm_get_wss reports are those of the sampled sequence (TAWSS = sum over the samples of |tau|
 * / 0.9), not the reference's every-iteration sums. */
int pvlbm_set_wss_every(pvlbm_handle h, int every);

/* Run n iterations of Simulation::Iterate's body (Simulation.cpp:577-586: collide, stream, inlet,
 * outlets, walls, WSS).  inlet_scale holds n values (NULL = all 1.0). */
int pvlbm_step(pvlbm_handle h, int n, const double* inlet_scale);

/* As pvlbm_step, also returning the elapsed milliseconds of the launches measured with HIP events
 * on the handle's stream. */
int pvlbm_step_timed(pvlbm_handle h, int n, const double* inlet_scale, double* ms);

/* Iterations completed so far. */
int pvlbm_iteration(pvlbm_handle h, int64_t* it);

/* Density and velocity of the current iteration n, computed as Simulation::Collision does from
 * the populations after iteration n-1 (Simulation.cpp:52-105).  Dense box arrays; u holds three
 * planes.  Solid points get 0.  Either pointer may be NULL. */
int pvlbm_get_macros(pvlbm_handle h, double* rho, double* u);

/* The same fields as FP32 into caller-owned device memory on this handle's device: rho[nbox],
 * u[3][nbox] (dense box order); points that are not nodes get rho = 1, u = 0, so the arrays can
 * be sampled as a coarse field by pvlbm_init_fields_coarse of another handle.  No counterpart in
 * the reference (which has no device memory); used for coarse-grid pre-runs. */
int pvlbm_get_macros_device(pvlbm_handle h, float* rho, float* u);

/* Per next-to-wall node, in list order: x, y, z, WSS, OSI, TAWSS — the columns of
 * FlowHemodynamics/<iter>_FlowHemodynamics.csv (OutputFile.cpp:867-938), from the last WSS
 * evaluation (Simulation.cpp:880-905). */
int pvlbm_get_wss(pvlbm_handle h, double* out6);

/* Unit normals of the next-to-wall nodes, n3[3][nw] in next-to-wall list order: the
 * area/distance^2-weighted sum over every triangle of SubDomain::calculateBoundaryNormals
 * (Subdomain.cpp:224-370), as the WSS evaluation uses them; zero for a handle created without
 * triangles. */
int pvlbm_get_normals(pvlbm_handle h, double* n3);

/* The WSS column of pvlbm_get_wss alone (nw doubles in next-to-wall list order: the magnitude of
 * Simulation.cpp:880-898 from the last WSS evaluation), without coordinates, OSI and TAWSS, for
 * callers that sample WSS often. */
int pvlbm_get_wss_magnitude(pvlbm_handle h, double* out);

/* Raw WSS state per next-to-wall node: tau[3][nw] of the last evaluation and the accumulators
 * acc[4][nw] = (sum tau_x, sum tau_y, sum tau_z, sum |tau|). Either pointer may be NULL. */
int pvlbm_get_wss_raw(pvlbm_handle h, double* tau3, double* acc4);

/* FP64 global sums over the current iteration's macroscopic state (no reference counterpart):
 * out[0] = sum rho over all nodes, out[1] = sum rho*u_x over inlet nodes, out[2] = sum rho*u_x over
 * outlet nodes, out[3] = number of nodes whose rho or u is not finite. */
int pvlbm_reduce(pvlbm_handle h, double out[4]);

/* Number of gathered populations that met the reference's "f_new != 0.0" stream gate
 * (Simulation.cpp:222) on a link whose stale value is not kept (FP64); nonzero means a run left the
 * reference's semantics. */
int pvlbm_gate_events(pvlbm_handle h, int64_t* n);

/* Reference inlet waveform: piecewise polynomial of period `period`
 * (Simulation::getInletVelocity, Simulation.cpp:548-563; pulsatile.h:12-28). */
double pvlbm_pulsatile_velocity(double t, double period);

int pvlbm_destroy(pvlbm_handle h);

/* Measurement helper (no reference counterpart): run `reps` launches of a plain streaming copy of
 * `bytes` bytes (rounded down to 16) between two device buffers and return the elapsed
 * milliseconds.  Its known traffic (bytes read + bytes written per launch) calibrates hardware
 * traffic counters and gives the attainable copy bandwidth. */
int pvlbm_copy_bench(int64_t bytes, int reps, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* PVLBM_H */
