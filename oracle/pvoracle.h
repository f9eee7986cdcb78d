/*
 * pvoracle.h — TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference solver (Resacd/PV_Dec_Original) used to check the HIP library.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and only as the
 * checker or the timed CPU baseline; the product path never links or calls it.
 *
 * Every function follows the reference file:line cited beside it, in source operation order, and
 * is compiled with -ffp-contract=off so that its rounding is the strict-IEEE evaluation of that
 * order.  Parity pins (checks that do not restate the formula) live in tests/test_oracle_*.py.
 */
#ifndef PVORACLE_H
#define PVORACLE_H

#include <stdint.h>
#include "../include/pvlbm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Reader::readSTLASCII (GeometryReader.cpp:33-117). */
int64_t orc_read_stl_ascii(const char* path, double* tri, int64_t capacity);

/* Reader::readSTLBinary (GeometryReader.cpp:119-176); -1 when the file cannot be opened. */
int64_t orc_read_stl_binary(const char* path, double* tri, int64_t capacity);

/* bbox rule (ComplexGeometry.cpp:142-154) and grid dims (:163-165). */
void orc_bbox(const double* tri, int64_t ntri, double bb[6]);
void orc_grid_dims(const double bb[6], double h, int32_t dims[3]);

/* isPointInsideTriangle (ComplexGeometry.cpp:620-674). */
int orc_point_in_triangle(double x, double y, double z, const double* tri12);

/* countIntersectinsTriangle (ComplexGeometry.cpp:803-848) for one line. */
int orc_count_crossings(const double* tri, int64_t ntri, const double start[3],
                        const double end[3]);

/* labelPoints (ComplexGeometry.cpp:978-1053) with adjustLineCoordinates (:853-884), testing
 * every triangle for every point. fluid: dims[0]*dims[1]*dims[2] bytes, x fastest. */
void orc_label_points(const double* tri, int64_t ntri, const double bb[6], const int32_t dims[3],
                      const double origin[3], double h, uint8_t* fluid, int threads);

/* Same labels with candidate triangles binned per (i,j) column by their xy extent (a CPU
 * speed-up of the same per-triangle test; tests check it equals orc_label_points). */
void orc_label_points_binned(const double* tri, int64_t ntri, const double bb[6],
                             const int32_t dims[3], const double origin[3], double h,
                             uint8_t* fluid, int threads);

/* classifyBoundaryPoints and the CreateNodes precedence (ComplexGeometry.cpp:1369-1409,
 * Subdomain.cpp:124-155) for one boundary point at c: the index of the face it is assigned to
 * (first matching role in the order inlet, outlet 1, 2, 3; first face of that role in array
 * order), or -1 for a wall point. */
int orc_classify_point(const pvlbm_face* faces, int nfaces, const double c[3]);

typedef struct orc_sim orc_sim;

/* SubDomain construction (Subdomain.cpp:3-32, 91-218, 224-370, 549-596) from voxel flags, with
 * face specs in place of the hard-coded planes (ComplexGeometry.cpp:1369-1429).  normals = 0
 * skips calculateBoundaryNormals (for timing-only runs; WSS then reads zero normals). */
orc_sim* orc_sim_create(const uint8_t* fluid, const int32_t dims[3], const double origin[3],
                        double h, int q, const pvlbm_face* faces, int nfaces, const double* tri,
                        int64_t ntri, double tau, double dt, int normals);
void orc_sim_destroy(orc_sim* s);
/* out: nodes, inlet, outlet1, outlet2, outlet3, wall, next-to-wall */
void orc_sim_counts(const orc_sim* s, int64_t out[7]);
/* enum pvlbm_kind per box point */
void orc_sim_kinds(const orc_sim* s, uint8_t* kinds);
/* list members as box indices (list ids as enum pvlbm_list_id, LIST_FLUID..LIST_NTW) */
int64_t orc_sim_list(const orc_sim* s, int list, int64_t* idx);
/* inlet velocity per box point (read at inlet nodes); default 1.0 */
void orc_sim_set_inlet_profile(orc_sim* s, const double* prof_box);
void orc_sim_set_threads(orc_sim* s, int threads);

/* n reference iterations (Simulation.cpp:577-586), push streaming as the reference does.
 * inlet_scale: n values (NULL = 1.0). wss_every as pvlbm_set_wss_every. */
void orc_sim_step_push(orc_sim* s, int n, const double* inlet_scale, int wss_every);
/* The same iterations by the pull recipe the GPU uses (post-collision lattice + stale values for
 * boundary nodes + previous macros for next-to-wall nodes).  Must agree bit for bit with push. */
void orc_sim_step_pull(orc_sim* s, int n, const double* inlet_scale, int wss_every);

/* rho/u as Simulation::Collision computes them from the current populations (dense box). */
void orc_sim_macros(const orc_sim* s, double* rho, double* u);
/* populations of the current iteration, [n_nodes][q] in box-index order of the nodes */
void orc_sim_populations(const orc_sim* s, double* f);
/* overwrite populations (box-index order of nodes) — for property tests on random fields */
void orc_sim_set_populations(orc_sim* s, const double* f);
/* per next-to-wall node in list order: x,y,z,WSS,OSI,TAWSS; tau3 [3][nw]; acc4 [4][nw] */
void orc_sim_wss(const orc_sim* s, double* out6, double* tau3, double* acc4);
/* normals of the next-to-wall nodes [3][nw] */
void orc_sim_normals(const orc_sim* s, double* n3);
// replaces the next-to-wall normals ([3][nw]); with normals = 0 at create they start at zero
void orc_sim_set_normals(orc_sim* s, const double* n3);
// reference normal (Subdomain.cpp:301-369) at n points xyz[n][3] -> n3[3][n]
void orc_normal_at(const double* tri, int64_t ntri, const double* xyz, int64_t n, double* n3);
int64_t orc_sim_gate_events(const orc_sim* s);

/* Geometry::CalculateWallDistances (ComplexGeometry.cpp:550-608) with FindLineTriangleIntersection
 * (:689-730): per wall node (oracle list order) its box index and q per direction (-1: no link). */
void orc_sim_link_fractions(const orc_sim* s, const double* tri, int64_t ntri, int ndirs,
                            int64_t* box_out, double* q_out);

/* One node's Simulation::Collision (Simulation.cpp:52-205): out4 = (rho, ux, uy, uz) of f,
 * feq = its equilibrium, fnew = the relaxed populations. */
void orc_collide(int q, double tau, const double* f, double* feq, double* fnew, double* out4);
/* Zou-He closures on one node's populations (for closure tests). */
/* Wall shear stress vector of one node (Simulation::ComputeWallShearStress, Simulation.cpp:880-898):
 * tau[i] = C sum_alpha (f - feq)_alpha sum_j c_alpha,j n_j (c_alpha,i - sum_k c_alpha,k n_k n_i). */
void orc_wss_node(int q, const double* f, const double* feq, const double n[3], double C, double tau[3]);
void orc_zouhe(int q, int closure, double* f, double ux, double rho_out);
/* Simulation::getInletVelocity (Simulation.cpp:548-563). */
double orc_pulsatile_velocity(double t, double period);
/* Lattice tables (LatticeModel.cpp): c [q][3], w [q], opp [q]. */
void orc_lattice(int q, int32_t* c, double* w, int32_t* opp);

#ifdef __cplusplus
}
#endif

#endif
