/*
 * pvgen.h — synthetic vessel meshes for the benchmark and test workloads.
 *
 * No reference counterpart: the reference reads one patient STL (PV.stl, LBM.cpp:70).  These
 * generators produce closed, outward-oriented triangle meshes in the 12-double layout of pvlbm.h
 * (v1, v2, v3, unit facet normal) so that they can be written as ASCII STL and fed to the
 * voxeliser like any patient mesh.
 */
#ifndef PVGEN_H
#define PVGEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pvgen_mesh pvgen_mesh;

/* Straight tube along +x from x = 0 to x = length, axis at y = z = 0, nseg facets around with
 * ring vertices at angles (k + 1/2) * 360/nseg degrees, one axial segment, flat fan caps. */
int pvgen_tube(double radius, double length, int nseg, pvgen_mesh** out);

/* Symmetric Y-bifurcation: parent tube of radius rp along +x through (y0, z0) up to the branch
 * point x = xb, two daughters of radius rd at +-angle_deg in the x-y plane.  The solid is cut by
 * the planes x = xcap0 and x = xcap1 (flat caps).  Meshed by marching tetrahedra on a grid of the
 * given spacing whose origin is lo (choose lo off the lattice so no lattice ray meets a mesh edge);
 * the grid spans [lo, hi]. */
int pvgen_ybifurcation(double rp, double rd, double angle_deg, double xb, double y0, double z0,
                       double xcap0, double xcap1, const double lo[3], const double hi[3],
                       double spacing, pvgen_mesh** out);

/* Seeded random tortuous vessel in the box [0, n]^3: cubic-spline centre line with 4..6 control
 * points from x < 0 to x > n, radius varying in [rmin, rmax] along it, nbif (1 or 2) spline
 * branches leaving the main line and ending past x = n; cut by x = xcap0 / x = xcap1. */
int pvgen_vessel(uint64_t seed, double n, double rmin, double rmax, int nbif, double xcap0,
                 double xcap1, const double lo[3], const double hi[3], double spacing,
                 pvgen_mesh** out);

/* Seeded curved tube of radius rt along x through a box of size (nx, ny, nz) with a spherical
 * saccular bulge of radius rs on its side; cut by x = xcap0 / x = xcap1. */
int pvgen_aneurysm(uint64_t seed, double nx, double ny, double nz, double rt, double rs,
                   double xcap0, double xcap1, const double lo[3], const double hi[3],
                   double spacing, pvgen_mesh** out);

/* Vessel tree of straight segments: seg[11 * s + ...] = a (3), b (3), ra, rb, end_axis, end_val,
 * end_keep.  Segment s is the capsule around a -> b whose radius runs linearly from ra to rb; with
 * end_axis in 0..2 it is cut flat by the plane p[end_axis] = end_val, keeping p[end_axis] >= end_val
 * (end_keep > 0) or <= end_val (end_keep < 0); end_axis = -1 leaves the rounded end.  The union
 * is cut by the box [box_lo, box_hi] and meshed as pvgen_ybifurcation's shape is. */
int pvgen_tree(int nseg, const double* seg, const double box_lo[3], const double box_hi[3],
               const double lo[3], const double hi[3], double spacing, pvgen_mesh** out);

int64_t pvgen_mesh_size(const pvgen_mesh* m);
/* copy the triangles, 12 doubles each */
void pvgen_mesh_copy(const pvgen_mesh* m, double* tri);
void pvgen_mesh_free(pvgen_mesh* m);

/* Write triangles as ASCII STL with 17 significant digits (round-trips every double). */
int pvgen_write_stl_ascii(const char* path, const double* tri, int64_t ntri, const char* name);

/* Write triangles as binary STL (single precision, the format's own; name in the header). */
int pvgen_write_stl_binary(const char* path, const double* tri, int64_t ntri, const char* name);

/* Signed distance-like field of the last shape family on points (negative inside), for tests:
 * evaluates the same field the mesher contoured. shape: 1 = Y-bifurcation with the parameters
 * of pvgen_ybifurcation packed as {rp, rd, angle_deg, xb, y0, z0, xcap0, xcap1}. */
int pvgen_field(int shape, const double* params, int64_t npts, const double* xyz, double* phi);

#ifdef __cplusplus
}
#endif

#endif
