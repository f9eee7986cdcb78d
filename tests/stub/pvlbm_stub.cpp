// A stand-in for libpvlbm.so for the CPU test of lbm_run's loop (tests/test_cli.py): it records the
// order of the calls lbm_run makes -- steps, macro reads, WSS reads -- on stderr and returns
// small fixed data.  It computes nothing and is never shipped or loaded by the package.
#include <cstdio>
#include <cstring>

#include "../../include/pvlbm.h"

struct pvlbm_sim {
    long it = 0;
};
static pvlbm_sim g_sim;
static const int NX = 6, NY = 5, NZ = 4;

extern "C" {
const char* pvlbm_last_error(pvlbm_handle) { return ""; }
int pvlbm_read_stl_ascii(const char*, double* tri, int64_t, int64_t* n) {
    *n = 1;
    if (tri)
        for (int i = 0; i < 12; i++) tri[i] = i;
    return 0;
}
int pvlbm_read_stl_binary(const char*, double*, int64_t, int64_t* n) {
    *n = 0;
    return PVLBM_EIO;
}
int pvlbm_bbox(const double*, int64_t, double bb[6]) {
    for (int i = 0; i < 6; i++) bb[i] = i;
    return 0;
}
int pvlbm_voxelize(const double*, int64_t, double, const pvlbm_grid*, pvlbm_grid* g, uint8_t* f) {
    g->dims[0] = NX;
    g->dims[1] = NY;
    g->dims[2] = NZ;
    g->origin[0] = g->origin[1] = g->origin[2] = 0.0;
    if (f) std::memset(f, 1, NX * NY * NZ);
    return 0;
}
int pvlbm_create(const uint8_t*, const pvlbm_grid*, double, int, int, const pvlbm_face* faces, int nfaces,
                 const double*, int64_t, double, double, pvlbm_handle* out) {
    for (int i = 0; i < nfaces; i++)
        std::fprintf(stderr, "face %d %d %d %.21Lg\n", faces[i].role, faces[i].closure, faces[i].axis,
                     (long double)faces[i].plane + (long double)faces[i].plane_lo);
    *out = &g_sim;
    return 0;
}
int pvlbm_set_tau(pvlbm_handle, double) { return 0; }
int pvlbm_set_dt(pvlbm_handle, double) { return 0; }
int pvlbm_set_wall_closure(pvlbm_handle, int) { return 0; }
int pvlbm_set_streaming(pvlbm_handle, int) { return 0; }
int pvlbm_count(pvlbm_handle, int, int64_t* n) {
    *n = 1;
    return 0;
}
int pvlbm_list(pvlbm_handle, int, int64_t* v) {
    v[0] = 0;
    return 0;
}
int pvlbm_get_kinds(pvlbm_handle, uint8_t* k) {
    std::memset(k, PVLBM_FLUID, NX * NY * NZ);
    return 0;
}
int pvlbm_step(pvlbm_handle h, int n, const double* scale) {
    for (int s = 0; s < n; s++) std::fprintf(stderr, "iter %ld scale %.17g\n", h->it + s, scale ? scale[s] : 1.0);
    h->it += n;
    return 0;
}
int pvlbm_iteration(pvlbm_handle h, int64_t* it) {
    *it = h->it;
    return 0;
}
// the "state" reported is the number of completed iterations, so the test can tell which one a
// file holds
int pvlbm_get_macros(pvlbm_handle h, double* rho, double* u) {
    std::fprintf(stderr, "macros at %ld\n", h->it);
    const int n = NX * NY * NZ;
    for (int i = 0; i < n; i++) {
        rho[i] = (double)h->it;
        u[i] = u[n + i] = u[2 * n + i] = 0.0;
    }
    return 0;
}
int pvlbm_get_wss(pvlbm_handle h, double* out6) {
    std::fprintf(stderr, "wss at %ld\n", h->it);
    for (int c = 0; c < 6; c++) out6[c] = (double)h->it;
    return 0;
}
double pvlbm_pulsatile_velocity(double t, double) { return t; }
int pvlbm_destroy(pvlbm_handle) { return 0; }
}
