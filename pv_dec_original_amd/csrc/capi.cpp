// capi.cpp — host side of the C ABI declared in include/pvlbm.h.
//
// Set-up is split by cost: the O(N) and O(N*Q) work (voxelisation, boundary detection,
// next-to-wall marking, normals against every triangle) runs on the GPU; the O(boundary)
// classification runs on the host because the reference compares plane coordinates in x87
// long double (ComplexGeometry.cpp:1370-1408), which the GPU has no type for.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/pvlbm.h"
#include "kernels.hpp"
#include "lbm_device.hpp"

using namespace pvlbm;

namespace {

thread_local std::string g_err;

struct HipError {
    hipError_t e;
    const char* what;
};

#define HIPCHK(x)                                          \
    do {                                                   \
        hipError_t e_ = (x);                               \
        if (e_ != hipSuccess) throw HipError{e_, #x};      \
    } while (0)

struct Fail {
    int code;
    std::string msg;
};

[[noreturn]] void fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw Fail{code, buf};
}

template <class T>
T* dalloc(size_t count) {
    void* p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) fail(PVLBM_ENOMEM, "hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
    return static_cast<T*>(p);
}

void dfree(void* p) {
    if (p) (void)hipFree(p);
}

// Reference bbox rule (ComplexGeometry.cpp:142-154): x over v1..v3, y and z over v1 and v2.
void bbox_ref(const double* tri, int64_t ntri, double bb[6]) {
    const double inf = std::numeric_limits<double>::infinity();
    double lo[3] = {inf, inf, inf}, hi[3] = {-inf, -inf, -inf};
    for (int64_t t = 0; t < ntri; t++) {
        const double* T = tri + 12 * t;
        lo[0] = std::min({lo[0], T[0], T[3], T[6]});
        hi[0] = std::max({hi[0], T[0], T[3], T[6]});
        lo[1] = std::min({lo[1], T[1], T[4]});
        hi[1] = std::max({hi[1], T[1], T[4]});
        lo[2] = std::min({lo[2], T[2], T[5]});
        hi[2] = std::max({hi[2], T[2], T[5]});
    }
    bb[0] = lo[0]; bb[1] = hi[0]; bb[2] = lo[1]; bb[3] = hi[1]; bb[4] = lo[2]; bb[5] = hi[2];
}

// Ray (x, y) of a column after the bbox nudges of adjustLineCoordinates (ComplexGeometry.cpp:856-874).
double nudge(double v, double lo, double hi) {
    const double rayOffset = 0.02;
    if (v == lo) v += rayOffset;
    if (v == hi) v -= rayOffset;
    return v;
}

}  // namespace

// =============================================================================================
// The simulation handle
// =============================================================================================

struct pvlbm_sim {
    std::string err;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int device = 0;            // the device the handle's memory lives on
    void* scratch = nullptr;   // that device's scratch array (noalloc.hip), never freed
    int q = 19, storage = PVLBM_FP64;
    // padded box: nxp >= nx + 1 (multiple of 64) and nyp = ny + 1, so that every out-of-box
    // neighbour position of a node is a non-node slot (the home of a stale population)
    int64_t nx = 0, ny = 0, nz = 0, nxp = 0, nyp = 0, nsp = 0, guard = 0, ns = 0;
    double origin[3] = {0, 0, 0};
    double h = 1, tau = 1, dt = 1, C = 0;
    // device arrays
    uint8_t* d_kind = nullptr;
    int32_t* d_aux = nullptr;
    TileDesc* d_tdesc = nullptr;
    int32_t ntiles = 0;
    char* d_pop[2] = {nullptr, nullptr};  // allocation bases
    void* A = nullptr;                    // allocation base of the current lattice
    void* B = nullptr;
    int64_t nb = 0, nw = 0;
    uint32_t* d_bmask = nullptr;
    uint8_t* d_bclose = nullptr;
    double* d_bprof = nullptr;
    double* d_brho = nullptr;
    void* d_M = nullptr;
    double* d_nrm = nullptr;
    double* d_wss = nullptr;
    double* d_tau3 = nullptr;
    double* d_acc = nullptr;
    double* d_macro = nullptr;
    double* d_partials = nullptr;
    double* d_red = nullptr;
    unsigned long long* d_gate = nullptr;
    // host copy of the triangles (link fractions), boundary slots (padded node, closure)
    std::vector<double> tri;
    std::vector<int64_t> bnode;
    std::vector<uint8_t> bclose;
    int wall_mode = PVLBM_WALL_REFERENCE;
    int store_policy = PVLBM_STORES_AUTO;
    int streaming = PVLBM_STREAMING_AUTO;
    // kind of the launch that left the state (enum Scheme): SCH_PULL, state in A; SCH_EVEN /
    // SCH_ODD, in place in A
    int layout = SCH_PULL;
    size_t pop_bytes = 0;
    int64_t nfluid = 0;
    int64_t footprint = 0;  // bytes of the 128-byte lines of one lattice that hold nodes
    float* d_bq = nullptr;
    // host-side lists (unpadded box indices) and slots
    std::vector<int64_t> lists[8];
    std::vector<int32_t> inlet_slot;
    int64_t iter = 0;
    int wss_every = 1;
    bool red_valid = false;

    size_t store_bytes() const { return storage == PVLBM_FP64 ? 8 : 2; }
    bool stream_stores() const {
        if (store_policy != PVLBM_STORES_AUTO) return store_policy == PVLBM_STORES_STREAMING;
        if (in_place()) return false;  // in place the loads carry the cache policy
        return (int64_t)(2 * q * store_bytes()) * nfluid > PVLBM_STORES_AUTO_BYTES;
    }
    // nontemporal population loads under the automatic policy, in place (plain stores), by the
    // lines of the lattice that hold nodes: 2 = in every launch for a lattice that nearly fills
    // the last-level cache, 1 = in the even launches only for a larger one, 0 = none
    int stream_loads() const {
        if (store_policy != PVLBM_STORES_AUTO || !in_place()) return 0;
        if (footprint > PVLBM_LOADS_AUTO_EVEN_BYTES) return 1;
        return footprint > PVLBM_LOADS_AUTO_INPLACE_BYTES ? 2 : 0;
    }
    size_t real_bytes() const { return storage == PVLBM_FP64 ? 8 : 4; }
    // in place on one lattice when asked, or by default while that lattice is within the cache
    // budget; never under the interpolated closure, which reads the node's previous populations
    // from its own slots
    bool in_place() const {
        if (wall_mode != PVLBM_WALL_REFERENCE || streaming == PVLBM_STREAMING_TWO_LATTICES) return false;
        if (streaming == PVLBM_STREAMING_IN_PLACE) return true;
        return (int64_t)(q * store_bytes()) * nfluid <= PVLBM_STREAMING_AUTO_INPLACE_MAX_BYTES;
    }
    int blocks() const { return (ntiles + 3) / 4; }
    // padded row index of box row r = j + ny*k
    int64_t prow(int64_t r) const { return r % ny + nyp * (r / ny); }
    int64_t box_of_pad(int64_t n) const {
        int64_t i = n % nxp, pr = n / nxp;
        return i + nx * (pr % nyp + ny * (pr / nyp));
    }
    int64_t pad_of_box(int64_t b) const { return b % nx + nxp * prow(b / nx); }

    ~pvlbm_sim() {
        if (stream) (void)hipStreamSynchronize(stream);
        void* ptrs[] = {d_bq, d_kind, d_aux, d_tdesc, d_pop[0], d_pop[1], d_bmask, d_bclose, d_bprof, d_brho,
                        d_M, d_nrm, d_wss, d_tau3, d_acc, d_macro, d_partials, d_red, d_gate};
        for (void* p : ptrs) dfree(p);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }

    StepParams params(int flags, double scale) const {
        StepParams p{};
        p.A = A;
        p.B = B;
        p.ns = ns;
        p.guard = guard;
        p.sy = nxp;
        p.sz = nxp * nyp;
        {
            const int64_t S = (int64_t)store_bytes();
            const int* cx = (q == 19) ? Lat<19>::cx : Lat<27>::cx;
            const int* cy = (q == 19) ? Lat<19>::cy : Lat<27>::cy;
            const int* cz = (q == 19) ? Lat<19>::cz : Lat<27>::cz;
            p.poff_plane = ns * S;
            for (int d = 0; d < q; d++) {
                const int64_t o = (cx[d] + cy[d] * p.sy + cz[d] * p.sz) * S;
                p.poff_pull[d] = d * p.poff_plane - o;
                p.poff_push[d] = d * p.poff_plane + o;
            }
        }
        p.tdesc = d_tdesc;
        p.ntiles = ntiles;
        p.blocks = blocks();
        p.flags = flags;
        p.stream_stores = stream_stores() ? 1 : 0;
        p.bmask = d_bmask;
        p.bclose = d_bclose;
        p.bprof = d_bprof;
        p.brho = d_brho;
        p.bq = wall_mode == PVLBM_WALL_INTERPOLATED ? d_bq : nullptr;
        p.M = d_M;
        p.mst = std::max<int64_t>(nw, 1);
        p.nrm = d_nrm;
        p.wss = d_wss;
        p.tau3 = d_tau3;
        p.acc = d_acc;
        p.omega = 1.0 / tau;
        p.scale = scale;
        p.C = C;
        p.mrho = d_macro;
        p.partials = d_partials;
        p.gate = d_gate;
        return p;
    }

    // The second lattice, allocated when a scheme first needs it (the pull, or a change of layout).
    void ensure_second() {
        if (d_pop[1]) return;
        d_pop[1] = dalloc<char>(pop_bytes);
        HIPCHK(hipMemsetAsync(d_pop[1], 0, pop_bytes, stream));
        B = (A == d_pop[0]) ? d_pop[1] : d_pop[0];
    }

    // Brings the state into the layout of the scheme in use (values only move: fields unchanged).
    void sync_layout() {
        const bool want = in_place();
        if (want == (layout != SCH_PULL)) return;
        ensure_second();
        StepParams p = params(0, 1.0);
        const int to = want ? SCH_EVEN : SCH_PULL;
        HIPCHK(launch_relayout(q, storage, p, layout, to, stream));
        HIPCHK(hipStreamSynchronize(stream));
        std::swap(A, B);
        layout = to;
    }

    // Sets a mode that the scheme depends on and moves the state into its layout; if that fails
    // (no memory for the second lattice), the previous mode stays and the state is untouched.
    void set_mode_and_sync(int& field, int value) {
        const int old = field;
        field = value;
        try {
            sync_layout();
        } catch (...) {
            field = old;
            throw;
        }
    }

    void launch(int flags, double scale, const double* init_rho = nullptr, const double* init_u = nullptr) {
        StepParams p = params(flags, scale);
        p.init_rho = init_rho;
        p.init_u = init_u;
        if (in_place()) {
            // a start is an even launch without the gathers; then odd and even alternate
            p.scheme = ((flags & F_INIT) || layout == SCH_ODD) ? SCH_EVEN : SCH_ODD;
            if (!(flags & F_INIT) && layout == SCH_PULL) fail(PVLBM_ESTATE, "internal: state not in place");
            p.B = A;
            // with plain stores the lattice stays in the last-level cache between launches; odd
            // launches sweep it backwards, so each launch starts on the lines the previous one
            // touched last
            p.reverse = (p.scheme == SCH_ODD && !p.stream_stores) ? 1 : 0;
            const int loads = stream_loads();
            p.stream_loads = (loads == 2 || (loads == 1 && p.scheme == SCH_EVEN)) ? 1 : 0;
            HIPCHK(launch_step(q, storage, p, stream));
            layout = p.scheme;
        } else {
            ensure_second();
            if (!(flags & F_INIT) && layout != SCH_PULL) fail(PVLBM_ESTATE, "internal: state in place");
            p.B = B;
            p.scheme = SCH_PULL;
            HIPCHK(launch_step(q, storage, p, stream));
            std::swap(A, B);
            layout = SCH_PULL;
        }
        if (flags & F_REDUCE) {
            HIPCHK(launch_reduce_final(d_partials, blocks(), d_red, stream));
            red_valid = true;
        }
    }

    void reset_state() {
        HIPCHK(hipMemsetAsync(d_acc, 0, sizeof(double) * 4 * std::max<int64_t>(nw, 1), stream));
        HIPCHK(hipMemsetAsync(d_wss, 0, sizeof(double) * std::max<int64_t>(nw, 1), stream));
        HIPCHK(hipMemsetAsync(d_tau3, 0, sizeof(double) * 3 * std::max<int64_t>(nw, 1), stream));
        HIPCHK(hipMemsetAsync(d_gate, 0, sizeof(unsigned long long), stream));
        iter = 0;
    }
};

namespace {

template <class F>
int guarded(pvlbm_sim* h, F&& fn) {
    try {
        fn();
        if (h) h->err.clear();
        return PVLBM_OK;
    } catch (const Fail& f) {
        (h ? h->err : g_err) = f.msg;
        return f.code;
    } catch (const HipError& e) {
        (h ? h->err : g_err) = std::string(e.what) + ": " + hipGetErrorString(e.e);
        return PVLBM_EHIP;
    } catch (const std::bad_alloc&) {
        (h ? h->err : g_err) = "host allocation failed";
        return PVLBM_ENOMEM;
    } catch (...) {
        (h ? h->err : g_err) = "unexpected failure";
        return PVLBM_EINVAL;
    }
}

bool face_match(const pvlbm_face& F, const long double c[3]) {
    // ComplexGeometry.cpp:1380-1407: |c - plane| <= tol and open windows, in long double
    long double tolerance = F.tol;
    long double plane = (long double)F.plane + (long double)F.plane_lo;
    int a = (F.axis == 0) ? 1 : 0;
    int b = (F.axis == 2) ? 1 : 2;
    return (std::abs(c[F.axis] - plane) <= tolerance) && (F.lo[0] < c[a] && c[a] < F.hi[0]) &&
           (F.lo[1] < c[b] && c[b] < F.hi[1]);
}

}  // namespace

// =============================================================================================
// C ABI
// =============================================================================================

extern "C" const char* pvlbm_last_error(pvlbm_handle h) { return h ? h->err.c_str() : g_err.c_str(); }

extern "C" int pvlbm_device_count(int* count) {
    return guarded(nullptr, [&] {
        if (!count) fail(PVLBM_EINVAL, "count is NULL");
        int c = 0;
        hipError_t e = hipGetDeviceCount(&c);
        *count = (e == hipSuccess) ? c : 0;
    });
}

extern "C" int pvlbm_read_stl_ascii(const char* path, double* tri, int64_t capacity, int64_t* ntri) {
    return guarded(nullptr, [&] {
        if (!path || !ntri) fail(PVLBM_EINVAL, "path or ntri is NULL");
        std::ifstream file(path);
        if (!file.is_open()) fail(PVLBM_EIO, "cannot open file %s", path);
        // GeometryReader.cpp:46-112: a facet is "facet normal n n n", "outer loop", three
        // "vertex x y z", "endloop", "endfacet"; a facet whose lines do not parse is dropped.
        std::string line, w0, w1;
        int64_t count = 0;
        double T[12];
        auto next = [&](std::istringstream& ss) -> bool {
            if (!std::getline(file, line)) line.clear();
            ss.clear();
            ss.str(line);
            return true;
        };
        while (std::getline(file, line)) {
            std::istringstream ss(line);
            ss >> w0 >> w1 >> T[9] >> T[10] >> T[11];
            if (ss.fail() || w0 != "facet" || w1 != "normal") continue;
            next(ss);
            ss >> w0 >> w1;
            if (ss.fail() || w0 != "outer" || w1 != "loop") continue;
            for (int v = 0; v < 3; v++) {
                next(ss);
                double x, y, z;
                ss >> w0 >> x >> y >> z;
                if (ss.fail() || w0 != "vertex") break;
                T[3 * v] = x;
                T[3 * v + 1] = y;
                T[3 * v + 2] = z;
            }
            next(ss);
            ss >> w0;
            if (ss.fail() || w0 != "endloop") continue;
            next(ss);
            ss >> w0;
            if (ss.fail() || w0 != "endfacet") continue;
            if (tri && count < capacity) std::memcpy(tri + 12 * count, T, sizeof T);
            count++;
        }
        *ntri = count;
    });
}

extern "C" int pvlbm_read_stl_binary(const char* path, double* tri, int64_t capacity, int64_t* ntri) {
    return guarded(nullptr, [&] {
        if (!path || !ntri) fail(PVLBM_EINVAL, "path or ntri is NULL");
        std::ifstream file(path, std::ios::binary);
        if (!file.is_open()) fail(PVLBM_EIO, "cannot open file %s", path);
        // GeometryReader.cpp:119-176: 80-byte header, uint32 facet count, per facet float normal
        // (normalised in float: n / sqrtf(nx*nx + ny*ny + nz*nz)), three float vertices and a
        // uint16 attribute count that is ignored
        char header[80];
        uint32_t n = 0;
        if (!file.read(header, 80) || !file.read(reinterpret_cast<char*>(&n), 4))
            fail(PVLBM_EIO, "%s: shorter than a binary STL header", path);
        for (uint32_t t = 0; t < n; t++) {
            float f[12];
            uint16_t attr;
            if (!file.read(reinterpret_cast<char*>(f), 48) || !file.read(reinterpret_cast<char*>(&attr), 2))
                fail(PVLBM_EIO, "%s: truncated after %u of %u facets", path, t, n);
            float nx = f[0], ny = f[1], nz = f[2];
            const float norm = std::sqrt(nx * nx + ny * ny + nz * nz);
            nx /= norm;
            ny /= norm;
            nz /= norm;
            if (tri && (int64_t)t < capacity) {
                double* T = tri + 12 * (int64_t)t;
                for (int v = 0; v < 9; v++) T[v] = (double)f[3 + v];
                T[9] = nx;
                T[10] = ny;
                T[11] = nz;
            }
        }
        *ntri = (int64_t)n;
    });
}

extern "C" int pvlbm_bbox(const double* tri, int64_t ntri, double bbox[6]) {
    return guarded(nullptr, [&] {
        if (!tri || ntri <= 0 || !bbox) fail(PVLBM_EINVAL, "need at least one triangle");
        bbox_ref(tri, ntri, bbox);
    });
}

extern "C" int pvlbm_grid_from_bbox(const double bbox[6], double h, pvlbm_grid* grid) {
    return guarded(nullptr, [&] {
        if (!bbox || !grid || !(h > 0)) fail(PVLBM_EINVAL, "bad bbox/grid/h");
        // ComplexGeometry.cpp:163-165
        grid->origin[0] = bbox[0];
        grid->origin[1] = bbox[2];
        grid->origin[2] = bbox[4];
        grid->dims[0] = (int32_t)(std::ceil((bbox[1] - bbox[0]) / h) + 1);
        grid->dims[1] = (int32_t)(std::ceil((bbox[3] - bbox[2]) / h) + 1);
        grid->dims[2] = (int32_t)std::ceil((bbox[5] - bbox[4]) / h);
        grid->reserved = 0;
    });
}

namespace {
// work and kernel time of this thread's last labelling call (pvlbm_voxelize_stats)
thread_local int64_t g_vox_tests = 0;
thread_local double g_vox_ms = 0.0;
}  // namespace

extern "C" int pvlbm_voxelize_stats(int64_t* tests, double* kernel_ms) {
    if (tests) *tests = g_vox_tests;
    if (kernel_ms) *kernel_ms = g_vox_ms;
    return PVLBM_OK;
}

extern "C" int pvlbm_voxelize(const double* tri, int64_t ntri, double h, const pvlbm_grid* grid,
                              pvlbm_grid* grid_out, uint8_t* fluid_out) {
    return guarded(nullptr, [&] {
        if (!tri || ntri <= 0 || !(h > 0) || !grid_out) fail(PVLBM_EINVAL, "bad voxelize arguments");
        double bb[6];
        bbox_ref(tri, ntri, bb);
        pvlbm_grid G;
        if (!grid) {
            if (pvlbm_grid_from_bbox(bb, h, &G) != PVLBM_OK) fail(PVLBM_EINVAL, "%s", g_err.c_str());
        } else {
            G = *grid;
            for (int a = 0; a < 3; a++) {
                double lo = G.origin[a], hi = G.origin[a] + (double)(G.dims[a] - 1) * h;
                bb[2 * a] = std::min(bb[2 * a], lo);
                bb[2 * a + 1] = std::max(bb[2 * a + 1], hi);
            }
        }
        if (G.dims[0] <= 0 || G.dims[1] <= 0 || G.dims[2] <= 0) fail(PVLBM_EINVAL, "empty grid");
        *grid_out = G;
        if (!fluid_out) return;
        const int64_t nx = G.dims[0], ny = G.dims[1], nz = G.dims[2];
        // candidate triangles per column: those whose xy extent, widened by a margin far above
        // rounding, contains the column's (nudged) ray position.  The per-point test is the
        // reference's; the margin only keeps every triangle that test could accept.
        std::vector<double> xr(nx), yr(ny);
        for (int64_t i = 0; i < nx; i++) xr[i] = nudge(G.origin[0] + (double)i * h, bb[0], bb[1]);
        for (int64_t j = 0; j < ny; j++) yr[j] = nudge(G.origin[1] + (double)j * h, bb[2], bb[3]);
        double scale = 1.0;
        for (int a = 0; a < 6; a++) scale = std::max(scale, std::fabs(bb[a]));
        const double margin = 1e-6 * scale;
        const int64_t slack = 1 + (int64_t)std::ceil(0.02 / h);
        std::vector<int64_t> count((size_t)(nx * ny + 1), 0);
        std::vector<std::array<int64_t, 4>> ranges((size_t)ntri);
        auto col_range = [&](int64_t t, int64_t r[4], double box[4]) {
            const double* T = tri + 12 * t;
            box[0] = std::min({T[0], T[3], T[6]}) - margin;
            box[1] = std::max({T[0], T[3], T[6]}) + margin;
            box[2] = std::min({T[1], T[4], T[7]}) - margin;
            box[3] = std::max({T[1], T[4], T[7]}) + margin;
            r[0] = std::max<int64_t>(0, (int64_t)std::floor((box[0] - G.origin[0]) / h) - slack);
            r[1] = std::min<int64_t>(nx - 1, (int64_t)std::ceil((box[1] - G.origin[0]) / h) + slack);
            r[2] = std::max<int64_t>(0, (int64_t)std::floor((box[2] - G.origin[1]) / h) - slack);
            r[3] = std::min<int64_t>(ny - 1, (int64_t)std::ceil((box[3] - G.origin[1]) / h) + slack);
        };
        for (int pass = 0; pass < 2; pass++) {
            std::vector<int32_t> tri_ids;
            std::vector<int64_t> fill;
            if (pass == 1) {
                for (int64_t c = 0; c < nx * ny; c++) count[c + 1] += count[c];
                tri_ids.resize((size_t)count[nx * ny]);
                fill.assign(count.begin(), count.end() - 1);
            }
            for (int64_t t = 0; t < ntri; t++) {
                int64_t r[4];
                double box[4];
                col_range(t, r, box);
                for (int64_t j = r[2]; j <= r[3]; j++) {
                    if (yr[j] < box[2] || yr[j] > box[3]) continue;
                    for (int64_t i = r[0]; i <= r[1]; i++) {
                        if (xr[i] < box[0] || xr[i] > box[1]) continue;
                        if (pass == 0) count[i + nx * j + 1]++;
                        else tri_ids[(size_t)fill[i + nx * j]++] = (int32_t)t;
                    }
                }
            }
            if (pass == 1) {
                hipStream_t s;
                HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
                double* d_tri = nullptr;
                int64_t* d_off = nullptr;
                int32_t* d_ids = nullptr;
                uint8_t* d_fluid = nullptr;
                try {
                    d_tri = dalloc<double>((size_t)(12 * ntri));
                    d_off = dalloc<int64_t>(count.size());
                    d_ids = dalloc<int32_t>(tri_ids.size());
                    d_fluid = dalloc<uint8_t>((size_t)(nx * ny * nz));
                    HIPCHK(hipMemcpyAsync(d_tri, tri, sizeof(double) * 12 * ntri, hipMemcpyHostToDevice, s));
                    HIPCHK(hipMemcpyAsync(d_off, count.data(), sizeof(int64_t) * count.size(), hipMemcpyHostToDevice, s));
                    if (!tri_ids.empty())
                        HIPCHK(hipMemcpyAsync(d_ids, tri_ids.data(), sizeof(int32_t) * tri_ids.size(), hipMemcpyHostToDevice, s));
                    VoxParams p{};
                    p.tri = d_tri;
                    p.col_off = d_off;
                    p.col_tri = d_ids;
                    p.ox = G.origin[0];
                    p.oy = G.origin[1];
                    p.oz = G.origin[2];
                    p.h = h;
                    for (int a = 0; a < 6; a++) p.bb[a] = bb[a];
                    p.nx = (int32_t)nx;
                    p.ny = (int32_t)ny;
                    p.nz = (int32_t)nz;
                    p.fluid = d_fluid;
                    hipEvent_t e0 = nullptr, e1 = nullptr;
                    HIPCHK(hipEventCreate(&e0));
                    HIPCHK(hipEventCreate(&e1));
                    HIPCHK(hipEventRecord(e0, s));
                    hipError_t le = launch_voxelize(p, s);
                    if (le == hipSuccess) le = hipEventRecord(e1, s);
                    if (le == hipSuccess)
                        le = hipMemcpyAsync(fluid_out, d_fluid, (size_t)(nx * ny * nz), hipMemcpyDeviceToHost, s);
                    if (le == hipSuccess) le = hipStreamSynchronize(s);
                    float ms = 0.f;
                    if (le == hipSuccess) le = hipEventElapsedTime(&ms, e0, e1);
                    (void)hipEventDestroy(e0);
                    (void)hipEventDestroy(e1);
                    HIPCHK(le);
                    g_vox_tests = count[(size_t)(nx * ny)] * nz;
                    g_vox_ms = (double)ms;
                } catch (...) {
                    (void)hipStreamSynchronize(s);
                    dfree(d_tri); dfree(d_off); dfree(d_ids); dfree(d_fluid);
                    (void)hipStreamDestroy(s);
                    throw;
                }
                dfree(d_tri); dfree(d_off); dfree(d_ids); dfree(d_fluid);
                (void)hipStreamDestroy(s);
            }
        }
    });
}

extern "C" int pvlbm_create(const uint8_t* fluid, const pvlbm_grid* grid, double h, int lattice,
                            int storage, const pvlbm_face* faces, int nfaces, const double* tri,
                            int64_t ntri, double tau, double dt, pvlbm_handle* out) {
    pvlbm_sim* S = nullptr;
    int rc = guarded(nullptr, [&] {
        if (!fluid || !grid || !out) fail(PVLBM_EINVAL, "fluid, grid and out are required");
        if (lattice != 19 && lattice != 27) fail(PVLBM_EINVAL, "lattice must be 19 or 27, got %d", lattice);
        if (storage < PVLBM_FP64 || storage > PVLBM_BF16S) fail(PVLBM_EINVAL, "bad storage %d", storage);
        if (!(h > 0) || !(tau > 0.5) || !(dt > 0)) fail(PVLBM_EINVAL, "need h > 0, tau > 0.5, dt > 0");
        if (nfaces < 0 || (nfaces > 0 && !faces)) fail(PVLBM_EINVAL, "bad faces");
        if (ntri < 0 || (ntri > 0 && !tri)) fail(PVLBM_EINVAL, "bad triangles");
        for (int f = 0; f < nfaces; f++) {
            const pvlbm_face& F = faces[f];
            if (F.role < PVLBM_INLET || F.role > PVLBM_OUTLET3) fail(PVLBM_EINVAL, "face %d: bad role", f);
            if (F.axis < 0 || F.axis > 2) fail(PVLBM_EINVAL, "face %d: bad axis", f);
            bool inlet = F.role == PVLBM_INLET;
            if (inlet != (F.closure == PVLBM_INLET_XMIN)) fail(PVLBM_EINVAL, "face %d: inlets take INLET_XMIN, outlets a pressure closure", f);
            if (F.closure < PVLBM_INLET_XMIN || F.closure > PVLBM_OUTLET_XMAX) fail(PVLBM_EINVAL, "face %d: bad closure", f);
            if (lattice == 27 && (F.closure == PVLBM_OUTLET_YMAX || F.closure == PVLBM_OUTLET_ZMIN))
                fail(PVLBM_EINVAL, "face %d: the y-max/z-min outlet closures exist for D3Q19 only", f);
            if (!inlet && !(F.rho > 0)) fail(PVLBM_EINVAL, "face %d: outlet density must be > 0", f);
        }
        const int64_t nx = grid->dims[0], ny = grid->dims[1], nz = grid->dims[2];
        if (nx <= 0 || ny <= 0 || nz <= 0) fail(PVLBM_EINVAL, "empty grid");
        const int64_t nxp = (nx + 1 + 63) / 64 * 64;
        const int64_t nyp = ny + 1;
        const int64_t nsp = nxp * nyp * nz;
        if (nsp / 64 > (int64_t)INT32_MAX) fail(PVLBM_EINVAL, "grid too large");

        S = new pvlbm_sim();
        S->q = lattice;
        S->storage = storage;
        S->nx = nx; S->ny = ny; S->nz = nz; S->nxp = nxp; S->nyp = nyp; S->nsp = nsp;
        S->guard = nxp * nyp + nxp + 64;
        S->ns = nsp + 2 * S->guard;
        for (int a = 0; a < 3; a++) S->origin[a] = grid->origin[a];
        S->h = h; S->tau = tau; S->dt = dt;
        {
            // Simulation.cpp:880-881
            double phys_den = 1060;
            double dtdt = dt * dt;
            double dxdx = h * h * 1e-6;
            S->C = phys_den * dtdt / dxdx;
        }
        HIPCHK(hipGetDevice(&S->device));
        HIPCHK(noalloc_scratch(S->device, &S->scratch));
        HIPCHK(hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking));
        HIPCHK(hipEventCreate(&S->ev0));
        HIPCHK(hipEventCreate(&S->ev1));
        hipStream_t st = S->stream;
        const int q = lattice;
        const int (*cx) = (q == 19) ? Lat<19>::cx : Lat<27>::cx;
        const int (*cy) = (q == 19) ? Lat<19>::cy : Lat<27>::cy;
        const int (*cz) = (q == 19) ? Lat<19>::cz : Lat<27>::cz;

        // padded fluid flags -> boundary flags on the GPU
        std::vector<uint8_t> fpad((size_t)nsp, 0);
        for (int64_t r = 0; r < ny * nz; r++) std::memcpy(&fpad[S->prow(r) * nxp], fluid + r * nx, (size_t)nx);
        uint8_t* d_fluid = dalloc<uint8_t>((size_t)nsp);
        uint8_t* d_bflag = nullptr;
        std::vector<uint8_t> bflag((size_t)nsp);
        try {
            d_bflag = dalloc<uint8_t>((size_t)nsp);
            HIPCHK(hipMemcpyAsync(d_fluid, fpad.data(), (size_t)nsp, hipMemcpyHostToDevice, st));
            HIPCHK(launch_boundary(q, d_fluid, nxp, nyp, nx, ny, nz, d_bflag, st));
            HIPCHK(hipMemcpyAsync(bflag.data(), d_bflag, (size_t)nsp, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
        } catch (...) {
            dfree(d_fluid); dfree(d_bflag);
            throw;
        }
        dfree(d_fluid);
        dfree(d_bflag);

        // classification of boundary nodes (host, long double) and the boundary table
        std::vector<uint8_t> kind((size_t)nsp, K_SOLID);
        std::vector<int32_t> aux((size_t)nsp, -1);
        std::vector<uint32_t> bmask;
        std::vector<uint8_t> bclose;
        std::vector<double> bprof, brho;
        std::vector<int64_t> bnode;
        std::vector<int> brole;
        for (int64_t n = 0; n < nsp; n++) {
            if (!fpad[n]) continue;
            kind[n] = K_FLUID;
            if (!bflag[n]) continue;
            const int64_t i = n % nxp, j = (n / nxp) % nyp, k = n / (nxp * nyp);
            long double c3[3] = {(long double)(grid->origin[0] + (double)i * h),
                                 (long double)(grid->origin[1] + (double)j * h),
                                 (long double)(grid->origin[2] + (double)k * h)};
            int matched = -1;
            for (int role = PVLBM_INLET; role <= PVLBM_OUTLET3 && matched < 0; role++)
                for (int f = 0; f < nfaces; f++)
                    if (faces[f].role == role && face_match(faces[f], c3)) {
                        matched = f;
                        break;
                    }
            uint32_t m = 0;
            for (int d = 0; d < q; d++) {
                int64_t a = i + cx[d], b = j + cy[d], c = k + cz[d];
                bool inside = a >= 0 && a < nx && b >= 0 && b < ny && c >= 0 && c < nz;
                if (!inside || !fpad[a + nxp * (b + nyp * c)]) m |= 1u << d;
            }
            aux[n] = (int32_t)bnode.size();
            bnode.push_back(n);
            // bits 0..q-1: the missing links; bits 28..31: the closure (the step reads both at once)
            const uint8_t closure = matched < 0 ? (uint8_t)BC_WALL : (uint8_t)(faces[matched].closure + 1);
            bmask.push_back(m | ((uint32_t)closure << 28));
            if (matched < 0) {
                kind[n] = K_WALL;
                bclose.push_back(BC_WALL);
                brole.push_back(-1);
                bprof.push_back(0.0);
                brho.push_back(1.0);
            } else {
                const pvlbm_face& F = faces[matched];
                kind[n] = (F.role == PVLBM_INLET) ? K_INLET : K_OUTLET;
                bclose.push_back((uint8_t)(F.closure + 1));
                brole.push_back(F.role);
                bprof.push_back(1.0);
                brho.push_back(F.rho);
            }
        }
        S->nb = (int64_t)bnode.size();
        S->bnode = bnode;
        S->bclose = bclose;

        // next-to-wall marking on the GPU
        S->d_kind = dalloc<uint8_t>((size_t)nsp);
        HIPCHK(hipMemcpyAsync(S->d_kind, kind.data(), (size_t)nsp, hipMemcpyHostToDevice, st));
        HIPCHK(launch_ntw(q, S->d_kind, nxp, nyp, nz, st));
        HIPCHK(hipMemcpyAsync(kind.data(), S->d_kind, (size_t)nsp, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));

        // lists, next-to-wall slots, active tiles
        std::vector<int32_t> tiles;
        std::vector<double> xyz;
        std::vector<int64_t> wnode;
        for (int64_t n = 0; n < nsp; n++) {
            const uint8_t kd = kind[n];
            if (kd == K_SOLID) continue;
            const int64_t b = S->box_of_pad(n);
            S->lists[PVLBM_LIST_FLUID].push_back(b);
            if (tiles.empty() || tiles.back() != (int32_t)(n / 64)) tiles.push_back((int32_t)(n / 64));
            if (kd == K_NTW) {
                aux[n] = (int32_t)wnode.size();
                wnode.push_back(n);
                S->lists[PVLBM_LIST_NTW].push_back(b);
            } else if (kd >= K_INLET) {
                const int32_t slot = aux[n];
                if (kd == K_WALL) S->lists[PVLBM_LIST_WALL].push_back(b);
                else if (kd == K_INLET) {
                    S->lists[PVLBM_LIST_INLET].push_back(b);
                    S->inlet_slot.push_back(slot);
                } else {
                    S->lists[PVLBM_LIST_OUTLET1 + brole[slot] - 1].push_back(b);
                }
            }
        }
        S->nw = (int64_t)wnode.size();
        S->nfluid = (int64_t)S->lists[PVLBM_LIST_FLUID].size();
        S->ntiles = (int32_t)tiles.size();
        for (int32_t t : tiles) S->lists[PVLBM_LIST_TILES].push_back(t);
        // tile descriptors: lane masks and first side-table slots (slots follow node order)
        std::vector<TileDesc> tdesc(tiles.size());
        {
            int32_t bs = 0, wsl = 0;
            for (size_t ti = 0; ti < tiles.size(); ti++) {
                TileDesc& D = tdesc[ti];
                D.tile = tiles[ti];
                D.bslot = bs;
                D.wslot = wsl;
                D.pad = 0;
                D.node = D.ntw = D.bnd = 0;
                for (int l = 0; l < 64; l++) {
                    const uint8_t kd = kind[(size_t)tiles[ti] * 64 + l];
                    if (kd == K_SOLID) continue;
                    D.node |= 1ull << l;
                    if (kd == K_NTW) {
                        D.ntw |= 1ull << l;
                        wsl++;
                    } else if (kd >= K_INLET) {
                        D.bnd |= 1ull << l;
                        bs++;
                    }
                }
            }
            if (bs != S->nb || wsl != (int32_t)wnode.size()) fail(PVLBM_EINVAL, "internal: slot order mismatch");
            // cache footprint of one lattice: the 128-byte lines of a plane that hold a node, x Q
            const int per = (int)(128 / S->store_bytes());
            int64_t lines = 0;
            for (const TileDesc& D : tdesc)
                for (int l = 0; l < 64; l += per)
                    if ((D.node >> l) & (per == 64 ? ~0ull : ((1ull << per) - 1))) lines++;
            S->footprint = lines * 128 * q;
        }
        xyz.resize((size_t)(3 * S->nw));
        for (int64_t w = 0; w < S->nw; w++) {
            const int64_t n = wnode[w];
            const int64_t i = n % nxp, j = (n / nxp) % nyp, k = n / (nxp * nyp);
            xyz[w] = grid->origin[0] + (double)i * h;
            xyz[S->nw + w] = grid->origin[1] + (double)j * h;
            xyz[2 * S->nw + w] = grid->origin[2] + (double)k * h;
        }

        // device tables
        const int64_t nb1 = std::max<int64_t>(S->nb, 1), nw1 = std::max<int64_t>(S->nw, 1);
        S->d_aux = dalloc<int32_t>((size_t)nsp);
        S->d_tdesc = dalloc<TileDesc>(tdesc.size());
        S->d_bmask = dalloc<uint32_t>((size_t)nb1);
        S->d_bclose = dalloc<uint8_t>((size_t)nb1);
        S->d_bprof = dalloc<double>((size_t)nb1);
        S->d_brho = dalloc<double>((size_t)nb1);
        S->d_M = dalloc<char>((size_t)(4 * nw1) * S->real_bytes());
        S->d_nrm = dalloc<double>((size_t)(3 * nw1));
        S->d_wss = dalloc<double>((size_t)nw1);
        S->d_tau3 = dalloc<double>((size_t)(3 * nw1));
        S->d_acc = dalloc<double>((size_t)(4 * nw1));
        S->d_macro = dalloc<double>((size_t)(4 * S->ns));
        S->d_partials = dalloc<double>((size_t)(4 * std::max(S->blocks(), 1)));
        S->d_red = dalloc<double>(4);
        S->d_gate = dalloc<unsigned long long>(1);
        // one lattice for the in-place scheme; the second comes with the first launch of a scheme
        // that needs it
        S->pop_bytes = (size_t)q * (size_t)S->ns * S->store_bytes();
        S->d_pop[0] = dalloc<char>(S->pop_bytes);
        HIPCHK(hipMemsetAsync(S->d_pop[0], 0, S->pop_bytes, st));
        S->A = S->d_pop[0];
        S->B = nullptr;
        HIPCHK(hipMemcpyAsync(S->d_aux, aux.data(), sizeof(int32_t) * nsp, hipMemcpyHostToDevice, st));
        if (!tdesc.empty())
            HIPCHK(hipMemcpyAsync(S->d_tdesc, tdesc.data(), sizeof(TileDesc) * tdesc.size(), hipMemcpyHostToDevice, st));
        if (S->nb) {
            HIPCHK(hipMemcpyAsync(S->d_bmask, bmask.data(), sizeof(uint32_t) * S->nb, hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(S->d_bclose, bclose.data(), S->nb, hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(S->d_bprof, bprof.data(), sizeof(double) * S->nb, hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(S->d_brho, brho.data(), sizeof(double) * S->nb, hipMemcpyHostToDevice, st));
        }
        HIPCHK(hipMemsetAsync(S->d_nrm, 0, sizeof(double) * 3 * nw1, st));
        HIPCHK(hipMemsetAsync(S->d_M, 0, (size_t)(4 * nw1) * S->real_bytes(), st));
        HIPCHK(hipMemsetAsync(S->d_macro, 0, sizeof(double) * 4 * S->ns, st));
        // normals of the next-to-wall nodes against every triangle
        if (S->nw > 0 && ntri > 0) {
            double *d_tri = nullptr, *d_prep = nullptr, *d_xyz = nullptr;
            try {
                d_tri = dalloc<double>((size_t)(12 * ntri));
                d_prep = dalloc<double>((size_t)(7 * ntri));
                d_xyz = dalloc<double>((size_t)(3 * S->nw));
                HIPCHK(hipMemcpyAsync(d_tri, tri, sizeof(double) * 12 * ntri, hipMemcpyHostToDevice, st));
                HIPCHK(hipMemcpyAsync(d_xyz, xyz.data(), sizeof(double) * 3 * S->nw, hipMemcpyHostToDevice, st));
                S->tri.assign(tri, tri + 12 * ntri);
                HIPCHK(launch_tri_prep(d_tri, ntri, d_prep, st));
                HIPCHK(launch_normals(d_prep, ntri, d_xyz, S->nw, S->d_nrm, st));
                HIPCHK(hipStreamSynchronize(st));
            } catch (...) {
                (void)hipStreamSynchronize(st);
                dfree(d_tri); dfree(d_prep); dfree(d_xyz);
                throw;
            }
            dfree(d_tri); dfree(d_prep); dfree(d_xyz);
        }
        S->reset_state();
        // f = w everywhere, first collision (Subdomain.cpp:549-564, Simulation.cpp:43-208)
        S->launch(F_INIT | F_MACRO | F_REDUCE, 1.0);
        HIPCHK(hipStreamSynchronize(st));
        *out = S;
    });
    if (rc != PVLBM_OK) {
        delete S;
        if (out) *out = nullptr;
    }
    return rc;
}

extern "C" int pvlbm_set_tau(pvlbm_handle h, double tau) {
    if (!h) return PVLBM_EINVAL;
    return guarded(h, [&] {
        if (!(tau > 0.5)) fail(PVLBM_EINVAL, "tau must exceed 0.5");
        h->tau = tau;
    });
}

extern "C" int pvlbm_set_dt(pvlbm_handle h, double dt) {
    if (!h) return PVLBM_EINVAL;
    return guarded(h, [&] {
        if (!(dt > 0)) fail(PVLBM_EINVAL, "dt must be positive");
        h->dt = dt;
        // Simulation.cpp:880-881
        double phys_den = 1060;
        double dtdt = dt * dt;
        double dxdx = h->h * h->h * 1e-6;
        h->C = phys_den * dtdt / dxdx;
    });
}

extern "C" int pvlbm_count(pvlbm_handle h, int list, int64_t* n) {
    if (!h) return PVLBM_EINVAL;
    return guarded(h, [&] {
        if (list < 0 || list > PVLBM_LIST_TILES || !n) fail(PVLBM_EINVAL, "bad list id");
        *n = (int64_t)h->lists[list].size();
    });
}

extern "C" int pvlbm_list(pvlbm_handle h, int list, int64_t* idx) {
    if (!h) return PVLBM_EINVAL;
    return guarded(h, [&] {
        if (list < 0 || list > PVLBM_LIST_TILES || !idx) fail(PVLBM_EINVAL, "bad list id");
        std::copy(h->lists[list].begin(), h->lists[list].end(), idx);
    });
}

extern "C" int pvlbm_get_kinds(pvlbm_handle h, uint8_t* kinds) {
    if (!h) return PVLBM_EINVAL;
    return guarded(h, [&] {
        if (!kinds) fail(PVLBM_EINVAL, "kinds is NULL");
        std::vector<uint8_t> k((size_t)h->nsp);
        HIPCHK(hipMemcpyAsync(k.data(), h->d_kind, (size_t)h->nsp, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        for (int64_t r = 0; r < h->ny * h->nz; r++)
            std::memcpy(kinds + r * h->nx, &k[h->prow(r) * h->nxp], (size_t)h->nx);
    });
}

namespace {
// Checks a separable sample grid against the box and returns its point count; the product is
// bounded before it is formed.
size_t check_sample_grid(const pvlbm_sim* h, const int32_t* ix, int32_t nx, const int32_t* iy, int32_t ny,
                         const int32_t* iz, int32_t nz, const uint8_t* out) {
    if (!ix || !iy || !iz || !out || nx < 1 || ny < 1 || nz < 1) fail(PVLBM_EINVAL, "bad sample grid");
    if ((double)nx * (double)ny * (double)nz > 4.0e18) fail(PVLBM_EINVAL, "sample grid too large");
    const int32_t* idx[3] = {ix, iy, iz};
    const int32_t cnt[3] = {nx, ny, nz};
    const int64_t dims[3] = {h->nx, h->ny, h->nz};
    for (int a = 0; a < 3; a++)
        for (int32_t i = 0; i < cnt[a]; i++)
            if (idx[a][i] < 0 || idx[a][i] >= dims[a]) fail(PVLBM_EINVAL, "sample index outside the box");
    return (size_t)nx * (size_t)ny * (size_t)nz;
}
}  // namespace

extern "C" int pvlbm_sample_kinds(pvlbm_handle h, const int32_t* ix, int32_t nx, const int32_t* iy, int32_t ny,
                                  const int32_t* iz, int32_t nz, uint8_t* out) {
    if (!h) return PVLBM_EINVAL;
    return guarded(h, [&] {
        const size_t n = check_sample_grid(h, ix, nx, iy, ny, iz, nz, out);
        int32_t* d_idx = nullptr;
        uint8_t* d_out = nullptr;
        try {
            d_idx = dalloc<int32_t>((size_t)nx + ny + nz);
            d_out = dalloc<uint8_t>(n);
            HIPCHK(hipMemcpyAsync(d_idx, ix, sizeof(int32_t) * nx, hipMemcpyHostToDevice, h->stream));
            HIPCHK(hipMemcpyAsync(d_idx + nx, iy, sizeof(int32_t) * ny, hipMemcpyHostToDevice, h->stream));
            HIPCHK(hipMemcpyAsync(d_idx + nx + ny, iz, sizeof(int32_t) * nz, hipMemcpyHostToDevice, h->stream));
            HIPCHK(launch_sample_kinds(h->d_kind, h->nxp, h->nyp, d_idx, nx, d_idx + nx, ny, d_idx + nx + ny, nz,
                                       d_out, h->stream));
            HIPCHK(hipMemcpyAsync(out, d_out, n, hipMemcpyDeviceToHost, h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
        } catch (...) {
            (void)hipStreamSynchronize(h->stream);
            dfree(d_idx);
            dfree(d_out);
            throw;
        }
        dfree(d_idx);
        dfree(d_out);
    });
}

extern "C" int pvlbm_sample_kinds_noalloc(pvlbm_handle h, const int32_t* ix, int32_t nx, const int32_t* iy,
                                          int32_t ny, const int32_t* iz, int32_t nz, uint8_t* out) {
    if (!h) return PVLBM_EINVAL;
    return guarded(h, [&] {
        const size_t n = check_sample_grid(h, ix, nx, iy, ny, iz, nz, out);
        // [indices, padded to 256 bytes][sampled kinds] in the scratch array; every byte that is
        // read back is written by this call, so nothing carried from an earlier call is seen
        const size_t idx_bytes = ((size_t)nx + ny + nz) * sizeof(int32_t);
        const size_t out_off = (idx_bytes + 255) / 256 * 256;
        if (n > (size_t)PVLBM_NOALLOC_SCRATCH_BYTES || out_off + n > (size_t)PVLBM_NOALLOC_SCRATCH_BYTES)
            fail(PVLBM_EINVAL, "pvlbm_sample_kinds_noalloc: %zu samples and %zu index bytes exceed the %d-byte "
                 "scratch array", n, idx_bytes, (int)PVLBM_NOALLOC_SCRATCH_BYTES);
        if (!h->scratch) fail(PVLBM_ESTATE, "no scratch array on this handle's device");
        int32_t* d_idx = static_cast<int32_t*>(h->scratch);
        uint8_t* d_out = static_cast<uint8_t*>(h->scratch) + out_off;
        HIPCHK(hipMemcpyAsync(d_idx, ix, sizeof(int32_t) * nx, hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(d_idx + nx, iy, sizeof(int32_t) * ny, hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(d_idx + nx + ny, iz, sizeof(int32_t) * nz, hipMemcpyHostToDevice, h->stream));
        HIPCHK(launch_sample_kinds(h->d_kind, h->nxp, h->nyp, d_idx, nx, d_idx + nx, ny, d_idx + nx + ny, nz, d_out,
                                   h->stream));
        HIPCHK(hipMemcpyAsync(out, d_out, n, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    });
}

extern "C" int pvlbm_set_inlet_profile(pvlbm_handle h, const double* u) {
    if (!h) return PVLBM_EINVAL;
    return guarded(h, [&] {
        if (!u) fail(PVLBM_EINVAL, "u is NULL");
        std::vector<double> prof((size_t)std::max<int64_t>(h->nb, 1), 0.0);
        HIPCHK(hipMemcpyAsync(prof.data(), h->d_bprof, sizeof(double) * h->nb, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        for (size_t i = 0; i < h->inlet_slot.size(); i++) prof[h->inlet_slot[i]] = u[i];
        HIPCHK(hipMemcpyAsync(h->d_bprof, prof.data(), sizeof(double) * h->nb, hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    });
}

extern "C" int pvlbm_init_fields(pvlbm_handle h, const double* rho, const double* u) {
    if (!h) return PVLBM_EINVAL;
    return guarded(h, [&] {
        if (rho && !u) fail(PVLBM_EINVAL, "u is required with rho");
        double* d_in = nullptr;
        try {
            if (rho) {
                // padded [4][ns] staging: rho plane then three velocity planes
                std::vector<double> pad((size_t)(4 * h->ns), 0.0);
                for (int64_t r = 0; r < h->ny * h->nz; r++) {
                    const int64_t pr = h->prow(r) * h->nxp;
                    std::memcpy(&pad[pr], rho + r * h->nx, sizeof(double) * h->nx);
                    for (int c = 0; c < 3; c++)
                        std::memcpy(&pad[(c + 1) * h->ns + pr], u + c * h->nx * h->ny * h->nz + r * h->nx,
                                    sizeof(double) * h->nx);
                }
                d_in = dalloc<double>(pad.size());
                HIPCHK(hipMemcpyAsync(d_in, pad.data(), sizeof(double) * pad.size(), hipMemcpyHostToDevice, h->stream));
            }
            h->reset_state();
            h->launch(F_INIT | F_MACRO | F_REDUCE, 1.0, d_in, d_in ? d_in + h->ns : nullptr);
            HIPCHK(hipStreamSynchronize(h->stream));
        } catch (...) {
            (void)hipStreamSynchronize(h->stream);
            dfree(d_in);
            throw;
        }
        dfree(d_in);
    });
}

extern "C" int pvlbm_init_fields_coarse(pvlbm_handle h, const float* rho_c, const float* u_c,
                                        const int32_t cdims[3], double umax, double rho_lo, double rho_hi,
                                        double flux_x, const double* plane_flux, const double* plane_drho,
                                        int64_t* nonfinite) {
    if (!h) return PVLBM_EINVAL;
    return guarded(h, [&] {
        if (!rho_c || !u_c || !cdims) fail(PVLBM_EINVAL, "coarse fields and their dims are required");
        for (int a = 0; a < 3; a++)
            if (cdims[a] < 1) fail(PVLBM_EINVAL, "coarse dims must be positive");
        if (!(umax > 0.0) || !(rho_lo <= rho_hi)) fail(PVLBM_EINVAL, "bad clipping limits");
        // the coarse fields are the caller's device memory, possibly still being written on the
        // caller's stream: wait for all work on the device before this handle's stream reads them
        HIPCHK(hipDeviceSynchronize());
        double* d_in = nullptr;
        unsigned long long* d_cnt = nullptr;
        try {
            // [4][ns] fields, then nx each: plane flux scratch, flux targets, density offsets
            d_in = dalloc<double>((size_t)(4 * h->ns + 3 * h->nx));
            d_cnt = dalloc<unsigned long long>(1);
            double* d_tgt = d_in + 4 * h->ns + h->nx;
            double* d_drho = d_tgt + h->nx;
            if (plane_flux)
                HIPCHK(hipMemcpyAsync(d_tgt, plane_flux, sizeof(double) * h->nx, hipMemcpyHostToDevice, h->stream));
            if (plane_drho)
                HIPCHK(hipMemcpyAsync(d_drho, plane_drho, sizeof(double) * h->nx, hipMemcpyHostToDevice, h->stream));
            HIPCHK(hipMemsetAsync(d_cnt, 0, sizeof(unsigned long long), h->stream));
            HIPCHK(launch_expand_coarse(rho_c, u_c, cdims[0], cdims[1], cdims[2], h->nx, h->ny, h->nz, h->nxp,
                                        h->nyp, h->ns, umax, rho_lo, rho_hi, h->d_kind, flux_x,
                                        plane_flux ? d_tgt : nullptr, plane_drho ? d_drho : nullptr,
                                        d_in + 4 * h->ns, d_in, d_cnt, h->stream));
            h->reset_state();
            h->launch(F_INIT | F_MACRO | F_REDUCE, 1.0, d_in, d_in + h->ns);
            unsigned long long cnt = 0;
            HIPCHK(hipMemcpyAsync(&cnt, d_cnt, sizeof(cnt), hipMemcpyDeviceToHost, h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
            if (nonfinite) *nonfinite = (int64_t)cnt;
        } catch (...) {
            (void)hipStreamSynchronize(h->stream);
            dfree(d_in);
            dfree(d_cnt);
            throw;
        }
        dfree(d_in);
        dfree(d_cnt);
    });
}

namespace {

// Link fractions of (node, direction) pairs: nodes are unpadded box indices; for d < ndirs whose
// neighbour is in the box and solid, the segment from the node's point (z shifted by zshift) to
// point + c_d h; q_out[node][Q] = distance / h of the first triangle in file order (nearest = 0,
// the reference) or of the nearest intersection (nearest = 1), -1 without one.
void link_q(pvlbm_sim* h, const std::vector<int64_t>& nodes, int ndirs, double zshift, int nearest,
            double* q_out) {
    const int64_t ntri = (int64_t)h->tri.size() / 12;
    if (ntri == 0) fail(PVLBM_ESTATE, "the handle was created without triangles");
    const int Q = h->q;
    for (size_t i = 0; i < nodes.size() * (size_t)Q; i++) q_out[i] = -1.0;
    std::vector<uint8_t> k((size_t)h->nsp);
    HIPCHK(hipMemcpyAsync(k.data(), h->d_kind, (size_t)h->nsp, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    const int* cx = Q == 19 ? Lat<19>::cx : Lat<27>::cx;
    const int* cy = Q == 19 ? Lat<19>::cy : Lat<27>::cy;
    const int* cz = Q == 19 ? Lat<19>::cz : Lat<27>::cz;
    // (node, direction) pairs whose neighbour is in the box and solid (ComplexGeometry.cpp:559-572)
    std::vector<double> seg;
    std::vector<int64_t> where;
    for (size_t w = 0; w < nodes.size(); w++) {
        const int64_t b = nodes[w], i = b % h->nx, j = (b / h->nx) % h->ny, kk = b / (h->nx * h->ny);
        for (int d = 0; d < ndirs; d++) {
            const int64_t a = i + cx[d], bb = j + cy[d], c = kk + cz[d];
            if (a < 0 || a >= h->nx || bb < 0 || bb >= h->ny || c < 0 || c >= h->nz) continue;
            if (k[(size_t)(a + h->nxp * (bb + h->nyp * c))] != 0) continue;
            // line.start = the point (min + idx * h), line.end = start + c * h
            const double sx = h->origin[0] + (double)i * h->h, sy = h->origin[1] + (double)j * h->h,
                         sz = h->origin[2] + (double)kk * h->h + zshift;
            const double lx = cx[d] * h->h, ly = cy[d] * h->h, lz = cz[d] * h->h;
            seg.insert(seg.end(), {sx, sy, sz, sx + lx, sy + ly, sz + lz});
            where.push_back((int64_t)w * Q + d);
        }
    }
    const int64_t np = (int64_t)where.size();
    if (np == 0) return;
    // triangles binned into cells of 2h over the box with a margin (bounding boxes widened
    // by 1e-6 of the geometry's scale against rounding at cell borders)
    const double cs = 2.0 * h->h;
    const double og[3] = {h->origin[0] - cs, h->origin[1] - cs, h->origin[2] - cs};
    const int64_t dims[3] = {h->nx, h->ny, h->nz};
    int32_t nc[3];
    for (int a = 0; a < 3; a++) nc[a] = (int32_t)((double)(dims[a] + 3) * h->h / cs) + 2;
    double scale = 1.0;
    for (int64_t t = 0; t < ntri * 12; t++) scale = std::max(scale, std::fabs(h->tri[(size_t)t]));
    const double m = 1e-6 * scale;
    const int64_t ncell = (int64_t)nc[0] * nc[1] * nc[2];
    auto range = [&](int64_t t, int r0[3], int r1[3]) {
        const double* T = &h->tri[(size_t)(12 * t)];
        for (int a = 0; a < 3; a++) {
            double lo = std::min(T[a], std::min(T[3 + a], T[6 + a])) - m;
            double hi = std::max(T[a], std::max(T[3 + a], T[6 + a])) + m;
            r0[a] = (int)std::min<double>(std::max(std::floor((lo - og[a]) / cs), 0.0), nc[a] - 1);
            r1[a] = (int)std::min<double>(std::max(std::floor((hi - og[a]) / cs), 0.0), nc[a] - 1);
        }
    };
    // counting sort of (cell, triangle) entries; triangles ascending within a cell
    std::vector<int64_t> count((size_t)ncell + 1, 0);
    for (int64_t t = 0; t < ntri; t++) {
        int r0[3], r1[3];
        range(t, r0, r1);
        for (int z = r0[2]; z <= r1[2]; z++)
            for (int y = r0[1]; y <= r1[1]; y++)
                for (int x = r0[0]; x <= r1[0]; x++) count[(size_t)(((int64_t)z * nc[1] + y) * nc[0] + x) + 1]++;
    }
    for (int64_t c = 0; c < ncell; c++) count[(size_t)c + 1] += count[(size_t)c];
    if (count[(size_t)ncell] >= INT32_MAX) fail(PVLBM_EINVAL, "too many triangle-cell entries");
    std::vector<int32_t> start(count.begin(), count.end());
    std::vector<int32_t> cell_tri((size_t)std::max<int64_t>(count[(size_t)ncell], 1));
    std::vector<int64_t> fill(count.begin(), count.end() - 1);
    for (int64_t t = 0; t < ntri; t++) {
        int r0[3], r1[3];
        range(t, r0, r1);
        for (int z = r0[2]; z <= r1[2]; z++)
            for (int y = r0[1]; y <= r1[1]; y++)
                for (int x = r0[0]; x <= r1[0]; x++)
                    cell_tri[(size_t)fill[(size_t)(((int64_t)z * nc[1] + y) * nc[0] + x)]++] = (int32_t)t;
    }
    double *d_seg = nullptr, *d_tri = nullptr, *d_q = nullptr;
    int32_t *d_start = nullptr, *d_ct = nullptr;
    try {
        d_seg = dalloc<double>(seg.size());
        d_tri = dalloc<double>(h->tri.size());
        d_q = dalloc<double>((size_t)np);
        d_start = dalloc<int32_t>(start.size());
        d_ct = dalloc<int32_t>(cell_tri.size());
        HIPCHK(hipMemcpyAsync(d_seg, seg.data(), sizeof(double) * seg.size(), hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(d_tri, h->tri.data(), sizeof(double) * h->tri.size(), hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(d_start, start.data(), sizeof(int32_t) * start.size(), hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(d_ct, cell_tri.data(), sizeof(int32_t) * cell_tri.size(), hipMemcpyHostToDevice, h->stream));
        HIPCHK(launch_link_q(d_seg, np, d_tri, d_start, d_ct, og, cs, nc, h->h, nearest, d_q, h->stream));
        std::vector<double> q((size_t)np);
        HIPCHK(hipMemcpyAsync(q.data(), d_q, sizeof(double) * (size_t)np, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        for (int64_t p = 0; p < np; p++) q_out[where[(size_t)p]] = q[(size_t)p];
    } catch (...) {
        (void)hipStreamSynchronize(h->stream);
        dfree(d_seg);
        dfree(d_tri);
        dfree(d_q);
        dfree(d_start);
        dfree(d_ct);
        throw;
    }
    dfree(d_seg);
    dfree(d_tri);
    dfree(d_q);
    dfree(d_start);
    dfree(d_ct);
}

}  // namespace

extern "C" int pvlbm_link_fractions(pvlbm_handle h, int ndirs, double* q_out) {
    if (!h) return PVLBM_EINVAL;
    return guarded(h, [&] {
        if (!q_out) fail(PVLBM_EINVAL, "q_out is NULL");
        if (ndirs < 1 || ndirs > h->q) fail(PVLBM_EINVAL, "ndirs must be in [1, Q]");
        link_q(h, h->lists[PVLBM_LIST_WALL], ndirs, 0.0, 0, q_out);
    });
}

extern "C" int pvlbm_set_wall_closure(pvlbm_handle h, int mode) {
    if (!h) return PVLBM_EINVAL;
    return guarded(h, [&] {
        if (mode != PVLBM_WALL_REFERENCE && mode != PVLBM_WALL_INTERPOLATED) fail(PVLBM_EINVAL, "unknown wall closure");
        if (mode == PVLBM_WALL_INTERPOLATED && h->streaming == PVLBM_STREAMING_IN_PLACE)
            fail(PVLBM_EINVAL, "the interpolated wall closure needs two lattices (streaming is set to in place)");
        if (mode == PVLBM_WALL_INTERPOLATED && !h->d_bq) {
            // fractions of every link of every wall slot, per link length, nearest crossing of
            // the segment shifted like the voxelisation's test points
            const int Q = h->q;
            std::vector<int64_t> nodes;
            std::vector<int32_t> slot;
            for (size_t sl = 0; sl < h->bnode.size(); sl++)
                if (h->bclose[sl] == BC_WALL) {
                    nodes.push_back(h->box_of_pad(h->bnode[sl]));
                    slot.push_back((int32_t)sl);
                }
            std::vector<double> q(nodes.size() * (size_t)Q);
            if (!nodes.empty()) link_q(h, nodes, Q, 0.5 * h->h, 1, q.data());
            const int* cx = Q == 19 ? Lat<19>::cx : Lat<27>::cx;
            const int* cy = Q == 19 ? Lat<19>::cy : Lat<27>::cy;
            const int* cz = Q == 19 ? Lat<19>::cz : Lat<27>::cz;
            std::vector<float> bq(std::max<size_t>(h->bnode.size(), 1) * (size_t)Q, -1.0f);
            for (size_t w = 0; w < nodes.size(); w++)
                for (int d = 0; d < Q; d++) {
                    const double v = q[w * Q + d];
                    if (v < 0.0) continue;
                    const double len = std::sqrt((double)(cx[d] * cx[d] + cy[d] * cy[d] + cz[d] * cz[d]));
                    bq[(size_t)slot[w] * Q + d] = (float)std::min(1.0, v / len);
                }
            h->d_bq = dalloc<float>(bq.size());
            HIPCHK(hipMemcpyAsync(h->d_bq, bq.data(), sizeof(float) * bq.size(), hipMemcpyHostToDevice, h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
        }
        h->set_mode_and_sync(h->wall_mode, mode);
    });
}

extern "C" int pvlbm_set_streaming(pvlbm_handle h, int mode) {
    if (!h) return PVLBM_EINVAL;
    return guarded(h, [&] {
        if (mode != PVLBM_STREAMING_AUTO && mode != PVLBM_STREAMING_TWO_LATTICES && mode != PVLBM_STREAMING_IN_PLACE)
            fail(PVLBM_EINVAL, "unknown streaming scheme %d", mode);
        if (mode == PVLBM_STREAMING_IN_PLACE && h->wall_mode == PVLBM_WALL_INTERPOLATED)
            fail(PVLBM_EINVAL, "the interpolated wall closure needs two lattices");
        h->set_mode_and_sync(h->streaming, mode);
    });
}

extern "C" int pvlbm_get_streaming(pvlbm_handle h, int* in_place) {
    if (!h) return PVLBM_EINVAL;
    return guarded(h, [&] {
        if (!in_place) fail(PVLBM_EINVAL, "in_place is NULL");
        *in_place = h->in_place() ? 1 : 0;
    });
}

extern "C" int pvlbm_set_store_policy(pvlbm_handle h, int policy) {
    if (!h) return PVLBM_EINVAL;
    return guarded(h, [&] {
        if (policy < PVLBM_STORES_AUTO || policy > PVLBM_STORES_STREAMING) fail(PVLBM_EINVAL, "unknown store policy %d", policy);
        h->store_policy = policy;
    });
}

extern "C" int pvlbm_get_store_policy(pvlbm_handle h, int* streaming) {
    if (!h) return PVLBM_EINVAL;
    return guarded(h, [&] {
        if (!streaming) fail(PVLBM_EINVAL, "streaming is NULL");
        *streaming = h->stream_stores() ? 1 : 0;
    });
}

extern "C" int pvlbm_get_load_policy(pvlbm_handle h, int* streaming) {
    if (!h) return PVLBM_EINVAL;
    return guarded(h, [&] {
        if (!streaming) fail(PVLBM_EINVAL, "streaming is NULL");
        *streaming = h->stream_loads();
    });
}

extern "C" int pvlbm_set_wss_every(pvlbm_handle h, int every) {
    if (!h) return PVLBM_EINVAL;
    return guarded(h, [&] {
        if (every < 1) fail(PVLBM_EINVAL, "every must be >= 1");
        h->wss_every = every;
    });
}

namespace {
void run_steps(pvlbm_sim* h, int n, const double* scale) {
    for (int s = 0; s < n; s++) {
        int flags = (h->iter % h->wss_every == 0) ? F_WSS : 0;
        if (s == n - 1) flags |= F_MACRO | F_REDUCE;
        h->launch(flags, scale ? scale[s] : 1.0);
        h->iter++;
    }
}
}  // namespace

extern "C" int pvlbm_step(pvlbm_handle h, int n, const double* inlet_scale) {
    if (!h) return PVLBM_EINVAL;
    return guarded(h, [&] {
        if (n < 0) fail(PVLBM_EINVAL, "n < 0");
        run_steps(h, n, inlet_scale);
        HIPCHK(hipStreamSynchronize(h->stream));
    });
}

extern "C" int pvlbm_step_timed(pvlbm_handle h, int n, const double* inlet_scale, double* ms) {
    if (!h) return PVLBM_EINVAL;
    return guarded(h, [&] {
        if (n < 0 || !ms) fail(PVLBM_EINVAL, "bad arguments");
        HIPCHK(hipEventRecord(h->ev0, h->stream));
        run_steps(h, n, inlet_scale);
        HIPCHK(hipEventRecord(h->ev1, h->stream));
        HIPCHK(hipEventSynchronize(h->ev1));
        float f = 0;
        HIPCHK(hipEventElapsedTime(&f, h->ev0, h->ev1));
        *ms = f;
    });
}

extern "C" int pvlbm_iteration(pvlbm_handle h, int64_t* it) {
    if (!h || !it) return PVLBM_EINVAL;
    *it = h->iter;
    return PVLBM_OK;
}

extern "C" int pvlbm_get_macros(pvlbm_handle h, double* rho, double* u) {
    if (!h) return PVLBM_EINVAL;
    return guarded(h, [&] {
        std::vector<double> m((size_t)(4 * h->ns));
        HIPCHK(hipMemcpyAsync(m.data(), h->d_macro, sizeof(double) * m.size(), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        std::vector<uint8_t> k((size_t)h->nsp);
        HIPCHK(hipMemcpy(k.data(), h->d_kind, (size_t)h->nsp, hipMemcpyDeviceToHost));
        const int64_t nbox = h->nx * h->ny * h->nz;
        for (int64_t r = 0; r < h->ny * h->nz; r++)
            for (int64_t i = 0; i < h->nx; i++) {
                const int64_t n = h->prow(r) * h->nxp + i, b = r * h->nx + i;
                const bool node = k[n] != K_SOLID;
                if (rho) rho[b] = node ? m[n] : 0.0;
                if (u)
                    for (int c = 0; c < 3; c++) u[c * nbox + b] = node ? m[(c + 1) * h->ns + n] : 0.0;
            }
    });
}

extern "C" int pvlbm_get_macros_device(pvlbm_handle h, float* rho, float* u) {
    if (!h) return PVLBM_EINVAL;
    return guarded(h, [&] {
        if (!rho || !u) fail(PVLBM_EINVAL, "rho and u must be device pointers");
        HIPCHK(launch_macros_f32(h->d_macro, h->d_kind, h->nx, h->ny, h->nz, h->nxp, h->nyp, h->ns, rho, u,
                                 h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    });
}

extern "C" int pvlbm_get_normals(pvlbm_handle h, double* n3) {
    if (!h) return PVLBM_EINVAL;
    return guarded(h, [&] {
        if (!n3) fail(PVLBM_EINVAL, "n3 is NULL");
        if (h->nw > 0) {
            // device rows are [3][max(nw, 1)]; nw == max(nw, 1) here, so one contiguous copy
            HIPCHK(hipMemcpyAsync(n3, h->d_nrm, sizeof(double) * 3 * h->nw, hipMemcpyDeviceToHost, h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
        }
    });
}

extern "C" int pvlbm_get_wss_magnitude(pvlbm_handle h, double* out) {
    if (!h) return PVLBM_EINVAL;
    return guarded(h, [&] {
        if (!out) fail(PVLBM_EINVAL, "out is NULL");
        if (h->nw > 0) {
            HIPCHK(hipMemcpyAsync(out, h->d_wss, sizeof(double) * h->nw, hipMemcpyDeviceToHost, h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
        }
    });
}

extern "C" int pvlbm_get_wss(pvlbm_handle h, double* out6) {
    if (!h) return PVLBM_EINVAL;
    return guarded(h, [&] {
        if (!out6) fail(PVLBM_EINVAL, "out6 is NULL");
        const int64_t nw = h->nw, nw1 = std::max<int64_t>(nw, 1);
        std::vector<double> wss((size_t)nw1), acc((size_t)(4 * nw1));
        HIPCHK(hipMemcpyAsync(wss.data(), h->d_wss, sizeof(double) * nw1, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(acc.data(), h->d_acc, sizeof(double) * 4 * nw1, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        for (int64_t w = 0; w < nw; w++) {
            const int64_t b = h->lists[PVLBM_LIST_NTW][w];
            const int64_t i = b % h->nx, j = (b / h->nx) % h->ny, k = b / (h->nx * h->ny);
            const double a0 = acc[w], a1 = acc[nw1 + w], a2 = acc[2 * nw1 + w], am = acc[3 * nw1 + w];
            // ComputeOSIAndTAWSS (Simulation.cpp:899-905), period 0.9 (pulsatile.h:12)
            const double tawss = am / 0.9;
            const double norm_acc = std::sqrt(a0 * a0 + a1 * a1 + a2 * a2);
            const double osi = 0.5 * (1.0 - (norm_acc / am));
            double* o = out6 + 6 * w;
            o[0] = h->origin[0] + (double)i * h->h;
            o[1] = h->origin[1] + (double)j * h->h;
            o[2] = h->origin[2] + (double)k * h->h;
            o[3] = wss[w];
            o[4] = osi;
            o[5] = tawss;
        }
    });
}

extern "C" int pvlbm_get_wss_raw(pvlbm_handle h, double* tau3, double* acc4) {
    if (!h) return PVLBM_EINVAL;
    return guarded(h, [&] {
        const int64_t nw = h->nw, nw1 = std::max<int64_t>(nw, 1);
        std::vector<double> t((size_t)(3 * nw1)), a((size_t)(4 * nw1));
        HIPCHK(hipMemcpyAsync(t.data(), h->d_tau3, sizeof(double) * t.size(), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(a.data(), h->d_acc, sizeof(double) * a.size(), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        for (int64_t w = 0; w < nw; w++) {
            if (tau3)
                for (int c = 0; c < 3; c++) tau3[c * nw + w] = t[c * nw1 + w];
            if (acc4)
                for (int c = 0; c < 4; c++) acc4[c * nw + w] = a[c * nw1 + w];
        }
    });
}

extern "C" int pvlbm_reduce(pvlbm_handle h, double out[4]) {
    if (!h) return PVLBM_EINVAL;
    return guarded(h, [&] {
        if (!out) fail(PVLBM_EINVAL, "out is NULL");
        if (!h->red_valid) fail(PVLBM_ESTATE, "no reduction computed yet");
        HIPCHK(hipMemcpyAsync(out, h->d_red, sizeof(double) * 4, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    });
}

extern "C" int pvlbm_gate_events(pvlbm_handle h, int64_t* n) {
    if (!h) return PVLBM_EINVAL;
    return guarded(h, [&] {
        if (!n) fail(PVLBM_EINVAL, "n is NULL");
        unsigned long long v = 0;
        HIPCHK(hipMemcpyAsync(&v, h->d_gate, sizeof v, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        *n = (int64_t)v;
    });
}

extern "C" double pvlbm_pulsatile_velocity(double x, double period) {
    // Simulation::getInletVelocity (Simulation.cpp:548-563) with the Pulsatile polynomials
    // (pulsatile.h:16-28); x^n through pow as the reference writes it.
    double vel = 0.0;
    double t = std::fmod(x, period);
    if (t >= 0 && t < (0.1 * period)) {
        vel = (-8.579e+08 * std::pow(t, 8) + 4.059e+08 * std::pow(t, 7) - 7.817e+07 * std::pow(t, 6) +
               7.994e+06 * std::pow(t, 5) - 4.835e+05 * std::pow(t, 4) + 1.782e+04 * std::pow(t, 3) -
               321.6 * std::pow(t, 2) + 3.215 * t + 0.1423);
    } else if (t >= (0.1 * period) && t < (0.3 * period)) {
        vel = (-4.32e+06 * std::pow(t, 8) + 7.94e+06 * std::pow(t, 7) - 6.298e+06 * std::pow(t, 6) +
               2.82e+06 * std::pow(t, 5) - 7.809e+05 * std::pow(t, 4) + 1.37e+05 * std::pow(t, 3) -
               1.484e+04 * std::pow(t, 2) + 901.9 * t - 22.87);
    } else if (t >= (0.3 * period) && t < (0.9 * period)) {
        vel = (-5.123 * std::pow(t, 4) + 14.15 * std::pow(t, 3) - 14.21 * std::pow(t, 2) + 5.865 * t - 0.5696);
    }
    return vel;
}

extern "C" int pvlbm_destroy(pvlbm_handle h) {
    delete h;
    return PVLBM_OK;
}

extern "C" int pvlbm_copy_bench(int64_t bytes, int reps, double* ms) {
    return guarded(nullptr, [&] {
        if (bytes < 16 || reps < 1 || !ms) fail(PVLBM_EINVAL, "bad copy arguments");
        bytes = bytes / 16 * 16;
        hipStream_t s;
        HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        char *a = nullptr, *b = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        try {
            a = dalloc<char>((size_t)bytes);
            b = dalloc<char>((size_t)bytes);
            HIPCHK(hipMemsetAsync(a, 1, (size_t)bytes, s));
            HIPCHK(hipEventCreate(&e0));
            HIPCHK(hipEventCreate(&e1));
            HIPCHK(launch_copy(a, b, bytes, s));
            HIPCHK(hipEventRecord(e0, s));
            for (int r = 0; r < reps; r++) HIPCHK(launch_copy(a, b, bytes, s));
            HIPCHK(hipEventRecord(e1, s));
            HIPCHK(hipEventSynchronize(e1));
            float f = 0;
            HIPCHK(hipEventElapsedTime(&f, e0, e1));
            *ms = f;
        } catch (...) {
            (void)hipStreamSynchronize(s);
            dfree(a); dfree(b);
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            (void)hipStreamDestroy(s);
            throw;
        }
        dfree(a); dfree(b);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(s);
    });
}
