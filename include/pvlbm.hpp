/*
 * pvlbm.hpp — C++ facade with the reference solver's class names over the C ABI of pvlbm.h.
 *
 * The reference builds its solver as  LatticeModel -> Geometry -> SubDomain -> Simulation
 * (LBM.cpp:69-72).  These classes keep those names, constructor arguments and call order; the
 * state lives on the GPU behind a pvlbm_handle.  Errors throw pvlbm::Error (the reference prints
 * and continues, Simulation.cpp:298-300; a device solver cannot continue after a failed call).
 * Header-only; link with libpvlbm.so.
 */
#ifndef PVLBM_HPP
#define PVLBM_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "pvlbm.h"

namespace pvlbm {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc, pvlbm_handle h = nullptr) {
    if (rc != PVLBM_OK) throw Error(rc, pvlbm_last_error(h));
}

/* LatticeModel (LatticeModel.h:9): the reference picks D3Q19 or D3Q27 at compile time
 * (Macro.h:6); here it is a constructor argument. */
class LatticeModel {
public:
    explicit LatticeModel(int q = 19) : q_(q) {
        if (q != 19 && q != 27) throw Error(PVLBM_EINVAL, "lattice must be 19 or 27");
    }
    int Get_nDirections() const { return q_; }
    int Get_nDimensions() const { return 3; }

private:
    int q_;
};

/* Geometry(FileName, velSet, spacing, raydirection) (ComplexGeometry.h:91): reads the ASCII STL,
 * takes the reference bbox and grid and labels the points by +z ray parity on the GPU.  A second
 * constructor takes triangles already in memory and an optional explicit grid. */
class Geometry {
public:
    /* ASCII or binary STL (GeometryReader.cpp:33-176): binary when the file is exactly the
     * 84-byte header plus 50 bytes per facet its count declares. */
    Geometry(const std::string& FileName, const LatticeModel& velSet, double spacing, char raydirection = 'z')
        : velSet_(velSet), spacing(spacing), raydirection(raydirection) {
        const bool binary = is_binary_stl(FileName);
        auto read = binary ? pvlbm_read_stl_binary : pvlbm_read_stl_ascii;
        int64_t n = 0;
        check(read(FileName.c_str(), nullptr, 0, &n));
        if (n == 0) throw Error(PVLBM_EIO, "no facets read from " + FileName);
        triangles.resize((size_t)(12 * n));
        check(read(FileName.c_str(), triangles.data(), n, &n));
        label(nullptr);
    }
    static bool is_binary_stl(const std::string& path) {
        FILE* f = std::fopen(path.c_str(), "rb");
        if (!f) return false;
        unsigned char hdr[84];
        const bool full = std::fread(hdr, 1, 84, f) == 84;
        std::fseek(f, 0, SEEK_END);
        const long size = std::ftell(f);
        std::fclose(f);
        if (!full) return false;
        const uint32_t n = (uint32_t)hdr[80] | ((uint32_t)hdr[81] << 8) | ((uint32_t)hdr[82] << 16) | ((uint32_t)hdr[83] << 24);
        return size == 84L + 50L * (long)n;
    }
    Geometry(std::vector<double> tri, const LatticeModel& velSet, double spacing,
             const pvlbm_grid* grid = nullptr)
        : velSet_(velSet), spacing(spacing), raydirection('z'), triangles(std::move(tri)) {
        label(grid);
    }

    const LatticeModel& velSet() const { return velSet_; }
    int64_t ntri() const { return (int64_t)(triangles.size() / 12); }
    double coord(int axis, int64_t idx) const { return grid.origin[axis] + (double)idx * spacing; }

    LatticeModel velSet_;
    double spacing;
    char raydirection;  // stored; labelling always casts +z rays (ComplexGeometry.cpp:994)
    std::vector<double> triangles;
    double boundingbox[6] = {0, 0, 0, 0, 0, 0};
    pvlbm_grid grid{};
    std::vector<uint8_t> fluid;  // nx*ny*nz, x fastest

private:
    void label(const pvlbm_grid* g) {
        check(pvlbm_bbox(triangles.data(), ntri(), boundingbox));
        check(pvlbm_voxelize(triangles.data(), ntri(), spacing, g, &grid, nullptr));
        fluid.resize((size_t)grid.dims[0] * grid.dims[1] * grid.dims[2]);
        check(pvlbm_voxelize(triangles.data(), ntri(), spacing, g, &grid, fluid.data()));
    }
};

/* SubDomain(Geometry&) (Subdomain.h:66) with the boundary faces that the reference hard-codes
 * (ComplexGeometry.cpp:1378-1408) passed in; storage selects FP64 or the shifted 16-bit forms.
 * tau and del_t take the reference defaults until a Simulation sets them. */
class SubDomain {
public:
    SubDomain(const Geometry& geo, const std::vector<pvlbm_face>& faces, int storage = PVLBM_FP64,
              double tau = 1.0, double del_t = 8e-6)
        : geo_(&geo) {
        check(pvlbm_create(geo.fluid.data(), &geo.grid, geo.spacing, geo.velSet().Get_nDirections(),
                           storage, faces.empty() ? nullptr : faces.data(), (int)faces.size(),
                           geo.triangles.data(), geo.ntri(), tau, del_t, &h_));
    }
    ~SubDomain() {
        if (h_) pvlbm_destroy(h_);
    }
    SubDomain(const SubDomain&) = delete;
    SubDomain& operator=(const SubDomain&) = delete;

    std::vector<int64_t> list(int which) const {
        int64_t n = 0;
        check(pvlbm_count(h_, which, &n), h_);
        std::vector<int64_t> v((size_t)n);
        if (n) check(pvlbm_list(h_, which, v.data()), h_);
        return v;
    }
    int64_t count(int which) const {
        int64_t n = 0;
        check(pvlbm_count(h_, which, &n), h_);
        return n;
    }
    // node kind per grid point (PVLBM_SOLID .. PVLBM_WALL), x fastest
    std::vector<uint8_t> kinds() const {
        const size_t n = (size_t)geo_->grid.dims[0] * geo_->grid.dims[1] * geo_->grid.dims[2];
        std::vector<uint8_t> k(n);
        check(pvlbm_get_kinds(h_, k.data()), h_);
        return k;
    }
    /* Geometry::CalculateWallDistances (ComplexGeometry.cpp:550-608): q per wall node (list order)
     * and direction, -1 without a link; ndirs = 19 is the reference's loop. */
    std::vector<double> linkFractions(int ndirs = 19) const {
        std::vector<double> q((size_t)(count(PVLBM_LIST_WALL) * geo_->velSet().Get_nDirections()));
        check(pvlbm_link_fractions(h_, ndirs, q.data()), h_);
        return q;
    }
    /* SubDomain::calculateBoundaryNormals (Subdomain.cpp:224-370): unit normals [3][nw] of the
     * next-to-wall nodes in list order. */
    std::vector<double> normals() const {
        std::vector<double> n3((size_t)(3 * count(PVLBM_LIST_NTW)));
        if (!n3.empty()) check(pvlbm_get_normals(h_, n3.data()), h_);
        return n3;
    }
    std::vector<int64_t> InletBoundaryNode_() const { return list(PVLBM_LIST_INLET); }
    std::vector<int64_t> WallBoundaryNode_() const { return list(PVLBM_LIST_WALL); }
    std::vector<int64_t> nextToTheWall() const { return list(PVLBM_LIST_NTW); }
    const Geometry& Geometry_() const { return *geo_; }
    pvlbm_handle handle() const { return h_; }

private:
    const Geometry* geo_;
    pvlbm_handle h_ = nullptr;
};

/* Simulation(SubDomain&, tau, pressure, del_t) (Simulation.h:26).  The reference runs its whole
 * loop inside the constructor (Simulation.cpp:13); here Iterate() does, and step() advances any
 * number of iterations of the loop body (Simulation.cpp:577-586). */
class Simulation {
public:
    Simulation(SubDomain& sd, double tau, double pressure, double del_t)
        : sd_(sd), h_(sd.handle()), pressure_(pressure), m_dt(del_t) {
        SetRelaxation(tau);
        check(pvlbm_set_dt(h_, del_t), h_);
    }
    void SetRelaxation(double tau) {
        m_relaxation = tau;
        check(pvlbm_set_tau(h_, tau), h_);
    }
    void setInletProfile(const std::vector<double>& u) { check(pvlbm_set_inlet_profile(h_, u.data()), h_); }
    void setWssEvery(int every) { check(pvlbm_set_wss_every(h_, every), h_); }
    // PVLBM_WALL_REFERENCE (the reference's wall rule) or PVLBM_WALL_INTERPOLATED
    void setWallClosure(int mode) { check(pvlbm_set_wall_closure(h_, mode), h_); }
    // PVLBM_STREAMING_AUTO (in place on one lattice up to 512 MiB, two lattices beyond),
    // PVLBM_STREAMING_TWO_LATTICES or PVLBM_STREAMING_IN_PLACE; the fields are the same
    void setStreaming(int mode) { check(pvlbm_set_streaming(h_, mode), h_); }
    // restart from fields (rho [nz][ny][nx], u three such planes); empty vectors: from rest
    void initFields(const std::vector<double>& rho, const std::vector<double>& u) {
        check(pvlbm_init_fields(h_, rho.empty() ? nullptr : rho.data(), u.empty() ? nullptr : u.data()), h_);
    }
    int64_t iteration() const {
        int64_t it = 0;
        check(pvlbm_iteration(h_, &it), h_);
        return it;
    }
    // FP64 (mass, inlet flux, outlet flux, non-finite count) of the last step
    std::vector<double> reduce() const {
        std::vector<double> r(4);
        check(pvlbm_reduce(h_, r.data()), h_);
        return r;
    }
    void step(int n, const double* inlet_scale = nullptr) { check(pvlbm_step(h_, n, inlet_scale), h_); }
    static double getInletVelocity(double t, double period) { return pvlbm_pulsatile_velocity(t, period); }

    /* Iterations 0 .. n-1 of the loop body with InletVelocity_ = scale[iter].  output(iter) is
     * called after every iteration with iter % output_every == 0, iteration 0 included (WriteOutput,
     * Simulation.cpp:589-591); inside it surfaceAttributes() holds that iteration's WSS, OSI and
     * TAWSS and, when fields is true, outputRho() / outputU() hold the density and velocity its
     * Collision computed (Simulation.cpp:52-105).  One engine launch finishes an iteration and runs
     * the next one's collision, so those fields are read before the launch that finishes it. */
    template <class F>
    void Run(const std::vector<double>& scale, int output_every, F output, bool fields = false) {
        if (output_every < 1) throw std::runtime_error("output_every must be >= 1");
        const int64_t n = (int64_t)scale.size();
        for (int64_t it = 0; it < n;) {
            // iterations it .. last, ending at the next output iteration
            const int64_t next = (it + output_every - 1) / output_every * output_every;
            const int64_t last = std::min(n - 1, next);
            const bool out = last % output_every == 0;
            const int64_t m = last - it + 1;
            if (out && fields) {
                if (m > 1) step((int)(m - 1), scale.data() + it);
                macros(out_rho_, out_u_);
                step(1, scale.data() + last);
            } else {
                step((int)m, scale.data() + it);
            }
            it = last + 1;
            if (out) output(last);
        }
    }

    /* Simulation::Iterate (Simulation.cpp:569-601): total_runtime starts at 0; while
     * total_runtime <= t_end, one iteration with the pulsatile inlet at total_runtime, then
     * total_runtime += dt (an accumulated sum, as the reference does); outputs as Run. */
    template <class F>
    void Iterate(double t_end, double period, int output_every, F output, bool fields = false) {
        std::vector<double> sc;
        double t = 0.0;
        while (t <= t_end) {
            sc.push_back(getInletVelocity(t, period));
            t += m_dt;
        }
        Run(sc, output_every, output, fields);
        total_runtime = t;
    }

    // density [nz][ny][nx] and velocity (three such planes) written at the last output of Run
    const std::vector<double>& outputRho() const { return out_rho_; }
    const std::vector<double>& outputU() const { return out_u_; }

    /* x, y, z, WSS, OSI, TAWSS per next-to-wall node (OutputFile.cpp:884-894) */
    // (rho, u) of every grid point, rho [nz][ny][nx] then u as three such planes (x, y, z)
    void macros(std::vector<double>& rho, std::vector<double>& u) const {
        const Geometry& g = sd_.Geometry_();
        const size_t n = (size_t)g.grid.dims[0] * g.grid.dims[1] * g.grid.dims[2];
        rho.resize(n);
        u.resize(3 * n);
        check(pvlbm_get_macros(h_, rho.data(), u.data()), h_);
    }
    std::vector<double> surfaceAttributes() const {
        std::vector<double> out((size_t)(6 * sd_.count(PVLBM_LIST_NTW)));
        if (!out.empty()) check(pvlbm_get_wss(h_, out.data()), h_);
        return out;
    }
    // the WSS column of surfaceAttributes alone
    std::vector<double> wssMagnitude() const {
        std::vector<double> out((size_t)sd_.count(PVLBM_LIST_NTW));
        if (!out.empty()) check(pvlbm_get_wss_magnitude(h_, out.data()), h_);
        return out;
    }

    double total_runtime = 0.0;

private:
    SubDomain& sd_;
    pvlbm_handle h_;
    double pressure_;  // stored; the outlets use their face density (Simulation.cpp:630)
    double m_dt;
    double m_relaxation = 1.0;
    std::vector<double> out_rho_, out_u_;
};

}  // namespace pvlbm

#endif
