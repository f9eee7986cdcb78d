// Test infrastructure: drives the C++ facade's reference loop (include/pvlbm.hpp,
// pvlbm::Simulation::Iterate, which follows Simulation::Iterate, Simulation.cpp:569-601) inside the
// calling process, so that the device test of the loop needs no second process on the GPU.  Built
// in-tree next to the library by the package build; loaded by tests/test_gpu_facade.py only.
#include <cstdint>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <vector>

#include "pvlbm.hpp"

extern "C" {

// Geometry from the triangles (12 doubles each: v1, v2, v3, file normal), SubDomain with the given
// faces (FP64), Simulation(tau, 1.0, dt), then Iterate(t_end, period, output_every) with field
// capture.  iters[0 .. *n_out) receives the iteration index of every output; rho_box / u_box
// (nx*ny*nz and 3*nx*ny*nz) and wss6 (6 per next-to-wall node) receive the last output's fields and
// surface rows; *runtime and *iterations the loop's total_runtime and iteration count.  Returns 0,
// or 1 with the message in err.
int facade_iterate(const double* tri, int64_t ntri, double spacing, int q, const pvlbm_face* faces, int nfaces,
                   double tau, double dt, double t_end, double period, int output_every, int64_t* iters,
                   int max_out, int* n_out, double* rho_box, double* u_box, double* wss6, int64_t nwss,
                   double* runtime, int64_t* iterations, char* err, int errlen) {
    try {
        pvlbm::LatticeModel velSet(q);
        pvlbm::Geometry geo(std::vector<double>(tri, tri + 12 * ntri), velSet, spacing);
        pvlbm::SubDomain sd(geo, std::vector<pvlbm_face>(faces, faces + nfaces), PVLBM_FP64, tau, dt);
        pvlbm::Simulation sim(sd, tau, 1.0, dt);
        *n_out = 0;
        sim.Iterate(t_end, period, output_every, [&](int64_t it) {
            if (*n_out < max_out) iters[*n_out] = it;
            ++*n_out;
            const std::vector<double>& r = sim.outputRho();
            const std::vector<double>& u = sim.outputU();
            std::memcpy(rho_box, r.data(), sizeof(double) * r.size());
            std::memcpy(u_box, u.data(), sizeof(double) * u.size());
            const std::vector<double> w = sim.surfaceAttributes();
            if ((int64_t)w.size() != 6 * nwss) throw std::runtime_error("unexpected number of surface rows");
            std::memcpy(wss6, w.data(), sizeof(double) * w.size());
        }, true);
        *runtime = sim.total_runtime;
        *iterations = sim.iteration();
        return 0;
    } catch (const std::exception& e) {
        std::strncpy(err, e.what(), (size_t)errlen - 1);
        err[errlen - 1] = 0;
        return 1;
    }
}
}
