// lbm_run — command-line driver replacing the reference's main() (LBM.cpp:20-85).
//
// The reference hard-codes its inputs (PV.stl, spacing 0.05, tau 1.5, dt 8e-6, the PV.stl boundary
// planes, 2.7 s of the pulsatile waveform, output every 100 iterations).  Here each is a flag with
// the reference value as default where one exists:
//
//   lbm_run --stl FILE [--spacing 0.05] [--grid ox,oy,oz,nx,ny,nz] [--lattice 19|27]
//           [--storage fp64|fp16|bf16] [--tau 1.5] [--dt 8e-6]
//           [--faces pv|x] [--face role,closure,axis,plane,tol,lo0,lo1,hi0,hi1,rho]...
//           [--t-end 2.7] [--period 0.9] [--steps N --inlet const:V|sin:MEAN,AMP,PERIOD]
//           [--output-every 100] [--outdir .] [--vtu] [--wall reference|interpolated]
//
// --faces pv uses the reference's PV.stl planes (ComplexGeometry.cpp:1378-1408); --faces x puts the
// inlet on the grid's x-min plane and a pressure outlet (rho 1) on its x-max plane.  Without
// --steps the run is the reference loop: the pulsatile waveform until --t-end.  After every
// iteration whose index is a multiple of --output-every, iteration 0 included (Simulation.cpp:589-591),
// FlowHemodynamics/<iter>_FlowHemodynamics.csv is written in the reference's format
// (OutputFile.cpp:867-938); with --vtu also Fluid_Variables/<iter>.vtu with the density and velocity
// of that iteration's collision (OutputFile.cpp:259-429, named as at :275).

#include <sys/stat.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/pvlbm.hpp"

namespace {

std::vector<double> parse_list(const std::string& s) {
    std::vector<double> v;
    std::stringstream ss(s);
    std::string tok;
    while (std::getline(ss, tok, ',')) v.push_back(std::stod(tok));
    return v;
}

// plane is a long double so that the reference's long double literals (ComplexGeometry.cpp:1376,
// 1385) reach the face match exactly, as the pair plane + plane_lo
pvlbm_face make_face(int role, int closure, int axis, long double plane, double tol, double lo0, double lo1,
                     double hi0, double hi1, double rho) {
    pvlbm_face f{};
    f.role = role;
    f.closure = closure;
    f.axis = axis;
    f.plane = (double)plane;
    f.plane_lo = (double)(plane - (long double)f.plane);
    f.tol = tol;
    f.lo[0] = lo0;
    f.lo[1] = lo1;
    f.hi[0] = hi0;
    f.hi[1] = hi1;
    f.rho = rho;
    return f;
}

void write_csv(const std::string& outdir, int64_t iter, const std::vector<double>& rows) {
    std::string folder = outdir + "/FlowHemodynamics";
    mkdir(outdir.c_str(), 0777);
    mkdir(folder.c_str(), 0777);
    std::ofstream out(folder + "/" + std::to_string(iter) + "_FlowHemodynamics.csv");
    out << "X,Y,Z,WSS,OSI,TAWSS\n";
    for (size_t i = 0; i + 5 < rows.size(); i += 6)
        out << rows[i] << "," << rows[i + 1] << "," << rows[i + 2] << "," << rows[i + 3] << ","
            << rows[i + 4] << "," << rows[i + 5] << "\n";
}

// Fluid_Variables/<idx>.vtu in the content of OutputFile::WriteVTK_Collective
// (OutputFile.cpp:259-429): fluid nodes at integer lattice indices in z-y-x loop order, point data
// Velocity, Density, VelocityMagnitude, a hexahedron per cube of eight fluid corners; XML with raw
// appended binary blocks (UInt64 byte count, then little-endian data).
void write_vtu(const std::string& outdir, int64_t idx, const std::vector<uint8_t>& kind, int nx, int ny, int nz,
               const std::vector<double>& rho, const std::vector<double>& u) {
    const size_t n = (size_t)nx * ny * nz;
    std::vector<int64_t> pid(n, -1);
    std::vector<double> vel, den, mag, pts;
    for (size_t b = 0; b < n; b++)
        if (kind[b] != PVLBM_SOLID) {
            pid[b] = (int64_t)den.size();
            const double v0 = u[b], v1 = u[n + b], v2 = u[2 * n + b];
            vel.insert(vel.end(), {v0, v1, v2});
            den.push_back(rho[b]);
            mag.push_back(std::sqrt(v0 * v0 + v1 * v1 + v2 * v2));
            pts.insert(pts.end(), {(double)(b % nx), (double)((b / nx) % ny), (double)(b / ((size_t)nx * ny))});
        }
    std::vector<int64_t> conn;
    const int cx[8] = {0, 1, 1, 0, 0, 1, 1, 0}, cy[8] = {0, 0, 1, 1, 0, 0, 1, 1}, cz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    for (int z = 0; z + 1 < nz; z++)
        for (int y = 0; y + 1 < ny; y++)
            for (int x = 0; x + 1 < nx; x++) {
                int64_t c[8];
                bool ok = true;
                for (int i = 0; i < 8 && ok; i++) {
                    c[i] = pid[((size_t)(z + cz[i]) * ny + (y + cy[i])) * nx + (x + cx[i])];
                    ok = c[i] >= 0;
                }
                if (ok) conn.insert(conn.end(), c, c + 8);
            }
    const size_t m = conn.size() / 8;
    std::vector<int64_t> offs(m);
    for (size_t i = 0; i < m; i++) offs[i] = 8 * (int64_t)(i + 1);
    std::vector<uint8_t> types(m, 12);  // VTK_HEXAHEDRON
    struct Blk { const char* name; const char* type; int nc; const void* data; size_t bytes; };
    const Blk blk[7] = {{"Velocity", "Float64", 3, vel.data(), vel.size() * 8},
                        {"Density", "Float64", 1, den.data(), den.size() * 8},
                        {"VelocityMagnitude", "Float64", 1, mag.data(), mag.size() * 8},
                        {nullptr, "Float64", 3, pts.data(), pts.size() * 8},
                        {"connectivity", "Int64", 1, conn.data(), conn.size() * 8},
                        {"offsets", "Int64", 1, offs.data(), offs.size() * 8},
                        {"types", "UInt8", 1, types.data(), types.size()}};
    size_t off[7], o = 0;
    for (int i = 0; i < 7; i++) {
        off[i] = o;
        o += 8 + blk[i].bytes;
    }
    auto da = [&](int i) {
        std::string s = std::string("<DataArray type=\"") + blk[i].type + "\"";
        if (blk[i].name) s += std::string(" Name=\"") + blk[i].name + "\"";
        if (blk[i].nc > 1) s += " NumberOfComponents=\"" + std::to_string(blk[i].nc) + "\"";
        return s + " format=\"appended\" offset=\"" + std::to_string(off[i]) + "\"/>";
    };
    std::string folder = outdir + "/Fluid_Variables";
    mkdir(outdir.c_str(), 0777);
    mkdir(folder.c_str(), 0777);
    std::ofstream out(folder + "/" + std::to_string(idx) + ".vtu", std::ios::binary);
    out << "<?xml version=\"1.0\"?>\n"
        << "<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n"
        << "  <UnstructuredGrid>\n"
        << "    <Piece NumberOfPoints=\"" << den.size() << "\" NumberOfCells=\"" << m << "\">\n"
        << "      <PointData>\n        " << da(0) << "\n        " << da(1) << "\n        " << da(2)
        << "\n      </PointData>\n      <Points>\n        " << da(3) << "\n      </Points>\n      <Cells>\n        "
        << da(4) << "\n        " << da(5) << "\n        " << da(6) << "\n      </Cells>\n    </Piece>\n"
        << "  </UnstructuredGrid>\n  <AppendedData encoding=\"raw\">\n   _";
    for (int i = 0; i < 7; i++) {
        const uint64_t nb = blk[i].bytes;
        out.write(reinterpret_cast<const char*>(&nb), 8);
        out.write(static_cast<const char*>(blk[i].data), (std::streamsize)nb);
    }
    out << "\n  </AppendedData>\n</VTKFile>\n";
}

int usage() {
    std::cerr << "usage: lbm_run --stl FILE [--spacing H] [--grid ox,oy,oz,nx,ny,nz] [--lattice 19|27]\n"
                 "               [--storage fp64|fp16|bf16] [--tau T] [--dt DT] [--faces pv|x]\n"
                 "               [--face role,closure,axis,plane,tol,lo0,lo1,hi0,hi1,rho]...\n"
                 "               [--t-end T] [--period P] [--steps N --inlet const:V|sin:M,A,P]\n"
                 "               [--output-every K] [--outdir DIR] [--vtu] [--wall reference|interpolated]\n"
                 "               [--streaming auto|two-lattices|in-place]\n";
    return 2;
}

}  // namespace

int main(int argc, char** argv) {
    std::string stl, storage = "fp64", faces_kind, inlet = "const:0.05", outdir = ".";
    double spacing = 0.05, tau = 1.5, dt = 8e-6, t_end = 2.7, period = 0.9;
    int lattice = 19, output_every = 100;
    bool vtu = false;
    std::string wall = "reference";
    std::string streaming = "auto";
    long steps = -1;
    std::vector<double> grid;
    std::vector<pvlbm_face> faces;
    for (int a = 1; a < argc; a++) {
        std::string k = argv[a];
        auto val = [&]() -> std::string {
            if (a + 1 >= argc) throw std::runtime_error("missing value for " + k);
            return argv[++a];
        };
        try {
            if (k == "--stl") stl = val();
            else if (k == "--spacing") spacing = std::stod(val());
            else if (k == "--grid") grid = parse_list(val());
            else if (k == "--lattice") lattice = std::stoi(val());
            else if (k == "--storage") storage = val();
            else if (k == "--tau") tau = std::stod(val());
            else if (k == "--dt") dt = std::stod(val());
            else if (k == "--faces") faces_kind = val();
            else if (k == "--face") {
                std::vector<double> f = parse_list(val());
                if (f.size() != 10) return usage();
                faces.push_back(make_face((int)f[0], (int)f[1], (int)f[2], f[3], f[4], f[5], f[6], f[7], f[8], f[9]));
            } else if (k == "--t-end") t_end = std::stod(val());
            else if (k == "--period") period = std::stod(val());
            else if (k == "--steps") steps = std::stol(val());
            else if (k == "--inlet") inlet = val();
            else if (k == "--output-every") output_every = std::stoi(val());
            else if (k == "--outdir") outdir = val();
            else if (k == "--vtu") vtu = true;
            else if (k == "--wall") wall = val();
            else if (k == "--streaming") streaming = val();
            else if (k == "--help" || k == "-h") return usage(), 0;
            else return usage();
        } catch (const std::exception& e) {
            std::cerr << e.what() << "\n";
            return usage();
        }
    }
    if (stl.empty()) return usage();
    const int st = storage == "fp16" ? PVLBM_FP16S : storage == "bf16" ? PVLBM_BF16S : PVLBM_FP64;
    try {
        auto t0 = std::chrono::steady_clock::now();
        pvlbm::LatticeModel velSet(lattice);
        pvlbm::Geometry* geo = nullptr;
        if (grid.empty()) {
            geo = new pvlbm::Geometry(stl, velSet, spacing, 'z');
        } else {
            if (grid.size() != 6) return usage();
            pvlbm::Geometry tmp(stl, velSet, spacing, 'z');
            pvlbm_grid g{};
            for (int i = 0; i < 3; i++) {
                g.origin[i] = grid[i];
                g.dims[i] = (int32_t)grid[3 + i];
            }
            geo = new pvlbm::Geometry(tmp.triangles, velSet, spacing, &g);
        }
        const double big = 1e300;
        if (faces_kind == "pv") {
            // the reference's PV.stl planes (ComplexGeometry.cpp:1378-1408)
            faces.push_back(make_face(PVLBM_INLET, PVLBM_INLET_XMIN, 0, -16.9538L, 1e-2, -191.0, -118.0, -187.0, -114.0, 1.0));
            faces.push_back(make_face(PVLBM_OUTLET1, PVLBM_OUTLET_YMAX, 1, -170.016L, 1e-2, 2.0, -109.0, 5.0, -106.0, 1.0));
            faces.push_back(make_face(PVLBM_OUTLET2, PVLBM_OUTLET_ZMIN, 2, -110.0L, 1e-2, -27.0, -198.0, -25.0, -195.0, 1.0));
            faces.push_back(make_face(PVLBM_OUTLET3, PVLBM_OUTLET_ZMIN, 2, -123.0L, 1e-2, -33.0, -189.0, -30.0, -186.0, 1.0));
        } else if (faces_kind == "x") {
            const double x0 = geo->grid.origin[0], x1 = x0 + (geo->grid.dims[0] - 1) * spacing;
            faces.push_back(make_face(PVLBM_INLET, PVLBM_INLET_XMIN, 0, x0, 1e-2 * spacing, -big, -big, big, big, 1.0));
            faces.push_back(make_face(PVLBM_OUTLET1, PVLBM_OUTLET_XMAX, 0, x1, 1e-2 * spacing, -big, -big, big, big, 1.0));
        }
        pvlbm::SubDomain sd(*geo, faces, st, tau, dt);
        pvlbm::Simulation sim(sd, tau, 1.0, dt);
        if (wall == "interpolated") sim.setWallClosure(PVLBM_WALL_INTERPOLATED);
        else if (wall != "reference") throw std::runtime_error("--wall takes reference or interpolated");
        if (streaming == "two-lattices") sim.setStreaming(PVLBM_STREAMING_TWO_LATTICES);
        else if (streaming == "in-place") sim.setStreaming(PVLBM_STREAMING_IN_PLACE);
        else if (streaming != "auto") throw std::runtime_error("--streaming takes auto, two-lattices or in-place");
        std::cerr << "grid " << geo->grid.dims[0] << "x" << geo->grid.dims[1] << "x" << geo->grid.dims[2]
                  << ", nodes " << sd.count(PVLBM_LIST_FLUID) << ", inlet " << sd.count(PVLBM_LIST_INLET)
                  << ", wall " << sd.count(PVLBM_LIST_WALL) << ", next-to-wall " << sd.count(PVLBM_LIST_NTW) << "\n";
        std::vector<uint8_t> kinds;
        if (vtu) kinds = sd.kinds();
        // WriteOutput (Simulation.cpp:872-876): both files are named by the iteration
        auto out = [&](int64_t it) {
            write_csv(outdir, it, sim.surfaceAttributes());
            if (vtu)
                write_vtu(outdir, it, kinds, geo->grid.dims[0], geo->grid.dims[1], geo->grid.dims[2], sim.outputRho(),
                          sim.outputU());
            std::cerr << "iteration " << it << " written\n";
        };
        if (steps < 0) {
            sim.Iterate(t_end, period, output_every, out, vtu);
        } else {
            std::vector<double> sc((size_t)steps);
            if (inlet.rfind("const:", 0) == 0) {
                double v = std::stod(inlet.substr(6));
                for (auto& s : sc) s = v;
            } else if (inlet.rfind("sin:", 0) == 0) {
                std::vector<double> p = parse_list(inlet.substr(4));
                if (p.size() != 3) return usage();
                for (long s = 0; s < steps; s++) sc[(size_t)s] = p[0] + p[1] * std::sin(2.0 * M_PI * (double)s / p[2]);
            } else {
                return usage();
            }
            sim.Run(sc, output_every, out, vtu);
        }
        delete geo;
        double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::cerr << "Total Elapsed time: " << el << " seconds.\n";
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
