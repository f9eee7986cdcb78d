"""The lbm_run driver (C++ facade over the C ABI): CPU usage check and a GPU run against the oracle."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "pv_dec_original_amd", "lbm_run")


def test_cli_usage():
    r = subprocess.run([EXE, "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and "--stl" in r.stderr
    r = subprocess.run([EXE, "--bogus"], capture_output=True, text=True)
    assert r.returncode == 2


@pytest.mark.gpu
def test_cli_tube_matches_oracle(tmp_path):
    from pv_dec_original_amd import _native as N
    from pv_dec_original_amd.geometry import write_stl_ascii, voxelize
    from pv_dec_original_amd.output import read_surface_csv
    from pv_dec_original_amd.simulation import face
    from pv_dec_original_amd.workloads import tube_mesh
    from tests import oracle_ctypes as O

    tri = tube_mesh()
    stl = tmp_path / "tube.stl"
    write_stl_ascii(stl, tri)
    r = subprocess.run([EXE, "--stl", str(stl), "--spacing", "1", "--lattice", "19", "--tau", "0.8",
                        "--faces", "x", "--steps", "200", "--inlet", "const:0.05", "--output-every", "100",
                        "--outdir", str(tmp_path), "--vtu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    # iterations 0..199: WriteOutput after iterations 0 and 100 (Simulation.cpp:589-591)
    assert sorted(os.listdir(tmp_path / "FlowHemodynamics")) == ["0_FlowHemodynamics.csv",
                                                                 "100_FlowHemodynamics.csv"]
    assert sorted(os.listdir(tmp_path / "Fluid_Variables")) == ["0.vtu", "100.vtu"]
    got = read_surface_csv(tmp_path / "FlowHemodynamics" / "100_FlowHemodynamics.csv")
    fluid, origin, dims = voxelize(tri, 1.0)
    x1 = origin[0] + dims[0] - 1
    faces = [face(N.INLET, N.INLET_XMIN, 0, origin[0]), face(N.OUTLET1, N.OUTLET_XMAX, 0, x1)]
    orc = O.OracleSim(fluid, origin, 1.0, 19, faces, tri, 0.8, 8e-6)
    orc.step(100, np.full(100, 0.05))
    rho_o, u_o = orc.macros()  # iteration 100's collision state, what Fluid_Variables/100.vtu holds
    orc.step(1, np.full(1, 0.05))
    ref = orc.wss()[0]
    assert got.shape == ref.shape
    # rows carry 6 significant digits (std::ostream defaults, OutputFile.cpp:927-928)
    assert np.allclose(got[:, :3], ref[:, :3], rtol=1e-5, atol=1e-5)
    assert np.allclose(got[:, 3], ref[:, 3], rtol=1e-5, atol=0)
    assert np.allclose(got[:, 5], ref[:, 5], rtol=1e-5, atol=0)
    # the field output of iteration 100 against the Python writer on the oracle's fields
    from pv_dec_original_amd.output import read_fluid_vtu, write_fluid_vtu
    d = read_fluid_vtu(tmp_path / "Fluid_Variables" / "100.vtu")
    write_fluid_vtu(tmp_path / "py.vtu", fluid.reshape(rho_o.shape) > 0, rho_o, u_o)
    e = read_fluid_vtu(tmp_path / "py.vtu")
    for k in ("Points", "connectivity", "offsets", "types"):
        assert np.array_equal(d[k], e[k]), k
    for k in ("Velocity", "Density", "VelocityMagnitude"):
        assert np.abs(d[k] - e[k]).max() <= 1e-12, k


@pytest.mark.gpu
def test_cli_binary_stl_and_interpolated_wall(tmp_path):
    """lbm_run reads a binary STL (float32 vertices, GeometryReader.cpp:119-176) and runs the
    interpolated wall closure; the surface CSV has one row per next-to-wall node of the binary
    mesh's voxelisation, and the closure changes the WSS against the reference wall rule."""
    from pv_dec_original_amd.geometry import read_stl, voxelize, write_stl_binary
    from pv_dec_original_amd.output import read_surface_csv
    from pv_dec_original_amd.workloads import tube_mesh

    stl = tmp_path / "tube_bin.stl"
    write_stl_binary(stl, tube_mesh())
    out = {}
    for wall in ("reference", "interpolated"):
        d = tmp_path / wall
        r = subprocess.run([EXE, "--stl", str(stl), "--spacing", "1", "--lattice", "19", "--tau", "0.8",
                            "--faces", "x", "--steps", "301", "--inlet", "const:0.05", "--output-every", "300",
                            "--outdir", str(d), "--wall", wall], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        out[wall] = read_surface_csv(d / "FlowHemodynamics" / "300_FlowHemodynamics.csv")
    fluid, origin, dims = voxelize(read_stl(stl), 1.0)
    assert out["reference"].shape == out["interpolated"].shape and len(out["reference"]) > 100
    assert np.array_equal(out["reference"][:, :3], out["interpolated"][:, :3])
    assert np.isfinite(out["interpolated"]).all()
    rel = np.linalg.norm(out["interpolated"][:, 3] - out["reference"][:, 3]) / np.linalg.norm(out["reference"][:, 3])
    assert rel > 1e-3
    r = subprocess.run([EXE, "--stl", str(stl), "--wall", "other"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0


def test_fluid_vtu_round_trip_and_cells(tmp_path):
    """The fluid-variable VTU: points in z-y-x order at integer indices, hexahedra exactly for the
    cubes whose eight corners are fluid (counted by brute force), arrays read back bit for bit."""
    from pv_dec_original_amd.output import read_fluid_vtu, write_fluid_vtu
    rng = np.random.default_rng(2)
    fluid = rng.random((5, 6, 7)) < 0.8
    rho = 1.0 + 0.01 * rng.standard_normal(fluid.shape)
    u = 0.05 * rng.standard_normal((3,) + fluid.shape)
    path = tmp_path / "0.vtu"
    write_fluid_vtu(path, fluid, rho, u)
    d = read_fluid_vtu(path)
    z, y, x = np.nonzero(fluid)
    assert d["n_points"] == fluid.sum()
    assert np.array_equal(d["Points"], np.stack([x, y, z], 1).astype(float))
    assert np.array_equal(d["Density"], rho[fluid])
    assert np.array_equal(d["Velocity"], np.stack([u[c][fluid] for c in range(3)], 1))
    assert np.allclose(d["VelocityMagnitude"], np.sqrt((u[:, fluid] ** 2).sum(axis=0)), rtol=0, atol=0)
    cubes = []
    for k in range(4):
        for j in range(5):
            for i in range(6):
                corners = [(i, j, k), (i + 1, j, k), (i + 1, j + 1, k), (i, j + 1, k),
                           (i, j, k + 1), (i + 1, j, k + 1), (i + 1, j + 1, k + 1), (i, j + 1, k + 1)]
                if all(fluid[c[2], c[1], c[0]] for c in corners):
                    cubes.append(corners)
    assert d["n_cells"] == len(cubes)
    pts = d["Points"].astype(int)
    conn = d["connectivity"].reshape(-1, 8)
    for cell, corners in zip(conn, cubes):
        assert [tuple(pts[p]) for p in cell] == corners
    assert np.array_equal(d["offsets"], 8 * np.arange(1, len(cubes) + 1)) and (d["types"] == 12).all()


def test_lbm_run_loop_against_stub(tmp_path):
    """lbm_run's loop on the CPU, linked against tests/stub/pvlbm_stub.cpp, a stand-in for the
    library that logs the calls it receives: the reference loop (no --steps) runs iterations
    0, 1, ... with InletVelocity_ = getInletVelocity(total_runtime) for total_runtime = 0, dt,
    2 dt, ... while <= t_end (Simulation.cpp:569-601); after iterations 0, 100, 200 it writes
    <iter>_FlowHemodynamics.csv from the WSS read after that iteration and <iter>.vtu from the
    fields read before it (the state its Collision saw); --faces pv passes the reference's long
    double planes exactly."""
    import shutil
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    inc = os.path.join(ROOT, "include")
    stub = os.path.join(ROOT, "tests", "stub", "pvlbm_stub.cpp")
    lib = tmp_path / "libpvlbm.so"
    exe = tmp_path / "lbm_run_stub"
    subprocess.run(["g++", "-O1", "-std=c++17", "-shared", "-fPIC", "-I", inc, stub, "-o", str(lib)], check=True)
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", inc, os.path.join(ROOT, "pv_dec_original_amd", "csrc", "lbm_run.cpp"),
                    "-L", str(tmp_path), "-lpvlbm", "-Wl,-rpath," + str(tmp_path), "-o", str(exe)], check=True)
    stl = tmp_path / "x.stl"
    stl.write_text("solid x\nendsolid x\n")
    dt, t_end = 8e-6, 200 * 8e-6
    r = subprocess.run([str(exe), "--stl", str(stl), "--dt", repr(dt), "--t-end", repr(t_end), "--faces", "pv",
                        "--output-every", "100", "--outdir", str(tmp_path / "o"), "--vtu"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    lines = r.stderr.splitlines()
    times = []
    t = 0.0
    while t <= t_end:
        times.append(t)
        t += dt
    its = [ln.split() for ln in lines if ln.startswith("iter ")]
    assert [int(v[1]) for v in its] == list(range(len(times)))
    assert [float(v[3]) for v in its] == times  # the stub's waveform returns t itself
    events = [ln for ln in lines if ln.startswith(("macros at", "wss at"))]
    outs = list(range(0, len(times), 100))
    assert events == sum([[f"macros at {k}", f"wss at {k + 1}"] for k in outs], [])
    assert sorted(os.listdir(tmp_path / "o" / "FlowHemodynamics")) == sorted(f"{k}_FlowHemodynamics.csv" for k in outs)
    assert sorted(os.listdir(tmp_path / "o" / "Fluid_Variables")) == sorted(f"{k}.vtu" for k in outs)
    from pv_dec_original_amd.output import read_fluid_vtu, read_surface_csv
    for k in outs:
        # the stub reports the completed-iteration count as the state
        assert read_surface_csv(tmp_path / "o" / "FlowHemodynamics" / f"{k}_FlowHemodynamics.csv")[0, 3] == k + 1
        assert np.all(read_fluid_vtu(tmp_path / "o" / "Fluid_Variables" / f"{k}.vtu")["Density"] == k)
    faces = [ln.split() for ln in lines if ln.startswith("face ")]
    # 21 significant digits round-trip an x87 long double
    assert [np.longdouble(f[4]) for f in faces] == [np.longdouble(v) for v in ("-16.9538", "-170.016", "-110", "-123")]
    assert np.longdouble(faces[0][4]) != np.longdouble(-16.9538)  # not the double nearest to it
    # --steps N: the same cadence over N iterations with the given inlet
    r = subprocess.run([str(exe), "--stl", str(stl), "--faces", "x", "--steps", "201", "--inlet", "const:0.5",
                        "--output-every", "100", "--outdir", str(tmp_path / "s")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    assert sorted(os.listdir(tmp_path / "s" / "FlowHemodynamics")) == ["0_FlowHemodynamics.csv",
                                                                      "100_FlowHemodynamics.csv",
                                                                      "200_FlowHemodynamics.csv"]
