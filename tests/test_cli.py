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
    got = read_surface_csv(tmp_path / "FlowHemodynamics" / "199_FlowHemodynamics.csv")
    assert os.path.exists(tmp_path / "FlowHemodynamics" / "99_FlowHemodynamics.csv")
    fluid, origin, dims = voxelize(tri, 1.0)
    x1 = origin[0] + dims[0] - 1
    faces = [face(N.INLET, N.INLET_XMIN, 0, origin[0]), face(N.OUTLET1, N.OUTLET_XMAX, 0, x1)]
    orc = O.OracleSim(fluid, origin, 1.0, 19, faces, tri, 0.8, 8e-6)
    orc.step(200, np.full(200, 0.05))
    ref = orc.wss()[0]
    assert got.shape == ref.shape
    # rows carry 6 significant digits (std::ostream defaults, OutputFile.cpp:927-928)
    assert np.allclose(got[:, :3], ref[:, :3], rtol=1e-5, atol=1e-5)
    assert np.allclose(got[:, 3], ref[:, 3], rtol=1e-5, atol=0)
    assert np.allclose(got[:, 5], ref[:, 5], rtol=1e-5, atol=0)
    # the second field output (after 200 steps) against the Python writer on the oracle's fields
    from pv_dec_original_amd.output import read_fluid_vtu, write_fluid_vtu
    d = read_fluid_vtu(tmp_path / "Fluid_Variables" / "1.vtu")
    rho_o, u_o = orc.macros()
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
                            "--faces", "x", "--steps", "300", "--inlet", "const:0.05", "--output-every", "300",
                            "--outdir", str(d), "--wall", wall], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        out[wall] = read_surface_csv(d / "FlowHemodynamics" / "299_FlowHemodynamics.csv")
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
