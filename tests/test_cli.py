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
                        "--outdir", str(tmp_path)], capture_output=True, text=True, timeout=120)
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
