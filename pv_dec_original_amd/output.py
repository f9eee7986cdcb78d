"""On-disk outputs in the reference's formats.

FlowHemodynamics CSV (OutputFile::WriteSurfaceAttribute, OutputFile.cpp:867-938): header
``X,Y,Z,WSS,OSI,TAWSS`` and one row per next-to-wall node, each value written by a default
``std::ostream`` (6 significant digits, ``%g`` style).  The reference's rows follow an
unordered_set (Subdomain.cpp:202-217), so row order carries no meaning; ours are in box-index
order.

Fluid-variable VTU (OutputFile::WriteVTK_Collective, OutputFile.cpp:259-429): an XML unstructured
grid whose points are the fluid nodes at their integer lattice indices (x, y, z), in the order of
the reference's z, y, x loops; point data Velocity (3 components), Density and
VelocityMagnitude; one hexahedron for every lattice cube whose eight corners are all fluid nodes,
corners in the order (x,y,z), (x+1,y,z), (x+1,y+1,z), (x,y+1,z) and the same at z+1.  The
reference writes through VTK's XML writer (encoding chosen by the VTK version); this file carries
the same dataset as raw appended binary, which VTK's reader and ParaView read alike.
"""
import numpy as np


def _g(v):
    return "%g" % v


def write_surface_csv(path, rows):
    """rows: [n, 6] array of x, y, z, WSS, OSI, TAWSS."""
    with open(path, "w") as fp:
        fp.write("X,Y,Z,WSS,OSI,TAWSS\n")
        for r in rows:
            fp.write(",".join(_g(float(v)) for v in r) + "\n")


def read_surface_csv(path):
    with open(path) as fp:
        header = fp.readline().strip()
        if header != "X,Y,Z,WSS,OSI,TAWSS":
            raise ValueError(f"unexpected header {header!r}")
        rows = [[float(v) for v in line.split(",")] for line in fp if line.strip()]
    return np.array(rows, dtype=np.float64).reshape(-1, 6)


VTK_HEXAHEDRON = 12


def fluid_grid(fluid):
    """Points (integer lattice indices, z-y-x loop order) and hexahedral cells of a fluid mask
    fluid[z, y, x]: returns (points [n, 3] float64, connectivity [m, 8] int64)."""
    fluid = np.asarray(fluid, dtype=bool)
    nz, ny, nx = fluid.shape
    idx = np.full(fluid.shape, -1, dtype=np.int64)
    z, y, x = np.nonzero(fluid)
    idx[z, y, x] = np.arange(len(z))
    points = np.stack([x, y, z], axis=1).astype(np.float64)
    f = fluid
    cube = (f[:-1, :-1, :-1] & f[:-1, :-1, 1:] & f[:-1, 1:, 1:] & f[:-1, 1:, :-1] &
            f[1:, :-1, :-1] & f[1:, :-1, 1:] & f[1:, 1:, 1:] & f[1:, 1:, :-1])
    cz, cy, cx = np.nonzero(cube)
    corners = [(0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0), (1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0)]
    conn = np.stack([idx[cz + dz, cy + dy, cx + dx] for dz, dy, dx in corners], axis=1)
    return points, conn


def write_fluid_vtu(path, fluid, rho, u):
    """The fluid-variable VTU of one output step.  fluid: mask [nz, ny, nx]; rho: [nz, ny, nx];
    u: [3, nz, ny, nx]."""
    fluid = np.asarray(fluid, dtype=bool)
    points, conn = fluid_grid(fluid)
    vel = np.stack([np.asarray(u[c])[fluid] for c in range(3)], axis=1).astype(np.float64)
    den = np.asarray(rho)[fluid].astype(np.float64)
    mag = np.sqrt(vel[:, 0] * vel[:, 0] + vel[:, 1] * vel[:, 1] + vel[:, 2] * vel[:, 2])
    m = len(conn)
    blocks = [("Velocity", "Float64", 3, vel), ("Density", "Float64", 1, den),
              ("VelocityMagnitude", "Float64", 1, mag), (None, "Float64", 3, points),
              ("connectivity", "Int64", 1, conn.astype(np.int64)),
              ("offsets", "Int64", 1, 8 * np.arange(1, m + 1, dtype=np.int64)),
              ("types", "UInt8", 1, np.full(m, VTK_HEXAHEDRON, dtype=np.uint8))]
    offs, off = [], 0
    for _, _, _, a in blocks:
        offs.append(off)
        off += 8 + np.ascontiguousarray(a).nbytes

    def da(i):
        name, ty, nc, _ = blocks[i]
        nm = f' Name="{name}"' if name else ""
        comp = f' NumberOfComponents="{nc}"' if nc > 1 else ""
        return f'<DataArray type="{ty}"{nm}{comp} format="appended" offset="{offs[i]}"/>'

    head = "\n".join([
        '<?xml version="1.0"?>',
        '<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" header_type="UInt64">',
        "  <UnstructuredGrid>",
        f'    <Piece NumberOfPoints="{len(points)}" NumberOfCells="{m}">',
        "      <PointData>", "        " + da(0), "        " + da(1), "        " + da(2), "      </PointData>",
        "      <Points>", "        " + da(3), "      </Points>",
        "      <Cells>", "        " + da(4), "        " + da(5), "        " + da(6), "      </Cells>",
        "    </Piece>", "  </UnstructuredGrid>", '  <AppendedData encoding="raw">', "   _"])
    with open(path, "wb") as fp:
        fp.write(head.encode())
        for _, _, _, a in blocks:
            b = np.ascontiguousarray(a).astype(a.dtype.newbyteorder("<"), copy=False).tobytes()
            fp.write(np.uint64(len(b)).tobytes())
            fp.write(b)
        fp.write(b"\n  </AppendedData>\n</VTKFile>\n")


def read_fluid_vtu(path):
    """Reads a VTU written with raw appended data (this module's or lbm_run's): returns a dict of
    the arrays by name ("Points" for the coordinates) reshaped by their component counts."""
    import re
    raw = open(path, "rb").read()
    start = raw.index(b"<AppendedData")
    head = raw[:start].decode()
    base = raw.index(b"_", start) + 1
    types = {"Float64": np.float64, "Int64": np.int64, "UInt8": np.uint8}
    out = {}
    in_points = False
    for line in head.splitlines():
        if "<Points>" in line:
            in_points = True
        if "</Points>" in line:
            in_points = False
        mt = re.search(r'<DataArray type="(\w+)"(?: Name="(\w+)")?(?: NumberOfComponents="(\d+)")? '
                       r'format="appended" offset="(\d+)"', line)
        if not mt:
            continue
        ty, name, nc, off = mt.group(1), mt.group(2), int(mt.group(3) or 1), int(mt.group(4))
        p = base + off
        nbytes = int(np.frombuffer(raw[p:p + 8], dtype="<u8")[0])
        a = np.frombuffer(raw[p + 8:p + 8 + nbytes], dtype=np.dtype(types[ty]).newbyteorder("<"))
        out["Points" if in_points else name] = a.reshape(-1, nc) if nc > 1 else a
    mt = re.search(r'NumberOfPoints="(\d+)" NumberOfCells="(\d+)"', head)
    out["n_points"], out["n_cells"] = int(mt.group(1)), int(mt.group(2))
    return out
