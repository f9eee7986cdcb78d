"""On-disk outputs in the reference's formats.

FlowHemodynamics CSV (OutputFile::WriteSurfaceAttribute, OutputFile.cpp:867-938): header
``X,Y,Z,WSS,OSI,TAWSS`` and one row per next-to-wall node, each value written by a default
``std::ostream`` (6 significant digits, ``%g`` style).  The reference's rows follow an
unordered_set (Subdomain.cpp:202-217), so row order carries no meaning; ours are in box-index
order.
"""


def _g(v):
    return "%g" % v


def write_surface_csv(path, rows):
    """rows: [n, 6] array of x, y, z, WSS, OSI, TAWSS."""
    with open(path, "w") as fp:
        fp.write("X,Y,Z,WSS,OSI,TAWSS\n")
        for r in rows:
            fp.write(",".join(_g(float(v)) for v in r) + "\n")


def read_surface_csv(path):
    import numpy as np
    with open(path) as fp:
        header = fp.readline().strip()
        if header != "X,Y,Z,WSS,OSI,TAWSS":
            raise ValueError(f"unexpected header {header!r}")
        rows = [[float(v) for v in line.split(",")] for line in fp if line.strip()]
    return np.array(rows, dtype=np.float64).reshape(-1, 6)
