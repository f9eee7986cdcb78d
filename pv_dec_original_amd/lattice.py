"""LatticeModel: the D3Q19 / D3Q27 velocity sets in the reference's index order.

D3Q19: LatticeModel.cpp:53-92 (rest direction 18), opposites :108-151, face unknown sets :96-101.
D3Q27: LatticeModel.cpp:210-266 (rest direction 0), opposites :274-334.
The index order is part of the interface: the boundary closures are written against it.
"""
import numpy as np

_C19 = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 0), (1, -1, 0),
        (1, 0, 1), (1, 0, -1), (-1, 1, 0), (-1, -1, 0), (-1, 0, 1), (-1, 0, -1), (0, 1, 1),
        (0, 1, -1), (0, -1, 1), (0, -1, -1), (0, 0, 0)]
_W19 = [2.0 / 36.0] * 6 + [1.0 / 36.0] * 12 + [12.0 / 36.0]
_O19 = [1, 0, 3, 2, 5, 4, 11, 10, 13, 12, 7, 6, 9, 8, 17, 16, 15, 14, 18]

_C27 = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 0),
        (1, -1, 0), (-1, 1, 0), (-1, -1, 0), (1, 0, 1), (-1, 0, 1), (1, 0, -1), (-1, 0, -1),
        (0, 1, 1), (0, 1, -1), (0, -1, 1), (0, -1, -1), (1, 1, 1), (-1, -1, -1), (1, 1, -1),
        (-1, -1, 1), (1, -1, 1), (-1, 1, -1), (1, -1, -1), (-1, 1, 1)]
_W27 = [8.0 / 27.0] + [2.0 / 27.0] * 6 + [1.0 / 54.0] * 12 + [1.0 / 216.0] * 8
_O27 = [0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15, 20, 19, 22, 21, 24, 23,
        26, 25]

# D3Q19 unknown populations of each face, labelled by the face's inward normal
# (LatticeModel.cpp:96-101 / Pair0fNormalUnkDis :154-202).
UNKNOWN19 = {(1, 0, 0): (0, 6, 7, 8, 9), (-1, 0, 0): (1, 10, 11, 12, 13),
             (0, -1, 0): (3, 7, 11, 16, 17), (0, 1, 0): (2, 6, 10, 14, 15),
             (0, 0, 1): (4, 8, 12, 14, 16), (0, 0, -1): (5, 9, 13, 15, 17)}


class LatticeModel:
    """Velocity set with weights, directions and opposites (``LatticeModel()``, LatticeModel.h:9;
    the reference selects the lattice at compile time, Macro.h:6; here it is an argument)."""

    def __init__(self, q=19):
        if q == 19:
            c, w, o = _C19, _W19, _O19
        elif q == 27:
            c, w, o = _C27, _W27, _O27
        else:
            raise ValueError("lattice must be 19 or 27")
        self.q = q
        self.c = np.array(c, dtype=np.int64)
        self.w = np.array(w, dtype=np.float64)
        self.opp = np.array(o, dtype=np.int64)
        self.cs2 = 1.0 / 3.0

    def GetWeight(self, d):
        return float(self.w[d])

    def GetDirection(self, d):
        return [int(v) for v in self.c[d]]

    def OppositeDirection(self, d):
        return int(self.opp[d])

    def Get_nDirections(self):
        return self.q

    def Get_nDimensions(self):
        return 3

    def Get_c_square(self, d=0):
        return self.cs2

    @property
    def rest(self):
        return 18 if self.q == 19 else 0
