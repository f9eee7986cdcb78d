"""CPU check of the slot addressing of the step's streaming schemes (test infrastructure).

The kernel's offset helpers (pv_dec_original_amd/csrc/k_step.hip: load_off, store_off, park_off;
kernels.hpp: enum Scheme) restated on a flat array with the same plane-and-guard layout, on a
random sparse lattice where most nodes miss links:
  - in each kind of in-place launch (even, odd) every slot a node reads, writes or parks into is
    addressed by that node alone, so waves need no ordering (a device run cannot show this: a
    race may go unseen);
  - the pull from one lattice into another, and the in-place scheme from either start parity, give
    the push streaming of the reference (Simulation.cpp:211-237; a link with no upstream node keeps
    its previous post-closure value, :218-224) value for value, with a stand-in collision and an
    order-dependent stand-in closure;
  - moving the state between the layouts (k_relayout) in the middle of a run changes nothing.
The stand-ins are arbitrary functions of a node's populations: what is checked is where values go.
"""
import numpy as np
import pytest

from pv_dec_original_amd.lattice import LatticeModel

PULL, EVEN, ODD = 0, 1, 2


def _lattice(q):
    lm = LatticeModel(q)  # the kernel's direction order (lattice.py mirrors lbm_device.hpp)
    c = np.asarray(lm.c, dtype=int).reshape(q, 3)
    opp = np.asarray(lm.opp, dtype=int)
    assert all((c[opp[i]] == -c[i]).all() for i in range(q))
    return c, opp


class Layout:
    def __init__(self, q, shape, density, seed):
        rng = np.random.default_rng(seed)
        nz, ny, nx = shape
        node = rng.random(shape) < density
        node[2:-2, 2:-2, 2:-2] = True  # a core of interior nodes inside the random shell
        node[0] = node[-1] = False
        node[:, 0] = node[:, -1] = False
        node[:, :, 0] = node[:, :, -1] = False
        self.q, self.c, self.opp = q, *_lattice(q)
        self.n = nx * ny * nz
        self.nodef = node.ravel()
        self.off = self.c[:, 0] + self.c[:, 1] * nx + self.c[:, 2] * nx * ny
        self.idx = np.nonzero(self.nodef)[0]
        self.guard = nx * ny + nx + 1
        self.ns = self.n + 2 * self.guard
        # missing links: d with x - c_d not a node (the kernel tests bit opp d of the word whose
        # bit k says x + c_k is not a node)
        self.miss = {x: [d for d in range(q) if not self.nodef[x - self.off[d]]] for x in self.idx}

    def slot(self, d, pos):
        return d * self.ns + self.guard + pos

    # k_step.hip load_off / store_off / park_off, per kind of launch
    def load(self, k, x, d):
        if k == PULL:
            return self.slot(d, x - self.off[d])
        if k == EVEN:
            return self.slot(d, x)
        return self.slot(self.opp[d], x - self.off[d])

    def store(self, k, x, d):
        if k == PULL:
            return self.slot(d, x)
        if k == EVEN:
            return self.slot(self.opp[d], x)
        return self.slot(d, x + self.off[d])

    def park(self, k, x, d):
        if k == PULL:
            return self.slot(d, x - self.off[d])
        return self.load(ODD if k == EVEN else EVEN, x, d)


def _closure(lay, x, f):
    f = f.copy()
    for d in lay.miss[x]:
        f[lay.opp[d]] = 0.7 * f[d] + 0.3 * f[lay.opp[d]] + 0.01 * d
    return f


def _collide(M, f):
    return np.tanh(M @ f) + 0.05 * f * f


@pytest.fixture(scope="module", params=[19, 27])
def lay(request):
    return Layout(request.param, (9, 11, 14), 0.8, seed=request.param)


def test_every_slot_has_one_owner_per_launch(lay):
    for k in (EVEN, ODD):
        owner = {}
        for x in lay.idx:
            addrs = {lay.load(k, x, d) for d in range(lay.q)} | {lay.store(k, x, d) for d in range(lay.q)}
            addrs |= {lay.park(k, x, d) for d in lay.miss[x]}
            for a in addrs:
                assert 0 <= a < lay.q * lay.ns
                assert owner.setdefault(a, x) == x, (k, a, owner[a], x)
        # a node's relaxed populations go back to the slots it read (so the update is in place);
        # its parked links go to slots it alone addresses (checked above)
        for x in lay.idx:
            reads = {lay.load(k, x, d) for d in range(lay.q)}
            assert reads == {lay.store(k, x, d) for d in range(lay.q)}, (k, x)


def _push_reference(lay, f0, M, steps):
    f, out = dict(f0), []
    for _ in range(steps):
        fpost = {x: _closure(lay, x, f[x]) for x in lay.idx}
        fstar = {x: _collide(M, fpost[x]) for x in lay.idx}
        out.append(fstar)
        nf = {}
        for x in lay.idx:
            v = fpost[x].copy()
            for i in range(lay.q):
                if lay.nodef[x - lay.off[i]]:
                    v[i] = fstar[x - lay.off[i]][i]
            nf[x] = v
        f = nf
    return out


def _run(lay, f0, M, plan, rng):
    """plan: per launch 'pull' or 'inplace'; the state moves between layouts when the plan
    changes, as the library does when the scheme changes."""
    L = np.full(lay.q * lay.ns, np.nan)
    for x in lay.idx:
        for i in range(lay.q):
            L[lay.load(PULL, x, i)] = f0[x][i]
    last, out = PULL, []
    for want in plan:
        if (want == "pull") != (last == PULL):
            to = PULL if want == "pull" else EVEN
            B = np.full_like(L, np.nan)
            for x in lay.idx:
                for d in range(lay.q):
                    B[lay.store(to, x, d)] = L[lay.store(last, x, d)]
                for d in lay.miss[x]:
                    B[lay.park(to, x, d)] = L[lay.park(last, x, d)]
            L, last = B, to
        k = PULL if want == "pull" else (ODD if last == EVEN else EVEN)
        dst = np.full_like(L, np.nan) if k == PULL else L
        fs = {}
        for x in rng.permutation(lay.idx):
            v = _closure(lay, x, np.array([L[lay.load(k, x, i)] for i in range(lay.q)]))
            for d in lay.miss[x]:
                dst[lay.park(k, x, d)] = v[d]
            s = _collide(M, v)
            fs[x] = s
            for d in range(lay.q):
                dst[lay.store(k, x, d)] = s[d]
        L, last = dst, k
        out.append(fs)
    return out


def test_schemes_reproduce_push_streaming(lay):
    rng = np.random.default_rng(3)
    M = rng.random((lay.q, lay.q)) * 0.1 + np.eye(lay.q)
    f0 = {x: rng.random(lay.q) for x in lay.idx}
    steps = 6
    ref = _push_reference(lay, f0, M, steps)
    for plan in (["pull"] * steps, ["inplace"] * steps, ["pull"] + ["inplace"] * (steps - 1),
                 ["inplace", "inplace", "pull", "inplace", "inplace", "pull"],
                 ["inplace", "pull", "pull", "inplace", "pull", "inplace"]):
        out = _run(lay, f0, M, plan, rng)
        for t in range(steps):
            for x in lay.idx:
                assert np.array_equal(ref[t][x], out[t][x]), (plan, t, x)
    assert sum(len(v) > 0 for v in lay.miss.values()) > 0.5 * len(lay.idx)
    assert sum(len(v) == 0 for v in lay.miss.values()) > 0
