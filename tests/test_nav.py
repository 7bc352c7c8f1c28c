"""Volume-tree navigation on the host (no GPU): the oracle's navigator and the
library's tree validation.

The reference navigates with VecGeom (an un-vendored submodule; its calls are
geometry/NavigationInterface.hpp:42-98), so the navigator is pinned by what
the geometry fixes rather than by reference outputs:
  - on the reference's own calorimeter (50 layers of 10 mm Pb + 10 mm lAr
    along x, 1 m square, ProxyDetectorConstruction.cpp:150-164, 225-325) the
    touchable, boundary distance, next touchable and safety have closed forms;
  - on random nested trees: locate agrees with the deepest box holding the
    point; the boundary distance and next touchable agree with a brute-force
    scan of every face plane crossed by the ray; no point closer than the
    safety lies in another touchable;
  - surface points go to the box the direction enters; a ray walked boundary
    to boundary through the calorimeter visits every layer in order;
  - the library rejects exactly the trees the oracle rejects (daughters outside
    their mother, overlapping daughters, cycles), before any device call.
"""
import ctypes as C
import os

import numpy as np
import pytest

from geantexascalepilot_amd import abi
import oracle_harness as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "geantexascalepilot_amd", "libgxt.so")


def _unit(rs, n):
    d = rs.normal(size=(3, n))
    return d / np.linalg.norm(d, axis=0)


def calo_points(rs, n):
    """Points inside the reference calorimeter away from every face, random directions."""
    pos = np.stack([rs.uniform(-500, 500, n), rs.uniform(-500, 500, n), rs.uniform(-500, 500, n)])
    frac = np.mod(pos[0] + 500.0, 10.0)
    keep = (frac > 1e-6) & (frac < 10 - 1e-6) & (np.abs(pos[1]) < 500 - 1e-6) & (np.abs(pos[2]) < 500 - 1e-6)
    return pos[:, keep], _unit(rs, n)[:, keep]


def test_reference_calorimeter_closed_form():
    v, p, w = abi.reference_calorimeter_tree()
    nav = O.Navigator(v, p, w)
    info = nav.info()
    # world, calorimeter, 50 layers, 100 absorbers, breadth-first
    assert len(info["volume"]) == 152
    assert list(info["volume"][:3]) == [0, 2, 1] and (info["volume"][52:] == np.tile([3, 4], 50)).all()
    np.testing.assert_array_equal(info["centre"][2:52, 0], -490.0 + 20.0 * np.arange(50))
    rs = np.random.default_rng(1)
    pos, d = calo_points(rs, 20000)
    t = nav.locate(pos, d)
    slab = np.floor((pos[0] + 500.0) / 10.0).astype(int)  # 0..99: Pb, lAr, Pb, ...
    np.testing.assert_array_equal(t, 52 + slab)
    s, nxt, saf = nav.step(t, pos, d)
    lo = -500.0 + 10.0 * slab
    with np.errstate(divide="ignore"):
        sx = np.where(d[0] > 0, (lo + 10.0 - pos[0]) / d[0], (lo - pos[0]) / d[0])
        sy = np.where(d[1] > 0, (500.0 - pos[1]) / d[1], (-500.0 - pos[1]) / d[1])
        sz = np.where(d[2] > 0, (500.0 - pos[2]) / d[2], (-500.0 - pos[2]) / d[2])
    want = np.minimum(sx, np.minimum(sy, sz))
    np.testing.assert_allclose(s, want, rtol=1e-12, atol=1e-12)
    # across a layer face: the neighbouring slab (or out of the calorimeter into the world)
    via_x = sx < np.minimum(sy, sz)
    nslab = slab + np.where(d[0] > 0, 1, -1)
    want_next = np.where(via_x & (nslab >= 0) & (nslab < 100), 52 + nslab, 0)
    np.testing.assert_array_equal(nxt, want_next)
    want_saf = np.minimum.reduce([pos[0] - lo, lo + 10.0 - pos[0], 500 - np.abs(pos[1]), 500 - np.abs(pos[2])])
    np.testing.assert_allclose(saf, want_saf, rtol=1e-12, atol=1e-12)


def test_surface_points_follow_direction():
    v, p, w = abi.reference_calorimeter_tree()
    nav = O.Navigator(v, p, w)
    faces = -500.0 + 10.0 * np.arange(101)
    n = len(faces)
    pos = np.stack([faces, np.zeros(n), np.zeros(n)])
    fwd = np.stack([np.ones(n), np.zeros(n), np.zeros(n)])
    t_fwd = nav.locate(pos, fwd)
    t_bwd = nav.locate(pos, -fwd)
    np.testing.assert_array_equal(t_fwd[:-1], 52 + np.arange(100))
    np.testing.assert_array_equal(t_bwd[1:], 52 + np.arange(100))
    assert t_fwd[-1] == 0 and t_bwd[0] == 0  # leaving the calorimeter into the world
    # moving along a face enters neither box that shares it: the point on the
    # face between layers 24 and 25 moving along y stays in the calorimeter
    tan = nav.locate(np.array([[0.0], [0.0], [0.0]]), np.array([[0.0], [1.0], [0.0]]))
    assert tan[0] == 1
    # outside the world
    assert nav.locate(np.array([[700.0], [0.0], [0.0]]), np.array([[1.0], [0.0], [0.0]]))[0] == -1
    # on the world face moving out
    assert nav.locate(np.array([[600.0], [0.0], [0.0]]), np.array([[1.0], [0.0], [0.0]]))[0] == -1
    assert nav.locate(np.array([[600.0], [0.0], [0.0]]), np.array([[-1.0], [0.0], [0.0]]))[0] == 0


def test_walk_through_calorimeter():
    """A ray along +x from the world face crosses the 100 slabs in order, 10 mm each."""
    v, p, w = abi.reference_calorimeter_tree()
    nav = O.Navigator(v, p, w)
    pos = np.array([[-600.0], [3.0], [-7.0]])
    d = np.array([[1.0], [0.0], [0.0]])
    t = nav.locate(pos, d)
    seen, lengths = [], []
    while t[0] >= 0:
        s, nxt, _ = nav.step(t, pos, d)
        seen.append(int(t[0]))
        lengths.append(float(s[0]))
        pos = pos + s * d
        t = nxt
    assert seen == [0] + list(range(52, 152)) + [0]
    np.testing.assert_allclose(lengths, [100.0] + [10.0] * 100 + [100.0], rtol=1e-12)


def random_tree(rs, nvol=7, max_daughters=4):
    """Random nested boxes: volume i may hold smaller volumes j > i, placed
    without overlap by rejection."""
    half = [rs.uniform(40, 60, 3)]
    for _ in range(1, nvol):
        half.append(half[-1] * rs.uniform(0.3, 0.6, 3))
    vols = [(h, int(rs.integers(-1, 3))) for h in half]
    pls = []
    for m in range(nvol - 1):
        boxes = []
        for _ in range(int(rs.integers(1, max_daughters + 1))):
            dv = int(rs.integers(m + 1, min(m + 3, nvol)))
            for _try in range(50):
                lim = half[m] - half[dv]
                t = rs.uniform(-lim, lim) if rs.random() < 0.8 else np.sign(rs.uniform(-1, 1, 3)) * lim  # touch walls
                if all((np.abs(t - t2) >= half[dv] + h2).any() for t2, h2 in boxes):
                    boxes.append((t, half[dv]))
                    pls.append((m, dv, t))
                    break
    return vols, pls


def many_daughters_tree(rs, n=300, layers=False):
    """One mother holding up to n small boxes placed without overlap by
    rejection (layers: slabs stacked along y, each holding boxes), so that
    volumes have far more daughters than the device's index threshold."""
    world = np.array([100.0, 100.0, 100.0])
    vols = [(world, -1)]
    pls = []
    if layers:
        vols.append(((100.0, 4.0, 100.0), -1))
        for i in range(25):
            pls.append((0, 1, (0.0, -96.0 + 8.0 * i, 0.0)))
        mother, mhalf = 1, np.array([100.0, 4.0, 100.0])
    else:
        mother, mhalf = 0, world
    boxes = []
    for _ in range(n):
        h = rs.uniform(0.5, 6.0, 3)
        h = np.minimum(h, mhalf)
        lim = mhalf - h
        for _try in range(20):
            t = rs.uniform(-lim, lim)
            if rs.random() < 0.3:  # snap to a neighbour's face
                if boxes:
                    t2, h2 = boxes[int(rs.integers(len(boxes)))]
                    k = int(rs.integers(3))
                    t[k] = t2[k] + h2[k] + h[k]
                    if abs(t[k]) > lim[k]:
                        continue
            if all((np.abs(t - t2) >= h + h2).any() for t2, h2 in boxes):
                boxes.append((t, h))
                vols.append((h, 0))
                pls.append((mother, len(vols) - 1, t))
                break
    return vols, pls


class BruteTree:
    """Touchables as plain boxes: the deepest box holding a point, by testing
    every box."""

    def __init__(self, info, half):
        self.lo = info["centre"] - half
        self.hi = info["centre"] + half
        self.level = np.zeros(len(half), dtype=int)
        for t in range(1, len(half)):
            self.level[t] = self.level[info["parent"][t]] + 1

    def deepest(self, pts, eps=0.0):
        """pts: [3, m]; -1 where no box holds the point (strictly, by eps)."""
        inb = ((pts.T[:, None, :] > self.lo[None] + eps) & (pts.T[:, None, :] < self.hi[None] - eps)).all(axis=2)
        lvl = np.where(inb, self.level[None], -1)
        best = lvl.argmax(axis=1)
        return np.where(lvl.max(axis=1) >= 0, best, -1)


@pytest.mark.parametrize("seed,shape", [(3, "nested"), (4, "nested"), (5, "nested"), (6, "nested"),
                                        (7, "many"), (8, "layers")])
def test_random_tree_against_brute_force(seed, shape):
    rs = np.random.default_rng(seed)
    if shape == "nested":
        vols, pls = random_tree(rs)
    else:
        vols, pls = many_daughters_tree(rs, n=16 if shape == "layers" else 120, layers=shape == "layers")
    nav = O.Navigator(vols, pls, 0)
    info = nav.info()
    half = np.array([vols[v][0] for v in info["volume"]])
    bt = BruteTree(info, half)
    n = 300
    pos = rs.uniform(-1.1, 1.1, (3, n)) * half[0][:, None]
    d = _unit(rs, n)
    t = nav.locate(pos, d)
    np.testing.assert_array_equal(t, bt.deepest(pos))
    inside = np.nonzero(t >= 0)[0]
    s, nxt, saf = nav.step(t[inside], pos[:, inside], d[:, inside])
    for k, i in enumerate(inside):
        p, v = pos[:, i], d[:, i]
        # every face plane the ray meets; the first at which the deepest box changes
        with np.errstate(divide="ignore", invalid="ignore"):
            cand = np.concatenate([(bt.lo - p) / v, (bt.hi - p) / v]).ravel()
        cand = np.sort(cand[np.isfinite(cand) & (cand > 1e-7)])
        after = bt.deepest(p[:, None] + (cand + 1e-6)[None, :] * v[:, None])
        first = np.nonzero(after != t[i])[0][0]
        assert s[k] == pytest.approx(cand[first], abs=1e-9), (i, s[k], cand[first])
        assert nxt[k] == after[first], (i, nxt[k], after[first])
        # nothing within the safety lies elsewhere; the safety is attained
        ball = p[:, None] + 0.999 * saf[k] * _unit(rs, 64)
        assert (bt.deepest(ball, eps=-1e-9) == t[i]).all()
        dist = [np.min(np.minimum(p - bt.lo[t[i]], bt.hi[t[i]] - p))]
        for dd in np.nonzero(info["parent"] == t[i])[0]:
            dist.append(np.linalg.norm(p - np.clip(p, bt.lo[dd], bt.hi[dd])))
        assert saf[k] == pytest.approx(min(dist), rel=1e-12, abs=1e-12)


def _lib():
    lib = C.CDLL(LIB)
    lib.gxt_last_error.restype = C.c_char_p
    return lib


def _create_rc(lib, vols, pls, world=0):
    vol, pl = abi.box_tree(vols, pls)
    h = C.c_void_p()
    rc = lib.gxt_nav_create(vol, len(vols), pl, len(pls), world, C.byref(h))
    if rc == 0:
        lib.gxt_nav_free(h)
    return rc


def _oracle_rc(vols, pls, world=0):
    vol, pl = abi.box_tree(vols, pls)
    h = C.c_void_p()
    rc = O.lib().gxref_nav_create(vol, len(vols), pl, len(pls), world, C.byref(h))
    if rc == 0:
        O.lib().gxref_nav_free(h)
    return rc


BAD_TREES = {
    "outside": ([((10, 10, 10), -1), ((2, 2, 2), 0)], [(0, 1, (9.0, 0, 0))], b"not inside"),
    "overlap": ([((10, 10, 10), -1), ((2, 2, 2), 0)], [(0, 1, (0, 0, 0)), (0, 1, (3.9, 1.0, 0))], b"overlap"),
    "cycle": ([((10, 10, 10), -1), ((5, 5, 5), 0)], [(0, 1, (0, 0, 0)), (1, 1, (0, 0, 0))], b"cycle"),
    "half": ([((10, 0, 10), -1)], [], b"half"),
    "material": ([((10, 10, 10), 7)], [], b"material"),
    "index": ([((10, 10, 10), -1)], [(0, 2, (0, 0, 0))], b"range"),
}


@pytest.mark.parametrize("case", sorted(BAD_TREES))
def test_invalid_trees_rejected_without_device(case):
    vols, pls, needle = BAD_TREES[case]
    lib = _lib()
    assert _create_rc(lib, vols, pls) == -1
    assert needle in lib.gxt_last_error()
    assert _oracle_rc(vols, pls) == -1


def test_touching_daughters_accepted():
    vols = [((10, 10, 10), -1), ((2, 2, 2), 0)]
    pls = [(0, 1, (0, 0, 0)), (0, 1, (4.0, 0, 0)), (0, 1, (8.0, 8.0, -8.0))]
    lib = _lib()
    assert _create_rc(lib, vols, pls) != -1
    assert _oracle_rc(vols, pls) == 0


def test_overlap_sweep_matches_pairwise_check():
    """The library's sweep finds the same overlaps as the oracle's all-pairs test."""
    rs = np.random.default_rng(11)
    lib = _lib()
    verdicts = []
    for _ in range(300):
        k = int(rs.integers(2, 12))
        vols = [((50, 50, 50), -1)] + [(tuple(rs.uniform(1, 8, 3)), 0) for _ in range(k)]
        pls = []
        for j in range(k):
            lim = 50 - np.array(vols[1 + j][0])
            pls.append((0, 1 + j, tuple(np.round(rs.uniform(-lim, lim) / 4) * 4)))
        a, b = _create_rc(lib, vols, pls), _oracle_rc(vols, pls)
        assert (a == -1) == (b == -1)
        verdicts.append(a == -1)
    assert 0.2 < np.mean(verdicts) < 0.9


def test_find_next_boundary_cases():
    """NavigationInterface::FindNextBoundary (NavigationInterface.hpp:64-98)
    on the reference calorimeter: a point 5 mm into the first Pb slab moving
    +x has boundary distance 5 and safety 5; a point on the Pb/lAr face moving
    +x is in lAr with distance 10 and safety 0."""
    nav = O.Navigator(*abi.reference_calorimeter_tree())
    tol2 = 2e-9
    pos = np.array([[-495.0] * 4 + [-490.0] * 2, [0.0] * 6, [0.0] * 6])
    d = np.array([[1.0] * 6, [0.0] * 6, [0.0] * 6])
    t = nav.locate(pos, d)
    np.testing.assert_array_equal(t, [52, 52, 52, 52, 53, 53])
    pstep = np.array([1.0, 5.0, 7.0, 1e20, 0.0, 3.0])
    s, nxt, saf, b = nav.find_next_boundary(t, pos, d, pstep)
    # safety 5 > pstep 1: the physics step, no boundary
    assert (s[0], nxt[0], b[0]) == (1.0, 52, 0)
    # safety 5 <= pstep 5 = distance: no boundary, pushed past the step
    assert (s[1], nxt[1], b[1]) == (5.0 + tol2, 52, 0)
    # boundary within the step: pushed past the face, into the lAr slab
    assert (s[2], nxt[2], b[2]) == (5.0 + tol2, 53, 1)
    assert (s[3], nxt[3], b[3]) == (5.0 + tol2, 53, 1)
    # on the face (safety 0): pstep 0 gives the minimum push; pstep 3 < 10 no boundary
    assert (s[4], nxt[4], b[4]) == (tol2, 53, 0)
    assert (s[5], nxt[5], b[5]) == (3.0 + tol2, 53, 0)
    np.testing.assert_array_equal(saf, [5.0, 5.0, 5.0, 5.0, 0.0, 0.0])


def test_query_argument_checks_without_device():
    """Navigation queries and the other [3n] entry points reject bad
    arguments, and more than 2^29 points, before any device call."""
    lib = _lib()
    pos = np.zeros(3)
    d = np.array([1.0, 0.0, 0.0])
    t = np.zeros(1, dtype=np.int32)
    out = np.zeros(1)
    assert lib.gxt_nav_locate(None, abi.ptr(pos), abi.ptr(d), abi.ptr(t, C.c_int32), 1) == -1
    assert lib.gxt_nav_step(None, abi.ptr(t, C.c_int32), abi.ptr(pos), abi.ptr(d), abi.ptr(out),
                            abi.ptr(t, C.c_int32), abi.ptr(out), 1) == -1
    assert lib.gxt_nav_find_next_boundary(None, abi.ptr(t, C.c_int32), abi.ptr(pos), abi.ptr(d), abi.ptr(out),
                                          abi.ptr(out), abi.ptr(t, C.c_int32), abi.ptr(out), abi.ptr(t, C.c_int32),
                                          1) == -1
    assert lib.gxt_nav_free(None) == -1
    assert lib.gxt_nav_info(None, None, None, None, None) == -1
    f = (C.c_double * 3)(0.0, 0.0, 4.0)
    q, p, s = np.ones(1), np.ones(1), np.ones(1)
    big = (1 << 29) + 1
    assert lib.gxt_helix_step(f, abi.FP64, abi.ptr(q), abi.ptr(p), abi.ptr(s), abi.ptr(pos), abi.ptr(d), big) == -1
    assert b"2^29" in lib.gxt_last_error()
