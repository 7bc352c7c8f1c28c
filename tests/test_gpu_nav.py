"""GPU parity of the volume-tree navigator: libgxt.so against the CPU oracle,
through the C ABI.

Locate, boundary distance, next touchable and safety are bit-exact: both
sides do the same correctly rounded FP64 operations (differences, quotients,
min/max, one square root).  Cases: the reference's calorimeter (152
touchables) over the whole world, points on every layer face with directions
that enter, leave and run along it, random nested trees, and rays walked
boundary to boundary through the calorimeter.
"""
import numpy as np
import pytest

from geantexascalepilot_amd import abi
import oracle_harness as O
from test_gpu_parity import _flush
from test_nav import _unit, many_daughters_tree, random_tree

pytestmark = pytest.mark.gpu


def _pair(gpu_lib, vols, pls, world=0):
    return gpu_lib.Navigator(vols, pls, world), O.Navigator(vols, pls, world)


def _same(a, b):
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


@pytest.fixture(scope="module")
def calo(gpu_lib):
    g, c = _pair(gpu_lib, *abi.reference_calorimeter_tree())
    yield g, c
    g.close()


def test_info_matches(calo):
    _flush("nav info: reference calorimeter")
    g, c = calo
    gi, ci = g.info(), c.info()
    for k in gi:
        np.testing.assert_array_equal(gi[k], ci[k])


def test_calorimeter_bit_exact(calo):
    _flush("nav locate/step: reference calorimeter, 200k points")
    g, c = calo
    rs = np.random.default_rng(21)
    n = 200000
    pos = rs.uniform(-1.0, 1.0, (3, n)) * np.array([[600.0], [600.0], [600.0]])
    d = _unit(rs, n)
    t = g.locate(pos, d)
    np.testing.assert_array_equal(t, c.locate(pos, d))
    ins = t >= 0
    _same(g.step(t[ins], pos[:, ins], d[:, ins]), c.step(t[ins], pos[:, ins], d[:, ins]))


def test_surface_points_bit_exact(calo):
    _flush("nav surface points: every layer face")
    g, c = calo
    rs = np.random.default_rng(22)
    faces = -500.0 + 10.0 * np.arange(101)
    m = 200
    x = np.repeat(faces, m)
    n = len(x)
    pos = np.stack([x, rs.uniform(-500, 500, n), rs.uniform(-500, 500, n)])
    d = _unit(rs, n)
    # a third run along the face (no x component), some along an axis
    d[0, ::3] = 0.0
    d[:, ::3] /= np.linalg.norm(d[:, ::3], axis=0)
    d[:, 1::7] = np.array([[1.0], [0.0], [0.0]])
    d[:, 2::7] = np.array([[-1.0], [0.0], [0.0]])
    t = g.locate(pos, d)
    np.testing.assert_array_equal(t, c.locate(pos, d))
    ins = t >= 0
    _same(g.step(t[ins], pos[:, ins], d[:, ins]), c.step(t[ins], pos[:, ins], d[:, ins]))


@pytest.mark.parametrize("seed", [3, 4, 5, 6, 7, 8])
def test_random_trees_bit_exact(gpu_lib, seed):
    _flush("nav random tree seed %d" % seed)
    rs = np.random.default_rng(seed)
    vols, pls = random_tree(rs)
    g, c = _pair(gpu_lib, vols, pls)
    try:
        gi, ci = g.info(), c.info()
        for k in gi:
            np.testing.assert_array_equal(gi[k], ci[k])
        half = np.array(vols[0][0])[:, None]
        n = 50000
        pos = rs.uniform(-1.1, 1.1, (3, n)) * half
        d = _unit(rs, n)
        t = g.locate(pos, d)
        np.testing.assert_array_equal(t, c.locate(pos, d))
        ins = t >= 0
        _same(g.step(t[ins], pos[:, ins], d[:, ins]), c.step(t[ins], pos[:, ins], d[:, ins]))
    finally:
        g.close()


@pytest.mark.parametrize("seed,layers", [(31, False), (32, True), (33, False)])
def test_many_daughters_bit_exact(gpu_lib, seed, layers):
    """Volumes with hundreds of daughters take the device's daughter index;
    the oracle scans every daughter.  Points include every daughter's faces."""
    _flush("nav many daughters seed %d layers %d" % (seed, layers))
    rs = np.random.default_rng(seed)
    vols, pls = many_daughters_tree(rs, n=400, layers=layers)
    g, c = _pair(gpu_lib, vols, pls)
    try:
        info = c.info()
        assert len(info["volume"]) > 300
        half = np.array([vols[v][0] for v in info["volume"]])
        n = 60000
        pos = rs.uniform(-1.02, 1.02, (3, n)) * 100.0
        # points on faces: a random face of a random touchable
        m = n // 2
        tt = rs.integers(1, len(half), m)
        ax = rs.integers(0, 3, m)
        sign = rs.choice([-1.0, 1.0], m)
        pos[:, :m] = info["centre"][tt].T + rs.uniform(-1, 1, (3, m)) * half[tt].T
        pos[ax, np.arange(m)] = info["centre"][tt, ax] + sign * half[tt, ax]
        d = _unit(rs, n)
        # along the face normal (in and out), and tangent to the face
        normal = np.arange(0, m, 5)
        d[:, normal] = 0.0
        d[ax[normal], normal] = rs.choice([-1.0, 1.0], len(normal))
        tangent = np.arange(1, m, 5)
        d[ax[tangent], tangent] = 0.0
        d[:, tangent] /= np.linalg.norm(d[:, tangent], axis=0)
        t = g.locate(pos, d)
        np.testing.assert_array_equal(t, c.locate(pos, d))
        ins = t >= 0
        _same(g.step(t[ins], pos[:, ins], d[:, ins]), c.step(t[ins], pos[:, ins], d[:, ins]))
    finally:
        g.close()


def test_walk_bit_exact(calo):
    """4096 rays from the world, stepped boundary to boundary until they leave it."""
    _flush("nav walk: 4096 rays through the reference calorimeter")
    g, c = calo
    rs = np.random.default_rng(23)
    n = 4096
    pos = np.stack([np.full(n, -600.0), rs.uniform(-300, 300, n), rs.uniform(-300, 300, n)])
    d = _unit(rs, n)
    d[0] = np.abs(d[0]) + 0.2
    d /= np.linalg.norm(d, axis=0)
    t = g.locate(pos, d)
    np.testing.assert_array_equal(t, c.locate(pos, d))
    crossings = 0
    for _ in range(400):
        live = np.nonzero(t >= 0)[0]
        if len(live) == 0:
            break
        gs = g.step(t[live], pos[:, live], d[:, live])
        _same(gs, c.step(t[live], pos[:, live], d[:, live]))
        s, nxt, _ = gs
        pos[:, live] += s * d[:, live]
        t[live] = nxt
        crossings += len(live)
    assert (t == -1).all()
    assert crossings > 100 * n // 4


def test_device_arrays_match_host_arrays(calo, gpu_lib):
    """Arrays already in HBM are used in place (stream-ordered on the null
    stream); a touchable out of range gives next = -2 and NaN distances."""
    import ctypes as C
    _flush("nav device-resident arrays")
    g, c = calo
    rs = np.random.default_rng(24)
    n = 100000
    pos = rs.uniform(-1.0, 1.0, (3, n)) * 600.0
    d = _unit(rs, n)
    t = c.locate(pos, d)
    t[::1000] = 152
    t[1::1000] = -7
    D = gpu_lib.DeviceArray
    dp, dd, dl = D(pos), D(d), D(np.zeros(n, dtype=np.int32))
    lib = gpu_lib.lib
    gpu_lib._check(lib.gxt_nav_locate(g.h, dp.ptr, dd.ptr, dl.ptr, n))
    np.testing.assert_array_equal(dl.numpy(), c.locate(pos, d))
    dt, ds, dn, dsaf = D(t), D(np.zeros(n)), D(np.zeros(n, dtype=np.int32)), D(np.zeros(n))
    gpu_lib._check(lib.gxt_nav_step(g.h, dt.ptr, dp.ptr, dd.ptr, ds.ptr, dn.ptr, dsaf.ptr, n))
    bad = (t < 0) | (t >= 152)
    ok = ~bad
    want = c.step(t[ok], pos[:, ok], d[:, ok])
    got = (ds.numpy(), dn.numpy(), dsaf.numpy())
    for x, y in zip(got, want):
        np.testing.assert_array_equal(x[ok], y)
    assert (got[1][bad] == -2).all() and np.isnan(got[0][bad]).all() and np.isnan(got[2][bad]).all()
    # host and device arrays mixed
    hs = np.zeros(n)
    assert lib.gxt_nav_step(g.h, dt.ptr, dp.ptr, dd.ptr, abi.ptr(hs), dn.ptr, dsaf.ptr, n) == -1
    for a in (dp, dd, dl, dt, ds, dn, dsaf):
        a.free()


def test_find_next_boundary_bit_exact(calo, gpu_lib):
    """FindNextBoundary over random points and proposed steps that fall below
    the safety, between safety and boundary, and beyond the boundary; host
    arrays and device arrays."""
    _flush("nav find_next_boundary")
    g, c = calo
    rs = np.random.default_rng(25)
    n = 100000
    pos = rs.uniform(-1.0, 1.0, (3, n)) * 600.0
    d = _unit(rs, n)
    t = c.locate(pos, d)
    ins = t >= 0
    t, pos, d = t[ins], pos[:, ins], d[:, ins]
    n = len(t)
    s0, _, saf0 = c.step(t, pos, d)
    u = rs.uniform(0.0, 1.0, n)
    pstep = np.select([u < 0.25, u < 0.5, u < 0.75, u < 0.8], [saf0 * rs.uniform(0, 1, n), s0,
                      saf0 + (s0 - saf0) * rs.uniform(0, 1, n), np.zeros(n)], s0 * 3 + 1)
    want = c.find_next_boundary(t, pos, d, pstep)
    _same(g.find_next_boundary(t, pos, d, pstep), want)
    assert 0.1 < want[3].mean() < 0.5
    D = gpu_lib.DeviceArray
    arrs = [D(t), D(pos), D(d), D(pstep), D(np.zeros(n)), D(np.zeros(n, dtype=np.int32)), D(np.zeros(n)),
            D(np.zeros(n, dtype=np.int32))]
    gpu_lib._check(gpu_lib.lib.gxt_nav_find_next_boundary(g.h, *[a.ptr for a in arrs], n))
    _same([arrs[4].numpy(), arrs[5].numpy(), arrs[6].numpy(), arrs[7].numpy()], want)
    for a in arrs:
        a.free()
    with pytest.raises(gpu_lib.GxtError, match="pstep"):
        g.find_next_boundary(t[:1], pos[:, :1], d[:, :1], np.array([-1.0]))


def test_bad_touchable_rejected(calo, gpu_lib):
    g, _ = calo
    pos = np.zeros((3, 1))
    d = np.array([[1.0], [0.0], [0.0]])
    with pytest.raises(gpu_lib.GxtError, match="touchable"):
        g.step(np.array([152]), pos, d)
    with pytest.raises(gpu_lib.GxtError, match="unit direction"):
        g.locate(pos, np.array([[1.0], [1.0], [0.0]]))
