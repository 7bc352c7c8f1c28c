"""Non-uniform field (host only, no GPU): the cylindrical (r, z) field map and
the Cash-Karp RK45 stepper of the oracle, and the library's map reader.

The map lookup is pinned by the reference's own checks of its CMS map
(examples/magneticfield/simplifiedCMS/test/MainMagFieldTest.cpp:84-184):
at every node the field is the node's value, at the midpoint of a cell edge
the mean of the edge's two nodes, at the centre of a cell the mean of its four
nodes; here all three hold exactly (the last to an ulp) over the whole
181 x 161 map.  A bilinear function on a synthetic grid is reproduced to
rounding (a transposed or swapped weight fails), points outside the grid take
the edge values, the axis takes azimuth 0, and the cylindrical components turn
with the azimuth.

The stepper is pinned by the exact helix: in a uniform field the Cash-Karp
steps follow it to 1e-9 mm over the reference's stepper test track
(magneticfield/test/testConstFieldStepper.cpp:57-126: B = (1, 0, 0) T,
p = (0, 1/sqrt 2, 1/sqrt 2) GeV, q = +1, 20 steps of 200 mm), the error over
a fixed arc falls by 2^5 per halving of the step (fifth order), and the
embedded error estimate falls as h^5 and bounds the actual error.  Through
the transport, a map holding a constant axial field reproduces the uniform
field run step for step (its bilinear weights sum to 1 only to rounding, so
energies agree to 1e-12, not bitwise)."""
import math
import os

import numpy as np
import pytest

from geantexascalepilot_amd import abi
import oracle_harness as O

ME = 0.510998910
KB2C = 0.299792458


@pytest.fixture(scope="module")
def cms():
    return O.FieldMap(abi.CMS_FIELD_MAP)


@pytest.fixture(scope="module")
def nodes():
    """(r, z, Br, Bphi, Bz) of every node, parsed here independently of the oracle."""
    t = np.loadtxt(abi.CMS_FIELD_MAP, usecols=(1, 5, 9, 11, 7))
    return t.reshape(181, 161, 5)


def test_cms_grid(cms, nodes):
    info = cms.info()
    assert (info["nr"], info["nz"]) == (181, 161)
    assert (info["r0"], info["z0"], info["dr"], info["dz"]) == (0.0, -16000.0, 50.0, 200.0)
    b = nodes[:, :, 2:]
    assert info["bmax"] == pytest.approx(np.sqrt((b * b).sum(axis=2)).max(), rel=1e-15)
    assert 3.8 < info["bmax"] < 4.5


def _at(cms, r, z, y=None):
    r = np.asarray(r, dtype=np.float64).ravel()
    z = np.asarray(z, dtype=np.float64).ravel()
    y = np.zeros_like(r) if y is None else np.asarray(y, dtype=np.float64).ravel()
    return cms.value(np.stack([r, y, z]))


def test_nodes_exact(cms, nodes):
    b = _at(cms, nodes[:, :, 0], nodes[:, :, 1])
    want = nodes[:, :, 2:].reshape(-1, 3).T
    assert np.array_equal(b, want)


def test_edge_midpoints_exact(cms, nodes):
    n = nodes
    # midpoint in r: mean of (ir, iz) and (ir + 1, iz)
    b = _at(cms, n[:-1, :, 0] + 25.0, n[:-1, :, 1])
    want = ((n[:-1, :, 2:] + n[1:, :, 2:]) * 0.5).reshape(-1, 3).T
    assert np.array_equal(b, want)
    # midpoint in z
    b = _at(cms, n[:, :-1, 0], n[:, :-1, 1] + 100.0)
    want = ((n[:, :-1, 2:] + n[:, 1:, 2:]) * 0.5).reshape(-1, 3).T
    assert np.array_equal(b, want)


def test_cell_centres(cms, nodes):
    n = nodes
    b = _at(cms, n[:-1, :-1, 0] + 25.0, n[:-1, :-1, 1] + 100.0)
    want = ((n[:-1, :-1, 2:] + n[1:, :-1, 2:] + n[:-1, 1:, 2:] + n[1:, 1:, 2:]) * 0.25).reshape(-1, 3).T
    np.testing.assert_allclose(b, want, rtol=0, atol=1e-15 * np.abs(want).max())


def test_outside_grid_takes_edge_values(cms, nodes):
    n = nodes
    # beyond r = 9 m: the r = 9 m nodes; beyond |z| = 16 m: the z edge nodes
    b = _at(cms, np.full(161, 2.0e4), n[-1, :, 1])
    assert np.array_equal(b, n[-1, :, 2:].T)
    b = _at(cms, n[:, 0, 0], np.full(181, -4.0e4))
    assert np.array_equal(b, n[:, 0, 2:].T)
    b = _at(cms, n[:, -1, 0], np.full(181, 4.0e4))
    assert np.array_equal(b, n[:, -1, 2:].T)
    # the upper corner node itself
    b = _at(cms, [9000.0], [16000.0])
    assert np.array_equal(b[:, 0], n[-1, -1, 2:])


def test_axis_and_azimuth(cms, nodes):
    rs = np.random.default_rng(3)
    m = 2000
    r = rs.uniform(0.0, 9000.0, m)
    z = rs.uniform(-16000.0, 16000.0, m)
    phi = rs.uniform(-math.pi, math.pi, m)
    cyl = _at(cms, r, z)  # y = 0: (Br, Bphi, Bz)
    b = cms.value(np.stack([r * np.cos(phi), r * np.sin(phi), z]))
    c, s = np.cos(phi), np.sin(phi)
    np.testing.assert_allclose(b[0], cyl[0] * c - cyl[1] * s, rtol=0, atol=1e-13)
    np.testing.assert_allclose(b[1], cyl[0] * s + cyl[1] * c, rtol=0, atol=1e-13)
    np.testing.assert_allclose(b[2], cyl[2], rtol=0, atol=1e-13)
    # on the axis the azimuth is 0: (Bx, By) = (Br, Bphi)
    b = _at(cms, np.zeros(161), nodes[0, :, 1])
    assert np.array_equal(b, nodes[0, :, 2:].T)


def _write_map(path, r, z, f):
    with open(path, "w") as out:
        for rv in r:
            for zv in z:
                bz, br, bp = f(rv, zv)
                out.write("R(mm)= %.17g phi(deg)= 0 Z(mm)= %.17g   Bz(tesla)= %.17g Br(tesla)= %.17g "
                          "Bphi(tesla)= %.17g\n" % (rv, zv, bz, br, bp))
    return str(path)


def test_bilinear_function_reproduced(tmp_path):
    """Bilinear interpolation is exact on a function bilinear in (r, z); a
    grid that does not start at r = 0 and has nr != nz is read from the file."""
    coef = dict(bz=(1.5, 2e-3, -1e-3, 4e-6), br=(-0.5, 1e-3, 3e-3, -2e-6), bp=(0.25, -4e-3, 1e-3, 1e-6))

    def f(r, z):
        return tuple(a + b * r + c * z + d * r * z for a, b, c, d in (coef["bz"], coef["br"], coef["bp"]))

    r = 100.0 + 30.0 * np.arange(7)
    z = -250.0 + 40.0 * np.arange(11)
    m = O.FieldMap(_write_map(tmp_path / "lin.txt", r, z, f))
    info = m.info()
    assert (info["nr"], info["nz"], info["r0"], info["z0"], info["dr"], info["dz"]) == (7, 11, 100.0, -250.0,
                                                                                     30.0, 40.0)
    rs = np.random.default_rng(11)
    rr = rs.uniform(100.0, 280.0, 500)
    zz = rs.uniform(-250.0, 150.0, 500)
    b = m.value(np.stack([rr, np.zeros(500), zz]))
    bz, br, bp = f(rr, zz)
    np.testing.assert_allclose(b[0], br, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(b[1], bp, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(b[2], bz, rtol=1e-12, atol=1e-12)
    # r below the first node clamps to it
    b = m.value(np.array([[10.0], [0.0], [0.0]]))
    np.testing.assert_allclose(b[:, 0], [f(100.0, 0.0)[1], f(100.0, 0.0)[2], f(100.0, 0.0)[0]], rtol=1e-12)


BAD_MAPS = {
    "label": "R(mm)= 0 phi(deg)= 0 Z(mm)= 0 Bz(tesla)= 1 Br(tesla)= 0 Bphi= 0\n",
    "number": "R(mm)= 0 phi(deg)= 0 Z(mm)= 0 Bz(tesla)= x1 Br(tesla)= 0 Bphi(tesla)= 0\n",
    "extra": "R(mm)= 0 phi(deg)= 0 Z(mm)= 0 Bz(tesla)= 1 Br(tesla)= 0 Bphi(tesla)= 0 7\n",
}


def _grid_lines(nr, nz, skip=None, move=None):
    out = []
    for i in range(nr):
        for j in range(nz):
            if skip == (i, j):
                continue
            z = 10.0 * j + (3.0 if move == (i, j) else 0.0)
            out.append("R(mm)= %g phi(deg)= 0 Z(mm)= %g Bz(tesla)= 1 Br(tesla)= 0 Bphi(tesla)= 0\n" % (5.0 * i, z))
    return "".join(out)


@pytest.mark.parametrize("case", ["label", "number", "extra", "truncated", "offgrid", "single_r", "missing"])
def test_malformed_maps_rejected(tmp_path, case):
    """Oracle and library readers reject the same files (the library's reader
    runs on the host before any device call, so this needs no GPU)."""
    import geantexascalepilot_amd.gxt as g
    p = tmp_path / ("%s.txt" % case)
    if case in BAD_MAPS:
        p.write_text(_grid_lines(2, 3) + BAD_MAPS[case])
    elif case == "truncated":
        p.write_text(_grid_lines(3, 4, skip=(2, 3)))
    elif case == "offgrid":
        p.write_text(_grid_lines(3, 4, move=(1, 2)))
    elif case == "single_r":
        p.write_text(_grid_lines(1, 4))
    with pytest.raises(RuntimeError, match="-3"):
        O.FieldMap(str(p))
    with pytest.raises(g.GxtError) as e:
        g.FieldMap(str(p))
    assert e.value.code == -3


def _helix(field, q, mom, step, pos, d):
    p, dd = O.helix_step(field, [q], [mom], [step], np.array(pos, dtype=float).reshape(3, 1),
                         np.array(d, dtype=float).reshape(3, 1))
    return p[:, 0], dd[:, 0]


def _rk(field, q, mom, step, pos, d, fmap=None, stepper=abi.STEPPER_RK45, origin=(0.0, 0.0, 0.0)):
    p, dd, err = O.field_step(fmap, field, [q], [mom], [step], np.array(pos, dtype=float).reshape(3, 1),
                              np.array(d, dtype=float).reshape(3, 1), stepper=stepper, origin=origin)
    return p[:, 0], dd[:, 0], err[:, 0]


def test_rk45_follows_helix_on_reference_track():
    field = (1.0, 0.0, 0.0)
    mom = 1000.0
    d0 = np.array([0.0, 0.5 * math.sqrt(2.0), 0.5 * math.sqrt(2.0)])
    ph, dh = np.zeros(3), d0.copy()
    pr, dr = np.zeros(3), d0.copy()
    for _ in range(20):
        ph, dh = _helix(field, 1.0, mom, 200.0, ph, dh)
        pr, dr, err = _rk(field, 1.0, mom, 200.0, pr, dr)
        assert np.abs(pr - ph).max() < 1e-9 * 4000.0
        assert np.abs(dr - dh).max() < 1e-11
        assert abs(np.linalg.norm(dr) - 1.0) < 1e-15
        assert 0.0 < err[0] < 1e-6 and 0.0 < err[1] < 1e-9
    # the track turned about the field axis: the drift along x is zero
    assert abs(pr[0]) < 1e-9
    assert np.linalg.norm(pr) > 1000.0


def test_rk45_bending_sense():
    """A positive charge moving along +x in B = +z bends towards -y (F = q v x B)."""
    for stepper in (abi.STEPPER_HELIX, abi.STEPPER_RK45):
        p, d, _ = _rk((0.0, 0.0, 2.0), 1.0, 100.0, 10.0, (0, 0, 0), (1, 0, 0), stepper=stepper)
        assert d[1] < 0.0 and p[1] < 0.0
        p, d, _ = _rk((0.0, 0.0, 2.0), -1.0, 100.0, 10.0, (0, 0, 0), (1, 0, 0), stepper=stepper)
        assert d[1] > 0.0 and p[1] > 0.0


def test_rk45_fifth_order_and_error_estimate():
    field = (0.3, -0.2, 1.7)
    bm = math.sqrt(sum(b * b for b in field))
    mom = 50.0
    d0 = np.array([0.6, 0.0, 0.8])
    k = KB2C * bm / mom            # curvature of the full momentum, 1/mm
    arc = 1.0 / k                  # one radian of turning
    pe, de = _helix(field, -1.0, mom, arc, np.zeros(3), d0)
    errs, ests = [], []
    for n in (4, 8, 16, 32):
        p, d = np.zeros(3), d0.copy()
        est = 0.0
        for _ in range(n):
            p, d, e = _rk(field, -1.0, mom, arc / n, p, d)
            est += e[0]
        errs.append(np.linalg.norm(p - pe))
        ests.append(est)
    ratios = [errs[i] / errs[i + 1] for i in range(3)]
    assert all(20.0 < r < 50.0 for r in ratios[:2]), ratios   # global error O(h^5)
    eratios = [ests[i] / ests[i + 1] for i in range(3)]
    assert all(12.0 < r < 40.0 for r in eratios), eratios     # summed local estimate O(h^4)
    assert all(e > a for e, a in zip(ests, errs))              # the estimate bounds the actual error


def test_rk45_substeps_are_chained_single_steps():
    """One call over 0.9 rad takes ceil(0.9 / 0.25) = 4 equal Cash-Karp steps:
    the same as four calls of a quarter of the length."""
    field = (0.0, 1.0, 3.0)
    mom = 20.0
    s = 0.9 * mom / (KB2C * math.sqrt(10.0))
    p1, d1, _ = _rk(field, 1.0, mom, s, (1, 2, 3), (0.0, 0.6, 0.8))
    p, d = np.array([1.0, 2, 3]), np.array([0.0, 0.6, 0.8])
    for _ in range(4):
        p, d, _ = _rk(field, 1.0, mom, s / 4, p, d)
    assert np.array_equal(p1, p) and np.array_equal(d1, d)


def test_rk45_in_cms_map_follows_local_helix(cms):
    """In the slowly varying central field the RK45 and local-helix steppers
    agree over short steps; a 10 GeV track crosses the 3.8 T core nearly as a
    helix in the central field."""
    rs = np.random.default_rng(5)
    n = 200
    pos = np.stack([rs.uniform(-1000, 1000, n), rs.uniform(-1000, 1000, n), rs.uniform(-3000, 3000, n)])
    d = rs.normal(size=(3, n))
    d /= np.linalg.norm(d, axis=0)
    mom = rs.uniform(100.0, 5000.0, n)
    q = rs.choice([-1.0, 1.0], n)
    step = np.full(n, 5.0)
    ph, dh, _ = O.field_step(cms, None, q, mom, step, pos, d, stepper=abi.STEPPER_HELIX)
    pr, dr, err = O.field_step(cms, None, q, mom, step, pos, d, stepper=abi.STEPPER_RK45)
    # the local helix ignores the field gradient over 5 mm: a few 1e-7 mm
    assert np.abs(pr - ph).max() < 1e-5
    assert np.abs(dr - dh).max() < 1e-5
    # the bilinear map has kinks at cell edges, where a Cash-Karp step loses
    # order: estimates stay below 1e-7 mm over 5 mm
    assert (err[0] < 1e-7).all()


def test_uniform_map_reproduces_uniform_field_run(tmp_path):
    """The general propagation with a map holding a constant axial field and
    the helix stepper takes the same steps as the constant-field path (the
    interpolated field is 3 T to an ulp)."""
    path = _write_map(tmp_path / "const.txt", 100.0 * np.arange(11), -1000.0 + 200.0 * np.arange(11),
                      lambda r, z: (3.0, 0.0, 0.0))
    n = 6
    prim = O.make_primaries(4321, abi.ELECTRON, 2000.0, 2000.0, 10.0, n)
    runs = []
    for kw in (dict(field=(0.0, 0.0, 3.0)), dict(field=(0.0, 0.0, 0.0), field_map=path)):
        cfg = abi.config_calorimeter(max_events=n, seed=4321, **kw)
        o = O.Oracle(cfg)
        o.run(prim, nthreads=2)
        runs.append(o.event_tallies(n))
        o.close()
    for k in runs[0]:
        if runs[0][k].dtype.kind == "f":
            np.testing.assert_allclose(runs[0][k], runs[1][k], rtol=1e-12, atol=1e-12, err_msg=k)
        else:
            assert np.array_equal(runs[0][k], runs[1][k]), k


def test_rk45_transport_matches_helix_transport():
    """In a uniform field the RK45 and exact-helix transports take the same
    steps to ~1e-9 mm.  A track whose boundary or step decision flips on that
    difference follows another history from there, so an event either agrees
    to ~1e-8 or (rarely) differs like two statistically equivalent events."""
    n = 8
    prim = O.make_primaries(97, abi.ELECTRON, 1000.0, 3000.0, 10.0, n)
    runs = []
    for st in (abi.STEPPER_HELIX, abi.STEPPER_RK45):
        cfg = abi.config_calorimeter(max_events=n, seed=97, field_stepper=st)
        o = O.Oracle(cfg)
        o.run(prim, nthreads=2)
        runs.append(o.event_tallies(n))
        o.close()
    np.testing.assert_allclose(runs[0]["edep"], runs[1]["edep"], rtol=1e-6)
    vis = np.abs(runs[1]["visible"] / runs[0]["visible"] - 1.0)
    assert (vis < 1e-6).sum() >= n // 2
    assert vis.max() < 0.02
    np.testing.assert_allclose(runs[0]["n_steps"], runs[1]["n_steps"], rtol=5e-3)
