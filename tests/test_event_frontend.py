"""Event front end (host only, no GPU): the HEPEVT reader and the egamma gun.

The reference ships no HEPEVT files, so the reader is pinned by small
hand-written events whose primaries follow from the format's definition:
which entries are primaries (daughters without daughters, and stable
top-level entries), their kinetic energies (closed form from momentum and
mass, in MeV), directions and substreams.  The library and the oracle must
agree on every field.  The gun is pinned against the oracle draw for draw
(directions to libm precision) and by the distributions it samples (gamma
fraction, flat momentum, flat pseudorapidity and azimuth)."""
import math

import numpy as np
import pytest
from scipy.stats import kstest

import geantexascalepilot_amd.gxt as g
import oracle_harness as O

ME = 0.510998910


def write(tmp_path, name, events):
    """events: list of lists of (ISTHEP, IDHEP, JDAHEP1, JDAHEP2, PX, PY, PZ, M) in GeV."""
    p = tmp_path / name
    lines = []
    for ev in events:
        lines.append("%d" % len(ev))
        for e in ev:
            lines.append("%d %d %d %d %.17g %.17g %.17g %.17g" % e)
    p.write_text("\n".join(lines) + "\n")
    return str(p)


def fields(prims):
    return [(q.species, q.event, q.ekin, tuple(q.dir), tuple(q.pos), q.substream) for q in prims]


def test_flat_event_list(tmp_path):
    """Stable top-level entries are primaries; Ekin = sqrt(p^2 + m^2) - m."""
    path = write(tmp_path, "flat.hepevt", [
        [(1, 22, 0, 0, 0.0, 0.0, 2.0, 0.0), (1, 11, 0, 0, 0.003, 0.004, 0.0, ME * 1e-3)],
        [(1, -11, 0, 0, 0.0, -1.0, 0.0, ME * 1e-3)],
    ])
    p, ne, nu = g.read_hepevt(path, vertex=(1.0, 2.0, 3.0), first=100)
    assert (len(p), ne, nu) == (3, 2, 0)
    assert [q.species for q in p] == [0, 1, 2]
    assert [q.event for q in p] == [0, 0, 1]
    assert [q.substream for q in p] == [100, 101, 102]
    assert p[0].ekin == 2000.0 and list(p[0].dir) == [0.0, 0.0, 1.0]
    pe = 5.0  # MeV
    assert p[1].ekin == pytest.approx(math.sqrt(pe * pe + ME * ME) - ME, rel=1e-14)
    assert list(p[1].dir) == pytest.approx([0.6, 0.8, 0.0], abs=1e-15)
    assert p[2].ekin == pytest.approx(math.sqrt(1e6 + ME * ME) - ME, rel=1e-14)
    assert all(list(q.pos) == [1.0, 2.0, 3.0] for q in p)


def test_mass_is_not_squared(tmp_path):
    """An electron of 1 MeV/c: the mass column enters as the mass (the
    reference squares it, which would give almost the massless energy)."""
    path = write(tmp_path, "slow.hepevt", [[(1, 11, 0, 0, 0.0, 0.0, 1e-3, ME * 1e-3)]])
    p, _, _ = g.read_hepevt(path)
    right = math.sqrt(1.0 + ME * ME) - ME
    squared = 1.0 / (math.sqrt(1.0 + (ME * 1e-3) ** 4 * 1e6) + (ME * 1e-3) ** 2 * 1e3)
    assert p[0].ekin == pytest.approx(right, rel=1e-14)
    assert abs(p[0].ekin - squared) > 0.3


def test_decay_tree(tmp_path):
    """A decaying pi0 (status 2, daughters 2-3) gives its two photons; a
    decayed intermediate with daughters of its own is not a primary; a stable
    charged pion is counted as unsupported."""
    path = write(tmp_path, "tree.hepevt", [[
        (2, 111, 2, 3, 0.0, 0.0, 1.0, 0.135),
        (1, 22, 0, 0, 0.1, 0.0, 0.5, 0.0),
        (1, 22, 0, 0, -0.1, 0.0, 0.5, 0.0),
        (2, 22, 5, 6, 0.0, 0.0, 3.0, 0.0),
        (1, 11, 0, 0, 0.0, 0.0, 1.5, ME * 1e-3),
        (1, -11, 0, 0, 0.0, 0.0, 1.5, ME * 1e-3),
        (1, 211, 0, 0, 0.0, 0.3, 0.0, 0.1396),
    ]])
    p, ne, nu = g.read_hepevt(path)
    assert ne == 1 and nu == 1
    assert [q.species for q in p] == [0, 0, 1, 2]
    assert p[0].ekin == pytest.approx(1000.0 * math.hypot(0.1, 0.5), rel=1e-14)


def test_max_events_and_counting(tmp_path):
    ev = [(1, 22, 0, 0, 0.0, 0.0, 1.0, 0.0)]
    path = write(tmp_path, "many.hepevt", [ev] * 5)
    p, ne, _ = g.read_hepevt(path, max_events=3)
    assert (len(p), ne) == (3, 3)
    p, ne, _ = g.read_hepevt(path)
    assert (len(p), ne) == (5, 5)


def test_malformed_files(tmp_path):
    bad = tmp_path / "trunc.hepevt"
    bad.write_text("2\n1 22 0 0 0 0 1 0\n")
    with pytest.raises(g.GxtError):
        g.read_hepevt(str(bad))
    out_of_range = write(tmp_path, "range.hepevt", [[(2, 111, 2, 5, 0.0, 0.0, 1.0, 0.135),
                                                     (1, 22, 0, 0, 0.0, 0.0, 0.5, 0.0)]])
    with pytest.raises(g.GxtError):
        g.read_hepevt(out_of_range)
    with pytest.raises(g.GxtError):
        g.read_hepevt(str(tmp_path / "missing.hepevt"))


def test_reader_matches_oracle(tmp_path):
    rs = np.random.default_rng(5)
    events = []
    for k in range(40):
        ev = []
        n = int(rs.integers(1, 8))
        for i in range(n):
            pdg = int(rs.choice([22, 11, -11, 211, 2212]))
            m = {22: 0.0, 11: ME * 1e-3, -11: ME * 1e-3, 211: 0.1396, 2212: 0.938}[pdg]
            ev.append((1, pdg, 0, 0, *rs.normal(0, 2, 3), m))
        if k % 3 == 0:  # a decaying parent in front
            ev = [(2, 111, 2, 1 + len(ev), 0.0, 0.0, 1.0, 0.135)] + ev
        events.append(ev)
    path = write(tmp_path, "mix.hepevt", events)
    a = g.read_hepevt(path, vertex=(0.5, -0.5, 0.0), first=7)
    b = O.read_hepevt(path, vertex=(0.5, -0.5, 0.0), first=7)
    assert a[1:] == b[1:]
    assert fields(a[0]) == fields(b[0])


def test_egamma_gun_matches_oracle_and_distributions():
    n_ev, per = 2000, 10
    kw = dict(gamma_fraction=0.3, pmin=1000.0, pmax=20000.0, eta=(-1.0, 2.0), phi=(0.5, 2.5))
    a = g.egamma_gun(4242, n_ev, per, vertex=(0.0, 0.0, -1.0), **kw)
    b = O.egamma_gun(4242, n_ev, per, kw["gamma_fraction"], kw["pmin"], kw["pmax"], kw["eta"], kw["phi"],
                     vertex=(0.0, 0.0, -1.0))
    # draws, species, events, substreams and energies exactly; directions to
    # libm precision (the compilers may reach sin/cos through different entry
    # points, e.g. sincos)
    fa, fb = fields(a), fields(b)
    assert [(x[0], x[1], x[2], x[4], x[5]) for x in fa] == [(x[0], x[1], x[2], x[4], x[5]) for x in fb]
    np.testing.assert_allclose([x[3] for x in fa], [x[3] for x in fb], rtol=0, atol=1e-15)
    sp = np.array([q.species for q in a])
    frac = (sp == 0).mean()
    assert abs(frac - 0.3) < 4 * math.sqrt(0.3 * 0.7 / len(a))
    assert set(np.unique(sp)) <= {0, 1}
    e = np.array([q.ekin for q in a])
    m = np.where(sp == 0, 0.0, ME)
    p = np.sqrt(e * (e + 2 * m))
    assert kstest((p - 1000.0) / 19000.0, "uniform").pvalue > 0.01
    d = np.array([list(q.dir) for q in a])
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-14)
    eta = -np.log(np.tan(0.5 * np.arccos(d[:, 2])))
    phi = np.arctan2(d[:, 1], d[:, 0])
    assert kstest((eta + 1.0) / 3.0, "uniform").pvalue > 0.01
    assert kstest((phi - 0.5) / 2.0, "uniform").pvalue > 0.01
    assert [q.event for q in a[:12]] == [0] * 10 + [1] * 2
    assert [q.substream for q in a[:3]] == [0, 1, 2]
