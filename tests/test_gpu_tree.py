"""Transport through a volume tree on the device (GXT_GEOM_TREE) against the
oracle: the reference's calorimeter
(proxy/test/BasicCpuTransport/ProxyDetectorConstruction.cpp:38-60, 225-325:
50 layers of 10 mm Pb + 10 mm lAr along x, 1 m square, in a vacuum world)
with the primaries of its transport test (test_BasicCpuTransport.cpp:556-586,
677-698: 10 GeV, position in [0, 1) mm per coordinate, direction in the
positive octant).

- Event parity (FP64): secondary, step, pair and annihilation counts exact,
  energies within 1e-9 relative (FP64 atomic sums in another order), the
  per-cell deposits likewise; with no field, in a 4 T field, and with
  multiple scattering.
- The first 50 steps of every track of whole events bit-identical to the
  oracle's (process sequences and FP64 kinematics).
- Flight through the world's vacuum: the device gives the oracle's events
  and the identity of the oracle's own test (one more step, nothing else).
- The slab calorimeter restated as a tree gives the device's slab events
  (counts exact, no field).
- A basket handed over mid-shower continues as the oracle continues it.
- FP32 form: the total deposit within 0.5 % of the FP64 run.
"""
import numpy as np
import pytest

from geantexascalepilot_amd import abi
import oracle_harness as O
from test_oracle_tree import slab_as_tree

pytestmark = pytest.mark.gpu

COUNTS = ("n_steps", "n_pairs", "n_annihil", "n_secondaries")
ENERGIES = ("edep", "visible", "leak", "charged_length")


def _flush(msg):
    print(msg, flush=True)


def assert_events(a, b, what):
    for k in COUNTS:
        assert np.array_equal(a[k], b[k]), (what, k, np.nonzero(np.atleast_2d(a[k] != b[k]))[0][:5])
    for k in ENERGIES:
        np.testing.assert_allclose(a[k], b[k], rtol=1e-9, atol=1e-9, err_msg="%s %s" % (what, k))


def device_run(gpu_lib, cfg, prims, n, job=1):
    with gpu_lib.Transport(cfg) as t:
        t.submit(prims, job)
        t.run(max_iterations=500000)
        ev = t.event_tallies(job, n)
        prof = t.profiles()
        trace = t.trace() if cfg.trace_tracks > 0 else None
        st = t.stats()
    return ev, prof, trace, st


def oracle_run(cfg, prims, n):
    o = O.Oracle(cfg)
    o.run(prims, 16)
    return o.event_tallies(n), o.profiles(), (o.trace() if cfg.trace_tracks > 0 else None)


CASES = {
    "e-": dict(species=abi.ELECTRON),
    "gamma": dict(species=abi.GAMMA),
    "e- 4T": dict(species=abi.ELECTRON, field=(0.0, 0.0, 4.0)),
    "e- msc": dict(species=abi.ELECTRON, msc=1),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_reference_calorimeter_event_parity(gpu_lib, case):
    kw = dict(CASES[case])
    species = kw.pop("species")
    n = 8
    _flush("tree parity: reference calorimeter, 10 GeV %s, %d events" % (case, n))
    cfg = abi.config_reference_calorimeter(max_events=n, **kw)
    prims = abi.reference_calorimeter_primaries(n, species=species)
    g, gprof, _, st = device_run(gpu_lib, cfg, prims, n)
    c, cprof, _ = oracle_run(cfg, prims, n)
    assert_events(g, c, case)
    for a, b in zip(gprof, cprof):
        np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-9)
    assert len(gprof[2]) == 101 and gprof[2][100] == 0.0
    assert st.track_steps == int(c["n_steps"].sum())


@pytest.mark.parametrize("field", [(0.0, 0.0, 0.0), (0.0, 0.0, 4.0)])
def test_reference_calorimeter_trace_first_50_steps(gpu_lib, field):
    n = 2
    _flush("tree trace 50 steps, field %s" % (field,))
    cfg = abi.config_reference_calorimeter(max_events=n, field=field, trace_tracks=n, trace_steps=50)
    prims = abi.reference_calorimeter_primaries(n, ekin=4000.0, seed=3)
    g, _, gtr, _ = device_run(gpu_lib, cfg, prims, n)
    c, _, ctr = oracle_run(cfg, prims, n)
    assert len(gtr) == len(ctr) and len(gtr) > 1000

    def key(t):
        u = t[:, 0].copy().view(np.uint64)
        return np.lexsort((t[:, 1], u))

    a, b = gtr[key(gtr)], ctr[key(ctr)]
    assert np.array_equal(a[:, 0].view(np.uint64), b[:, 0].view(np.uint64))
    assert np.array_equal(a[:, 1:4], b[:, 1:4]), "step/species/process sequences differ"
    for j, nm in enumerate(["ekin", "x", "y", "z", "dx", "dy", "dz", "edep"]):
        assert np.array_equal(a[:, 4 + j], b[:, 4 + j]), nm
    assert_events(g, c, "trace run")


def test_flight_through_vacuum(gpu_lib):
    _flush("tree: vacuum flight")
    cfg = abi.config_reference_calorimeter(max_events=6)
    p = abi.primaries_array(6)
    for i in range(6):
        sp = (abi.ELECTRON, abi.GAMMA, abi.POSITRON)[i // 2]
        p[i].species, p[i].event, p[i].ekin = sp, i, 2000.0
        p[i].pos[:] = (-550.0 if i % 2 == 0 else -500.0, 0.3, -0.2)
        p[i].dir[:] = (1.0, 0.0, 0.0)
        p[i].substream = 7 + i // 2
    g, _, _, _ = device_run(gpu_lib, cfg, p, 6)
    c, _, _ = oracle_run(cfg, p, 6)
    assert_events(g, c, "vacuum flight")
    assert np.array_equal(g["n_steps"][0::2], g["n_steps"][1::2] + 1)
    for k in ENERGIES:
        np.testing.assert_allclose(g[k][0::2], g[k][1::2], rtol=1e-9, atol=1e-9, err_msg=k)


def test_slab_restated_as_tree_gives_the_slab_events(gpu_lib):
    n = 32
    _flush("tree: slab calorimeter restated, %d events" % n)
    slab = abi.config_calorimeter(max_events=n, field=(0.0, 0.0, 0.0))
    vols, pl, shift = slab_as_tree(slab)
    tree = abi.set_tree(abi.config_calorimeter(max_events=n, field=(0.0, 0.0, 0.0)), vols, pl)
    ps = gpu_lib.make_primaries(slab.seed, abi.ELECTRON, 1000.0, 3000.0, 10.0, n)
    pt = gpu_lib.make_primaries(slab.seed, abi.ELECTRON, 1000.0, 3000.0, 10.0, n)
    for q in pt:
        q.pos[2] = -shift
    a, pa, _, _ = device_run(gpu_lib, slab, ps, n)
    b, pb, _, _ = device_run(gpu_lib, tree, pt, n)
    assert_events(b, a, "tree vs slab")
    np.testing.assert_allclose(pb[2][:-1], pa[2], rtol=1e-9, atol=1e-9)


def test_basket_continuation_in_the_tree(gpu_lib):
    n = 8
    _flush("tree: basket continuation")
    cfg = abi.config_reference_calorimeter(max_events=n, field=(0.0, 0.0, 4.0))
    prims = abi.reference_calorimeter_primaries(n, ekin=3000.0, seed=5)
    part = O.Oracle(cfg)
    basket = part.run_partial(prims, 600)
    assert any(t.volume > 51 for t in basket)  # tracks in flight carry their touchable
    cpu = O.Oracle(cfg)
    cpu.run_tracks(basket, 16)
    C = cpu.event_tallies(n)
    with gpu_lib.Transport(cfg) as t:
        t.submit_tracks(basket, n, 3)
        t.run()
        G = t.event_tallies(3, n)
        b = abi.tracks_array(1)
        b[0] = basket[0]
        b[0].volume = 152  # one past the last touchable
        with pytest.raises(gpu_lib.GxtError, match="volume out of range"):
            t.submit_tracks(b, n, 4)
        b[0].pos[:] = (0.5, 0.5, 0.5)
        b[0].volume = 60  # a leaf of layer 4, far from the origin
        with pytest.raises(gpu_lib.GxtError, match="outside its volume"):
            t.submit_tracks(b, n, 4)
    assert_events(G, C, "tree basket")


def test_fp32_form_within_half_a_percent(gpu_lib):
    n = 64
    _flush("tree: FP32 vs FP64, %d events" % n)
    prims = abi.reference_calorimeter_primaries(n, seed=11)
    e64, _, _, _ = device_run(gpu_lib, abi.config_reference_calorimeter(max_events=n), prims, n)
    e32, _, _, _ = device_run(gpu_lib, abi.config_reference_calorimeter(max_events=n, precision=abi.FP32), prims, n)
    assert abs(e32["edep"].sum() / e64["edep"].sum() - 1.0) < 5e-3
    # visible energy: sampling fluctuations dominate; means within 4 standard errors
    d = e32["visible"].mean() - e64["visible"].mean()
    err = np.sqrt(e32["visible"].var() / n + e64["visible"].var() / n)
    assert abs(d) < 4 * err


def test_tree_errors(gpu_lib):
    vols, pl, _ = abi.reference_calorimeter_tree(materials=(abi.MAT_PB, abi.MAT_PBWO4))
    with pytest.raises(gpu_lib.GxtError, match="neither the absorber"):
        gpu_lib.Transport(abi.set_tree(abi.config_reference_calorimeter(max_events=2), vols, pl))
    with pytest.raises(gpu_lib.GxtError, match="uniform field only"):
        gpu_lib.Transport(abi.config_reference_calorimeter(max_events=2, field_stepper=abi.STEPPER_RK45))
    with gpu_lib.Transport(abi.config_reference_calorimeter(max_events=2)) as t:
        with pytest.raises(gpu_lib.GxtError, match="volume tree"):
            t.slab_distance(np.zeros((3, 4)), np.tile([[0.0], [0.0], [1.0]], (1, 4)))
