"""ctypes mirror of the C ABI declared in include/gxt.h.

The structures here are byte-for-byte the ones in the header; the helper
functions build the configurations of the five transport workloads (a
homogeneous lead block, the PbWO4 block, the lead/liquid-argon sampling
calorimeter in a 4 T field, its FP32 precision form, and the serving mix).
"""
import ctypes as C
import os

import numpy as np

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")

GEOM_BOX, GEOM_CALO, GEOM_TREE = 0, 1, 2
MAT_PB, MAT_LAR, MAT_PBWO4 = 0, 1, 2
FP64, FP32 = 0, 1
STEPPER_HELIX, STEPPER_RK45 = 0, 1
RK_ANGLE = 0.25
CMS_FIELD_MAP = os.path.join(DATA_DIR, "field", "cmsmagfield2015.txt")
GAMMA, ELECTRON, POSITRON = 0, 1, 2
SAMPLE_NOALLOC_MAX = 65536  # GXT_SAMPLE_NOALLOC_MAX

TAB = dict(LAMBDA_COMPT=0, LAMBDAPRIM_COMPT=1, LAMBDA_CONV=2, LAMBDAPRIM_PHOT=3, LAMBDA_BREM_EM=4,
           LAMBDA_IONI_EM=5, DEDX_EM=6, RANGE_EM=7, INVRANGE_EM=8, LAMBDA_BREM_EP=9, LAMBDA_IONI_EP=10,
           DEDX_EP=11, RANGE_EP=12, INVRANGE_EP=13, MSC_EM=14, MSC_EP=15)
NTAB = 16
MODEL = dict(KLEIN_NISHINA=0, BETHE_HEITLER=1, SAUTER_GAVRILA=2, SELTZER_BERGER_EM=3, MOLLER=4,
             SELTZER_BERGER_EP=5, BHABHA=6, ANNIHIL=7)

ERRORS = {-1: "GXT_EINVAL", -2: "GXT_EHIP", -3: "GXT_EDATA", -4: "GXT_ECAPACITY", -5: "GXT_ESTATE",
          -6: "GXT_ENOTFOUND"}


class BoxVolume(C.Structure):
    _fields_ = [("half", C.c_double * 3), ("material", C.c_int32), ("pad", C.c_int32)]


class Placement(C.Structure):
    _fields_ = [("mother", C.c_int32), ("daughter", C.c_int32), ("translation", C.c_double * 3)]


class Config(C.Structure):
    _fields_ = [
        ("geometry", C.c_int32), ("material", C.c_int32), ("box_xy", C.c_double), ("box_depth", C.c_double),
        ("n_layers", C.c_int32), ("absorber", C.c_int32), ("gap", C.c_int32), ("pad0", C.c_int32),
        ("absorber_mm", C.c_double), ("gap_mm", C.c_double), ("lar_density", C.c_double),
        ("field", C.c_double * 3), ("tracking_cut", C.c_double), ("seed", C.c_uint32),
        ("precision", C.c_int32), ("zbin_mm", C.c_double), ("rbin_mm", C.c_double), ("n_zbins", C.c_int32),
        ("n_rbins", C.c_int32), ("pool_capacity", C.c_int64), ("max_steps", C.c_int32),
        ("max_events", C.c_int32), ("max_jobs", C.c_int32), ("trace_tracks", C.c_int32),
        ("trace_steps", C.c_int32), ("profile_kernels", C.c_int32), ("data_dir", C.c_char_p),
        ("field_map", C.c_char_p), ("field_origin", C.c_double * 3), ("field_stepper", C.c_int32),
        ("msc", C.c_int32), ("tree_volumes", C.POINTER(BoxVolume)), ("tree_placements", C.POINTER(Placement)),
        ("n_tree_volumes", C.c_int32), ("n_tree_placements", C.c_int32), ("tree_world", C.c_int32),
        ("pad1", C.c_int32),
    ]


class Primary(C.Structure):
    _fields_ = [("species", C.c_int32), ("event", C.c_int32), ("ekin", C.c_double), ("pos", C.c_double * 3),
                ("dir", C.c_double * 3), ("substream", C.c_uint64)]


class Track(C.Structure):
    """A track in flight (gxt_track): the hand-off state of a basket."""
    _fields_ = [("species", C.c_int32), ("event", C.c_int32), ("ekin", C.c_double), ("pos", C.c_double * 3),
                ("dir", C.c_double * 3), ("nil", C.c_double * 3), ("rng", C.c_uint32 * 6), ("volume", C.c_int32),
                ("nsteps", C.c_int32), ("substream", C.c_uint64)]


class EventTally(C.Structure):
    _fields_ = [("edep", C.c_double), ("visible", C.c_double), ("leak", C.c_double),
                ("charged_length", C.c_double), ("n_secondaries", C.c_int64 * 3), ("n_steps", C.c_int64),
                ("n_pairs", C.c_int64), ("n_annihil", C.c_int64)]


class Stats(C.Structure):
    _fields_ = [("track_steps", C.c_int64), ("iterations", C.c_int64), ("max_live", C.c_int64),
                ("killed_by_limit", C.c_int64), ("wall_ms", C.c_double)]


class KernelStats(C.Structure):
    _fields_ = [("transport_neutral_ms", C.c_double), ("transport_charged_ms", C.c_double),
                ("interact_neutral_ms", C.c_double), ("interact_charged_ms", C.c_double),
                ("derive_ms", C.c_double),
                ("advance_ms", C.c_double), ("launches", C.c_int64), ("track_steps", C.c_int64),
                ("neutral_steps", C.c_int64), ("charged_steps", C.c_int64), ("interactions", C.c_int64),
                ("created", C.c_int64), ("record_bytes", C.c_int32), ("pad", C.c_int32),
                ("neutral_loaded", C.c_int64), ("charged_loaded", C.c_int64),
                ("neutral_interactions", C.c_int64), ("neutral_created", C.c_int64)]


class JobResult(C.Structure):
    _fields_ = [("job_id", C.c_uint64), ("n_events", C.c_int32), ("pad", C.c_int32), ("edep", C.c_double),
                ("visible", C.c_double), ("submit_ms", C.c_double), ("admit_ms", C.c_double),
                ("done_ms", C.c_double), ("track_steps", C.c_int64),
                ("iterations", C.c_int64)]


class SampleIO(C.Structure):
    _fields_ = [("n", C.c_int), ("ekin", C.POINTER(C.c_double)), ("dir", C.POINTER(C.c_double)),
                ("rng", C.POINTER(C.c_uint32)), ("out_ekin", C.POINTER(C.c_double)),
                ("out_dir", C.POINTER(C.c_double)), ("out_edep", C.POINTER(C.c_double)),
                ("nsec", C.POINTER(C.c_int32)), ("sec_species", C.POINTER(C.c_int32)),
                ("sec_ekin", C.POINTER(C.c_double)), ("sec_dir", C.POINTER(C.c_double))]


def box_tree(volumes, placements):
    """ctypes arrays from [(half_xyz, material)] and [(mother, daughter, translation_xyz)]."""
    vol = (BoxVolume * len(volumes))()
    for v, (half, mat) in zip(vol, volumes):
        v.half[:] = [float(x) for x in half]
        v.material = mat
    pl = (Placement * max(len(placements), 1))()
    for q, (m, d, t) in zip(pl, placements):
        q.mother, q.daughter = m, d
        q.translation[:] = [float(x) for x in t]
    return vol, pl


def reference_calorimeter_tree(n_layers=50, thickness=(10.0, 10.0), materials=(MAT_PB, MAT_LAR), size_yz=1000.0):
    """The reference's calorimeter as a volume tree
    (proxy/test/BasicCpuTransport/ProxyDetectorConstruction.cpp:150-164, 225-325):
    world (1.2 x the calorimeter, vacuum) > calorimeter > n_layers layers along x
    > one box per absorber.  Returns (volumes, placements, world index)."""
    layer_x = float(sum(thickness))
    calo_x = n_layers * layer_x
    volumes = [((0.6 * calo_x, 0.6 * size_yz, 0.6 * size_yz), -1),
               ((0.5 * layer_x, 0.5 * size_yz, 0.5 * size_yz), -1),
               ((0.5 * calo_x, 0.5 * size_yz, 0.5 * size_yz), -1)]
    placements = []
    start = -0.5 * layer_x
    for i, (t, m) in enumerate(zip(thickness, materials)):
        volumes.append(((0.5 * t, 0.5 * size_yz, 0.5 * size_yz), m))
        placements.append((1, 3 + i, (start + 0.5 * t, 0.0, 0.0)))
        start += t
    start = -0.5 * calo_x
    for i in range(n_layers):
        placements.append((2, 1, (start + 0.5 * layer_x, 0.0, 0.0)))
        start += layer_x
    placements.append((0, 2, (0.0, 0.0, 0.0)))
    return volumes, placements, 0


def ptr(a, ctype=C.c_double):
    return a.ctypes.data_as(C.POINTER(ctype))


def base_config(**kw):
    """Defaults shared by every workload: 1 keV tracking cut, 1 cm bins."""
    c = Config()
    c.geometry = GEOM_BOX
    c.material = MAT_PB
    c.box_xy = 200.0
    c.box_depth = 200.0
    c.n_layers = 50
    c.absorber = MAT_PB
    c.gap = MAT_LAR
    c.absorber_mm = 2.3
    c.gap_mm = 5.7
    c.lar_density = 1.396
    c.tracking_cut = 1.0e-3
    c.seed = 12345
    c.precision = FP64
    c.zbin_mm = 10.0
    c.rbin_mm = 10.0
    c.n_zbins = 20
    c.n_rbins = 20
    c.pool_capacity = 2_000_000
    c.max_steps = 100000
    c.max_events = 1000
    c.max_jobs = 64
    c.trace_tracks = 0
    c.trace_steps = 0
    c.profile_kernels = 0
    c._data = DATA_DIR.encode()
    c.data_dir = c._data
    for k, v in kw.items():
        if k == "field":
            c.field[:] = v
        elif k == "field_origin":
            c.field_origin[:] = v
        elif k == "field_map":
            c._field_map = None if v is None else os.fsencode(v)
            c.field_map = c._field_map
        else:
            setattr(c, k, v)
    return c


def config_lead_block(**kw):
    """1 GeV gammas into a 20x20x20 cm lead block (11.35 g/cm3), seed 12345."""
    d = dict(geometry=GEOM_BOX, material=MAT_PB, box_xy=200.0, box_depth=200.0, seed=12345, n_zbins=20,
             n_rbins=10, max_events=1000)
    d.update(kw)
    return base_config(**d)


def config_pbwo4_block(**kw):
    """10 GeV electrons into a 30x30x30 cm PbWO4 block (8.28 g/cm3), seed 777."""
    d = dict(geometry=GEOM_BOX, material=MAT_PBWO4, box_xy=300.0, box_depth=300.0, seed=777, n_zbins=30,
             n_rbins=15, max_events=10000, pool_capacity=2_000_000)
    d.update(kw)
    return base_config(**d)


def config_calorimeter(precision=FP64, **kw):
    """50 x (2.3 mm Pb + 5.7 mm lAr 1.39 g/cm3), 40x40 cm, 4 T along +x."""
    d = dict(geometry=GEOM_CALO, n_layers=50, absorber=MAT_PB, gap=MAT_LAR, absorber_mm=2.3, gap_mm=5.7,
             lar_density=1.39, box_xy=400.0, field=(4.0, 0.0, 0.0), seed=2468, n_zbins=40, n_rbins=20,
             max_events=20000, precision=precision)
    d.update(kw)
    return base_config(**d)


def config_calorimeter_cms(precision=FP64, stepper=STEPPER_HELIX, **kw):
    """The sampling calorimeter in the CMS 2015 solenoid map, its front face
    centre at r = 600 mm, z = 6 m (the end region: Bz falls from 2.8 to 2.0 T
    across the calorimeter and Br reaches 0.9 T)."""
    d = dict(field=(0.0, 0.0, 0.0), field_map=CMS_FIELD_MAP, field_origin=(600.0, 0.0, 6000.0),
             field_stepper=stepper)
    d.update(kw)
    return config_calorimeter(precision=precision, **d)


def set_tree(c, volumes, placements, world=0):
    """Make config c a GXT_GEOM_TREE setup from [(half_xyz, material)] and
    [(mother, daughter, translation_xyz)] (the arrays stay referenced by c)."""
    vol, pl = box_tree(volumes, placements)
    c._tree = (vol, pl)
    c.geometry = GEOM_TREE
    c.tree_volumes = C.cast(vol, C.POINTER(BoxVolume))
    c.tree_placements = C.cast(pl, C.POINTER(Placement))
    c.n_tree_volumes = len(volumes)
    c.n_tree_placements = len(placements)
    c.tree_world = world
    return c


def config_reference_calorimeter(precision=FP64, **kw):
    """The reference's calorimeter as a volume tree
    (proxy/test/BasicCpuTransport/ProxyDetectorConstruction.cpp:38-60,
    225-325): 50 layers of 10 mm Pb + 10 mm lAr along x, 1 m square, in a
    vacuum world 1.2 times its size; no field unless given; 0.7 mm cuts,
    1 keV tracking cut.  lAr at the reference tables' 1.396 g/cm3."""
    tree = kw.pop("tree", None) or reference_calorimeter_tree()
    d = dict(absorber=MAT_PB, gap=MAT_LAR, lar_density=1.396, seed=12345, n_zbins=20, n_rbins=20,
             zbin_mm=50.0, rbin_mm=25.0, max_events=1000, precision=precision)
    d.update(kw)
    return set_tree(base_config(**d), *tree)


def reference_calorimeter_primaries(n, species=ELECTRON, ekin=10000.0, seed=0):
    """Primaries of the reference's transport test (test_BasicCpuTransport.cpp:
    556-586, 677-698): 10 GeV, position with each coordinate uniform in
    [0, 1) mm, direction (u1, u2, u3) normalised with u uniform in [0, 1);
    here from numpy's generator with the given seed, one event per primary,
    substream = event index."""
    rs = np.random.default_rng(seed)
    p = primaries_array(n)
    for i in range(n):
        u = rs.random(6)
        d = u[3:] / np.linalg.norm(u[3:])
        p[i].species, p[i].event, p[i].ekin = species, i, ekin
        p[i].pos[:] = [float(x) for x in u[:3]]
        p[i].dir[:] = [float(x) for x in d]
        p[i].substream = i
    return p


def primaries_array(n):
    return (Primary * n)()


def tracks_array(n):
    return (Track * n)()


def tallies_to_numpy(t):
    n = len(t)
    out = {k: np.zeros(n) for k in ("edep", "visible", "leak", "charged_length")}
    out["n_secondaries"] = np.zeros((n, 3), dtype=np.int64)
    out["n_steps"] = np.zeros(n, dtype=np.int64)
    out["n_pairs"] = np.zeros(n, dtype=np.int64)
    out["n_annihil"] = np.zeros(n, dtype=np.int64)
    for i in range(n):
        for k in ("edep", "visible", "leak", "charged_length"):
            out[k][i] = getattr(t[i], k)
        out["n_secondaries"][i] = list(t[i].n_secondaries)
        out["n_steps"][i] = t[i].n_steps
        out["n_pairs"][i] = t[i].n_pairs
        out["n_annihil"][i] = t[i].n_annihil
    return out
