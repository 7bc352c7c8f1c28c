"""The C ABI without a GPU: libgxt.so loads and exports every function that
include/gxt.h declares; the ctypes mirror has the header's struct layout
(checked against a C compiler's sizeof/offsetof); argument validation fails
with the documented codes before any device work."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from geantexascalepilot_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gxt.h")
LIB = os.path.join(ROOT, "geantexascalepilot_amd", "libgxt.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gxt_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    assert os.path.exists(LIB), "libgxt.so not built"
    lib = C.CDLL(LIB)
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


STRUCTS = {
    "gxt_config": abi.Config,
    "gxt_primary": abi.Primary,
    "gxt_track": abi.Track,
    "gxt_event_tally": abi.EventTally,
    "gxt_stats": abi.Stats,
    "gxt_job_result": abi.JobResult,
    "gxt_sample_io": abi.SampleIO,
    "gxt_kernel_stats": abi.KernelStats,
    "gxt_box_volume": abi.BoxVolume,
    "gxt_placement": abi.Placement,
}


def test_struct_layout_matches_header(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "%s"' % HEADER, "int main(void){"]
    for cname, py in STRUCTS.items():
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in py._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-o", str(exe), str(src)])
    out = subprocess.check_output([str(exe)]).decode().split("\n")
    got = {}
    for line in out:
        if line:
            a, b, v = line.split()
            got[(a, b)] = int(v)
    for cname, py in STRUCTS.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)


def test_constants_match_header():
    text = open(HEADER).read()
    consts = dict((k, int(v)) for k, v in re.findall(r"#define (GXT_[A-Z0-9_]+) \(?(-?\d+)\)?", text))
    for name, v in abi.TAB.items():
        assert consts["GXT_TAB_" + name] == v
    for name, v in abi.MODEL.items():
        assert consts["GXT_MODEL_" + name] == v
    assert consts["GXT_NTAB"] == abi.NTAB
    for code, name in abi.ERRORS.items():
        assert consts[name] == code


def test_validation_without_device():
    lib = C.CDLL(LIB)
    lib.gxt_last_error.restype = C.c_char_p
    cfg = abi.config_lead_block()
    cfg.n_zbins = 0
    ctx = C.c_void_p()
    assert lib.gxt_create(C.byref(cfg), C.byref(ctx)) == -1
    assert b"bins" in lib.gxt_last_error()
    assert lib.gxt_create(None, C.byref(ctx)) == -1
    d = abi.Config()
    assert lib.gxt_default_config(C.byref(d)) == 0
    assert d.tracking_cut == 1e-3 and d.pool_capacity == 2_000_000


@pytest.mark.parametrize("field,value,needle", [
    ("field_stepper", 2, b"stepper"),
    ("field_origin", (0.0, float("nan"), 0.0), b"field"),
    ("field", (float("inf"), 0.0, 0.0), b"field"),
    ("precision", 3, b"precision"),
    ("pool_capacity", 10, b"pool_capacity"),
])
def test_config_validation_without_device(field, value, needle):
    lib = C.CDLL(LIB)
    lib.gxt_last_error.restype = C.c_char_p
    cfg = abi.config_calorimeter(**{field: value})
    ctx = C.c_void_p()
    assert lib.gxt_create(C.byref(cfg), C.byref(ctx)) == -1
    assert needle in lib.gxt_last_error()


def test_field_step_argument_checks_without_device():
    """Argument checks of the field stepper come before any device call."""
    lib = C.CDLL(LIB)
    lib.gxt_last_error.restype = C.c_char_p
    pos = np.zeros(3)
    d = np.array([1.0, 0.0, 0.0])
    f = (C.c_double * 3)(0.0, 0.0, 4.0)

    def step(stepper=abi.STEPPER_RK45, charge=1.0, mom=100.0, length=1.0, field=f):
        q, p, s = np.array([charge]), np.array([mom]), np.array([length])
        return lib.gxt_field_step(None, field, None, stepper, abi.FP64, abi.ptr(q), abi.ptr(p), abi.ptr(s),
                                  abi.ptr(pos), abi.ptr(d), None, 1)

    assert step(stepper=5) == -1 and b"stepper" in lib.gxt_last_error()
    assert step(charge=0.5) == -1 and b"charge" in lib.gxt_last_error()
    assert step(mom=0.0) == -1 and b"momentum" in lib.gxt_last_error()
    assert step(length=float("nan")) == -1
    assert step(mom=1.0, length=1e9) == -1 and b"2^16" in lib.gxt_last_error()
    assert step(field=None) == -1
    assert lib.gxt_field_map_value(None, abi.ptr(pos), abi.ptr(pos), 1) == -1
    assert lib.gxt_field_map_free(None) == -1


def test_host_gun_matches_oracle():
    import oracle_harness as O
    lib = C.CDLL(LIB)
    n = 50
    a = abi.primaries_array(n)
    assert lib.gxt_make_primaries(C.c_uint32(2468), 1, C.c_double(1000.0), C.c_double(50000.0), C.c_double(10.0),
                                  n, C.c_uint64(0), C.c_uint64(n), a) == 0
    b = O.make_primaries(2468, 1, 1000.0, 50000.0, 10.0, n)
    for x, y in zip(a, b):
        assert x.ekin == y.ekin and list(x.pos) == list(y.pos) and x.substream == y.substream
    e = [x.ekin for x in a]
    assert 1000.0 <= min(e) and max(e) <= 50000.0


def test_submit_checks_primaries_without_device():
    """gxt_submit rejects a bad primary before anything is queued (and before
    it needs a context): non-finite or non-positive energy, non-finite
    position or direction, a direction that is not a unit vector, a bad
    species or a negative event index."""
    lib = C.CDLL(LIB)
    lib.gxt_last_error.restype = C.c_char_p

    def submit(**kw):
        p = abi.primaries_array(2)
        for q in p:
            q.species, q.ekin = abi.ELECTRON, 100.0
            q.dir[:] = (0.0, 0.0, 1.0)
        for k, v in kw.items():
            if k in ("pos", "dir"):
                getattr(p[1], k)[:] = v
            else:
                setattr(p[1], k, v)
        return lib.gxt_submit(None, p, 2, C.c_uint64(1))

    cases = [(dict(ekin=float("inf")), b"ekin"), (dict(ekin=0.0), b"ekin"), (dict(ekin=float("nan")), b"ekin"),
             (dict(pos=(float("nan"), 0.0, 0.0)), b"finite"), (dict(dir=(0.0, float("inf"), 0.0)), b"finite"),
             (dict(dir=(0.0, 0.0, 2.0)), b"unit"), (dict(dir=(0.0, 0.0, 1.0 + 1e-5)), b"unit"),
             (dict(species=3), b"species"), (dict(event=-1), b"event")]
    for kw, needle in cases:
        assert submit(**kw) == -1, kw
        msg = lib.gxt_last_error()
        assert needle in msg and b"primary 1" in msg, (kw, msg)
    # good primaries pass the checks and stop at the missing context
    assert submit() == -1 and b"null context" in lib.gxt_last_error()
    assert submit(dir=(0.6, 0.0, 0.8)) == -1 and b"null context" in lib.gxt_last_error()


def test_submit_tracks_checks_tracks_without_device():
    """gxt_submit_tracks rejects a bad track in flight before anything is
    queued, and what a track carries before it needs a context: the
    kinematic checks of a primary, an event outside [0, n_events), a
    non-finite number of interaction lengths, a negative step count, an
    MRG32k3a state with a word at or above its modulus or an all-zero
    triple."""
    lib = C.CDLL(LIB)
    lib.gxt_last_error.restype = C.c_char_p

    def submit(n_events=2, **kw):
        t = abi.tracks_array(2)
        for q in t:
            q.species, q.ekin, q.volume = abi.GAMMA, 5.0, -1
            q.dir[:] = (0.0, 0.0, 1.0)
            q.nil[:] = (-1.0, 0.3, 2.0)
            q.rng[:] = (1, 2, 3, 4, 5, 6)
            q.event = 1
        for k, v in kw.items():
            if k in ("pos", "dir", "nil", "rng"):
                getattr(t[1], k)[:] = v
            else:
                setattr(t[1], k, v)
        return lib.gxt_submit_tracks(None, t, 2, n_events, C.c_uint64(1))

    cases = [(dict(ekin=float("inf")), b"ekin"), (dict(ekin=-1.0), b"ekin"),
             (dict(pos=(0.0, float("nan"), 0.0)), b"finite"), (dict(dir=(0.0, 0.0, 1.5)), b"unit"),
             (dict(species=-1), b"species"), (dict(event=2), b"n_events"), (dict(event=-3), b"event"),
             (dict(nil=(0.0, float("nan"), 1.0)), b"interaction lengths"), (dict(nsteps=-1), b"nsteps"),
             (dict(rng=(4294967087, 1, 1, 1, 1, 1)), b"MRG32k3a"), (dict(rng=(1, 1, 1, 1, 4294944443, 1)), b"MRG32k3a"),
             (dict(rng=(0, 0, 0, 1, 1, 1)), b"MRG32k3a"), (dict(rng=(1, 0, 0, 0, 0, 0)), b"MRG32k3a")]
    for kw, needle in cases:
        assert submit(**kw) == -1, kw
        msg = lib.gxt_last_error()
        assert needle in msg and b"track 1" in msg, (kw, msg)
    assert submit(n_events=0) == -1 and b"bad submit_tracks arguments" in lib.gxt_last_error()
    # good tracks pass the checks and stop at the missing context
    assert submit() == -1 and b"null context" in lib.gxt_last_error()
    assert submit(rng=(4294967086, 0, 0, 0, 0, 4294944442), nsteps=7) == -1
    assert b"null context" in lib.gxt_last_error()
