"""Python host interface of libgxt.so (the MI355X step loop).

Mirrors the reference's device-manager flow (ProxyDeviceManager:
UploadTrackState -> DoStep -> DownloadTrackData, proxy/ProxyDeviceManager.hpp:46-65)
as Transport.submit -> Transport.run -> Transport.event_tallies / profiles,
and exposes the component entry points (RNG, tables, samplers, helix) that
the parity tests drive.  Every call goes through the C ABI of include/gxt.h;
there is no Python or CPU fallback: if libgxt.so is missing this module
raises at import.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgxt.so")

if not os.path.exists(LIB_PATH):
    raise ImportError("libgxt.so not built (run __graft_entry__.build() or geantexascalepilot_amd/build.py)")

lib = C.CDLL(LIB_PATH)
lib.gxt_last_error.restype = C.c_char_p
lib.gxt_ctx_error.restype = C.c_char_p
lib.gxt_now_ms.restype = C.c_double
lib.gxt_now_ms.argtypes = [C.c_void_p]


class GxtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (abi.ERRORS.get(code, "error"), code, msg))
        self.code = code


def _check(rc, ctx=None):
    if rc != 0:
        msg = lib.gxt_ctx_error(ctx).decode() if ctx else lib.gxt_last_error().decode()
        raise GxtError(rc, msg)


def make_primaries(seed, species, emin, emax, spot_mm, n, first=0, n_total=None):
    """Primary gun (ProxyEgammaGun::GenerateOneEvent, proxy/src/ProxyEgammaGun.cpp:43-85)."""
    p = abi.primaries_array(n)
    _check(lib.gxt_make_primaries(C.c_uint32(seed), species, C.c_double(emin), C.c_double(emax),
                                  C.c_double(spot_mm), n, C.c_uint64(first),
                                  C.c_uint64(n if n_total is None else n_total), p))
    return p


def read_hepevt(path, max_events=-1, vertex=(0.0, 0.0, 0.0), first=0):
    """HEPEVT file -> (primaries, n_events, n_unsupported) (ProxyHepEvtGenerator,
    proxy/src/ProxyHepEvtGenerator.cpp:48-132; contract in include/gxt.h)."""
    v = (C.c_double * 3)(*vertex)
    n_p, n_e, n_u = C.c_int(0), C.c_int(0), C.c_int(0)
    path_b = os.fsencode(path)
    _check(lib.gxt_read_hepevt(path_b, max_events, v, C.c_uint64(first), None, 0, C.byref(n_p), C.byref(n_e),
                               C.byref(n_u)))
    p = abi.primaries_array(n_p.value)
    _check(lib.gxt_read_hepevt(path_b, max_events, v, C.c_uint64(first), p, n_p.value, C.byref(n_p), C.byref(n_e),
                               C.byref(n_u)))
    return p, n_e.value, n_u.value


def egamma_gun(seed, n_events, per_event, gamma_fraction=0.5, pmin=1000.0, pmax=1.0e6, eta=(-3.0, 3.0),
               phi=(0.0, 2.0 * 3.141592653589793), vertex=(0.0, 0.0, 0.0), first=0, n_total=None):
    """Egamma gun (ProxyEgammaGun::GenerateOneEvent, proxy/src/ProxyEgammaGun.cpp:43-85)."""
    n = n_events * per_event
    p = abi.primaries_array(n)
    v = (C.c_double * 3)(*vertex)
    _check(lib.gxt_egamma_gun(C.c_uint32(seed), n_events, per_event, C.c_double(gamma_fraction), C.c_double(pmin),
                              C.c_double(pmax), C.c_double(eta[0]), C.c_double(eta[1]), C.c_double(phi[0]),
                              C.c_double(phi[1]), v, C.c_uint64(first), C.c_uint64(n if n_total is None else n_total),
                              p))
    return p


def rng_init(seed6, substreams):
    seed6 = np.ascontiguousarray(seed6, dtype=np.uint32)
    idx = np.ascontiguousarray(substreams, dtype=np.uint64)
    out = np.zeros((len(idx), 6), dtype=np.uint32)
    _check(lib.gxt_rng_init(abi.ptr(seed6, C.c_uint32), abi.ptr(idx, C.c_uint64), len(idx),
                            abi.ptr(out, C.c_uint32)))
    return out


def rng_draw(states, k):
    st = np.ascontiguousarray(states, dtype=np.uint32).copy()
    n = st.shape[0]
    out = np.zeros((n, k))
    _check(lib.gxt_rng_draw(abi.ptr(st, C.c_uint32), n, k, abi.ptr(out)))
    return out, st


def helix_step(field, charge, mom, step, pos, dirs, precision=abi.FP64):
    field = np.ascontiguousarray(field, dtype=np.float64)
    charge = np.ascontiguousarray(charge, dtype=np.float64)
    mom = np.ascontiguousarray(mom, dtype=np.float64)
    step = np.ascontiguousarray(step, dtype=np.float64)
    pos = np.ascontiguousarray(pos, dtype=np.float64).copy()
    dirs = np.ascontiguousarray(dirs, dtype=np.float64).copy()
    _check(lib.gxt_helix_step(abi.ptr(field), precision, abi.ptr(charge), abi.ptr(mom), abi.ptr(step),
                              abi.ptr(pos), abi.ptr(dirs), len(charge)))
    return pos, dirs


class FieldMap:
    """A cylindrical (r, z) field map resident on the current HIP device."""

    def __init__(self, path=abi.CMS_FIELD_MAP):
        self.h = C.c_void_p()
        _check(lib.gxt_field_map_load(os.fsencode(path), C.byref(self.h)))

    def info(self):
        nr, nz = C.c_int(), C.c_int()
        d = [C.c_double() for _ in range(5)]
        _check(lib.gxt_field_map_info(self.h, C.byref(nr), C.byref(nz), *[C.byref(x) for x in d]))
        return dict(nr=nr.value, nz=nz.value, r0=d[0].value, z0=d[1].value, dr=d[2].value, dz=d[3].value,
                    bmax=d[4].value)

    def value(self, pos):
        """Cartesian field (tesla) at points of the map frame; pos: [3, n] (SoA)."""
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        n = pos.shape[1]
        b = np.zeros((3, n))
        _check(lib.gxt_field_map_value(self.h, abi.ptr(pos), abi.ptr(b), n))
        return b

    def close(self):
        if self.h:
            _check(lib.gxt_field_map_free(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def hip_runtime():
    """The HIP runtime libgxt.so runs on (already loaded; found by its soname)."""
    h = C.CDLL("libamdhip64.so.7")
    h.hipGetErrorString.restype = C.c_char_p
    return h


class DeviceArray:
    """A device copy of a numpy array, allocated with the HIP runtime libgxt.so
    uses, so that it can be passed to the entry points that take device memory
    (ptr)."""

    def __init__(self, a):
        self.hip = hip_runtime()
        self.a = np.ascontiguousarray(a)
        self.ptr = C.c_void_p()
        self._ok(self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(max(self.a.nbytes, 1))))
        self._ok(self.hip.hipMemcpy(self.ptr, self.a.ctypes.data_as(C.c_void_p), C.c_size_t(self.a.nbytes), 1))

    def _ok(self, rc):
        if rc != 0:
            raise GxtError(-2, "HIP: " + self.hip.hipGetErrorString(rc).decode())

    def numpy(self):
        """Copies the device contents back (after the device is synchronised)."""
        self._ok(self.hip.hipDeviceSynchronize())
        out = np.empty_like(self.a)
        self._ok(self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(self.a.nbytes), 2))
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Navigator:
    """A volume tree of placed boxes resident on the current HIP device.

    volumes: [(half_xyz, material)]; placements: [(mother, daughter,
    translation_xyz)] (abi.reference_calorimeter_tree builds the reference's
    calorimeter).  Touchables are numbered breadth-first from the world (0)."""

    def __init__(self, volumes, placements, world=0):
        vol, pl = abi.box_tree(volumes, placements)
        self.h = C.c_void_p()
        _check(lib.gxt_nav_create(vol, len(volumes), pl, len(placements), world, C.byref(self.h)))

    def info(self):
        n = C.c_int()
        _check(lib.gxt_nav_info(self.h, C.byref(n), None, None, None))
        vol = np.zeros(n.value, dtype=np.int32)
        par = np.zeros(n.value, dtype=np.int32)
        cen = np.zeros((n.value, 3))
        _check(lib.gxt_nav_info(self.h, None, abi.ptr(vol, C.c_int32), abi.ptr(par, C.c_int32), abi.ptr(cen)))
        return dict(volume=vol, parent=par, centre=cen)

    def locate(self, pos, dirs):
        """Deepest touchable holding each point (-1 outside); pos, dirs: [3, n]."""
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        dirs = np.ascontiguousarray(dirs, dtype=np.float64)
        n = pos.shape[1]
        out = np.zeros(n, dtype=np.int32)
        _check(lib.gxt_nav_locate(self.h, abi.ptr(pos), abi.ptr(dirs), abi.ptr(out, C.c_int32), n))
        return out

    def step(self, touchable, pos, dirs):
        """(snext, next touchable, safety) of each point inside touchable[i]."""
        t = np.ascontiguousarray(touchable, dtype=np.int32)
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        dirs = np.ascontiguousarray(dirs, dtype=np.float64)
        n = len(t)
        s, saf = np.zeros(n), np.zeros(n)
        nxt = np.zeros(n, dtype=np.int32)
        _check(lib.gxt_nav_step(self.h, abi.ptr(t, C.c_int32), abi.ptr(pos), abi.ptr(dirs), abi.ptr(s),
                                abi.ptr(nxt, C.c_int32), abi.ptr(saf), n))
        return s, nxt, saf

    def find_next_boundary(self, touchable, pos, dirs, pstep):
        """FindNextBoundary: (snext, next touchable, safety, boundary flag) for proposed steps pstep."""
        t = np.ascontiguousarray(touchable, dtype=np.int32)
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        dirs = np.ascontiguousarray(dirs, dtype=np.float64)
        ps = np.ascontiguousarray(pstep, dtype=np.float64)
        n = len(t)
        s, saf = np.zeros(n), np.zeros(n)
        nxt = np.zeros(n, dtype=np.int32)
        b = np.zeros(n, dtype=np.int32)
        _check(lib.gxt_nav_find_next_boundary(self.h, abi.ptr(t, C.c_int32), abi.ptr(pos), abi.ptr(dirs), abi.ptr(ps), abi.ptr(s),
                 abi.ptr(nxt, C.c_int32), abi.ptr(saf), abi.ptr(b, C.c_int32), n))
        return s, nxt, saf, b

    def close(self):
        if self.h:
            _check(lib.gxt_nav_free(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def field_step(field_map, field, charge, mom, step, pos, dirs, stepper=abi.STEPPER_HELIX, precision=abi.FP64,
               origin=(0.0, 0.0, 0.0)):
    """One propagation sub-step per track in a map (field_map) or a uniform field; returns (pos, dir, err)."""
    charge = np.ascontiguousarray(charge, dtype=np.float64)
    mom = np.ascontiguousarray(mom, dtype=np.float64)
    step = np.ascontiguousarray(step, dtype=np.float64)
    pos = np.ascontiguousarray(pos, dtype=np.float64).copy()
    dirs = np.ascontiguousarray(dirs, dtype=np.float64).copy()
    n = len(charge)
    err = np.zeros(2 * n)
    f = None if field is None else (C.c_double * 3)(*field)
    o = (C.c_double * 3)(*origin)
    h = field_map.h if field_map is not None else None
    _check(lib.gxt_field_step(h, f, o, stepper, precision, abi.ptr(charge), abi.ptr(mom), abi.ptr(step),
                              abi.ptr(pos), abi.ptr(dirs), abi.ptr(err), n))
    return pos, dirs, err.reshape(2, n)


class Transport:
    """One device context: physics tables, track pool, tallies, one HIP stream."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.ctx = C.c_void_p()
        _check(lib.gxt_create(C.byref(cfg), C.byref(self.ctx)))
        nb = C.c_int()
        _check(lib.gxt_volume_bins(self.ctx, C.byref(nb), None))
        self.nvol = nb.value

    def close(self):
        if self.ctx:
            lib.gxt_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc):
        _check(rc, self.ctx)

    def submit(self, prims, job_id):
        self._c(lib.gxt_submit(self.ctx, prims, len(prims), C.c_uint64(job_id)))

    def submit_tracks(self, tracks, n_events, job_id):
        """A basket of tracks in flight (abi.Track array) over n_events events."""
        self._c(lib.gxt_submit_tracks(self.ctx, tracks, len(tracks), n_events, C.c_uint64(job_id)))

    def run(self, max_iterations=0):
        self._c(lib.gxt_run(self.ctx, C.c_int64(max_iterations)))

    def poll(self, max_jobs=256):
        out = (abi.JobResult * max_jobs)()
        n = C.c_int(0)
        self._c(lib.gxt_poll(self.ctx, out, max_jobs, C.byref(n)))
        return [out[i] for i in range(n.value)]

    def event_tallies(self, job_id, n):
        t = (abi.EventTally * n)()
        self._c(lib.gxt_get_event_tallies(self.ctx, C.c_uint64(job_id), t, n))
        return abi.tallies_to_numpy(t)

    def profiles(self):
        ez = np.zeros(self.cfg.n_zbins)
        er = np.zeros(self.cfg.n_rbins)
        ev = np.zeros(self.nvol)
        self._c(lib.gxt_get_profiles(self.ctx, abi.ptr(ez), abi.ptr(er), abi.ptr(ev)))
        return ez, er, ev

    def spectra(self):
        s = np.zeros(3 * 64, dtype=np.int64)
        self._c(lib.gxt_get_spectra(self.ctx, abi.ptr(s, C.c_int64)))
        return s.reshape(3, 64)

    def stats(self):
        s = abi.Stats()
        self._c(lib.gxt_get_stats(self.ctx, C.byref(s)))
        return s

    def kernel_stats(self):
        s = abi.KernelStats()
        self._c(lib.gxt_get_kernel_stats(self.ctx, C.byref(s)))
        return s

    def reset_tallies(self):
        self._c(lib.gxt_reset_tallies(self.ctx))

    def trace(self):
        n = C.c_int64()
        self._c(lib.gxt_get_trace(self.ctx, None, 0, C.byref(n)))
        out = np.zeros((max(n.value, 1), 12))
        self._c(lib.gxt_get_trace(self.ctx, abi.ptr(out), n.value, C.byref(n)))
        return out[: n.value]

    def now_ms(self):
        return lib.gxt_now_ms(self.ctx)

    def serve_start(self):
        self._c(lib.gxt_serve_start(self.ctx))

    def serve_stop(self):
        self._c(lib.gxt_serve_stop(self.ctx))

    def table_value(self, slot, table, x, precision=abi.FP64):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros_like(x)
        self._c(lib.gxt_table_value(self.ctx, slot, table, precision, abi.ptr(x), abi.ptr(out), len(x)))
        return out

    def macro_xs(self, slot, species, process, e, precision=abi.FP64):
        e = np.ascontiguousarray(e, dtype=np.float64)
        out = np.zeros_like(e)
        self._c(lib.gxt_macro_xs(self.ctx, slot, species, process, precision, abi.ptr(e), abi.ptr(out), len(e)))
        return out

    def sample(self, model, slot, ekin, dirs, states, precision=abi.FP64, noalloc=False):
        """Final-state sampler of a model on n tracks; noalloc stages through the
        library's fixed device array (gxt_sample_noalloc, n <= SAMPLE_NOALLOC_MAX)."""
        fn = lib.gxt_sample_noalloc if noalloc else lib.gxt_sample
        return _sample_with(lambda io: fn(self.ctx, model, slot, precision, C.byref(io)), ekin, dirs, states)

    def slab_distance(self, pos, dirs, vol=None, pstep=None):
        """Straight-line navigation of the loop's geometry (FindNextBoundary,
        geometry/NavigationInterface.hpp:64-98): returns (vol, snext, next, safety,
        boundary) for points pos, dirs ([3, n] SoA); vol None locates the points."""
        return _slab_with(lambda *a: lib.gxt_slab_distance(self.ctx, *a), pos, dirs, vol, pstep, self._c)


def _slab_with(call, pos, dirs, vol, pstep, check):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    dirs = np.ascontiguousarray(dirs, dtype=np.float64)
    n = pos.shape[1]
    v = np.full(n, -1, dtype=np.int32) if vol is None else np.ascontiguousarray(vol, dtype=np.int32).copy()
    ps = None if pstep is None else np.ascontiguousarray(pstep, dtype=np.float64)
    s, saf = np.zeros(n), np.zeros(n)
    nxt = np.zeros(n, dtype=np.int32)
    b = np.zeros(n, dtype=np.int32)
    check(call(abi.ptr(pos), abi.ptr(dirs), None if ps is None else abi.ptr(ps), abi.ptr(v, C.c_int32), abi.ptr(s),
               abi.ptr(nxt, C.c_int32), abi.ptr(saf), abi.ptr(b, C.c_int32), n))
    return v, s, nxt, saf, b


def _sample_with(call, ekin, dirs, states):
    n = len(ekin)
    ekin = np.ascontiguousarray(ekin, dtype=np.float64)
    dirs = np.ascontiguousarray(dirs, dtype=np.float64).reshape(3, n)
    st = np.ascontiguousarray(states, dtype=np.uint32).reshape(n, 6).copy()
    out = dict(ekin=np.zeros(n), dir=np.zeros((3, n)), edep=np.zeros(n), nsec=np.zeros(n, dtype=np.int32),
               sec_species=np.zeros((n, 2), dtype=np.int32), sec_ekin=np.zeros((n, 2)),
               sec_dir=np.zeros((n, 2, 3)))
    io = abi.SampleIO()
    io.n = n
    io.ekin = abi.ptr(ekin)
    io.dir = abi.ptr(dirs)
    io.rng = abi.ptr(st, C.c_uint32)
    io.out_ekin = abi.ptr(out["ekin"])
    io.out_dir = abi.ptr(out["dir"])
    io.out_edep = abi.ptr(out["edep"])
    io.nsec = abi.ptr(out["nsec"], C.c_int32)
    io.sec_species = abi.ptr(out["sec_species"], C.c_int32)
    io.sec_ekin = abi.ptr(out["sec_ekin"])
    io.sec_dir = abi.ptr(out["sec_dir"])
    _check(call(io))
    out["rng"] = st
    return out
