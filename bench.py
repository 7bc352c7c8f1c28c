"""Benchmark of the MI355X EM step loop on one GPU.

Workload (the headline configuration): 10,000 primary electrons of 10 GeV
along +z into a 30x30x30 cm PbWO4 block (8.28 g/cm3), no field, Compton,
photoelectric, conversion, e-/e+ bremsstrahlung, ionisation and e+
annihilation, 0.7 mm production cut, 1 keV tracking cut, live-track pool of
2e6 tracks, MRG32k3a seed 777 with one substream per primary.  A "step" of
this benchmark is one transport of the whole batch (every shower run to
completion); value = events/s of the FP64 form.  The same line reports
track-steps/s, the FP32 precision form on the same batch, a short serving
run (Poisson job mix, FP32, p50/p99 job latency), the roofline of the
dominant kernel from HIP events around each launch, and the CPU oracle
timed on the host cores on a bounded sample.

python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
ME = 0.510998910


def profiled_traffic(kernel, summary="hbm_traffic_c1_fp64.md"):
    """HBM bytes per launch of `kernel` as measured by the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of its workload (calibrated against a copy
    kernel of known size); None when the summary does not list the kernel."""
    import re
    path = os.path.join(ROOT, "profiles", summary)
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"^\| %s \|(?:[^|\n]*\|){3}\s*([0-9.eE+]+)\s*\|" % re.escape(kernel), text, re.M)
    return float(m.group(1)) if m else None


def log(msg):
    sys.stderr.write("[bench] %s\n" % msg)
    sys.stderr.flush()


def run_batch(g, abi, precision, n_events, steps, warmup, seed=777, pool=2_000_000):
    cfg = abi.config_pbwo4_block(max_events=n_events, precision=precision, profile_kernels=2,
                                 pool_capacity=pool, seed=seed)
    prims = g.make_primaries(seed, abi.ELECTRON, 10000.0, 10000.0, 0.0, n_events)
    import torch
    with g.Transport(cfg) as t:
        job = 0
        for w in range(warmup):
            t.submit(prims, job)
            t.run()
            job += 1
            log("warmup %d/%d done (precision %d)" % (w + 1, warmup, precision))
        t.reset_tallies()
        # each batch is submitted (its primaries copied to HBM) before its
        # timed region; the timed region is the transport of the batch
        elapsed = 0.0
        for k in range(steps):
            t.submit(prims, job)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            t.run()
            torch.cuda.synchronize()
            elapsed += time.perf_counter() - t0
            job += 1
        ks = t.kernel_stats()
        st = t.stats()
        edep = t.event_tallies(job - 1, n_events)
    log("timed %d batches in %.3f s (precision %d): %d iterations, peak %d live tracks"
        % (steps, elapsed, precision, st.iterations, st.max_live))
    return dict(seconds=elapsed, ks=ks, stats=st, edep=edep)


def precision_form(g, abi, n_events=20000, seed=2468):
    """FP32 precision form against FP64 on the sampling calorimeter: 50 x (2.3 mm
    Pb + 5.7 mm lAr), 40x40 cm, 4 T along +x, n_events electrons of 1-50 GeV
    (uniform) started over a 1 cm square at the front face.  Per-observable
    error of the FP32 run against the FP64 run on the same primaries."""
    import numpy as np
    prims = g.make_primaries(seed, abi.ELECTRON, 1000.0, 50000.0, 10.0, n_events)
    res = {}
    for prec in (abi.FP64, abi.FP32):
        cfg = abi.config_calorimeter(precision=prec, max_events=n_events, n_zbins=40, zbin_mm=10.0)
        with g.Transport(cfg) as t:
            t0 = time.perf_counter()
            t.submit(prims, 1)
            t.run()
            dt = time.perf_counter() - t0
            ev = t.event_tallies(1, n_events)
            ez, er, evol = t.profiles()
            st = t.stats()
        zc = (np.arange(len(ez)) + 0.5) * 10.0
        res[prec] = dict(seconds=dt, visible=ev["visible"], layers=evol[1::2][:50], ez=ez,
                         shower_max_mm=float(zc[np.argmax(ez)]), centroid_mm=float((zc * ez).sum() / ez.sum()),
                         steps=int(st.track_steps))
    a, b = res[abi.FP64], res[abi.FP32]
    lay = (b["layers"] - a["layers"]) / a["layers"]
    log("precision form: FP64 %.2f s, FP32 %.2f s on %d calorimeter events" % (a["seconds"], b["seconds"], n_events))
    return dict(workload="e- 1-50 GeV x %d into 50 x (2.3 mm Pb + 5.7 mm lAr), 4 T along +x" % n_events,
                fp64_events_per_s=n_events / a["seconds"], fp32_events_per_s=n_events / b["seconds"],
                visible_mean_mev=dict(fp64=float(a["visible"].mean()), fp32=float(b["visible"].mean()),
                                      rel_err=float(b["visible"].mean() / a["visible"].mean() - 1)),
                visible_rms_mev=dict(fp64=float(a["visible"].std()), fp32=float(b["visible"].std()),
                                     rel_err=float(b["visible"].std() / a["visible"].std() - 1)),
                layer_deposit_rel_err=dict(max_abs=float(np.abs(lay).max()), rms=float(np.sqrt((lay ** 2).mean()))),
                shower_max_depth_mm=dict(fp64=a["shower_max_mm"], fp32=b["shower_max_mm"]),
                shower_centroid_mm=dict(fp64=a["centroid_mm"], fp32=b["centroid_mm"],
                                        diff=b["centroid_mm"] - a["centroid_mm"]),
                track_steps=dict(fp64=a["steps"], fp32=b["steps"]))


def field_forms(g, abi, n_events=20000, seed=2468, cpu=False):
    """The calorimeter workload of the precision form (FP64) in three fields:
    uniform 4 T along +x with the exact helix; the CMS 2015 solenoid map with
    the calorimeter's front face at r = 600 mm, z = 6 m, with the helix in the
    local field and with Cash-Karp RK45; the uniform field with RK45; and the
    uniform field with multiple scattering of e-/e+.  Throughput and device
    time per track-step of each."""
    prims = g.make_primaries(seed, abi.ELECTRON, 1000.0, 50000.0, 10.0, n_events)
    kw = dict(max_events=n_events, profile_kernels=2)
    forms = [("uniform_helix", lambda: abi.config_calorimeter(**kw)),
             ("cms_map_helix", lambda: abi.config_calorimeter_cms(stepper=abi.STEPPER_HELIX, **kw)),
             ("cms_map_rk45", lambda: abi.config_calorimeter_cms(stepper=abi.STEPPER_RK45, **kw)),
             ("uniform_rk45", lambda: abi.config_calorimeter(field_stepper=abi.STEPPER_RK45, **kw)),
             ("uniform_helix_msc", lambda: abi.config_calorimeter(msc=1, **kw))]
    out = dict(workload="e- 1-50 GeV x %d into 50 x (2.3 mm Pb + 5.7 mm lAr), FP64" % n_events)
    for name, make in forms:
        with g.Transport(make()) as t:
            t0 = time.perf_counter()
            t.submit(prims, 1)
            t.run()
            dt = time.perf_counter() - t0
            ev = t.event_tallies(1, n_events)
            st = t.stats()
            ks = t.kernel_stats()
        # the e-/e+ transport against the HBM roofline, as for the headline
        # (2 x 148 B per e-/e+ step over its HIP-event-timed launches)
        gbs = 2 * 148 * ks.charged_steps / max(ks.transport_charged_ms * 1e-3, 1e-12) / 1e9
        out[name] = dict(events_per_s=n_events / dt, seconds=dt, track_steps=int(st.track_steps),
                         ns_per_track_step=1e9 * dt / max(int(st.track_steps), 1),
                         visible_mean_mev=float(ev["visible"].mean()),
                         charged_transport=dict(kernel_ms=ks.transport_charged_ms, launches=int(ks.launches),
                                                steps=int(ks.charged_steps), achieved_gbs=gbs,
                                                frac=gbs / PEAK_HBM_GBS))
        log("field %s: %.2f s, %d track-steps" % (name, dt, st.track_steps))
    if cpu:
        # the oracle (FP64, host threads) on the first primaries of the same batch in the map
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_harness as O
        threads = max(1, min(16, os.cpu_count() or 1))
        n = 16 * threads
        o = O.Oracle(abi.config_calorimeter_cms(stepper=abi.STEPPER_HELIX, max_events=n))
        p = O.make_primaries(seed, abi.ELECTRON, 1000.0, 50000.0, 10.0, n)
        t0 = time.perf_counter()
        o.run(p, threads)
        dt = time.perf_counter() - t0
        steps = int(o.stats().track_steps)
        o.close()
        out["cpu_oracle_cms_map_helix"] = dict(events_per_s=n / dt, cores=threads, track_steps_per_s=steps / dt,
                                               sample="first %d primaries, %d track-steps, %.1f s" % (n, steps, dt))
        log("field cpu oracle: %d events on %d threads in %.2f s" % (n, threads, dt))
    return out


def tree_form(g, abi, n_events=10000, cpu=False):
    """Transport through the volume tree of the reference's calorimeter (50
    layers of 10 mm Pb + 10 mm lAr along x, 1 m square, in a vacuum world;
    ProxyDetectorConstruction.cpp:38-60, 225-325) with the gun of its
    transport test (10 GeV e-, position in [0, 1) mm per coordinate,
    direction in the positive octant; test_BasicCpuTransport.cpp:556-586),
    no field and 4 T along z, FP64 and FP32."""
    prims = abi.reference_calorimeter_primaries(n_events)
    out = dict(workload="10 GeV e- x %d into the reference calorimeter as a volume tree (152 touchables)" % n_events)
    for name, kw in (("no_field_fp64", {}), ("no_field_fp32", dict(precision=abi.FP32)),
                     ("4T_fp64", dict(field=(0.0, 0.0, 4.0))), ("4T_fp32", dict(field=(0.0, 0.0, 4.0), precision=abi.FP32))):
        cfg = abi.config_reference_calorimeter(max_events=n_events, **kw)
        with g.Transport(cfg) as t:
            t0 = time.perf_counter()
            t.submit(prims, 1)
            t.run()
            dt = time.perf_counter() - t0
            ev = t.event_tallies(1, n_events)
            st = t.stats()
        out[name] = dict(events_per_s=n_events / dt, seconds=dt, track_steps_per_s=int(st.track_steps) / dt,
                         visible_mean_mev=float(ev["visible"].mean()), edep_mean_mev=float(ev["edep"].mean()))
        log("tree %s: %.2f s, %d track-steps" % (name, dt, st.track_steps))
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_harness as O
        threads = max(1, min(16, os.cpu_count() or 1))
        n = 32 * threads
        o = O.Oracle(abi.config_reference_calorimeter(max_events=n))
        t0 = time.perf_counter()
        o.run(abi.reference_calorimeter_primaries(n), threads)
        dt = time.perf_counter() - t0
        steps = int(o.stats().track_steps)
        o.close()
        out["cpu_oracle_no_field"] = dict(events_per_s=n / dt, cores=threads, track_steps_per_s=steps / dt,
                                          sample="first %d primaries, %d track-steps, %.1f s" % (n, steps, dt))
    return out


def navigation(g, abi, n=1 << 24, reps=20, cpu=False, seed=99):
    """Volume-tree navigation in the reference's calorimeter (50 layers of
    10 mm Pb + 10 mm lAr, 1 m square, 152 touchables): n points uniform in the
    calorimeter with isotropic directions, resident in HBM, located once, then
    the boundary step (distance, next touchable, safety) of every point, timed
    with HIP events on the null stream, where the launches go.  Algorithmic
    bytes per query: position and direction 48 + touchable 4 read, distance
    and safety 16 + next touchable 4 written = 72 B."""
    import ctypes as C
    import numpy as np
    rs = np.random.default_rng(seed)
    pos = rs.uniform(-500.0, 500.0, (3, n))
    d = rs.normal(size=(3, n))
    d /= np.linalg.norm(d, axis=0)
    nav = g.Navigator(*abi.reference_calorimeter_tree())
    D = g.DeviceArray
    dp, dd = D(pos), D(d)
    dt, dn = D(np.zeros(n, dtype=np.int32)), D(np.zeros(n, dtype=np.int32))
    ds, dsaf = D(np.zeros(n)), D(np.zeros(n))
    hip = g.hip_runtime()
    ev = [C.c_void_p() for _ in range(4)]
    for e in ev:
        assert hip.hipEventCreate(C.byref(e)) == 0

    def locate():
        g._check(g.lib.gxt_nav_locate(nav.h, dp.ptr, dd.ptr, dt.ptr, n))

    def step():
        g._check(g.lib.gxt_nav_step(nav.h, dt.ptr, dp.ptr, dd.ptr, ds.ptr, dn.ptr, dsaf.ptr, n))

    def timed(fn, e0, e1):
        fn()
        assert hip.hipDeviceSynchronize() == 0
        assert hip.hipEventRecord(e0, None) == 0
        for _ in range(reps):
            fn()
        assert hip.hipEventRecord(e1, None) == 0
        assert hip.hipEventSynchronize(e1) == 0
        ms = C.c_float()
        assert hip.hipEventElapsedTime(C.byref(ms), e0, e1) == 0
        return 1e3 * ms.value / reps

    loc_us = timed(locate, ev[0], ev[1])
    us = timed(step, ev[2], ev[3])
    nxt, s = dn.numpy(), ds.numpy()
    achieved = 72.0 * n / (us * 1e-6) / 1e9
    out = dict(workload="%d points uniform in the reference calorimeter (50 x (10 mm Pb + 10 mm lAr), 1 m "
                        "square, 152 touchables), isotropic directions, FP64" % n,
               step_queries_per_s=n / (us * 1e-6), step_launch_us=us, locate_queries_per_s=n / (loc_us * 1e-6),
               locate_launch_us=loc_us, mean_snext_mm=float(s.mean()), into_world=int((nxt == 0).sum()),
               roofline=dict(bound="hbm", kernel="k_nav_step", achieved=achieved, peak=PEAK_HBM_GBS, unit="GB/s",
                             frac=achieved / PEAK_HBM_GBS, per_unit="72 B per query",
                             traffic=profiled_traffic("k_nav_step", "hbm_traffic_nav_reference_calorimeter.md")
                             if n == 1 << 24 else None))
    log("navigation: step %.1f us, locate %.1f us for %d points" % (us, loc_us, n))
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_harness as O
        m = 1 << 22
        o = O.Navigator(*abi.reference_calorimeter_tree())
        ht = dt.numpy()[:m]
        t0 = time.perf_counter()
        o.step(ht, pos[:, :m], d[:, :m])
        dt_s = time.perf_counter() - t0
        out["cpu_oracle"] = dict(step_queries_per_s=m / dt_s, cores=1,
                                 sample="first %d points, one thread, %.2f s" % (m, dt_s))
    for e in ev:
        hip.hipEventDestroy(e)
    for a in (dp, dd, dt, dn, ds, dsaf):
        a.free()
    nav.close()
    return out


MIXES = {
    # kind: (probability, events per job, species, energy MeV)
    "c4": [("gamma1", 0.4, 1, 0, 1000.0), ("e10x10", 0.4, 10, 1, 10000.0), ("e100x50", 0.2, 100, 1, 50000.0)],
    # 10 GeV e- in jobs of 100 events: the headline workload as a stream (the
    # job size keeps the submitting thread far below its limit of a few
    # thousand jobs/s)
    "e10": [("e10x100", 1.0, 100, 1, 10000.0)],
}


def mix_gev_per_job(mix):
    return sum(p * n * e * 1e-3 for _, p, n, _, e in MIXES[mix])


def serving(g, abi, seconds, rate_jobs_s, mix="c4", warm_s=2.0, drain_s=60.0, seed=4242):
    """Poisson stream of jobs into the PbWO4 block, FP32 form, one shared pool,
    submitted for `seconds`; mix "c4": 40 % one 1 GeV gamma, 40 % 10 x 10 GeV
    e-, 20 % 100 x 50 GeV e-; mix "e10": every job 100 x 10 GeV e-.  The
    throughput is the completion rate inside the submission window after
    warm_s (so it saturates at the capacity when the offered rate exceeds
    it); latencies are those of the jobs submitted after warm_s, with a
    drain of at most drain_s after the window."""
    import numpy as np
    rs = np.random.default_rng(seed)
    cfg = abi.config_pbwo4_block(max_events=40000, precision=abi.FP32, max_jobs=4096, pool_capacity=8_000_000)
    kinds, sub_t, done = [], [], []
    with g.Transport(cfg) as t:
        t.serve_start()
        t_start = time.perf_counter()
        next_sub, first, submitted = 0.0, 0, 0
        next_log = 60.0
        while True:
            now = time.perf_counter() - t_start
            if now >= seconds:
                break
            # every job whose arrival time has passed is submitted now, so the
            # arrivals keep the offered rate when a pass of this loop is slow
            while now >= next_sub:
                u, acc = rs.random(), 0.0
                for kind, p, n, sp, e in MIXES[mix]:
                    acc += p
                    if u < acc:
                        break
                prims = g.make_primaries(777, sp, e, e, 0.0, n, first=first, n_total=1 << 40)
                first += n
                t.submit(prims, submitted)
                kinds.append(kind)
                sub_t.append(now)
                submitted += 1
                next_sub += rs.exponential(1.0 / rate_jobs_s)
            for r in t.poll():
                done.append((time.perf_counter() - t_start, r))
            if now >= next_log:
                log("serving %s: %.0f s, %d jobs submitted, %d completed" % (mix, now, submitted, len(done)))
                next_log += 60.0
            time.sleep(0.0002)
        t_window = time.perf_counter() - t_start
        while len(done) < submitted and time.perf_counter() - t_start < seconds + drain_s:
            for r in t.poll():
                done.append((time.perf_counter() - t_start, r))
            time.sleep(0.002)
        st = t.stats()
        t.serve_stop()
    inwin = [r for tc, r in done if warm_s <= tc <= t_window]
    events = sum(r.n_events for r in inwin)
    steps = sum(r.track_steps for r in inwin)
    win = t_window - warm_s
    late = [r for tc, r in done if sub_t[r.job_id] >= warm_s]
    lat = np.array([r.done_ms - r.submit_ms for r in late]) if late else np.zeros(1)
    single = np.array([r.done_ms - r.submit_ms for r in late if kinds[r.job_id] == "gamma1"])
    backlog = sum(1 for j in range(submitted) if sub_t[j] <= t_window) - sum(1 for tc, _ in done if tc <= t_window)
    # a job's latency over the step-loop iterations it spent in flight: the
    # host finds a finished job at the next counter snapshot, at most one
    # round (4-8 iterations) after its last track died
    its = np.array([max(r.iterations, 1) for r in late]) if late else np.ones(1)
    ms_per_it = float(np.median(lat / its))
    single_its = np.array([r.iterations for r in late if kinds[r.job_id] == "gamma1"])
    return dict(mix=mix, offered_jobs_per_s=rate_jobs_s, offered_gev_per_s=rate_jobs_s * mix_gev_per_job(mix),
                seconds=seconds, jobs_submitted=submitted, jobs_completed=len(done), backlog_at_window_end=backlog,
                events_per_s=events / win, track_steps_per_s=steps / win, gev_per_s=events / win * (
                    sum(p * n * e for _, p, n, _, e in MIXES[mix]) / sum(p * n for _, p, n, _, _ in MIXES[mix])) * 1e-3,
                p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                single_event_p50_ms=float(np.percentile(single, 50)) if len(single) else None,
                single_event_p99_ms=float(np.percentile(single, 99)) if len(single) else None,
                max_live=int(st.max_live), ms_per_iteration=ms_per_it,
                single_event_iterations_p50=float(np.percentile(single_its, 50)) if len(single_its) else None)


def serving_sweep(g, abi, capacity_gev_s, mix, loads, seconds, drain_s=15.0):
    """The serving loop at offered loads given as fractions of capacity_gev_s
    (GeV/s of primaries); the sustained capacity is the highest completion
    rate among the loads the loop kept up with (backlog at the window's end
    below 5 % of the jobs submitted)."""
    runs = []
    for f in loads:
        rate = f * capacity_gev_s / mix_gev_per_job(mix)
        r = serving(g, abi, seconds, rate, mix=mix, drain_s=drain_s if len(loads) > 1 else 60.0)
        r["load"] = f
        runs.append(r)
        log("serving %s at %.2f x: %.0f events/s, backlog %d of %d, p99 %.0f ms"
            % (mix, f, r["events_per_s"], r["backlog_at_window_end"], r["jobs_submitted"], r["p99_ms"]))
    kept = [r for r in runs if r["backlog_at_window_end"] <= max(2, 0.05 * r["jobs_submitted"])]
    best = max(kept, key=lambda r: r["events_per_s"]) if kept else None
    return dict(runs=runs, sustained=dict(events_per_s=best["events_per_s"], track_steps_per_s=best["track_steps_per_s"],
                                          load=best["load"], p50_ms=best["p50_ms"], p99_ms=best["p99_ms"],
                                          single_event_p99_ms=best["single_event_p99_ms"]) if best else None,
                saturated_events_per_s=max(r["events_per_s"] for r in runs))


def c0_block(g, abi, cpu=True):
    """The reference's own CPU-runnable case: 1000 gammas of 1 GeV into the
    20 cm lead block, seed 12345, in full on the GPU and on the host (all
    threads, then one), FP64."""
    n = 1000
    cfg = abi.config_lead_block(max_events=n)
    prims = g.make_primaries(12345, abi.GAMMA, 1000.0, 1000.0, 0.0, n)
    import torch
    with g.Transport(cfg) as t:
        t.submit(prims, 0)
        t.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        t.submit(prims, 1)
        t.run()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        st = t.stats()
        ev = t.event_tallies(1, n)
    out = dict(workload="1 GeV gamma x %d into the 20x20x20 cm Pb block, FP64" % n, gpu_events_per_s=n / dt,
               gpu_seconds=dt, edep_mean_mev=float(ev["edep"].mean()))
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_harness as O
        threads = max(1, min(16, os.cpu_count() or 1))
        for nt, key in ((threads, "cpu_all_threads"), (1, "cpu_one_core")):
            o = O.Oracle(cfg)
            p = O.make_primaries(12345, abi.GAMMA, 1000.0, 1000.0, 0.0, n)
            t0 = time.perf_counter()
            o.run(p, nt)
            dtc = time.perf_counter() - t0
            steps = int(o.stats().track_steps)
            cev = o.event_tallies(n)
            o.close()
            out[key] = dict(events_per_s=n / dtc, track_steps_per_s=steps / dtc, threads=nt, seconds=dtc)
        out["edep_mean_rel_diff_gpu_vs_oracle"] = float(ev["edep"].mean() / cev["edep"].mean() - 1.0)
    log("C0: GPU %.3f s for %d events" % (dt, n))
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(abi, target_s=15.0, single_s=4.0):
    """The oracle on the host: one worker per core (up to the 16 the GPU host
    allots this run), events handed out dynamically, on a bounded first slice
    of the same batch; then on one core."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_harness as O
    threads = max(1, min(16, os.cpu_count() or 1))
    cfg = abi.config_pbwo4_block(max_events=100000)

    def timed(n, nt):
        o = O.Oracle(cfg)
        p = O.make_primaries(777, abi.ELECTRON, 10000.0, 10000.0, 0.0, n)
        t0 = time.perf_counter()
        o.run(p, nt)
        dt = time.perf_counter() - t0
        return dt, o.stats().track_steps

    dt, _ = timed(threads, threads)
    n = int(max(threads, min(10000, threads * target_s / max(dt, 1e-3))))
    dt, steps = timed(n, threads)
    log("cpu oracle: %d events on %d threads in %.2f s" % (n, threads, dt))
    per_event_1 = dt * threads / n  # estimate of one core's time per event
    n1 = int(max(1, min(n, single_s / max(per_event_1, 1e-6))))
    dt1, steps1 = timed(n1, 1)
    log("cpu oracle: %d events on 1 thread in %.2f s" % (n1, dt1))
    return dict(value=n / dt, unit="events/s", cores=threads, kind="port",
                sample="first %d primaries of the same 10 GeV e- / PbWO4 batch (FP64 oracle, %d track-steps, %.1f s)"
                % (n, steps, dt), track_steps_per_s=steps / dt, cpu_model=cpu_model(),
                host_cores=os.cpu_count(),
                single_core=dict(value=n1 / dt1, track_steps_per_s=steps1 / dt1,
                                 sample="first %d primaries (%d track-steps, %.1f s)" % (n1, steps1, dt1)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--events", type=int, default=10000)
    ap.add_argument("--serve-seconds", type=float, default=8.0, help="seconds of submission per offered load")
    ap.add_argument("--serve-loads", default="0.6,0.9,1.1,1.4,1.8",
                    help="offered loads of the serving sweep, as fractions of the FP32 batch's GeV/s")
    ap.add_argument("--serve-mix", default="c4,e10", help="job mixes of the serving sweep (c4, e10)")
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the serving sweep (offered loads relative to --serve-capacity-gev-s)")
    ap.add_argument("--serve-capacity-gev-s", type=float, default=0.0,
                    help="with --serve-only: the GeV/s the loads are fractions of (0: measure the FP32 batch)")
    ap.add_argument("--skip-serving", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-fp32", action="store_true")
    ap.add_argument("--skip-precision", action="store_true", help="skip the FP32-vs-FP64 calorimeter comparison")
    ap.add_argument("--skip-field", action="store_true", help="skip the field-map and RK45 calorimeter runs")
    ap.add_argument("--skip-c0", action="store_true", help="skip the 1000 x 1 GeV gamma lead-block run")
    ap.add_argument("--skip-nav", action="store_true", help="skip the volume-tree navigation queries")
    ap.add_argument("--skip-tree", action="store_true", help="skip the transport through the volume tree")
    ap.add_argument("--tree-only", action="store_true", help="run only the transport through the volume tree")
    ap.add_argument("--nav-only", action="store_true", help="run only the volume-tree navigation queries")
    ap.add_argument("--field-only", action="store_true",
                    help="run only the field-map, RK45 and multiple-scattering calorimeter runs and print them")
    ap.add_argument("--pool", type=int, default=2_000_000, help="live-track pool capacity")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first run one device-to-device copy of MIB MiB (a kernel of known traffic that "
                         "calibrates the FETCH_SIZE / WRITE_SIZE counters of a profiler run)")
    a = ap.parse_args()
    if a.gpus != 1:
        sys.stderr.write("bench.py: this build runs the step loop on one GPU; multi-GPU runs are not supported "
                         "(events shard independently; see the design notes)\n")
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "needs a GPU"
    torch.cuda.set_device(0)
    from geantexascalepilot_amd import abi
    import geantexascalepilot_amd.gxt as g

    if a.copy_probe > 0:
        src = torch.ones(a.copy_probe << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copy probe: %d MiB read, %d MiB written" % (a.copy_probe, a.copy_probe))
        del src, dst
    if a.nav_only:
        print(json.dumps(dict(navigation=navigation(g, abi, cpu=not a.skip_cpu))), flush=True)
        return
    if a.tree_only:
        print(json.dumps(dict(tree_form=tree_form(g, abi, cpu=not a.skip_cpu))), flush=True)
        return
    if a.field_only:
        print(json.dumps(dict(field_forms=field_forms(g, abi, cpu=not a.skip_cpu))), flush=True)
        return
    n = a.events
    if a.serve_only:
        cap = a.serve_capacity_gev_s
        if cap <= 0:
            r32 = run_batch(g, abi, abi.FP32, n, a.steps, a.warmup, pool=a.pool)
            cap = a.steps * n / r32["seconds"] * 10.0
        loads = [float(x) for x in a.serve_loads.split(",")]
        out = dict(capacity_gev_per_s=cap)
        for mix in a.serve_mix.split(","):
            out["serving_" + mix] = serving_sweep(g, abi, cap, mix, loads, a.serve_seconds)
        print(json.dumps(out), flush=True)
        return
    r64 = run_batch(g, abi, abi.FP64, n, a.steps, a.warmup, pool=a.pool)
    value = a.steps * n / r64["seconds"]
    ks = r64["ks"]
    steps_per_s = ks.track_steps / r64["seconds"]
    # Algorithmic bytes of the dominant kernel: one read and one write of the
    # pool's track record (140 B in the FP64 form: position 24, direction 24,
    # energy 8, 3 interaction lengths 24, RNG 24, event, volume|species, step
    # count 4 each, 3 per-track tallies 24) for every track a launch loads.  A
    # track takes up to inner_steps steps in one launch while it neither
    # interacts nor ends, with one load and one store, so the unit is the
    # track loaded, not the track-step.  (The per-track-step figure of 2 x 148 B,
    # a nominal record with inverse mean free paths and safety, is reported
    # beside it.)
    rec = int(ks.record_bytes)
    # The dominant kernel (the rocprofv3 summary in profiles/kernels_c1_fp64.md)
    # is the photon interaction kernel, then the e-/e+ transport; only their
    # launches carry HIP events during the timed batches (profile_kernels = 2,
    # the events taken from the kernels' own dispatches), which keeps the
    # instrumentation out of the measured throughput.
    launches = max(ks.launches, 1)
    name = "k_interact<f64,photon>"
    ms, units, made = ks.interact_neutral_ms, ks.neutral_interactions, ks.neutral_created
    # a sampled photon's record read and written back (or its slot left:
    # the record is read whole either way, and a survivor is written whole to
    # the next pool), plus each created track's record written
    nbytes = units * 2 * rec + made * rec
    per_unit = ("2 x %d B per photon sampled (its record read, the survivor written) + %d B per track it "
                "creates; %d photons sampled, %d tracks created" % (rec, rec, units, made))
    avg_us = 1e3 * ms / launches
    achieved = (nbytes / launches) / (avg_us * 1e-6) / 1e9
    # the e-/e+ transport, the dominant kernel of earlier builds, for continuity
    tname = "k_transport<f64,e-/e+>"
    tms, tunits = ks.transport_charged_ms, ks.charged_loaded
    tbytes = tunits * 2 * rec
    tavg = 1e3 * tms / launches
    tach = (tbytes / launches) / (tavg * 1e-6) / 1e9
    transport = dict(kernel=tname, achieved=tach, frac=tach / PEAK_HBM_GBS, traffic=profiled_traffic(tname),
                     avg_launch_us=tavg, bytes_per_launch=tbytes / launches, kernel_ms=tms,
                     per_unit="2 x %d B per e-/e+ track loaded by a launch (one read and one write of its record), "
                              "%d tracks loaded for %d e-/e+ steps" % (rec, tunits, ks.charged_steps),
                     per_track_step_296b_gbs=(ks.charged_steps * 2 * 148 / launches) / (tavg * 1e-6) / 1e9)
    roofline = dict(bound="hbm", achieved=achieved, peak=PEAK_HBM_GBS, unit="GB/s", frac=achieved / PEAK_HBM_GBS,
                    traffic=profiled_traffic(name), kernel=name, launches=int(ks.launches), avg_launch_us=avg_us,
                    bytes_per_launch=nbytes / launches, per_unit=per_unit, record_bytes_in_pool=rec,
                    kernel_ms=ms, e_transport=transport,
                    # every track loaded by a transport launch (photons included) over the wall time
                    whole_run_gbs=(ks.charged_loaded + ks.neutral_loaded) * 2 * rec / r64["seconds"] / 1e9)
    e64 = r64["edep"]["edep"]
    out = dict(metric="events/s and track-steps/s, FP32 vs FP64 on 1 MI355X; p99 job latency", value=value,
               unit="events/s", n_gpus=1, steps=a.steps, warmup=a.warmup,
               ms_per_step=1e3 * r64["seconds"] / a.steps, higher_is_better=True, scaling="weak", vs_baseline=None,
               dtype="f64", data="synthetic",
               config=dict(workload="10 GeV e- x %d into 30x30x30 cm PbWO4 (8.28 g/cm3), no field, FP64, "
                                    "pool %.0e tracks, seed 777" % (n, a.pool), primaries=n, energy_gev=10.0,
                           material="PbWO4", pool_capacity=a.pool),
               track_steps_per_s=steps_per_s, track_steps_per_event=ks.track_steps / (a.steps * n),
               counts=dict(track_steps=int(ks.track_steps), photon_steps=int(ks.neutral_steps),
                           charged_steps=int(ks.charged_steps), interactions=int(ks.interactions),
                           created=int(ks.created), iterations=int(r64["stats"].iterations),
                           peak_live=int(r64["stats"].max_live)),
               mean_edep_mev=float(e64.mean()), roofline=roofline)
    if not a.skip_fp32:
        r32 = run_batch(g, abi, abi.FP32, n, a.steps, a.warmup, pool=a.pool)
        e32 = r32["edep"]["edep"]
        out["fp32"] = dict(value=a.steps * n / r32["seconds"], unit="events/s",
                           track_steps_per_s=r32["ks"].track_steps / r32["seconds"],
                           speedup_vs_fp64=(a.steps * n / r32["seconds"]) / value,
                           edep_rel_diff_vs_fp64=float(e32.sum() / e64.sum() - 1.0))
    if not a.skip_precision:
        out["precision_form"] = precision_form(g, abi)
    if not a.skip_field:
        out["field_forms"] = field_forms(g, abi, cpu=not a.skip_cpu)
    if not a.skip_nav:
        out["navigation"] = navigation(g, abi, cpu=not a.skip_cpu)
    if not a.skip_tree:
        out["tree_form"] = tree_form(g, abi, cpu=not a.skip_cpu)
    if not a.skip_serving:
        # offered loads relative to the FP32 batch's energy throughput on the
        # same block (GeV/s of primaries)
        gev_per_s = (out.get("fp32", {}).get("value", value)) * 10.0
        loads = [float(x) for x in a.serve_loads.split(",")]
        mixes = a.serve_mix.split(",")
        if "c4" in mixes:
            out["serving"] = serving_sweep(g, abi, gev_per_s, "c4", loads, a.serve_seconds)
        e10 = serving_sweep(g, abi, gev_per_s, "e10", loads, a.serve_seconds) if "e10" in mixes else None
        if e10:
            out["serving_e10"] = e10
        if e10 and e10["sustained"]:
            # the serving form (FP32) on 10 GeV e- jobs against the plain FP64 batch
            out["serving_e10"]["ratio_to_plain_fp64"] = e10["sustained"]["events_per_s"] / value
    if not a.skip_c0:
        out["c0"] = c0_block(g, abi, cpu=not a.skip_cpu)
    if not a.skip_cpu:
        out["cpu_baseline"] = cpu_baseline(abi)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
