"""The serving form on its job mix: a Poisson stream of transport jobs,
40 % one 1 GeV gamma, 40 % 10 x 10 GeV e-, 20 % 100 x 50 GeV e-, into the
30 cm PbWO4 block, FP32, admitted by the serving loop into one shared pool
while earlier jobs are on the device (continuous batching).

Stated tolerances (the workload's acceptance):
- every job completes, with submit <= admit <= done on the host clock;
- every FP32 event balances energy to 1e-5 relative;
- the total deposit of the mix, and of each kind of job, within 0.5 % of the
  FP64 oracle on the same primaries (same substreams);
- per-event deposit / primary energy: two-sample KS p > 0.01 against the
  oracle, for all events and for each electron job kind;
- a job of each kind gives the same events run alone (FP32, batch mode) as
  in the mix: counts exact, FP64 tallies within 1e-9 (atomic-sum order).
"""
import time

import numpy as np
import pytest
from scipy.stats import ks_2samp

from geantexascalepilot_amd import abi
import oracle_harness as O

pytestmark = pytest.mark.gpu

ME = 0.510998910
KINDS = {"gamma1": (1, abi.GAMMA, 1000.0), "e10x10": (10, abi.ELECTRON, 10000.0),
         "e100x50": (100, abi.ELECTRON, 50000.0)}


def _flush(msg):
    print(msg, flush=True)


def make_jobs(g, n_jobs=30, seed=4242):
    """n_jobs of the mix in the 40/40/20 proportions, in a seeded random order;
    job j's primaries take the next substreams after job j-1's."""
    rs = np.random.default_rng(seed)
    kinds = ["gamma1"] * (4 * n_jobs // 10) + ["e10x10"] * (4 * n_jobs // 10)
    kinds += ["e100x50"] * (n_jobs - len(kinds))
    rs.shuffle(kinds)
    jobs, first = [], 0
    for kind in kinds:
        n, sp, e = KINDS[kind]
        jobs.append((kind, g.make_primaries(777, sp, e, e, 0.0, n, first=first, n_total=1 << 40)))
        first += n
    return jobs, rs


def test_serving_mix_fp32_against_fp64_oracle(gpu_lib):
    g = gpu_lib
    jobs, rs = make_jobs(g)
    n_ev = sum(len(p) for _, p in jobs)
    _flush("C4: %d jobs, %d events, FP32 serving" % (len(jobs), n_ev))
    cfg = abi.config_pbwo4_block(max_events=4000, precision=abi.FP32, max_jobs=256, pool_capacity=4_000_000)
    done = []
    with g.Transport(cfg) as t:
        t.serve_start()
        for j, (_, p) in enumerate(jobs):
            t.submit(p, j)
            time.sleep(rs.exponential(0.004))
            done += t.poll()
        t0 = time.time()
        while len(done) < len(jobs) and time.time() - t0 < 120:
            done += t.poll()
            time.sleep(0.001)
        t.serve_stop()
        assert sorted(r.job_id for r in done) == list(range(len(jobs)))
        tallies = {j: t.event_tallies(j, len(p)) for j, (_, p) in enumerate(jobs)}
    for r in done:
        assert r.submit_ms <= r.admit_ms <= r.done_ms
    # the oracle (FP64) on the same primaries, events numbered over the mix
    allp = abi.primaries_array(n_ev)
    k = 0
    for _, p in jobs:
        for q in p:
            allp[k] = q
            allp[k].event = k
            k += 1
    ref = O.Oracle(abi.config_pbwo4_block(max_events=n_ev))
    t0 = time.time()
    ref.run(allp, nthreads=16)
    _flush("  oracle: %d events in %.1f s" % (n_ev, time.time() - t0))
    cpu = ref.event_tallies(n_ev)
    ref.close()
    gpu_edep = np.concatenate([tallies[j]["edep"] for j in range(len(jobs))])
    e0 = np.array([q.ekin for q in allp])
    bal = np.concatenate([tallies[j]["edep"] + tallies[j]["leak"]
                          + 2 * ME * (tallies[j]["n_pairs"] - tallies[j]["n_annihil"]) for j in range(len(jobs))])
    np.testing.assert_allclose(bal, e0, rtol=1e-5)
    rel = gpu_edep.sum() / cpu["edep"].sum() - 1.0
    _flush("  total deposit FP32 serving vs FP64 oracle: %.2e" % rel)
    assert abs(rel) < 5e-3
    p = ks_2samp(gpu_edep / e0, cpu["edep"] / e0).pvalue
    assert p > 0.01, p
    kind_of_event = np.concatenate([[kind] * len(p) for kind, p in jobs])
    for kind in KINDS:
        sel = kind_of_event == kind
        relk = gpu_edep[sel].sum() / cpu["edep"][sel].sum() - 1.0
        _flush("  %s: %d events, deposit %.2e" % (kind, sel.sum(), relk))
        assert abs(relk) < 5e-3, kind
        if sel.sum() >= 50:
            pk = ks_2samp(gpu_edep[sel] / e0[sel], cpu["edep"][sel] / e0[sel]).pvalue
            assert pk > 0.01, (kind, pk)
    # a job of each kind, alone
    for kind in KINDS:
        j = next(i for i, (kd, _) in enumerate(jobs) if kd == kind)
        p = jobs[j][1]
        with g.Transport(cfg) as t:
            t.submit(p, 99)
            t.run()
            alone = t.event_tallies(99, len(p))
        mixed = tallies[j]
        for key in ("n_steps", "n_secondaries", "n_pairs", "n_annihil"):
            assert np.array_equal(alone[key], mixed[key]), (kind, key)
        for key in ("edep", "visible", "leak", "charged_length"):
            np.testing.assert_allclose(mixed[key], alone[key], rtol=1e-9, atol=1e-9, err_msg=kind)
