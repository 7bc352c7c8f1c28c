"""The precision form of the sampling-calorimeter workload at full size:
20,000 e- of 1-50 GeV (10 mm spot, seed 2468) into 50 x (2.3 mm Pb + 5.7 mm
lAr 1.39 g/cm3), 40x40 cm, uniform 4 T along +x, exact helix, no multiple
scattering; FP32 table interpolation, sampling and helix, FP64 positions,
safety and tallies.  The same primaries run in the FP64 form.

Stated tolerances of the FP32 form against the FP64 run:
- every FP32 event balances energy (deposit + leakage + 2 m_e per pair not
  annihilated = the primary's energy) to 1e-5 relative (FP64: 1e-9);
- mean visible energy within 1e-3, visible-energy RMS within 5e-3;
- argon deposit per layer within 1e-2 for every layer holding more than
  0.1 % of the visible energy;
- shower maximum (1 cm bins) in the same bin, energy-weighted centroid within 0.1 mm.
Against the oracle (FP64, the first 256 primaries of the batch, run alone):
two-sample KS p > 0.01 on the per-event deposit, visible deposit, charged
length and step count, each divided by the primary's energy.
"""
import numpy as np
import pytest
from scipy.stats import ks_2samp

from geantexascalepilot_amd import abi
import oracle_harness as O

pytestmark = pytest.mark.gpu

ME = 0.510998910
N = 20000
SEED = 2468


def _flush(msg):
    print(msg, flush=True)


def run(g, precision, prims):
    cfg = abi.config_calorimeter(precision=precision, max_events=N, n_zbins=40, zbin_mm=10.0)
    with g.Transport(cfg) as t:
        t.submit(prims, 1)
        t.run()
        ev = t.event_tallies(1, N)
        ez, er, evol = t.profiles()
        st = t.stats()
    return ev, ez, evol, st


def test_precision_form_calorimeter_20k(gpu_lib):
    _flush("C3: 20000 e- 1-50 GeV in the 4 T calorimeter, FP32 against FP64")
    prims = gpu_lib.make_primaries(SEED, abi.ELECTRON, 1000.0, 50000.0, 10.0, N)
    e0 = np.array([p.ekin for p in prims])
    ev64, ez64, vol64, st64 = run(gpu_lib, abi.FP64, prims)
    ev32, ez32, vol32, st32 = run(gpu_lib, abi.FP32, prims)
    for ev, tol in ((ev64, 1e-9), (ev32, 1e-5)):
        bal = ev["edep"] + ev["leak"] + 2.0 * ME * (ev["n_pairs"] - ev["n_annihil"])
        np.testing.assert_allclose(bal, e0, rtol=tol)
    assert st32.track_steps == int(ev32["n_steps"].sum()) and st32.killed_by_limit == 0
    vis64, vis32 = ev64["visible"], ev32["visible"]
    mean_err = vis32.mean() / vis64.mean() - 1.0
    rms_err = vis32.std() / vis64.std() - 1.0
    lay64, lay32 = vol64[1::2], vol32[1::2]
    big = lay64 > 1e-3 * lay64.sum()
    lay_err = np.abs(lay32[big] / lay64[big] - 1.0)
    zc = (np.arange(len(ez64)) + 0.5) * 10.0
    cen64, cen32 = (zc * ez64).sum() / ez64.sum(), (zc * ez32).sum() / ez32.sum()
    _flush("  visible mean %.2e, rms %.2e, layers max %.2e, centroid %+.4f mm, steps %d / %d"
           % (mean_err, rms_err, lay_err.max(), cen32 - cen64, st32.track_steps, st64.track_steps))
    assert abs(mean_err) < 1e-3
    assert abs(rms_err) < 5e-3
    assert big.sum() >= 30 and lay_err.max() < 1e-2
    assert np.argmax(ez32) == np.argmax(ez64)
    assert abs(cen32 - cen64) < 0.1
    # FP32 against the FP64 oracle on the batch's first 256 primaries
    m = 256
    ref = O.Oracle(abi.config_calorimeter(max_events=m, n_zbins=40, zbin_mm=10.0))
    ref.run(O.make_primaries(SEED, abi.ELECTRON, 1000.0, 50000.0, 10.0, m, first=0, n_total=N), nthreads=16)
    cpu = ref.event_tallies(m)
    ref.close()
    for k in ("edep", "visible", "charged_length", "n_steps"):
        p = ks_2samp(ev32[k][:m] / e0[:m], cpu[k] / e0[:m]).pvalue
        _flush("  KS fp32 vs oracle %s: p = %.3g" % (k, p))
        assert p > 0.01, (k, p)
