"""GPU parity of e-/e+ multiple scattering: libgxt.so against the CPU oracle,
through the C ABI.  The model (the oracle's, pinned by its first moment and
closed forms in test_msc.py) has no counterpart in the reference.  With or
without a field, events agree in every count and within 1e-9 in every energy,
as for the rest of the transport (both sides compute the same correctly
rounded FP64 operations, sin/cos/log/exp included); the FP32 precision form
agrees with FP64 in distribution."""
import numpy as np
import pytest

from geantexascalepilot_amd import abi
import oracle_harness as O
from test_gpu_parity import _compare_events, _flush, _transport_pair

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("field", ["none", "4T", "cms"])
def test_transport_msc_event_parity(gpu_lib, field):
    _flush("transport calorimeter msc field %s" % field)
    n = 8
    if field == "cms":
        cfg = abi.config_calorimeter_cms(max_events=n, msc=1)
    else:
        cfg = abi.config_calorimeter(max_events=n, msc=1,
                                     field=(0.0, 0.0, 0.0) if field == "none" else (4.0, 0.0, 0.0))
    pg = gpu_lib.make_primaries(cfg.seed, abi.ELECTRON, 1000.0, 3000.0, 10.0, n)
    pc = O.make_primaries(cfg.seed, abi.ELECTRON, 1000.0, 3000.0, 10.0, n)
    gpu, cpu, gprof, cprof, *_ = _transport_pair(cfg, pg, pc, n)
    _compare_events(gpu, cpu)
    for a, b in zip(gprof, cprof):
        np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-9)


def test_transport_msc_fp32(gpu_lib):
    """Precision form with multiple scattering: means of 200 events agree with
    FP64 on the same primaries within their statistical error."""
    _flush("transport calorimeter msc FP32")
    n = 200
    res = []
    for prec in (abi.FP64, abi.FP32):
        cfg = abi.config_calorimeter(precision=prec, max_events=n, msc=1)
        p = gpu_lib.make_primaries(cfg.seed, abi.ELECTRON, 1000.0, 5000.0, 10.0, n)
        with gpu_lib.Transport(cfg) as t:
            t.submit(p, 1)
            t.run(max_iterations=200000)
            res.append(t.event_tallies(1, n))
    for k in ("edep", "visible", "charged_length"):
        a, b = res[0][k], res[1][k]
        err = np.sqrt((a.var() + b.var()) / n)
        assert abs(a.mean() - b.mean()) < 4 * err + 1e-6 * abs(a.mean()), k


def test_msc_needs_tables(gpu_lib):
    with pytest.raises(gpu_lib.GxtError, match="LambdaMod1"):
        gpu_lib.Transport(abi.config_pbwo4_block(max_events=4, msc=1))
