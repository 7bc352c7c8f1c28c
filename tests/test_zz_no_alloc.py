"""gxt_sample_noalloc, the sampler entry point that allocates no device
memory (its inputs and outputs are staged through the library's fixed device
array), against gxt_sample through the C ABI: the same kernel on the same
inputs gives bit-identical outputs and RNG states, on two calls in a row
(the second smaller, so stale staging from the first would show), for every
model in FP64 and FP32; n beyond GXT_SAMPLE_NOALLOC_MAX is rejected."""
import numpy as np
import pytest

from geantexascalepilot_amd import abi
import oracle_harness as O

pytestmark = pytest.mark.gpu

MODEL_E = {"KLEIN_NISHINA": (1e-3, 1e4), "BETHE_HEITLER": (1.1, 1e5), "SAUTER_GAVRILA": (2e-3, 5.0),
           "SELTZER_BERGER_EM": (0.2, 5e4), "MOLLER": (2.5, 5e4), "SELTZER_BERGER_EP": (0.2, 5e4),
           "BHABHA": (1.5, 5e4), "ANNIHIL": (1e-3, 1e4)}


def inputs(model, n, seed):
    rs = np.random.default_rng(seed)
    lo, hi = MODEL_E[model]
    ekin = np.exp(rs.uniform(np.log(lo), np.log(hi), n))
    d = rs.normal(size=(3, n))
    d /= np.linalg.norm(d, axis=0)
    st = O.rng_init(np.full(6, 99, dtype=np.uint32), np.arange(seed * n, seed * n + n, dtype=np.uint64))
    return ekin, d, st


@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32])
@pytest.mark.parametrize("model", sorted(MODEL_E))
def test_sample_noalloc_equals_sample(gpu_lib, model, precision):
    print("noalloc %s precision %d" % (model, precision), flush=True)
    with gpu_lib.Transport(abi.config_pbwo4_block(max_events=1)) as t:
        m = abi.MODEL[model]
        for n, seed in ((abi.SAMPLE_NOALLOC_MAX, 1), (1000, 2)):
            ekin, d, st = inputs(model, n, seed)
            a = t.sample(m, 0, ekin, d, st, precision)
            b = t.sample(m, 0, ekin, d, st, precision, noalloc=True)
            for k in a:
                assert np.array_equal(a[k].view(np.int64) if a[k].dtype == np.float64 else a[k],
                                      b[k].view(np.int64) if b[k].dtype == np.float64 else b[k]), (model, n, k)


def test_sample_noalloc_rejects_oversize(gpu_lib):
    with gpu_lib.Transport(abi.config_pbwo4_block(max_events=1)) as t:
        ekin, d, st = inputs("KLEIN_NISHINA", abi.SAMPLE_NOALLOC_MAX + 1, 3)
        with pytest.raises(gpu_lib.GxtError, match="staging limit"):
            t.sample(abi.MODEL["KLEIN_NISHINA"], 0, ekin, d, st, noalloc=True)
