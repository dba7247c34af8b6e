"""Caller-owned output buffers of the device ensemble are validated before any launch: a
wrong dtype, shape, device, layout or an alias of the input raises ValueError (the kernels
write the whole extent through a raw pointer)."""
import numpy as np
import pytest

import fixtures
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def test_bad_output_buffers_are_refused():
    import torch
    from tdscha_amd.ensemble import DeviceEnsemble, SymmetryBlocks
    print("case: output-buffer validation, 12-mode random ensemble", flush=True)
    ens = fixtures.random_ensemble(71, 12, 20, 3)
    s = ens.sym
    dev = DeviceEnsemble(ens.X, ens.Y, ens.w, ens.rho, ens.T,
                         SymmetryBlocks(s.symmetries, s.N_degeneracy, s.degenerate_space,
                                        s.block_id))
    n_psi = orc.psi_len(12)
    psi = torch.as_tensor(np.random.default_rng(72).standard_normal(n_psi), device=dev.device)
    good = torch.empty_like(psi)
    bad = [torch.empty(n_psi, dtype=torch.float32, device=dev.device),
           torch.empty(n_psi - 1, dtype=torch.float64, device=dev.device),
           torch.empty(2 * n_psi, dtype=torch.float64, device=dev.device)[::2],
           torch.empty(n_psi, dtype=torch.float64),
           psi]
    for o in bad:
        with pytest.raises(ValueError):
            dev.apply_full_L(psi, out=o)
        with pytest.raises(ValueError):
            dev.apply_full_L_batched([psi, psi], outs=[good, o])
    with pytest.raises(ValueError):
        dev.apply_full_L_batched([psi, psi], outs=[good])
    with pytest.raises(ValueError):
        dev.apply(np.zeros(12), np.zeros((12, 12)),
                  f=torch.empty(11, dtype=torch.float64, device=dev.device))
    with pytest.raises(ValueError):
        dev.apply(np.zeros(12), np.zeros((12, 12)),
                  d2v=torch.empty((12, 12), dtype=torch.float32, device=dev.device))
    # the good buffer still works and nothing was written by the refused calls
    ref = dev.apply_full_L(psi).clone()
    dev.apply_full_L(psi, out=good)
    assert torch.equal(good, ref)


def test_mfma_stream_rates_are_plausible():
    """The sustained-rate probe returns a rate between a quarter of and the published peak
    (bf16 2500, fp64 78.6 TFLOP/s) on a short run."""
    import ctypes
    import torch
    from tdscha_amd import _lib as L
    torch.cuda.init()
    for prec, peak in ((L.PREC_BF16X3, 2500.0), (L.PREC_FP64, 78.6)):
        tf, ms = ctypes.c_double(), ctypes.c_double()
        L.check(L.lib().tdscha_mfma_stream(prec, 0.6, ctypes.byref(tf), ctypes.byref(ms)),
                "tdscha_mfma_stream")
        print("precision %d: %.1f TFLOP/s, %.2f ms per launch" % (prec, tf.value, ms.value),
              flush=True)
        assert 0.25 * peak < tf.value <= 1.02 * peak
