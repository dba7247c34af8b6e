"""CPU-side checks of the C ABI: the library loads, exports every declared symbol, and its
host-side argument validation rejects bad input with an error (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

import fixtures

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tdscha_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)) -
                  {"if", "defined", "sizeof"})


def test_library_exports_every_declared_symbol():
    from tdscha_amd import _lib as L
    lib = L.lib()
    syms = declared_symbols()
    assert len(syms) >= 14
    for s in syms:
        assert hasattr(lib, s), s
        assert s in L.SIGNATURES, s
    assert lib.tdscha_abi_version() == 2


def _create(ens, **over):
    from tdscha_amd import _lib as L
    from tdscha_amd.ensemble import SymmetryBlocks
    s = ens.sym
    sym = SymmetryBlocks(over.get("symmetries", s.symmetries),
                         over.get("N_degeneracy", s.N_degeneracy),
                         over.get("degenerate_space", s.degenerate_space),
                         over.get("block_id", s.block_id))
    syms, degs = sym.pointer_arrays()
    w = np.ascontiguousarray(over.get("w", ens.w))
    X, Y, rho = (np.ascontiguousarray(a) for a in (ens.X, ens.Y, ens.rho))
    h = ctypes.c_void_p()
    rc = L.lib().tdscha_ensemble_create(
        ctypes.byref(h), 0, X.ctypes.data_as(L._dp), Y.ctypes.data_as(L._dp),
        w.ctypes.data_as(L._dp), rho.ctypes.data_as(L._dp), ens.T, X.shape[1], X.shape[0],
        sym.n_syms, sym.n_blocks, syms, sym.N_degeneracy.ctypes.data_as(L._ip), degs,
        sym.block_id.ctypes.data_as(L._ip), 1, 0)
    msg = L.lib().tdscha_last_error().decode()
    if rc == 0:
        L.lib().tdscha_ensemble_destroy(h)
    return rc, msg


def test_rejects_non_contiguous_blocks():
    ens = fixtures.random_ensemble(1, 6, 4, 2, max_block=1)
    degs = [d.copy() for d in ens.sym.degenerate_space]
    degs[0], degs[1] = degs[1], degs[0]
    rc, msg = _create(ens, degenerate_space=degs)
    assert rc == 2 and "blocks_ids" in msg or "contiguous" in msg


def test_rejects_nonpositive_frequency():
    ens = fixtures.random_ensemble(3, 5, 4, 2)
    w = ens.w.copy()
    w[2] = 0.0
    rc, msg = _create(ens, w=w)
    assert rc == 3 and "frequency" in msg


def test_shim_rejects_bad_apply_d4_before_device_work(mode10):
    from tdscha_amd.sscha_HP_odd import GetPerturbAverageSym
    f, d = np.zeros(45), np.zeros((45, 45))
    with pytest.raises(ValueError):
        GetPerturbAverageSym(mode10.X, mode10.Y, mode10.w, mode10.rho, np.zeros(45),
                             np.zeros((45, 45)), mode10.T, 3, 384, mode10.sym.symmetries,
                             mode10.sym.N_degeneracy, mode10.sym.degenerate_space,
                             mode10.sym.block_id, f, d)
    with pytest.raises(ValueError):
        GetPerturbAverageSym(mode10.X, mode10.Y, mode10.w[:44], mode10.rho, np.zeros(45),
                             np.zeros((45, 45)), mode10.T, 1, 384, mode10.sym.symmetries,
                             mode10.sym.N_degeneracy, mode10.sym.degenerate_space,
                             mode10.sym.block_id, f, d)


def test_torch_operators_load_and_declare_their_schemas():
    """libtdscha_torch.so (the PyTorch-ROCm operators over the C ABI) loads without a GPU and
    registers the per-step operators (no compute call here)."""
    from tdscha_amd import _lib as L
    ops = L.ops()
    schemas = {name: str(getattr(ops, name).default._schema) for name in
               ("apply_full_L_batched", "ensemble_apply", "force_grad", "projector_apply")}
    assert "Tensor[] psis" in schemas["apply_full_L_batched"]
    assert "Tensor grad" in schemas["force_grad"]
    assert "bool adjoint" in schemas["projector_apply"]


def test_mfma_stream_rejects_bad_arguments_without_a_device():
    """tdscha_mfma_stream checks its arguments before touching the device."""
    from tdscha_amd import _lib as L
    tf = ctypes.c_double()
    assert L.lib().tdscha_mfma_stream(L.PREC_FP32, 1.0, ctypes.byref(tf), None) == 1
    assert L.lib().tdscha_mfma_stream(L.PREC_BF16X3, 0.0, ctypes.byref(tf), None) == 1
    assert L.lib().tdscha_mfma_stream(L.PREC_FP64, 1.0, None, None) == 1
