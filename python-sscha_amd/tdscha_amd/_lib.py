"""Bindings of libtdscha_hip.so (the C ABI declared in include/tdscha_hip.h).

Two layers, both loaded from the package's lib/ directory, with no fallback (if a shared
object is missing, the entry points that need it raise):
  * ctypes on the C ABI for set-up, teardown and the host-pointer drop-ins (lib());
  * the PyTorch-ROCm operators torch.ops.tdscha.* (libtdscha_torch.so, csrc/torch_ops.cpp)
    for the per-step device calls: they take device tensors, check them, and enqueue the
    kernels on the caller's current HIP stream (ops()).
"""
import ctypes
import os

from ._build import LIB_PATH, TORCH_OPS_PATH

PREC_FP64, PREC_FP32, PREC_BF16X3 = 0, 1, 2
PRECISIONS = {"fp64": PREC_FP64, "fp32": PREC_FP32, "bf16x3": PREC_BF16X3}
TERM_F, TERM_D2V_R, TERM_D2V_Y = 1, 2, 4
OP_TRANSPOSE, OP_WIGNER, OP_IGNORE_V3, OP_IGNORE_V4, OP_IGNORE_HARMONIC, OP_REVERSE = (
    1, 2, 4, 8, 16, 32)

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_dpp = ctypes.POINTER(_dp)
_ipp = ctypes.POINTER(_ip)
_vp = ctypes.c_void_p

# name -> (restype, argtypes); every symbol declared in include/tdscha_hip.h
# the Julia-level range entries: X, Y, omega, rho, R1, Y1, temperature, apply_v4, symmetries,
# n_degeneracies, degenerate_space, blocks, start_index, end_index, n_modes, n_configs,
# n_syms, n_blocks, d_max (then the outputs)
_JL = [_dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int, _dp, _ip, _ip, _ip,
       ctypes.c_longlong, ctypes.c_longlong, ctypes.c_int, ctypes.c_int, ctypes.c_int,
       ctypes.c_int, ctypes.c_int]
_TRIO = [_dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, _dpp, ctypes.c_int,
         _ip, _ipp, _ip, _dp]
SIGNATURES = {
    "tdscha_last_error": (ctypes.c_char_p, []),
    "tdscha_abi_version": (ctypes.c_int, []),
    "get_f_average_from_Y_pert_sym_fast": (None, _TRIO),
    "get_d2v_dR2_from_R_pert_sym_fast": (None, _TRIO),
    "get_d2v_dR2_from_Y_pert_sym_fast": (None, _TRIO),
    "tdscha_get_f_average_from_Y_pert_sym_fast": (ctypes.c_int, _TRIO),
    "tdscha_get_d2v_dR2_from_R_pert_sym_fast": (ctypes.c_int, _TRIO),
    "tdscha_get_d2v_dR2_from_Y_pert_sym_fast": (ctypes.c_int, _TRIO),
    "tdscha_get_perturb_average_sym": (ctypes.c_int, [
        _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
        _dpp, _ip, _ipp, _ip, ctypes.c_int, ctypes.c_int, _dp, _dp]),
    "tdscha_get_perturb_average_sym_noalloc": (ctypes.c_int, [
        _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
        _dpp, _ip, _ipp, _ip, ctypes.c_int, ctypes.c_int, _dp, _dp]),
    "tdscha_noalloc_scratch_bytes": (ctypes.c_size_t, []),
    "tdscha_drop_in_cache": (ctypes.c_int, [ctypes.c_int]),
    "tdscha_mfma_stream": (ctypes.c_int, [ctypes.c_int, ctypes.c_double,
                                          ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_double)]),
    "tdscha_jl_get_perturb_averages_sym": (ctypes.c_int, _JL + [_dp, _dp]),
    "tdscha_jl_get_perturb_d2v_averages_sym": (ctypes.c_int, _JL + [_dp]),
    "tdscha_jl_get_perturb_f_averages_sym": (ctypes.c_int, _JL + [_dp]),
    "tdscha_ensemble_create": (ctypes.c_int, [
        ctypes.POINTER(_vp), ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int,
        ctypes.c_int, ctypes.c_int, ctypes.c_int, _dpp, _ip, _ipp, _ip, ctypes.c_int, ctypes.c_int]),
    "tdscha_ensemble_destroy": (ctypes.c_int, [_vp]),
    "tdscha_ensemble_dims": (ctypes.c_int, [_vp, _ip]),
    "tdscha_ensemble_bytes": (ctypes.c_longlong, [_vp]),
    "tdscha_ensemble_apply": (ctypes.c_int, [
        _vp, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp, ctypes.c_int, _vp]),
    "tdscha_apply_full_L": (ctypes.c_int, [
        _vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _vp, _vp, _vp]),
    "tdscha_apply_full_L_batched": (ctypes.c_int, [
        _vp, ctypes.c_int, ctypes.c_int, _ip, _dp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp]),
    "tdscha_ensemble_reserve": (ctypes.c_int, [_vp, ctypes.c_int]),
    "tdscha_ensemble_autotune": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _ip, _dp]),
    "tdscha_ensemble_set_tuning": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "tdscha_ensemble_get_tuning": (ctypes.c_int, [_vp, _ip, _ip]),
    "tdscha_ensemble_force_grad": (ctypes.c_int, [
        _vp, ctypes.c_int, ctypes.c_int, _ip, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp,
        ctypes.c_int, _vp]),
    "tdscha_lanczos_create": (ctypes.c_int, [
        ctypes.POINTER(_vp), _vp, ctypes.c_int, ctypes.c_int, _ip, _dp, ctypes.c_int, ctypes.c_int,
        ctypes.c_int, ctypes.c_int, _dp]),
    "tdscha_lanczos_read_gs": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _ip]),
    "tdscha_lanczos_destroy": (ctypes.c_int, [_vp]),
    "tdscha_lanczos_set_projector": (ctypes.c_int, [_vp, _vp]),
    "tdscha_projector_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int,
                                               ctypes.c_int, _dp, ctypes.c_int, _ip]),
    "tdscha_projector_destroy": (ctypes.c_int, [_vp]),
    "tdscha_projector_apply": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp]),
    "tdscha_lanczos_start": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), _vp]),
    "tdscha_lanczos_run": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp]),
    "tdscha_lanczos_read": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp, _dp, _ip]),
    "tdscha_lanczos_basis": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_vp),
                                            ctypes.POINTER(_vp)]),
    "tdscha_lanczos_group_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                                   ctypes.c_int]),
    "tdscha_lanczos_group_run": (ctypes.c_int, [_vp, ctypes.c_int, _vp]),
    "tdscha_lanczos_group_destroy": (ctypes.c_int, [_vp]),
    "tdscha_project_ensemble": (ctypes.c_int, [
        ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_double,
        _vp, _vp, _vp]),
    "tdscha_symmetries_on_modes": (ctypes.c_int, [
        ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, ctypes.c_int, _ip, _vp, _vp]),
    "tdscha_weighted_tn_product": (ctypes.c_int, [
        ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp, ctypes.c_int,
        ctypes.c_double, _vp, ctypes.c_int, _vp]),
    "tdscha_profile_enable": (ctypes.c_int, [ctypes.c_int]),
    "tdscha_profile_reset": (ctypes.c_int, []),
    "tdscha_profile_query": (ctypes.c_int, [ctypes.c_int, _dp,
                                            ctypes.POINTER(ctypes.c_longlong)]),
}

KGROUP_GEMM1, KGROUP_GEMM2, KGROUP_ROTATE, KGROUP_FAV = 0, 1, 2, 3
KGROUP_NAMES = {KGROUP_GEMM1: "gemm1", KGROUP_GEMM2: "gemm2", KGROUP_ROTATE: "rotate",
                KGROUP_FAV: "fav"}


def profile_query(group):
    """(total milliseconds, launches) of a kernel group since the last reset."""
    ms = ctypes.c_double()
    cnt = ctypes.c_longlong()
    check(lib().tdscha_profile_query(int(group), ctypes.byref(ms), ctypes.byref(cnt)),
          "tdscha_profile_query")
    return ms.value, cnt.value

_lib = None
_ops = None


def ops():
    """torch.ops.tdscha, after loading libtdscha_torch.so (raises if it has not been built)."""
    global _ops
    if _ops is None:
        if not os.path.exists(TORCH_OPS_PATH):
            raise ImportError(
                "libtdscha_torch.so is not built (%s); run __graft_entry__.build() or "
                "python -m tdscha_amd._build" % TORCH_OPS_PATH)
        import torch
        lib()                      # the C library first (the operators link against it)
        torch.ops.load_library(TORCH_OPS_PATH)
        _ops = torch.ops.tdscha
    return _ops


class TdschaError(RuntimeError):
    """An entry point of libtdscha_hip returned an error code."""


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                "libtdscha_hip.so is not built (%s); run __graft_entry__.build() or "
                "python -m tdscha_amd._build" % LIB_PATH)
        # PyTorch first: its wheel ships the HIP runtime under the same soname
        # (libamdhip64.so.7) as the one the library links, and the process keeps the first
        # copy it loads. Loaded before torch, the system runtime would serve torch too, and
        # torch then sees no device.
        import torch  # noqa: F401
        L = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_LOCAL)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc, what):
    if rc != 0:
        msg = lib().tdscha_last_error().decode(errors="replace")
        if rc in (1, 2, 3, 4):
            raise ValueError("%s: %s" % (what, msg))
        raise TdschaError("%s failed (code %d): %s" % (what, rc, msg))
