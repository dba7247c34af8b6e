"""tdscha_amd: the TD-SCHA ensemble L-application on MI355X (gfx950).

Host-side mirror of the reference's ``tdscha`` interface for the Lanczos hot path:

* ``sscha_HP_odd.GetPerturbAverageSym`` - drop-in for the reference's CPython binding
  (CModules/odd_corr_module.c:491-634), computed by libtdscha_hip on the GPU.
* ``DynamicalLanczos.Lanczos`` - apply_full_L / run_FT / continued fraction with the new
  device modes MODE_GPU_FP64, MODE_GPU_FP32, MODE_GPU_BF16X3.
* ``ensemble.DeviceEnsemble`` - the device-resident ensemble behind both.
"""
from ._lib import (PREC_BF16X3, PREC_FP32, PREC_FP64, TERM_D2V_R, TERM_D2V_Y, TERM_F,
                   TdschaError, lib)
from .ensemble import DeviceEnsemble, SymmetryBlocks

__all__ = ["lib", "TdschaError", "DeviceEnsemble", "SymmetryBlocks", "PREC_FP64", "PREC_FP32",
           "PREC_BF16X3", "TERM_F", "TERM_D2V_R", "TERM_D2V_Y"]
