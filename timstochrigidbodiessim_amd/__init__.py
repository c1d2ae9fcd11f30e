"""MI355X-native FCM mobility solve (force-coupling method, rigid bodies).

The compute path is libfcm_hip.so (hand-written HIP kernels for gfx950 behind
the C ABI of include/fcm_hip.h); this package is its Python binding.
"""
from ._lib import (FCM_PREC_FP32, FCM_PREC_FP64, EXPORTS, PASS_NAMES, FcmBatch, FcmConfig, FcmError, FcmSolver, GmresOpts,
                   GmresReport, clear_tuning_cache, footprint, lib, make_config, synth_rods)

__all__ = ["FCM_PREC_FP32", "FCM_PREC_FP64", "EXPORTS", "PASS_NAMES", "FcmBatch", "FcmConfig", "FcmError", "FcmSolver", "GmresOpts",
           "GmresReport", "clear_tuning_cache", "footprint", "lib", "make_config", "synth_rods"]
