"""fastpcc_amd: interstation phase cross-correlation on MI355X (gfx950).

The compute path is libfastpcc.so (hand-written HIP kernels behind a C ABI, see
include/fastpcc.h); this package loads it with ctypes and mirrors the reference's
correlation interface on numpy arrays (api.py).
"""
from .api import (FastPCCError, DeviceRunner, Server, pinned_empty, analytic_signal, allpairs, preprocess,
                  make_preproc, make_job, output_shape, pcc2_set, ccgn_set,  # noqa: F401
                  cc1b_set, pcc1_set, pcc_set, pcc1_highlevel, pcc_highlevel, pcc1_highlevel2, tspcc2_set,
                  cuda_tspcc2_set_freq, release_buffers, set_kept_limits)
from . import synth  # noqa: F401
