"""ctypes entry to the Fortran host (libvolcanor_host.so, volcanor_run).

Runs a whole `.case` directory (config.nml + geomXX.nml) through the
Fortran object model and the HIP library, returning the CL/CT history.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import vlm

_HERE = os.path.dirname(os.path.abspath(__file__))
HOST_LIB = os.path.join(_HERE, "lib", "libvolcanor_host.so")


class Stats(C.Structure):
    _fields_ = [("wall_total", C.c_double), ("wall_loop", C.c_double),
                ("wall_steps_tail", C.c_double), ("bs_kernel_ms", C.c_double),
                ("bs_pairs", C.c_double), ("bs_launches", C.c_int64),
                ("nt", C.c_int), ("tail_steps", C.c_int),
                ("pred_full", C.c_int64), ("pred_learned", C.c_int64),
                ("pred_fallbacks", C.c_int64), ("pred_last_err", C.c_double),
                ("pred_mlp_rms", C.c_double), ("pred_base_rms", C.c_double),
                ("pred_judged", C.c_int64), ("bs_pairs_all", C.c_double),
                ("bs_launches_all", C.c_int64)]


_host = None


def load():
    global _host
    if _host is None:
        vlm.load_library()
        if not os.path.exists(HOST_LIB):
            raise vlm.VLMError(f"Fortran host not built at {HOST_LIB}")
        lib = C.CDLL(HOST_LIB)
        lib.volcanor_run.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                     C.c_int, C.POINTER(Stats), C.c_int, C.c_int, C.c_double]
        lib.volcanor_run.restype = C.c_int
        lib.volcanor_last_error.argtypes = []
        lib.volcanor_last_error.restype = C.c_char_p
        _host = lib
    return _host


def run_case(casedir, precision="fp64", nt=0, verbose=0, nrotors=1, max_steps=20000,
             tail_steps=0, learned=False, tol=5e-4, fixed_ploop=False):
    """Run the case; returns (ct_history[nrotors, nt+1], Stats).  A case the
    host rejects, or a failing library call, raises VLMError with the
    host's message; the process carries on.  fixed_ploop convects every
    row of the predicted near wake (the reference convects row 1 only,
    classdef.f90:1552-1560, kept by default)."""
    lib = load()
    hist = np.zeros((max_steps + 1, nrotors))  # Fortran (ct_cap, hist_len) order
    st = Stats()
    p = 1 if precision == "mixed" else 0
    rc = lib.volcanor_run(os.path.abspath(casedir).encode(), p, nt, verbose,
                          hist.ctypes.data_as(C.c_void_p), nrotors, max_steps + 1,
                          C.byref(st), tail_steps,
                          (1 if learned else 0) | (2 if fixed_ploop else 0), tol)
    if rc != 0:
        raise vlm.VLMError(f"volcanor_run failed: {lib.volcanor_last_error().decode()}")
    return hist[: st.nt + 1].T.copy(), st
