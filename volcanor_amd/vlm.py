"""ctypes binding of libvlm.so (include/vlm_c.h) for Python callers.

The product path is the HIP library; this module only marshals numpy
arrays across the C ABI.  It raises when the library is missing instead of
falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libvlm.so")

VLM_OK = 0
FP64, MIXED = 0, 1
SET_WING, SET_WING_BOUND, SET_WAKE, SET_WAKE_PRED = 0, 1, 2, 3
WAKE_C, WAKE_P = 0, 1
EDGE_LE, EDGE_TE = 0, 1
TERM_NEGATE = 1
SCHEME_FT1, SCHEME_PC2, SCHEME_AB2, SCHEME_AM2, SCHEME_AM3, SCHEME_AM4 = 0, 1, 2, 3, 4, 5
VEL, VEL_PREV, VEL_PRED, VEL_STEP = 0, 1, 2, 3


class RotorDesc(C.Structure):
    _fields_ = [
        ("nb", C.c_int), ("nc", C.c_int), ("ns", C.c_int),
        ("nNwake", C.c_int), ("nFwake", C.c_int), ("nbConvect", C.c_int),
        ("rollupStart", C.c_int), ("rollupEnd", C.c_int),
        ("rollupMin", C.c_int), ("suppressFwake", C.c_int),
    ]


class VLMError(RuntimeError):
    pass


_P = C.c_void_p
_I = C.c_int
_L = C.c_int64
_D = C.c_double

_SIGS = {
    "vlm_create": [C.POINTER(_P), _I],
    "vlm_destroy": [_P],
    "vlm_set_precision": [_P, _I],
    "vlm_set_bug_compat": [_P, _I],
    "vlm_synchronize": [_P],
    "vlm_filament_vind": [_P, _L, _P, _P, _P, _P, _L, _P, _P, _I],
    "vlm_filament_vind_noalloc": [_P, _L, _P, _P, _P, _P, _L, _P, _P, _I],
    "vlm_fset_create": [_P, _L, _P, _P, _P, _P, C.POINTER(_I)],
    "vlm_pset_create": [_P, _L, _P, C.POINTER(_I)],
    "vlm_fset_vind": [_P, _I, _I],
    "vlm_pset_velocity": [_P, _I, _P],
    "vlm_rotor_create": [_P, C.POINTER(RotorDesc), _D, _P, C.POINTER(_I)],
    "vlm_rotor_set_wing": [_P, _I, _P, _P, _P],
    "vlm_rotor_set_axis": [_P, _I, _P, _P],
    "vlm_rotor_set_init_wake_vel": [_P, _I, _D, _I],
    "vlm_rotor_set_prescribed_wake": [_P, _I, _I, _I, _D],
    "vlm_rotor_get_prescribed_wake": [_P, _I, _I, _P, _P, _P, _P],
    "vlm_rotor_init_cursors": [_P, _I],
    "vlm_rotor_begin_step": [_P, _I, _I],
    "vlm_rotor_cursors": [_P, _I, C.POINTER(_I), C.POINTER(_I)],
    "vlm_rotor_assignshed": [_P, _I, _I],
    "vlm_rotor_age_wake": [_P, _I, _D],
    "vlm_rotor_dissipate_wake": [_P, _I, _D, _D, _D, _D],
    "vlm_rotor_strain_wake": [_P, _I],
    "vlm_rotor_set_image": [_P, _I, _I, _I],
    "vlm_rotor_mirror_gamma": [_P, _I],
    "vlm_rotor_mirror_wake": [_P, _I, _I],
    "vlm_rotor_burst_wake": [_P, _I, _D, _D],
    "vlm_rotor_state_bytes": [_P, _I, C.POINTER(_L)],
    "vlm_rotor_get_state": [_P, _I, _P, _L],
    "vlm_rotor_set_state": [_P, _I, _P, _L],
    "vlm_rotor_rollup": [_P, _I],
    "vlm_rotor_convectwake": [_P, _I, _D, _I],
    "vlm_rotor_induced_velocity": [_P, _I, _I, _L, _P, _P, _I],
    "vlm_induced_velocity_terms": [_P, _I, _P, _P, _P, _L, _P, _P, _I],
    "vlm_induced_velocity_terms_deferred": [_P, _I, _P, _P, _P, _L, _P, _P, _I],
    "vlm_rotor_solve_deferred": [_P, _I, _P, _P, _I, _D, _P],
    "vlm_rotor_set_panels": [_P, _I, _P, _P, _P],
    "vlm_rotor_rhs": [_P, _I, _I, _P, _P, _P, _P, _P],
    "vlm_rotor_mirror_velcp": [_P, _I],
    "vlm_rotor_loads": [_P, _I, _I, _P, _P, _P, _D, _D, _D, _I, _P],
    "vlm_wait": [_P],
    "vlm_wake_velocity": [_P, _I],
    "vlm_convect": [_P, _I, _I, _D],
    "vlm_rotor_get_wake": [_P, _I, _I, _P, _P, _P, _P, _P, _P, _P],
    "vlm_rotor_set_wake": [_P, _I, _P, _P, _P, _P, _P, _P, _P, _I, _I],
    "vlm_rotor_get_velocity": [_P, _I, _I, _P, _P],
    "vlm_rotor_calc_aic": [_P, _I, _P, _P, _P],
    "vlm_rotor_solve": [_P, _I, _P, _P, _I, _D, C.POINTER(_I)],
    "vlm_lu_factor": [_P, _I, _P, _P],
    "vlm_lu_solve_refine": [_P, _I, _P, _P, _P, _P, _P, _I, _D, C.POINTER(_I)],
    "vlm_mlp_num_params": [],
    "vlm_mlp_set_params": [_P, _P],
    "vlm_mlp_get_params": [_P, _P],
    "vlm_mlp_forward": [_P, _L, _P, _P],
    "vlm_mlp_loss_grad": [_P, _L, _P, _P, _P, C.POINTER(C.c_float)],
    "vlm_mlp_train_step": [_P, _L, _P, _P, C.c_float, C.POINTER(C.c_float)],
    "vlm_set_predictor": [_P, _I, _D, _I, _I, _I, _D],
    "vlm_predictor_stats": [_P, C.POINTER(_L), C.POINTER(_L), C.POINTER(_L), C.POINTER(_D),
                            C.POINTER(_D)],
    "vlm_predictor_errors": [_P, C.POINTER(_D), C.POINTER(_D), C.POINTER(_L)],
    "vlm_profile_enable": [_P, _I],
    "vlm_profile_reset": [_P],
    "vlm_profile_get": [_P, C.POINTER(_D), C.POINTER(_L), C.POINTER(_D)],
    "vlm_profile_get_all": [_P, C.POINTER(_D), C.POINTER(_L)],
    "vlm_profile_get_lu": [_P, C.POINTER(_D), C.POINTER(_L), C.POINTER(_D)],
    "vlm_fp64_issue_rate": [_P, C.POINTER(_D), C.POINTER(_D)],
}

_lib = None


def load_library(path: str = LIB_PATH):
    """Load libvlm.so and declare every entry point; raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise VLMError(f"libvlm.so not built at {path}; run __graft_entry__.build()")
    lib = C.CDLL(path)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.vlm_last_error.argtypes = [_P]
    lib.vlm_last_error.restype = C.c_char_p
    _lib = lib
    return lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


class Context:
    """One HIP device context (vlm_ctx)."""

    def __init__(self, device: int = 0):
        self._pending = []  # outputs of deferred calls, alive until filled
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.vlm_create(C.byref(h), device)
        if rc != VLM_OK:
            raise VLMError(f"vlm_create failed with code {rc}")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.vlm_destroy(self.h)
            self.h = None
        self._pending = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != VLM_OK:
            msg = self.lib.vlm_last_error(self.h)
            raise VLMError(f"code {rc}: {msg.decode() if msg else ''}")

    # -- configuration -------------------------------------------------
    def set_precision(self, p):
        self._chk(self.lib.vlm_set_precision(self.h, p))

    def set_bug_compat(self, on):
        self._chk(self.lib.vlm_set_bug_compat(self.h, int(on)))

    def synchronize(self):
        self._chk(self.lib.vlm_synchronize(self.h))
        self._pending = []

    # -- explicit filament lists -----------------------------------------
    def filament_vind(self, a, b, gam, rc, p, v=None):
        a, b = _f64(a, (-1, 3)), _f64(b, (-1, 3))
        gam, rc = _f64(gam).ravel(), _f64(rc).ravel()
        p = _f64(p, (-1, 3))
        acc = v is not None
        out = _f64(v, (-1, 3)).copy() if acc else np.zeros_like(p)
        self._chk(self.lib.vlm_filament_vind(self.h, a.shape[0], _ptr(a), _ptr(b), _ptr(gam),
                                             _ptr(rc), p.shape[0], _ptr(p), _ptr(out), int(acc)))
        return out

    def fset_create(self, a, b, gam, rc):
        a, b = _f64(a, (-1, 3)), _f64(b, (-1, 3))
        gam, rc = _f64(gam).ravel(), _f64(rc).ravel()
        k = C.c_int()
        self._chk(self.lib.vlm_fset_create(self.h, a.shape[0], _ptr(a), _ptr(b), _ptr(gam),
                                           _ptr(rc), C.byref(k)))
        return k.value

    def pset_create(self, p):
        p = _f64(p, (-1, 3))
        k = C.c_int()
        self._chk(self.lib.vlm_pset_create(self.h, p.shape[0], _ptr(p), C.byref(k)))
        return k.value, p.shape[0]

    def fset_vind(self, fset, pset):
        self._chk(self.lib.vlm_fset_vind(self.h, fset, pset))

    def pset_velocity(self, pset, m):
        v = np.zeros((m, 3))
        self._chk(self.lib.vlm_pset_velocity(self.h, pset, _ptr(v)))
        return v

    # -- rotors ----------------------------------------------------------
    def rotor_create(self, nb, nc, ns, nNwake, nFwake, nbConvect=None, rollupStart=1,
                     rollupEnd=None, rollupMin=1, suppressFwake=0, spanwiseCore=0.0,
                     streamwiseCoreVec=None):
        d = RotorDesc(nb, nc, ns, nNwake, nFwake, nb if nbConvect is None else nbConvect,
                      rollupStart, ns if rollupEnd is None else rollupEnd, rollupMin,
                      suppressFwake)
        scv = _f64(np.zeros(ns + 1) if streamwiseCoreVec is None else streamwiseCoreVec)
        k = C.c_int()
        self._chk(self.lib.vlm_rotor_create(self.h, C.byref(d), spanwiseCore, _ptr(scv),
                                            C.byref(k)))
        return k.value

    def rotor_set_wing(self, ir, corners=None, rc=None, gam=None):
        c = None if corners is None else _f64(corners).ravel()
        r = None if rc is None else _f64(rc).ravel()
        g = None if gam is None else _f64(gam).ravel()
        self._chk(self.lib.vlm_rotor_set_wing(self.h, ir, _ptr(c), _ptr(r), _ptr(g)))

    def rotor_set_axis(self, ir, shaft, hub):
        s, h = _f64(shaft).ravel(), _f64(hub).ravel()
        self._chk(self.lib.vlm_rotor_set_axis(self.h, ir, _ptr(s), _ptr(h)))

    def rotor_set_init_wake_vel(self, ir, init_wake_vel, init_wake_vel_nt):
        self._chk(self.lib.vlm_rotor_set_init_wake_vel(self.h, ir, init_wake_vel,
                                                       init_wake_vel_nt))

    def rotor_init_cursors(self, ir):
        self._chk(self.lib.vlm_rotor_init_cursors(self.h, ir))

    def rotor_set_prescribed_wake(self, ir, presc_nt, presc_gen_nt, dpsi):
        self._chk(self.lib.vlm_rotor_set_prescribed_wake(self.h, ir, presc_nt, presc_gen_nt,
                                                         dpsi))

    def rotor_get_prescribed_wake(self, ir, nb, wake_type=WAKE_C):
        """(fc (nb, 240, 2, 3), gam (nb, 240), rc (nb, 240), built)"""
        fc = np.zeros((nb, 240, 2, 3))
        g = np.zeros((nb, 240))
        rc = np.zeros((nb, 240))
        built = C.c_int(0)
        self._chk(self.lib.vlm_rotor_get_prescribed_wake(self.h, ir, wake_type, _ptr(fc),
                                                         _ptr(g), _ptr(rc), C.byref(built)))
        return fc, g, rc, bool(built.value)

    def rotor_begin_step(self, ir, it):
        self._chk(self.lib.vlm_rotor_begin_step(self.h, ir, it))

    def rotor_cursors(self, ir):
        a, b = C.c_int(), C.c_int()
        self._chk(self.lib.vlm_rotor_cursors(self.h, ir, C.byref(a), C.byref(b)))
        return a.value, b.value

    def rotor_assignshed(self, ir, edge):
        self._chk(self.lib.vlm_rotor_assignshed(self.h, ir, edge))

    def rotor_age_wake(self, ir, dt):
        self._chk(self.lib.vlm_rotor_age_wake(self.h, ir, dt))

    def rotor_dissipate_wake(self, ir, dt, nu, avc, decay=0.0):
        self._chk(self.lib.vlm_rotor_dissipate_wake(self.h, ir, dt, nu, avc, decay))

    def rotor_strain_wake(self, ir):
        self._chk(self.lib.vlm_rotor_strain_wake(self.h, ir))

    def rotor_set_image(self, ir, source, plane):
        """Rotor ir mirrors rotor source in the plane where coordinate
        plane (1..3) is zero."""
        self._chk(self.lib.vlm_rotor_set_image(self.h, ir, source, plane))

    def rotor_mirror_gamma(self, ir):
        self._chk(self.lib.vlm_rotor_mirror_gamma(self.h, ir))

    def rotor_mirror_wake(self, ir, mode):
        """mode 0 current rows, 1 predicted rows, 2 every row (VLM_WAKE_A)."""
        self._chk(self.lib.vlm_rotor_mirror_wake(self.h, ir, mode))

    def rotor_burst_wake(self, ir, skew_limit, large_core):
        self._chk(self.lib.vlm_rotor_burst_wake(self.h, ir, skew_limit, large_core))

    def rotor_get_state(self, ir):
        """Opaque snapshot (bytes) of the rotor's device state, for restarts."""
        n = C.c_int64()
        self._chk(self.lib.vlm_rotor_state_bytes(self.h, ir, C.byref(n)))
        buf = np.zeros(n.value, dtype=np.uint8)
        self._chk(self.lib.vlm_rotor_get_state(self.h, ir, _ptr(buf), n.value))
        return buf

    def rotor_set_state(self, ir, buf):
        buf = np.ascontiguousarray(buf, dtype=np.uint8)
        self._chk(self.lib.vlm_rotor_set_state(self.h, ir, _ptr(buf), buf.size))

    def rotor_rollup(self, ir):
        self._chk(self.lib.vlm_rotor_rollup(self.h, ir))

    def rotor_convectwake(self, ir, dt, wake_type):
        self._chk(self.lib.vlm_rotor_convectwake(self.h, ir, dt, wake_type))

    def rotor_induced_velocity(self, ir, sset, p, v=None):
        p = _f64(p, (-1, 3))
        acc = v is not None
        out = _f64(v, (-1, 3)).copy() if acc else np.zeros_like(p)
        self._chk(self.lib.vlm_rotor_induced_velocity(self.h, ir, sset, p.shape[0], _ptr(p),
                                                      _ptr(out), int(acc)))
        return out

    def induced_velocity_terms(self, terms, p, v=None):
        """terms: list of (ir, set, negate_before); see vlm_induced_velocity_terms."""
        p = _f64(p, (-1, 3))
        acc = v is not None
        out = _f64(v, (-1, 3)).copy() if acc else np.zeros_like(p)
        ir = np.array([t[0] for t in terms], dtype=np.int32)
        st = np.array([t[1] for t in terms], dtype=np.int32)
        fl = np.array([TERM_NEGATE if t[2] else 0 for t in terms], dtype=np.int32)
        self._chk(self.lib.vlm_induced_velocity_terms(self.h, len(terms), _ptr(ir), _ptr(st),
                                                      _ptr(fl), p.shape[0], _ptr(p), _ptr(out),
                                                      int(acc)))
        return out

    def wake_velocity(self, wake_type=WAKE_C):
        self._chk(self.lib.vlm_wake_velocity(self.h, wake_type))

    def convect(self, scheme, it, dt):
        self._chk(self.lib.vlm_convect(self.h, scheme, it, dt))

    def rotor_get_wake(self, ir, dims, wake_type=WAKE_C):
        nb, ns, nN, nF = dims
        nodes = np.zeros((nb, nN, ns + 1, 3))
        le = np.zeros((nb, ns + 1, 3))
        gam = np.zeros((nb, nN, ns))
        rc4 = np.zeros((nb, nN, ns, 4))
        ffc = np.zeros((nb, max(nF, 0), 2, 3))
        fg = np.zeros((nb, max(nF, 0)))
        frc = np.zeros((nb, max(nF, 0)))
        self._chk(self.lib.vlm_rotor_get_wake(self.h, ir, wake_type, _ptr(nodes), _ptr(le),
                                              _ptr(gam), _ptr(rc4), _ptr(ffc), _ptr(fg),
                                              _ptr(frc)))
        return dict(nodes=nodes, le=le, gam=gam, rc4=rc4, far_fc=ffc, far_gam=fg, far_rc=frc)

    def rotor_set_wake(self, ir, w, rowNear, rowFar):
        arrs = [_f64(w[k]) for k in ("nodes", "le", "gam", "rc4", "far_fc", "far_gam",
                                      "far_rc")]
        self._chk(self.lib.vlm_rotor_set_wake(self.h, ir, *[_ptr(a) for a in arrs], rowNear,
                                              rowFar))

    def rotor_get_velocity(self, ir, which, dims):
        nb, ns, nN, nF = dims
        vn = np.zeros((nb, nN, ns + 1, 3))
        vf = np.zeros((nb, max(nF, 0), 3))
        self._chk(self.lib.vlm_rotor_get_velocity(self.h, ir, which, _ptr(vn), _ptr(vf)))
        return vn, vf

    # -- dense solve -----------------------------------------------------
    def rotor_calc_aic(self, ir, cp, ncap, n):
        cp, ncap = _f64(cp, (-1, 3)), _f64(ncap, (-1, 3))
        A = np.zeros((n, n), order="F")
        self._chk(self.lib.vlm_rotor_calc_aic(self.h, ir, _ptr(cp), _ptr(ncap), _ptr(A)))
        return A

    def rotor_solve(self, ir, rhs, max_iter=3, tol=1e-15):
        rhs = _f64(rhs).ravel()
        g = np.zeros_like(rhs)
        it = C.c_int()
        self._chk(self.lib.vlm_rotor_solve(self.h, ir, _ptr(rhs), _ptr(g), max_iter, tol,
                                           C.byref(it)))
        return g, it.value

    # -- deferred results: outputs are filled by wait() (or any call that
    #    synchronises); the arrays are kept alive here until then
    def rotor_solve_deferred(self, ir, rhs, max_iter=3, tol=1e-15):
        rhs = _f64(rhs).ravel()
        g = np.zeros_like(rhs)
        it = np.zeros(1, dtype=np.int32)
        self._chk(self.lib.vlm_rotor_solve_deferred(self.h, ir, _ptr(rhs), _ptr(g), max_iter,
                                                    tol, _ptr(it)))
        self._pending.extend([g, it])
        return g, it

    def rotor_set_panels(self, ir, panel, section, blade):
        """Panel (15), section (16) and blade (9) records after a rigid motion."""
        pn, sc, bl = _f64(panel, (-1, 15)), _f64(section, (-1, 16)), _f64(blade, (-1, 9))
        self._chk(self.lib.vlm_rotor_set_panels(self.h, ir, _ptr(pn), _ptr(sc), _ptr(bl)))

    @staticmethod
    def _terms(terms):
        ir = np.array([t[0] for t in terms], dtype=np.int32)
        st = np.array([t[1] for t in terms], dtype=np.int32)
        fl = np.array([TERM_NEGATE if t[2] else 0 for t in terms], dtype=np.int32)
        return ir, st, fl

    def rotor_rhs(self, ir, terms, kin, n):
        """Right-hand side of rotor ir (n panels) on the device; returns its
        host copy (waits).  kin: velBody, omegaBody, cgCoords, omegaSlow."""
        a, b, c = self._terms(terms)
        k = _f64(kin, (10,))
        out = np.zeros(n)
        self._chk(self.lib.vlm_rotor_rhs(self.h, ir, len(terms), _ptr(a), _ptr(b), _ptr(c),
                                         _ptr(k), _ptr(out)))
        self.wait()
        return out

    def rotor_mirror_velcp(self, ir):
        self._chk(self.lib.vlm_rotor_mirror_velcp(self.h, ir))

    def rotor_loads(self, ir, terms, density, dt, omega_sign, span_lift, nbc, ns):
        """Loads of rotor ir (vlm_rotor_loads records; waits): returns the
        (nbc*ns, 19) section records and the 9 rotor totals."""
        a, b, c = self._terms(terms)
        out = np.zeros(19 * nbc * ns + 9)
        self._chk(self.lib.vlm_rotor_loads(self.h, ir, len(terms), _ptr(a), _ptr(b), _ptr(c),
                                           density, dt, omega_sign, int(span_lift), _ptr(out)))
        self.wait()
        return out[:19 * nbc * ns].reshape(nbc * ns, 19), out[19 * nbc * ns:]

    def induced_velocity_terms_deferred(self, terms, p, v=None):
        p = _f64(p, (-1, 3))
        acc = v is not None
        out = _f64(v, (-1, 3)).copy() if acc else np.zeros_like(p)
        ir = np.array([t[0] for t in terms], dtype=np.int32)
        st = np.array([t[1] for t in terms], dtype=np.int32)
        fl = np.array([TERM_NEGATE if t[2] else 0 for t in terms], dtype=np.int32)
        self._chk(self.lib.vlm_induced_velocity_terms_deferred(
            self.h, len(terms), _ptr(ir), _ptr(st), _ptr(fl), p.shape[0], _ptr(p), _ptr(out),
            int(acc)))
        self._pending.append(out)
        return out

    def wait(self):
        self._chk(self.lib.vlm_wait(self.h))
        self._pending = []

    def lu_factor(self, A):
        A = np.asfortranarray(A, dtype=np.float64).copy(order="F")
        n = A.shape[0]
        piv = np.zeros(n, dtype=np.int32)
        self._chk(self.lib.vlm_lu_factor(self.h, n, _ptr(A), _ptr(piv)))
        return A, piv

    def lu_solve_refine(self, A, LU, piv, rhs, max_iter=3, tol=1e-15):
        A = np.asfortranarray(A, dtype=np.float64)
        LU = np.asfortranarray(LU, dtype=np.float64)
        piv = np.ascontiguousarray(piv, dtype=np.int32)
        rhs = _f64(rhs).ravel()
        x = np.zeros_like(rhs)
        it = C.c_int()
        self._chk(self.lib.vlm_lu_solve_refine(self.h, A.shape[0], _ptr(A), _ptr(LU), _ptr(piv),
                                               _ptr(rhs), _ptr(x), max_iter, tol, C.byref(it)))
        return x, it.value

    # -- learned predictor ------------------------------------------------
    def mlp_num_params(self):
        return self.lib.vlm_mlp_num_params()

    def mlp_set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self._chk(self.lib.vlm_mlp_set_params(self.h, _ptr(p)))

    def mlp_get_params(self):
        p = np.zeros(self.mlp_num_params(), dtype=np.float32)
        self._chk(self.lib.vlm_mlp_get_params(self.h, _ptr(p)))
        return p

    def mlp_forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 45)
        y = np.zeros((x.shape[0], 3), dtype=np.float32)
        self._chk(self.lib.vlm_mlp_forward(self.h, x.shape[0], _ptr(x), _ptr(y)))
        return y

    def mlp_loss_grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 45)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1, 3)
        g = np.zeros(self.mlp_num_params(), dtype=np.float32)
        loss = C.c_float()
        self._chk(self.lib.vlm_mlp_loss_grad(self.h, x.shape[0], _ptr(x), _ptr(t), _ptr(g),
                                             C.byref(loss)))
        return loss.value, g

    def mlp_train_step(self, x, t, lr=1e-3):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 45)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1, 3)
        loss = C.c_float()
        self._chk(self.lib.vlm_mlp_train_step(self.h, x.shape[0], _ptr(x), _ptr(t), lr,
                                              C.byref(loss)))
        return loss.value

    def set_predictor(self, mode, tol=5e-4, sample_stride=32, warmup=8, train_iters=4, lref=1.0):
        self._chk(self.lib.vlm_set_predictor(self.h, mode, tol, sample_stride, warmup,
                                             train_iters, lref))

    def predictor_stats(self):
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        e, l = C.c_double(), C.c_double()
        self._chk(self.lib.vlm_predictor_stats(self.h, C.byref(a), C.byref(b), C.byref(c),
                                               C.byref(e), C.byref(l)))
        return dict(full=a.value, learned=b.value, fallbacks=c.value, last_err=e.value,
                    last_loss=l.value)

    def predictor_errors(self):
        """Relative RMS error of the learned pass 2 and of the network-free
        substitute on the judge samples, and the number of judged steps."""
        a, b, n = C.c_double(), C.c_double(), C.c_int64()
        self._chk(self.lib.vlm_predictor_errors(self.h, C.byref(a), C.byref(b), C.byref(n)))
        return {"mlp": a.value, "base": b.value, "judged": n.value}

    # -- timing ----------------------------------------------------------
    def profile_enable(self, on=True):
        self._chk(self.lib.vlm_profile_enable(self.h, int(on)))

    def profile_reset(self):
        self._chk(self.lib.vlm_profile_reset(self.h))

    def profile_get(self):
        ms, n, pairs = C.c_double(), C.c_int64(), C.c_double()
        self._chk(self.lib.vlm_profile_get(self.h, C.byref(ms), C.byref(n), C.byref(pairs)))
        return ms.value, n.value, pairs.value

    def profile_get_all(self):
        """(pairs, launches) of every Biot-Savart launch since the reset,
        small-target launches included."""
        pairs, n = C.c_double(), C.c_int64()
        self._chk(self.lib.vlm_profile_get_all(self.h, C.byref(pairs), C.byref(n)))
        return pairs.value, n.value

    def fp64_issue_rate(self):
        """(TFLOP/s of a stream of independent FP64 FMAs on this device,
        time of the same stream of v_rsq_f64 relative to it)."""
        t, r = C.c_double(), C.c_double()
        self._chk(self.lib.vlm_fp64_issue_rate(self.h, C.byref(t), C.byref(r)))
        return t.value, r.value

    def profile_get_lu(self):
        """(device ms, factorisations, nominal flops) of LU factorisations."""
        ms, n, fl = C.c_double(), C.c_int64(), C.c_double()
        self._chk(self.lib.vlm_profile_get_lu(self.h, C.byref(ms), C.byref(n), C.byref(fl)))
        return ms.value, n.value, fl.value
