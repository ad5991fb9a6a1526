"""Parity of the device-resident free wake (lattice form, ring buffer,
unique segments) with the ring-form CPU oracle, step by step."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests import wake_harness as H

pytestmark = pytest.mark.gpu


def _ctx(precision=0):
    from volcanor_amd import vlm
    c = vlm.Context(0)
    c.set_precision(precision)
    return c


def _cmp(a, b, scale, tol, what):
    err = np.max(np.abs(a - b)) if a.size else 0.0
    assert err <= tol * max(scale, 1e-300), f"{what}: max err {err:.3e} scale {scale:.3e}"


def compare(sa, sb, case, tol=1e-11):
    for k, (ra, rb) in enumerate(zip(sa, sb)):
        c = case.surface(k)
        assert ra["cursors"] == rb["cursors"]
        rn, rf = ra["cursors"][0] - 1, ra["cursors"][1] - 1
        if rn < c.nNwake:
            sc = np.max(np.abs(ra["nodes"][:, rn:])) + 1.0
            _cmp(ra["nodes"][:, rn:], rb["nodes"][:, rn:], sc, tol, "near nodes")
            _cmp(ra["le"], rb["le"], sc, tol, "shed line")
            sg = np.max(np.abs(ra["gam"][:, rn:])) + 1e-30
            _cmp(ra["gam"][:, rn:], rb["gam"][:, rn:], sg, tol, "near gam")
            _cmp(ra["rc4"][:, rn:], rb["rc4"][:, rn:], 1.0, tol, "near cores")
        if rf < c.nFwake:
            sc = np.max(np.abs(ra["far_fc"][:, rf:])) + 1.0
            _cmp(ra["far_fc"][:, rf:], rb["far_fc"][:, rf:], sc, tol, "far ends")
            _cmp(ra["far_gam"][:, rf:], rb["far_gam"][:, rf:], 1.0, tol, "far gam")
            _cmp(ra["far_rc"][:, rf:], rb["far_rc"][:, rf:], 1.0, tol, "far cores")
        for k in ("vn", "vn1", "vf", "vf1"):
            sv = np.max(np.abs(rb[k])) + 1e-30 if rb[k].size else 1.0
            _cmp(ra[k], rb[k], sv, tol, k)
        for t in (0, 1):
            if f"presc{t}" in rb:
                fa, ga, ca, ba = ra[f"presc{t}"]
                fb, gb, cb, bb = rb[f"presc{t}"]
                assert ba == bb, f"prescribed wake {t} built"
                _cmp(fa, fb, np.max(np.abs(fb)) + 1.0, tol, f"prescribed wake {t} ends")
                _cmp(ga, gb, np.max(np.abs(gb)) + 1e-30, tol, f"prescribed wake {t} gam")
                _cmp(ca, cb, 1.0, tol, f"prescribed wake {t} cores")


CASES = {
    "wing_am2": dict(nb=1, nc=3, ns=5, nNwake=5, nFwake=4, nt=12, scheme=3, vel=(10, 0, 0)),
    "rotor_am2": dict(nb=2, nc=1, ns=6, nNwake=6, nFwake=4, nt=13, scheme=3, omega=30.0,
                      vel=(0, 0, 1.0)),
    "rotor_am2_axisym_diss_strain": dict(nb=3, nc=1, ns=5, nNwake=4, nFwake=5, nt=12, scheme=3,
                                         omega=25.0, vel=(0, 0, 2.0), axisym=True,
                                         dissipation=True, strain=True),
    "rotor_am2_fixed_ploop": dict(nb=2, nc=2, ns=4, nNwake=5, nFwake=3, nt=11, scheme=3,
                                  omega=30.0, vel=(0, 0, 1.0), bug_compat=False,
                                  dissipation=True),
    "wing_ft1": dict(nb=1, nc=2, ns=4, nNwake=4, nFwake=3, nt=9, scheme=0),
    "wing_pc2_strain": dict(nb=1, nc=2, ns=4, nNwake=4, nFwake=4, nt=10, scheme=1, strain=True),
    "rotor_ab2": dict(nb=2, nc=1, ns=4, nNwake=4, nFwake=3, nt=9, scheme=2, omega=20.0,
                      vel=(0, 0, 1.0)),
    "two_rotors_am2": dict(nb=2, nc=1, ns=4, nNwake=4, nFwake=3, nt=9, scheme=3, omega=25.0,
                           vel=(0, 0, 1.0), nr=2, hub_dx=2.5, dissipation=True),
    "nonuniform_cores": dict(nb=1, nc=2, ns=4, nNwake=4, nFwake=2, nt=8, scheme=3,
                             scv=np.array([0.01, 0.02, 0.03, 0.02, 0.01]), dissipation=True),
    "truncated_no_far": dict(nb=1, nc=1, ns=3, nNwake=3, nFwake=0, nt=7, scheme=3, rollup=(2, 3)),
    # prescribed far wake rebuilt after both AM2 convections from step 9:
    # from every active far filament, and with axisymmetric blades from the
    # last four (prescWakeGenNt 3)
    "rotor_am2_presc": dict(nb=2, nc=1, ns=4, nNwake=4, nFwake=5, nt=14, scheme=3, omega=30.0,
                            vel=(0, 0, 1.0), presc=(8, 0)),
    "rotor_am2_axisym_presc": dict(nb=3, nc=1, ns=4, nNwake=4, nFwake=6, nt=14, scheme=3,
                                   omega=25.0, vel=(0, 0, 2.0), axisym=True, dissipation=True,
                                   presc=(8, 3)),
    # image rotors (rotor_mirrorGamma / rotor_mirrorWake, classdef.f90:4990-5131):
    # the mirror after every convection stage of the scheme and of the whole
    # wake at the end of the step, beside the lifting source
    "image_ft1_z": dict(nb=2, nc=1, ns=4, nNwake=4, nFwake=5, nt=14, scheme=0, omega=30.0,
                        vel=(0, 0, 1.0), hub=[0.0, 0.0, 0.5],
                        surfaces=[{}, {"image": (0, 3)}]),
    "image_pc2_z": dict(nb=2, nc=1, ns=4, nNwake=4, nFwake=5, nt=14, scheme=1, omega=30.0,
                        vel=(0, 0, 1.0), hub=[0.0, 0.0, 0.5],
                        surfaces=[{}, {"image": (0, 3)}]),
    "image_am2_z": dict(nb=2, nc=1, ns=4, nNwake=4, nFwake=5, nt=14, scheme=3, omega=30.0,
                        vel=(0, 0, 1.0), hub=[0.0, 0.0, 0.5], dissipation=True, strain=True,
                        surfaces=[{}, {"image": (0, 3)}]),
    "image_ab2_xy_axisym": dict(nb=3, nc=1, ns=4, nNwake=4, nFwake=5, nt=14, scheme=2,
                                omega=25.0, vel=(0, 0, 2.0), hub=[1.5, 1.8, 0.0], axisym=True,
                                surfaces=[{}, {"image": (0, 1)}, {"image": (0, 2)}]),
}


@pytest.mark.parametrize("name", list(CASES))
def test_wake_steps_match_oracle(name):
    case = H.Case(**CASES[name])
    dev, ref = _ctx(), orc.System()
    ia, ib = H.setup(dev, case), H.setup(ref, case)
    for it in range(1, case.nt + 1):
        H.step(dev, case, ia, it)
        H.step(ref, case, ib, it)
        compare(H.snapshot(dev, case, ia), H.snapshot(ref, case, ib), case)


def test_mixed_mode_small_launches_stay_fp64():
    """In MIXED mode, launches with fewer than 1024 targets (here every wake
    pass: at most 84 nodes) run the FP64 small-target kernel, so the wake
    stays at FP64 parity with the oracle and no work-queue launch happens."""
    case = H.Case(**CASES["rotor_am2"])
    dev, ref = _ctx(precision=1), orc.System()
    dev.profile_enable(True)
    dev.profile_reset()
    ia, ib = H.setup(dev, case), H.setup(ref, case)
    for it in range(1, case.nt + 1):
        H.step(dev, case, ia, it)
        H.step(ref, case, ib, it)
    compare(H.snapshot(dev, case, ia), H.snapshot(ref, case, ib), case, tol=1e-11)
    assert dev.profile_get()[1] == 0
    assert dev.profile_get_all()[1] > 0


def test_mixed_precision_work_queue_wake_close_to_oracle():
    """MIXED mode on wakes large enough for the work-queue kernel
    (bs_mixed_kernel, >= 1024 targets): a two-bladed rotor climbing at
    3 m/s, 2 x 12 panels, 60 near rows, AM2.

    Per call: the oracle's own wake states after steps 45 and 60 (1196 and
    1560 node targets on the segments they are summed against) are loaded
    into the device, and one MIXED wake pass is compared with the oracle's
    FP64 pass on the same state: max error within 3e-5 of the largest
    speed and RMS error within 1e-6 of the RMS speed.

    Over the run: the device steps in MIXED beside the oracle for 60 steps;
    node positions stay within 3e-6 of their scale and velocities within
    1e-4 of the largest.  The climb keeps the wake below the rotor, where
    the oracle's own history moves by 3e-11 under a 1e-12 change of the
    circulation; the same rotor descending at 1 m/s into its wake
    amplifies that change to 2.5e-4 in 60 steps."""
    from tests.geometry import rings_from_grids
    case = H.Case(nb=2, nc=1, ns=12, nNwake=60, nFwake=0, nt=60, scheme=3, omega=30.0,
                  vel=(0.0, 0.0, -3.0), rollup=(9, 12))
    dev, ref = _ctx(precision=1), orc.System()
    probe = _ctx(precision=1)
    dev.profile_enable(True)
    dev.profile_reset()
    ia, ib, ip = H.setup(dev, case), H.setup(ref, case), H.setup(probe, case)
    worst_node = worst_vel = 0.0
    for it in range(1, case.nt + 1):
        H.step(dev, case, ia, it)
        H.step(ref, case, ib, it)
        ra, rb = H.snapshot(dev, case, ia)[0], H.snapshot(ref, case, ib)[0]
        assert ra["cursors"] == rb["cursors"]
        rn = ra["cursors"][0] - 1
        sc = np.max(np.abs(rb["nodes"][:, rn:])) + 1.0
        worst_node = max(worst_node, np.max(np.abs(ra["nodes"][:, rn:] - rb["nodes"][:, rn:])) / sc)
        for k in ("vn", "vn1"):
            sv = np.max(np.abs(rb[k])) + 1e-30
            worst_vel = max(worst_vel, np.max(np.abs(ra[k] - rb[k])) / sv)
        if it in (45, 60):
            # one pass on the oracle's state: the oracle's wake and wing
            w = ref.rotor_get_wake(ib[0], case.dims)
            rnr, rfr = ref.rotor_cursors(ib[0])
            probe.rotor_set_wake(ip[0], w, rnr, rfr)
            grids, _ = H.blade_grids(case, it * case.dt, 0)
            probe.rotor_set_wing(ip[0], rings_from_grids(grids), gam=H.wing_gamma(case, it, 0))
            probe.wake_velocity(0)
            ref.wake_velocity(0)
            vp = probe.rotor_get_velocity(ip[0], 0, case.dims)[0][:, rnr - 1:]
            vr = ref.rotor_get_velocity(ib[0], 0, case.dims)[0][:, rnr - 1:]
            d = np.linalg.norm(vp - vr, axis=-1)
            sp = np.linalg.norm(vr, axis=-1)
            e_max = np.max(d) / np.max(sp)
            e_rms = np.sqrt(np.mean(d ** 2) / np.mean(sp ** 2))
            print(f"mixed pass on the oracle state at step {it}: {vr.size // 3} targets, "
                  f"max err {e_max:.2e}, rms err {e_rms:.2e}")
            assert e_max <= 3e-5 and e_rms <= 1e-6
    launches = dev.profile_get()[1]
    print(f"mixed work-queue wake over 60 steps: node err {worst_node:.2e}, velocity err "
          f"{worst_vel:.2e}, work-queue launches {launches}")
    assert launches >= 2 * 20          # both AM2 passes of the last 20 steps at least
    assert worst_node <= 3e-6 and worst_vel <= 1e-4


def test_surfaces_with_different_spanwise_counts_match_oracle():
    """A three-bladed rotor with 30 spanwise panels above a 4 x 40 wing
    (cfg5's discretisation), both with free wakes that roll up into far
    wakes, coupled through every Biot-Savart pass (main.f90:549-561,
    libCommon.f90:114-211), step by step against the oracle."""
    case = H.Case(scheme=3, nt=9, dt=0.005, vel=(13.6, 0.0, 0.0), surfaces=[
        dict(nb=3, nc=1, ns=30, nNwake=5, nFwake=3, omega=136.0, span=1.0, root=0.1,
             chord=1 / 6, pitch=10.0, rollup=(22, 30), hub=(0.0, 0.0, 0.0)),
        dict(nb=1, nc=4, ns=40, nNwake=5, nFwake=3, omega=0.0, span=2.0, root=-1.0,
             chord=0.25, pitch=3.0, rollup=(30, 40), hub=(0.0, 0.0, -0.5))])
    dev, ref = _ctx(), orc.System()
    ia, ib = H.setup(dev, case), H.setup(ref, case)
    for it in range(1, case.nt + 1):
        H.step(dev, case, ia, it)
        H.step(ref, case, ib, it)
        compare(H.snapshot(dev, case, ia), H.snapshot(ref, case, ib), case)


def test_induced_velocity_sets_match_oracle():
    case = H.Case(**CASES["rotor_am2"])
    dev, ref = _ctx(), orc.System()
    ia, ib = H.setup(dev, case), H.setup(ref, case)
    for it in range(1, 9):
        H.step(dev, case, ia, it)
        H.step(ref, case, ib, it)
    p = np.random.default_rng(0).normal(size=(37, 3)) * 0.8
    for s in (0, 1, 2):
        a = dev.rotor_induced_velocity(ia[0], s, p)
        b = ref.rotor_induced_velocity(ib[0], s, p)
        np.testing.assert_allclose(a, b, rtol=1e-11, atol=1e-13 * np.max(np.abs(b)))
    # accumulate adds into the given array
    a2 = dev.rotor_induced_velocity(ia[0], 2, p, v=np.ones_like(p))
    np.testing.assert_allclose(a2, 1.0 + ref.rotor_induced_velocity(ib[0], 2, p), rtol=1e-11)


def test_induced_velocity_terms_equal_sequential_calls():
    """vlm_induced_velocity_terms against one vlm_rotor_induced_velocity call
    per term with the negation done on the host: bitwise equal, for a
    target count on the small-target path (37) and on the work-queue path
    (1500), with and without accumulate."""
    case = H.Case(**CASES["rotor_am2"])
    dev = _ctx()
    ia = H.setup(dev, case)
    for it in range(1, 9):
        H.step(dev, case, ia, it)
    r = ia[0]
    terms = [(r, 2, False), (r, 1, False), (r, 0, True), (r, 2, True)]
    rng = np.random.default_rng(5)
    for m in (37, 1500):
        p = rng.normal(size=(m, 3)) * 0.8
        for v0 in (None, rng.normal(size=(m, 3))):
            seq = None if v0 is None else v0.copy()
            for ir, st, neg in terms:
                if neg:
                    seq = -seq
                seq = dev.rotor_induced_velocity(ir, st, p, v=seq if seq is not None else np.zeros_like(p))
            got = dev.induced_velocity_terms(terms, p, v=v0)
            assert np.array_equal(got, seq), (m, v0 is None, np.max(np.abs(got - seq)))
    assert np.array_equal(dev.induced_velocity_terms([], p), np.zeros_like(p))
    with pytest.raises(Exception):
        dev.induced_velocity_terms([(r + 7, 2, False)], p)
    with pytest.raises(Exception):
        dev.induced_velocity_terms([(r, 9, False)], p)


def test_deferred_results_equal_plain_calls():
    """vlm_rotor_solve_deferred and vlm_induced_velocity_terms_deferred give
    the plain calls' results once vlm_wait returns, with device work
    enqueued in between; a synchronising call also completes them."""
    case = H.Case(**CASES["rotor_am2"])
    dev = _ctx()
    ia = H.setup(dev, case)
    for it in range(1, 6):
        H.step(dev, case, ia, it)
    r = ia[0]
    terms = [(r, 1, False), (r, 0, True)]
    p = np.random.default_rng(9).normal(size=(1200, 3)) * 0.8
    n = case.nb * case.nc * case.ns
    cp = np.random.default_rng(11).normal(size=(n, 3))
    ncap = np.random.default_rng(12).normal(size=(n, 3))
    ncap /= np.linalg.norm(ncap, axis=1)[:, None]
    dev.rotor_calc_aic(r, cp, ncap, n)
    rhs = np.random.default_rng(10).normal(size=n)
    g0, it0 = dev.rotor_solve(r, rhs)    # also sets the wing circulation
    plain = dev.induced_velocity_terms(terms, p)
    v = dev.induced_velocity_terms_deferred(terms, p)
    g, it = dev.rotor_solve_deferred(r, rhs)
    dev.convect(3, 6, case.dt)           # device work after the deferred copies
    dev.wait()
    assert np.array_equal(v, plain) and np.array_equal(g, g0) and int(it[0]) == it0
    v2 = dev.induced_velocity_terms_deferred(terms, p)
    dev.synchronize()
    assert np.array_equal(v2, dev.induced_velocity_terms(terms, p))


def test_set_wake_roundtrip_and_errors():
    from volcanor_amd import vlm
    case = H.Case(**CASES["wing_am2"])
    dev = _ctx()
    ia = H.setup(dev, case)
    for it in range(1, 8):
        H.step(dev, case, ia, it)
    w = dev.rotor_get_wake(ia[0], case.dims)
    rn, rf = dev.rotor_cursors(ia[0])
    dev2 = _ctx()
    ib = H.setup(dev2, case)
    dev2.rotor_set_wake(ib[0], w, rn, rf)
    w2 = dev2.rotor_get_wake(ib[0], case.dims)
    for k in ("nodes", "le", "gam", "rc4"):
        assert np.array_equal(w[k][:, rn - 1:], w2[k][:, rn - 1:]), k
    with pytest.raises(vlm.VLMError):
        dev2.rotor_set_wake(ib[0], w, 0, rf)
    with pytest.raises(vlm.VLMError):
        dev2.rotor_assignshed(ib[0], 7)
    with pytest.raises(vlm.VLMError):
        dev2.convect(9, 1, 0.1)
