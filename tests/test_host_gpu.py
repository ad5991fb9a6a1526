"""End-to-end parity: the Fortran host + HIP library against the reference's
own fixtures and against histories of the reference solver compiled from
its sources (tests/golden/)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
CASES = os.path.join(ROOT, "cases")


def read_force(path):
    rows = [l.split() for l in open(path) if l.strip() and not l.lstrip().startswith("iter")]
    return np.array([[float(x) for x in r[1:]] for r in rows])


def run(name, precision="fp64", nt=0, nrotors=1):
    from volcanor_amd import host
    return host.run_case(os.path.join(CASES, name + ".case"), precision=precision, nt=nt,
                         nrotors=nrotors)


def reference_histories(name, nrotors):
    """CL/CT column of rXXForceNonDim.csv of the reference solver, per rotor."""
    out = []
    for ir in range(nrotors):
        f = f"{name}_ForceNonDim.csv" if ir == 0 else f"{name}_r{ir + 1:02d}ForceNonDim.csv"
        out.append(read_force(os.path.join(GOLD, f))[:, 0])
    return np.array(out)


def test_katz_plotkin_fixture_cl_history():
    """tests/test_katzNplotkinAR04.py of the reference: CL of the last step and
    the mean of the last 10 steps at places=6, here on the whole history."""
    ref = read_force(os.path.join(GOLD, "katz_plotkin_AR4_ForceNonDim.csv"))[:, 0]
    ct, st = run("katz_plotkin_AR4")
    cl = ct[0]
    assert len(cl) == len(ref) == 161
    assert round(cl[-1] - ref[-1], 6) == 0
    assert round(np.mean(cl[-10:]) - np.mean(ref[-10:]), 6) == 0
    err = np.abs(cl - ref)
    assert np.all(err <= 6e-7 * np.maximum(1.0, np.abs(ref))), (np.argmax(err), err.max())


def read_table(path):
    """Whitespace table with a header row -> {column: array}."""
    with open(path) as fh:
        names = fh.readline().split()
        rows = np.array([[float(x) for x in l.split()] for l in fh if l.strip()])
    return {n: rows[:, k] for k, n in enumerate(names)}


@pytest.mark.parametrize("case,fixture", [("katz_plotkin_AR4", "katz_plotkin_AR4_ForceDist00160.csv"),
                                          ("elevate", "elevate_ForceDist00150.csv")])
def test_fixture_sectional_loads(case, fixture):
    """testSect of the reference's tests/test_katzNplotkinAR04.py and
    tests/test_elevate.py: secSpan, secCL, secArea, secVel and secAlpha of
    the last step's rXXbYYForceDistNNNNN.csv at places=6; the file carries
    the reference's 16 columns.  The fixture is printed with 7 significant
    digits (E15.7), coarser than places=6 for values above 1: there a
    difference of one unit in the last printed digit is accepted, since two
    values 1e-8 apart print one unit apart when they straddle a rounding
    boundary (measured: elevate secAlpha 7.040416 against 7.040415)."""
    ref = read_table(os.path.join(GOLD, fixture))
    run(case)
    got = read_table(os.path.join(CASES, case + ".case", "Results",
                                  "r01b01" + fixture.split("_")[-1]))
    assert list(got) == list(ref)
    for k in ["secSpan", "secCL", "secArea", "secVel", "secAlpha"]:
        assert len(got[k]) == len(ref[k])
        for a, b in zip(got[k], ref[k]):
            unit = 10.0 ** (np.floor(np.log10(abs(b))) - 6) if b != 0 else 0.0
            assert round(a - b, 6) == 0 or abs(a - b) <= 1.01 * unit, (k, a, b)
    # the columns the reference's tests do not check, to the printed digits
    for k in ["secCLu", "secLift", "secLiftInPl", "secLiftOutPl", "secChord", "secTheta",
              "secPhi", "secViz", "secVix"]:
        np.testing.assert_allclose(got[k], ref[k], rtol=1e-5, atol=1e-6, err_msg=k)


@pytest.mark.parametrize("case,fixture", [("katz_plotkin_AR4", "katz_plotkin_AR4_Params.json"),
                                          ("elevate", "elevate_Params.json")])
def test_fixture_params(case, fixture):
    """testParams of the reference's Python tests: every key of the
    reference's rXXParams.json, floats at places=6, the rest equal."""
    import json
    ref = json.load(open(os.path.join(GOLD, fixture)))
    run(case)
    got = json.load(open(os.path.join(CASES, case + ".case", "Results", "r01Params.json")))
    for k, v in ref.items():
        assert k in got, k
        if isinstance(v, float):
            assert round(got[k] - v, 6) == 0, (k, got[k], v)
        else:
            assert got[k] == v, (k, got[k], v)


@pytest.fixture(scope="module")
def probes_run():
    run("probes_wing")
    return os.path.join(CASES, "probes_wing.case", "Results")


@pytest.mark.parametrize("step", ["00001", "00010", "00020"])
def test_velocity_probes_match_reference_solver(step, probes_run):
    """probes2file of the reference solver (libPostprocess.f90:333-361) on a
    short wing run with three fixed and three moving probes: velocity and
    position of every probe, to the 7 printed digits."""
    ref = read_table(os.path.join(GOLD, f"probes_wing_probes{step}.csv"))
    got = read_table(os.path.join(probes_run, f"probes{step}.csv"))
    assert list(got) == ["u", "v", "w", "x", "y", "z"] == list(ref)
    for k in got:
        np.testing.assert_allclose(got[k], ref[k], rtol=2e-6, atol=2e-7, err_msg=k)


def test_elevate_fixture_ct_history():
    """tests/test_elevate.py of the reference (axisymmetric rotor, PLOT3D
    blade, dissipation, truncated far wake)."""
    ref = read_force(os.path.join(GOLD, "elevate_ForceNonDim.csv"))[:, 0]
    ct, st = run("elevate")
    c = ct[0]
    assert len(c) == len(ref) == 151
    assert round(c[-1] - ref[-1], 6) == 0
    assert round(np.mean(c[-10:]) - np.mean(ref[-10:]), 6) == 0
    err = np.abs(c - ref)
    assert np.all(err <= 2e-8 + 1e-6 * np.abs(ref)), (np.argmax(err), err.max())


def test_cfg1_wing_matches_reference_solver():
    ref = read_force(os.path.join(GOLD, "cfg1_wing_ForceNonDim.csv"))[:, 0]
    ct, _ = run("cfg1_wing")
    np.testing.assert_allclose(ct[0], ref, rtol=1e-6, atol=1e-7)


def test_cfg2_two_revs_matches_reference_solver():
    path = os.path.join(GOLD, "cfg2_rotor_2rev_ForceNonDim.csv")
    ref = read_force(path)[:, 0]
    ct, _ = run("cfg2_rotor_2rev")
    assert len(ct[0]) == len(ref) == 145
    # printed digits of the reference output: 7 significant
    np.testing.assert_allclose(ct[0], ref, rtol=2e-6, atol=1e-9)


def test_cfg2_far_wake_matches_reference_solver():
    """cfg2 with a one-revolution near wake: from step 72 on, the oldest row
    rolls up into a far wake every step and strain (far wake only) stretches
    its cores; two revolutions against the reference solver."""
    ref = read_force(os.path.join(GOLD, "cfg2_rotor_farwake_2rev_ForceNonDim.csv"))[:, 0]
    ct, _ = run("cfg2_rotor_farwake_2rev")
    assert len(ct[0]) == len(ref) == 145
    np.testing.assert_allclose(ct[0], ref, rtol=2e-6, atol=1e-9)


def test_cfg2_mixed_precision_ct_within_1e4():
    """Mixed FP32/FP64 Biot-Savart against the FP64 path, two revolutions."""
    a, _ = run("cfg2_rotor_2rev", "fp64")
    b, _ = run("cfg2_rotor_2rev", "mixed")
    rel = np.abs(b[0] - a[0]) / np.abs(a[0])
    assert np.max(rel) < 1e-4


def test_learned_predictor_gain_over_extrapolation():
    """Learned pass 2 on the two-revolution cfg2 rotor (mixed precision,
    the bench's judge tolerance 5e-4).  The reference convects only row 1
    of the predicted near wake (classdef.f90:1552-1560), so before the
    near wake is full (here all 144 steps) pass 2 sees pass 1's geometry
    and returns pass 1's velocities; the recorded corrections are zero, the
    network's output is scaled by their RMS, and its prediction is pass 1
    itself.  At the last step the near wake is full (rowNear = 1), row 1
    moves, pass 2 differs from pass 1 and the judge sends that pass back to
    the full sum.  The judge scores the network and the network-free
    substitute of the same cost (linear extrapolation of the pass-1
    velocity, i.e. AM2 degraded to AB2) on the same exact samples: at most
    that one pass falls back, the network's accumulated error stays far
    below the extrapolation's (measured 1.8e-4 against 3.0e-1), and CT
    stays within 1e-3 of max |CT| of the reference solver."""
    from volcanor_amd import host
    ref = read_force(os.path.join(GOLD, "cfg2_rotor_2rev_ForceNonDim.csv"))[:, 0]
    ct, st = host.run_case(os.path.join(CASES, "cfg2_rotor_2rev.case"), precision="mixed",
                           learned=True, tol=5e-4)
    print(f"judged {st.pred_judged} learned {st.pred_learned} fallbacks {st.pred_fallbacks} "
          f"mlp {st.pred_mlp_rms:.3e} extrapolation {st.pred_base_rms:.3e}", flush=True)
    assert st.pred_judged > 100 and st.pred_fallbacks <= 1
    assert 0.0 <= st.pred_mlp_rms < 1e-3 and st.pred_mlp_rms < 0.01 * st.pred_base_rms
    assert len(ct[0]) == len(ref)
    assert np.max(np.abs(ct[0] - ref)) < 1e-3 * np.max(np.abs(ref))


@pytest.mark.parametrize("name", [f"scheme_{s}_{g}" for s in ("ft1", "ab2", "am3", "am4")
                                  for g in ("wing", "rotor")])
def test_other_schemes_match_reference_solver(name):
    """FT1 (main.f90:846-859), AB2 (951-1000), AM3 (1117-1248, whose
    start-up test iter == 0 is never true) and AM4 (1250-1404) on a wing and
    on a rotor in forward flight with dissipation and strain, each reaching
    a far wake with rollup and truncation: the CL/CT history against the
    reference solver's (7 printed significant digits)."""
    ref = read_force(os.path.join(GOLD, name + "_ForceNonDim.csv"))[:, 0]
    ct, _ = run(name)
    assert len(ct[0]) == len(ref)
    np.testing.assert_allclose(ct[0], ref, rtol=2e-6, atol=1e-9)


@pytest.mark.parametrize("name", ["initwake_am2_rotor", "initwake_am4_rotor"])
def test_initial_wake_velocity_matches_reference_solver(name):
    """initWakeVel for the first 30 of 48 steps (main.f90:829-838 in pass 1:
    added to the near wake, subtracted from the far wake; 1070-1081 and
    1353-1362 in pass 2: subtracted from both) on a forward-flight rotor
    whose far wake forms at step 12, under AM2 and AM4.  Without the
    start-up velocity the reference's history differs by 2%."""
    ref = read_force(os.path.join(GOLD, name + "_ForceNonDim.csv"))[:, 0]
    ct, _ = run(name)
    assert len(ct[0]) == len(ref)
    np.testing.assert_allclose(ct[0], ref, rtol=2e-6, atol=1e-9)


def read_tec(path):
    """Zones of a geom2file .tec file: [(title, I, J, X, Y, Z, GAM, skew)]."""
    import re
    zones, cur, vals = [], None, []

    def close():
        if cur is not None:
            t, I, J = cur
            nn, nc = I * J, (I - 1) * max(J - 1, 1)
            v = np.array(vals)
            assert len(v) == 3 * nn + 2 * nc, (t, len(v))
            zones.append((t, I, J, v[:nn], v[nn:2 * nn], v[2 * nn:3 * nn],
                          v[3 * nn:3 * nn + nc], v[3 * nn + nc:]))

    for line in open(path):
        w = line.strip()
        if w.startswith("Zone"):
            close()
            m = re.match(r'Zone I=\s*(\d+) J=\s*(\d+)\s+K=1\s+T="([^"]*)"', w)
            cur, vals = (m.group(3), int(m.group(1)), int(m.group(2))), []
        elif w and not w.startswith(("Title", "VARIABLES", "DATAPACKING", "VARLOCATION")):
            vals.extend(float(x) for x in w.split())
    close()
    return zones


@pytest.mark.parametrize("step", ["00020", "00040"])
def test_wake_plot_matches_reference_solver(step):
    """geom2file (libPostprocess.f90:143-331) at steps 20 and 40 of a
    forward-flight rotor whose far wake forms at step 12: every zone (blade
    panels, active near-wake lattice, far-wake tip line, the redundant
    prescribed-wake point) with the reference's sizes, node coordinates and
    -Γ node by node.  The near-wake skew column is left out: the reference
    fills it only for skew plots and otherwise writes uninitialised memory."""
    ref = read_tec(os.path.join(GOLD, f"tec_rotor_wingNwake{step}.tec"))
    run("tec_rotor")
    got = read_tec(os.path.join(CASES, "tec_rotor.case", "Results", f"r01wingNwake{step}.tec"))
    assert [z[:3] for z in got] == [z[:3] for z in ref]
    for zg, zr in zip(got, ref):
        for k in range(3, 7):
            np.testing.assert_allclose(zg[k], zr[k], rtol=1e-10, atol=1e-12,
                                       err_msg=f"{zg[0]} column {k - 2}")
        if zg[0] != "NearWake":
            np.testing.assert_array_equal(zg[7], zr[7])


@pytest.mark.parametrize("name", ["tip_rotor", "tip_rotor_cw"])
def test_wake_tip_plots_match_reference_solver(name):
    """tip2file (libPostprocess.f90:617-784) every 12 steps of a
    forward-flight rotor in slow start whose far wake forms at step 12, fills
    at step 30 and then shifts on every rollup.  Per blade: the wing mesh and
    -Γ, the Γ-weighted tip centroid of every active near-wake row with the
    row's extreme Γ, and the far-wake tip line, each with its azimuthal age
    in degrees (grown by dt*omegaSlow on active slots, carried through both
    shifts and never reset), compared zone by zone, and the tipXYZ point
    lists.  The CL/CT history matches as well.  tip_rotor_cw turns the other
    way: Γ changes sign, each row's highest Γ is taken, and the ages are
    negative."""
    ref_ct = read_force(os.path.join(GOLD, f"{name}_ForceNonDim.csv"))[:, 0]
    ct, _ = run(name)
    np.testing.assert_allclose(ct[0], ref_ct, rtol=2e-6, atol=1e-9)
    res = os.path.join(CASES, f"{name}.case", "Results")
    ages = []
    for step in ("00012", "00024", "00036", "00048"):
        ref = read_tec(os.path.join(GOLD, f"{name}_r01tip{step}.tec"))
        got = read_tec(os.path.join(res, f"r01tip{step}.tec"))
        assert [z[:3] for z in got] == [z[:3] for z in ref]
        for zg, zr in zip(got, ref):
            for k in range(3, 8):
                np.testing.assert_allclose(zg[k], zr[k], rtol=1e-10, atol=1e-12,
                                           err_msg=f"{step} {zg[0]} column {k - 2}")
            if zg[0] == "FarWake":
                ages.append(zr[7])

        def xyz(path):
            rows = [ln.split() for ln in open(path)]
            return sum(r[0] == "X" for r in rows), \
                np.array([[float(x) for x in r] for r in rows if r[0] != "X"])
        hg, pg = xyz(os.path.join(res, f"r01tipXYZ{step}.csv"))
        hr, pr = xyz(os.path.join(GOLD, f"{name}_r01tipXYZ{step}.csv"))
        assert hg == hr and pg.shape == pr.shape
        np.testing.assert_allclose(pg, pr, rtol=1e-10, atol=1e-12, err_msg=step)
    # the far-wake ages vary along the line at step 36 (blade 1's zone);
    # by step 48 the shifts have copied slot 1's age into every slot
    sgn = -1.0 if name.endswith("_cw") else 1.0
    assert len(ages[4]) == 18 and np.ptp(ages[4]) > 0
    assert np.ptp(ages[6]) == 0 and sgn * ages[6][0] > sgn * ages[4][-1] > 0


def test_grid_velocity_field_matches_reference_gridgen():
    """The Eulerian grid (SURVEY's gridgen row): the host writes
    filaments2file's dumps at steps 20 and 40 (libPostprocess.f90:363-473),
    byte for byte the size of the reference's dumps of the same steps (the
    same unformatted records), and volcanor_amd/bin/gridgen_gpu evaluates
    every filament at the 6x6x6 cell centres on the GPU, with the
    reference's wing-loop assignment (gridgen.f90:122: only the last wing
    ring counts) and the free stream added.  Compared with the grid files
    of the reference's own gridgen run on the reference's dumps."""
    import subprocess
    run("grid_rotor")
    case = os.path.join(CASES, "grid_rotor.case")
    res = os.path.join(case, "Results")
    ref_dump_bytes = {"00020": 93000, "00040": 96328}
    for st, n in ref_dump_bytes.items():
        assert os.path.getsize(os.path.join(res, f"filaments{st}.dat")) == n
    subprocess.run([os.path.join(ROOT, "volcanor_amd", "bin", "gridgen_gpu"), case], check=True,
                   timeout=120, stdout=subprocess.DEVNULL)

    def split(path):
        text, nums = [], []
        for line in open(path):
            w = line.strip()
            if w and w[0].isalpha():
                text.append(w)
            else:
                nums.extend(float(x) for x in w.split())
        return text, np.array(nums)

    for st in ref_dump_bytes:
        tr, vr = split(os.path.join(GOLD, f"grid_rotor_grid{st}.tec"))
        tg, vg = split(os.path.join(res, f"grid{st}.tec"))
        assert tg == tr and vg.shape == vr.shape == (3 * 343 + 3 * 216,)
        np.testing.assert_array_equal(vg[:3 * 343], vr[:3 * 343])      # grid nodes
        np.testing.assert_allclose(vg[3 * 343:], vr[3 * 343:], rtol=1e-10,
                                   atol=1e-12 * np.max(np.abs(vr[3 * 343:])))


def test_inflow_and_circulation_plots_match_reference_solver():
    """inflow2file (libPostprocess.f90:900-946: every rotor's wing less its
    bound vortices plus its wake, along -z at the section CPs; the span
    column is secCP . (hubCoords - yAxis) as the reference computes it) and
    gamma2file (949-978) on a rotor above a wing: the rotor plots every 12
    steps, the wing every 24.  At step 24 the reference calls gamma2file
    with rotor 1 for both switches (main.f90:777-783), so rotor 1's files
    hold two copies; the wing's own circulation is never written."""
    import glob
    res = os.path.join(CASES, "plots_rotor_wing.case", "Results")
    run("plots_rotor_wing", nrotors=2)
    refs = sorted(glob.glob(os.path.join(GOLD, "plots_rotor_wing_*.curve")))
    names = sorted(os.path.basename(f) for f in glob.glob(os.path.join(res, "*.curve")))
    assert names == [os.path.basename(f)[len("plots_rotor_wing_"):] for f in refs]
    for f in refs:
        mine = os.path.join(res, os.path.basename(f)[len("plots_rotor_wing_"):])
        if "inflowDist" in f:
            r, g = read_table(f), read_table(mine)
            assert list(g) == list(r) == ["secSpan", "inflowVel"]
            for k in r:
                np.testing.assert_allclose(g[k], r[k], rtol=2e-6, atol=2e-7, err_msg=f)
        else:
            lr = [l.split() for l in open(f)]
            lg = [l.split() for l in open(mine)]
            assert [x for x in lg if x[0] == "#"] == [x for x in lr if x[0] == "#"]
            vr = np.array([[float(v) for v in x] for x in lr if x[0] != "#"])
            vg = np.array([[float(v) for v in x] for x in lg if x[0] != "#"])
            assert vg.shape == vr.shape
            np.testing.assert_allclose(vg, vr, rtol=1e-10, atol=1e-12 * np.max(np.abs(vr)),
                                       err_msg=f)


@pytest.mark.parametrize("blade", ["b01", "b02"])
def test_surface_geometry_matches_reference_solver(blade):
    """geomSurface2file (libPostprocess.f90:475-615, written once after the
    initial pitch, main.f90:60-63): panel corners, vortex-ring corners, CPs
    with nCap, tauCapChord and tauCapSpan, and the sectional CPs with their
    normal and tangent vectors of both blades of the forward-flight rotor.
    The text lines (titles, zone sizes, sharing lists) must be identical and
    every number must agree."""
    run("tec_rotor")

    def split(path):
        text, nums = [], []
        for line in open(path):
            w = line.strip()
            if w and (w[0].isalpha() or w[0] == "_"):
                text.append(w)
            else:
                nums.extend(float(x) for x in w.split())
        return text, np.array(nums)

    tr, vr = split(os.path.join(GOLD, f"tec_rotor_{blade}Surface.tec"))
    tg, vg = split(os.path.join(CASES, "tec_rotor.case", "Results", f"r01{blade}Surface.tec"))
    assert tg == tr
    assert vg.shape == vr.shape
    np.testing.assert_allclose(vg, vr, rtol=1e-12, atol=1e-14)


def test_force_outputs_match_reference_solver():
    """force2file (libPostprocess.f90:814-882) at every force step of the
    forward-flight rotor: every column of rXXForceNonDim.csv and
    rXXForceDim.csv, the ForceDist tables of both blades at steps 0, 1 and
    40, and the header of rXXbladedynamics.csv (init_plots, 10-49), against
    the reference solver's files (7 printed significant digits; components
    that are round-off in the reference, like 1e-18, within 2e-6 of the
    column's scale)."""
    res = os.path.join(CASES, "tec_rotor.case", "Results")
    run("tec_rotor")

    def same(name, mine):
        ref, got = read_table(os.path.join(GOLD, "tec_rotor_" + name)), read_table(mine)
        assert list(got) == list(ref), name
        for k in ref:
            assert len(got[k]) == len(ref[k]), (name, k)
            scale = max(np.max(np.abs(ref[k])), 1e-300)
            np.testing.assert_allclose(got[k], ref[k], rtol=2e-6, atol=2e-6 * scale,
                                       err_msg=f"{name} {k}")

    same("ForceNonDim.csv", os.path.join(res, "r01ForceNonDim.csv"))
    same("ForceDim.csv", os.path.join(res, "r01ForceDim.csv"))
    for f in ("b01ForceDist00000.csv", "b02ForceDist00001.csv", "b01ForceDist00040.csv",
              "b02ForceDist00040.csv"):
        same(f, os.path.join(res, "r01" + f))
    assert (open(os.path.join(res, "r01bladedynamics.csv")).read()
            == open(os.path.join(GOLD, "tec_rotor_bladedynamics.csv")).read())


def test_custom_trajectory_matches_reference_solver():
    """customTrajectorySwitch = 1 (classdef.f90:3064-3080; main.f90:112-117
    and 419-426): the AR-4 wing's body velocity and rate come from
    trajectory01.in, one line per step (speed varying by 20%, a plunge
    velocity, a pitch rate about the c.g.).  Without the trajectory the
    reference's CL history differs by more than 100%."""
    ref = read_force(os.path.join(GOLD, "traj_wing_ForceNonDim.csv"))[:, 0]
    ct, _ = run("traj_wing")
    assert len(ct[0]) == len(ref)
    np.testing.assert_allclose(ct[0], ref, rtol=2e-6, atol=1e-9)


@pytest.mark.parametrize("f", ["r01b01skew00012.csv", "r01b02skew00012.csv",
                               "r01b01skew00024.csv", "r01b02skew00024.csv"])
def test_skew_plots_match_reference_solver(f):
    """calc_skew and skew2file (classdef.f90:704-747, libPostprocess.f90:
    786-812) at steps 12 and 24 of a forward-flight rotor: the maximum and
    the spanwise mean of the bimedian skew of every active near-wake row,
    from the device wake, against the reference solver's files."""
    run("skew_rotor")
    r = read_table(os.path.join(GOLD, "skew_rotor_" + f))
    g = read_table(os.path.join(CASES, "skew_rotor.case", "Results", f))
    assert list(g) == list(r) == ["max", "avg"]
    for k in r:
        np.testing.assert_allclose(g[k], r[k], rtol=2e-6, atol=1e-9, err_msg=k)


def resume(name, step, nrotors=1):
    """Run a copy of the case resumed from its restart file of `step`."""
    import shutil
    import tempfile
    from volcanor_amd import host
    case = os.path.join(CASES, name + ".case")
    tmp = tempfile.mkdtemp(prefix="vlmrestart_")
    try:
        for f in os.listdir(case):
            if os.path.isfile(os.path.join(case, f)):
                shutil.copy(os.path.join(case, f), tmp)
        os.makedirs(os.path.join(tmp, "Restart"))
        shutil.copy(os.path.join(case, "Restart", f"restart{step:05d}.dat"),
                    os.path.join(tmp, "Restart"))
        cfg = os.path.join(tmp, "config.nml")
        text = open(cfg).read()
        assert "restartFromNt = 0" in text
        open(cfg, "w").write(text.replace("restartFromNt = 0", f"restartFromNt = {step}"))
        ct, _ = host.run_case(tmp, nrotors=nrotors)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return ct


def test_restart_two_surfaces_resume_bitwise():
    """Restart of the rotor above the wing under AM2 with sub-iterations,
    both surfaces with a truncated far wake: resumed from step 18, both
    histories of steps 19-36 equal the uninterrupted run's bit for bit.  The
    uninterrupted run matches the reference solver's run of the case with
    the restart write switched off (its write aborts under AM2)."""
    ct, _ = run("restart_rotor_wing", nrotors=2)
    np.testing.assert_allclose(ct, reference_histories("restart_rotor_wing", 2), rtol=2e-6,
                               atol=1e-9)
    ct2 = resume("restart_rotor_wing", 18, nrotors=2)
    assert ct2.shape == ct.shape
    np.testing.assert_array_equal(ct2[:, 19:], ct[:, 19:])


def test_restart_resumes_bitwise():
    """restartWriteNt / restartFromNt (main.f90:390-398, 1442-1450): the run
    writes its state at steps 20 and 40; a copy of the case resumed from
    step 20 reproduces steps 21-40 of the uninterrupted run bit for bit
    (host frames, panels and circulation, and the device wake, cursors and
    velocity histories of vlm_rotor_get_state).  The history also matches
    the reference solver's uninterrupted run."""
    case = os.path.join(CASES, "restart_rotor.case")
    ref = read_force(os.path.join(GOLD, "restart_rotor_ForceNonDim.csv"))[:, 0]
    ct, _ = run("restart_rotor")
    np.testing.assert_allclose(ct[0], ref, rtol=2e-6, atol=1e-9)
    assert os.path.exists(os.path.join(case, "Restart", "restart00020.dat"))
    ct2 = resume("restart_rotor", 20)
    assert ct2.shape == ct.shape
    np.testing.assert_array_equal(ct2[0][21:], ct[0][21:])


def test_camber_line_matches_reference_solver():
    """A camber file (rotor_getCamber, classdef.f90:4060-4112, with the
    order-2 interp1d of libMath.f90:520-575): the AR-4 wing with a 4%
    parabolic camber line.  The cambered grid is checked node by node
    through the blade surface file, and the CL history against the
    reference solver's (the camber raises the final CL from 0.33 to 0.57)."""
    ref = read_force(os.path.join(GOLD, "camber_wing_ForceNonDim.csv"))[:, 0]
    ct, _ = run("camber_wing")
    assert len(ct[0]) == len(ref)
    np.testing.assert_allclose(ct[0], ref, rtol=2e-6, atol=1e-9)

    def nums(path):
        return np.array([float(x) for l in open(path) if l.strip() and not l.strip()[0].isalpha()
                         for x in l.split()])

    np.testing.assert_allclose(
        nums(os.path.join(CASES, "camber_wing.case", "Results", "r01b01Surface.tec")),
        nums(os.path.join(GOLD, "camber_wing_b01Surface.tec")), rtol=1e-12, atol=1e-14)


def test_wake_burst_matches_reference_solver():
    """wakeBurst = 1 with skewLimit 0.02 (main.f90:490-496, classdef.f90:
    2306-2339): every step, far filaments whose angle with a neighbour
    exceeds the limit take the chord as core radius; strain is off so the
    cores persist.  Without the burst the reference's history differs by
    4.6% from step 14 on."""
    ref = read_force(os.path.join(GOLD, "burst_rotor_ForceNonDim.csv"))[:, 0]
    ct, _ = run("burst_rotor")
    assert len(ct[0]) == len(ref)
    np.testing.assert_allclose(ct[0], ref, rtol=2e-6, atol=1e-9)


@pytest.mark.parametrize("k", [1, 2, 3])
def test_slow_start_matches_reference_solver(k):
    """slowStart 1-3 (main.f90:97-105, 428-452): the initial solution at
    rest, then the rotor speed ramps over 20 steps, linearly, as
    tanh(5 n/N) or as a centred tanh; hover rotor, 48 steps."""
    ref = read_force(os.path.join(GOLD, f"slowstart{k}_rotor_ForceNonDim.csv"))[:, 0]
    ct, _ = run(f"slowstart{k}_rotor")
    assert len(ct[0]) == len(ref)
    np.testing.assert_allclose(ct[0], ref, rtol=2e-6, atol=1e-9)


def test_sub_iterations_match_reference_solver():
    """ntSubInit = 5 and ntSub = 3 (main.f90:120-211, 522-615) on the rotor
    above a wing: the Jacobi coupling of the two surfaces is iterated to the
    limit every step (the reference's residual never reaches eps here).
    Without the sub-iterations the wing's history differs by 56%."""
    ref = reference_histories("subiter_rotor_wing", 2)
    ct, _ = run("subiter_rotor_wing", nrotors=2)
    assert ct.shape == ref.shape
    np.testing.assert_allclose(ct, ref, rtol=2e-6, atol=2e-9)


@pytest.mark.parametrize("name,nr", [("cfg4_quad_short", 4), ("cfg5_rotor_wing_short", 2)])
def test_multi_rotor_matches_reference_solver(name, nr):
    """Several lifting surfaces coupled through the right-hand side (every
    wake and the other surfaces' wings, main.f90:521-603) and through the
    wake passes (vind_onNwake_byRotor over all rotors): the quad-rotor over
    one revolution (far wake, rollup and truncation reached) and the rotor in
    forward flight above a wing, every CT/CL of every surface against the
    reference solver's printed history (7 significant digits)."""
    ref = reference_histories(name, nr)
    ct, _ = run(name, nrotors=nr)
    assert ct.shape == ref.shape
    np.testing.assert_allclose(ct, ref, rtol=2e-6, atol=2e-9)


def test_error_after_device_setup_returns_and_next_run_works():
    """A restart file that does not exist is found missing after the device
    context and the rotors are set up: volcanor_run returns an error with
    the message, destroys that context, and the next run in the same
    process gives the reference history."""
    import shutil
    import tempfile
    from volcanor_amd import host, vlm
    d = tempfile.mkdtemp(prefix="vlmerr_")
    try:
        src = os.path.join(CASES, "cfg1_wing.case")
        for f in os.listdir(src):
            if os.path.isfile(os.path.join(src, f)):
                shutil.copy(os.path.join(src, f), d)
        cfg = os.path.join(d, "config.nml")
        text = open(cfg).read()
        assert "restartFromNt = 0" in text
        with open(cfg, "w") as fh:
            fh.write(text.replace("restartFromNt = 0", "restartFromNt = 7"))
        with pytest.raises(vlm.VLMError, match="restart00007.dat"):
            host.run_case(d, nt=10)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    ref = read_force(os.path.join(GOLD, "cfg1_wing_ForceNonDim.csv"))[:, 0]
    ct, _ = run("cfg1_wing", nt=20)
    np.testing.assert_allclose(ct[0], ref[:21], rtol=1e-6, atol=1e-7)


def test_cfg5_as_specified_one_revolution():
    """cfg5 as specified: the three-bladed rotor (30 spanwise panels) in
    forward flight 0.5 R above the 4 x 40 wing, coupled through the
    right-hand side and every wake pass, for one revolution (72 steps).
    The reference solver turns NaN on surfaces with unequal spanwise counts,
    so this run is held to properties: every CT/CL finite, the rotor's CT
    positive and below 0.05, the wing's CL within (-2, 2) once the
    impulsive start has passed (from step 5; the reference's equal-count
    stand-in, cfg5_rotor_wing_short, prints 3.18 at step 0 and -0.72 to
    -0.24 from step 5), and the MIXED path within 1e-3 of FP64 relative to
    the largest value (measured 6.9e-5 on the rotor: the wing sits in the
    rotor's wake, the most sensitive of the cases).  The wake-level parity
    of unequal spanwise counts against the oracle is
    test_wake_gpu.test_surfaces_with_different_spanwise_counts_match_oracle."""
    ct, st = run("cfg5_rotor_wing", nt=72, nrotors=2)
    assert ct.shape == (2, 73) and np.all(np.isfinite(ct))
    assert np.all(ct[0, 1:] > 0.0) and np.all(ct[0, 1:] < 0.05)
    assert np.all(np.abs(ct[1, 5:]) < 2.0)
    cm, _ = run("cfg5_rotor_wing", precision="mixed", nt=72, nrotors=2)
    scale = np.max(np.abs(ct), axis=1, keepdims=True)
    assert np.max(np.abs(cm - ct) / scale) < 1e-3


def test_restart_carries_tip_plot_ages():
    """A restart file carries the tip plot's azimuthal ages: tip_rotor (AM2,
    slow start, truncated far wake) run with restarts every 24 steps, and a
    copy resumed from step 24, write byte-identical tip files at steps 36
    and 48 and the same CL/CT history from step 25."""
    import shutil
    import tempfile
    from volcanor_amd import host
    src = os.path.join(CASES, "tip_rotor.case")
    dirs = [tempfile.mkdtemp(prefix="vlmtip_") for _ in range(2)]
    try:
        for d in dirs:
            for f in os.listdir(src):
                if os.path.isfile(os.path.join(src, f)):
                    shutil.copy(os.path.join(src, f), d)
            cfg = os.path.join(d, "config.nml")
            text = open(cfg).read()
            assert "restartWriteNt = 0" in text
            open(cfg, "w").write(text.replace("restartWriteNt = 0", "restartWriteNt = 24"))
        full, _ = host.run_case(dirs[0])
        os.makedirs(os.path.join(dirs[1], "Restart"))
        shutil.copy(os.path.join(dirs[0], "Restart", "restart00024.dat"),
                    os.path.join(dirs[1], "Restart"))
        cfg = os.path.join(dirs[1], "config.nml")
        text = open(cfg).read()
        with open(cfg, "w") as fh:
            fh.write(text.replace("restartFromNt = 0", "restartFromNt = 24"))
        resumed, _ = host.run_case(dirs[1])
        np.testing.assert_array_equal(resumed[:, 25:], full[:, 25:])
        for step in ("00036", "00048"):
            for f in (f"r01tip{step}.tec", f"r01tipXYZ{step}.csv"):
                a = open(os.path.join(dirs[0], "Results", f)).read()
                b = open(os.path.join(dirs[1], "Results", f)).read()
                assert a == b, f
    finally:
        for d in dirs:
            shutil.rmtree(d, ignore_errors=True)


@pytest.mark.parametrize("name", ["presc_rotor", "presc_gen_rotor"])
def test_prescribed_far_wake_matches_reference_solver(name):
    """The prescribed far wake (pFwake_update, classdef.f90:998-1066,
    rotor_updatePrescribedWake 5172-5217) on the device: a climbing rotor
    whose 240-filament helix is rebuilt after both AM2 convections from
    step 31 and joins every wake source set, from every far filament
    (presc_rotor) or, with three axisymmetric blades, from the last six
    (presc_gen_rotor).  The CL/CT history against the reference solver's
    and every zone of the wing and wake plots at steps 24 and 48, the helix
    (PrescFarWake) included, to 1e-8 relative or 1e-10 absolute (the bound
    of the CPU build of the same case: the helix switches on with a jump in
    CT, after which the wake amplifies round-off to about 4e-9 by step 48)."""
    ref = read_force(os.path.join(GOLD, f"{name}_ForceNonDim.csv"))[:, 0]
    ct, _ = run(name)
    np.testing.assert_allclose(ct[0], ref, rtol=2e-6, atol=1e-9)
    res = os.path.join(CASES, f"{name}.case", "Results")
    for step in ("00024", "00048"):
        got = read_tec(os.path.join(res, f"r01wingNwake{step}.tec"))
        exp = read_tec(os.path.join(GOLD, f"{name}_wingNwake{step}.tec"))
        assert [z[:3] for z in got] == [z[:3] for z in exp]
        for zg, ze in zip(got, exp):
            for k in range(3, 8):
                np.testing.assert_allclose(zg[k], ze[k], rtol=1e-8, atol=1e-10,
                                           err_msg=f"{step} {zg[0]} column {k - 2}")
    assert any(z[0] == "PrescFarWake" and z[1] == 241 for z in got)


IMAGE_CASES = [("image_rotor", 2, False), ("image_rotor_am4", 2, False),
               ("image_rotor_xy", 3, True)]


@pytest.mark.parametrize("name,nr,tips", IMAGE_CASES)
def test_image_rotors_match_reference_solver(name, nr, tips):
    """Image rotors (surfaceType -1: rotor_init classdef.f90:2798-2960,
    rotor_mirrorVelCP / rotor_mirrorGamma / rotor_mirrorWake 4990-5131,
    main.f90:181-193, 595-599, 844-1437) on the device: image_rotor, a
    two-bladed axisymmetric rotor in ground effect (0.6 R above z = 0,
    descending, AM2, far wake truncated at 24 steps) and its image in z = 0,
    and the same pair under AM4 (image_rotor_am4); image_rotor_xy, a three-bladed rotor with cyclic pitch and its images in
    x = 0 (rotation reversed) and y = 0 (rotation reversed, start azimuth +
    pi), AB2.  Every column of every rotor's force history against the
    reference solver's, and every zone of every rotor's wing and wake plots
    (and tip plots) at steps 24 and 48 to 1e-8 relative or 1e-10 absolute."""
    ct, _ = run(name, nrotors=nr)
    res = os.path.join(CASES, f"{name}.case", "Results")
    for k in range(1, nr + 1):
        ref = read_force(os.path.join(GOLD, f"{name}_r{k:02d}ForceNonDim.csv"))
        got = read_force(os.path.join(res, f"r{k:02d}ForceNonDim.csv"))
        np.testing.assert_allclose(ct[k - 1], ref[:, 0], rtol=2e-6, atol=1e-9, err_msg=f"r{k:02d}")
        np.testing.assert_allclose(got, ref, rtol=2e-6, atol=1e-9, err_msg=f"r{k:02d} columns")
        plots = ["wingNwake"] + (["tip"] if tips else [])
        for step in ("00024", "00048"):
            for p in plots:
                g = read_tec(os.path.join(res, f"r{k:02d}{p}{step}.tec"))
                e = read_tec(os.path.join(GOLD, f"{name}_r{k:02d}{p}{step}.tec"))
                assert [z[:3] for z in g] == [z[:3] for z in e]
                for zg, ze in zip(g, e):
                    for c in range(3, 8):
                        np.testing.assert_allclose(zg[c], ze[c], rtol=1e-8, atol=1e-10,
                                                   err_msg=f"r{k:02d} {p}{step} {zg[0]} column {c - 2}")
