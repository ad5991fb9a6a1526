#!/usr/bin/env python3
"""Writes the `.case` directories (config.nml v0.5 + geomXX.nml v0.15 and any
PLOT3D blade file) used by the tests and the benchmark.

  cfg1_wing          flat AR-4 wing, 4 x 13, 5 deg, 200 steps, FP64
  cfg2_rotor         two-bladed hover rotor, 1 x 20, -8 deg linear twist,
                     8 deg collective, tip Mach 0.4, 5 deg steps, 10 revs
  cfg4_quad          four three-bladed rotors, 1 x 30, 10 deg steps, 20 revs,
                     15-rev truncated lattice
  cfg5_rotor_wing    three-bladed rotor in forward flight above an AR-8 wing
  refprobe_rotor     2 x (2 x 3) rotor in forward flight, 14 steps: the case on
                     which oracle/refprobe.f90 dumps the reference's wake state
  probes_wing        a short wing run with six velocity probes (probes.in)
  scheme_*_wing/rotor  FT1, AB2, AM3 and AM4 on a wing and a forward-flight
                     rotor
  cfg2_rotor_farwake_2rev  cfg2 with a 72-row near wake and a far wake (strain)
  initwake_*_rotor   the start-up wake velocity initWakeVel under AM2 and AM4
  tec_rotor          wing and wake plots (.tec) of a forward-flight rotor
  plots_rotor_wing   inflow and bound-circulation plots of a rotor above a wing
  grid_rotor         filament dumps and gridconfig.nml for the grid velocity field
  burst_rotor        far-wake burst every step with a small skew limit
  traj_wing          the AR-4 wing on a custom trajectory (trajectory01.in)
  camber_wing        the AR-4 wing with a camber line
  skew_rotor         near-wake skew plots of a forward-flight rotor
  presc_rotor, presc_gen_rotor  the prescribed far wake (helix) of a climbing
                     rotor, from every far filament and, with axisymmetric
                     blades, from the last six
  image_rotor(_am4), image_rotor_xy  image rotors: an axisymmetric rotor in
                     ground effect and its image in z = 0 (AM2, AM4); a rotor
                     with cyclic pitch and its images in x = 0 and y = 0
  restart_rotor      restart files of a forward-flight rotor
  restart_rotor_wing restart files of the rotor above the wing (AM2, sub-iterations)
  tip_rotor(_cw)     wake tip plots of a forward-flight rotor in slow start, turning
                     anticlockwise and clockwise
  slowstart*_rotor   the three slow-start ramps of the rotor speed
  subiter_rotor_wing cfg5_rotor_wing_short with sub-iterations (ntSubInit 5, ntSub 3)
  katz_plotkin_AR4   the reference's AR-4 wing test case (4 x 26, 160 steps)
  elevate            the reference's 5-bladed axisymmetric test rotor
  (the last two reproduce tests/katzNplotkin-AR04.case and
  tests/elevateTest.case of cibinjoseph/VOLCANOR; their blade grid and
  fixture CSVs are under tests/golden/)

Usage: python cases/make_cases.py [outdir]   (default: this directory)
"""
import math
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))

# Vortex core of the rotor cases: 0.2 chord with Oseen-type growth at
# apparentViscCoeff = 1000.  With 0.04 chord and coefficient 1 the hovering
# wake of cfg2 blows up within four revolutions (CT -0.49) and FP32 and FP64
# runs separate after one; with these values they agree to 3e-6 over four.
ROTOR_CORE = 0.2
ROTOR_AVC = 1000.0

CONFIG_KEYS = ["restartWriteNt", "restartFromNt", "ntSub", "ntSubInit", "wakePlot",
               "wakeTipPlot", "rotorForcePlot", "gridPlot", "wakeDissipation", "wakeStrain",
               "wakeBurst", "wakeSuppress", "slowStart", "slowStartNt", "fdScheme",
               "initWakeVelNt", "probe"]


def fmt(v):
    if isinstance(v, str):
        return v
    if isinstance(v, (list, tuple)):
        return ", ".join(fmt(x) for x in v)
    if isinstance(v, float):
        return repr(v)
    return str(v)


def write_config(path, nt, dt, nr, density=1.0, velSound=330.0, kinematicVisc=1.8e-5, **opts):
    o = {k: 0 for k in CONFIG_KEYS}
    o["rotorForcePlot"] = 1
    o["fdScheme"] = 3
    o.update(opts)
    with open(path, "w") as f:
        f.write("&VERSION\nfileFormatVersion = 0.5\n/\n")
        f.write(f"&PARAMS\nnt = {nt}\ndt = {fmt(dt)}\nnr = {nr}\ndensity = {fmt(density)}\n"
                f"velSound = {fmt(velSound)}\nkinematicVisc = {fmt(kinematicVisc)}\n/\n")
        f.write("&OPTIONS\n" + "".join(f"{k} = {fmt(o[k])}\n" for k in CONFIG_KEYS) + "/\n")


def write_geom(path, **g):
    d = dict(surfaceType=0, imagePlane=3, imageRotorNum=1, nb=1, propConvention=0,
             spanSpacing=1, chordSpacing=1, geometryFile="0", nCamberFiles=0, nc=4, ns=13,
             nNwake=200, hubCoords=[0.0, 0.0, 0.0], cgCoords=[0.0, 0.0, 0.0],
             fromCoords=[0.0, 0.0, 0.0], phiThetaPsi=[0.0, 0.0, 0.0], span=4.0, rootcut=0.0,
             chord=1.0, preconeAngle=0.0, Omega=0.0, shaftAxis=[0.0, 0.0, 1.0], theta0=5.0,
             thetaC=0.0, thetaS=0.0, thetaTwist=0.0, ductSwitch=0, axisymmetrySwitch=0,
             pivotLE=0.0, flapHinge=0.0, spanwiseLiftSwitch=0, symmetricTau=1,
             customTrajectorySwitch=0, velBody=[-10.0, 0.0, 0.0], omegaBody=[0.0, 0.0, 0.0],
             forceCalcSwitch=0, nAirfoils=0, apparentViscCoeff=5.0, decayCoeff=0.0,
             wakeTruncateNt=0, prescWakeAfterTruncNt=0, prescWakeGenNt=0, spanwiseCore=0.03,
             streamwiseCoreVec=0.03, rollupStartRadius=0.75, rollupEndRadius=1.0,
             initWakeVel=0.0, psiStart=0.0, skewLimit=0.5, inflowPlotSwitch=0,
             gammaPlotSwitch=0, skewPlotSwitch=0, camberSectionLimit=[1.0], camberFile=["0"])
    d.update(g)
    groups = [
        ("SURFACE", ["surfaceType", "imagePlane", "imageRotorNum"]),
        ("PANELS", ["nb", "propConvention", "spanSpacing", "chordSpacing", "geometryFile",
                    "nCamberFiles", "nc", "ns", "nNwake"]),
        ("CAMBERSECTIONS", []),
        ("ORIENT", ["hubCoords", "cgCoords", "fromCoords", "phiThetaPsi"]),
        ("GEOMPARAMS", ["span", "rootcut", "chord", "preconeAngle", "Omega", "shaftAxis",
                        "theta0", "thetaC", "thetaS", "thetaTwist", "ductSwitch",
                        "axisymmetrySwitch", "pivotLE", "flapHinge", "spanwiseLiftSwitch",
                        "symmetricTau", "customTrajectorySwitch", "velBody", "omegaBody",
                        "forceCalcSwitch", "nAirfoils"]),
        ("WAKEPARAMS", ["apparentViscCoeff", "decayCoeff", "wakeTruncateNt",
                        "prescWakeAfterTruncNt", "prescWakeGenNt", "spanwiseCore",
                        "streamwiseCoreVec", "rollupStartRadius", "rollupEndRadius",
                        "initWakeVel", "psiStart", "skewLimit"]),
    ]
    with open(path, "w") as f:
        f.write("&VERSION\nfileFormatVersion = 0.15\n/\n")
        for name, keys in groups:
            f.write(f"&{name}\n")
            if name == "CAMBERSECTIONS":
                f.write(f"camberSectionLimit = {fmt(d['camberSectionLimit'])}\n"
                        f"camberFile = {fmt(d['camberFile'])}\n")
            for k in keys:
                f.write(f"{k} = {fmt(d[k])}\n")
            f.write("/\n")
        f.write("&DYNAMICS\nbladeDynamicsSwitch = 0\nflapInitial = 0.0\ndflapInitial = 0.0\n"
                "Iflap = 0.0\ncflap = 0.0\nkflap = 0.0\nMflapConstant = 0.0\n"
                "pitchDynamicsSwitch = 0\ndpitch = 0.0\nbodyDynamicsSwitch = 0\n"
                "bodyDynamicsIOVars = 1\n/\n")
        f.write("&WINDFRAME\ndragUnitVec = 0.0, 0.0, 0.0\nsideUnitVec = 0.0, 0.0, 0.0\n"
                "liftUnitVec = 0.0, 0.0, 0.0\n/\n")
        f.write(f"&PLOTS\ninflowPlotSwitch = {d['inflowPlotSwitch']}\n"
                f"gammaPlotSwitch = {d['gammaPlotSwitch']}\n"
                f"skewPlotSwitch = {d['skewPlotSwitch']}\n/\n")
        if d["nAirfoils"] > 0:
            # section tables named '0' are not read (classdef.f90:3697-3699);
            # with forceCalcSwitch = 0 alpha0 never enters the loads
            n = d["nAirfoils"]
            none = ["'0'"] * n
            f.write(f"&AIRFOILS\nairfoilSectionLimit = {fmt([1.0] * n)}\n"
                    f"alpha0 = {fmt([0.0] * n)}\nairfoilFile = {fmt(none)}\n/\n")


def twisted_blade_plot3d(path, nc, ns, R, chord, twist_deg, omega_sign=1.0, r_zero=0.75,
                         pivot=0.25):
    """Flat blade grid with linear twist (zero at r_zero*R) applied by
    rotation about the pivot line at `pivot` chord from the leading edge."""
    xs = [(-chord if omega_sign >= 0 else chord) * (1 - i / nc) for i in range(nc + 1)]
    ys = [R * j / ns for j in range(ns + 1)]
    xle = xs[0]
    xp = xle + pivot * (0.0 - xle)
    X, Y, Z = [], [], []
    for j in range(ns + 1):
        th = math.radians(twist_deg * (ys[j] / R - r_zero))
        for i in range(nc + 1):
            dx = xs[i] - xp
            X.append(xp + dx * math.cos(th))
            Y.append(ys[j])
            Z.append(-dx * math.sin(th) * (1 if omega_sign >= 0 else -1))
    with open(path, "w") as f:
        f.write(f"{nc + 1} {ns + 1} 1\n")
        for arr in (X, Y, Z):
            f.write(" ".join(f"{v:.16e}" for v in arr) + "\n")


def make_all(out):
    def case(name):
        d = os.path.join(out, name + ".case")
        os.makedirs(d, exist_ok=True)
        return d

    # cfg1: flat wing
    d = case("cfg1_wing")
    write_config(os.path.join(d, "config.nml"), nt=200, dt=0.025, nr=1)
    write_geom(os.path.join(d, "geom01.nml"), nc=4, ns=13, nNwake=200, spanwiseCore=0.03,
               streamwiseCoreVec=0.03)

    # velocity probes on a short wing run: three fixed and three moving probes
    d = case("probes_wing")
    write_config(os.path.join(d, "config.nml"), nt=20, dt=0.025, nr=1, probe=1)
    write_geom(os.path.join(d, "geom01.nml"), nc=4, ns=13, nNwake=20, spanwiseCore=0.03,
               streamwiseCoreVec=0.03)
    with open(os.path.join(d, "probes.in"), "w") as f:
        f.write("6\n")
        for p in ((0.5, 0.3, 0.2, 0.0, 0.0, 0.0), (2.0, 1.0, -0.1, 0.0, 0.0, 0.0),
                  (1.0, 2.1, 0.0, 0.0, 0.0, 0.0), (0.2, 0.5, 0.15, 2.0, 0.0, 0.0),
                  (-0.5, 1.5, -0.3, 5.0, 0.5, 0.1), (3.0, 0.0, 0.5, 1.0, 0.0, -0.2)):
            f.write(" ".join(repr(x) for x in p) + "\n")

    # reference fixture case: Katz & Plotkin AR-4 wing
    d = case("katz_plotkin_AR4")
    write_config(os.path.join(d, "config.nml"), nt=160, dt=0.00625, nr=1, wakePlot=32)
    write_geom(os.path.join(d, "geom01.nml"), nc=4, ns=26, spanSpacing=2, nNwake=160,
               spanwiseCore=0.04, streamwiseCoreVec=0.04)

    # reference fixture case: 5-bladed axisymmetric rotor with a PLOT3D blade
    d = case("elevate")
    write_config(os.path.join(d, "config.nml"), nt=150, dt=-0.04, nr=1,
                 kinematicVisc=0.0000181, wakeDissipation=1, wakePlot=-1)
    write_geom(os.path.join(d, "geom01.nml"), nb=5, nc=16, ns=16, spanSpacing=2,
               geometryFile="elevate.xyz", nNwake=30, span=5.0, Omega=89.98, theta0=0.0,
               axisymmetrySwitch=1, velBody=[0.0, 0.0, 10.0], apparentViscCoeff=5000.0,
               wakeTruncateNt=-3, spanwiseCore=0.04, streamwiseCoreVec=0.04)
    src = os.path.join(HERE, "..", "tests", "golden", "elevate.xyz")
    if os.path.exists(src):
        shutil.copy(src, os.path.join(d, "elevate.xyz"))

    # cfg2: two-bladed rotor in hover, 10 revolutions
    R, c, om = 1.0, 1.0 / 6.0, 0.4 * 340.0
    for name, revs in (("cfg2_rotor", 10), ("cfg2_rotor_2rev", 2)):
        d = case(name)
        write_config(os.path.join(d, "config.nml"), nt=-revs, dt=0.0, nr=1, velSound=340.0,
                     wakeDissipation=1, wakeStrain=1)
        twisted_blade_plot3d(os.path.join(d, "blade.xyz"), 1, 20, R, c, -8.0)
        write_geom(os.path.join(d, "geom01.nml"), nb=2, nc=1, ns=20, nNwake=-revs, span=R,
                   chord=c, Omega=om, theta0=8.0, pivotLE=0.25, geometryFile="blade.xyz",
                   velBody=[0.0, 0.0, 0.0], apparentViscCoeff=ROTOR_AVC,
                   spanwiseCore=ROTOR_CORE, streamwiseCoreVec=ROTOR_CORE)
    # cfg2 with a one-revolution near wake rolled up into a far wake (the
    # variant that exercises strain, which acts on the far wake only); two
    # revolutions against the reference solver
    d = case("cfg2_rotor_farwake_2rev")
    write_config(os.path.join(d, "config.nml"), nt=-2, dt=0.0, nr=1, velSound=340.0,
                 wakeDissipation=1, wakeStrain=1)
    twisted_blade_plot3d(os.path.join(d, "blade.xyz"), 1, 20, R, c, -8.0)
    write_geom(os.path.join(d, "geom01.nml"), nb=2, nc=1, ns=20, nNwake=72, wakeTruncateNt=-2,
               span=R, chord=c, Omega=om, theta0=8.0, pivotLE=0.25, geometryFile="blade.xyz",
               velBody=[0.0, 0.0, 0.0], apparentViscCoeff=ROTOR_AVC,
               spanwiseCore=ROTOR_CORE, streamwiseCoreVec=ROTOR_CORE)

    # small forward-flight rotor for the per-node wake comparison with the
    # reference (oracle/refprobe.f90): far wake, rollup, truncation shift,
    # dissipation and strain within 14 steps
    d = case("refprobe_rotor")
    write_config(os.path.join(d, "config.nml"), nt=14, dt=0.004, nr=1, wakeDissipation=1,
                 wakeStrain=1, rotorForcePlot=0)
    write_geom(os.path.join(d, "geom01.nml"), nb=2, nc=2, ns=3, nNwake=5, span=1.0,
               rootcut=0.2, chord=0.2, Omega=60.0, theta0=8.0, pivotLE=0.25,
               velBody=[-2.0, 0.0, 0.0], wakeTruncateNt=9, apparentViscCoeff=50.0,
               spanwiseCore=0.2, streamwiseCoreVec=0.2, rollupStartRadius=0.6,
               rollupEndRadius=1.0)

    # cfg4: quad-rotor, 20 revolutions, 15-rev truncated near wake; the
    # _short variant (one revolution, 12-step truncation) is compared with
    # the reference solver step by step
    dt4 = math.radians(10.0) / om
    hubs = [(1.25, 1.25), (-1.25, 1.25), (-1.25, -1.25), (1.25, -1.25)]
    for name, nt4, nN4, tr4 in (("cfg4_quad", 720, 540, 540), ("cfg4_quad_short", 36, 12, 24)):
        d = case(name)
        write_config(os.path.join(d, "config.nml"), nt=nt4, dt=dt4, nr=4, velSound=340.0)
        for k, (hx, hy) in enumerate(hubs):
            sgn = 1.0 if k % 2 == 0 else -1.0
            write_geom(os.path.join(d, f"geom{k + 1:02d}.nml"), nb=3, nc=1, ns=30, nNwake=nN4,
                       span=R, chord=c, Omega=sgn * om, theta0=8.0, pivotLE=0.25,
                       hubCoords=[hx, hy, 0.0], velBody=[0.0, 0.0, 0.0], wakeTruncateNt=tr4,
                       apparentViscCoeff=ROTOR_AVC, spanwiseCore=ROTOR_CORE,
                       streamwiseCoreVec=ROTOR_CORE)

    # the other wake schemes (FT1, AB2, AM3, AM4) against the reference solver: a wing
    # and a two-bladed rotor in forward flight, both reaching a far wake with
    # rollup and truncation within the run.  PC2 is not among them: the
    # reference's vel_order2_Fwake (libCommon.f90:236-258) writes elements 1
    # and 0 of its zero-length result while the far-wake slice is empty, which
    # corrupts the heap (free(): invalid pointer at step 2).
    for sch, tag in ((0, "ft1"), (2, "ab2"), (4, "am3"), (5, "am4")):
        d = case(f"scheme_{tag}_wing")
        write_config(os.path.join(d, "config.nml"), nt=60, dt=0.025, nr=1, fdScheme=sch)
        write_geom(os.path.join(d, "geom01.nml"), nc=2, ns=8, nNwake=20, wakeTruncateNt=45,
                   spanwiseCore=0.03, streamwiseCoreVec=0.03)
        d = case(f"scheme_{tag}_rotor")
        write_config(os.path.join(d, "config.nml"), nt=72, dt=math.radians(5.0) / om, nr=1,
                     velSound=340.0, fdScheme=sch, wakeDissipation=1, wakeStrain=1)
        write_geom(os.path.join(d, "geom01.nml"), nb=2, nc=1, ns=10, nNwake=24,
                   wakeTruncateNt=60, span=R, chord=c, Omega=om, theta0=8.0, pivotLE=0.25,
                   velBody=[-13.6, 0.0, -2.0], apparentViscCoeff=ROTOR_AVC,
                   spanwiseCore=ROTOR_CORE, streamwiseCoreVec=ROTOR_CORE)

    # initWakeVel: for the first 30 of 48 steps every pass adds the
    # reference's start-up wake velocity (main.f90:829-838, pass-2 blocks),
    # with a 12-row near wake so that far-wake rows exist while it applies;
    # under AM2 and under AM4
    for sch, tag in ((3, "am2"), (5, "am4")):
        d = case(f"initwake_{tag}_rotor")
        write_config(os.path.join(d, "config.nml"), nt=48, dt=math.radians(5.0) / om, nr=1,
                     velSound=340.0, fdScheme=sch, wakeDissipation=1, wakeStrain=1,
                     initWakeVelNt=30)
        write_geom(os.path.join(d, "geom01.nml"), nb=2, nc=1, ns=10, nNwake=12,
                   wakeTruncateNt=40, span=R, chord=c, Omega=om, theta0=8.0, pivotLE=0.25,
                   velBody=[-13.6, 0.0, -2.0], apparentViscCoeff=ROTOR_AVC,
                   spanwiseCore=ROTOR_CORE, streamwiseCoreVec=ROTOR_CORE, initWakeVel=2.0)

    # wing and wake plots (.tec, geom2file) at steps 20 and 40 of a
    # forward-flight rotor whose far wake forms at step 12
    d = case("tec_rotor")
    write_config(os.path.join(d, "config.nml"), nt=40, dt=math.radians(5.0) / om, nr=1,
                 velSound=340.0, wakeDissipation=1, wakeStrain=1, wakePlot=20)
    write_geom(os.path.join(d, "geom01.nml"), nb=2, nc=2, ns=8, nNwake=12, wakeTruncateNt=36,
               span=R, chord=c, Omega=om, theta0=8.0, pivotLE=0.25, velBody=[-13.6, 0.0, -2.0],
               apparentViscCoeff=ROTOR_AVC, spanwiseCore=ROTOR_CORE,
               streamwiseCoreVec=ROTOR_CORE)

    # prescribed far wake (pFwake_update, classdef.f90:998-1066, 5172-5217):
    # far wake from step 12, truncated at 24, the helix rebuilt after every
    # convection from step 31 (prescWakeAfterTruncNt 6), wake plots at
    # steps 24 and 48.  presc_rotor: two blades convected each, the helix
    # from every active far filament; presc_gen_rotor: three axisymmetric
    # blades, the helix from the last 6 far filaments (prescWakeGenNt 5)
    for name, nb_p, axi, gen in (("presc_rotor", 2, 0, 0), ("presc_gen_rotor", 3, 1, 5)):
        d = case(name)
        write_config(os.path.join(d, "config.nml"), nt=48, dt=math.radians(5.0) / om, nr=1,
                     velSound=340.0, wakeDissipation=1, wakeStrain=1, wakePlot=24)
        write_geom(os.path.join(d, "geom01.nml"), nb=nb_p, nc=2, ns=8, nNwake=12,
                   wakeTruncateNt=24, prescWakeAfterTruncNt=6, prescWakeGenNt=gen,
                   axisymmetrySwitch=axi, span=R, chord=c, Omega=om, theta0=8.0, pivotLE=0.25,
                   velBody=[0.0, 0.0, -2.0], apparentViscCoeff=ROTOR_AVC,
                   spanwiseCore=ROTOR_CORE, streamwiseCoreVec=ROTOR_CORE)

    # image rotors (surfaceType -1; rotor_init classdef.f90:2798-2960,
    # rotor_mirrorGamma / rotor_mirrorWake 4990-5131): image_rotor, a
    # two-bladed axisymmetric rotor 0.6 R above the ground plane z = 0,
    # descending, and its image (AM2, far wake from step 12 truncated at 24,
    # wake plots at steps 24 and 48); image_rotor_xy, a three-bladed rotor
    # with cyclic pitch and its images in the planes x = 0 and y = 0 (AB2,
    # wake and tip plots at steps 24 and 48; PC2 on this rotor stops the
    # reference solver with "free(): invalid pointer", images or not).  The
    # image files hold only the entries an image reads (surface and panel
    # counts).
    # image_rotor_am4: the same pair under AM4 (start-up steps, predictor
    # and corrector mirrors)
    for name, scheme in (("image_rotor", 3), ("image_rotor_am4", 5)):
        d = case(name)
        write_config(os.path.join(d, "config.nml"), nt=48, dt=math.radians(5.0) / om, nr=2,
                     velSound=340.0, wakeDissipation=1, wakeStrain=1, wakePlot=24,
                     fdScheme=scheme)
        write_geom(os.path.join(d, "geom01.nml"), nb=2, nc=2, ns=8, nNwake=12,
                   wakeTruncateNt=24, axisymmetrySwitch=1, span=R, chord=c, Omega=om,
                   theta0=8.0, pivotLE=0.25, hubCoords=[0.0, 0.0, 0.6],
                   cgCoords=[0.0, 0.0, 0.6], velBody=[0.0, 0.0, -2.0],
                   apparentViscCoeff=ROTOR_AVC, spanwiseCore=ROTOR_CORE,
                   streamwiseCoreVec=ROTOR_CORE)
        write_geom(os.path.join(d, "geom02.nml"), surfaceType=-1, imagePlane=3,
                   imageRotorNum=1, nb=2, nc=2, ns=8, nNwake=12)
    d = case("image_rotor_xy")
    write_config(os.path.join(d, "config.nml"), nt=48, dt=math.radians(5.0) / om, nr=3,
                 velSound=340.0, wakeDissipation=1, fdScheme=2, wakePlot=24, wakeTipPlot=24)
    write_geom(os.path.join(d, "geom01.nml"), nb=3, nc=2, ns=8, nNwake=12, wakeTruncateNt=24,
               span=R, chord=c, Omega=om, theta0=8.0, thetaC=2.0,
               pivotLE=0.25, hubCoords=[1.3, 1.6, 0.0], cgCoords=[1.3, 1.6, 0.0],
               velBody=[0.0, 0.0, -2.0], apparentViscCoeff=ROTOR_AVC,
               spanwiseCore=ROTOR_CORE, streamwiseCoreVec=ROTOR_CORE)
    for k, plane in ((2, 1), (3, 2)):
        write_geom(os.path.join(d, f"geom0{k}.nml"), surfaceType=-1, imagePlane=plane,
                   imageRotorNum=1, nb=3, nc=2, ns=8, nNwake=12)

    # custom trajectory (classdef.f90:3064-3080, main.f90:112-117 and
    # 419-426): the AR-4 wing's body velocity and rate follow
    # trajectory01.in, one line per step: speed varying by 20%, a plunge
    # velocity and a pitch rate
    nt_t = 60
    d = case("traj_wing")
    write_config(os.path.join(d, "config.nml"), nt=nt_t, dt=0.025, nr=1)
    write_geom(os.path.join(d, "geom01.nml"), nc=4, ns=13, nNwake=200,
               customTrajectorySwitch=1)
    with open(os.path.join(d, "trajectory01.in"), "w") as f:
        for i in range(1, nt_t + 1):
            u = -10.0 * (1.0 + 0.2 * math.sin(2.0 * math.pi * i / 30.0))
            w = -0.5 * math.sin(2.0 * math.pi * i / 20.0)
            q = 0.1 * math.cos(2.0 * math.pi * i / 30.0)
            f.write(f"{u:.15g} 0.0 {w:.15g} 0.0 {q:.15g} 0.0\n")

    # wake tip plots every 12 steps of a forward-flight rotor in slow start
    # (omegaSlow ramps over 20 steps): far wake from step 12, full at step
    # 30 (wakeTruncateNt 30), then shifted on every rollup
    # tip_rotor_cw: the same rotor turning clockwise (negative Omega): Γ
    # changes sign, so the tip plot takes each row's highest Γ, and the
    # azimuthal ages are negative
    for name, sgn in (("tip_rotor", 1.0), ("tip_rotor_cw", -1.0)):
        d = case(name)
        write_config(os.path.join(d, "config.nml"), nt=48, dt=math.radians(5.0) / om, nr=1,
                     velSound=340.0, wakeDissipation=1, wakeStrain=1, wakeTipPlot=12,
                     slowStart=1, slowStartNt=20)
        write_geom(os.path.join(d, "geom01.nml"), nb=2, nc=2, ns=8, nNwake=12,
                   wakeTruncateNt=30, span=R, chord=c, Omega=sgn * om, theta0=8.0,
                   pivotLE=0.25, velBody=[-13.6, 0.0, -2.0], apparentViscCoeff=ROTOR_AVC,
                   spanwiseCore=ROTOR_CORE, streamwiseCoreVec=ROTOR_CORE)

    # restart files at steps 20 and 40 of a forward-flight rotor (far wake
    # from step 12); the test resumes a copy from step 20.  The reference's
    # restart write lists every scheme's arrays and the airfoil alpha0
    # (classdef.f90:2485-2505), so it succeeds only under AM4, the one
    # scheme that allocates them all, and with nAirfoils > 0 (one dummy
    # table, unused by forceCalcSwitch = 0)
    d = case("restart_rotor")
    write_config(os.path.join(d, "config.nml"), nt=40, dt=math.radians(5.0) / om, nr=1,
                 velSound=340.0, fdScheme=5, wakeDissipation=1, wakeStrain=1,
                 restartWriteNt=20)
    write_geom(os.path.join(d, "geom01.nml"), nb=2, nc=2, ns=8, nNwake=12, wakeTruncateNt=36,
               nAirfoils=1,
               span=R, chord=c, Omega=om, theta0=8.0, pivotLE=0.25, velBody=[-13.6, 0.0, -2.0],
               apparentViscCoeff=ROTOR_AVC, spanwiseCore=ROTOR_CORE,
               streamwiseCoreVec=ROTOR_CORE)

    # near-wake skew plots at steps 12 and 24 of a forward-flight rotor
    d = case("skew_rotor")
    write_config(os.path.join(d, "config.nml"), nt=24, dt=math.radians(5.0) / om, nr=1,
                 velSound=340.0, wakeDissipation=1)
    write_geom(os.path.join(d, "geom01.nml"), nb=2, nc=1, ns=10, nNwake=30, span=R, chord=c,
               Omega=om, theta0=8.0, pivotLE=0.25, velBody=[-13.6, 0.0, -2.0],
               apparentViscCoeff=ROTOR_AVC, spanwiseCore=ROTOR_CORE,
               streamwiseCoreVec=ROTOR_CORE, skewPlotSwitch=12)

    # camber lines (rotor_getCamber, classdef.f90:4060-4112): the AR-4 wing
    # with a 4% parabolic camber.  One file only: the reference sizes its
    # camber arrays from this%nCamberFiles before setting it
    # (classdef.f90:2618 against 2668), so with two or more files its
    # namelist read aborts.
    d = case("camber_wing")
    write_config(os.path.join(d, "config.nml"), nt=60, dt=0.025, nr=1)
    write_geom(os.path.join(d, "geom01.nml"), nc=4, ns=13, nNwake=200, nCamberFiles=1,
               camberSectionLimit=[1.0], camberFile=["'camber1.dat'"])
    for name, m, n in (("camber1.dat", 0.04, 21),):
        with open(os.path.join(d, name), "w") as f:
            f.write(f"{n}\n")
            for k in range(n):
                xk = 2.0 * k / (n - 1)           # chord of 2 units, normalised on reading
                f.write(f"{xk:.15g} {4.0 * m * xk * (1.0 - xk / 2.0):.15g}\n")

    # far-wake burst every step (main.f90:490-496) with a small skew limit,
    # strain off so that the burst cores persist
    d = case("burst_rotor")
    write_config(os.path.join(d, "config.nml"), nt=48, dt=math.radians(5.0) / om, nr=1,
                 velSound=340.0, wakeDissipation=1, wakeBurst=1)
    write_geom(os.path.join(d, "geom01.nml"), nb=2, nc=1, ns=10, nNwake=12, wakeTruncateNt=44,
               span=R, chord=c, Omega=om, theta0=8.0, pivotLE=0.25, velBody=[-13.6, 0.0, -2.0],
               apparentViscCoeff=ROTOR_AVC, spanwiseCore=ROTOR_CORE,
               streamwiseCoreVec=ROTOR_CORE, skewLimit=0.02)

    # filament dumps for gridgen at steps 20 and 40 of the same rotor, and
    # the grid the velocity field is evaluated on (gridconfig.nml 0.2)
    d = case("grid_rotor")
    write_config(os.path.join(d, "config.nml"), nt=40, dt=math.radians(5.0) / om, nr=1,
                 velSound=340.0, wakeDissipation=1, wakeStrain=1, gridPlot=20)
    write_geom(os.path.join(d, "geom01.nml"), nb=2, nc=2, ns=8, nNwake=12, wakeTruncateNt=36,
               span=R, chord=c, Omega=om, theta0=8.0, pivotLE=0.25, velBody=[-13.6, 0.0, -2.0],
               apparentViscCoeff=ROTOR_AVC, spanwiseCore=ROTOR_CORE,
               streamwiseCoreVec=ROTOR_CORE)
    with open(os.path.join(d, "gridconfig.nml"), "w") as f:
        f.write("&VERSION\nfileFormatVersion = 0.2\n/\n")
        f.write(f"&INPUTS\nnx = 7\nny = 7\nnz = 7\n"
                f"xyzMin = {fmt([-1.2 * R, -1.2 * R, -0.6 * R])}\n"
                f"xyzMax = {fmt([1.2 * R, 1.2 * R, 0.2 * R])}\n"
                f"vel = 13.6, 0.0, 2.0\nfileRangeStart = 20\nfileRangeStep = 20\n"
                f"fileRangeEnd = 40\n/\n")

    # inflow and bound-circulation plots of a rotor above a wing: the rotor
    # plots every 12 steps, the wing every 24; at step 24 the reference
    # writes rotor 1's circulation twice (gamma2file is called with rotor 1)
    d = case("plots_rotor_wing")
    write_config(os.path.join(d, "config.nml"), nt=24, dt=math.radians(5.0) / om, nr=2,
                 velSound=340.0)
    write_geom(os.path.join(d, "geom01.nml"), nb=3, nc=1, ns=30, nNwake=24, span=R, chord=c,
               Omega=om, theta0=10.0, pivotLE=0.25, velBody=[-13.6, 0.0, 0.0],
               apparentViscCoeff=ROTOR_AVC, spanwiseCore=ROTOR_CORE,
               streamwiseCoreVec=ROTOR_CORE, inflowPlotSwitch=12, gammaPlotSwitch=12)
    write_geom(os.path.join(d, "geom02.nml"), nb=1, nc=4, ns=30, nNwake=24, span=2.0,
               chord=0.25, Omega=0.0, theta0=3.0, hubCoords=[0.0, 0.0, -0.5],
               fromCoords=[0.0, 1.0, 0.0], velBody=[-13.6, 0.0, 0.0],
               apparentViscCoeff=ROTOR_AVC, spanwiseCore=ROTOR_CORE,
               streamwiseCoreVec=ROTOR_CORE, inflowPlotSwitch=24, gammaPlotSwitch=24)

    # slow start (main.f90:97-105, 428-452): the rotor speed ramps from rest
    # over 20 steps, linearly (1), as tanh(5 n/N) (2) or as a centred tanh (3)
    for k in (1, 2, 3):
        d = case(f"slowstart{k}_rotor")
        write_config(os.path.join(d, "config.nml"), nt=48, dt=math.radians(5.0) / om, nr=1,
                     velSound=340.0, slowStart=k, slowStartNt=20)
        write_geom(os.path.join(d, "geom01.nml"), nb=2, nc=1, ns=10, nNwake=48, span=R,
                   chord=c, Omega=om, theta0=8.0, pivotLE=0.25, velBody=[0.0, 0.0, 0.0],
                   apparentViscCoeff=ROTOR_AVC, spanwiseCore=ROTOR_CORE,
                   streamwiseCoreVec=ROTOR_CORE)

    # cfg5: rotor in forward flight above a wing (15 revolutions); the
    # _short variant (half a revolution) is compared with the reference.  The
    # reference solver turns NaN from the first step whenever two coupled
    # surfaces have different spanwise panel counts (rotor 30, wing 40 here;
    # each alone, or both with 30, run normally), so the short parity case
    # gives the wing 30 spanwise panels.
    dt5 = math.radians(5.0) / om
    for name, nt5, ns_w, opts in (("cfg5_rotor_wing", 1080, 40, {}),
                                  ("cfg5_rotor_wing_short", 36, 30, {}),
                                  ("subiter_rotor_wing", 36, 30, dict(ntSubInit=5, ntSub=3))):
        d = case(name)
        write_config(os.path.join(d, "config.nml"), nt=nt5, dt=dt5, nr=2, velSound=340.0,
                     **opts)
        write_geom(os.path.join(d, "geom01.nml"), nb=3, nc=1, ns=30, nNwake=nt5, span=R,
                   chord=c, Omega=om, theta0=10.0, pivotLE=0.25, velBody=[-13.6, 0.0, 0.0],
                   apparentViscCoeff=ROTOR_AVC, spanwiseCore=ROTOR_CORE,
                   streamwiseCoreVec=ROTOR_CORE)
        write_geom(os.path.join(d, "geom02.nml"), nb=1, nc=4, ns=ns_w, nNwake=nt5, span=2.0,
                   chord=0.25, Omega=0.0, theta0=3.0, hubCoords=[0.0, 0.0, -0.5],
                   fromCoords=[0.0, 1.0, 0.0], velBody=[-13.6, 0.0, 0.0],
                   apparentViscCoeff=ROTOR_AVC, spanwiseCore=ROTOR_CORE,
                   streamwiseCoreVec=ROTOR_CORE)

    # restart files of the rotor above the wing under AM2 with sub-iterations,
    # both surfaces rolling up into a truncated far wake; written at step 18
    # (the reference's restart write aborts for AM2, see restart_rotor)
    d = case("restart_rotor_wing")
    write_config(os.path.join(d, "config.nml"), nt=36, dt=dt5, nr=2, velSound=340.0,
                 ntSubInit=5, ntSub=3, wakeDissipation=1, wakeStrain=1, restartWriteNt=18)
    write_geom(os.path.join(d, "geom01.nml"), nb=3, nc=1, ns=30, nNwake=12, wakeTruncateNt=30,
               span=R, chord=c, Omega=om, theta0=10.0, pivotLE=0.25, velBody=[-13.6, 0.0, 0.0],
               apparentViscCoeff=ROTOR_AVC, spanwiseCore=ROTOR_CORE,
               streamwiseCoreVec=ROTOR_CORE)
    write_geom(os.path.join(d, "geom02.nml"), nb=1, nc=4, ns=30, nNwake=12, wakeTruncateNt=30,
               span=2.0, chord=0.25, Omega=0.0, theta0=3.0, hubCoords=[0.0, 0.0, -0.5],
               fromCoords=[0.0, 1.0, 0.0], velBody=[-13.6, 0.0, 0.0],
               apparentViscCoeff=ROTOR_AVC, spanwiseCore=ROTOR_CORE,
               streamwiseCoreVec=ROTOR_CORE)


if __name__ == "__main__":
    make_all(sys.argv[1] if len(sys.argv) > 1 else HERE)
