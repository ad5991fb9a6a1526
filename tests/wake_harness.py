"""Drives one call sequence of main.f90's time loop (400-1452) against any
object with the vlm.Context rotor API (the HIP library or the CPU oracle).

The circulation solve is replaced by a prescribed, step-dependent wing
circulation so that the wake machinery is exercised on its own; the solve
has its own tests.
"""
import numpy as np

from tests.geometry import blade_grid, rings_from_grids, rot_z


class Case:
    """surfaces: optional list of per-rotor overrides (nb, nc, ns, nNwake,
    nFwake, omega, span, chord, pitch, rollup, hub, ...); rotor ir then
    uses the case's values with surfaces[ir] applied (nr = len(surfaces)).
    presc: (prescWakeNt, prescWakeGenNt) of a prescribed far wake.
    image: (source rotor, imagePlane) of an image rotor, whose wing is the
    source's mirrored and whose circulation and wake are mirrored copies."""

    def __init__(self, nb=2, nc=2, ns=4, nNwake=6, nFwake=4, nt=14, scheme=3, axisym=False,
                 dissipation=False, strain=False, omega=0.0, vel=(10.0, 0.0, 0.0), dt=0.02,
                 span=1.0, root=0.2, chord=0.25, pitch=6.0, rollup=(3, 4), rollupMin=1,
                 spanwiseCore=0.02, scv=None, bug_compat=True, nr=1, hub_dx=0.0, hub=None,
                 surfaces=None, presc=None, image=None):
        self.__dict__.update(locals())
        del self.__dict__["self"]
        if scv is None:
            self.scv = np.full(ns + 1, 0.02)
        if surfaces is not None:
            self.nr = len(surfaces)

    def surface(self, ir):
        """The case as rotor ir sees it."""
        if self.surfaces is None:
            return self
        d = {k: v for k, v in self.__dict__.items() if k not in ("surfaces", "nr", "scv")}
        d.update(self.surfaces[ir])
        if "scv" not in self.surfaces[ir]:
            d["scv"] = np.full(d["ns"] + 1, 0.02)
        return Case(**d)

    @property
    def dims(self):
        return (self.nb, self.ns, self.nNwake, self.nFwake)


def blade_grids(case, t, ir=0):
    """Rigid motion at time t: blades rotate about z through the hub and
    the whole rotor translates with -vel."""
    if case.surface(ir).image is not None:
        src, plane = case.surface(ir).image
        grids, hub = blade_grids(case, t, src)
        for g in grids:
            g[..., plane - 1] *= -1.0
        hub = hub.copy()
        hub[plane - 1] *= -1.0
        return grids, hub
    case = case.surface(ir)
    hub = np.array(case.hub if case.hub is not None else [case.hub_dx * ir, 0.0, 0.0], float)
    g0 = blade_grid(case.nc, case.ns, chord=case.chord, span=case.span,
                    root=case.root, pitch_deg=case.pitch, seed=100 + ir)
    grids = []
    for b in range(case.nb):
        ang = 2 * np.pi * b / case.nb + case.omega * t
        grids.append(rot_z(g0, ang) + hub - np.asarray(case.vel) * t)
    return grids, hub - np.asarray(case.vel) * t


def wing_gamma(case, it, ir=0):
    """Smooth prescribed circulation, equal on every blade (as map_gam
    copies blade 1 under axisymmetry)."""
    case = case.surface(ir)
    ic = np.arange(case.nc)[None, :]
    js = np.arange(case.ns)[:, None]
    g = -(1.0 + 0.3 * np.sin(0.7 * it + 0.5 * ir)) * np.sin(np.pi * (js + 0.5) / case.ns) * (ic + 1) / case.nc
    return np.broadcast_to(g[None], (case.nb, case.ns, case.nc)).copy()


def setup(sys_, case):
    irs = []
    for ir in range(case.nr):
        c = case.surface(ir)
        k = sys_.rotor_create(c.nb, c.nc, c.ns, c.nNwake, c.nFwake,
                              nbConvect=1 if c.axisym else c.nb,
                              rollupStart=c.rollup[0], rollupEnd=c.rollup[1],
                              rollupMin=c.rollupMin, spanwiseCore=c.spanwiseCore,
                              streamwiseCoreVec=c.scv)
        irs.append(k)
        grids, hub = blade_grids(case, 0.0, ir)
        corners = rings_from_grids(grids)
        rc = np.full(corners.shape[:3] + (4,), 0.01)
        sys_.rotor_set_wing(k, corners, rc, wing_gamma(case, 0, ir))
        sys_.rotor_set_axis(k, [0.0, 0.0, 1.0], hub)
        if c.presc is not None:
            sys_.rotor_set_prescribed_wake(k, c.presc[0], c.presc[1], c.omega * c.dt)
        if c.image is not None:
            sys_.rotor_set_image(k, irs[c.image[0]], c.image[1])
        sys_.rotor_init_cursors(k)
        sys_.rotor_assignshed(k, 1)  # main.f90:227-233: shed as TE
    sys_.set_bug_compat(case.bug_compat)
    return irs


def step(sys_, case, irs, it):
    t = it * case.dt
    for ir in irs:
        sys_.rotor_begin_step(ir, it)
    for ir in irs:
        grids, hub = blade_grids(case, t, ir)
        sys_.rotor_set_wing(ir, rings_from_grids(grids))
        sys_.rotor_set_axis(ir, [0.0, 0.0, 1.0], hub)
        sys_.rotor_assignshed(ir, 0)
        sys_.rotor_age_wake(ir, case.dt)
        if case.dissipation:
            sys_.rotor_dissipate_wake(ir, case.dt, 1.8e-5, 50.0, 0.0)
    for k, ir in enumerate(irs):  # stands in for the circulation solve
        if case.surface(k).image is not None:
            sys_.rotor_mirror_gamma(ir)
        else:
            sys_.rotor_set_wing(ir, gam=wing_gamma(case, it, k))
    sys_.convect(case.scheme, it, case.dt)
    if case.strain:
        for ir in irs:
            sys_.rotor_strain_wake(ir)
    for k, ir in enumerate(irs):
        if case.surface(k).image is not None:  # main.f90:1433-1436
            sys_.rotor_mirror_wake(ir, 2)
            continue
        rn, _ = sys_.rotor_cursors(ir)
        if rn == 1:
            sys_.rotor_rollup(ir)
        sys_.rotor_assignshed(ir, 1)


def snapshot(sys_, case, irs):
    out = []
    for k, ir in enumerate(irs):
        dims = case.surface(k).dims
        w = sys_.rotor_get_wake(ir, dims)
        vn, vf = sys_.rotor_get_velocity(ir, 0, dims)
        vn1, vf1 = sys_.rotor_get_velocity(ir, 1, dims)
        extra = {}
        if case.surface(k).presc is not None:
            for t in (0, 1):
                extra[f"presc{t}"] = sys_.rotor_get_prescribed_wake(ir, dims[0], t)
        out.append(dict(cursors=sys_.rotor_cursors(ir), vn=vn, vf=vf, vn1=vn1, vf1=vf1, **w,
                        **extra))
    return out
