"""Synthetic lifting-surface lattices for parity tests.

A blade is an (nc+1) x (ns+1) grid of points; ring (ic, is) has corners
c1 = X[ic, is], c2 = X[ic+1, is], c3 = X[ic+1, is+1], c4 = X[ic, is+1]
(reference panel numbering, classdef.f90:113-123), so shared corners are
bitwise identical as in the reference.
"""
import numpy as np


def blade_grid(nc, ns, chord=1.0, span=4.0, root=0.0, pitch_deg=5.0, seed=None):
    x = np.linspace(-chord, 0.0, nc + 1)
    y = np.linspace(root, span, ns + 1)
    X = np.zeros((nc + 1, ns + 1, 3))
    X[:, :, 0] = x[:, None]
    X[:, :, 1] = y[None, :]
    th = np.deg2rad(pitch_deg)
    # pitch about the leading edge (y axis)
    xr = X[:, :, 0] * np.cos(th)
    zr = -X[:, :, 0] * np.sin(th)
    X[:, :, 0], X[:, :, 2] = xr, zr
    if seed is not None:
        rng = np.random.default_rng(seed)
        X = X + 1e-3 * rng.standard_normal(X.shape)
    return X


def rot_z(X, ang, hub=(0.0, 0.0, 0.0)):
    c, s = np.cos(ang), np.sin(ang)
    h = np.asarray(hub)
    Y = X - h
    out = np.empty_like(Y)
    out[..., 0] = c * Y[..., 0] - s * Y[..., 1]
    out[..., 1] = s * Y[..., 0] + c * Y[..., 1]
    out[..., 2] = Y[..., 2]
    return out + h


def rings_from_grids(grids):
    """grids: list over blades of (nc+1, ns+1, 3) -> corners (nb, ns, nc, 4, 3)."""
    out = []
    for X in grids:
        nc, ns = X.shape[0] - 1, X.shape[1] - 1
        c = np.zeros((ns, nc, 4, 3))
        for js in range(ns):
            for ic in range(nc):
                c[js, ic, 0] = X[ic, js]
                c[js, ic, 1] = X[ic + 1, js]
                c[js, ic, 2] = X[ic + 1, js + 1]
                c[js, ic, 3] = X[ic, js + 1]
        out.append(c)
    return np.stack(out)


def cp_and_normals(corners):
    """Collocation points and normals as calcCP/calcN (classdef.f90:658-697)."""
    nb, ns, nc = corners.shape[:3]
    cp = np.zeros((nb, ns, nc, 3))
    nv = np.zeros((nb, ns, nc, 3))
    for b in range(nb):
        for js in range(ns):
            for ic in range(nc):
                pc = corners[b, js, ic]
                cp[b, js, ic] = ((pc[0] + pc[3]) * 0.25 + (pc[1] + pc[2]) * 0.75) * 0.5
                n = np.cross(pc[2] - pc[0], pc[3] - pc[1])
                nv[b, js, ic] = n / np.linalg.norm(n)
    return cp, nv
