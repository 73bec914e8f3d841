"""Solid particles of the mixture scenes (a helper of the tests, not a test
module): the vertices of a flat cloth sheet, each carrying its share of the
sheet's area as rest volume (area x thickness) and mass (area x areal
density), a seeded rest volume fraction 1 - porosity with the porosity
uniform in [lo, hi], an orientation and a shape factor.  The liquid path
sees the cloth through these per-vertex fields only (the reference's P2G of
elasto particles, TwoDScene.cpp:6686-6699)."""
import numpy as np


def sheet(nv, y, **kw):
    """libwetcloth_amd.scenes.cloth_sheet (a host-side builder, no GPU)."""
    from libwetcloth_amd.scenes import cloth_sheet
    return cloth_sheet(nv, y, **kw)


def wet_sheet(seed=4):
    """A 33 x 33 sheet at y = 0.3 over the unit square, 2 cm thick (psi
    reaches ~0.15 in 6 cm cells), jittered by 0.3 spacings, with seeded
    orientations and shape factors and a fifth of its pores filled."""
    return sheet(33, 0.3, thickness=0.02, ori=None, sf=-1, jitter=0.3, fluid_fraction=0.2,
                 seed=seed)


def config2_sheet():
    """The cloth of the 128^3 wet-cloth scene (scenes.config2_cloth)."""
    from libwetcloth_amd.scenes import config2_cloth
    return config2_cloth()


def lattice(n, dx, rest_vf, m=1e-3):
    """One solid particle at every cell centre of an n^3 grid of spacing dx,
    each of rest volume dx^3 and rest volume fraction rest_vf: the volume
    fraction psi is rest_vf on every face whose stencil the lattice covers
    (the quadratic B-spline weights of a lattice sum to one)."""
    c = (np.arange(n) + 0.5) * dx
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    x = np.stack([X.ravel(), Y.ravel(), Z.ravel()])
    ne = x.shape[1]
    return dict(x=x, m=np.full(ne, m), fm=np.zeros(ne), fvol=np.zeros(ne),
                rest_vol=np.full(ne, dx ** 3), rest_vf=np.full(ne, float(rest_vf)),
                ori=np.repeat(np.array([[0.0], [1.0], [0.0]]), ne, axis=1),
                sf=np.full(ne, 0.25))


def args(s):
    """The arrays in set_elasto's order."""
    return s["x"], s["m"], s["fm"], s["fvol"], s["rest_vol"], s["rest_vf"], s["ori"], s["sf"]
