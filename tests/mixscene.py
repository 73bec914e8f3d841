"""Solid particles of the mixture scenes (a helper of the tests, not a test
module): the vertices of a flat cloth sheet, each carrying its share of the
sheet's area as rest volume (area x thickness) and mass (area x areal
density), a seeded rest volume fraction 1 - porosity with the porosity
uniform in [lo, hi], an orientation and a shape factor.  The liquid path
sees the cloth through these per-vertex fields only (the reference's P2G of
elasto particles, TwoDScene.cpp:6686-6699)."""
import numpy as np


def sheet(nv, y, lo=(0.0, 0.0), hi=(1.0, 1.0), thickness=2e-3, areal_density=0.2,
          porosity=(0.3, 0.7), seed=2, ori=(0.0, 1.0, 0.0), sf=0.25, jitter=0.0,
          fluid_fraction=0.0):
    """An nv x nv vertex sheet over [lo, hi] in x and z at height y.  `ori`
    is one orientation for every vertex, or None for seeded unit vectors;
    `jitter` moves each vertex by up to that fraction of the spacing;
    `fluid_fraction` fills that fraction of each vertex's pore volume with
    liquid (its fluid volume and mass at 1000 kg/m^3)."""
    rng = np.random.default_rng(seed)
    hx = (hi[0] - lo[0]) / (nv - 1)
    hz = (hi[1] - lo[1]) / (nv - 1)
    gi, gk = np.meshgrid(np.arange(nv), np.arange(nv), indexing="ij")
    gi, gk = gi.ravel(), gk.ravel()
    ne = gi.size
    x = np.stack([lo[0] + gi * hx, np.full(ne, float(y)), lo[1] + gk * hz])
    if jitter > 0:  # kept inside the sheet's square (and so the container)
        x += rng.uniform(-jitter, jitter, x.shape) * np.array([[hx], [hx], [hz]])
        x[0] = np.clip(x[0], lo[0], hi[0])
        x[2] = np.clip(x[2], lo[1], hi[1])
    # each vertex's share of the area: halved on an edge, quartered at a corner
    share = np.where((gi == 0) | (gi == nv - 1), 0.5, 1.0) * \
        np.where((gk == 0) | (gk == nv - 1), 0.5, 1.0)
    area = share * hx * hz
    rvol = area * thickness
    rvf = 1.0 - rng.uniform(porosity[0], porosity[1], ne)
    fvol = fluid_fraction * rvol * (1.0 - rvf)
    out = dict(x=x, m=area * areal_density, fm=1000.0 * fvol, fvol=fvol, rest_vol=rvol,
               rest_vf=rvf, sf=np.full(ne, float(sf)))
    if ori is None:
        o = rng.standard_normal((3, ne))
        out["ori"] = o / np.linalg.norm(o, axis=0)
    else:
        out["ori"] = np.repeat(np.asarray(ori, dtype=np.float64)[:, None], ne, axis=1)
    if np.ndim(sf) == 0 and sf < 0:
        out["sf"] = rng.uniform(0.0, 1.0, ne)
    return out


def wet_sheet(seed=4):
    """A 33 x 33 sheet at y = 0.3 over the unit square, 2 cm thick (psi
    reaches ~0.15 in 6 cm cells), jittered by 0.3 spacings, with seeded
    orientations and shape factors and a fifth of its pores filled."""
    return sheet(33, 0.3, thickness=0.02, ori=None, sf=-1, jitter=0.3, fluid_fraction=0.2,
                 seed=seed)


def config2_sheet():
    """The cloth of the 128^3 wet-cloth scene: 257 x 257 vertices over the
    unit square at y = 0.3, 2 mm thick, 0.2 kg/m^2, porosity uniform in
    [0.3, 0.7] (seed 2), orientation along the sheet normal, shape factor
    0.25."""
    return sheet(257, 0.3, seed=2, ori=(0.0, 1.0, 0.0), sf=0.25)


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
