"""Host-side builders of the solid particles of the wet-cloth scenes.

A cloth enters the liquid path as its vertices (wc_set_elasto_particles):
each vertex carries its share of the sheet's area as rest volume (area x
thickness) and as mass (area x areal density), a rest volume fraction
1 - porosity, an orientation and a shape factor (the elasto fields the
reference's P2G reads, TwoDScene.cpp:6686-6699).
"""
from __future__ import annotations

import numpy as np


def cloth_sheet(nv, y, lo=(0.0, 0.0), hi=(1.0, 1.0), thickness=2e-3, areal_density=0.2,
                porosity=(0.3, 0.7), seed=2, ori=(0.0, 1.0, 0.0), sf=0.25, jitter=0.0,
                fluid_fraction=0.0, liquid_density=1000.0):
    """An nv x nv vertex sheet over [lo, hi] in x and z at height y, as a
    dict of the arrays Scene.set_elasto takes (x and ori (3, ne), the rest
    (ne,)).  The porosity of each vertex is seeded uniform in `porosity`;
    `ori` is one orientation for every vertex, or None for seeded unit
    vectors; sf < 0 seeds each shape factor uniform in [0, 1]; `jitter`
    moves each vertex by up to that fraction of the spacing (kept inside
    the sheet's square); `fluid_fraction` fills that fraction of each
    vertex's pores with liquid."""
    rng = np.random.default_rng(seed)
    hx = (hi[0] - lo[0]) / (nv - 1)
    hz = (hi[1] - lo[1]) / (nv - 1)
    gi, gk = np.meshgrid(np.arange(nv), np.arange(nv), indexing="ij")
    gi, gk = gi.ravel(), gk.ravel()
    ne = gi.size
    x = np.stack([lo[0] + gi * hx, np.full(ne, float(y)), lo[1] + gk * hz])
    if jitter > 0:
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
    out = dict(x=x, m=area * areal_density, fm=liquid_density * fvol, fvol=fvol, rest_vol=rvol,
               rest_vf=rvf, sf=np.full(ne, float(sf)))
    if ori is None:
        o = rng.standard_normal((3, ne))
        out["ori"] = o / np.linalg.norm(o, axis=0)
    else:
        out["ori"] = np.repeat(np.asarray(ori, dtype=np.float64)[:, None], ne, axis=1)
    if np.ndim(sf) == 0 and sf < 0:
        out["sf"] = rng.uniform(0.0, 1.0, ne)
    return out


def config2_cloth():
    """The cloth of the 128^3 wet-cloth scene: 257 x 257 vertices over the
    unit square at y = 0.3, 2 mm thick, 0.2 kg/m^2, porosity uniform in
    [0.3, 0.7] (seed 2), orientation along the sheet normal, shape factor
    0.25."""
    return cloth_sheet(257, 0.3, seed=2, ori=(0.0, 1.0, 0.0), sf=0.25)


def elasto_args(s):
    """The arrays of a sheet in Scene.set_elasto's order."""
    return s["x"], s["m"], s["fm"], s["fvol"], s["rest_vol"], s["rest_vf"], s["ori"], s["sf"]
