"""Pins of the oracle's grid set-up, extrapolation, grid update and G2P flip
that do not restate the oracle's own code.

* The activated buckets, the pressure DOF set and its numbering are
  recomputed by a model of the reference's *sparse* bucket grid, in its own
  coordinates: the grid origin and size from the particles' bounding box
  (TwoDScene.cpp:1588-1611, 1657-1690), findNodes' bucket-local stencils
  (:3683-3752), expandFluidNodesMarked(1) (:5021-5064), the corner order of
  connectSolidPhiNodes (:4433-4597) and updateSolidWeights (:1871-1962),
  updateLiquidPhi (:3286-3373), connectPressureNodes' face lookup
  (:4598-4747) and the DOF numbering of solveNodePressure (Pressure.cpp:
  348-439).  The oracle works on a fixed dense grid instead; comparing the
  two in world coordinates catches anything the dense form gets wrong about
  extent, clipping or order (it caught a ghost band one bucket too thin).
  fraction_inside comes from the reference's own MathUtilities.h where the
  reference tree is present (oracle/_ref), else from the oracle, whose
  fraction_inside is bitwise the reference's (test_oracle_math.py).
* A lattice-seeded cube of liquid away from the walls has a closed-form DOF
  count, m^3 + 6 m^2, fixed by the liquid-phi radius 0.883644 dx.
* Extrapolation: a lone valid face spreads its value, exactly, to its
  two-step 6-neighbourhood and no further; a uniform valid field stays
  uniform.
* stepVelocity: u* = u + g dt at faces with mass, 0 elsewhere.
* G2P with flip = 0.5 on an affine field: v = a + B x_p (unscaled) and
  C = 0.5 B.

Each pin is also run against the oracle built with a planted mistake
(WCO_PLANT in oracle/wc_oracle.c), and must fail there.
"""
import numpy as np
import pytest
import scipy.ndimage as ndi

DT = 1.0 / 480
OFFSETS5 = [(0.0, 0.5, 0.5), (0.5, 0.0, 0.5), (0.5, 0.5, 0.0), (0.0, 0.0, 0.0), (0.5, 0.5, 0.5)]


# ------------------------------------------------------------ helpers

@pytest.fixture(scope="module")
def variants(oracle, tmp_path_factory):
    """Oracle libraries with planted mistakes, built once per module."""
    d = str(tmp_path_factory.mktemp("oracle_plants"))
    cache = {}

    def get(plant, ghost=None):
        key = (plant, ghost)
        if key not in cache:
            cache[key] = oracle.load_variant(oracle.build_variant(d, plant, ghost))
        return cache[key]
    return get


def fraction_inside4(oracle):
    """The reference's fraction_inside (compiled from its sources) or the
    oracle's bitwise restatement of it."""
    R = oracle.ref_lib()
    if R is not None:
        return R.ref_fraction_inside4
    return oracle.lib().wco_fraction_inside4


def box_phi(p, n, dx):
    """DistanceFields.cpp:68-78 box_phi of the container [0, n dx]^3 with
    zero corner radius, negated for a container (:486): positive inside.
    p: (..., 3)."""
    lo, hi = 0.0, n * dx
    c, e = 0.5 * (lo + hi), 0.5 * (hi - lo)
    dd = np.abs(p - c) - e
    da = np.maximum(dd, 0.0)
    m = np.minimum(dd.max(axis=-1), 0.0)
    return -(m + np.sqrt((da * da).sum(axis=-1)))


def reference_dofs(x, n, dx, fi4, radius):
    """DOF cells (world cell coordinates, in DOF order) and activated buckets
    (world bucket coordinates) of the reference's sparse grid."""
    B = 4 * dx
    D = 2 * B
    bbmin, bbmax = x.min(axis=1), x.max(axis=1)
    bbx_min = np.floor(bbmin / D) * D
    bbx_max = np.ceil(bbmax / D) * D
    content = bbx_max - bbx_min + B * 3.0 * 2.0
    cells = np.maximum(1, np.ceil(content / dx).astype(int))
    minc = bbx_min - B * 3.0
    nbk = np.maximum(1, np.ceil(cells * dx / B).astype(int))
    origin = np.rint(minc / dx).astype(int)  # world cell of reference cell 0
    handle = np.floor((x - minc[:, None]) / B).astype(int)
    touched = np.zeros(nbk, dtype=bool)  # [bi, bj, bk]
    stencil_p = None
    for off in OFFSETS5:
        corner = handle * B + minc[:, None] + np.array(off)[:, None] * dx
        lp = (x - corner) / dx
        il = np.floor(lp).astype(int)
        fr = lp - np.floor(lp)
        low = np.where(fr > 0.5, 0, -1)
        base = handle * 4 + il + low  # global reference node of the stencil's first node
        for c in np.ndindex(3, 3, 3):
            node = base + np.array(c)[:, None]
            b = node // 4
            touched[b[0], b[1], b[2]] = True
        if off == (0.5, 0.5, 0.5):
            stencil_p = base
    act = ndi.binary_dilation(touched, structure=np.ones((3, 3, 3), bool))
    N = nbk * 4

    def active(I):  # I: (..., 3) global reference node indices
        ok = np.all((I >= 0) & (I < N), axis=-1)
        Ic = np.clip(I, 0, N - 1) // 4
        return ok & act[Ic[..., 0], Ic[..., 1], Ic[..., 2]]

    # liquid phi at cell centres (default 3 B; stencil minimum; solid override)
    r0 = max(dx * 0.883644, radius)
    lphi = np.full(N, 3.0 * B)
    for c in np.ndindex(3, 3, 3):
        I = (stencil_p + np.array(c)[:, None]).T
        on = active(I)
        npos = minc + (I + 0.5) * dx
        d = x.T - npos
        phi = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]) - r0
        I, phi = I[on], phi[on]
        np.minimum.at(lphi, (I[:, 0], I[:, 1], I[:, 2]), phi)
    grid = np.stack(np.meshgrid(*[np.arange(m) for m in N], indexing="ij"), axis=-1)
    cell_on = active(grid)
    solid = box_phi(minc + (grid + 0.5) * dx, n, dx) < 0.0
    lphi[cell_on & solid] = -0.5 * dx

    def corner_phi(I):
        return box_phi(minc + I * dx, n, dx) if active(I) else 0.5 * dx

    corners = {0: [(0, 0, 0), (0, 1, 0), (0, 0, 1), (0, 1, 1)],
               1: [(0, 0, 0), (0, 0, 1), (1, 0, 0), (1, 0, 1)],
               2: [(0, 0, 0), (0, 1, 0), (1, 0, 0), (1, 1, 0)]}
    wcache = {}

    def weight(d, I):
        key = (d, tuple(I))
        if key not in wcache:
            ph = [corner_phi(np.array(I) + np.array(o)) for o in corners[d]]
            f = 1.0 - fi4(*[float(p) for p in ph])
            wcache[key] = min(max(f, 0.0), 1.0)
        return wcache[key]

    # candidate cells: activated, phi < 0; a DOF has a face with w > 0
    cand = np.argwhere(cell_on & (lphi < 0))
    dofs = []
    for I in cand:
        for f in range(6):
            d, hi = f >> 1, f & 1
            q = I.copy()
            q[d] += hi
            if active(q) and weight(d, q) > 0:
                dofs.append(I)
                break
    dofs = np.array(dofs, dtype=np.int64).reshape(-1, 3)
    # numbering: buckets in Sorter::bucket_index order ((bk nj + bj) ni + bi,
    # Sorter.h:66-68), then the bucket-local node k 16 + j 4 + i
    bkt = dofs // 4
    loc = dofs % 4
    key = ((bkt[:, 2] * nbk[1] + bkt[:, 1]) * nbk[0] + bkt[:, 0]) * 64 + \
        loc[:, 2] * 16 + loc[:, 1] * 4 + loc[:, 0]
    dofs = dofs[np.argsort(key, kind="stable")] + origin
    ab = np.argwhere(act) + origin // 4
    return dofs, ab


def oracle_dofs(s, ghost):
    """DOF cells of an oracle scene in world coordinates, in DOF order, and its
    activated buckets in world bucket coordinates."""
    g = s.grid
    dof = s.field("dof")
    idx = np.nonzero(dof >= 0)[0]
    order = np.argsort(dof[idx])
    idx = idx[order]
    b = idx >> 6
    l = idx & 63
    bi = b % g.nb[0]
    bj = (b // g.nb[0]) % g.nb[1]
    bk = b // (g.nb[0] * g.nb[1])
    ijk = np.stack([bi * 4 + (l & 3), bj * 4 + ((l >> 2) & 3), bk * 4 + (l >> 4)], axis=1) - ghost
    act = s.field("act")
    a = np.nonzero(act)[0]
    ab = np.stack([a % g.nb[0], (a // g.nb[0]) % g.nb[1], a // (g.nb[0] * g.nb[1])], axis=1) \
        - ghost // 4
    return ijk, dof[idx], ab


def scenes(oracle):
    """(name, n, x) of the scenes the grid pins run on: dx = 1/n is a power of
    two, so node positions are exact in both grid forms."""
    out = []
    g, x, _ = oracle.dam_break(8, block_hi=(1.0, 1.0, 1.0), seed=1)
    out.append(("full box 8^3", 8, x))
    g, x, _ = oracle.dam_break(16, block_hi=(0.4, 0.6, 1.0), seed=2)
    out.append(("dam break 16^3", 16, x))
    g, x, _ = oracle.dam_break(16, block_hi=(0.3, 0.25, 0.5), seed=3)
    out.append(("corner block 16^3", 16, x))
    rng = np.random.default_rng(4)
    out.append(("scattered 16^3", 16, rng.uniform(0.02, 0.98, (3, 300))))
    return out


_MODEL = {}


def reference_model(oracle, name, n, x, radius):
    """reference_dofs of a scene, computed once per session."""
    if name not in _MODEL:
        _MODEL[name] = reference_dofs(x, n, 1.0 / n, fraction_inside4(oracle), radius)
    return _MODEL[name]


def check_grid_setup(oracle, L, ghost):
    for name, n, x in scenes(oracle):
        s = oracle.OracleScene((n, n, n), 1.0 / n, x, L=L)
        s.pre_solve(DT)
        ref, ref_act = reference_model(oracle, name, n, x, s.st.radius)
        got, num, got_act = oracle_dofs(s, ghost)
        ra = {tuple(r) for r in ref_act}
        ga = {tuple(r) for r in got_act}
        assert ra == ga, f"{name}: activated buckets differ ({len(ra)} vs {len(ga)})"
        assert got.shape == ref.shape, f"{name}: {len(got)} DOFs, the reference's grid has {len(ref)}"
        assert np.array_equal(num, np.arange(len(num))), f"{name}: DOF numbers are not 0..n-1"
        assert np.array_equal(got, ref), f"{name}: DOF cells or their order differ"


def lattice_block(lo, m, dx):
    """Particles at the centres of the 2x2x2 sub-cells of an m^3 block of
    cells starting at cell lo."""
    c = np.arange(m)
    sub = (np.arange(2) + 0.5) / 2
    P = np.stack(np.meshgrid(c, c, c, sub, sub, sub, indexing="ij"), axis=0).reshape(6, -1)
    return np.stack([(lo[a] + P[a] + P[3 + a]) * dx for a in range(3)])


def check_closed_form_count(oracle, L):
    """A lattice cube of m^3 cells in the middle of a 32^3 box, far enough
    from the walls that its activated buckets hold no solid cell: the liquid
    phi of a cell centre half a cell outside a face is sqrt(0.75^2 + 2 *
    0.25^2) dx - 0.883644 dx < 0, outside an edge sqrt(2 * 0.75^2 + 0.25^2)
    dx - 0.883644 dx > 0, so the DOFs are the cube and one layer on each of
    its six faces, every face open: m^3 + 6 m^2."""
    n = 32
    dx = 1.0 / n
    for m in (2, 3, 6):
        x = lattice_block((12, 13, 11), m, dx)
        s = oracle.OracleScene((n, n, n), dx, x, L=L)
        nd = s.pre_solve(DT)
        assert nd == m ** 3 + 6 * m ** 2, (m, nd)


# ------------------------------------------------------------ grid set-up

def test_grid_setup_matches_sparse_reference_grid(oracle):
    check_grid_setup(oracle, oracle.lib(), oracle.GHOST)


def test_closed_form_dof_count(oracle):
    check_closed_form_count(oracle, oracle.lib())


@pytest.mark.parametrize("plant,what", [
    (1, "activation without the one-bucket ring"),
    (2, "activation with two rings"),
    (3, "stencil buckets from the base node only"),
    (4, "no solid override of the liquid phi"),
    (5, "DOF test w >= 0"),
    (6, "DOFs numbered i-fastest"),
])
def test_grid_pin_fails_on_planted_mistake(oracle, variants, plant, what):
    with pytest.raises(AssertionError):
        check_grid_setup(oracle, variants(plant), oracle.GHOST)


def test_grid_pin_fails_with_one_ghost_bucket(oracle, variants):
    """The dense grid with one bucket of ghost cells below the container
    clips the activation's dilation there, and loses the reference's
    zero-row DOFs at the activated region's edges."""
    with pytest.raises(AssertionError):
        check_grid_setup(oracle, variants(0, 4), 4)


def test_closed_form_fails_on_planted_mistake(oracle, variants):
    """With the particle radius alone (0.31 dx) in place of max(0.883644 dx,
    radius) the liquid layer and the cube itself turn to air."""
    with pytest.raises(AssertionError):
        check_closed_form_count(oracle, variants(12))


# ------------------------------------------------------------ extrapolation

def interior_scene(oracle, L, n=16):
    g, x, _ = oracle.dam_break(n, block_hi=(0.5, 0.5, 0.5), seed=6)
    s = oracle.OracleScene((n, n, n), 1.0 / n, x, L=L)
    s.activate()
    return s


def node_ijk(g):
    idx = np.arange(g.nnode)
    b = idx >> 6
    l = idx & 63
    return (b % g.nb[0]) * 4 + (l & 3), ((b // g.nb[0]) % g.nb[1]) * 4 + ((l >> 2) & 3), \
        (b // (g.nb[0] * g.nb[1])) * 4 + (l >> 4)


def check_extrapolation(oracle, L):
    s = interior_scene(oracle, L)
    g = s.grid
    i, j, k = node_ijk(g)
    act = s.field("act")[np.arange(g.nnode) >> 6] == 1
    G = oracle.GHOST
    c = np.array([G + 4, G + 5, G + 3])
    n0 = oracle.node_index(g, *c)
    u0 = 0.5
    u = np.zeros(g.nnode)
    v = np.zeros(g.nnode, dtype=np.uint8)
    u[n0] = u0
    v[n0] = 1
    for d in range(3):  # the same lone valid face on every axis
        s.set_field("unew", d, u)
        s.set_field("valid", d, v)
    s.extrapolate()
    dist = np.abs(i - c[0]) + np.abs(j - c[1]) + np.abs(k - c[2])
    for d in range(3):
        uu, vv = s.field("unew", d), s.field("valid", d)
        reach = act & (dist <= 2)
        assert np.array_equal(vv.astype(bool), reach), f"axis {d}: valid set is not the 2-ball"
        assert np.all(uu[reach] == u0) and np.all(uu[~reach] == 0.0), f"axis {d}: values"
    # a uniform field on a half space stays exactly uniform and grows by
    # two layers
    s = interior_scene(oracle, L)
    u1 = 0.75
    half = act & (i <= G + 6)
    for d in range(3):
        s.set_field("unew", d, np.where(half, u1, 0.0))
        s.set_field("valid", d, half.astype(np.uint8))
    s.extrapolate()
    for d in range(3):
        uu, vv = s.field("unew", d), s.field("valid", d).astype(bool)
        assert np.array_equal(vv, act & (i <= G + 8)), f"axis {d}: uniform field's valid set"
        assert np.all(uu[vv] == u1), f"axis {d}: uniform field changed"


def test_extrapolation_pins(oracle):
    check_extrapolation(oracle, oracle.lib())


@pytest.mark.parametrize("plant", [7, 8])
def test_extrapolation_pin_fails_on_planted_mistake(oracle, variants, plant):
    with pytest.raises(AssertionError):
        check_extrapolation(oracle, variants(plant))


# ------------------------------------------------------------ grid update

def check_step_velocity(oracle, L):
    n = 16
    g, x, v = oracle.dam_break(n, block_hi=(0.5, 0.5, 0.5), seed=7, sigma=0.4)
    grav = (0.3, -9.81, 1.7)
    s = oracle.OracleScene((n, n, n), 1.0 / n, x, v, gravity=grav, L=L)
    s.activate()
    s.p2g()
    s.step_velocity(DT)
    for d in range(3):
        M = s.field("mass", d)
        u = s.field("vel", d)
        us = s.field("ustar", d)
        has = M > 1e-20
        exp = u + grav[d] * DT
        assert np.abs(us[has] - exp[has]).max() <= 1e-13 * (np.abs(u).max() + abs(grav[d]) * DT), d
        assert np.all(us[~has] == 0.0), d


def test_step_velocity_pin(oracle):
    check_step_velocity(oracle, oracle.lib())


@pytest.mark.parametrize("plant", [10, 11])
def test_step_velocity_pin_fails_on_planted_mistake(oracle, variants, plant):
    with pytest.raises(AssertionError):
        check_step_velocity(oracle, variants(plant))


# ------------------------------------------------------------ G2P flip

def check_flip(oracle, L):
    """APIC with flip_coeff = 0.5 on an affine field: G2P returns v = a + B
    x_p and C = 0.5 B (TwoDScene.cpp:7018 scales only the affine matrix)."""
    n = 16
    g, x, _ = oracle.dam_break(n, block_hi=(0.75, 0.75, 0.75))
    a = np.array([0.3, -0.2, 0.1])
    Bm = np.array([[0.5, -1.0, 0.25], [2.0, 0.1, -0.3], [-0.7, 0.4, 1.1]])
    v = a[:, None] + Bm @ x
    Cm = np.repeat(Bm.reshape(9, 1), x.shape[1], axis=1)
    s = oracle.OracleScene((n, n, n), 1.0 / n, x, v, Cm, flip=0.5, L=L)
    s.set_solid_projection(False)
    s.activate()
    s.p2g()
    s.copy_vel_to_unew()
    dx = 1.0 / n
    inner = np.all((x > 2.5 * dx) & (x < 0.75 - 2.5 * dx), axis=0)
    s.g2p_advect(0.0)
    _, vg, Cg = s.particles()
    assert np.abs(vg[:, inner] - v[:, inner]).max() < 1e-12
    assert np.abs(Cg[:, inner] - 0.5 * Cm[:, inner]).max() < 1e-10


def test_flip_pin(oracle):
    check_flip(oracle, oracle.lib())


def test_flip_pin_fails_on_planted_mistake(oracle, variants):
    with pytest.raises(AssertionError):
        check_flip(oracle, variants(9))
