"""The oracle's mixture terms: a static porous solid (the cloth's vertices,
LiquidInfo::solve_solid off) in the liquid path.  Each part is pinned by a
check that does not retype its formula:

* the rotated drag and pore coefficients (MathUtilities.h:792-878) against
  the reference's own header compiled into oracle/_ref, and against the row
  sums of R diag(t, t, u) R^T built from a rotation taking z to the
  orientation;
* the drag coefficient (TwoDScene.cpp:4042-4097): its early returns, the
  linear part against Drummond & Tahir's (1984) permeabilities of fibre
  arrays in their fibre-radius form, the nonlinear part's homogeneity in
  |rho dv| (power yazdchi_power) and its absence at dv = 0;
* the capillary pressure (TwoDScene.cpp:563-571): early returns, its value
  sigma cos(theta) / r at psi = 1/2 and its growth with psi;
* the solid's P2G and the cell-centred saturation / volume fraction pass
  (TwoDScene.cpp:5952-5992, 6686-6736) against a brute-force numpy
  scatter over every node with the quadratic B-spline;
* a solid away from the liquid leaves the liquid-only substep bit for bit;
* with the drag and surface tension off, a uniform volume fraction psi
  scales the pressure system by (1 - psi) and leaves the substep's
  particles where the liquid-only substep puts them;
* the projected mixture flux (1 - psi) w u_f is divergence free on every
  DOF cell to the solve's tolerance (assembly and pressure application
  agree);
* each face's new liquid velocity is the liquid row of the implicit
  two-phase drag system [[M_f + h, -h], [-h, M_s + h]] (u_f, u_s) =
  (rhs_f - dt V_f grad p, M_s u_s - dt V_s grad p), solved by numpy;
* the system is a symmetric M-matrix;
* the pore pressure gradient draws liquid into the dry cloth and is linear
  in the surface tension;
* planted mistakes (WCO_PLANT 20-24) each fail one of these checks.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

from tests import mixscene

DT = 1.0 / 480


@pytest.fixture(scope="module")
def ref(oracle):
    import os
    if not os.path.exists(oracle.REF_LIB_PATH):
        pytest.skip("oracle/_ref not built (the reference tree is absent)")
    L = oracle.ref_lib()
    if not hasattr(L, "ref_rotated_drag"):
        pytest.skip("oracle/_ref predates the rotated-drag harness")
    return L


def vec(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.POINTER(C.c_double))


# ------------------------------------------------------------ pointwise terms

def rotated(L, a, d, k):
    return L.wco_rotated_drag(vec(a), vec(d), k)


def test_rotated_terms_match_reference(oracle, ref):
    L = oracle.lib()
    rng = np.random.default_rng(11)
    worst = 0.0
    cases = [np.zeros(3), np.array([0.0, 2.0, 0.0]), np.array([1e-12, 0.0, 0.0])]
    cases += [rng.standard_normal(3) * 10.0 ** rng.uniform(-3, 3) for _ in range(2000)]
    for a in cases:
        d = np.array([rng.uniform(0.1, 10)] * 2 + [rng.uniform(0.1, 10)])
        sf = rng.uniform(0, 1)
        for k in range(3):
            r1 = ref.ref_rotated_drag(vec(a), vec(d), k)
            o1 = L.wco_rotated_drag(vec(a), vec(d), k)
            r2 = ref.ref_rotated_pore_coeff(vec(a), sf, k)
            o2 = L.wco_rotated_pore_coeff(vec(a), sf, k)
            for r, o in ((r1, o1), (r2, o2)):
                worst = max(worst, abs(o - r) / max(abs(r), 1e-300) if r else abs(o))
    assert worst < 1e-12, worst


def rotation_to(n):
    """An orthonormal R whose third column is the unit vector n."""
    t = np.array([1.0, 0.0, 0.0]) if abs(n[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    e1 = np.cross(n, t)
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(n, e1)
    return np.stack([e1, e2, n], axis=1)


def test_rotated_terms_are_row_sums_of_the_rotated_tensor(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(5)
    for _ in range(300):
        a = rng.standard_normal(3) * rng.uniform(0.1, 5)
        n = a / np.linalg.norm(a)
        R = rotation_to(n)
        c, e, sf = rng.uniform(0.1, 3), rng.uniform(0.1, 3), rng.uniform(0, 1)
        Td = R @ np.diag([c, c, e]) @ R.T
        Tp = R @ np.diag([2 - sf, 2 - sf, 1 + sf]) @ R.T
        for k in range(3):
            assert abs(rotated(L, a, [c, c, e], k) - Td[k].sum()) < 1e-13 * np.abs(Td).sum()
            assert abs(L.wco_rotated_pore_coeff(vec(a), sf, k) - Tp[k].sum()) < 1e-13 * 9
    # an orientation along y: the y row is the along-orientation value
    assert rotated(L, [0, 3, 0], [2.0, 2.0, 5.0], 1) == 5.0
    assert rotated(L, [0, 3, 0], [2.0, 2.0, 5.0], 0) == 2.0
    assert rotated(L, [0, 0, 0], [2.0, 2.0, 5.0], 2) == 0.0


def test_capillary_pressure(oracle):
    L = oracle.lib()
    p = oracle.mix_params(surf_tension_coeff=0.07, pore_radius=1e-4,
                          rest_contact_angle=np.pi / 3)
    cap = lambda psi, prm=p: L.wco_capillary_pressure(C.byref(prm), psi)  # noqa: E731
    assert cap(0.0) == 0.0 and cap(1.0) == 0.0
    assert abs(cap(0.5) - 0.07 * 0.5 / 1e-4) < 1e-12 * 350.0
    vals = [cap(x) for x in np.linspace(0.05, 0.95, 19)]
    assert all(b > a for a, b in zip(vals, vals[1:]))
    q = oracle.mix_params(pore_radius=0.0)
    assert cap(0.5, q) == 0.0
    # the parser's derived pore radius (yarn diameter and volume fraction)
    r = oracle.mix_params(yarn_diameter=1e-4, rest_volume_fraction=0.2).pore_radius
    assert abs(r - 1e-4) < 1e-18


def drag(L, prm, psi, sat=1.0, dv=0.0, ori=(0.0, 1.0, 0.0), sf=0.0, k=0, rho=1000.0):
    return L.wco_drag_coeff(C.byref(prm), rho, psi, sat, dv, vec(ori), sf, k)


def drummond_tahir(psi, radius):
    """Permeabilities of a fibre array of volume fraction psi and fibre
    radius a (Drummond & Tahir 1984): along the fibres
    a^2 / (4 psi) (-ln psi - 1.476 + 2 psi - psi^2 / 2), across them
    a^2 / (8 psi) (-ln psi - 1.476 + 2 psi - 1.774 psi^2 + 4.078 psi^3)."""
    lp = -np.log(psi) - 1.476 + 2 * psi
    k_par = radius ** 2 / (4 * psi) * (lp - psi ** 2 / 2)
    k_perp = radius ** 2 / (8 * psi) * (lp - 1.774 * psi ** 2 + 4.078 * psi ** 3)
    return k_par, k_perp


def test_drag_coefficient(oracle):
    L = oracle.lib()
    lin = oracle.mix_params(use_nonlinear_drag=0)
    non = oracle.mix_params()
    off = oracle.mix_params(use_drag=0)
    assert drag(L, off, 0.3) == 0.0
    assert drag(L, lin, 0.0) == 0.0 and drag(L, lin, 0.3, sat=0.0) == 0.0
    assert drag(L, lin, 0.3, ori=(0.0, 0.0, 0.0)) == 0.0
    assert drag(L, lin, 0.3, ori=(1e-11, 0.0, 0.0)) == 0.0  # |o|^2 < 1e-20
    mu, di = lin.viscosity, lin.yarn_diameter
    for psi in (0.05, 0.1, 0.2, 0.3, 0.45):
        kp, kt = drummond_tahir(psi, di / 2)
        assert kp > kt > 0  # flow along the fibres meets less resistance
        # orientation along y and shape factor 0: across the orientation the
        # coefficient is mu / k_along, along it mu / k_across
        assert abs(drag(L, lin, psi, k=0) - mu / kp) < 1e-12 * mu / kp
        assert abs(drag(L, lin, psi, k=2) - mu / kp) < 1e-12 * mu / kp
        assert abs(drag(L, lin, psi, k=1) - mu / kt) < 1e-12 * mu / kt
        # shape factor 1 swaps the two
        assert abs(drag(L, lin, psi, k=0, sf=1.0) - mu / kt) < 1e-12 * mu / kt
        # the saturation only gates the drag, and the linear drag has no dv
        assert drag(L, lin, psi, sat=0.2, dv=3.0) == drag(L, lin, psi)
        # the nonlinear part vanishes at dv = 0 and is homogeneous of degree
        # yazdchi_power in |dv|
        assert drag(L, non, psi) == drag(L, lin, psi)
        e1 = drag(L, non, psi, dv=0.01) - drag(L, lin, psi)
        e2 = drag(L, non, psi, dv=-0.02) - drag(L, lin, psi)
        assert e1 > 0 and abs(e2 / e1 - 2 ** non.yazdchi_power) < 1e-9


# --------------------------------------------------------------- scene checks

def bspline(r):
    a = np.abs(r)
    return np.where(a < 0.5, 0.75 - a * a, np.where(a < 1.5, 0.5 * (1.5 - a) ** 2, 0.0))


def dense_blocked(oracle, g):
    nx = [g.nx[a] for a in range(3)]
    I, J, K = np.meshgrid(*(np.arange(m) for m in nx), indexing="ij")
    return oracle.node_index_array(g, I, J, K)


def brute_scatter(oracle, g, X, vals, off):
    """sum_p val_p N((x_n - x_p)/dx) over every node of the lattice with
    offsets `off` (in cells), as dense (nx, ny, nz) arrays."""
    dx = g.dx
    W = [bspline(((np.arange(g.nx[a]) - oracle.GHOST + off[a]) * dx - X[a][:, None]) / dx)
         for a in range(3)]
    out = []
    for v in vals:
        out.append(np.einsum("pi,pj,pk->ijk", v[:, None] * W[0], W[1], W[2], optimize=True))
    return out


def make_scene(oracle, n, liquid_lo, liquid_hi, solid, prm=None, sigma=0.0, seed=1,
               gravity=(0.0, -9.81, 0.0), L=None):
    g = oracle.grid((n, n, n), 1.0 / n)
    x = oracle.seed_block(g, liquid_lo, liquid_hi, seed)
    v = oracle.seed_velocity(g, seed, sigma, x) if sigma > 0 else np.zeros_like(x)
    s = oracle.OracleScene((n, n, n), 1.0 / n, x, v, gravity=gravity, L=L)
    if solid is not None:
        s.set_elasto(*mixscene.args(solid), prm=prm)
    return s


wet_sheet = mixscene.wet_sheet


def test_solid_p2g_matches_brute_force(oracle):
    n = 12
    rng = np.random.default_rng(3)
    ne = 150
    solid = dict(x=rng.uniform(0.2, 0.8, (3, ne)), m=rng.uniform(1e-4, 1e-3, ne),
                 fm=rng.uniform(0, 1e-4, ne), fvol=rng.uniform(0, 1e-7, ne),
                 rest_vol=rng.uniform(1e-7, 4e-6, ne), rest_vf=rng.uniform(0.3, 0.7, ne),
                 ori=rng.standard_normal((3, ne)), sf=rng.uniform(0, 1, ne))
    assert check_solid_p2g(oracle, oracle.lib(), n, solid) < 1e-12


def check_solid_p2g(oracle, L, n, solid):
    s = make_scene(oracle, n, (0.1, 0.1, 0.1), (0.5, 0.4, 0.5), solid, L=L)
    s.activate()
    s.update_liquid_phi()
    s.update_solid_phi()
    s.update_solid_weights()
    s.p2g()
    g = s.grid
    idx = dense_blocked(oracle, g)
    act = s.field("act")[idx >> 6].astype(bool)
    dV = g.dx ** 3
    X = solid["x"]
    vsol_p = solid["rest_vol"] * solid["rest_vf"]
    worst = 0.0

    def cmp(name, got, want, scale=None):
        nonlocal worst
        sc = scale if scale is not None else max(np.abs(want).max(), 1e-300)
        if name == "smass":  # every node the solid reaches is in an active bucket
            assert np.all(want[~act] == 0.0)
        worst = max(worst, float(np.abs(got[act] - want[act]).max() / sc))

    for d in range(3):
        off = [0.5, 0.5, 0.5]
        off[d] = 0.0
        ms, vs, vfe, sfw, rw, o0, o1, o2 = brute_scatter(
            oracle, g, X, [solid["m"] + solid["fm"], vsol_p, solid["fvol"], solid["sf"],
                           np.ones(X.shape[1]), *solid["ori"]], off)
        cmp("smass", s.mix_field("smass", d)[idx], ms)
        cmp("svol", s.mix_field("svol", d)[idx], vs + vfe)
        cmp("psi", s.mix_field("psi", d)[idx], np.clip(vs / dV, 0, 1), 1.0)
        vl = s.field("vol", d)[idx]
        cmp("sat", s.mix_field("sat", d)[idx],
            np.clip((vl + vfe) / np.maximum(1e-20, dV - vs), 0, 1), 1.0)
        cmp("sf", s.mix_field("sf", d)[idx], np.where(rw > 1e-20, sfw / np.where(rw > 1e-20, rw, 1), sfw), 1.0)
        o = np.stack([o0, o1, o2])
        lo = np.sqrt((o * o).sum(axis=0))
        o = np.where(lo > 1e-20, o / np.where(lo > 1e-20, lo, 1.0), o)
        for c in range(3):
            cmp("ori", s.mix_field("ori", 3 * d + c)[idx], o[c], 1.0)
    # cell centres: solid and liquid particles together
    xl = s.particles()[0]
    pvol = s.st.pvol
    vs, vfe = brute_scatter(oracle, g, np.concatenate([X, xl], axis=1),
                            [np.concatenate([vsol_p, np.zeros(xl.shape[1])]),
                             np.concatenate([solid["fvol"], np.full(xl.shape[1], pvol)])],
                            [0.5, 0.5, 0.5])
    psi_p = np.clip(vs / dV, 0, 1)
    sat_p = np.clip(vfe / np.maximum(1e-20, dV - vs), 0, 1)
    cmp("psi_p", s.mix_field("psi_p")[idx], psi_p, 1.0)
    cmp("sat_p", s.mix_field("sat_p")[idx], sat_p, 1.0)
    prm = s._mixprm
    cap = np.vectorize(lambda q: L.wco_capillary_pressure(C.byref(prm), q))(psi_p)
    cmp("pp", s.mix_field("pp")[idx], cap * (1.0 - sat_p))
    return worst


def test_far_solid_leaves_the_liquid_substep_bitwise(oracle):
    n = 16
    solid = mixscene.sheet(17, 0.85, ori=None, sf=-1)
    a = make_scene(oracle, n, (0, 0, 0), (0.5, 0.5, 0.5), None, sigma=0.5)
    b = make_scene(oracle, n, (0, 0, 0), (0.5, 0.5, 0.5), solid, sigma=0.5)
    for _ in range(3):
        assert a.substep(DT, 1e-10, 1000) and b.substep(DT, 1e-10, 1000)
    xa, va, Ca = a.particles()
    xb, vb, Cb = b.particles()
    assert np.array_equal(xa, xb) and np.array_equal(va, vb) and np.array_equal(Ca, Cb)
    # the sheet was seen: its faces carry the solid
    assert b.mix_field("psi", 1).max() > 0.01


def cell_matrix(s):
    """The pressure matrix and right-hand side keyed by cell node index."""
    rs, col, val, b, _ = s.system()
    dof = s.field("dof")
    cells = np.flatnonzero(dof >= 0)
    cell_of = np.empty(len(cells), dtype=np.int64)
    cell_of[dof[cells]] = cells
    A = sp.csr_matrix((val, col, rs), shape=(len(b), len(b))).tocoo()
    return {(int(cell_of[r]), int(cell_of[c])): v for r, c, v in zip(A.row, A.col, A.data)}, \
        dict(zip(cell_of.tolist(), b.tolist()))


def check_uniform_psi(oracle, L):
    """Drag and surface tension off, psi = 0.3 on every face near the liquid:
    returns (matrix error, rhs error, position error in dx)."""
    n, psi = 32, 0.3
    dx = 1.0 / n
    lat = mixscene.lattice(20, dx, psi)
    lat["x"] += 6 * dx  # cells 6..25: the stencils stay clear of the ghost buckets
    prm = oracle.mix_params(use_drag=0, surf_tension_coeff=0.0)
    lo, hi = (0.375,) * 3, (0.625,) * 3
    a = make_scene(oracle, n, lo, hi, None, sigma=0.3, L=L)
    b = make_scene(oracle, n, lo, hi, lat, prm=prm, sigma=0.3, L=L)
    a.pre_solve(DT)
    b.pre_solve(DT)
    Aa, ba = cell_matrix(a)
    Ab, bb = cell_matrix(b)
    assert Aa.keys() == Ab.keys() and ba.keys() == bb.keys()
    sa = max(abs(v) for v in Aa.values())
    ea = max(abs(Ab[k] - (1 - psi) * v) for k, v in Aa.items()) / sa
    sb = max(abs(v) for v in ba.values())
    eb = max(abs(bb[k] - (1 - psi) * v) for k, v in ba.items()) / sb
    a2 = make_scene(oracle, n, lo, hi, None, sigma=0.3, L=L)
    b2 = make_scene(oracle, n, lo, hi, lat, prm=prm, sigma=0.3, L=L)
    for _ in range(2):
        assert a2.substep(DT, 1e-13, 2000) and b2.substep(DT, 1e-13, 2000)
    ex = float(np.abs(a2.particles()[0] - b2.particles()[0]).max()) / dx
    return ea, eb, ex


def test_uniform_psi_without_drag_scales_the_system(oracle):
    ea, eb, ex = check_uniform_psi(oracle, oracle.lib())
    print(f"A vs (1-psi) A_liquid {ea:.2e}, b {eb:.2e}, positions {ex:.2e} dx")
    assert ea < 1e-14 and eb < 1e-14 and ex < 1e-11


def projected(oracle, L, solid=None, prm=None, tol=1e-12):
    s = make_scene(oracle, 16, (0.2, 0.32, 0.2), (0.8, 0.6, 0.8),
                   wet_sheet() if solid is None else solid, prm=prm, sigma=0.5, L=L)
    s.pre_solve(DT)
    ok = s.solve_pressure(tol, 2000)
    assert ok
    s.apply_pressure(DT)
    return s


def check_divergence(oracle, s):
    """max over DOF cells of |div((1 - psi) w u_new)| relative to max |b|."""
    g = s.grid
    idx = dense_blocked(oracle, g)
    dof = s.field("dof")[idx]
    flux = []
    for d in range(3):
        f = (1.0 - s.mix_field("psi", d)) * s.field("w", d) * s.field("unew", d)
        flux.append(f[idx] / g.dx)
    div = np.zeros(dof.shape)
    div[:-1, :, :] += flux[0][1:, :, :]
    div -= flux[0]
    div[:, :-1, :] += flux[1][:, 1:, :]
    div -= flux[1]
    div[:, :, :-1] += flux[2][:, :, 1:]
    div -= flux[2]
    b = s.system()[3]
    return float(np.abs(div[dof >= 0]).max() / np.abs(b).max())


def test_projected_mixture_flux_is_divergence_free(oracle):
    s = projected(oracle, oracle.lib())
    e = check_divergence(oracle, s)
    print(f"max |div (1-psi) w u| / max |b| = {e:.2e}")
    assert e < 1e-10


def check_two_phase(oracle, s):
    """max relative difference between u_new and the liquid row of the
    two-phase implicit drag system, over faces between two liquid DOF cells
    that carry both phases and drag."""
    g = s.grid
    idx = dense_blocked(oracle, g)
    dof = s.field("dof")[idx]
    lphi = s.field("lphi")[idx]
    p = s.field("pressure")[idx]
    worst, count = 0.0, 0
    for d in range(3):
        f = {k: v[idx] for k, v in (("w", s.field("w", d)), ("Mf", s.field("mass", d)),
                                    ("Vf", s.field("vol", d)), ("rhs", s.field("rhs", d)),
                                    ("u", s.field("unew", d)), ("Ms", s.mix_field("smass", d)),
                                    ("Vs", s.mix_field("svol", d)), ("h", s.mix_field("hdv", d)))}
        sl = [slice(None)] * 3
        sl[d] = slice(1, None)
        sr = tuple(sl)
        sl[d] = slice(0, -1)
        sL = tuple(sl)
        sel = ((dof[sr] >= 0) & (dof[sL] >= 0) & (lphi[sr] < 0) & (lphi[sL] < 0) &
               (f["w"][sr] > 0) & (f["Mf"][sr] > 1e-20) & (f["Ms"][sr] > 0) & (f["h"][sr] > 0))
        gp = (p[sr] - p[sL])[sel] / g.dx
        Mf, Ms, h = f["Mf"][sr][sel], f["Ms"][sr][sel], f["h"][sr][sel]
        r1 = f["rhs"][sr][sel] - DT * f["Vf"][sr][sel] * gp
        r2 = 0.0 - DT * f["Vs"][sr][sel] * gp  # the solid at rest: M_s u_s = 0
        for t in range(Mf.size):
            K = np.array([[Mf[t] + h[t], -h[t]], [-h[t], Ms[t] + h[t]]])
            uf = np.linalg.solve(K, [r1[t], r2[t]])[0]
            u = f["u"][sr][sel][t]
            worst = max(worst, abs(u - uf) / max(abs(uf), 1e-3))
        count += Mf.size
    return worst, count


def test_new_velocity_is_the_two_phase_drag_solve(oracle):
    s = projected(oracle, oracle.lib())
    worst, count = check_two_phase(oracle, s)
    hq = [s.mix_field("Q", d) for d in range(3)]
    print(f"{count} faces with both phases and drag: max rel diff {worst:.2e}; "
          f"Q in [{min(q[q > 0].min() for q in hq):.3f}, {max(q.max() for q in hq):.3f}]")
    assert count > 100 and worst < 1e-12


def test_mixture_system_is_a_symmetric_m_matrix(oracle):
    s = projected(oracle, oracle.lib())
    rs, col, val, b, _ = s.system()
    A = sp.csr_matrix((val, col, rs), shape=(len(b), len(b)))
    assert (A - A.T).count_nonzero() == 0
    D = A.diagonal()
    off = A - sp.diags(D)
    assert off.max() <= 0.0 and np.all(D >= 0)
    rowsum = np.asarray(abs(off).sum(axis=1)).ravel()
    assert np.all(D >= rowsum * (1 - 1e-12))
    # the solid enters: rows touching it differ from a liquid-only system
    assert s.mix_field("psi", 1).max() > 0.05


def wicking(oracle, L, sigma):
    prm = oracle.mix_params(use_drag=0, surf_tension_coeff=sigma)
    s = make_scene(oracle, 16, (0.2, 0.33, 0.2), (0.8, 0.5, 0.8), mixscene.sheet(33, 0.3),
                   prm=prm, gravity=(0.0, 0.0, 0.0), L=L)
    s.activate()
    s.update_liquid_phi()
    s.update_solid_phi()
    s.update_solid_weights()
    s.p2g()
    s.step_velocity(DT)
    return s.field("rhs", 1)


def test_pore_pressure_draws_liquid_into_the_dry_cloth(oracle):
    L = oracle.lib()
    r0 = wicking(oracle, L, 0.0)
    r1 = wicking(oracle, L, 0.0728)
    r2 = wicking(oracle, L, 2 * 0.0728)
    assert np.all(r0 == 0.0)  # at rest, no gravity, no surface tension
    assert r1.sum() < 0 and r1.min() < 0  # toward the sheet below the liquid
    assert np.abs(r2 - 2 * r1).max() <= 1e-13 * np.abs(r1).max()


# ----------------------------------------------------------------- plants

@pytest.fixture(scope="module")
def planted(oracle, tmp_path_factory):
    d = str(tmp_path_factory.mktemp("mixplant"))
    return lambda k: oracle.load_variant(oracle.build_variant(d, k))


def test_planted_mixture_mistakes_fail(oracle, planted):
    # 20: assembly without (1 - psi)
    ea, eb, _ = check_uniform_psi(oracle, planted(20))
    assert ea > 1e-3
    assert check_divergence(oracle, projected(oracle, planted(20))) > 1e-6
    # 21: pressure application without V_s Q
    assert check_divergence(oracle, projected(oracle, planted(21))) > 1e-6
    assert check_two_phase(oracle, projected(oracle, planted(21)))[0] > 1e-6
    # 22: the pore pressure gradient's sign
    assert wicking(oracle, planted(22), 0.0728).sum() > 0
    # 23: Q from the liquid mass
    assert check_two_phase(oracle, projected(oracle, planted(23)))[0] > 1e-6
    # 24: the solid's P2G on the cell-centre stencil
    rng = np.random.default_rng(3)
    ne = 60
    solid = dict(x=rng.uniform(0.2, 0.8, (3, ne)), m=rng.uniform(1e-4, 1e-3, ne),
                 fm=np.zeros(ne), fvol=np.zeros(ne), rest_vol=rng.uniform(1e-7, 4e-6, ne),
                 rest_vf=rng.uniform(0.3, 0.7, ne), ori=rng.standard_normal((3, ne)),
                 sf=rng.uniform(0, 1, ne))
    assert check_solid_p2g(oracle, planted(24), 12, solid) > 1e-3
