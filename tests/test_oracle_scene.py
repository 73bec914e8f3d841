"""Pin the oracle's scene phases with independent restatements and invariants.

Each check is chosen so that a plausible mistake fails it: conservation and
affine reproduction catch a wrong weight, offset or transposed affine matrix
in P2G/G2P; the numpy restatements of activation, liquid phi and the
extrapolation catch index and default-value errors; the textbook Laplacian
and the divergence identity (div u_new == b - A p, the reference's own
check_divergence diagnostic, LinearizedImplicitEuler.cpp:1150-1175) catch
sign and coefficient errors in the assembly and the pressure application.
"""
import numpy as np
import pytest
import scipy.ndimage as ndi
import scipy.sparse as sp

GHOST = 8
DT = 1.0 / 480


def scene(oracle, n, block_hi=(0.4, 0.6, 1.0), seed=0, sigma=0.0, **kw):
    g, x, v = oracle.dam_break(n, block_hi=block_hi, seed=seed, sigma=sigma)
    return oracle.OracleScene((n, n, n), 1.0 / n, x, v, **kw), x, v


def grid_coords(g):
    nn = g.nnode
    idx = np.arange(nn)
    b = idx >> 6
    l = idx & 63
    bi = b % g.nb[0]
    bj = (b // g.nb[0]) % g.nb[1]
    bk = b // (g.nb[0] * g.nb[1])
    return bi * 4 + (l & 3), bj * 4 + ((l >> 2) & 3), bk * 4 + (l >> 4)


# ------------------------------------------------------------------ seeding

def test_seeding_fills_subcells(oracle):
    n = 16
    g, x, v = oracle.dam_break(n, block_hi=(0.5, 0.5, 0.5))
    assert x.shape[1] == 8 * 8 * 8 * 8
    dx = 1.0 / n
    cell = np.floor(x / dx).astype(int)
    sub = np.floor((x / dx - cell) * 2).astype(int)
    key = ((cell[2] * n + cell[1]) * n + cell[0]) * 8 + sub[2] * 4 + sub[1] * 2 + sub[0]
    assert len(np.unique(key)) == x.shape[1]  # one particle per 2x2x2 sub-cell
    g2, x2, _ = oracle.dam_break(n, block_hi=(0.5, 0.5, 0.5))
    assert np.array_equal(x, x2)


def test_seeded_velocity_statistics(oracle):
    n = 32
    g, x, v = oracle.dam_break(n, block_hi=(1.0, 1.0, 1.0), seed=7, sigma=0.5)
    # the filtered field has std sigma by construction; B2 interpolation
    # smooths it further, so the particle std is below sigma but of its order
    s = v.std(axis=1)
    assert np.all(s > 0.15) and np.all(s < 0.5)
    assert np.abs(v.mean(axis=1)).max() < 0.05


# ------------------------------------------------------------ P2G / G2P

def test_p2g_conserves_mass_and_momentum(oracle):
    n = 16
    s, x, v = scene(oracle, n, sigma=0.4, seed=2)
    rng = np.random.default_rng(0)
    Cm = rng.standard_normal((9, x.shape[1])) * 3.0
    s = oracle.OracleScene((n, n, n), 1.0 / n, x, v, Cm)
    s.activate()
    s.p2g()
    m = s.st.pmass
    for d in range(3):
        M = s.field("mass", d)
        assert abs(M.sum() - m * x.shape[1]) < 1e-12 * m * x.shape[1]
        P = (M * s.field("vel", d)).sum()
        # sum_n w (x_n - x_p) = 0, so the affine part carries no net momentum
        assert abs(P - m * v[d].sum()) < 1e-11 * m * np.abs(v[d]).sum()


def test_apic_affine_reproduction(oracle):
    """u(x) = a + B x survives P2G exactly at every node and G2P (flip = 1)
    returns v = a + B x_p and C = B."""
    n = 16
    g, x, _ = oracle.dam_break(n, block_hi=(0.75, 0.75, 0.75))
    a = np.array([0.3, -0.2, 0.1])
    B = np.array([[0.5, -1.0, 0.25], [2.0, 0.1, -0.3], [-0.7, 0.4, 1.1]])
    v = a[:, None] + B @ x
    Cm = np.repeat(B.reshape(9, 1), x.shape[1], axis=1)
    s = oracle.OracleScene((n, n, n), 1.0 / n, x, v, Cm, flip=1.0)
    s.set_solid_projection(False)
    s.activate()
    s.p2g()
    i, j, k = grid_coords(s.grid)
    dx = 1.0 / n
    for d in range(3):
        off = [0.5, 0.5, 0.5]
        off[d] = 0.0
        pos = np.stack([(i - GHOST + off[0]) * dx, (j - GHOST + off[1]) * dx,
                        (k - GHOST + off[2]) * dx])
        M = s.field("mass", d)
        has = M > 1e-20
        u = s.field("vel", d)
        exact = a[d] + B[d] @ pos
        assert np.abs(u[has] - exact[has]).max() < 1e-12
    s.copy_vel_to_unew()
    # particles whose whole stencil has mass: interior of the block
    inner = np.all((x > 2.5 * dx) & (x < 0.75 - 2.5 * dx), axis=0)
    s.g2p_advect(0.0)
    xg, vg, Cg = s.particles()
    assert np.abs(vg[:, inner] - v[:, inner]).max() < 1e-12
    assert np.abs(Cg[:, inner] - Cm[:, inner]).max() < 1e-10


# -------------------------------------------------- activation and phi

def expected_activation(g, x):
    nb = np.array(g.nb)
    touched = np.zeros(tuple(nb[::-1]), dtype=bool)  # [bk, bj, bi]
    dx = g.dx
    offs = [(0, .5, .5), (.5, 0, .5), (.5, .5, 0), (0, 0, 0), (.5, .5, .5)]
    for p in range(x.shape[1]):
        for off in offs:
            lo, hi = [], []
            for a in range(3):
                t = x[a, p] / dx + (GHOST - off[a])
                fl = np.floor(t)
                base = int(fl) + (0 if t - fl > 0.5 else -1)
                lo.append(base >> 2)
                hi.append((base + 2) >> 2)
            touched[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = True
    act = ndi.binary_dilation(touched, structure=np.ones((3, 3, 3), bool))
    return act.ravel().astype(np.uint8)


def test_activation_matches_numpy(oracle):
    n = 16
    g, x, _ = oracle.dam_break(n, block_hi=(0.25, 0.3, 0.5))
    s = oracle.OracleScene((n, n, n), 1.0 / n, x)
    s.activate()
    assert np.array_equal(s.field("act"), expected_activation(s.grid, x))


def test_liquid_phi_matches_brute_force(oracle):
    n = 12
    rng = np.random.default_rng(3)
    x = rng.uniform(0.05, 0.95, (3, 40))
    s = oracle.OracleScene((n, n, n), 1.0 / n, x)
    s.activate()
    s.update_liquid_phi()
    g = s.grid
    dx = 1.0 / n
    act = s.field("act")
    i, j, k = grid_coords(g)
    b = ((k >> 2) * g.nb[1] + (j >> 2)) * g.nb[0] + (i >> 2)
    exp = np.full(g.nnode, 3.0 * 4 * dx)
    r0 = max(0.883644 * dx, s.st.radius)
    for p in range(x.shape[1]):
        base = [int(np.floor(x[a, p] / dx + GHOST - 0.5)) for a in range(3)]
        base = [bb + (0 if (x[a, p] / dx + GHOST - 0.5) - np.floor(x[a, p] / dx + GHOST - 0.5) > 0.5 else -1)
                for a, bb in enumerate(base)]
        for c2 in range(3):
            for c1 in range(3):
                for c0 in range(3):
                    ii, jj, kk = base[0] + c0, base[1] + c1, base[2] + c2
                    nidx = oracle.node_index(g, ii, jj, kk)
                    c = np.array([(ii - GHOST + .5) * dx, (jj - GHOST + .5) * dx,
                                  (kk - GHOST + .5) * dx])
                    d = x[:, p] - c
                    phi = np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) - r0
                    exp[nidx] = min(exp[nidx], phi)
    centre = np.stack([(i - GHOST + .5) * dx, (j - GHOST + .5) * dx, (k - GHOST + .5) * dx])
    outside = np.any((centre < 0) | (centre > 1.0), axis=0)
    exp[outside & (act[b] == 1)] = -0.5 * dx
    assert np.array_equal(s.field("lphi"), exp)


def test_wall_face_weights(oracle):
    n = 8
    s, x, _ = scene(oracle, n, block_hi=(1.0, 1.0, 1.0))
    s.activate()
    s.update_solid_phi()
    s.update_solid_weights()
    g = s.grid
    w = s.field("w", 0)
    ni = oracle.node_index
    # interior face, the wall plane x = 0 and the far wall x = 1: open
    assert w[ni(g, GHOST + 3, GHOST + 3, GHOST + 3)] == 1.0
    assert w[ni(g, GHOST + 0, GHOST + 3, GHOST + 3)] == 1.0
    assert w[ni(g, GHOST + n, GHOST + 3, GHOST + 3)] == 1.0
    # one face into the solid: closed; y-face of a ghost cell: closed
    assert w[ni(g, GHOST - 1, GHOST + 3, GHOST + 3)] == 0.0
    assert s.field("w", 1)[ni(g, GHOST - 1, GHOST + 3, GHOST + 3)] == 0.0


# ------------------------------------------------------------ assembly

def test_full_box_is_textbook_laplacian(oracle):
    n = 8
    s, x, _ = scene(oracle, n, block_hi=(1.0, 1.0, 1.0), gravity=(0.0, 0.0, 0.0))
    s.pre_solve(DT)
    rs, col, val, b, _ = s.system()
    nd = len(rs) - 1
    A = sp.csr_matrix((val, col, rs), shape=(nd, nd))
    dof = s.field("dof")
    g = s.grid
    dx = 1.0 / n
    c = DT / (1000.0 * dx * dx)
    # a cell two away from every wall has six liquid neighbours
    me = dof[oracle.node_index(g, GHOST + 3, GHOST + 4, GHOST + 2)]
    row = A[me].toarray().ravel()
    assert abs(row[me] - 6 * c) < 1e-12 * c
    offd = row[np.arange(nd) != me]
    assert np.sum(offd != 0) == 6
    assert np.abs(offd[offd != 0] + c).max() < 1e-12 * c
    assert abs(row.sum()) < 1e-12 * c
    # zero velocity and no gravity: zero right-hand side
    assert np.abs(b).max() == 0.0
    # every interior liquid row sums to zero (pure Neumann-like coupling)
    assert abs(A - A.T).max() == 0.0


def test_projection_divergence_is_pcg_residual(oracle):
    """After projection the weighted face divergence of each DOF cell equals
    b - A p (the PCG residual), and so is below tol * |b|_inf."""
    n = 16
    tol = 1e-8
    s, x, v = scene(oracle, n, sigma=0.5, seed=4)
    s.pre_solve(DT)
    assert s.solve_pressure(tol, 2000)
    s.apply_pressure(DT)
    rs, col, val, b, p = s.system()
    nd = len(rs) - 1
    A = sp.csr_matrix((val, col, rs), shape=(nd, nd))
    r = b - A @ p
    g = s.grid
    dx = 1.0 / n
    i, j, k = grid_coords(g)
    dof = s.field("dof")
    u = [s.field("unew", d) for d in range(3)]
    psi = [s.field("psi_fs", d) for d in range(3)]
    w = [s.field("w", d) for d in range(3)]
    act = s.field("act")
    nx = np.array(g.nx)
    div = np.zeros(nd)
    cells = np.nonzero(dof >= 0)[0]
    for d in range(3):
        for hi, sign in ((0, 1.0), (1, -1.0)):
            q = [i[cells], j[cells], k[cells]]
            q[d] = q[d] + hi
            # a face exists only inside the grid and in an activated bucket
            ok = q[d] < nx[d]
            q[d] = np.minimum(q[d], nx[d] - 1)
            f = oracle.node_index_array(g, *q)
            ok &= act[f >> 6] == 1
            contrib = np.where(ok, sign * (u[d][f] * psi[d][f] * w[d][f]) / dx, 0.0)
            np.add.at(div, dof[cells], contrib)
    assert np.abs(div - r).max() < 1e-9 * np.abs(b).max()
    assert np.abs(div).max() <= 1.001 * tol * np.abs(b).max()


def test_extrapolation_and_constraint_restated(oracle):
    """Two extrapolation layers (Pressure.cpp:1826-1903) and the w == 0
    constraint, recomputed in numpy from the directly projected faces."""
    n = 12
    s, x, v = scene(oracle, n, block_hi=(0.5, 0.5, 0.5), sigma=0.5, seed=5)
    s.pre_solve(DT)
    s.solve_pressure(1e-8, 1000)
    s.apply_pressure(DT)
    g = s.grid
    act = s.field("act")
    lphi = s.field("lphi")
    nx = np.array(g.nx)
    i, j, k = grid_coords(g)
    for d in range(3):
        w = s.field("w", d)
        u = s.field("unew", d)
        # direct faces: w > 0 with a liquid side (Pressure.cpp:1184-1211)
        li = [i.copy(), j.copy(), k.copy()]
        li[d] -= 1
        inr = (li[0] >= 0) & (li[1] >= 0) & (li[2] >= 0)
        lidx = oracle.node_index_array(g, np.maximum(li[0], 0), np.maximum(li[1], 0),
                                       np.maximum(li[2], 0))
        lb = ((np.maximum(li[2], 0) >> 2) * g.nb[1] + (np.maximum(li[1], 0) >> 2)) * g.nb[0] \
            + (np.maximum(li[0], 0) >> 2)
        lex = inr & (act[lb] == 1)
        lphi_l = np.where(lex, lphi[lidx], 0.0)
        me_act = act[np.arange(g.nnode) >> 6] == 1
        direct = me_act & (w > 0) & ((lphi_l < 0) | (lphi < 0))
        valid = direct.copy()
        vals = np.where(direct, u, 0.0)
        for layer in range(2):
            new_valid = valid.copy()
            new_vals = vals.copy()
            s_sum = np.zeros(g.nnode)
            cnt = np.zeros(g.nnode, dtype=int)
            for r in range(6):
                q = [i.copy(), j.copy(), k.copy()]
                q[r // 2] += 1 if r % 2 else -1
                ok = np.all([(q[a] >= 0) & (q[a] < nx[a]) for a in range(3)], axis=0)
                qc = [np.clip(q[a], 0, nx[a] - 1) for a in range(3)]
                m = oracle.node_index_array(g, *qc)
                qb = m >> 6
                ok &= act[qb] == 1
                ok &= valid[m]
                s_sum += np.where(ok, vals[m], 0.0)
                cnt += ok
            fill = me_act & ~valid & (cnt > 0)
            new_vals[fill] = s_sum[fill] / cnt[fill]
            new_valid[fill] = True
            valid, vals = new_valid, new_vals
        vals = np.where(me_act & (w == 0.0), 0.0, vals)
        vals = np.where(me_act, vals, 0.0)
        check = me_act & (valid | (w == 0.0))
        assert np.abs(vals[check] - u[check]).max() < 1e-14 * max(np.abs(u).max(), 1.0)
        assert np.array_equal(valid.astype(np.uint8)[me_act], s.field("valid", d)[me_act])


def test_dam_break_stays_in_box_and_falls(oracle):
    n = 16
    s, x, v = scene(oracle, n)
    for _ in range(10):
        assert s.substep(DT, 1e-8, 1000)
    xs, vs, _ = s.particles()
    assert xs.min() >= -1e-12 and xs.max() <= 1.0 + 1e-12
    # the column accelerates downward at about g early on
    assert -9.81 * 10 * DT * 1.5 < vs[1].mean() < 0.0
