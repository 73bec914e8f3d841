"""The sub-tile MIC(0) sweeps' index arithmetic, restated on the CPU (no GPU).

k_sweep_sub (libwetcloth_amd/csrc/wc_sweep.hip) runs each 8x8 quarter of a
16x16 tile of the skewed box layout as one unit, one wave of compute lanes,
fed by a loader wave (records through an LDS ring) and a face wave (the
upstream values beside the unit's low x and y faces, in the sweep's frame).
This module restates every slot index those waves form and checks, over many
box shapes and both passes:
  - every address is inside the box arrays, and the form without the face
    guard is not (so the check is sharp);
  - the compute lanes write every slot exactly once per pass;
  - the record the loader puts at ring position p of step w is the one the
    compute lane that reads position p stores at step w;
  - each needed face read is the slot of the upstream neighbour's cell;
  - the halo cells the face wave writes in the exchange plane are the cells
    the operand words point at, and never a cell the compute lanes write;
  - a unit's upstream units hold earlier tickets, and the z . r partial
    indices cover the units once.
"""
import numpy as np
import pytest

KT = 16
THREADS = KT * KT
KS = 8
KNB = 4
KLA = 9
ROW = KT + 1  # exchange-plane row stride


def plan(B0, B1, Bz):
    nta, ntb = B0 // KT, B1 // KT
    nsteps = Bz + 2 * (KT - 1)
    ncell = nta * ntb * nsteps * THREADS
    kdx = -nsteps * THREADS + (KT - 1) * THREADS + (KT - 1)
    kdy = -nta * nsteps * THREADS + (KT - 1) * THREADS + (KT - 1) * KT
    return dict(nta=nta, ntb=ntb, nsteps=nsteps, ncell=ncell, kdx=kdx, kdy=kdy, Bz=Bz,
                nb=(nsteps + KNB - 1) // KNB)


def skew_slot(P, a, b, li, lj, w):
    return ((b * P["nta"] + a) * P["nsteps"] + w) * THREADS + lj * KT + li


def order_sub(P):
    nua, nub = 2 * P["nta"], 2 * P["ntb"]
    out = []
    for s in range(nua + nub - 1):
        for A in range(nua):
            B = s - A
            if 0 <= B < nub:
                out.append((A, B))
    return out


def decode_unit(P, A, B, bwd):
    sa, sb = A & 1, B & 1
    a = P["nta"] - 1 - (A >> 1) if bwd else A >> 1
    b = P["ntb"] - 1 - (B >> 1) if bwd else B >> 1
    psa, psb = (1 - sa, 1 - sb) if bwd else (sa, sb)
    return sa, sb, a, b, psa, psb


def base_of(P, a, b, li, lj, bwd):
    lip = KT - 1 - li if bwd else li
    ljp = KT - 1 - lj if bwd else lj
    return skew_slot(P, a, b, lip, ljp, P["nsteps"] - 1 if bwd else 0)


def loader_slot(P, a, b, psa, psb, w, p, bwd):
    wc = np.minimum(w, P["nsteps"] - 1)
    wp = P["nsteps"] - 1 - wc if bwd else wc
    return skew_slot(P, a, b, 0, 0, 0) + wp * THREADS + (KS * psb + (p >> 3)) * KT + KS * psa + (p & 7)


def unit_faces(P, A, B, bwd, guarded=True):
    """(x, y) face reads of unit (A, B) for every step the face wave fetches:
    per face, arrays (address, needed, step, m)."""
    sa, sb, a, b, psa, psb = decode_unit(P, A, B, bwd)
    sgn = -THREADS if bwd else THREADS
    hx, hy = A > 0, B > 0
    if not guarded:
        hx = hy = True
    dxs = (THREADS + 1 if bwd else -(THREADS + 1)) if sa else (-P["kdx"] if bwd else P["kdx"])
    dys = (THREADS + KT if bwd else -(THREADS + KT)) if sb else (-P["kdy"] if bwd else P["kdy"])
    xd = dxs if hx else 0
    yd = dys if hy else 0
    nb = P["nb"]
    w = np.arange(nb * KNB)          # every step of every fetched batch
    out = {}
    for face in ("x", "y"):
        addrs, needs, steps, ms = [], [], [], []
        for m in range(KS):
            if face == "x":
                li, lj, h, d = KS * sa, KS * sb + m, hx, xd
            else:
                li, lj, h, d = KS * sa + m, KS * sb, hy, yd
            base = base_of(P, a, b, li, lj, bwd)
            off = li + lj
            kc = np.clip(w - off, 0, P["Bz"] - 1)
            addrs.append(base + sgn * (kc + off) + d)
            kk = w - off
            needs.append(h & (w < P["nsteps"]) & (kk >= 0) & (kk < P["Bz"]))
            steps.append(w)
            ms.append(np.full_like(w, m))
        out[face] = tuple(np.concatenate(v) for v in (addrs, needs, steps, ms))
    return out


def cell_of(P, t):
    """Physical (px, py, kk) of slot t."""
    lane = t % THREADS
    q = t // THREADS
    w = q % P["nsteps"]
    tile = q // P["nsteps"]
    a, b = tile % P["nta"], tile // P["nta"]
    li, lj = lane % KT, lane // KT
    return a * KT + li, b * KT + lj, w - li - lj


SHAPES = [(B0, B1, Bz) for B0 in (16, 32, 48) for B1 in (16, 32) for Bz in (1, 5, 17, 40)]


@pytest.mark.parametrize("bwd", [False, True])
def test_every_address_in_range_and_every_slot_written_once(bwd):
    for B0, B1, Bz in SHAPES:
        P = plan(B0, B1, Bz)
        n = P["ncell"]
        sgn = -THREADS if bwd else THREADS
        written = np.zeros(n, dtype=np.int64)
        steps = np.arange(P["nsteps"])
        for A, B in order_sub(P):
            sa, sb, a, b, psa, psb = decode_unit(P, A, B, bwd)
            # compute lanes: one store per step
            for lane in range(64):
                li, lj = KS * sa + (lane & 7), KS * sb + (lane >> 3)
                st = base_of(P, a, b, li, lj, bwd) + sgn * steps
                assert st.min() >= 0 and st.max() < n, (B0, B1, Bz, bwd)
                np.add.at(written, st, 1)
            # loader: records, rhs, words, r and the clears, up to the last
            # batch issued ahead (clamped to the last step)
            w = np.arange((P["nb"] + KLA) * KNB)
            for p in range(64):
                ls = loader_slot(P, a, b, psa, psb, w, p, bwd)
                assert ls.min() >= 0 and ls.max() < n
            # face wave
            for face, (addr, need, _, _) in unit_faces(P, A, B, bwd).items():
                assert addr.min() >= 0 and addr.max() < n, (B0, B1, Bz, bwd, A, B, face)
        assert np.all(written == 1), (B0, B1, Bz, bwd, np.unique(written))


@pytest.mark.parametrize("bwd", [False, True])
def test_unguarded_face_address_leaves_the_array(bwd):
    """Adding the face offset on the units of the first row or column of the
    sweep (no upstream unit) reads outside the box on every shape."""
    for B0, B1, Bz in SHAPES:
        P = plan(B0, B1, Bz)
        lo, hi = 0, 0
        for A, B in order_sub(P):
            for addr, *_ in unit_faces(P, A, B, bwd, guarded=False).values():
                lo, hi = min(lo, addr.min()), max(hi, addr.max())
        assert lo < 0 or hi >= P["ncell"], (B0, B1, Bz, bwd)


@pytest.mark.parametrize("bwd", [False, True])
def test_ring_position_matches_compute_lane(bwd):
    """The loader fills ring position p of step w with the slot the compute
    lane reading position p (lane p forward, 63 - p backward) stores at w."""
    P = plan(32, 48, 11)
    sgn = -THREADS if bwd else THREADS
    w = np.arange(P["nsteps"])
    for A, B in order_sub(P):
        sa, sb, a, b, psa, psb = decode_unit(P, A, B, bwd)
        for lane in range(64):
            li, lj = KS * sa + (lane & 7), KS * sb + (lane >> 3)
            pos = 63 - lane if bwd else lane
            st = base_of(P, a, b, li, lj, bwd) + sgn * w
            assert np.array_equal(loader_slot(P, a, b, psa, psb, w, pos, bwd), st)


@pytest.mark.parametrize("bwd", [False, True])
def test_face_reads_hit_the_upstream_cell(bwd):
    """A needed face value of step w is the cell beside the lane's cell of
    step w, in the sweep's upstream direction (-x / -y forward, +x / +y
    backward): across the quarter boundary inside a tile and across tiles."""
    P = plan(48, 32, 9)
    sgn = -THREADS if bwd else THREADS
    step = 1 if not bwd else -1
    for A, B in order_sub(P):
        sa, sb, a, b, psa, psb = decode_unit(P, A, B, bwd)
        faces = unit_faces(P, A, B, bwd)
        for face, (addr, need, w, m) in faces.items():
            if not need.any():
                continue
            mm = m[need]
            li = np.full_like(mm, KS * sa) if face == "x" else KS * sa + mm
            lj = KS * sb + mm if face == "x" else np.full_like(mm, KS * sb)
            own = base_of(P, a, b, li, lj, bwd)
            own = own + sgn * w[need]
            ox, oy, ok = cell_of(P, own)
            ux, uy, uk = cell_of(P, addr[need])
            assert np.all((ok >= 0) & (ok < P["Bz"]))
            if face == "x":
                assert np.array_equal(ux, ox - step) and np.array_equal(uy, oy)
            else:
                assert np.array_equal(uy, oy - step) and np.array_equal(ux, ox)
            assert np.array_equal(uk, ok)


def operand_offsets(li, lj):
    """The operand words' offsets in the exchange plane: -1 (x), -ROW (y),
    0 (own previous value) and -pos (the +0 at [0][0])."""
    pos = (lj + 1) * ROW + li + 1
    return pos, {"x": pos - 1, "y": pos - ROW, "z": pos, "none": 0}


def test_plane_halos_are_the_operands():
    """For every quarter, the x operand of a lane on the unit's low x face
    and the y operand of a lane on its low y face are the halo cells the face
    wave writes; no halo cell is a cell a compute lane writes, and no one
    writes [0][0]."""
    for sa in (0, 1):
        for sb in (0, 1):
            own = set()
            x_halo, y_halo = set(), set()
            for m in range(KS):
                x_halo.add((KS * sb + m + 1) * ROW + KS * sa)
                y_halo.add(KS * sb * ROW + KS * sa + m + 1)
            for lane in range(64):
                li, lj = KS * sa + (lane & 7), KS * sb + (lane >> 3)
                pos, ops = operand_offsets(li, lj)
                own.add(pos)
                if li == KS * sa:
                    assert ops["x"] in x_halo
                else:
                    assert ops["x"] in {(lj + 1) * ROW + li}  # the lane to the left
                if lj == KS * sb:
                    assert ops["y"] in y_halo
            assert not (own & x_halo) and not (own & y_halo)
            assert 0 not in own | x_halo | y_halo


def test_tickets_respect_dependencies_and_partials_cover_units():
    for B0, B1 in ((16, 16), (32, 48), (64, 32)):
        P = plan(B0, B1, 4)
        order = order_sub(P)
        ticket = {u: t for t, u in enumerate(order)}
        assert len(order) == 4 * P["nta"] * P["ntb"]
        for (A, B), t in ticket.items():
            if A > 0:
                assert ticket[(A - 1, B)] < t
            if B > 0:
                assert ticket[(A, B - 1)] < t
        for bwd in (False, True):
            idx = []
            for A, B in order:
                sa, sb, a, b, psa, psb = decode_unit(P, A, B, bwd)
                idx.append((2 * b + psb) * 2 * P["nta"] + 2 * a + psa)
            assert sorted(idx) == list(range(len(order)))
