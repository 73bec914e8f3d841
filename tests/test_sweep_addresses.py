"""Every global address the pipelined MIC(0) sweep kernel forms stays inside
its arrays (CPU enumeration; no GPU).

The kernel (k_sweep in libwetcloth_amd/csrc/wc_sweep.hip) stores the box in
the skewed tile layout slot = ((b * ntile_a + a) * nsteps + w) * 256 + l and
reads, per tile, (1) its own records through the loader waves, (2) its own
inputs in the factor pass, clamped into the pencil, and (3) through the face
wave, the upstream tiles' outputs beside its low x and y faces.  A face lane
whose tile has no upstream tile must read an in-range slot (its own), not
the slot one tile before the box: a form that adds the face offset on such
lanes read before the start of the output array and faulted the GPU.

This test restates that index arithmetic over many box shapes and checks
that every address is in [0, nslot) -- and that the unguarded form of the
face address does leave the array, so the check would have caught it.
"""
import numpy as np
import pytest

KT = 16
THREADS = KT * KT
KFB = 2          # face batch (steps)
KCH_FAC = 4      # factor-pass prefetch depth
KCH_APPLY = 8    # sweep prefetch depth (unused by the ring kernels; bounded the same way)


def plan(B0, B1, Bz):
    nta, ntb = B0 // KT, B1 // KT
    nsteps = Bz + 2 * (KT - 1)
    ncell = nta * ntb * nsteps * THREADS
    kdx = -nsteps * THREADS + (KT - 1) * THREADS + (KT - 1)
    kdy = -nta * nsteps * THREADS + (KT - 1) * THREADS + (KT - 1) * KT
    return dict(nta=nta, ntb=ntb, nsteps=nsteps, ncell=ncell, kdx=kdx, kdy=kdy, Bz=Bz)


def skew_slot(P, a, b, li, lj, w):
    return ((b * P["nta"] + a) * P["nsteps"] + w) * THREADS + lj * KT + li


def in_index(P, base, off, kk, sgn):
    kc = np.clip(kk, 0, P["Bz"] - 1)
    return base + sgn * (kc + off)


def addresses(P, bwd, guarded=True):
    """All slot indices one pass of k_sweep reads or writes (as int64)."""
    nta, ntb, nsteps = P["nta"], P["ntb"], P["nsteps"]
    sgn = -THREADS if bwd else THREADS
    out = []
    ws = np.arange(nsteps + 2 * KFB + max(KCH_FAC, KCH_APPLY))
    for ta in range(nta):
        for tb in range(ntb):
            a = nta - 1 - ta if bwd else ta
            b = ntb - 1 - tb if bwd else tb
            has_a = (a + 1 < nta) if bwd else (a > 0)
            has_b = (b + 1 < ntb) if bwd else (b > 0)

            def base_of(li, lj):
                lip = KT - 1 - li if bwd else li
                ljp = KT - 1 - lj if bwd else lj
                return skew_slot(P, a, b, lip, ljp, nsteps - 1 if bwd else 0)

            # compute lanes: inputs (clamped) and one store per step
            for lane in range(THREADS):
                li, lj = lane % KT, lane // KT
                base = base_of(li, lj)
                off = li + lj
                out.append(in_index(P, base, off, ws - off, sgn))
                out.append(base + sgn * np.arange(nsteps))
            # face wave: lanes 0..15 x face (pencil (0, lj)), 16..31 y face
            # (pencil (li, 0)); lanes 32..63 repeat them and publish nothing
            for t in range(64):
                fl = t & 31
                xf = fl < KT
                li, lj = (0, fl) if xf else (fl - KT, 0)
                pub = t < 32
                on = pub and (has_a if xf else has_b)
                if not guarded:
                    on = True  # the faulting form: face offset on every lane
                base = base_of(li, lj)
                off = li + lj
                d = (-P["kdx"] if bwd else P["kdx"]) if xf else (-P["kdy"] if bwd else P["kdy"])
                dstep = d if on else 0
                out.append(in_index(P, base, off, ws - off, sgn) + dstep)
            # loader waves: step blocks of 256 consecutive slots
            wc = np.minimum(ws, nsteps - 1)
            s0 = skew_slot(P, a, b, 0, 0, (nsteps - 1 - wc) if bwd else wc)
            out.append(s0)
            out.append(s0 + THREADS - 1)
    return np.concatenate([np.asarray(o, dtype=np.int64).ravel() for o in out])


SHAPES = [(B0, B1, Bz) for B0 in (16, 32, 48) for B1 in (16, 32, 64) for Bz in (1, 2, 5, 17, 31, 40)]


@pytest.mark.parametrize("bwd", [False, True])
def test_every_sweep_address_in_range(bwd):
    for B0, B1, Bz in SHAPES:
        P = plan(B0, B1, Bz)
        addr = addresses(P, bwd)
        assert addr.min() >= 0, (B0, B1, Bz, bwd, addr.min())
        assert addr.max() < P["ncell"], (B0, B1, Bz, bwd, addr.max(), P["ncell"])


@pytest.mark.parametrize("bwd", [False, True])
def test_unguarded_face_address_leaves_the_array(bwd):
    """The check is sharp: adding the face offset on lanes of a tile with no
    upstream tile reads outside the output array on every shape."""
    for B0, B1, Bz in SHAPES:
        P = plan(B0, B1, Bz)
        addr = addresses(P, bwd, guarded=False)
        assert addr.min() < 0 or addr.max() >= P["ncell"], (B0, B1, Bz, bwd)


def test_face_reads_hit_the_upstream_tile_cell():
    """In range is not enough: a guarded face read of step w must be the
    upstream pencil's cell of the same k (tile a-1, lane (15, lj))."""
    P = plan(48, 32, 9)
    nsteps = P["nsteps"]
    for a in range(1, P["nta"]):
        for b in range(P["ntb"]):
            for lj in range(KT):
                base = skew_slot(P, a, b, 0, lj, 0)
                for kk in range(P["Bz"]):
                    addr = in_index(P, base, lj, kk, THREADS) + P["kdx"]
                    # the upstream pencil (15, lj) computes its cell kk at step kk + 15 + lj
                    assert addr == skew_slot(P, a - 1, b, 15, lj, kk + 15 + lj)
                    assert kk + 15 + lj < nsteps


def decode(P, t):
    lane = t % THREADS
    q = t // THREADS
    w = q % P["nsteps"]
    tile = q // P["nsteps"]
    a, b = tile % P["nta"], tile // P["nta"]
    li, lj = lane % KT, lane // KT
    return a * KT + li, b * KT + lj, w - li - lj


@pytest.mark.parametrize("B0,B1,Bz", [(16, 16, 8), (32, 48, 20), (64, 32, 37)])
def test_spmv_lower_neighbour_slots(B0, B1, Bz):
    """The symmetric SpMV (k_box_spmv_sym) reads A(k_d, me) and x at the slot
    of the lower neighbour k_d of every DOF slot that has one: the previous
    step of the lane one to the left (below, or the same lane for z), or
    across the tile face by the face offset.  Wherever the neighbour cell is
    inside the box that slot must be in range and hold exactly that cell."""
    P = plan(B0, B1, Bz)
    t = np.arange(P["ncell"], dtype=np.int64)
    px, py, kk = decode(P, t)
    inside = (kk >= 0) & (kk < Bz)
    lane = t % THREADS
    li, lj = lane % KT, lane // KT
    forms = [np.where(li > 0, t - THREADS - 1, t + P["kdx"]),
             np.where(lj > 0, t - THREADS - KT, t + P["kdy"]),
             t - THREADS]
    shifts = [(1, 0, 0), (0, 1, 0), (0, 0, 1)]
    for d in range(3):
        sx, sy, sz = shifts[d]
        has = inside & (px - sx >= 0) & (py - sy >= 0) & (kk - sz >= 0)
        n = forms[d][has]
        assert n.min() >= 0 and n.max() < P["ncell"]
        qx, qy, qk = decode(P, n)
        assert np.array_equal(qx, px[has] - sx)
        assert np.array_equal(qy, py[has] - sy)
        assert np.array_equal(qk, kk[has] - sz)
