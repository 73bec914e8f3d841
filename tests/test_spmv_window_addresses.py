"""The windowed box SpMV's index arithmetic, restated on the CPU (no GPU).

k_box_spmv_win (libwetcloth_amd/csrc/wc_sweep.hip) walks each 16x16 tile of
the skewed box layout in runs of consecutive steps and takes a row's six
neighbours from LDS rings (x of steps w - 1 .. w + 1, the upper coefficient
planes of step w - 1) or, on the tile's faces, from the next tile in global
memory.  This module restates where each term comes from and checks, over
many box shapes and run lengths:
  - every global address is inside the box arrays, and the runs cover every
    slot once;
  - for every cell of the box and each of its six neighbours inside the box,
    the ring entry or face slot the term reads is that neighbour's slot;
  - the ring indices of the values a step reads are never the ones the next
    step overwrites before its barrier (x ring of 4, planes ring of 3);
  - the compact x-face columns it reads hold the face slots' values.
"""
import numpy as np
import pytest

KT = 16
THREADS = KT * KT


def plan(B0, B1, Bz):
    nta, ntb = B0 // KT, B1 // KT
    nsteps = Bz + 2 * (KT - 1)
    ncell = nta * ntb * nsteps * THREADS
    kdx = -nsteps * THREADS + (KT - 1) * THREADS + (KT - 1)
    kdy = -nta * nsteps * THREADS + (KT - 1) * THREADS + (KT - 1) * KT
    return dict(nta=nta, ntb=ntb, nsteps=nsteps, ncell=ncell, kdx=kdx, kdy=kdy, Bz=Bz)


def slot(P, tile, w, lane):
    return (tile * P["nsteps"] + w) * THREADS + lane


def cell_of(P, t):
    lane = t % THREADS
    q = t // THREADS
    w = q % P["nsteps"]
    tile = q // P["nsteps"]
    a, b = tile % P["nta"], tile // P["nta"]
    li, lj = lane % KT, lane // KT
    return a * KT + li, b * KT + lj, w - li - lj


def clampn(P, n):
    """The kernel's clamp of a 32-bit slot index (negatives wrap high)."""
    n = np.asarray(n) % (1 << 32)
    return np.minimum(n, P["ncell"] - 1)


def sources(P, tile, w, lane):
    """Slot each term of row (tile, w, lane) reads, or None for a ring entry
    outside the tile's steps (read as +0): lower x, y, z, upper x, y, z."""
    li, lj = lane % KT, lane // KT
    t = slot(P, tile, w, lane)
    ring = lambda ww, ll: slot(P, tile, ww, ll) if 0 <= ww < P["nsteps"] else None
    lo = [ring(w - 1, lane - 1) if li > 0 else int(clampn(P, t + P["kdx"])),
          ring(w - 1, lane - KT) if lj > 0 else int(clampn(P, t + P["kdy"])),
          ring(w - 1, lane)]
    up = [ring(w + 1, lane + 1) if li < KT - 1 else int(clampn(P, t - P["kdx"])),
          ring(w + 1, lane + KT) if lj < KT - 1 else int(clampn(P, t - P["kdy"])),
          ring(w + 1, lane)]
    return lo, up


SHAPES = [(16, 16, 1), (16, 32, 5), (32, 16, 17), (48, 32, 9), (32, 48, 40)]


@pytest.mark.parametrize("B0,B1,Bz", SHAPES)
def test_terms_read_the_neighbours_slots(B0, B1, Bz):
    P = plan(B0, B1, Bz)
    dirs = [(-1, 0, 0), (0, -1, 0), (0, 0, -1), (1, 0, 0), (0, 1, 0), (0, 0, 1)]
    checked = 0
    for tile in range(P["nta"] * P["ntb"]):
        for w in range(P["nsteps"]):
            for lane in range(THREADS):
                px, py, kk = cell_of(P, slot(P, tile, w, lane))
                if not 0 <= kk < Bz:
                    continue
                lo, up = sources(P, tile, w, lane)
                for (ddx, ddy, ddz), src in zip(dirs, lo + up):
                    nx, ny, nz = px + ddx, py + ddy, kk + ddz
                    if not (0 <= nx < B0 and 0 <= ny < B1 and 0 <= nz < Bz):
                        continue
                    assert src is not None, (tile, w, lane, ddx, ddy, ddz)
                    assert 0 <= src < P["ncell"]
                    assert cell_of(P, src) == (nx, ny, nz), (tile, w, lane, ddx, ddy, ddz)
                    checked += 1
    assert checked > 0


@pytest.mark.parametrize("run", [1, 5, 16, 64, 300])
def test_runs_cover_every_slot_once_in_range(run):
    for B0, B1, Bz in SHAPES:
        P = plan(B0, B1, Bz)
        nrun = (P["nsteps"] + run - 1) // run
        written = np.zeros(P["ncell"], dtype=np.int64)
        for block in range(P["nta"] * P["ntb"] * nrun):
            tile, w0 = block // nrun, (block % nrun) * run
            w1 = min(w0 + run, P["nsteps"])
            lanes = np.arange(THREADS)
            # the prologue: x of w0 - 1 and w0, planes of w0 - 1 (guarded)
            for ww in (w0 - 1, w0):
                if 0 <= ww < P["nsteps"]:
                    assert 0 <= slot(P, tile, ww, 0) and slot(P, tile, ww, THREADS - 1) < P["ncell"]
            for w in range(w0, w1):
                t = slot(P, tile, w, lanes)
                np.add.at(written, t, 1)
                if w + 1 < P["nsteps"]:
                    assert slot(P, tile, w + 1, THREADS - 1) < P["ncell"]
                for off in (P["kdx"], P["kdy"], -P["kdx"], -P["kdy"]):
                    c = clampn(P, t + off)
                    assert c.min() >= 0 and c.max() < P["ncell"]
        assert np.all(written == 1), (B0, B1, Bz, run)


def test_ring_entries_live_until_read():
    """Step w writes x of w + 1 into xs[(w + 1) & 3] and its planes into
    us[w % 3], then reads x of w - 1, w, w + 1 and the planes of w - 1; step
    w + 1 writes before its barrier, possibly while a slower wave still reads
    step w's values: those writes must miss everything step w reads."""
    for w in range(0, 40):
        reads_x = {(w - 1) & 3, w & 3, (w + 1) & 3}
        reads_u = {(w - 1) % 3}
        assert len(reads_x) == 3
        assert ((w + 2) & 3) not in reads_x
        assert ((w + 1) % 3) not in reads_u
        # the planes step w reads are the ones step w - 1 wrote
        assert (w + 2) % 3 == (w - 1) % 3


@pytest.mark.parametrize("B0,B1,Bz", SHAPES)
def test_face_columns_hold_the_face_slots(B0, B1, Bz):
    """The x faces come from compact columns: f15[i] = x[16 i + 15] and
    f0[i] = x[16 i].  A lane with li = 0 reads f15[clampn(t + dx) >> 4]: the
    clamped slot has li = 15, so that entry is exactly x at the slot the
    row needs.  A lane with li = 15 reads f0[clampn(t - dx) >> 4]: where
    the slot is inside the box it has li = 0, so the entry is x there (a
    clamped slot belongs to no neighbour, and its entry is still a value of
    x).  Every index is inside the columns."""
    P = plan(B0, B1, Bz)
    n16 = P["ncell"] // 16
    t = np.arange(P["ncell"], dtype=np.int64)
    li = (t % THREADS) % KT
    lo = clampn(P, t[li == 0] + P["kdx"]).astype(np.int64)
    assert np.all(lo % 16 == 15)
    assert np.array_equal((lo >> 4) * 16 + 15, lo)
    assert (lo >> 4).max() < n16
    raw = t[li == KT - 1] - P["kdx"]
    up = clampn(P, raw).astype(np.int64)
    inside = (raw >= 0) & (raw < P["ncell"])
    assert np.all(up[inside] % 16 == 0)
    assert np.array_equal((up[inside] >> 4) * 16, up[inside])
    assert (up >> 4).max() < n16
    # the s update writes f15 / f0 at exactly the slots with t & 15 == 15 / 0
    assert np.array_equal(np.unique(t[(t & 15) == 15] >> 4), np.arange(n16))
    assert np.array_equal(np.unique(t[(t & 15) == 0] >> 4), np.arange(n16))
