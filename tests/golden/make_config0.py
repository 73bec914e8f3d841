"""Write tests/golden/dam_break_64_100.npz: the FP64 oracle over 100 substeps
of the 64^3 dam break (block [0,0.4]x[0,0.6]x[0,1], 8 particles per cell,
503,312 particles, at rest, dt = 1/480, MIC(0)-PCG to 1e-8).

The full state is 12 MB per field, so the fixture keeps what a GPU run can
be checked against: the positions and velocities of every 997th particle
(505 of them), the mean position and velocity over all particles, and the
PCG iteration count of every substep.  About three minutes on one core.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402

N = 64
DT = 1.0 / 480
TOL = 1e-8
STEPS = 100
STRIDE = 997


def generate():
    g, x, v = O.dam_break(N, seed=0, sigma=0.0)
    s = O.OracleScene((N, N, N), 1.0 / N, x, v)
    iters = []
    for _ in range(STEPS):
        assert s.substep(DT, TOL, 2000)
        iters.append(s.last_solve[1])
    xf, vf, _ = s.particles()
    idx = np.arange(0, x.shape[1], STRIDE)
    return dict(n=N, dt=DT, tol=TOL, steps=STEPS, np=x.shape[1], idx=idx, x=xf[:, idx],
                v=vf[:, idx], x_mean=xf.mean(axis=1), v_mean=vf.mean(axis=1),
                iters=np.array(iters, dtype=np.int32))


if __name__ == "__main__":
    out = os.path.join(HERE, "dam_break_64_100.npz")
    np.savez_compressed(out, **generate())
    print(out, os.path.getsize(out))
