"""Write tests/golden/dam_break_12.npz from the FP64 oracle.

The reference ships no tests or golden vectors for this path, so the
fixture is the oracle's own output on a seeded 12^3 dam break: it pins the
oracle against regressions and gives the GPU tests a stored target.
Inputs: seed 0, block [0,0.4]x[0,0.6]x[0,1], sigma 0.5 m/s, dt 1/480.
Outputs: the step-1 pressure system and solution, particles after 3 steps.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402

N = 12
DT = 1.0 / 480
TOL = 1e-10


def generate():
    g, x, v = O.dam_break(N, seed=0, sigma=0.5)
    s = O.OracleScene((N, N, N), 1.0 / N, x, v)
    s.pre_solve(DT)
    ok = s.solve_pressure(TOL, 2000)
    assert ok
    rs, col, val, b, p = s.system()
    dof = s.field("dof")
    s.apply_pressure(DT)
    s.g2p_advect(DT)
    for _ in range(2):
        assert s.substep(DT, TOL, 2000)
    x3, v3, C3 = s.particles()
    return dict(n=N, dt=DT, tol=TOL, x0=x, v0=v, rowstart=rs, col=col, val=val, rhs=b,
                pressure=p, dof=dof, x3=x3, v3=v3)


if __name__ == "__main__":
    out = os.path.join(HERE, "dam_break_12.npz")
    np.savez_compressed(out, **generate())
    print(out, os.path.getsize(out))
