"""Pin the oracle's math kernels.

1. Bitwise against the reference's own Core/MathUtilities.h and
   Core/PCGSolver/BLASWrapper.h, compiled from the reference sources into
   oracle/_ref (skipped where the reference tree is absent).
2. Against closed forms that a transposed operand, dropped term or wrong
   branch would break: B-spline partition of unity and moments, the
   marching-squares areas of fraction_inside, the box distance.
"""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def ref(oracle):
    oracle.build()
    r = oracle.ref_lib()
    if r is None:
        pytest.skip("reference tree not present: oracle/_ref not built")
    return r


def sample_values(rng, n=4000):
    v = list(rng.uniform(-2.0, 2.0, n))
    v += [0.0, -0.0, 0.5, -0.5, 1.5, -1.5, 1.0, -1.0, 0.49999999999999994,
          1.4999999999999998, 2.0, -3.0]
    return v


def test_quad_kernel_bitwise_vs_reference(oracle, ref):
    L = oracle.lib()
    for x in sample_values(np.random.default_rng(0)):
        assert L.wco_quad_kernel(x) == ref.ref_quad_kernel(x)


def test_N_kernel_is_product_in_reference_order(oracle, ref):
    L = oracle.lib()
    rng = np.random.default_rng(1)
    for _ in range(2000):
        a, b, c = rng.uniform(-1.6, 1.6, 3)
        mine = L.wco_quad_kernel(a) * L.wco_quad_kernel(b) * L.wco_quad_kernel(c)
        assert mine == ref.ref_N_kernel2(a, b, c)


def test_fraction_inside2_bitwise(oracle, ref):
    L = oracle.lib()
    rng = np.random.default_rng(2)
    vals = list(rng.uniform(-1, 1, 200)) + [0.0, -0.0, 1e-300, -1e-300]
    for a in vals:
        for b in vals[::7]:
            assert L.wco_fraction_inside2(a, b) == ref.ref_fraction_inside2(a, b)


def test_fraction_inside4_bitwise_all_sign_patterns(oracle, ref):
    L = oracle.lib()
    rng = np.random.default_rng(3)
    for pattern in range(16):
        for _ in range(200):
            mag = rng.uniform(0.01, 1.0, 4)
            c = [(-m if (pattern >> i) & 1 else m) for i, m in enumerate(mag)]
            assert L.wco_fraction_inside4(*c) == ref.ref_fraction_inside4(*c)
    # exact zeros (wall planes) and mixed zero/negative corners
    for c in ([0, 0, 0, 0], [-1, 0, 0, 0], [0, -1, -1, 0], [-1, -1, 0, 0], [0, 0, -1, -1],
              [-0.5, 0, -0.5, 0], [-1, -1, -1, 0]):
        c = [float(t) for t in c]
        assert L.wco_fraction_inside4(*c) == ref.ref_fraction_inside4(*c)


def test_inverse_D_and_lerp_and_clamp(ref):
    for h in (1 / 64, 1 / 128, 1 / 256, 0.013):
        assert ref.ref_inverse_D_coeff(h, 2) == 4.0 / (h * h)
    # constrainLiquidVelocity with friction 1 replaces u by the solid velocity
    for u in (1.5, -2.25, 0.0):
        assert ref.ref_lerp(u, 0.0, 1.0) == 0.0
    assert ref.ref_clamp(1.5, 0.0, 1.0) == 1.0 and ref.ref_clamp(-0.5, 0.0, 1.0) == 0.0


def test_blas_wrappers(oracle, ref):
    rng = np.random.default_rng(4)
    for n in (1, 7, 1000, 4099):
        a = rng.standard_normal(n)
        b = rng.standard_normal(n)
        d_ref = ref.ref_blas_dot(oracle.ptr(a), oracle.ptr(b), n)
        d_or = oracle.lib().wco_dot(n, oracle.ptr(a), oracle.ptr(b))
        # Eigen's dot is SIMD-blocked; the oracle sums sequentially
        assert abs(d_ref - d_or) <= 1e-13 * np.abs(a * b).sum()
        assert oracle.lib().wco_abs_max(n, oracle.ptr(a)) == ref.ref_blas_abs_max(oracle.ptr(a), n)


# ---- closed forms (independent of the reference build) ----

def test_quad_kernel_closed_form(oracle):
    L = oracle.lib()
    assert L.wco_quad_kernel(0.0) == 0.75
    assert L.wco_quad_kernel(0.5) == 0.5 * 1.0 * 1.0
    assert L.wco_quad_kernel(1.5) == 0.0
    assert L.wco_quad_kernel(-1.0) == 0.125


@pytest.mark.parametrize("off", [0.0, 0.5])
def test_bspline_partition_and_moments(oracle, off):
    """Sum w = 1, sum w (x_n - x) = 0, sum w (x_n - x)^2 = dx^2/4 (this is
    why D^-1 = 4/dx^2, MathUtilities.h:728-737)."""
    L = oracle.lib()
    dx = 1.0 / 64
    rng = np.random.default_rng(5)
    for x in rng.uniform(0.1, 0.9, 500):
        t = x / dx + (oracle.GHOST - off)
        fl = math.floor(t)
        base = fl + (0 if t - fl > 0.5 else -1)
        nodes = [((base + a - oracle.GHOST) + off) * dx for a in range(3)]
        w = [L.wco_quad_kernel((x - xn) / dx) for xn in nodes]
        assert abs(sum(w) - 1.0) < 1e-14
        assert abs(sum(wi * (xn - x) for wi, xn in zip(w, nodes))) < 1e-15
        m2 = sum(wi * (xn - x) ** 2 for wi, xn in zip(w, nodes))
        assert abs(m2 - dx * dx / 4) < 1e-16


def test_fraction_inside_textbook_areas(oracle):
    L = oracle.lib()
    f4 = L.wco_fraction_inside4
    assert f4(-1.0, -1.0, -1.0, -1.0) == 1.0
    assert f4(1.0, 1.0, 1.0, 1.0) == 0.0
    # one corner inside: triangle with legs a, b (fractions along the edges)
    phi = -0.3
    a = phi / (phi - 0.7)   # along bl->br
    b = phi / (phi - 0.9)   # along bl->tl
    assert abs(f4(phi, 0.7, 0.9, 2.0) - 0.5 * a * b) < 1e-15
    # two adjacent corners inside: trapezoid (mean of the two edge fractions)
    fl = -0.2 / (-0.2 - 0.6)
    fr = -0.4 / (-0.4 - 0.4)
    assert abs(f4(-0.2, -0.4, 0.6, 0.4) - 0.5 * (fl + fr)) < 1e-15
    # complementary: three inside = 1 - one outside triangle
    assert abs(f4(0.3, -0.7, -0.9, -2.0) - (1 - 0.5 * (0.3 / 1.0) * (0.3 / 1.2))) < 1e-15
    # symmetric in the 1-D form
    for l, r in ((-0.3, 0.7), (0.7, -0.3)):
        assert L.wco_fraction_inside2(l, r) == 0.3 / 1.0


def test_box_container_phi(oracle):
    L = oracle.lib()
    lo = np.zeros(3)
    hi = np.ones(3)

    def phi(p):
        p = np.asarray(p, dtype=np.float64)
        return L.wco_box_container_phi(oracle.ptr(p), oracle.ptr(lo), oracle.ptr(hi))

    assert phi([0.5, 0.5, 0.5]) == 0.5
    assert phi([0.125, 0.5, 0.75]) == 0.125  # nearest wall
    assert phi([0.0, 0.5, 0.5]) == 0.0      # on the wall plane
    assert phi([-0.25, 0.5, 0.5]) == -0.25  # in the solid
    assert abs(phi([-0.3, -0.4, 0.5]) + 0.5) < 1e-15  # outside an edge: -sqrt(.09+.16)
