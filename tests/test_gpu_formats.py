"""The serializer entry points (wc_serialize_position_only,
wc_load_position_only, wc_serialize_fluid) through the C ABI, byte for byte
against the oracle's restatement of the reference's files
(oracle/serializer.py, pinned in tests/test_formats.py), on the particle
state the context holds after a few substeps (FP64 and mixed)."""
import numpy as np
import pytest

from oracle import serializer as OS

pytestmark = pytest.mark.gpu

N = 16
DT = 1.0 / 480


def _scene(wc, precision):
    s = wc.Scene((N, N, N), 1.0 / N, precision=precision)
    s.seed_block((0.0, 0.0, 0.0), (0.4, 0.6, 1.0), seed=5, sigma=0.5)
    for _ in range(3):
        s.step(DT, 1e-8, 500)
    return s


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_serialize_matches_oracle_bytes(wc, tmp_path, prec):
    s = _scene(wc, wc.WC_FP64 if prec == "fp64" else wc.WC_MIXED)
    x4, _, _ = s.particles_reference_layout()
    p_pos = tmp_path / "pos.bin"
    p_fluid = tmp_path / "fluid.obj"
    s.serializePositionOnly(str(p_pos))
    s.serializeFluid(str(p_fluid))
    assert p_pos.read_bytes() == OS.position_only_bytes(x4)
    radius = s.desc.particle_radius
    assert p_fluid.read_bytes() == OS.fluid_text(x4, radius).encode()
    s.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_load_position_only_round_trip(wc, tmp_path, prec):
    s = _scene(wc, wc.WC_FP64 if prec == "fp64" else wc.WC_MIXED)
    x4, v4, fB = s.particles_reference_layout()
    n = x4.size // 4
    rng = np.random.default_rng(11)
    new = x4.copy()
    new[0::4] = rng.uniform(0.1, 0.9, size=n)
    new[1::4] = rng.uniform(0.1, 0.5, size=n)
    new[2::4] = rng.uniform(0.1, 0.9, size=n)
    p = tmp_path / "pos.bin"
    p.write_bytes(OS.position_only_bytes(new))
    s.loadPosOnly(str(p))
    x4b, v4b, fBb = s.particles_reference_layout()
    want = new if prec == "fp64" else new.astype(np.float32).astype(np.float64)
    assert np.array_equal(x4b, want)
    assert np.array_equal(v4b, v4) and np.array_equal(fBb, fB)
    # and written back, the same bytes as the oracle's for the loaded state
    p2 = tmp_path / "pos2.bin"
    s.serializePositionOnly(str(p2))
    assert p2.read_bytes() == OS.position_only_bytes(want)
    # the context still steps from the loaded positions
    ok, it, res = s.step(DT, 1e-8, 500)
    assert ok
    s.close()


def test_load_position_only_rejects_mismatched_files(wc, tmp_path):
    s = _scene(wc, wc.WC_FP64)
    x4, _, _ = s.particles_reference_layout()
    bad_count = tmp_path / "short.bin"
    bad_count.write_bytes(OS.position_only_bytes(x4[:-4]))
    with pytest.raises(wc.WetclothError):
        s.loadPosOnly(str(bad_count))
    gauss = tmp_path / "gauss.bin"
    gauss.write_bytes(OS.position_only_bytes(x4, np.eye(3).ravel()))
    with pytest.raises(wc.WetclothError):
        s.loadPosOnly(str(gauss))
    trunc = tmp_path / "trunc.bin"
    trunc.write_bytes(OS.position_only_bytes(x4)[:-6])
    with pytest.raises(wc.WetclothError):
        s.loadPosOnly(str(trunc))
    with pytest.raises(wc.WetclothError):
        s.serializeFluid(str(tmp_path / "no_such_dir" / "fluid.obj"))
    # the positions are untouched by the rejected loads
    x4b, _, _ = s.particles_reference_layout()
    assert np.array_equal(x4b, x4)
    s.close()
