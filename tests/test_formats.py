"""The reference serializer's particle files: the oracle's restatement
(oracle/serializer.py) pinned on the CPU.

* Number text: the fluid file writes each double with ``std::ostream <<
  std::setprecision(8)`` (TwoDSceneSerializer.cpp:42-45).  The oracle's
  ``%.8g`` is checked against libstdc++'s own operator<< on ~2,500 values
  (every decade from 1e-310 to 1e300, both signs, values that round up to a
  new decade, the switch to exponent form at 1e-5 and 1e8, exact integers,
  +-0), through a helper compiled here with g++.
* Binary layout: the byte counts are m_x.size() * sizeof(scalar) and
  m_d_gauss.size() * sizeof(scalar) as native ints (:25-31); the test
  states those sizes from the reference's definition and checks the
  oracle's bytes and its reader against them.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import serializer as OS

HELPER = r"""
#include <cstdio>
#include <iomanip>
#include <iostream>
int main(int argc, char** argv) {
  FILE* f = std::fopen(argv[1], "rb");
  double v;
  while (std::fread(&v, sizeof(double), 1, f) == 1) std::cout << std::setprecision(8) << v << "\n";
  std::fclose(f);
  return 0;
}
"""


def _values():
    rng = np.random.default_rng(7)
    vals = [0.0, -0.0, 1.0, -1.0, 0.1, 1.0 / 3.0, 2.0 / 3.0, 0.5, 100.0, 12345678.0,
            123456789.0, 99999999.5, 9.9999999949, 9.999999995, 0.0001, 0.00001,
            0.000099999999, 1e-5, 1e8, 99999999.0, 5e-324, 2.2250738585072014e-308,
            1.7976931348623157e308, 0.999999995, 0.9999999949]
    for k in range(-310, 301, 3):
        m = rng.uniform(1.0, 10.0, size=4)
        vals += list(m * 10.0 ** k) + list(-m * 10.0 ** k)
    vals += list(rng.uniform(0.0, 1.0, size=800))  # positions in a unit box
    vals += list(np.float32(rng.uniform(0.0, 1.0, size=400)).astype(np.float64))
    return np.array([v for v in vals if np.isfinite(v)], dtype=np.float64)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++ for the libstdc++ helper")
def test_fmt_double_matches_libstdcxx_ostream(tmp_path):
    src = tmp_path / "fmt.cpp"
    src.write_text(HELPER)
    exe = tmp_path / "fmt"
    subprocess.run(["g++", "-O1", "-o", str(exe), str(src)], check=True)
    vals = _values()
    data = tmp_path / "vals.bin"
    vals.tofile(data)
    out = subprocess.run([str(exe), str(data)], capture_output=True, text=True, check=True).stdout
    ref = out.splitlines()
    assert len(ref) == len(vals)
    bad = [(v, r, OS.fmt_double(v)) for v, r in zip(vals, ref) if OS.fmt_double(v) != r]
    assert not bad, bad[:5]


def test_fluid_text_lines():
    x4 = np.array([0.25, 0.5, 0.75, 0.0, 1.0 / 3.0, 1e-5, 123456789.0, 0.0])
    t = OS.fluid_text(x4, 0.0025)
    assert t == "v 0.25 0.5 0.75 0.0025\nv 0.33333333 1e-05 1.2345679e+08 0.0025\n"


def test_position_only_layout_and_reader():
    rng = np.random.default_rng(3)
    n = 37
    x4 = rng.uniform(size=4 * n)
    x4[3::4] = 0.0
    b = OS.position_only_bytes(x4)
    # int bufsize = m_pos.size() * sizeof(scalar), the data, then the same for
    # m_d_gauss (TwoDSceneSerializer.cpp:25-31); empty without elastic bodies
    assert len(b) == 4 + 4 * n * 8 + 4
    assert int(np.frombuffer(b[:4], dtype="<i4")[0]) == 4 * n * 8
    assert np.array_equal(np.frombuffer(b[4:4 + 32 * n], dtype="<f8"), x4)
    assert int(np.frombuffer(b[-4:], dtype="<i4")[0]) == 0
    x, g = OS.parse_position_only(b)
    assert np.array_equal(x, x4) and g.size == 0
    gauss = rng.uniform(size=18)
    b2 = OS.position_only_bytes(x4, gauss)
    x2, g2 = OS.parse_position_only(b2)
    assert np.array_equal(x2, x4) and np.array_equal(g2, gauss)
    assert len(b2) == len(b) + 18 * 8
