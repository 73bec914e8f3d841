"""The reference serializer's particle files, restated -- TEST INFRASTRUCTURE.

Only tests/ import this module; the product package never does.

- ``position_only_bytes`` / ``parse_position_only``: serialize_pos_subprog
  (App/TwoDSceneSerializer.cpp:23-37) and loadPosOnly (:172-183): an int
  (native, 4 bytes) byte count, m_x as raw doubles (4 per particle:
  x, y, z, theta), an int byte count and m_d_gauss (3 x 3 blocks of the
  Gauss points, column-major Eigen storage; none in a liquid-only scene).
- ``fluid_text``: serialize_subprog's fluid loop (:39-47) over updateFluid
  (:242-255): ``"v " << setprecision(8) << x << " " << y << " " << z << " "
  << r << std::endl`` per fluid particle.  An ostream inserts a double with
  the default float field as printf ``%.8g``; ``fmt_double`` is Python's
  ``%.8g``, and tests/test_formats.py pins it against libstdc++'s own
  ``operator<<`` (a helper compiled with g++), so the restatement is checked
  against the real conversion, not retyped from it.
"""
from __future__ import annotations

import numpy as np


def fmt_double(v: float) -> str:
    """std::ostream << std::setprecision(8) << v (default float field)."""
    return "%.8g" % v


def position_only_bytes(x4: np.ndarray, gauss: np.ndarray | None = None) -> bytes:
    """serialize_pos_subprog (TwoDSceneSerializer.cpp:23-37)."""
    x4 = np.ascontiguousarray(x4, dtype="<f8")
    g = np.zeros(0, dtype="<f8") if gauss is None else np.ascontiguousarray(gauss, dtype="<f8")
    out = bytearray()
    out += np.int32(x4.nbytes).tobytes()
    out += x4.tobytes()
    out += np.int32(g.nbytes).tobytes()
    out += g.tobytes()
    return bytes(out)


def parse_position_only(data: bytes):
    """loadPosOnly (TwoDSceneSerializer.cpp:172-183): (m_x, m_d_gauss) as flat arrays."""
    b0 = int(np.frombuffer(data[:4], dtype="<i4")[0])
    x = np.frombuffer(data[4:4 + b0], dtype="<f8").copy()
    o = 4 + b0
    b1 = int(np.frombuffer(data[o:o + 4], dtype="<i4")[0])
    g = np.frombuffer(data[o + 4:o + 4 + b1], dtype="<f8").copy()
    return x, g


def fluid_text(x4: np.ndarray, radius: float) -> str:
    """The fluid file of serializeScene (TwoDSceneSerializer.cpp:39-47, 242-255)."""
    x = np.asarray(x4, dtype=np.float64).reshape(-1, 4)
    r = fmt_double(radius)
    return "".join("v %s %s %s %s\n" % (fmt_double(p[0]), fmt_double(p[1]), fmt_double(p[2]), r)
                   for p in x)
