"""ctypes wrapper over the FP64 CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product package never does.  See wc_oracle.h for what each
function restates and how it is pinned.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libref_math.so")
GHOST = 8

_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)
_i32p = C.POINTER(C.c_int32)
_u32p = C.POINTER(C.c_uint32)


class Grid(C.Structure):
    _fields_ = [("n", C.c_int * 3), ("dx", C.c_double), ("nx", C.c_int * 3),
                ("nb", C.c_int * 3), ("nnode", C.c_int64), ("nbucket", C.c_int64)]


class SceneStruct(C.Structure):
    _fields_ = [
        ("g", Grid),
        ("rho", C.c_double), ("flip", C.c_double), ("friction", C.c_double),
        ("radius", C.c_double), ("pmass", C.c_double), ("pvol", C.c_double),
        ("gravity", C.c_double * 3),
        ("solid_projection", C.c_int),
        ("box_lo", C.c_double * 3), ("box_hi", C.c_double * 3),
        ("np", C.c_int64),
        ("x", _dp), ("v", _dp), ("C", _dp),
        ("act", _u8p), ("lphi", _dp), ("sphi", _dp),
        ("w", _dp * 3), ("mass", _dp * 3), ("vel", _dp * 3), ("vol", _dp * 3),
        ("rhs", _dp * 3), ("invC", _dp * 3), ("ustar", _dp * 3),
        ("psi_fs", _dp * 3), ("unew", _dp * 3), ("valid", _u8p * 3),
        ("dof", _i32p), ("pressure", _dp),
        ("ndof", C.c_int64), ("rowstart", _u32p), ("colidx", _u32p),
        ("val", _dp), ("b", _dp), ("p", _dp), ("nnz_cap", C.c_int64),
        ("last_residual", C.c_double), ("last_iters", C.c_int), ("last_ok", C.c_int),
        ("frz_mode", C.c_int), ("lphi_arg", _i32p), ("frz_act", _u8p), ("frz_arg", _i32p),
        ("frz_liquid", _u8p), ("frz_dof", _i32p), ("frz_ndof", C.c_int64),
        ("frz_clamp", _u8p * 3), ("frz_mass", _u8p * 3), ("frz_proj", _u8p),
        ("mix", C.c_void_p),
    ]


class MixParams(C.Structure):
    """wco_mix_params: the LiquidInfo fields of the mixture terms."""
    _fields_ = [("use_drag", C.c_int), ("use_nonlinear_drag", C.c_int),
                ("yarn_diameter", C.c_double), ("viscosity", C.c_double),
                ("yazdchi_power", C.c_double), ("pore_radius", C.c_double),
                ("surf_tension_coeff", C.c_double), ("rest_contact_angle", C.c_double)]


def mix_params(**kw) -> MixParams:
    """The mixture parameters of the cloth scenes in SI units: the reference's
    LiquidInfo defaults (TwoDSceneXMLParser.cpp:1849-1904, given there in CGS)
    converted; a rest volume fraction, where given, derives the pore radius
    from the yarn diameter as the parser does (:1960-1963)."""
    yarn = kw.get("yarn_diameter", 1e-4)
    p = MixParams()
    p.use_drag = int(kw.get("use_drag", 1))
    p.use_nonlinear_drag = int(kw.get("use_nonlinear_drag", 1))
    p.yarn_diameter = yarn
    p.viscosity = kw.get("viscosity", 8.9e-4)
    p.yazdchi_power = kw.get("yazdchi_power", 1.6)
    if "rest_volume_fraction" in kw:
        rvf = kw["rest_volume_fraction"]
        p.pore_radius = yarn * 0.5 * np.sqrt(max(0.0, (1.0 - rvf) / rvf))
    else:
        p.pore_radius = kw.get("pore_radius", 5e-5)
    p.surf_tension_coeff = kw.get("surf_tension_coeff", 0.0728)
    p.rest_contact_angle = kw.get("rest_contact_angle", 40.8 / 180.0 * np.pi)
    return p


_lib = None


def build(force: bool = False) -> None:
    """Compile liboracle.so (and the reference harness where the tree exists)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if os.path.isdir("/root/reference/libWetCloth") and (force or not os.path.exists(REF_LIB_PATH)):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _configure(C.CDLL(LIB_PATH))
    return _lib


def build_variant(out_dir: str, plant: int, ghost: int | None = None) -> str:
    """Compile the oracle with planted mistake ``plant`` (see WCO_PLANT in
    wc_oracle.c), or with ``ghost`` ghost cells below the container instead
    of WCO_GHOST, into ``out_dir``; returns the library's path.  For the
    tests that show each pin fails on a named mistake."""
    tag = f"plant{plant}" + (f"_ghost{ghost}" if ghost is not None else "")
    out = os.path.join(out_dir, f"liboracle_{tag}.so")
    defs = [f"-DWCO_PLANT={plant}"] + ([f"-DWCO_GHOST={ghost}"] if ghost is not None else [])
    subprocess.run(["gcc", "-O2", "-std=c99", "-ffp-contract=off", "-fPIC", "-shared", *defs,
                    "-o", out, os.path.join(HERE, "wc_oracle.c"),
                    os.path.join(HERE, "wc_oracle_amg.c"), "-lm"], check=True)
    return out


def load_variant(path: str):
    """A separately built oracle library (same interface as lib())."""
    return _configure(C.CDLL(path))


def _configure(L):
    """Set the argument and result types of an oracle library."""
    L.wco_quad_kernel.restype = C.c_double
    L.wco_quad_kernel.argtypes = [C.c_double]
    L.wco_fraction_inside2.restype = C.c_double
    L.wco_fraction_inside2.argtypes = [C.c_double] * 2
    L.wco_fraction_inside4.restype = C.c_double
    L.wco_fraction_inside4.argtypes = [C.c_double] * 4
    L.wco_box_container_phi.restype = C.c_double
    L.wco_box_container_phi.argtypes = [_dp, _dp, _dp]
    L.wco_hash3.restype = C.c_uint64
    L.wco_hash3.argtypes = [C.c_uint64] * 3
    L.wco_uniform.restype = C.c_double
    L.wco_uniform.argtypes = [C.c_uint64]
    L.wco_grid_init.argtypes = [C.POINTER(Grid), C.c_int, C.c_int, C.c_int, C.c_double]
    L.wco_node_index.restype = C.c_int64
    L.wco_node_index.argtypes = [C.POINTER(Grid), C.c_int, C.c_int, C.c_int]
    L.wco_seed_block.restype = C.c_int64
    L.wco_seed_block.argtypes = [C.POINTER(Grid), _dp, _dp, C.c_uint64, _dp, C.c_int64]
    L.wco_seed_velocity.argtypes = [C.POINTER(Grid), C.c_uint64, C.c_double,
                                    C.c_int64, _dp, _dp]
    L.wco_scene_create.restype = C.POINTER(SceneStruct)
    L.wco_scene_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                   C.c_double, C.c_double, C.c_double, C.c_double,
                                   C.c_int64, _dp, _dp, _dp, C.c_double, C.c_double,
                                   C.c_double]
    L.wco_scene_destroy.argtypes = [C.POINTER(SceneStruct)]
    for nm in ("wco_activate", "wco_update_liquid_phi", "wco_update_solid_phi",
               "wco_update_solid_weights", "wco_p2g", "wco_copy_vel_to_unew"):
        getattr(L, nm).argtypes = [C.POINTER(SceneStruct)]
    L.wco_step_velocity.argtypes = [C.POINTER(SceneStruct), C.c_double]
    L.wco_assemble.restype = C.c_int64
    L.wco_assemble.argtypes = [C.POINTER(SceneStruct), C.c_double]
    L.wco_solve_pressure.restype = C.c_int
    L.wco_solve_pressure.argtypes = [C.POINTER(SceneStruct), C.c_double, C.c_int]
    L.wco_apply_pressure.argtypes = [C.POINTER(SceneStruct), C.c_double]
    L.wco_g2p_advect.argtypes = [C.POINTER(SceneStruct), C.c_double]
    L.wco_substep.restype = C.c_int
    L.wco_substep.argtypes = [C.POINTER(SceneStruct), C.c_double, C.c_double, C.c_int]
    L.wco_spmv.argtypes = [C.c_int64, _u32p, _u32p, _dp, _dp, _dp]
    L.wco_mic0_factor.argtypes = [C.c_int64, _u32p, _u32p, _dp, C.c_double,
                                  C.c_double, _dp, _dp]
    L.wco_solve_lower.argtypes = [C.c_int64, _u32p, _u32p, _dp, _dp, _dp, _dp]
    L.wco_solve_lower_transpose.argtypes = [C.c_int64, _u32p, _u32p, _dp, _dp, _dp]
    L.wco_dot.restype = C.c_double
    L.wco_dot.argtypes = [C.c_int64, _dp, _dp]
    L.wco_abs_max.restype = C.c_double
    L.wco_abs_max.argtypes = [C.c_int64, _dp]
    L.wco_pcg_solve.restype = C.c_int
    L.wco_pcg_solve.argtypes = [C.c_int64, _u32p, _u32p, _dp, _dp, _dp, C.c_double,
                                C.c_int, C.c_double, C.c_double, _dp,
                                C.POINTER(C.c_int)]
    _i32p_ = C.POINTER(C.c_int32)
    L.wco_amg_levels.restype = C.c_int
    L.wco_amg_levels.argtypes = [C.c_int64, _u32p, _u32p, _dp, _i32p_, C.c_int, C.c_int,
                                 C.c_int, C.POINTER(C.c_int64), C.c_int]
    L.wco_amg_coarse.restype = C.c_int64
    L.wco_amg_coarse.argtypes = [C.c_int64, _u32p, _u32p, _dp, _i32p_, C.c_int, C.c_int,
                                 C.c_int, _u32p, _u32p, _u32p, _dp, C.c_int64]
    L.wco_amg_rbgs.argtypes = [C.c_int64, _u32p, _u32p, _dp, _i32p_, _dp, _dp, C.c_int]
    L.wco_amg_vcycle.argtypes = [C.c_int64, _u32p, _u32p, _dp, _i32p_, C.c_int, C.c_int,
                                 C.c_int, _dp, _dp]
    L.wco_amgpcg_solve.restype = C.c_int
    L.wco_amgpcg_solve.argtypes = [C.c_int64, _u32p, _u32p, _dp, _i32p_, C.c_int, C.c_int,
                                   C.c_int, _dp, _dp, C.c_double, C.c_int, _dp,
                                   C.POINTER(C.c_int)]
    for nm in ("wco_extrapolate", "wco_constrain_liquid_velocity"):
        getattr(L, nm).argtypes = [C.POINTER(SceneStruct)]
    L.wco_set_elasto.argtypes = [C.POINTER(SceneStruct), C.c_int64, _dp, _dp, _dp, _dp, _dp,
                                 _dp, _dp, _dp, C.POINTER(MixParams)]
    L.wco_mix_field.restype = _dp
    L.wco_mix_field.argtypes = [C.POINTER(SceneStruct), C.c_char_p, C.c_int]
    L.wco_drag_coeff.restype = C.c_double
    L.wco_drag_coeff.argtypes = [C.POINTER(MixParams), C.c_double, C.c_double, C.c_double,
                                 C.c_double, _dp, C.c_double, C.c_int]
    L.wco_rotated_drag.restype = C.c_double
    L.wco_rotated_drag.argtypes = [_dp, _dp, C.c_int]
    L.wco_rotated_pore_coeff.restype = C.c_double
    L.wco_rotated_pore_coeff.argtypes = [_dp, C.c_double, C.c_int]
    L.wco_capillary_pressure.restype = C.c_double
    L.wco_capillary_pressure.argtypes = [C.POINTER(MixParams), C.c_double]
    L.wco_frozen_set.argtypes = [C.POINTER(SceneStruct), C.c_int]
    L.wco_substep_adjoint.restype = C.c_int
    L.wco_substep_adjoint.argtypes = [C.POINTER(SceneStruct), C.c_double, C.c_double, C.c_int,
                                      _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    return L


def ref_lib():
    """The reference's MathUtilities/BLASWrapper compiled from its sources, or None."""
    if not os.path.exists(REF_LIB_PATH):
        return None
    L = C.CDLL(REF_LIB_PATH)
    for nm, n in (("ref_quad_kernel", 1), ("ref_N_kernel2", 3), ("ref_fraction_inside2", 2),
                  ("ref_fraction_inside4", 4), ("ref_lerp", 3), ("ref_clamp", 3)):
        f = getattr(L, nm)
        f.restype = C.c_double
        f.argtypes = [C.c_double] * n
    L.ref_inverse_D_coeff.restype = C.c_double
    L.ref_inverse_D_coeff.argtypes = [C.c_double, C.c_int]
    L.ref_blas_dot.restype = C.c_double
    L.ref_blas_dot.argtypes = [_dp, _dp, C.c_long]
    L.ref_blas_abs_max.restype = C.c_double
    L.ref_blas_abs_max.argtypes = [_dp, C.c_long]
    for nm in ("ref_rotated_drag", "ref_rotated_pore_coeff"):
        if hasattr(L, nm):
            getattr(L, nm).restype = C.c_double
    if hasattr(L, "ref_rotated_drag"):
        L.ref_rotated_drag.argtypes = [_dp, _dp, C.c_int]
        L.ref_rotated_pore_coeff.argtypes = [_dp, C.c_double, C.c_int]
    return L


def ptr(a: np.ndarray, t=_dp):
    return a.ctypes.data_as(t)


def grid(n, dx) -> Grid:
    g = Grid()
    lib().wco_grid_init(C.byref(g), int(n[0]), int(n[1]), int(n[2]), float(dx))
    return g


def node_index(g: Grid, i, j, k) -> int:
    return lib().wco_node_index(C.byref(g), i, j, k)


def node_index_array(g: Grid, i, j, k):
    """Vectorised blocked node index (same formula as wco_node_index)."""
    i = np.asarray(i, dtype=np.int64)
    j = np.asarray(j, dtype=np.int64)
    k = np.asarray(k, dtype=np.int64)
    b = ((k >> 2) * g.nb[1] + (j >> 2)) * g.nb[0] + (i >> 2)
    return (b << 6) | ((k & 3) << 4) | ((j & 3) << 2) | (i & 3)


def seed_block(g: Grid, lo, hi, seed: int) -> np.ndarray:
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    n = lib().wco_seed_block(C.byref(g), ptr(lo), ptr(hi), C.c_uint64(seed), None, 0)
    x = np.zeros(3 * n, dtype=np.float64)
    lib().wco_seed_block(C.byref(g), ptr(lo), ptr(hi), C.c_uint64(seed), ptr(x), n)
    return x.reshape(3, n)


def seed_velocity(g: Grid, seed: int, sigma: float, x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.shape[1]
    v = np.zeros((3, n), dtype=np.float64)
    lib().wco_seed_velocity(C.byref(g), C.c_uint64(seed), float(sigma), n, ptr(x), ptr(v))
    return v


class OracleScene:
    """FP64 CPU scene: one reference substep per :meth:`substep` call."""

    def __init__(self, n, dx, x, v=None, Cmat=None, rho=1000.0, flip=0.99,
                 gravity=(0.0, -9.81, 0.0), pvol=None, radius=None, L=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        npart = x.shape[1]
        v = np.zeros_like(x) if v is None else np.ascontiguousarray(v, dtype=np.float64)
        Cm = (np.zeros((9, npart)) if Cmat is None
              else np.ascontiguousarray(Cmat, dtype=np.float64))
        pvol = dx ** 3 / 8.0 if pvol is None else pvol
        pmass = rho * pvol
        radius = (3.0 * pvol / (4.0 * np.pi)) ** (1.0 / 3.0) if radius is None else radius
        self.n = tuple(int(a) for a in n)
        self.dx = float(dx)
        self.np = npart
        self.L = L if L is not None else lib()
        self.s = self.L.wco_scene_create(self.n[0], self.n[1], self.n[2], self.dx, rho, flip,
                                        gravity[0], gravity[1], gravity[2], npart,
                                        ptr(x), ptr(v), ptr(Cm), pmass, pvol, radius)

    def __del__(self):
        if getattr(self, "s", None) is not None and getattr(self, "L", None) is not None:
            self.L.wco_scene_destroy(self.s)
            self.s = None

    @property
    def st(self) -> SceneStruct:
        return self.s.contents

    @property
    def grid(self) -> Grid:
        return self.st.g

    def _arr(self, p, n, dtype=np.float64):
        if n == 0:
            return np.zeros(0, dtype=dtype)
        return np.ctypeslib.as_array(p, shape=(n,)).view(dtype).copy()

    def field(self, name: str, axis: int | None = None) -> np.ndarray:
        st = self.st
        nn = st.g.nnode
        if name in ("w", "mass", "vel", "vol", "rhs", "invC", "ustar", "psi_fs", "unew"):
            return self._arr(getattr(st, name)[axis], nn)
        if name == "valid":
            return self._arr(st.valid[axis], nn, np.uint8)
        if name in ("lphi", "sphi", "pressure"):
            return self._arr(getattr(st, name), nn)
        if name == "dof":
            return self._arr(st.dof, nn, np.int32)
        if name == "act":
            return self._arr(st.act, st.g.nbucket, np.uint8)
        raise KeyError(name)

    def particles(self):
        st = self.st
        n = self.np
        x = self._arr(st.x, 3 * n).reshape(3, n)
        v = self._arr(st.v, 3 * n).reshape(3, n)
        Cm = self._arr(st.C, 9 * n).reshape(9, n)
        return x, v, Cm

    def system(self):
        st = self.st
        n = st.ndof
        rs = self._arr(st.rowstart, n + 1, np.uint32)
        nnz = int(rs[-1]) if n > 0 else 0
        col = self._arr(st.colidx, nnz, np.uint32)
        val = self._arr(st.val, nnz)
        b = self._arr(st.b, n)
        p = self._arr(st.p, n)
        return rs, col, val, b, p

    def set_solid_projection(self, on: bool):
        self.st.solid_projection = 1 if on else 0

    # reference phases
    def activate(self): self.L.wco_activate(self.s)
    def update_liquid_phi(self): self.L.wco_update_liquid_phi(self.s)
    def update_solid_phi(self): self.L.wco_update_solid_phi(self.s)
    def update_solid_weights(self): self.L.wco_update_solid_weights(self.s)
    def p2g(self): self.L.wco_p2g(self.s)
    def copy_vel_to_unew(self): self.L.wco_copy_vel_to_unew(self.s)
    def step_velocity(self, dt): self.L.wco_step_velocity(self.s, dt)
    def assemble(self, dt): return self.L.wco_assemble(self.s, dt)
    def solve_pressure(self, tol, maxit): return self.L.wco_solve_pressure(self.s, tol, maxit)
    def apply_pressure(self, dt): self.L.wco_apply_pressure(self.s, dt)
    def extrapolate(self): self.L.wco_extrapolate(self.s)
    def constrain_liquid_velocity(self): self.L.wco_constrain_liquid_velocity(self.s)
    def g2p_advect(self, dt): self.L.wco_g2p_advect(self.s, dt)

    def set_field(self, name: str, axis: int | None, values: np.ndarray):
        """Overwrite a grid field of the scene (test hook)."""
        st = self.st
        nn = st.g.nnode
        if name in ("unew", "vel", "ustar", "w", "mass", "vol", "rhs", "invC"):
            dst = np.ctypeslib.as_array(getattr(st, name)[axis], shape=(nn,))
            dst[:] = np.asarray(values, dtype=np.float64)
        elif name == "valid":
            dst = np.ctypeslib.as_array(st.valid[axis], shape=(nn,))
            dst[:] = np.asarray(values, dtype=np.uint8)
        else:
            raise KeyError(name)

    def substep(self, dt, tol=1e-8, maxit=1000):
        return self.L.wco_substep(self.s, dt, tol, maxit)

    # ---- the adjoint (decisions recorded, replayed; wco_substep_adjoint) ----
    def set_particles(self, x, v, Cm):
        st = self.st
        n = self.np
        for name, a, k in (("x", x, 3), ("v", v, 3), ("C", Cm, 9)):
            dst = np.ctypeslib.as_array(getattr(st, name), shape=(k * n,))
            dst[:] = np.ascontiguousarray(a, dtype=np.float64).ravel()

    def frozen(self):
        """A copy of the recorded decisions of the last substep."""
        st = self.st
        nn, nb = st.g.nnode, st.g.nbucket
        return dict(act=self._arr(st.frz_act, nb, np.uint8), arg=self._arr(st.frz_arg, nn, np.int32),
                    liquid=self._arr(st.frz_liquid, nn, np.uint8),
                    dof=self._arr(st.frz_dof, nn, np.int32), ndof=int(st.frz_ndof),
                    clamp=[self._arr(st.frz_clamp[d], nn, np.uint8) for d in range(3)],
                    mass=[self._arr(st.frz_mass[d], nn, np.uint8) for d in range(3)],
                    proj=self._arr(st.frz_proj, self.np, np.uint8))

    def set_frozen(self, f):
        st = self.st
        nn, nb = st.g.nnode, st.g.nbucket

        def put(ptr, n, a):
            np.ctypeslib.as_array(ptr, shape=(n,))[:] = a
        put(st.frz_act, nb, f["act"])
        put(st.frz_arg, nn, f["arg"])
        put(st.frz_liquid, nn, f["liquid"])
        put(st.frz_dof, nn, f["dof"])
        st.frz_ndof = f["ndof"]
        for d in range(3):
            put(st.frz_clamp[d], nn, f["clamp"][d])
            put(st.frz_mass[d], nn, f["mass"][d])
        put(st.frz_proj, self.np, f["proj"])

    def frozen_mode(self, mode):
        self.L.wco_frozen_set(self.s, mode)

    def substep_adjoint(self, dt, tol, maxit, x0, v0, C0, xb, vb, Cb):
        """dL/d(step inputs) from dL/d(step outputs) and dL/d|g| of this step."""
        x0, v0, C0 = (np.ascontiguousarray(a, dtype=np.float64) for a in (x0, v0, C0))
        xb, vb, Cb = (np.array(a, dtype=np.float64, order="C") for a in (xb, vb, Cb))
        gb = C.c_double(0.0)
        ok = self.L.wco_substep_adjoint(self.s, dt, tol, maxit, ptr(x0), ptr(v0), ptr(C0),
                                        ptr(xb), ptr(vb), ptr(Cb), C.byref(gb))
        return bool(ok), xb, vb, Cb, gb.value

    def set_elasto(self, x, m, fm, fvol, rest_vol, rest_vf, ori, sf, prm=None):
        """The static porous solid: ne particles (SoA x and orientation)."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64)
                for a in (x, m, fm, fvol, rest_vol, rest_vf, ori, sf)]
        self._elasto = arrs  # kept alive: the oracle copies them anyway
        self._mixprm = prm if prm is not None else mix_params()
        ne = arrs[0].shape[1] if arrs[0].ndim == 2 else arrs[0].size // 3
        self.L.wco_set_elasto(self.s, ne, *(ptr(a) for a in arrs), C.byref(self._mixprm))

    def mix_field(self, name, axis=0):
        p = self.L.wco_mix_field(self.s, name.encode(), axis)
        if not p:
            raise KeyError(name)
        return self._arr(p, self.st.g.nnode)

    def pre_solve(self, dt):
        """Phases up to and including assembly."""
        self.activate()
        self.update_liquid_phi()
        self.update_solid_phi()
        self.update_solid_weights()
        self.p2g()
        self.step_velocity(dt)
        return self.assemble(dt)

    @property
    def last_solve(self):
        st = self.st
        return bool(st.last_ok), int(st.last_iters), float(st.last_residual)


# ---------------------------------------------------------------- CSR helpers

def csr_args(rs, col, val):
    rs = np.ascontiguousarray(rs, dtype=np.uint32)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    return rs, col, val


def spmv(rs, col, val, x):
    rs, col, val = csr_args(rs, col, val)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(len(rs) - 1)
    lib().wco_spmv(len(rs) - 1, ptr(rs, _u32p), ptr(col, _u32p), ptr(val), ptr(x), ptr(y))
    return y


def mic0(rs, col, val, tau=0.97, mdr=0.25):
    rs, col, val = csr_args(rs, col, val)
    n = len(rs) - 1
    invdiag = np.zeros(n)
    lv = np.zeros(max(len(val), 1))
    lib().wco_mic0_factor(n, ptr(rs, _u32p), ptr(col, _u32p), ptr(val), tau, mdr,
                          ptr(invdiag), ptr(lv))
    return invdiag, lv[: len(val)]


def apply_precond(rs, col, invdiag, lv, r):
    rs, col, _ = csr_args(rs, col, lv)
    n = len(rs) - 1
    r = np.ascontiguousarray(r, dtype=np.float64)
    lv = np.ascontiguousarray(lv, dtype=np.float64)
    invdiag = np.ascontiguousarray(invdiag, dtype=np.float64)
    z = np.zeros(n)
    lib().wco_solve_lower(n, ptr(rs, _u32p), ptr(col, _u32p), ptr(invdiag), ptr(lv),
                          ptr(r), ptr(z))
    lib().wco_solve_lower_transpose(n, ptr(rs, _u32p), ptr(col, _u32p), ptr(invdiag),
                                    ptr(lv), ptr(z))
    return z


def solve_lower(rs, col, invdiag, lv, r):
    rs, col, _ = csr_args(rs, col, lv)
    n = len(rs) - 1
    z = np.zeros(n)
    r = np.ascontiguousarray(r, dtype=np.float64)
    lib().wco_solve_lower(n, ptr(rs, _u32p), ptr(col, _u32p),
                          ptr(np.ascontiguousarray(invdiag, dtype=np.float64)),
                          ptr(np.ascontiguousarray(lv, dtype=np.float64)), ptr(r), ptr(z))
    return z


def pcg(rs, col, val, b, tol=1e-8, maxit=500, tau=0.97, mdr=0.25):
    rs, col, val = csr_args(rs, col, val)
    n = len(rs) - 1
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(n)
    res = C.c_double(0)
    it = C.c_int(0)
    ok = lib().wco_pcg_solve(n, ptr(rs, _u32p), ptr(col, _u32p), ptr(val), ptr(b), ptr(x),
                             tol, maxit, tau, mdr, C.byref(res), C.byref(it))
    return bool(ok), x, it.value, res.value


# ------------------------------------------------------------ AMG-PCG

def _amg_args(rs, col, val, ijk):
    rs = np.ascontiguousarray(rs, dtype=np.uint32)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    ijk = np.ascontiguousarray(ijk, dtype=np.int32)
    return rs, col, val, ijk


def amg_levels(rs, col, val, ijk, dims):
    rs, col, val, ijk = _amg_args(rs, col, val, ijk)
    sizes = np.zeros(32, dtype=np.int64)
    nl = lib().wco_amg_levels(len(rs) - 1, ptr(rs, _u32p), ptr(col, _u32p), ptr(val),
                              ptr(ijk, C.POINTER(C.c_int32)), *dims,
                              sizes.ctypes.data_as(C.POINTER(C.c_int64)), 32)
    return list(sizes[:nl])


def amg_coarse(rs, col, val, ijk, dims):
    """One coarsening step: (parent, coarse rowstart, col, val)."""
    rs, col, val, ijk = _amg_args(rs, col, val, ijk)
    n = len(rs) - 1
    cap = len(col)
    parent = np.zeros(n, dtype=np.uint32)
    crs = np.zeros(n + 1, dtype=np.uint32)
    ccol = np.zeros(cap, dtype=np.uint32)
    cval = np.zeros(cap)
    nc = lib().wco_amg_coarse(n, ptr(rs, _u32p), ptr(col, _u32p), ptr(val),
                              ptr(ijk, C.POINTER(C.c_int32)), *dims, ptr(parent, _u32p),
                              ptr(crs, _u32p), ptr(ccol, _u32p), ptr(cval), cap)
    nnz = int(crs[nc])
    return parent, crs[:nc + 1].copy(), ccol[:nnz].copy(), cval[:nnz].copy()


def amg_rbgs(rs, col, val, ijk, b, x, iters):
    rs, col, val, ijk = _amg_args(rs, col, val, ijk)
    x = np.array(x, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    lib().wco_amg_rbgs(len(rs) - 1, ptr(rs, _u32p), ptr(col, _u32p), ptr(val),
                       ptr(ijk, C.POINTER(C.c_int32)), ptr(b), ptr(x), iters)
    return x


def amg_vcycle(rs, col, val, ijk, dims, b):
    rs, col, val, ijk = _amg_args(rs, col, val, ijk)
    b = np.ascontiguousarray(b, dtype=np.float64)
    z = np.zeros(len(rs) - 1)
    lib().wco_amg_vcycle(len(rs) - 1, ptr(rs, _u32p), ptr(col, _u32p), ptr(val),
                         ptr(ijk, C.POINTER(C.c_int32)), *dims, ptr(b), ptr(z))
    return z


def amgpcg(rs, col, val, ijk, dims, b, tol=1e-8, maxit=500):
    rs, col, val, ijk = _amg_args(rs, col, val, ijk)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(len(rs) - 1)
    res = C.c_double(0)
    it = C.c_int(0)
    ok = lib().wco_amgpcg_solve(len(rs) - 1, ptr(rs, _u32p), ptr(col, _u32p), ptr(val),
                                ptr(ijk, C.POINTER(C.c_int32)), *dims, ptr(b), ptr(x), tol,
                                maxit, C.byref(res), C.byref(it))
    return bool(ok), x, it.value, res.value


# ------------------------------------------------------------ scene configs

def dam_break(n: int, block_hi=(0.4, 0.6, 1.0), seed: int = 0, sigma: float = 0.0):
    """Particles of the dam-break block at n^3 (dx = 1/n): (grid, x, v)."""
    dx = 1.0 / n
    g = grid((n, n, n), dx)
    x = seed_block(g, (0.0, 0.0, 0.0), block_hi, seed)
    v = seed_velocity(g, seed, sigma, x) if sigma > 0 else np.zeros_like(x)
    return g, x, v


# ------------------------------------------------------------ the adjoint

class AdjointRun:
    """A run of `steps` substeps whose decisions are recorded on the first
    forward pass and replayed afterwards, with the loss sum |x - target|^2 of
    the final positions and its gradient by the adjoint (wco_substep_adjoint)
    with respect to the initial velocities and the gravity magnitude."""

    def __init__(self, n, dx, x0, v0, C0, steps, dt, g=9.81, tol=1e-13, L=None):
        self.n, self.dx, self.steps, self.dt, self.tol = n, dx, steps, dt, tol
        self.x0, self.C0 = x0, C0
        self.scene = OracleScene(n, dx, x0, v0, C0, gravity=(0.0, -g, 0.0), L=L)
        self.scene.frozen_mode(1)
        self.states, self.records = [], []
        for _ in range(steps):
            self.states.append(self.scene.particles())
            assert self.scene.substep(dt, tol, 5000)
            self.records.append(self.scene.frozen())
        self.scene.frozen_mode(2)

    def forward(self, v0, g=None):
        """Final positions from initial velocities v0 (and gravity g), with
        the recorded decisions."""
        s = self.scene
        if g is not None:
            s.st.gravity[1] = -g
        s.set_particles(self.x0, v0, self.C0)
        for k in range(self.steps):
            s.set_frozen(self.records[k])
            assert s.substep(self.dt, self.tol, 5000)
        return s.particles()[0]

    def gradient(self, v0, target, g=None):
        """(loss, dL/dv0, dL/d|g|)."""
        s = self.scene
        xf = self.forward(v0, g)
        loss = float(((xf - target) ** 2).sum())
        xb = 2.0 * (xf - target)
        vb = np.zeros_like(xb)
        Cb = np.zeros((9, xb.shape[1]))
        gsum = 0.0
        states = [None] * self.steps
        # forward again, keeping each step's inputs (the stored record states
        # belong to the recorded velocities, not to v0)
        s.set_particles(self.x0, v0, self.C0)
        for k in range(self.steps):
            states[k] = s.particles()
            s.set_frozen(self.records[k])
            assert s.substep(self.dt, self.tol, 5000)
        for k in range(self.steps - 1, -1, -1):
            x, v, Cm = states[k]
            s.set_particles(x, v, Cm)
            s.set_frozen(self.records[k])
            assert s.substep(self.dt, self.tol, 5000)
            ok, xb, vb, Cb, gb = s.substep_adjoint(self.dt, self.tol, 5000, x, v, Cm, xb, vb, Cb)
            assert ok
            gsum += gb
        return loss, vb, gsum
