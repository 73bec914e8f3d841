"""libwetcloth_amd: MI355X (gfx950) liquid substep of libWetCloth.

Python binding of the C ABI in include/wetcloth_amd.h (libwcgpu.so, built
in-tree by libwetcloth_amd/build.py).  The binding never falls back to a CPU
implementation: if the shared library is missing or no GPU is present the
calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwcgpu.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "wetcloth_amd.h")

WC_OK = 0
WC_NOT_CONVERGED = 1
WC_FP64 = 0
WC_MIXED = 1

FIELD = {
    "act": 0, "lphi": 1, "w": 2, "mass": 3, "vel": 4, "vol": 5, "rhs": 6,
    "ustar": 7, "unew": 8, "dof": 9, "pressure": 10, "valid": 11,
    # with a solid (Scene.set_elasto)
    "psi": 12, "sat": 13, "smass": 14, "svol": 15, "sf": 16, "ori_x": 17, "ori_y": 18,
    "ori_z": 19, "Q": 20, "hdv": 21, "psi_p": 22, "sat_p": 23, "pp": 24,
}
_FIELD_DTYPE = {"act": np.uint8, "dof": np.int32, "valid": np.uint8}

KERNELS = ("sweep_fwd", "sweep_bwd", "spmv", "p2g", "g2p")  # WC_KERNEL_* order
PHASES = ("rebucketize", "p2g", "step_velocity", "assemble", "solve", "apply_pressure", "g2p")

_dp = C.POINTER(C.c_double)
_u32p = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)


class WcDesc(C.Structure):
    _fields_ = [
        ("ni", C.c_int32), ("nj", C.c_int32), ("nk", C.c_int32),
        ("dx", C.c_double),
        ("liquid_density", C.c_double),
        ("flip_coeff", C.c_double),
        ("liquid_boundary_friction", C.c_double),
        ("gravity", C.c_double * 3),
        ("particle_mass", C.c_double),
        ("particle_volume", C.c_double),
        ("particle_radius", C.c_double),
        ("solid_projection", C.c_int32),
        ("precision", C.c_int32),
        ("max_particles", C.c_int64),
    ]


class MixParams(C.Structure):
    """wc_mix_params: the LiquidInfo fields of the mixture terms."""
    _fields_ = [("use_drag", C.c_int32), ("use_nonlinear_drag", C.c_int32),
                ("yarn_diameter", C.c_double), ("viscosity", C.c_double),
                ("yazdchi_power", C.c_double), ("pore_radius", C.c_double),
                ("surf_tension_coeff", C.c_double), ("rest_contact_angle", C.c_double)]


def mix_params(**kw) -> MixParams:
    """LiquidInfo's mixture fields in SI units: the parser's defaults
    (TwoDSceneXMLParser.cpp:1849-1904, CGS there) converted, any of them
    overridden by keyword; a rest_volume_fraction keyword derives the pore
    radius from the yarn diameter as the parser does (:1960-1963)."""
    p = MixParams()
    p.use_drag = int(kw.get("use_drag", 1))
    p.use_nonlinear_drag = int(kw.get("use_nonlinear_drag", 1))
    p.yarn_diameter = float(kw.get("yarn_diameter", 1e-4))
    p.viscosity = float(kw.get("viscosity", 8.9e-4))
    p.yazdchi_power = float(kw.get("yazdchi_power", 1.6))
    if "rest_volume_fraction" in kw:
        rvf = float(kw["rest_volume_fraction"])
        p.pore_radius = p.yarn_diameter * 0.5 * float(np.sqrt(max(0.0, (1.0 - rvf) / rvf)))
    else:
        p.pore_radius = float(kw.get("pore_radius", 5e-5))
    p.surf_tension_coeff = float(kw.get("surf_tension_coeff", 0.0728))
    p.rest_contact_angle = float(kw.get("rest_contact_angle", 40.8 / 180.0 * np.pi))
    return p


class WetclothError(RuntimeError):
    pass


class NotConverged(WetclothError):
    pass


_lib = None


def lib():
    """Load libwcgpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise WetclothError(f"{LIB_PATH} is missing: run __graft_entry__.build() "
                                "or python libwetcloth_amd/build.py")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.wc_create.argtypes = [C.POINTER(WcDesc), C.POINTER(vp)]
        L.wc_destroy.argtypes = [vp]
        L.wc_destroy.restype = None
        L.wc_last_error.argtypes = [vp]
        L.wc_last_error.restype = C.c_char_p
        for nm in ("wc_set_particles", "wc_set_particles_soa"):
            getattr(L, nm).argtypes = [vp, C.c_int64, _dp, _dp, _dp]
        for nm in ("wc_get_particles", "wc_get_particles_soa"):
            getattr(L, nm).argtypes = [vp, _dp, _dp, _dp]
        L.wc_num_particles.argtypes = [vp, _i64p]
        L.wc_seed_block.argtypes = [vp, _dp, _dp, C.c_uint64, C.c_double, _i64p]
        L.wc_rebucketize.argtypes = [vp]
        L.wc_map_particle_nodes.argtypes = [vp]
        L.wc_step_velocity.argtypes = [vp, C.c_double]
        L.wc_project.argtypes = [vp, C.c_double, C.c_double, C.c_int32, _dp, _i32p]
        L.wc_apply_pressure.argtypes = [vp, C.c_double]
        L.wc_map_node_particles.argtypes = [vp]
        L.wc_advect_scene.argtypes = [vp, C.c_double]
        L.wc_solid_projection.argtypes = [vp, C.c_double]
        L.wc_map_node_particles_advect.argtypes = [vp, C.c_double]
        L.wc_max_fluid_velocity.argtypes = [vp, _dp]
        L.wc_substep.argtypes = [vp, C.c_double, C.c_double, C.c_int32, _dp, _i32p]
        L.wc_step.argtypes = [vp, C.c_double, C.c_double, C.c_int32, _dp, _i32p]
        L.wc_step_substeps.argtypes = [vp, _i32p]
        L.wc_grid_info.argtypes = [vp, _i32p, _i64p, _i64p]
        L.wc_download_field.argtypes = [vp, C.c_int32, C.c_int32, vp]
        L.wc_download_system.argtypes = [vp, _i64p, _i64p, _u32p, _u32p, _dp, _dp, _dp]
        L.wc_phase_times.argtypes = [vp, _dp]
        L.wc_set_option.argtypes = [vp, C.c_char_p, C.c_int64]
        L.wc_kernel_stats.argtypes = [vp, C.c_int32, _dp, _i64p, _i64p]
        L.wc_refinement_passes.argtypes = [vp, _i32p]
        for nm in ("wc_serialize_position_only", "wc_load_position_only", "wc_serialize_fluid"):
            getattr(L, nm).argtypes = [vp, C.c_char_p]
        L.wc_apply_preconditioner.argtypes = [vp, _dp, _dp]
        L.wc_system_multiply.argtypes = [vp, _dp, _dp]
        L.wc_pcg_solve_csr.argtypes = [C.c_int64, _u32p, _u32p, _dp, _dp, _dp, C.c_double,
                                       C.c_int32, _dp, _i32p]
        L.wc_pcg_solve_csr_noalloc.argtypes = L.wc_pcg_solve_csr.argtypes
        L.wc_set_gravity.argtypes = [vp, _dp]
        L.wc_set_elasto_particles.argtypes = [vp, C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                              _dp, C.POINTER(MixParams)]
        L.wc_set_frozen.argtypes = [vp, C.c_int32, C.c_int32]
        L.wc_checkpoint.argtypes = [vp, C.c_int32, C.c_int32]
        L.wc_frozen_departures.argtypes = [vp, _i64p]
        L.wc_substep_adjoint.argtypes = [vp, C.c_double, C.c_double, C.c_int32, _dp, _dp, _dp,
                                         _dp, _dp, _i32p]
        L.wc_rollout_forward.argtypes = [vp, C.c_int32, C.c_double, C.c_double, C.c_int32,
                                         C.c_int32, C.c_int32, _i32p]
        L.wc_rollout_backward.argtypes = [vp, C.c_double, C.c_double, C.c_int32, _dp, _dp, _dp,
                                          _dp, _i32p]
        L.wc_mic0_factor_csr.argtypes = [C.c_int64, _u32p, _u32p, _dp, C.c_double, C.c_double,
                                         _dp, _dp]
        L.wc_mic0_apply_csr.argtypes = [C.c_int64, _u32p, _u32p, _dp, _dp, _dp, _dp]
        L.wc_spmv_csr.argtypes = [C.c_int64, _u32p, _u32p, _dp, _dp, _dp]
        _i32p_ = C.POINTER(C.c_int32)
        L.wc_amgpcg_solve_csr.argtypes = [C.c_int64, _u32p, _u32p, _dp, _i32p_, C.c_int32,
                                          C.c_int32, C.c_int32, _dp, _dp, C.c_double,
                                          C.c_int32, _dp, _i32p]
        L.wc_amg_apply_csr.argtypes = [C.c_int64, _u32p, _u32p, _dp, _i32p_, C.c_int32,
                                       C.c_int32, C.c_int32, _dp, _dp]
        _lib = L
    return _lib


def declared_symbols() -> list[str]:
    """Entry points declared in include/wetcloth_amd.h."""
    import re
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:wc_status|void|const char\*)\s+(wc_\w+)\s*\(",
                                 txt, flags=re.M)))


def _ptr(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


def _check(st, ctx=None, allow_nc=False):
    if st == WC_OK or (allow_nc and st == WC_NOT_CONVERGED):
        return st
    msg = lib().wc_last_error(ctx).decode()
    if st == WC_NOT_CONVERGED:
        raise NotConverged(msg or "PCG did not converge")
    raise WetclothError(f"status {st}: {msg}")


class Scene:
    """A liquid scene resident on the GPU, stepped phase by phase in the
    reference's substep order (Core/WetClothCore.cpp:77-300)."""

    def __init__(self, n, dx, rho=1000.0, flip=0.99, gravity=(0.0, -9.81, 0.0),
                 pvol=None, radius=None, solid_projection=True, precision=WC_FP64,
                 friction=1.0):
        n = tuple(int(a) for a in n)
        pvol = dx ** 3 / 8.0 if pvol is None else pvol
        radius = (3.0 * pvol / (4.0 * np.pi)) ** (1.0 / 3.0) if radius is None else radius
        d = WcDesc()
        d.ni, d.nj, d.nk = n
        d.dx = dx
        d.liquid_density = rho
        d.flip_coeff = flip
        d.liquid_boundary_friction = friction
        for i in range(3):
            d.gravity[i] = gravity[i]
        d.particle_mass = rho * pvol
        d.particle_volume = pvol
        d.particle_radius = radius
        d.solid_projection = 1 if solid_projection else 0
        d.precision = precision
        d.max_particles = 0
        self.desc = d
        self.n = n
        self.dx = dx
        h = C.c_void_p()
        _check(lib().wc_create(C.byref(d), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().wc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- particles ----
    def set_particles(self, x, v=None, Cm=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        n = x.shape[1]
        v = None if v is None else np.ascontiguousarray(v, dtype=np.float64)
        Cm = None if Cm is None else np.ascontiguousarray(Cm, dtype=np.float64)
        _check(lib().wc_set_particles_soa(self.h, n, _ptr(x), _ptr(v), _ptr(Cm)), self.h)

    def set_particles_reference_layout(self, x4, v4=None, fB=None):
        n = len(x4) // 4
        _check(lib().wc_set_particles(self.h, n, _ptr(x4), _ptr(v4), _ptr(fB)), self.h)

    def num_particles(self) -> int:
        n = C.c_int64()
        _check(lib().wc_num_particles(self.h, C.byref(n)), self.h)
        return n.value

    def particles(self):
        n = self.num_particles()
        x = np.zeros((3, n))
        v = np.zeros((3, n))
        Cm = np.zeros((9, n))
        _check(lib().wc_get_particles_soa(self.h, _ptr(x), _ptr(v), _ptr(Cm)), self.h)
        return x, v, Cm

    def particles_reference_layout(self):
        n = self.num_particles()
        x4 = np.zeros(4 * n)
        v4 = np.zeros(4 * n)
        fB = np.zeros(9 * n)
        _check(lib().wc_get_particles(self.h, _ptr(x4), _ptr(v4), _ptr(fB)), self.h)
        return x4, v4, fB

    def seed_block(self, lo, hi, seed=0, sigma=0.0) -> int:
        lo = np.asarray(lo, dtype=np.float64)
        hi = np.asarray(hi, dtype=np.float64)
        n = C.c_int64()
        _check(lib().wc_seed_block(self.h, _ptr(lo), _ptr(hi), C.c_uint64(seed), sigma,
                                   C.byref(n)), self.h)
        return n.value

    # ---- the reference serializer's particle files (TwoDSceneSerializer.cpp) ----
    def serializePositionOnly(self, path):
        """TwoDSceneSerializer::serializePositionOnly (:161-170, :23-37)."""
        _check(lib().wc_serialize_position_only(self.h, os.fsencode(path)), self.h)

    def loadPosOnly(self, path):
        """TwoDSceneSerializer::loadPosOnly (:172-183): positions only."""
        _check(lib().wc_load_position_only(self.h, os.fsencode(path)), self.h)

    def serializeFluid(self, path):
        """The fluid file of TwoDSceneSerializer::serializeScene (:39-47, :242-255)."""
        _check(lib().wc_serialize_fluid(self.h, os.fsencode(path)), self.h)

    # ---- substep phases (reference names) ----
    def rebucketizeParticles(self):
        _check(lib().wc_rebucketize(self.h), self.h)

    def mapParticleNodesAPIC(self):
        _check(lib().wc_map_particle_nodes(self.h), self.h)

    def stepVelocity(self, dt):
        _check(lib().wc_step_velocity(self.h, dt), self.h)

    def projectFine(self, dt, tol=1e-8, maxit=1000):
        res = C.c_double()
        it = C.c_int32()
        st = lib().wc_project(self.h, dt, tol, maxit, C.byref(res), C.byref(it))
        _check(st, self.h, allow_nc=True)
        return st == WC_OK, it.value, res.value

    def applyPressureDragFluid(self, dt):
        _check(lib().wc_apply_pressure(self.h, dt), self.h)

    def mapNodeParticlesAPIC(self):
        """TwoDScene::mapNodeParticlesAPIC(): particle v and C from the grid."""
        _check(lib().wc_map_node_particles(self.h), self.h)

    def advectScene(self, dt):
        """SceneStepper::advectScene: x += v dt."""
        _check(lib().wc_advect_scene(self.h, dt), self.h)

    def solidProjection(self, dt):
        """TwoDScene::solidProjection(dt)."""
        _check(lib().wc_solid_projection(self.h, dt), self.h)

    def map_node_particles_advect(self, dt):
        """mapNodeParticlesAPIC, advectScene and solidProjection in one pass."""
        _check(lib().wc_map_node_particles_advect(self.h, dt), self.h)

    def getMaxFluidVelocity(self) -> float:
        v = C.c_double()
        _check(lib().wc_max_fluid_velocity(self.h, C.byref(v)), self.h)
        return v.value

    def substep(self, dt, tol=1e-8, maxit=1000):
        """One substep of dt: (converged, PCG iterations, residual)."""
        res = C.c_double()
        it = C.c_int32()
        st = lib().wc_substep(self.h, dt, tol, maxit, C.byref(res), C.byref(it))
        _check(st, self.h, allow_nc=True)
        return st == WC_OK, it.value, res.value

    def step(self, dt, tol=1e-8, maxit=1000):
        """WetClothCore::stepSystem(dt): the CFL-split substeps of dt;
        (all converged, PCG iterations summed, last residual)."""
        res = C.c_double()
        it = C.c_int32()
        st = lib().wc_step(self.h, dt, tol, maxit, C.byref(res), C.byref(it))
        _check(st, self.h, allow_nc=True)
        return st == WC_OK, it.value, res.value

    stepSystem = step

    def step_substeps(self) -> int:
        """Substeps the last step() split its dt into."""
        n = C.c_int32()
        _check(lib().wc_step_substeps(self.h, C.byref(n)), self.h)
        return n.value

    def set_gravity(self, g):
        """The gravity vector of the following substeps."""
        g = np.ascontiguousarray(g, dtype=np.float64)
        _check(lib().wc_set_gravity(self.h, _ptr(g)), self.h)

    def set_elasto(self, x, m, fm, fvol, rest_vol, rest_vf, ori, sf, prm=None):
        """The resting, unsolved solid (the cloth's vertices): ne particles
        with positions x (3, ne), mass m, carried liquid mass fm and volume
        fvol, rest volume and rest volume fraction, orientation (3, ne) and
        shape factor; prm = mix_params().  x = None removes the solid."""
        if x is None:
            _check(lib().wc_set_elasto_particles(self.h, 0, *([None] * 8), None), self.h)
            return
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(3, -1)
        ne = x.shape[1]
        per = [np.ascontiguousarray(a, dtype=np.float64).reshape(ne)
               for a in (m, fm, fvol, rest_vol, rest_vf)]
        ori = np.ascontiguousarray(ori, dtype=np.float64).reshape(3, ne)
        sf = np.ascontiguousarray(sf, dtype=np.float64).reshape(ne)
        self._mixprm = mix_params() if prm is None else prm
        _check(lib().wc_set_elasto_particles(self.h, ne, _ptr(x), *(_ptr(a) for a in per),
                                             _ptr(ori), _ptr(sf), C.byref(self._mixprm)),
               self.h)

    # ---- the adjoint (no reference counterpart) ----
    def set_frozen(self, mode, slot=0):
        """Record (1) the next substeps' decisions into slot, replay (2) a
        recorded slot, or neither (0)."""
        _check(lib().wc_set_frozen(self.h, int(mode), int(slot)), self.h)

    def frozen_departures(self) -> int:
        """Replayed decisions the run would have taken differently."""
        n = C.c_int64()
        _check(lib().wc_frozen_departures(self.h, C.byref(n)), self.h)
        return n.value

    def checkpoint(self, slot, restore=False):
        """Save the particle state to a device slot, or restore it."""
        _check(lib().wc_checkpoint(self.h, int(slot), 1 if restore else 0), self.h)

    def _grads(self, xb, vb, Cb):
        n = self.num_particles()
        out = []
        for a, rows in ((xb, 3), (vb, 3), (Cb, 9)):
            out.append(np.zeros((rows, n)) if a is None
                       else np.array(a, dtype=np.float64, order="C").reshape(rows, n))
        return out

    def substep_adjoint(self, dt, xb, vb=None, Cb=None, tol=1e-12, maxit=5000):
        """Adjoint of the substep just run in mode 1 or 2: from dL/d(x, v, C)
        after it, (converged, dL/dx, dL/dv, dL/dC before it, dL/d|g|)."""
        xb, vb, Cb = self._grads(xb, vb, Cb)
        gb = C.c_double(0.0)
        res = C.c_double()
        it = C.c_int32()
        st = lib().wc_substep_adjoint(self.h, dt, tol, maxit, _ptr(xb), _ptr(vb), _ptr(Cb),
                                      C.byref(gb), C.byref(res), C.byref(it))
        _check(st, self.h, allow_nc=True)
        return st == WC_OK, xb, vb, Cb, gb.value

    def rollout_forward(self, steps, dt, tol=1e-12, maxit=5000, every=10, replay=False):
        """`steps` substeps recording their decisions (checkpoints every
        `every`), or replaying the recorded ones: (converged, iterations)."""
        it = C.c_int32()
        st = lib().wc_rollout_forward(self.h, int(steps), dt, tol, maxit, int(every),
                                      1 if replay else 0, C.byref(it))
        _check(st, self.h, allow_nc=True)
        return st == WC_OK, it.value

    def rollout_backward(self, dt, xb, vb=None, Cb=None, tol=1e-12, maxit=5000):
        """From dL/d(final x, v, C): (converged, dL/dx0, dL/dv0, dL/dC0, dL/d|g|)."""
        xb, vb, Cb = self._grads(xb, vb, Cb)
        gb = C.c_double(0.0)
        it = C.c_int32()
        st = lib().wc_rollout_backward(self.h, dt, tol, maxit, _ptr(xb), _ptr(vb), _ptr(Cb),
                                       C.byref(gb), C.byref(it))
        _check(st, self.h, allow_nc=True)
        return st == WC_OK, xb, vb, Cb, gb.value

    # ---- diagnostics ----
    def grid_info(self):
        nx = (C.c_int32 * 3)()
        nn = C.c_int64()
        nb = C.c_int64()
        _check(lib().wc_grid_info(self.h, nx, C.byref(nn), C.byref(nb)), self.h)
        return tuple(nx), nn.value, nb.value

    def field(self, name, axis=0):
        _, nn, nb = self.grid_info()
        dt = _FIELD_DTYPE.get(name, np.float64)
        out = np.zeros(nb if name == "act" else nn, dtype=dt)
        _check(lib().wc_download_field(self.h, FIELD[name], axis,
                                       out.ctypes.data_as(C.c_void_p)), self.h)
        return out

    def system(self):
        n = C.c_int64()
        nnz = C.c_int64()
        L = lib()
        _check(L.wc_download_system(self.h, C.byref(n), C.byref(nnz), None, None, None, None,
                                    None), self.h)
        rs = np.zeros(n.value + 1, dtype=np.uint32)
        col = np.zeros(nnz.value, dtype=np.uint32)
        val = np.zeros(nnz.value)
        b = np.zeros(n.value)
        p = np.zeros(n.value)
        _check(L.wc_download_system(self.h, C.byref(n), C.byref(nnz), _ptr(rs, _u32p),
                                    _ptr(col, _u32p), _ptr(val), _ptr(b), _ptr(p)), self.h)
        return rs, col, val, b, p

    def set_option(self, name: str, value: int):
        _check(lib().wc_set_option(self.h, name.encode(), int(value)), self.h)

    def apply_preconditioner(self, r):
        """MIC(0) apply (PCGSolver.h:326-329) with the factor of the last
        projectFine, on the kernels it used; DOF-ordered vectors."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros_like(r)
        _check(lib().wc_apply_preconditioner(self.h, _ptr(r), _ptr(z)), self.h)
        return z

    def multiply(self, x):
        """y = A x for the last assembled pressure system."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        _check(lib().wc_system_multiply(self.h, _ptr(x), _ptr(y)), self.h)
        return y

    def refinement_passes(self) -> int:
        """Refinement passes of the last mixed-precision solve (0 in FP64)."""
        p = C.c_int32()
        _check(lib().wc_refinement_passes(self.h, C.byref(p)), self.h)
        return int(p.value)

    def kernel_stats(self) -> dict:
        """Per kernel class since set_option("kernel_timing", ...): {name:
        (total ms, launches, units processed)}; units are DOFs for the solver
        kernels and particles for "p2g" and "g2p"."""
        out = {}
        for k, name in enumerate(KERNELS):
            ms = C.c_double(0)
            nl = C.c_int64(0)
            nu = C.c_int64(0)
            _check(lib().wc_kernel_stats(self.h, k, C.byref(ms), C.byref(nl), C.byref(nu)),
                   self.h)
            out[name] = (ms.value, nl.value, nu.value)
        return out

    def phase_times(self) -> dict:
        ms = (C.c_double * 7)()
        _check(lib().wc_phase_times(self.h, ms), self.h)
        return dict(zip(PHASES, list(ms)))


# ---- robertbridson::PCGSolver<double> drop-in ----

def _csr(rs, col):
    return (np.ascontiguousarray(rs, dtype=np.uint32), np.ascontiguousarray(col, dtype=np.uint32))


def pcg_solve_csr(rs, col, val, rhs, tol=1e-8, maxit=500, no_alloc=False):
    """PCGSolver<double>::solve: returns (converged, x, iterations, residual).
    no_alloc=True calls wc_pcg_solve_csr_noalloc (the library's fixed device
    array instead of allocations; systems beyond it raise)."""
    rs, col = _csr(rs, col)
    val = np.ascontiguousarray(val, dtype=np.float64)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    n = len(rs) - 1
    x = np.zeros(n)
    res = C.c_double()
    it = C.c_int32()
    fn = lib().wc_pcg_solve_csr_noalloc if no_alloc else lib().wc_pcg_solve_csr
    st = fn(n, _ptr(rs, _u32p), _ptr(col, _u32p), _ptr(val), _ptr(rhs), _ptr(x), tol, maxit,
            C.byref(res), C.byref(it))
    _check(st, None, allow_nc=True)
    return st == WC_OK, x, it.value, res.value


def mic0_factor_csr(rs, col, val, tau=0.97, mdr=0.25):
    rs, col = _csr(rs, col)
    val = np.ascontiguousarray(val, dtype=np.float64)
    n = len(rs) - 1
    invdiag = np.zeros(n)
    lv = np.zeros(len(val))
    _check(lib().wc_mic0_factor_csr(n, _ptr(rs, _u32p), _ptr(col, _u32p), _ptr(val), tau, mdr,
                                    _ptr(invdiag), _ptr(lv)))
    return invdiag, lv


def mic0_apply_csr(rs, col, invdiag, lv, r):
    rs, col = _csr(rs, col)
    n = len(rs) - 1
    z = np.zeros(n)
    _check(lib().wc_mic0_apply_csr(n, _ptr(rs, _u32p), _ptr(col, _u32p),
                                   _ptr(np.ascontiguousarray(invdiag, dtype=np.float64)),
                                   _ptr(np.ascontiguousarray(lv, dtype=np.float64)),
                                   _ptr(np.ascontiguousarray(r, dtype=np.float64)), _ptr(z)))
    return z


def spmv_csr(rs, col, val, x):
    rs, col = _csr(rs, col)
    n = len(rs) - 1
    y = np.zeros(n)
    _check(lib().wc_spmv_csr(n, _ptr(rs, _u32p), _ptr(col, _u32p),
                             _ptr(np.ascontiguousarray(val, dtype=np.float64)),
                             _ptr(np.ascontiguousarray(x, dtype=np.float64)), _ptr(y)))
    return y


def amgpcg_solve_csr(rs, col, val, dof_ijk, dims, rhs, tol=1e-8, maxit=500):
    """AMGPCGSolveSparse (AlgebraicMultigrid.h:247-359) on a host CSR:
    (converged, x, iterations, residual)."""
    rs, col = _csr(rs, col)
    val = np.ascontiguousarray(val, dtype=np.float64)
    ijk = np.ascontiguousarray(dof_ijk, dtype=np.int32)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    n = len(rs) - 1
    x = np.zeros(n)
    res = C.c_double(0)
    it = C.c_int32(0)
    st = lib().wc_amgpcg_solve_csr(n, _ptr(rs, _u32p), _ptr(col, _u32p), _ptr(val),
                                   _ptr(ijk, C.POINTER(C.c_int32)), *dims, _ptr(rhs), _ptr(x),
                                   tol, maxit, C.byref(res), C.byref(it))
    _check(st, None, allow_nc=True)
    return st == WC_OK, x, it.value, res.value


def amg_apply_csr(rs, col, val, dof_ijk, dims, r):
    """amgPrecondCompressed: z = one V-cycle applied to r from z = 0."""
    rs, col = _csr(rs, col)
    val = np.ascontiguousarray(val, dtype=np.float64)
    ijk = np.ascontiguousarray(dof_ijk, dtype=np.int32)
    r = np.ascontiguousarray(r, dtype=np.float64)
    z = np.zeros(len(rs) - 1)
    _check(lib().wc_amg_apply_csr(len(rs) - 1, _ptr(rs, _u32p), _ptr(col, _u32p), _ptr(val),
                                  _ptr(ijk, C.POINTER(C.c_int32)), *dims, _ptr(r), _ptr(z)))
    return z
