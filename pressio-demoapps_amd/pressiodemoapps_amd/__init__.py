"""MI355X-native pressio-demoapps finite-volume engine: Python host mirror.

This module mirrors the reference Python API (pressiodemoapps/__init__.py and
src_py/main_binder.cc): ``load_cellcentered_uniform_mesh``, ``create_problem``
and the problem methods ``rightHandSide`` / ``jacobian`` /
``rightHandSideAndJacobian`` / ``applyJacobian`` / ``initialCondition``.
Everything evaluates through the C ABI of ``libpda.so`` (include/pda_c_api.h)
on HIP device memory; state, velocity and Jacobian values are torch tensors on
``cuda:0`` used as plain device buffers.  There is no CPU fallback: a missing
library or GPU raises.
"""
import ctypes
import enum
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(os.path.dirname(_HERE), "libpda.so")
_lib = None


class PdaError(RuntimeError):
    pass


def library_path():
    return _LIBPATH


class ImplicitStats(ctypes.Structure):
    """pda_implicit_stats"""
    _fields_ = [("newton_iters", ctypes.c_int), ("linear_iters", ctypes.c_int),
                ("converged", ctypes.c_int), ("residual_norm", ctypes.c_double),
                ("correction_norm", ctypes.c_double),
                ("last_linear_rel_residual", ctypes.c_double),
                ("initial_residual_norm", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def lib():
    """Load libpda.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIBPATH):
            raise PdaError(f"libpda.so not built at {_LIBPATH}; run __graft_entry__.build()")
        # torch owns the process's HIP runtime (device memory, streams); let it
        # initialise before the library's first HIP call so both use one
        # runtime instance
        try:
            import torch
            torch.cuda.is_available()
        except ImportError:
            pass
        L = ctypes.CDLL(_LIBPATH)
        i32, i64, dbl, vp = ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p
        P = ctypes.POINTER
        sig = {
            "pda_last_error": ([], ctypes.c_char_p),
            "pda_version": ([], ctypes.c_char_p),
            "pda_mesh_load": ([ctypes.c_char_p, P(vp)], i32),
            "pda_mesh_create_structured": ([i32, P(i64), P(dbl), i32, i32, P(vp)], i32),
            "pda_mesh_create_sample": ([vp, P(ctypes.c_int32), i64, P(vp)], i32),
            "pda_mesh_sample_gids": ([vp, P(ctypes.c_int32), P(ctypes.c_int32)], i32),
            "pda_mesh_write": ([vp, ctypes.c_char_p], i32),
            "pda_mesh_info": ([vp, P(i32), P(i32), P(i64), P(i64), P(dbl)], i32),
            "pda_mesh_graph": ([vp], P(ctypes.c_int32)),
            "pda_mesh_coords": ([vp, i32], P(dbl)),
            "pda_mesh_row_classes": ([vp, P(i64), P(i64)], i32),
            "pda_mesh_boundary_rows": ([vp, i32, P(ctypes.c_int32), P(i64)], i32),
            "pda_mesh_destroy": ([vp], None),
            "pda_problem_create": ([vp, i32, i32, i32, i32, P(ctypes.c_char_p), P(dbl), i32, P(vp)], i32),
            "pda_problem_destroy": ([vp], None),
            "pda_problem_sizes": ([vp, P(i32), P(i64), P(i64), P(i64)], i32),
            "pda_problem_parameter": ([vp, ctypes.c_char_p, P(dbl)], i32),
            "pda_initial_condition": ([vp, vp, vp], i32),
            "pda_initial_condition_host": ([vp, vp], i32),
            "pda_jacobian_pattern": ([vp, P(i64), P(i64), P(vp), P(vp)], i32),
            "pda_jacobian_pattern_host": ([vp, P(i64), P(ctypes.c_int32)], i32),
            "pda_velocity": ([vp, vp, dbl, vp, vp], i32),
            "pda_problem_autotune": ([vp, vp, vp, P(i32), P(dbl)], i32),
            "pda_problem_set_velocity_strip_height": ([vp, i32], i32),
            "pda_problem_set_jacobian_strip_height": ([vp, i32], i32),
            "pda_jacobian": ([vp, vp, dbl, vp, vp, vp], i32),
            "pda_apply_jacobian": ([vp, vp, dbl, vp, i64, i32, vp, vp], i32),
            "pda_jacobian_blocks": ([vp, P(i32), P(vp), P(vp)], i32),
            "pda_stepper_create": ([vp, i32, P(vp)], i32),
            "pda_stepper_advance": ([vp, vp, dbl, dbl, i64, i32, vp], i32),
            "pda_stepper_destroy": ([vp], None),
            "pda_implicit_create": ([vp, i32, P(vp)], i32),
            "pda_implicit_set_options": ([vp, i32, dbl, i32, dbl, i32], i32),
            "pda_implicit_step": ([vp, vp, vp, vp, dbl, dbl, vp, P(ImplicitStats)], i32),
            "pda_implicit_destroy": ([vp], None),
            "pda_guessnet_create": ([i32, i32, P(vp)], i32),
            "pda_guessnet_destroy": ([vp], None),
            "pda_guessnet_num_params": ([vp, P(i64)], i32),
            "pda_guessnet_set_params": ([vp, P(ctypes.c_float), i64], i32),
            "pda_guessnet_get_params": ([vp, P(ctypes.c_float), i64], i32),
            "pda_guessnet_set_scales": ([vp, dbl, P(dbl), dbl, dbl], i32),
            "pda_guessnet_predict": ([vp, vp, vp, P(vp), dbl, vp, vp], i32),
            "pda_guessnet_policy": ([vp, P(dbl), P(dbl), P(dbl)], i32),
            "pda_guessnet_gradients": ([vp, P(vp), dbl, P(dbl), P(ctypes.c_float), P(ctypes.c_float),
                                        P(ctypes.c_float), vp], i32),
            "pda_guessnet_train_step": ([vp, P(vp), dbl, P(dbl), ctypes.c_float, P(ctypes.c_float), vp], i32),
            "pda_batch_create": ([P(vp), i32, P(vp)], i32),
            "pda_batch_destroy": ([vp], None),
            "pda_batch_sizes": ([vp, P(i64), P(i64)], i32),
            "pda_batch_initial_condition": ([vp, vp, vp], i32),
            "pda_batch_velocity": ([vp, vp, dbl, vp, vp], i32),
            "pda_batch_autotune": ([vp, vp, vp, P(i32), P(dbl)], i32),
            "pda_batch_set_velocity_strip_height": ([vp, i32], i32),
            "pda_batch_advance": ([vp, i32, vp, dbl, dbl, i64, i32, vp], i32),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        msg = lib().pda_last_error().decode()
        raise PdaError(f"pda error {rc}: {msg}")


# ------------------------------------------------------------------ enums
class InviscidFluxReconstruction(enum.IntEnum):
    FirstOrder = 0
    Weno3 = 1
    Weno5 = 2


class ViscousFluxReconstruction(enum.IntEnum):
    """schemes_info.hpp: the viscous terms are second-order central differences."""
    FirstOrder = 0


class Euler1d(enum.IntEnum):
    PeriodicSmooth = 0
    Sod = 1
    Lax = 2
    ShuOsher = 3


class Euler2d(enum.IntEnum):
    PeriodicSmooth = 0
    KelvinHelmholtz = 1
    SedovFull = 2
    SedovSymmetry = 3
    Riemann = 4
    NormalShock = 5
    DoubleMachReflection = 6
    CrossShock = 7
    testingonlyneumann = 8


class Euler3d(enum.IntEnum):
    PeriodicSmooth = 0
    SedovSymmetry = 1


class Swe2d(enum.IntEnum):
    SlipWall = 0
    CustomBCs = 1


class AdvectionDiffusion2d(enum.IntEnum):
    BurgersPeriodic = 0
    BurgersOutflow = 1


class AdvectionDiffusionReaction2d(enum.IntEnum):
    ProblemA = 0


FAMILY = {Euler1d: 1, Euler2d: 2, Euler3d: 3, Swe2d: 4,
          AdvectionDiffusion2d: 5, AdvectionDiffusionReaction2d: 6}


# ------------------------------------------------------------------- mesh
class CellCenteredUniformMesh:
    """Cell-centred uniform mesh (impl/mesh_ccu.hpp:60-473)."""

    def __init__(self, handle):
        L = lib()
        if isinstance(handle, (str, os.PathLike)):
            # CellCenteredUniformMesh(path), src_py/main_binder.cc: read the directory
            h = ctypes.c_void_p()
            _check(L.pda_mesh_load(str(handle).encode(), ctypes.byref(h)))
            handle = h.value
        self._h = ctypes.c_void_p(handle)
        dim, st = ctypes.c_int(), ctypes.c_int()
        ns, nst = ctypes.c_int64(), ctypes.c_int64()
        d = (ctypes.c_double * 3)()
        _check(L.pda_mesh_info(self._h, ctypes.byref(dim), ctypes.byref(st), ctypes.byref(ns),
                               ctypes.byref(nst), d))
        self._dim, self._stencil = dim.value, st.value
        self._ns, self._nst = ns.value, nst.value
        self._d = tuple(d)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.pda_mesh_destroy(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def dimensionality(self):
        return self._dim

    def stencilSize(self):
        return self._stencil

    def sampleMeshSize(self):
        return self._ns

    def stencilMeshSize(self):
        return self._nst

    def dx(self):
        return self._d[0]

    def dy(self):
        return self._d[1]

    def dz(self):
        return self._d[2]

    # 1/d of the axes the mesh has, 0 for the others (impl/mesh_read_info.hpp:86-92)
    def dxInv(self):
        return 1.0 / self._d[0] if self._d[0] != 0.0 else 0.0

    def dyInv(self):
        return 1.0 / self._d[1] if self._d[1] != 0.0 else 0.0

    def dzInv(self):
        return 1.0 / self._d[2] if self._d[2] != 0.0 else 0.0

    def graph(self):
        ptr = lib().pda_mesh_graph(self._h)
        nc = 1 + (self._stencil - 1) * self._dim
        return np.ctypeslib.as_array(ptr, shape=(self._ns, nc)).copy()

    def _coords(self, axis):
        ptr = lib().pda_mesh_coords(self._h, axis)
        return np.ctypeslib.as_array(ptr, shape=(self._nst,)).copy()

    def viewX(self):
        return self._coords(0)

    def viewY(self):
        return self._coords(1)

    def viewZ(self):
        return self._coords(2)

    def row_classes(self):
        ni, nn = ctypes.c_int64(), ctypes.c_int64()
        _check(lib().pda_mesh_row_classes(self._h, ctypes.byref(ni), ctypes.byref(nn)))
        return ni.value, nn.value

    def _rows(self, which):
        n = ctypes.c_int64()
        _check(lib().pda_mesh_boundary_rows(self._h, which, None, ctypes.byref(n)))
        out = np.zeros(n.value, dtype=np.int32)
        _check(lib().pda_mesh_boundary_rows(self._h, which, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                            ctypes.byref(n)))
        return out

    def graphRowsOfCellsNearBd(self):
        """rows with a boundary within the mesh stencil width (impl/mesh_ccu.hpp:385-447)"""
        return self._rows(0)

    def graphRowsOfCellsAwayFromBd(self):
        return self._rows(1)

    def graphRowsOfCellsStrictlyOnBd(self):
        """2D: rows with a missing first-order neighbour (impl/mesh_ccu.hpp:276-305)"""
        return self._rows(2)

    def write(self, directory):
        os.makedirs(directory, exist_ok=True)
        _check(lib().pda_mesh_write(self._h, directory.encode()))

    def sample_gids(self):
        """(sample cells, stencil cells) as full-mesh ids, for sample meshes."""
        s = np.zeros(self._ns, dtype=np.int32)
        t = np.zeros(self._nst, dtype=np.int32)
        _check(lib().pda_mesh_sample_gids(
            self._h, s.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
            t.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return s, t


def load_cellcentered_uniform_mesh(path):
    h = ctypes.c_void_p()
    _check(lib().pda_mesh_load(str(path).encode(), ctypes.byref(h)))
    return CellCenteredUniformMesh(h.value)


def create_structured_mesh(n, bounds, stencil, periodic=()):
    """Full natural-order mesh as meshing_scripts/create_full_mesh.py writes it.

    n: cells per axis (len = dim); bounds: (xmin, xmax[, ymin, ymax[, zmin, zmax]]);
    periodic: axes (0,1,2) with periodic wrap, or True for all.
    """
    dim = len(n)
    nn = (ctypes.c_int64 * 3)(*(list(n) + [1] * (3 - dim)))
    b = (ctypes.c_double * 6)(*(list(bounds) + [0.0] * (6 - len(bounds))))
    if periodic is True:
        periodic = tuple(range(dim))
    mask = 0
    for a in periodic:
        mask |= 1 << a
    h = ctypes.c_void_p()
    _check(lib().pda_mesh_create_structured(dim, nn, b, stencil, mask, ctypes.byref(h)))
    return CellCenteredUniformMesh(h.value)


def create_sample_mesh(full_mesh, sample_gids):
    g = np.ascontiguousarray(np.asarray(sample_gids, dtype=np.int32))
    h = ctypes.c_void_p()
    _check(lib().pda_mesh_create_sample(full_mesh.handle,
                                        g.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                        len(g), ctypes.byref(h)))
    return CellCenteredUniformMesh(h.value)


# ---------------------------------------------------------------- problem
def _torch():
    import torch
    return torch


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)


def _dev_tensor(t, n, name):
    torch = _torch()
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float64:
        raise PdaError(f"{name}: expected a float64 CUDA (HIP) tensor")
    if not t.is_contiguous() or t.numel() != n:
        raise PdaError(f"{name}: expected a contiguous tensor of {n} entries, got {t.numel()}")
    return t


class JacobianCSR:
    """Row-major CSR Jacobian with a fixed pattern (device tensors)."""

    def __init__(self, rowptr, colidx, values, shape):
        self.rowptr, self.colidx, self.values, self.shape = rowptr, colidx, values, shape

    def nonZeros(self):
        return self.values.numel()

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.values.cpu().numpy(), self.colidx.cpu().numpy(),
                              self.rowptr.cpu().numpy()), shape=self.shape)


class Problem:
    """Mirror of PublicProblemEigenMixinCpp (adapter_cpp.hpp:60-264) on device."""

    def __init__(self, mesh, problem_enum, scheme=InviscidFluxReconstruction.FirstOrder,
                 ic_flag=0, params=None):
        fam = FAMILY[type(problem_enum)]
        self._mesh = mesh
        keys = list((params or {}).keys())
        vals = [float(params[k]) for k in keys]
        ck = (ctypes.c_char_p * max(len(keys), 1))(*[k.encode() for k in keys])
        cv = (ctypes.c_double * max(len(vals), 1))(*vals)
        h = ctypes.c_void_p()
        _check(lib().pda_problem_create(mesh.handle, fam, int(problem_enum), int(scheme),
                                        int(ic_flag), ck, cv, len(keys), ctypes.byref(h)))
        self._h = h
        m, ns, nst, nnz = ctypes.c_int(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _check(lib().pda_problem_sizes(h, ctypes.byref(m), ctypes.byref(ns), ctypes.byref(nst),
                                       ctypes.byref(nnz)))
        self._m, self._ns, self._nst, self._nnz = m.value, ns.value, nst.value, nnz.value
        self._pattern = None

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.pda_problem_destroy(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def numDofPerCell(self):
        return self._m

    def totalDofSampleMesh(self):
        return self._ns

    def totalDofStencilMesh(self):
        return self._nst

    def nonZeros(self):
        return self._nnz

    def queryParameter(self, name):
        v = ctypes.c_double()
        _check(lib().pda_problem_parameter(self._h, name.encode(), ctypes.byref(v)))
        return v.value

    def gamma(self):
        return self.queryParameter("gamma")

    def gravity(self):
        return self.queryParameter("gravity")

    def coriolis(self):
        return self.queryParameter("coriolis")

    def initialCondition(self):
        torch = _torch()
        U = torch.empty(self._nst, dtype=torch.float64, device="cuda")
        _check(lib().pda_initial_condition(self._h, _ptr(U), _stream()))
        return U

    def createState(self):
        return _torch().zeros(self._nst, dtype=_torch().float64, device="cuda")

    def createRightHandSide(self):
        return _torch().zeros(self._ns, dtype=_torch().float64, device="cuda")

    createVelocity = createRightHandSide
    createRhs = createRightHandSide  # adapter_cpp.hpp:116

    def jacobian_pattern_host(self):
        """(rowptr int64, colidx int32) as numpy, reference CSR order."""
        if self._pattern is None:
            rp = np.zeros(self._ns + 1, dtype=np.int64)
            ci = np.zeros(self._nnz, dtype=np.int32)
            _check(lib().pda_jacobian_pattern_host(
                self._h, rp.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                ci.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
            self._pattern = (rp, ci)
        return self._pattern

    def createJacobian(self):
        torch = _torch()
        rp, ci = self.jacobian_pattern_host()
        return JacobianCSR(torch.from_numpy(rp).cuda(), torch.from_numpy(ci).cuda(),
                           torch.zeros(self._nnz, dtype=torch.float64, device="cuda"),
                           (self._ns, self._nst))

    def rightHandSide(self, U, t, V):
        _dev_tensor(U, self._nst, "state")
        _dev_tensor(V, self._ns, "rhs")
        _check(lib().pda_velocity(self._h, _ptr(U), float(t), _ptr(V), _stream()))

    velocity = rightHandSide

    def autotune(self, U):
        """Time the velocity kernel's strip heights on state U and keep the
        fastest for this problem (pda_problem_autotune); returns (rows per
        strip, us per velocity call).  The velocity values do not depend on
        the choice."""
        _dev_tensor(U, self._nst, "state")
        r, us = ctypes.c_int(), ctypes.c_double()
        _check(lib().pda_problem_autotune(self._h, _ptr(U), _stream(), ctypes.byref(r), ctypes.byref(us)))
        return r.value, us.value

    def set_velocity_strip_height(self, rows_per_strip):
        """Fix the velocity kernel's strip height (rows per workgroup) for
        this problem, e.g. after timing heights in the caller's own loop
        (pda_problem_set_velocity_strip_height); 0 restores the grid-size
        rule.  The velocity values do not depend on the height."""
        _check(lib().pda_problem_set_velocity_strip_height(self._h, int(rows_per_strip)))

    def set_jacobian_strip_height(self, rows_per_strip):
        """Fix the strip height of the first-order 2D and 3D Euler Jacobian
        kernels (pda_problem_set_jacobian_strip_height); 0 restores the
        grid-size rule.  The Jacobian values do not depend on the height."""
        _check(lib().pda_problem_set_jacobian_strip_height(self._h, int(rows_per_strip)))

    rhs = rightHandSide
    __call__ = rightHandSide

    def _jvals(self, J):
        vals = J.values if isinstance(J, JacobianCSR) else J
        return _dev_tensor(vals, self._nnz, "jacobian values")

    def jacobian(self, U, t, J):
        _dev_tensor(U, self._nst, "state")
        _check(lib().pda_jacobian(self._h, _ptr(U), float(t), _ptr(self._jvals(J)), None,
                                  _stream()))

    def rightHandSideAndJacobian(self, U, t, V, J):
        _dev_tensor(U, self._nst, "state")
        _dev_tensor(V, self._ns, "rhs")
        _check(lib().pda_jacobian(self._h, _ptr(U), float(t), _ptr(self._jvals(J)), _ptr(V),
                                  _stream()))

    def rhsAndJacobian(self, U, t, V, J=None):
        """adapter_cpp.hpp:210-221: with J as rightHandSideAndJacobian,
        without it as rightHandSide."""
        if J is None:
            self.rightHandSide(U, t, V)
        else:
            self.rightHandSideAndJacobian(U, t, V, J)

    def createApplyJacobianResult(self, B):
        torch = _torch()
        if B.dim() == 1:
            return torch.zeros(self._ns, dtype=torch.float64, device="cuda")
        return torch.zeros(self._ns, B.shape[1], dtype=torch.float64, device="cuda")

    def applyJacobian(self, U, B, t, AJ):
        """AJ = J(U) B for a vector or a (column- or row-major) matrix B."""
        _dev_tensor(U, self._nst, "state")
        if B.dim() == 1:
            ncols, layout = 1, 0
            Bc = B.contiguous()
        else:
            ncols = B.shape[1]
            if B.is_contiguous():
                layout, Bc = 1, B
            else:
                layout, Bc = 0, B.t().contiguous()
        out = AJ if (AJ.is_contiguous() and (B.dim() == 1 or layout == 1)) else None
        tgt = out if out is not None else _torch().empty(
            (ncols, self._ns) if layout == 0 else (self._ns, ncols), dtype=_torch().float64,
            device="cuda")
        _check(lib().pda_apply_jacobian(self._h, _ptr(U), float(t), _ptr(Bc), ncols, layout,
                                        _ptr(tgt), _stream()))
        if out is None:
            AJ.copy_(tgt.t() if layout == 0 else tgt)


def create_problem(mesh, problem_enum, scheme=InviscidFluxReconstruction.FirstOrder,
                   ic_flag=0, params=None):
    """create_problem(mesh, ProblemEnum, InviscidFluxReconstruction[, icFlag][, params]);
    for the Burgers problems, as the reference's binding
    (src_py/main_binder.cc:387-392), the fourth argument may be the viscous
    scheme, ViscousFluxReconstruction.FirstOrder."""
    if isinstance(ic_flag, dict):
        params, ic_flag = ic_flag, 0
    if isinstance(ic_flag, ViscousFluxReconstruction):
        if not isinstance(problem_enum, AdvectionDiffusion2d):
            raise PdaError("create_problem: a viscous scheme applies to the Burgers problems only")
        ic_flag = 0
    return Problem(mesh, problem_enum, scheme, ic_flag, params)


def create_adv_diff_reac_2d_problem_A(mesh, scheme, ux, uy, diffusion, sigma):
    """Problem A with custom advection velocity, diffusion and reaction
    (advection_diffusion_reaction2d.hpp:148-171, src_py/main_binder.cc:365-370)."""
    return Problem(mesh, AdvectionDiffusionReaction2d.ProblemA, scheme, 0,
                   {"ux": ux, "uy": uy, "diffusion": diffusion, "sigma": sigma})


def create_slip_wall_swe_2d_problem(mesh, scheme, gravity, coriolis, pulseMagnitude):
    """Slip-wall shallow water with custom gravity, Coriolis and pulse
    (swe2d.hpp:276-302, src_py/main_binder.cc:508-513)."""
    return Problem(mesh, Swe2d.SlipWall, scheme, 1,
                   {"gravity": gravity, "coriolis": coriolis, "pulseMagnitude": pulseMagnitude})


def create_burgers_2d_problem(mesh, problem_enum, scheme, viscous_scheme, params):
    """Burgers 2D with a parameter map (advection_diffusion2d.hpp:113-148,
    src_py/main_binder.cc:394-399)."""
    if ViscousFluxReconstruction(viscous_scheme) != ViscousFluxReconstruction.FirstOrder:
        raise PdaError("create_burgers_2d_problem: invalid viscous flux reconstruction")
    return Problem(mesh, problem_enum, scheme, 1, dict(params))


# ------------------------------------------------------- explicit stepping
class StepScheme(enum.IntEnum):
    """pressio's stage formulas (ForwardEuler, SSPRK3, RK4) and the module's
    own integrators with their algebra (RK2, ModuleRK4, ModuleSSP3:
    pressiodemoapps/__init__.py:90-185, used by advanceRK2/RK4/SSP3)."""
    ForwardEuler = 0
    SSPRK3 = 1
    RK4 = 2
    RK2 = 3
    ModuleRK4 = 4
    ModuleSSP3 = 5


class ExplicitStepper:
    """Device time integrator for dU/dt = velocity(U) (pda_stepper_*):
    pressio's forward Euler / SSPRK3 / RK4 stage formulas
    (pressio/ode/impl/ode_explicit_stepper_without_mass_matrix.hpp:168-352)
    and the module's RK2 (pressiodemoapps/__init__.py:90-114).  With
    use_graph one step is recorded as a hipGraph and replayed."""

    def __init__(self, problem, scheme):
        self._p = problem
        h = ctypes.c_void_p()
        _check(lib().pda_stepper_create(problem.handle, int(scheme), ctypes.byref(h)))
        self._h = h.value

    def advance(self, state, dt, nsteps, start_time=0.0, use_graph=True):
        _dev_tensor(state, self._p.totalDofStencilMesh(), "state")
        _check(lib().pda_stepper_advance(self._h, _ptr(state), float(start_time), float(dt),
                                         int(nsteps), 1 if use_graph else 0, _stream()))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.pda_stepper_destroy(h)
            self._h = None


def _advance(scheme, appObj, state, dt, Nsteps, startTime, observer, use_graph):
    st = ExplicitStepper(appObj, scheme)
    if observer is None:
        st.advance(state, dt, Nsteps, startTime, use_graph)
        return
    v = appObj.createRightHandSide()
    t = startTime
    for step in range(1, Nsteps + 1):
        appObj.rightHandSide(state, t, v)
        observer(step - 1, state, v)
        st.advance(state, dt, 1, t, use_graph)
        t += dt


def advanceRK2(appObj, state, dt, Nsteps, startTime=0.0, observer=None, showProgress=False,
               use_graph=True):
    """pressiodemoapps.advanceRK2 (pressiodemoapps/__init__.py:90-114), on the device."""
    _advance(StepScheme.RK2, appObj, state, dt, Nsteps, startTime, observer, use_graph)


def advanceRK4(appObj, state, dt, Nsteps, startTime=0.0, observer=None, showProgress=False,
               use_graph=True):
    """pressiodemoapps.advanceRK4 (pressiodemoapps/__init__.py:117-152), on the
    device, with the module's own stage algebra (k_i = dt v_i, y + (k1 + 2 k2
    + 2 k3 + k4)/6) rounded as numpy rounds it."""
    _advance(StepScheme.ModuleRK4, appObj, state, dt, Nsteps, startTime, observer, use_graph)


def advanceSSP3(appObj, state, dt, Nsteps, startTime=0.0, observer=None, showProgress=False,
                use_graph=True):
    """pressiodemoapps.advanceSSP3 (pressiodemoapps/__init__.py:155-185), on the
    device, with the module's own stage algebra rounded as numpy rounds it."""
    _advance(StepScheme.ModuleSSP3, appObj, state, dt, Nsteps, startTime, observer, use_graph)


# ------------------------------------------------------- implicit stepping
class ImplicitScheme(enum.IntEnum):
    BDF1 = 1
    BDF2 = 2
    CrankNicolson = 3


class StopRule(enum.IntEnum):
    Correction = 0   # pressio default: ||dx|| < tol, dx not applied
    Residual = 1     # ||R(y)|| <= tol


class ImplicitStepper:
    """Device implicit stepper (pda_implicit_*): pressio's BDF1 / BDF2 /
    Crank-Nicolson residual and Jacobian, Newton on the exact Jacobian, a
    device BiCGSTAB with block-Jacobi preconditioning."""

    def __init__(self, problem, scheme):
        self._p = problem
        h = ctypes.c_void_p()
        _check(lib().pda_implicit_create(problem.handle, int(scheme), ctypes.byref(h)))
        self._h = h.value

    def set_options(self, stop=StopRule.Correction, tol=1e-6, max_newton=100, lin_rtol=1e-14,
                    lin_max=0):
        _check(lib().pda_implicit_set_options(self._h, int(stop), float(tol), int(max_newton),
                                              float(lin_rtol), int(lin_max)))

    def step(self, state, t, dt, prev=None, guess=None):
        """y_n (state, in place) -> y_{n+1}; prev = y_{n-1} for BDF2 (None:
        BDF1 step); guess = Newton's initial iterate (None: y_n).  Returns
        the step's statistics as a dict."""
        n = self._p.totalDofStencilMesh()
        _dev_tensor(state, n, "state")
        if prev is not None:
            _dev_tensor(prev, n, "prev")
        if guess is not None:
            _dev_tensor(guess, n, "guess")
        st = ImplicitStats()
        _check(lib().pda_implicit_step(self._h, _ptr(state), _ptr(prev) if prev is not None else None,
                                       _ptr(guess) if guess is not None else None, float(t),
                                       float(dt), _stream(), ctypes.byref(st)))
        return st.as_dict()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.pda_implicit_destroy(h)
            self._h = None


# ------------------------------------------------------ learned Newton guess
class GuessNet:
    """Learned initial guess for BDF2 steps of a periodic nx x ny
    two-component field (pda_guessnet_*): the network picks, per step, the
    extrapolation order t in [1, 4] of the FP64 predictor P(t) (the BDF2
    relation with f(y_{n+1}) extrapolated in time with order k from f_n,
    ..., f_{n-k}, linear in t between integer orders).  The grid means m_j
    of its output channels estimate the step's cost from P(T_j) at the
    orders ORDERS = (1, 1.5, ..., 4), relative to their mean; t = sum_j p_j
    T_j with p = softmax(-m / 0.25);
    a 4-layer 3x3 conv net with BF16 weights and FP32 accumulation on MFMA,
    trained on the device on sum_j (m_j - (c_j - mean c))^2 / 7 of samples
    with measured costs c.  f = [f_n, f_{n-1}, f_{n-2},
    f_{n-3}, f_{n-4}], the problem's velocity at y_n, ..., y_{n-4}."""

    ORDERS = tuple(1.0 + 0.5 * j for j in range(7))

    def __init__(self, nx, ny):
        h = ctypes.c_void_p()
        _check(lib().pda_guessnet_create(int(nx), int(ny), ctypes.byref(h)))
        self._h = h.value
        self.nx, self.ny = int(nx), int(ny)
        n = ctypes.c_int64()
        _check(lib().pda_guessnet_num_params(self._h, ctypes.byref(n)))
        self.nparams = n.value
        self.scales = (5e-3, (1e-3, 1e-4, 1e-5, 1e-6), 5.5e-3, 4.5e-3)

    def set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        _check(lib().pda_guessnet_set_params(self._h, w.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                             len(w)))

    def get_params(self):
        w = np.zeros(self.nparams, dtype=np.float32)
        _check(lib().pda_guessnet_get_params(self._h, w.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                             len(w)))
        return w

    def set_scales(self, dt, spreads, nu_mean, nu_scale):
        """dt; spreads of dt D^k f for k = 1..4; nu_mean, nu_scale of the
        diffusion input."""
        spreads = tuple(float(v) for v in spreads)
        if len(spreads) != 4:
            raise PdaError("GuessNet.set_scales: four spreads (of dt D^k f, k = 1..4)")
        self.scales = (float(dt), spreads, float(nu_mean), float(nu_scale))
        sp = (ctypes.c_double * 4)(*spreads)
        _check(lib().pda_guessnet_set_scales(self._h, float(dt), sp, float(nu_mean), float(nu_scale)))

    def _fields(self, f, **fields):
        n = 2 * self.nx * self.ny
        for name, t in fields.items():
            _dev_tensor(t, n, name)
        if len(f) != 5:
            raise PdaError("GuessNet: f must hold f_n, f_{n-1}, f_{n-2}, f_{n-3}, f_{n-4}")
        for k, t in enumerate(f):
            _dev_tensor(t, n, f"f_(n-{k})")
        return (ctypes.c_void_p * 5)(*[t.data_ptr() for t in f])

    def _costs(self, costs):
        c = np.ascontiguousarray(costs, dtype=np.float64)
        if c.shape != (len(self.ORDERS),) or not np.all(np.isfinite(c)):
            raise PdaError(f"GuessNet: costs must be {len(self.ORDERS)} finite values (one per order)")
        return c, c.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

    def predict(self, un, unm1, f, nu, out=None):
        fp = self._fields(f, un=un, unm1=unm1)
        if out is None:
            out = _torch().empty_like(un)
        _dev_tensor(out, 2 * self.nx * self.ny, "out")
        _check(lib().pda_guessnet_predict(self._h, _ptr(un), _ptr(unm1), fp, float(nu), _ptr(out), _stream()))
        return out

    def policy(self):
        """(t, p, m) of the last predict / gradients / train_step: the order,
        the weights over ORDERS and the cost estimates."""
        t = ctypes.c_double()
        p = (ctypes.c_double * len(self.ORDERS))()
        m = (ctypes.c_double * len(self.ORDERS))()
        _check(lib().pda_guessnet_policy(self._h, ctypes.byref(t), p, m))
        return t.value, np.array(p[:]), np.array(m[:])

    def gradients(self, f, nu, costs):
        """(loss, gradient, raw output [nx*ny*7]) of one sample with the
        given costs at ORDERS."""
        fp = self._fields(f)
        q, qp = self._costs(costs)
        g = np.zeros(self.nparams, dtype=np.float32)
        o = np.zeros(self.nx * self.ny * len(self.ORDERS), dtype=np.float32)
        loss = ctypes.c_float()
        F = ctypes.POINTER(ctypes.c_float)
        _check(lib().pda_guessnet_gradients(self._h, fp, float(nu), qp, g.ctypes.data_as(F), ctypes.byref(loss),
                                            o.ctypes.data_as(F), _stream()))
        return loss.value, g, o

    def train_step(self, f, nu, costs, lr, want_loss=False):
        fp = self._fields(f)
        q, qp = self._costs(costs)
        loss = ctypes.c_float()
        _check(lib().pda_guessnet_train_step(self._h, fp, float(nu), qp, float(lr),
                                             ctypes.byref(loss) if want_loss else None, _stream()))
        return loss.value if want_loss else None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.pda_guessnet_destroy(h)
            self._h = None


def _guessnet_layers():
    """(cin, cout) of the guess network's layers (pda_guessnet_set_params)."""
    return [(8, 32), (32, 32), (32, 32), (32, 7)]


# ----------------------------------------------------------------- batching
class Batch:
    """Many independent 2D instances of one family and scheme evaluated and
    stepped together (pda_batch_*): the state is the concatenation of the
    instances' states; one launch evaluates every instance's velocity."""

    def __init__(self, problems):
        self._probs = list(problems)  # the batch references them
        arr = (ctypes.c_void_p * len(self._probs))(*[p.handle for p in self._probs])
        h = ctypes.c_void_p()
        _check(lib().pda_batch_create(arr, len(self._probs), ctypes.byref(h)))
        self._h = h.value
        tot = ctypes.c_int64()
        offs = (ctypes.c_int64 * (len(self._probs) + 1))()
        _check(lib().pda_batch_sizes(self._h, ctypes.byref(tot), offs))
        self.total = tot.value
        self.offsets = np.array(offs[:], dtype=np.int64)

    def initialCondition(self):
        U = _torch().empty(self.total, dtype=_torch().float64, device="cuda")
        _check(lib().pda_batch_initial_condition(self._h, _ptr(U), _stream()))
        return U

    def createRightHandSide(self):
        return _torch().zeros(self.total, dtype=_torch().float64, device="cuda")

    def rightHandSide(self, U, t, V):
        _dev_tensor(U, self.total, "state")
        _dev_tensor(V, self.total, "velocity")
        _check(lib().pda_batch_velocity(self._h, _ptr(U), float(t), _ptr(V), _stream()))

    def autotune(self, U):
        _dev_tensor(U, self.total, "state")
        r, us = ctypes.c_int(), ctypes.c_double()
        _check(lib().pda_batch_autotune(self._h, _ptr(U), _stream(), ctypes.byref(r), ctypes.byref(us)))
        return r.value, us.value

    def set_velocity_strip_height(self, rows_per_strip):
        """Fix the batched velocity kernel's strip height (1..64), e.g. after
        timing heights in the caller's own stepping loop
        (pda_batch_set_velocity_strip_height); values do not depend on it."""
        _check(lib().pda_batch_set_velocity_strip_height(self._h, int(rows_per_strip)))

    def advance(self, U, scheme, dt, nsteps, t0=0.0, steps_per_graph=10):
        _dev_tensor(U, self.total, "state")
        _check(lib().pda_batch_advance(self._h, int(scheme), _ptr(U), float(t0), float(dt), int(nsteps),
                                       int(steps_per_graph), _stream()))

    def instance(self, U, i):
        return U[self.offsets[i]:self.offsets[i + 1]]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.pda_batch_destroy(h)
            self._h = None
