"""ctypes bindings of the CPU checkers.  TEST INFRASTRUCTURE ONLY.

- ``Oracle``: oracle/liboracle.so, the C restatement of the reference path.
- ``RefProblem``: oracle/_ref/libpda_ref*.so, the reference headers behind a
  C shim (present only where the reference sources were available at build
  time; callers skip when ``ref_available()`` is False).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libpda_ref.so")
REF_OMP_SO = os.path.join(HERE, "_ref", "libpda_ref_omp.so")

FAMILY_DIM = {1: 1, 2: 2, 3: 3, 4: 2, 5: 2, 6: 2}
FAMILY_M = {1: 3, 2: 4, 3: 5, 4: 3, 5: 2, 6: 1}

_P = ctypes.POINTER
_i32, _i64, _dbl, _vp = ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p


class OrcMesh(ctypes.Structure):
    _fields_ = [("dim", _i32), ("stencil", _i32), ("nrows", _i64), ("ncells", _i64),
                ("graph", _P(ctypes.c_int32)), ("dx", _dbl * 3),
                ("x", _P(_dbl)), ("y", _P(_dbl)), ("z", _P(_dbl))]


class OrcProblem(ctypes.Structure):
    _fields_ = [("family", _i32), ("prob", _i32), ("scheme", _i32), ("p", _dbl * 8)]


def build(quiet=True):
    """Compile liboracle.so (and oracle/_ref when the reference is present)."""
    r = subprocess.run(["make", "-C", HERE, "-j4"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    if not quiet:
        print(r.stdout)


def _dp(a):
    return a.ctypes.data_as(_P(_dbl))


class Oracle:
    _lib = None

    def __init__(self):
        if Oracle._lib is None:
            if not os.path.exists(ORACLE_SO):
                build()
            L = ctypes.CDLL(ORACLE_SO)
            L.orc_velocity.argtypes = [_P(OrcMesh), _P(OrcProblem), _P(_dbl), _dbl, _P(_dbl)]
            L.orc_jacobian_pattern.argtypes = [_P(OrcMesh), _P(OrcProblem), _P(_i64), _P(_i64),
                                               _P(ctypes.c_int32)]
            L.orc_velocity_and_jacobian.argtypes = [_P(OrcMesh), _P(OrcProblem), _P(_dbl), _dbl,
                                                    _P(_dbl), _P(_i64), _P(ctypes.c_int32),
                                                    _P(_dbl)]
            for n in ("orc_weno5", "orc_weno3"):
                getattr(L, n).argtypes = [_P(_dbl), _P(_dbl), _P(_dbl)]
            for n in ("orc_weno5_grad", "orc_weno3_grad"):
                getattr(L, n).argtypes = [_P(_dbl), _P(_dbl), _P(_dbl), _P(_dbl), _P(_dbl)]
            L.orc_face_flux.argtypes = [_P(OrcProblem), _i32, _P(_dbl), _P(_dbl), _P(_dbl),
                                        _P(_dbl)]
            L.orc_face_flux_jac.argtypes = [_P(OrcProblem), _i32, _P(_dbl), _P(_dbl), _P(_dbl),
                                            _P(_dbl), _P(_dbl)]
            Oracle._lib = L
        self.L = Oracle._lib

    # --- mesh / problem structs
    @staticmethod
    def mesh_struct(dim, stencil, graph, dx, x, y, z, ncells):
        g = np.ascontiguousarray(graph, dtype=np.int32)
        xs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z)]
        ms = OrcMesh(dim, stencil, g.shape[0], ncells, g.ctypes.data_as(_P(ctypes.c_int32)),
                     (_dbl * 3)(*dx), _dp(xs[0]), _dp(xs[1]), _dp(xs[2]))
        ms._keep = (g, xs)
        return ms

    @staticmethod
    def from_mesh(mesh):
        """Build the oracle mesh struct from a product mesh object (its arrays)."""
        return Oracle.mesh_struct(mesh.dimensionality(), mesh.stencilSize(), mesh.graph(),
                                  (mesh.dx(), mesh.dy(), mesh.dz()), mesh.viewX(), mesh.viewY(),
                                  mesh.viewZ(), mesh.stencilMeshSize())

    @staticmethod
    def problem(family, prob, scheme, params=()):
        p = (_dbl * 8)(*(list(params) + [0.0] * (8 - len(params))))
        return OrcProblem(family, prob, scheme, p)

    def velocity(self, ms, pb, U, t=0.0):
        m = FAMILY_M[pb.family]
        V = np.zeros(ms.nrows * m)
        U = np.ascontiguousarray(U, dtype=np.float64)
        rc = self.L.orc_velocity(ctypes.byref(ms), ctypes.byref(pb), _dp(U), t, _dp(V))
        if rc:
            raise RuntimeError(f"orc_velocity rc={rc}")
        return V

    def pattern(self, ms, pb):
        nnz = _i64()
        rc = self.L.orc_jacobian_pattern(ctypes.byref(ms), ctypes.byref(pb), ctypes.byref(nnz),
                                         None, None)
        if rc:
            raise RuntimeError(f"orc_jacobian_pattern rc={rc}")
        m = FAMILY_M[pb.family]
        rp = np.zeros(ms.nrows * m + 1, dtype=np.int64)
        ci = np.zeros(nnz.value, dtype=np.int32)
        self.L.orc_jacobian_pattern(ctypes.byref(ms), ctypes.byref(pb), ctypes.byref(nnz),
                                    rp.ctypes.data_as(_P(_i64)),
                                    ci.ctypes.data_as(_P(ctypes.c_int32)))
        return rp, ci

    def velocity_and_jacobian(self, ms, pb, U, t=0.0):
        rp, ci = self.pattern(ms, pb)
        m = FAMILY_M[pb.family]
        V = np.zeros(ms.nrows * m)
        vals = np.zeros(len(ci))
        U = np.ascontiguousarray(U, dtype=np.float64)
        rc = self.L.orc_velocity_and_jacobian(ctypes.byref(ms), ctypes.byref(pb), _dp(U), t,
                                              _dp(V), rp.ctypes.data_as(_P(_i64)),
                                              ci.ctypes.data_as(_P(ctypes.c_int32)), _dp(vals))
        if rc:
            raise RuntimeError(f"orc_velocity_and_jacobian rc={rc}")
        return V, rp, ci, vals

    def weno5(self, q, grad=False):
        q = np.ascontiguousarray(q, dtype=np.float64)
        un, up = _dbl(), _dbl()
        if not grad:
            self.L.orc_weno5(ctypes.byref(un), ctypes.byref(up), _dp(q))
            return un.value, up.value
        gn, gp = np.zeros(6), np.zeros(6)
        self.L.orc_weno5_grad(ctypes.byref(un), ctypes.byref(up), _dp(gn), _dp(gp), _dp(q))
        return un.value, up.value, gn, gp

    def weno3(self, q, grad=False):
        q = np.ascontiguousarray(q, dtype=np.float64)
        un, up = _dbl(), _dbl()
        if not grad:
            self.L.orc_weno3(ctypes.byref(un), ctypes.byref(up), _dp(q))
            return un.value, up.value
        gn, gp = np.zeros(4), np.zeros(4)
        self.L.orc_weno3_grad(ctypes.byref(un), ctypes.byref(up), _dp(gn), _dp(gp), _dp(q))
        return un.value, up.value, gn, gp

    def face_flux(self, pb, dim, qL, qR, n):
        m = FAMILY_M[pb.family]
        F = np.zeros(m)
        arr = [np.ascontiguousarray(a, dtype=np.float64) for a in (qL, qR, n)]
        self.L.orc_face_flux(ctypes.byref(pb), dim, _dp(arr[0]), _dp(arr[1]), _dp(arr[2]), _dp(F))
        return F

    def face_flux_jac(self, pb, dim, qL, qR, n):
        m = FAMILY_M[pb.family]
        JL, JR = np.zeros(m * m), np.zeros(m * m)
        arr = [np.ascontiguousarray(a, dtype=np.float64) for a in (qL, qR, n)]
        self.L.orc_face_flux_jac(ctypes.byref(pb), dim, _dp(arr[0]), _dp(arr[1]), _dp(arr[2]),
                                 _dp(JL), _dp(JR))
        return JL.reshape(m, m), JR.reshape(m, m)


def _cpu_flags():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return set()


# x86-64 micro-architecture levels (LZCNT is listed as "abm")
_V3 = {"avx", "avx2", "bmi1", "bmi2", "f16c", "fma", "abm", "movbe"}
_V4 = _V3 | {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"}


def ref_omp_build():
    """(path, level) of the fastest OpenMP reference build this host can
    run: x86-64-v4 (AVX-512), v3 (AVX2 + FMA) or the baseline x86-64."""
    flags = _cpu_flags()
    for level, need in (("v4", _V4), ("v3", _V3)):
        path = os.path.join(HERE, "_ref", f"libpda_ref_omp_{level}.so")
        if need <= flags and os.path.exists(path):
            return path, "x86-64-" + level
    return REF_OMP_SO, "x86-64"


def ref_available(omp=False):
    return os.path.exists(ref_omp_build()[0] if omp else REF_SO)


class RefProblem:
    """The reference's own problem object, loaded from oracle/_ref."""
    _libs = {}

    def __init__(self, meshdir, family, prob, scheme, icflag=0, params=None, omp=False):
        path = ref_omp_build()[0] if omp else REF_SO
        if path not in RefProblem._libs:
            L = ctypes.CDLL(path)
            L.ref_create.restype = _vp
            L.ref_create.argtypes = [ctypes.c_char_p, _i32, _i32, _i32, _i32,
                                     _P(ctypes.c_char_p), _P(_dbl), _i32]
            L.ref_destroy.argtypes = [_vp]
            L.ref_sizes.argtypes = [_vp, _P(_i32), _P(_i64), _P(_i64), _P(_i64)]
            L.ref_initial_condition.argtypes = [_vp, _P(_dbl)]
            L.ref_velocity.argtypes = [_vp, _P(_dbl), _dbl, _P(_dbl)]
            L.ref_velocity_and_jacobian.argtypes = [_vp, _P(_dbl), _dbl, _P(_dbl), _P(_i64),
                                                    _P(ctypes.c_int32), _P(_dbl)]
            L.ref_time_calls.argtypes = [_vp, _P(_dbl), _dbl, _i32, _i32, _P(_dbl)]
            L.ref_last_error.restype = ctypes.c_char_p
            RefProblem._libs[path] = L
        self.L = RefProblem._libs[path]
        params = params or {}
        keys = list(params.keys())
        ck = (ctypes.c_char_p * max(1, len(keys)))(*[k.encode() for k in keys])
        cv = (_dbl * max(1, len(keys)))(*[float(params[k]) for k in keys])
        self.h = self.L.ref_create(str(meshdir).encode(), family, prob, scheme, icflag, ck, cv,
                                   len(keys))
        if not self.h:
            raise RuntimeError("ref_create failed: " + self.L.ref_last_error().decode())
        m, r, c, nnz = _i32(), _i64(), _i64(), _i64()
        self.L.ref_sizes(self.h, ctypes.byref(m), ctypes.byref(r), ctypes.byref(c),
                         ctypes.byref(nnz))
        self.m, self.rows, self.cols, self.nnz = m.value, r.value, c.value, nnz.value

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ref_destroy(self.h)
            self.h = None

    def initial_condition(self):
        U = np.zeros(self.cols)
        if self.L.ref_initial_condition(self.h, _dp(U)):
            raise RuntimeError(self.L.ref_last_error().decode())
        return U

    def velocity(self, U, t=0.0):
        U = np.ascontiguousarray(U, dtype=np.float64)
        V = np.zeros(self.rows)
        if self.L.ref_velocity(self.h, _dp(U), t, _dp(V)):
            raise RuntimeError(self.L.ref_last_error().decode())
        return V

    def time_calls(self, U, which, reps, t=0.0):
        """Wall times (s) of `reps` calls of the reference's rightHandSide
        (which = 0), jacobian (1) or rightHandSideAndJacobian (2) on a state,
        velocity and Jacobian created once, after one warm-up call."""
        U = np.ascontiguousarray(U, dtype=np.float64)
        secs = np.zeros(reps)
        if self.L.ref_time_calls(self.h, _dp(U), t, which, reps, _dp(secs)):
            raise RuntimeError(self.L.ref_last_error().decode())
        return secs

    def velocity_and_jacobian(self, U, t=0.0):
        U = np.ascontiguousarray(U, dtype=np.float64)
        V = np.zeros(self.rows)
        rp = np.zeros(self.rows + 1, dtype=np.int64)
        ci = np.zeros(self.nnz, dtype=np.int32)
        vals = np.zeros(self.nnz)
        if self.L.ref_velocity_and_jacobian(self.h, _dp(U), t, _dp(V), rp.ctypes.data_as(_P(_i64)),
                                            ci.ctypes.data_as(_P(ctypes.c_int32)), _dp(vals)):
            raise RuntimeError(self.L.ref_last_error().decode())
        return V, rp, ci, vals
