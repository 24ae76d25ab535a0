"""pda_apply_jacobian_noalloc: the applyJacobian variant that allocates no
device memory (its Jacobian values live in a fixed 2^23-double device array).

Through the C ABI, its result is bit-identical to pda_apply_jacobian's (the
same kernels run on the same inputs) and matches the oracle's CSR product
within 1e-12, on two calls in a row whose second problem is smaller than the
first (the array keeps the first call's values, which the second must not
read), and on a problem with merged columns (its rows accumulate into a
cleared array).  Sizes beyond the array and problems whose device tables
were never built are refused before any device work."""
import ctypes

import numpy as np
import pytest

import cases
import pressiodemoapps_amd as pda

S = pda.InviscidFluxReconstruction
MAX_NNZ = 1 << 23  # PDA_NOALLOC_MAX_NNZ


def _noalloc():
    f = pda.lib().pda_apply_jacobian_noalloc
    vp = ctypes.c_void_p
    f.argtypes = [vp, vp, ctypes.c_double, vp, ctypes.c_int64, ctypes.c_int, vp, vp]
    f.restype = ctypes.c_int
    return f


def test_refuses_oversized_jacobian():
    mesh = pda.create_structured_mesh((210, 210), (0.0, 1.0, 0.0, 1.0), 7)
    prob = pda.create_problem(mesh, pda.Euler2d.Riemann, S.Weno5, 2)
    assert prob.nonZeros() > MAX_NNZ
    d = ctypes.c_void_p(16)  # never dereferenced: the size check comes first
    rc = _noalloc()(prob.handle, d, 0.0, d, 1, 0, d, None)
    assert rc == -1  # PDA_E_INVALID_ARG
    assert b"nonzeros" in pda.lib().pda_last_error()


def test_refuses_problem_without_device_tables():
    mesh = pda.create_structured_mesh((12, 10), (0.0, 1.0, 0.0, 1.0), 3)
    prob = pda.create_problem(mesh, pda.Euler2d.Riemann, S.FirstOrder, 2)
    d = ctypes.c_void_p(16)  # never dereferenced: the state check comes first
    rc = _noalloc()(prob.handle, d, 0.0, d, 1, 0, d, None)
    assert rc == -5  # PDA_E_STATE
    assert b"pda_jacobian_pattern" in pda.lib().pda_last_error()


def _case(name):
    return [c for c in cases.CASES + cases.BIG_CASES if c[0] == name][0]


# a periodic mesh narrower than the WENO5 stencil: rows map two graph columns
# to one block, so the assembly clears the values and accumulates
MERGED = ("burgers-merged-Weno5", pda.AdvectionDiffusion2d.BurgersPeriodic, S.Weno5, (5, 6), (-1, 1, -1, 1), 7,
          True, 1, {"diffusion": 1e-3})


@pytest.mark.gpu
def test_noalloc_matches_apply_jacobian():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import pyoracle
    import scipy.sparse as sp
    orc = pyoracle.Oracle()
    f = _noalloc()
    # first call on the largest problem, then smaller ones
    seq = [_case("euler2d-riemann-big-Weno5"), _case("swe-slipwall-FirstOrder"), MERGED]
    sizes = []
    for i, case in enumerate(seq):
        print(f"[no-alloc] call {i}: {case[0]}", flush=True)
        mesh, prob = cases.make(case)
        sizes.append(prob.nonZeros())
        U = cases.perturbed_state(prob, 11 + i)
        Ud = torch.from_numpy(U).cuda()
        prob.jacobian_pattern_host()
        rp = ctypes.c_int64()
        nz = ctypes.c_int64()
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        assert pda.lib().pda_jacobian_pattern(prob.handle, ctypes.byref(rp), ctypes.byref(nz), ctypes.byref(a),
                                              ctypes.byref(b)) == 0
        rng = np.random.default_rng(40 + i)
        nst, ns = prob.totalDofStencilMesh(), prob.totalDofSampleMesh()
        _, rpo, cio, Jo = orc.velocity_and_jacobian(orc.from_mesh(mesh), cases.oracle_problem(case), U)
        Jsp = sp.csr_matrix((Jo, cio, rpo), shape=(ns, nst))
        for ncols, layout in ((1, 0), (3, 0), (4, 1)):
            B = rng.standard_normal((nst, ncols))
            ref = Jsp @ B
            # layout 0 = column-major: column c at [c*rows, (c+1)*rows)
            Bd = torch.from_numpy(np.ascontiguousarray(B.T if layout == 0 else B).ravel()).cuda()
            out_a = torch.full((ns * ncols,), np.nan, dtype=torch.float64, device="cuda")
            out_b = torch.full((ns * ncols,), np.nan, dtype=torch.float64, device="cuda")
            s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
            rc = pda.lib().pda_apply_jacobian(prob.handle, ctypes.c_void_p(Ud.data_ptr()), 0.0,
                                              ctypes.c_void_p(Bd.data_ptr()), ncols, layout,
                                              ctypes.c_void_p(out_a.data_ptr()), s)
            assert rc == 0, pda.lib().pda_last_error()
            rc = f(prob.handle, ctypes.c_void_p(Ud.data_ptr()), 0.0, ctypes.c_void_p(Bd.data_ptr()), ncols, layout,
                   ctypes.c_void_p(out_b.data_ptr()), s)
            assert rc == 0, pda.lib().pda_last_error()
            torch.cuda.synchronize()
            assert torch.equal(out_a, out_b), (case[0], ncols, layout)
            got = out_b.cpu().numpy().reshape((ncols, ns) if layout == 0 else (ns, ncols))
            got = got.T if layout == 0 else got
            err = cases.rel_err(got, ref)
            print(f"[no-alloc] {case[0]} ncols {ncols} layout {layout}: vs oracle {err:.2e}", flush=True)
            assert err <= 1e-12, err
    assert sizes[0] > sizes[1] > sizes[2]
