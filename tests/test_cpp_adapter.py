"""The C++ drop-in facade (include/pda_hip_adapter.hpp: the reference's
create_problem_eigen / load_cellcentered_uniform_mesh_eigen and
PublicProblemEigenMixinCpp method names, adapter_cpp.hpp:60-264, over the C
ABI).

CPU: the header compiles with g++ against libpda.so, and the check program
(tests/cpp/adapter_check.cpp) reports the same mesh sizes, problem sizes and
Jacobian patterns as the Python module, the parameter values it was given,
and std::runtime_error for an unknown parameter and a missing mesh.
GPU: the program's rightHandSide / jacobian / rightHandSideAndJacobian /
rhsAndJacobian / operator() / applyJacobian on a 24 x 20 Riemann WENO5
problem against the oracle (1e-12)."""
import os
import subprocess

import numpy as np
import pytest

import cases
import pressiodemoapps_amd as pda

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = os.path.join(ROOT, "tests", "cpp", "_build", "adapter_check")
S = pda.InviscidFluxReconstruction


def _mesh_dir(tmp_path):
    mesh = pda.create_structured_mesh((24, 20), (0.0, 1.0, 0.0, 1.0), 7)
    d = str(tmp_path / "mesh")
    mesh.write(d)
    return mesh, d


def test_facade_compiles_and_matches_module_on_host(tmp_path):
    exe = str(tmp_path / "adapter_check")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-D__HIP_PLATFORM_AMD__",
           "-I" + os.path.join(ROOT, "include"), "-I/opt/rocm/include",
           os.path.join(ROOT, "tests", "cpp", "adapter_check.cpp"),
           "-L" + os.path.join(ROOT, "pressio-demoapps_amd"), "-lpda", "-L/opt/rocm/lib", "-lamdhip64",
           "-Wl,-rpath," + os.path.join(ROOT, "pressio-demoapps_amd"), "-Wl,-rpath,/opt/rocm/lib", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    mesh, d = _mesh_dir(tmp_path)
    r = subprocess.run([exe, "host", d], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[0] == f"mesh 2 7 {mesh.sampleMeshSize()} {mesh.stencilMeshSize()}"
    # the same three problems through the Python module
    lmesh = pda.load_cellcentered_uniform_mesh(d)
    probs = [pda.create_problem(lmesh, pda.Euler2d.Riemann, S.Weno5, 2),
             pda.create_problem(lmesh, pda.Swe2d.SlipWall, S.FirstOrder, 1, {"gravity": 7.5, "coriolis": -2.5}),
             pda.create_adv_diff_reac_2d_problem_A(lmesh, S.Weno3, 0.3, -0.2, 0.01, 0.5)]
    for line, p in zip(lines[1:4], probs):
        rp, ci = p.jacobian_pattern_host()
        want = (f"problem {p.numDofPerCell()} {p.totalDofSampleMesh()} {p.totalDofStencilMesh()} "
                f"{rp[-1]} {int(ci.astype(np.int64).sum())}")
        assert line == want
    assert lines[4] == "gravity 7.5 coriolis -2.5"
    assert lines[5].startswith("error parameter:") and "notAParameter" in lines[5]
    assert lines[6].startswith("error mesh:") and "not found" in lines[6]


@pytest.mark.gpu
def test_facade_evaluates_on_device(tmp_path, oracle):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    assert os.path.exists(PROG), f"{PROG} not built: run __graft_entry__.build()"
    mesh, d = _mesh_dir(tmp_path)
    lmesh = pda.load_cellcentered_uniform_mesh(d)
    prob = pda.create_problem(lmesh, pda.Euler2d.Riemann, S.Weno5, 2)
    U = cases.perturbed_state(prob, 9)
    U.tofile(str(tmp_path / "U.bin"))
    out = tmp_path / "out"
    out.mkdir()
    print("[gpu-cpp] adapter_check gpu", flush=True)
    r = subprocess.run([PROG, "gpu", d, str(tmp_path / "U.bin"), str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    rd = lambda name: np.fromfile(str(out / name), dtype=np.float64)  # noqa: E731
    ms = oracle.from_mesh(lmesh)
    pb = oracle.problem(2, int(pda.Euler2d.Riemann), int(S.Weno5), [1.4])
    Vo = oracle.velocity(ms, pb, U)
    V2o, rp, ci, Jo = oracle.velocity_and_jacobian(ms, pb, U)
    V, J, V2, J2 = rd("V.bin"), rd("J.bin"), rd("V2.bin"), rd("J2.bin")
    assert cases.rel_err(V, Vo) <= 1e-12
    assert cases.rel_err(V2, V2o) <= 1e-12
    assert cases.rel_err(J2, Jo) <= 1e-12
    assert cases.rel_err(J, Jo) <= 1e-12
    assert np.array_equal(rd("V3.bin"), V2) and np.array_equal(rd("J3.bin"), J2)  # rhsAndJacobian
    assert np.array_equal(rd("V4.bin"), V)                                          # operator()
    import scipy.sparse as sp
    Jsp = sp.csr_matrix((Jo, ci, rp), shape=(prob.totalDofSampleMesh(), prob.totalDofStencilMesh()))
    assert cases.rel_err(rd("AJ.bin"), Jsp @ rd("B.bin")) <= 1e-12
