// adapter_check.cpp -- drives include/pda_hip_adapter.hpp the way a caller of
// the reference's C++ API (adapter_cpp.hpp) does; the pytest suite compares
// what it prints and writes with the Python module and the oracle.
//
//   adapter_check host <meshdir>
//       no device: mesh, factories, sizes, pattern checksum, errors
//   adapter_check gpu <meshdir> <state.bin> <outdir>
//       Riemann WENO5 problem on the device: rightHandSide, jacobian,
//       rightHandSideAndJacobian, rhsAndJacobian, operator(), applyJacobian;
//       results written as raw doubles to <outdir>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <numeric>
#include <string>
#include <vector>

#include "pda_hip_adapter.hpp"

namespace pda = pressiodemoapps::hip;

static void write(const std::string &path, const std::vector<double> &v) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char *>(v.data()), std::streamsize(v.size() * sizeof(double)));
}

static int host_mode(const std::string &dir) {
  auto mesh = pda::load_cellcentered_uniform_mesh_eigen(dir);
  std::printf("mesh %d %d %lld %lld\n", mesh.dimensionality(), mesh.stencilSize(),
              (long long)mesh.sampleMeshSize(), (long long)mesh.stencilMeshSize());
  auto a = pda::create_problem_eigen(mesh, pda::Euler2d::Riemann, pda::InviscidFluxReconstruction::Weno5, 2);
  auto b = pda::create_problem_eigen(mesh, pda::Swe2d::SlipWall, pda::InviscidFluxReconstruction::FirstOrder, 1,
                                     {{"gravity", 7.5}, {"coriolis", -2.5}});
  auto c = pda::create_advdiffreac_2d_problem_A_eigen(mesh, pda::InviscidFluxReconstruction::Weno3, 0.3, -0.2,
                                                      0.01, 0.5);
  for (const pda::Problem *p : {&a, &b, &c}) {
    std::vector<int64_t> rp;
    std::vector<int32_t> ci;
    p->jacobianPatternHost(rp, ci);
    const long long csum = std::accumulate(ci.begin(), ci.end(), 0LL);
    std::printf("problem %d %lld %lld %lld %lld\n", p->numDofPerCell(), (long long)p->totalDofSampleMesh(),
                (long long)p->totalDofStencilMesh(), (long long)rp.back(), csum);
  }
  std::printf("gravity %.17g coriolis %.17g\n", b.queryParameter("gravity"), b.queryParameter("coriolis"));
  // the reference throws std::runtime_error for invalid input
  try {
    pda::create_problem_eigen(mesh, pda::Euler2d::Riemann, pda::InviscidFluxReconstruction::FirstOrder, 2,
                              {{"notAParameter", 1.0}});
    std::printf("error none\n");
  } catch (const std::runtime_error &e) {
    std::printf("error parameter: %s\n", e.what());
  }
  try {
    pda::load_cellcentered_uniform_mesh_eigen(dir + "/does-not-exist");
    std::printf("error none\n");
  } catch (const std::runtime_error &e) {
    std::printf("error mesh: %s\n", e.what());
  }
  return 0;
}

static int gpu_mode(const std::string &dir, const std::string &state, const std::string &out) {
  auto mesh = pda::load_cellcentered_uniform_mesh_eigen(dir);
  auto prob = pda::create_problem_eigen(mesh, pda::Euler2d::Riemann, pda::InviscidFluxReconstruction::Weno5, 2);
  std::vector<double> h(size_t(prob.totalDofStencilMesh()));
  std::ifstream f(state, std::ios::binary);
  f.read(reinterpret_cast<char *>(h.data()), std::streamsize(h.size() * sizeof(double)));
  if (!f) {
    std::fprintf(stderr, "cannot read %s\n", state.c_str());
    return 2;
  }
  auto U = prob.createState();
  U.copyFromHost(h.data());
  auto V = prob.createRightHandSide();
  prob.rightHandSide(U, 0.0, V);
  auto J = prob.createJacobian();
  prob.jacobian(U, 0.0, J);
  auto V2 = prob.createRhs();
  auto J2 = prob.createJacobian();
  prob.rightHandSideAndJacobian(U, 0.0, V2, J2);
  auto V3 = prob.createRightHandSide();
  auto J3 = J2.clone();
  prob.rhsAndJacobian(U, 0.0, V3, &J3);
  auto V4 = prob.createRightHandSide();
  prob(U, 0.0, V4);
  pda::DeviceVector B(prob.totalDofStencilMesh());
  std::vector<double> hb(h.size());
  for (size_t i = 0; i < hb.size(); ++i) hb[i] = 1.0 / double(1 + i % 7);
  B.copyFromHost(hb.data());
  auto AJ = prob.createApplyJacobianResult(B);
  prob.applyJacobian(U, B, 0.0, AJ);
  pda::check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
  write(out + "/V.bin", V.toHost());
  write(out + "/J.bin", J.values().toHost());
  write(out + "/V2.bin", V2.toHost());
  write(out + "/J2.bin", J2.values().toHost());
  write(out + "/V3.bin", V3.toHost());
  write(out + "/J3.bin", J3.values().toHost());
  write(out + "/V4.bin", V4.toHost());
  write(out + "/B.bin", hb);
  write(out + "/AJ.bin", AJ.toHost());
  std::printf("gpu ok %lld %lld\n", (long long)V.size(), (long long)J.nonZeros());
  return 0;
}

int main(int argc, char **argv) {
  try {
    if (argc >= 3 && std::string(argv[1]) == "host") return host_mode(argv[2]);
    if (argc >= 5 && std::string(argv[1]) == "gpu") return gpu_mode(argv[2], argv[3], argv[4]);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "exception: %s\n", e.what());
    return 1;
  }
  std::fprintf(stderr, "usage: adapter_check host <meshdir> | gpu <meshdir> <state.bin> <outdir>\n");
  return 2;
}
