// pda_batch.hip -- many independent problem instances evaluated and stepped
// together: each instance keeps its own mesh, parameters and state slice
// (the concatenation of the instances' states, in instance order), and one
// kernel launch evaluates the velocity of every instance.
//
// A batch accepts 2D structured problems of one family and scheme whose
// velocity box spans the whole grid (periodic or mirror-ghost boundaries,
// e.g. shallow water with slip walls), so an instance is exactly the
// workgroups of its row-marching velocity kernel.  The workgroup-to-instance
// table and the strip height (rows per workgroup) are built at creation;
// pda_batch_autotune times the candidate strip heights on the device and
// keeps the fastest.  Explicit steps use the stage formulas of pda_rk.hpp
// over the whole concatenated state, several steps recorded per hipGraph.
#include <algorithm>
#include <cmath>

#include "pda_rk.hpp"

struct pda_batch_s {
  std::vector<pda_problem_s *> probs;
  int family = 0, w = 0, m = 0;
  int64_t total = 0;                 // state doubles of the whole batch
  std::vector<int64_t> off;          // per instance state offset (count+1)
  int rps = 2;                       // rows per strip
  std::vector<pda::BatchInst> h_inst;
  std::vector<int32_t> h_blk;
  pda::DevBuf inst, blk;
  // stepping
  pda::DevBuf aux, k[4];
  hipStream_t cap = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  const double *g_U = nullptr;
  double g_dt = 0.;
  int g_scheme = -1, g_steps = 0;
  double tuned_us = 0.;
  ~pda_batch_s() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cap) (void)hipStreamDestroy(cap);
  }
};

namespace {
using namespace pda;

int build_tables(pda_batch_s &b, int rps) {
  b.h_inst.clear();
  b.h_blk.clear();
  int64_t blk = 0;
  for (size_t i = 0; i < b.probs.size(); ++i) {
    BatchInst I;
    I.P = b.probs[i]->view;
    I.B = b.probs[i]->vbox;
    I.off = b.off[i];
    I.blk0 = blk;
    I.rps = rps;
    const int64_t nb = velocity_march_blocks(I.B, 2, b.w, rps);
    b.h_inst.push_back(I);
    for (int64_t q = 0; q < nb; ++q) b.h_blk.push_back(int32_t(i));
    blk += nb;
  }
  if (blk >= (int64_t(1) << 31)) {
    set_error("batch: too many workgroups");
    return PDA_E_INVALID_ARG;
  }
  PDA_HIP_CHECK(b.inst.upload(b.h_inst));
  PDA_HIP_CHECK(b.blk.upload(b.h_blk));
  b.rps = rps;
  return PDA_OK;
}

int batch_velocity(const pda_batch_s &b, const double *U, double *V, hipStream_t s) {
  return launch_velocity_batch(b.family, b.w, b.inst.as<BatchInst>(), b.blk.as<int32_t>(),
                               int64_t(b.h_blk.size()), U, V, s);
}

// one SSPRK3 or RK4 step with every stage formed in the store of the
// batched velocity launch (the stage formulas and term order of rk_step,
// so the states equal rk_step's bit for bit; a stage never writes the
// vector its velocity reads); other schemes: rk_step
int batch_step(pda_batch_s &b, int scheme, double *U, double t, double dt, hipStream_t s) {
  double *k[4] = {b.k[0].as<double>(), b.k[1].as<double>(), b.k[2].as<double>(), b.k[3].as<double>()};
  double *aux = b.aux.as<double>();
  const int64_t nb = int64_t(b.h_blk.size());
  const BatchInst *inst = b.inst.as<BatchInst>();
  const int32_t *blk = b.blk.as<int32_t>();
  auto stage = [](double *kout, double *out, std::initializer_list<std::pair<double, const double *>> terms,
                  double cv) {
    RkStage S;
    S.kout = kout;
    S.out = out;
    for (auto &x : terms) {
      S.c[S.nx] = x.first;
      S.x[S.nx] = x.second;
      ++S.nx;
    }
    S.cv = cv;
    return S;
  };
  auto vs = [&](const double *y, const RkStage &S) {
    return launch_velocity_batch_stage(b.family, b.w, inst, blk, nb, y, S, s);
  };
  if (scheme == PDA_SSPRK3) {
    const double third = 1. / 3., twoThird = 2. / 3., quarter = 1. / 4., threeQuarter = 3. / 4.;
    double *aux2 = k[1];
    int rc = vs(U, stage(nullptr, aux, {{1., U}}, dt));
    if (!rc) rc = vs(aux, stage(nullptr, aux2, {{quarter, aux}, {threeQuarter, U}}, quarter * dt));
    if (!rc) rc = vs(aux2, stage(nullptr, U, {{third, U}, {twoThird, aux2}}, twoThird * dt));
    return rc;
  }
  if (scheme == PDA_RK4) {
    const double h2 = dt / 2., h6 = dt / 6., h3 = dt / 3.;
    double *aux2 = k[3];
    int rc = vs(U, stage(k[0], aux, {{1., U}}, h2));
    if (!rc) rc = vs(aux, stage(k[1], aux2, {{1., U}}, h2));
    if (!rc) rc = vs(aux2, stage(k[2], aux, {{1., U}}, dt));
    if (!rc) rc = vs(aux, stage(nullptr, U, {{1., U}, {h6, k[0]}, {h3, k[1]}, {h3, k[2]}}, h6));
    return rc;
  }
  return rk_step(
      scheme, [&](const double *y, double, double *f) { return batch_velocity(b, y, f, s); }, U, aux, k,
      b.total, t, dt, s);
}

void drop_graph(pda_batch_s &b) {
  if (b.exec) (void)hipGraphExecDestroy(b.exec);
  if (b.graph) (void)hipGraphDestroy(b.graph);
  b.exec = nullptr;
  b.graph = nullptr;
  b.g_U = nullptr;
}

}  // namespace

extern "C" {

int pda_batch_create(const pda_problem *probs, int count, pda_batch *out) {
  if (!probs || count < 1 || !out) {
    set_error("pda_batch_create: invalid argument");
    return PDA_E_INVALID_ARG;
  }
  auto *b = new pda_batch_s();
  b->family = probs[0] ? probs[0]->family : 0;
  b->w = probs[0] ? probs[0]->w : 0;
  b->m = probs[0] ? probs[0]->m : 0;
  b->off.push_back(0);
  for (int i = 0; i < count; ++i) {
    pda_problem_s *p = probs[i];
    if (!p || p->family != b->family || p->w != b->w || p->dim != 2 || !has_box_kernels(*p) ||
        !p->vbox.valid || !p->h_vrem_rows.empty() || p->mesh->n_sample != p->mesh->n_stencil) {
      set_error("pda_batch_create: instance " + std::to_string(i) +
                " is not a full-grid 2D structured problem of the batch's family and scheme");
      delete b;
      return PDA_E_INVALID_ARG;
    }
    const int rc = ensure_device(*p);
    if (rc) {
      delete b;
      return rc;
    }
    b->probs.push_back(p);
    b->off.push_back(b->off.back() + p->mesh->n_stencil * p->m);
  }
  b->total = b->off.back();
  int rc = build_tables(*b, 2);
  const size_t vb = sizeof(double) * size_t(b->total);
  hipError_t e = hipSuccess;
  if (!rc) {
    e = b->aux.alloc(vb);
    for (int i = 0; i < 4 && e == hipSuccess; ++i) e = b->k[i].alloc(vb);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking);
  }
  if (rc || e != hipSuccess) {
    if (e != hipSuccess) set_error(std::string("pda_batch_create: ") + hipGetErrorString(e));
    delete b;
    return rc ? rc : PDA_E_HIP;
  }
  *out = b;
  return PDA_OK;
}

void pda_batch_destroy(pda_batch b) { delete b; }

int pda_batch_sizes(pda_batch b, int64_t *total, int64_t *offsets) {
  if (!b) return PDA_E_INVALID_ARG;
  if (total) *total = b->total;
  if (offsets) std::copy(b->off.begin(), b->off.end(), offsets);
  return PDA_OK;
}

int pda_batch_initial_condition(pda_batch b, double *d_U, pda_stream stream) {
  if (!b || !d_U) return PDA_E_INVALID_ARG;
  for (size_t i = 0; i < b->probs.size(); ++i) {
    const int rc = pda_initial_condition(b->probs[i], d_U + b->off[i], stream);
    if (rc) return rc;
  }
  return PDA_OK;
}

int pda_batch_velocity(pda_batch b, const double *d_U, double, double *d_V, pda_stream stream) {
  if (!b || !d_U || !d_V) return PDA_E_INVALID_ARG;
  return batch_velocity(*b, d_U, d_V, hipStream_t(stream));
}

int pda_batch_autotune(pda_batch b, const double *d_U, pda_stream stream, int *rows_per_strip, double *us) {
  if (!b || !d_U) return PDA_E_INVALID_ARG;
  // the tables are re-uploaded below (new device buffers): a graph recorded
  // with the old ones must not be replayed, whatever the outcome
  drop_graph(*b);
  hipStream_t s = hipStream_t(stream);
  double *V = b->k[3].as<double>();
  // events owned for the whole call (every return path releases them)
  struct Events {
    hipEvent_t e[2] = {nullptr, nullptr};
    ~Events() {
      for (hipEvent_t x : e)
        if (x) (void)hipEventDestroy(x);
    }
  } ev;
  PDA_HIP_CHECK(hipEventCreate(&ev.e[0]));
  PDA_HIP_CHECK(hipEventCreate(&ev.e[1]));
  hipEvent_t e0 = ev.e[0], e1 = ev.e[1];
  int best = b->rps;
  double best_us = 1e300;
  for (int cand : {1, 2, 3, 4, 6, 8, 12, 16, 24, 32}) {
    int rc = build_tables(*b, cand);
    if (rc) return rc;
    rc = batch_velocity(*b, d_U, V, s); // warm-up
    if (rc) return rc;
    PDA_HIP_CHECK(hipEventRecord(e0, s));
    for (int r = 0; r < 5; ++r) {
      rc = batch_velocity(*b, d_U, V, s);
      if (rc) return rc;
    }
    PDA_HIP_CHECK(hipEventRecord(e1, s));
    PDA_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    PDA_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    const double t = 1e3 * ms / 5.;
    if (t < best_us) {
      best_us = t;
      best = cand;
    }
  }
  const int rc = build_tables(*b, best);
  if (rc) return rc;
  b->tuned_us = best_us;
  if (rows_per_strip) *rows_per_strip = best;
  if (us) *us = best_us;
  return PDA_OK;
}

int pda_batch_set_velocity_strip_height(pda_batch b, int rows_per_strip) {
  if (!b || rows_per_strip < 1 || rows_per_strip > 64) {
    set_error("pda_batch_set_velocity_strip_height: rows_per_strip must be 1..64");
    return PDA_E_INVALID_ARG;
  }
  // work in flight may still read the tables that are about to be replaced
  PDA_HIP_CHECK(hipDeviceSynchronize());
  drop_graph(*b);
  return build_tables(*b, rows_per_strip);
}

int pda_batch_advance(pda_batch b, int scheme, double *d_U, double t0, double dt, int64_t nsteps,
                      int steps_per_graph, pda_stream stream) {
  if (!b || !d_U || nsteps < 0 || scheme < PDA_FORWARD_EULER || scheme > PDA_MODULE_SSP3) {
    set_error("pda_batch_advance: invalid argument");
    return PDA_E_INVALID_ARG;
  }
  hipStream_t s = hipStream_t(stream);
  if (steps_per_graph <= 0) {
    for (int64_t i = 0; i < nsteps; ++i) {
      const int rc = batch_step(*b, scheme, d_U, t0 + double(i) * dt, dt, s);
      if (rc) return rc;
    }
    return PDA_OK;
  }
  const int spg = int(std::min<int64_t>(steps_per_graph, std::max<int64_t>(nsteps, 1)));
  if (!b->exec || b->g_U != d_U || b->g_dt != dt || b->g_scheme != scheme || b->g_steps != spg) {
    drop_graph(*b);
    PDA_HIP_CHECK(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
    int rc = PDA_OK;
    for (int i = 0; i < spg && !rc; ++i) rc = batch_step(*b, scheme, d_U, t0 + double(i) * dt, dt, b->cap);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->cap, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) {
      set_error(std::string("pda_batch_advance: capture: ") + hipGetErrorString(e));
      return PDA_E_HIP;
    }
    b->graph = g;
    PDA_HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->g_U = d_U;
    b->g_dt = dt;
    b->g_scheme = scheme;
    b->g_steps = spg;
  }
  int64_t done = 0;
  for (; done + spg <= nsteps; done += spg) PDA_HIP_CHECK(hipGraphLaunch(b->exec, s));
  for (; done < nsteps; ++done) {
    const int rc = batch_step(*b, scheme, d_U, t0 + double(done) * dt, dt, s);
    if (rc) return rc;
  }
  return PDA_OK;
}

}  // extern "C"
