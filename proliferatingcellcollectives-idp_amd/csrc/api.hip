// api.hip — the C ABI of libpcc (include/pcc.h) and the per-step orchestration.
//
// pcc_step() follows ParticleManager::step (ParticleManager.cpp:150-166) and
// dispatches to the hard-model ReLCP loop (solveHardModel, Hard.cpp:91-193) or
// the soft model (solveSoftPotential, Soft.cpp:21-114).  All device work runs
// on the context's stream; the host only reads the handful of scalars that
// steer the ReLCP loop (constraint count, max overlap, BBPGD flags).
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "common.hpp"
#include "internal.hpp"

namespace pcc {

void detect_batch(pcc_ctx* ctx, double cutoff, int32_t iteration);
void batch_to_host(pcc_ctx* ctx, pcc_constraints* out, int32_t iteration);
void build_grid(pcc_ctx* ctx);

BodyArrays body_arrays(double* base, int64_t N) {
  BodyArrays b;
  double** f[kBodyFields] = {&b.px, &b.py, &b.pz, &b.qw, &b.qx, &b.qy,
                             &b.qz, &b.len, &b.l0, &b.diam, &b.ldot};
  for (int k = 0; k < kBodyFields; ++k) *f[k] = base + (int64_t)k * N;
  return b;
}

KTimer::KTimer(pcc_ctx* c, int kernel, double bytes) : ctx(c), k(kernel) {
  Timing& T = ctx->timing;
  if (!T.enabled) return;
  auto get = [&]() {
    hipEvent_t e;
    if (!T.free_events.empty()) {
      e = T.free_events.back();
      T.free_events.pop_back();
    } else {
      PCC_HIP(hipEventCreate(&e));
    }
    return e;
  };
  a = get();
  b = get();
  PCC_HIP(hipEventRecord(a, ctx->stream));
  T.pending.push_back({a, b, kernel, bytes});
}

KTimer::~KTimer() {
  if (a) (void)hipEventRecord(b, ctx->stream);
}

static void resolve_timing(pcc_ctx* ctx) {
  Timing& T = ctx->timing;
  if (T.pending.empty()) return;
  PCC_HIP(hipStreamSynchronize(ctx->stream));
  for (auto& p : T.pending) {
    float ms = 0.f;
    PCC_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    T.launches[p.kernel] += 1;
    T.total_ms[p.kernel] += ms;
    T.bytes[p.kernel] += p.bytes;  // summed; reported per launch
    T.free_events.push_back(p.a);
    T.free_events.push_back(p.b);
  }
  T.pending.clear();
}

__global__ void k_reset(int64_t N, double* __restrict__ ldot, int32_t* __restrict__ nc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  ldot[i] = 0.0;  // Particle::reset (Particle.h:146-149)
  nc[i] = 0;
}

void reset_step(pcc_ctx* ctx) {
  const int64_t N = ctx->N;
  BodyArrays b = body_arrays(ctx->body.p, N);
  if (N > 0) k_reset<<<(unsigned)((N + 255) / 256), 256, 0, ctx->stream>>>(N, b.ldot, ctx->nc.p);
  ctx->S.C = 0;
}

}  // namespace pcc

using namespace pcc;

namespace {

pcc_status fail(pcc_ctx* ctx, pcc_status s, const std::string& msg) {
  if (ctx) ctx->last_error = msg;
  return s;
}

template <class F>
pcc_status guarded(pcc_ctx* ctx, F&& f) {
  try {
    return f();
  } catch (const HipError& e) {
    return fail(ctx, PCC_E_DEVICE, std::string(e.what) + ": " + hipGetErrorString(e.err));
  } catch (const StatusError& e) {
    return fail(ctx, e.status, e.msg);
  } catch (const std::exception& e) {
    return fail(ctx, PCC_E_DEVICE, e.what());
  }
}

pcc_status check_params(const pcc_params* p, std::string* why) {
  if (p->mode != PCC_MODE_HARD && p->mode != PCC_MODE_SOFT) {
    *why = "Invalid mode";
    return PCC_E_INVALID_MODE;
  }
  if (p->precision != PCC_PREC_FP64 && p->precision != PCC_PREC_MIXED) {
    *why = "invalid precision";
    return PCC_E_INVALID_ARGUMENT;
  }
  if (p->temperature != 0.0) {
    *why = "temperature > 0 (Brownian input) is not supported";
    return PCC_E_UNSUPPORTED;
  }
  if (!(p->linked_cell_size > 0) || !(p->dt > 0) || !(p->xi > 0) || !(p->tau != 0) ||
      p->max_bbpgd_iterations < 1 || p->max_recursive_iterations < 1) {
    *why = "invalid parameter value";
    return PCC_E_INVALID_ARGUMENT;
  }
  return PCC_OK;
}

void copy_set_to_host(pcc_ctx* ctx, pcc_constraints* out) {
  const ConstraintSet& S = ctx->S;
  const int64_t m = std::min<int64_t>(S.C, out->capacity);
  if (m <= 0) return;
  hipStream_t s = ctx->stream;
  auto cp = [&](void* dst, const void* src, size_t bytes) {
    if (dst) PCC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
  };
  cp(out->gidI, S.I.p, sizeof(int32_t) * m);
  cp(out->gidJ, S.J.p, sizeof(int32_t) * m);
  cp(out->signed_distance, S.phi.p, sizeof(double) * m);
  cp(out->normal, S.n.p, sizeof(double) * 3 * m);
  cp(out->rPosI, S.rI.p, sizeof(double) * 3 * m);
  cp(out->rPosJ, S.rJ.p, sizeof(double) * 3 * m);
  cp(out->contact_point, S.p.p, sizeof(double) * 3 * m);
  cp(out->stressI, S.sI.p, sizeof(double) * m);
  cp(out->stressJ, S.sJ.p, sizeof(double) * m);
  cp(out->gamma, S.gamma.p, sizeof(double) * m);
  cp(out->gid, S.gid.p, sizeof(int32_t) * m);
  cp(out->iteration, S.iter.p, sizeof(int32_t) * m);
  PCC_HIP(hipStreamSynchronize(s));
}

pcc_status hard_step(pcc_ctx* ctx, pcc_step_result* out) {
  const pcc_params& p = ctx->params;
  int32_t k = 0;
  int64_t total_it = 0;
  double res = 0.0, maxov = 0.0;
  reset_step(ctx);
  ctx->planar = ctx->allow_planar && bodies_planar(ctx);
  ctx->phase_iters.clear();
  ctx->phase_C.clear();
  pcc_status st = PCC_OK;
  while (k < p.max_recursive_iterations) {  // Hard.cpp:113
    const double cutoff = k == 0 ? p.cutoff_phase0 : p.cutoff_phaseK;  // Hard.cpp:121
    detect_batch(ctx, cutoff, k);
    append_batch(ctx, k);  // Hard.cpp:123
    maxov = max_overlap(ctx);  // Hard.cpp:125-130
    if (maxov <= p.relcp_tol && k > 0) break;  // Hard.cpp:132-134
    build_incidence(ctx, true);  // createMatrices, Hard.cpp:136
    int64_t its = 0;
    st = bbpgd_solve(ctx, &its, &res);  // Hard.cpp:153
    total_it += its;
    ctx->phase_iters.push_back(its);
    ctx->phase_C.push_back(ctx->S.C);
    if (st != PCC_OK) {
      ctx->last_error = "BBPGD did not converge";
      break;
    }
    integrate_hard(ctx, its);  // Hard.cpp:160-170
    ++k;
  }
  if (st == PCC_OK) {
    grow(ctx);  // Hard.cpp:185
    if (k == p.max_recursive_iterations) {  // Hard.cpp:187-190
      st = PCC_E_RELCP_MAX_ITERS;
      ctx->last_error = "Maximum number of constraint iterations reached";
    }
  }
  PCC_HIP(hipStreamSynchronize(ctx->stream));
  out->n_constraints = ctx->S.C;
  out->constraint_iterations = k;
  out->reserved = 0;
  out->bbpgd_iterations = total_it;
  out->residual = res;
  out->max_overlap = maxov;
  return st;
}

pcc_status soft_step(pcc_ctx* ctx, pcc_step_result* out) {
  reset_step(ctx);
  detect_batch(ctx, 0.0, 0);  // Soft.cpp:35
  append_batch(ctx, 0);
  const double maxov = max_overlap(ctx);  // Soft.cpp:37-41
  build_incidence(ctx, false);
  soft_forces_and_integrate(ctx);
  PCC_HIP(hipStreamSynchronize(ctx->stream));
  out->n_constraints = ctx->S.C;
  out->constraint_iterations = 0;
  out->reserved = 0;
  out->bbpgd_iterations = 0;
  out->residual = maxov;
  out->max_overlap = maxov;
  return PCC_OK;
}

}  // namespace

extern "C" {

void pcc_default_params(pcc_params* p) {
  memset(p, 0, sizeof(*p));
  p->mode = PCC_MODE_HARD;
  p->precision = PCC_PREC_FP64;
  p->dt = 1e-2;
  p->xi = 1.0;
  p->tau = 1.0;
  p->lambda = 1e-2;
  p->temperature = 0.0;
  p->kcc = 20000.0;
  p->bbpgd_tol = 1e-3;
  p->relcp_tol = 1e-3;
  p->cutoff_phase0 = 0.25;
  p->cutoff_phaseK = 1e-3;
  p->max_bbpgd_iterations = 500000;
  p->max_recursive_iterations = 50;
  p->linked_cell_size = 2.2;
  p->grid_origin = -1.0;
  p->warm_start_factor = 500.0;
}

const char* pcc_status_string(pcc_status s) {
  switch (s) {
    case PCC_OK: return "ok";
    case PCC_E_BBPGD_NOT_CONVERGED: return "BBPGD did not converge";
    case PCC_E_RELCP_MAX_ITERS: return "Maximum number of constraint iterations reached";
    case PCC_E_INVALID_MODE: return "Invalid mode";
    case PCC_E_INVALID_ARGUMENT: return "invalid argument";
    case PCC_E_DEVICE: return "HIP runtime error";
    case PCC_E_OUT_OF_MEMORY: return "out of device memory";
    case PCC_E_UNSUPPORTED: return "unsupported parameter";
    case PCC_E_NO_STATE: return "no body state set";
  }
  return "unknown status";
}

pcc_status pcc_create(const pcc_params* params, int device, pcc_ctx** out) {
  if (!params || !out) return PCC_E_INVALID_ARGUMENT;
  *out = nullptr;
  std::string why;
  pcc_status ps = check_params(params, &why);
  if (ps != PCC_OK) return ps;
  pcc_ctx* ctx = new pcc_ctx();
  ctx->params = *params;
  ctx->device = device;
  const char* planar_env = getenv("PCC_PLANAR");
  ctx->allow_planar = !(planar_env && planar_env[0] == '0');
  pcc_status s = guarded(ctx, [&]() {
    PCC_HIP(hipSetDevice(device));
    PCC_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    PCC_HIP(hipHostMalloc((void**)&ctx->state_host, 4 * sizeof(SolverState), hipHostMallocDefault));
    ctx->state.ensure(1);
    ctx->ticket.ensure(1);
    PCC_HIP(hipMemsetAsync(ctx->ticket.p, 0, sizeof(uint32_t) * ctx->ticket.cap, ctx->stream));
    PCC_HIP(hipStreamSynchronize(ctx->stream));
    return PCC_OK;
  });
  if (s != PCC_OK) {
    pcc_destroy(ctx);
    return s;
  }
  *out = ctx;
  return PCC_OK;
}

void pcc_destroy(pcc_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto& p : ctx->timing.pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : ctx->timing.free_events) (void)hipEventDestroy(e);
  if (ctx->state_host) (void)hipHostFree(ctx->state_host);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;  // DArr members free their device memory
}

pcc_status pcc_set_params(pcc_ctx* ctx, const pcc_params* params) {
  if (!ctx || !params) return PCC_E_INVALID_ARGUMENT;
  std::string why;
  pcc_status s = check_params(params, &why);
  if (s != PCC_OK) return fail(ctx, s, why);
  ctx->params = *params;
  return PCC_OK;
}

const char* pcc_last_error(const pcc_ctx* ctx) { return ctx ? ctx->last_error.c_str() : ""; }

pcc_status pcc_set_state(pcc_ctx* ctx, const pcc_bodies* b) {
  if (!ctx || !b || b->n < 0) return PCC_E_INVALID_ARGUMENT;
  if (b->n > 0 && (!b->pos || !b->quat || !b->length || !b->l0 || !b->diameter || !b->ldot))
    return fail(ctx, PCC_E_INVALID_ARGUMENT, "missing body field");
  if (b->n >= (int64_t)1 << 30) return fail(ctx, PCC_E_INVALID_ARGUMENT, "too many bodies");
  return guarded(ctx, [&]() {
    PCC_HIP(hipSetDevice(ctx->device));
    const int64_t N = b->n;
    std::vector<double> h((size_t)kBodyFields * N);
    for (int64_t i = 0; i < N; ++i) {
      for (int k = 0; k < 3; ++k) h[(size_t)k * N + i] = b->pos[3 * i + k];
      for (int k = 0; k < 4; ++k) h[(size_t)(3 + k) * N + i] = b->quat[4 * i + k];
      h[(size_t)7 * N + i] = b->length[i];
      h[(size_t)8 * N + i] = b->l0[i];
      h[(size_t)9 * N + i] = b->diameter[i];
      h[(size_t)10 * N + i] = b->ldot[i];
    }
    ctx->N = N;
    ctx->body.ensure(kBodyFields * N + 1);
    ctx->F.ensure(6 * N + 6);
    ctx->U.ensure(6 * N + 6);
    ctx->sig.ensure(N + 1);
    ctx->imp.ensure(N + 1);
    ctx->ldot_new.ensure(N + 1);
    ctx->nc.ensure(N + 1);
    if (N > 0)
      PCC_HIP(hipMemcpyAsync(ctx->body.p, h.data(), sizeof(double) * h.size(),
                             hipMemcpyHostToDevice, ctx->stream));
    PCC_HIP(hipMemsetAsync(ctx->F.p, 0, sizeof(double) * (6 * N + 6), ctx->stream));
    PCC_HIP(hipMemsetAsync(ctx->U.p, 0, sizeof(double) * (6 * N + 6), ctx->stream));
    PCC_HIP(hipMemsetAsync(ctx->sig.p, 0, sizeof(double) * (N + 1), ctx->stream));
    PCC_HIP(hipMemsetAsync(ctx->imp.p, 0, sizeof(double) * (N + 1), ctx->stream));
    PCC_HIP(hipMemsetAsync(ctx->nc.p, 0, sizeof(int32_t) * (N + 1), ctx->stream));
    PCC_HIP(hipStreamSynchronize(ctx->stream));
    ctx->S.C = 0;
    ctx->has_state = true;
    ctx->has_snapshot = false;
    return PCC_OK;
  });
}

pcc_status pcc_get_state(pcc_ctx* ctx, pcc_bodies* b) {
  if (!ctx || !b) return PCC_E_INVALID_ARGUMENT;
  if (!ctx->has_state) return fail(ctx, PCC_E_NO_STATE, "no state");
  if (b->n != ctx->N) return fail(ctx, PCC_E_INVALID_ARGUMENT, "body count mismatch");
  return guarded(ctx, [&]() {
    PCC_HIP(hipSetDevice(ctx->device));
    const int64_t N = ctx->N;
    std::vector<double> h((size_t)kBodyFields * N), F(6 * N), U(6 * N), sg(N), im(N);
    std::vector<int32_t> nc(N);
    hipStream_t s = ctx->stream;
    if (N > 0) {
      PCC_HIP(hipMemcpyAsync(h.data(), ctx->body.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
      PCC_HIP(hipMemcpyAsync(F.data(), ctx->F.p, sizeof(double) * 6 * N, hipMemcpyDeviceToHost, s));
      PCC_HIP(hipMemcpyAsync(U.data(), ctx->U.p, sizeof(double) * 6 * N, hipMemcpyDeviceToHost, s));
      PCC_HIP(hipMemcpyAsync(sg.data(), ctx->sig.p, sizeof(double) * N, hipMemcpyDeviceToHost, s));
      PCC_HIP(hipMemcpyAsync(im.data(), ctx->imp.p, sizeof(double) * N, hipMemcpyDeviceToHost, s));
      PCC_HIP(hipMemcpyAsync(nc.data(), ctx->nc.p, sizeof(int32_t) * N, hipMemcpyDeviceToHost, s));
    }
    PCC_HIP(hipStreamSynchronize(s));
    for (int64_t i = 0; i < N; ++i) {
      if (b->pos)
        for (int k = 0; k < 3; ++k) b->pos[3 * i + k] = h[(size_t)k * N + i];
      if (b->quat)
        for (int k = 0; k < 4; ++k) b->quat[4 * i + k] = h[(size_t)(3 + k) * N + i];
      if (b->length) b->length[i] = h[(size_t)7 * N + i];
      if (b->l0) b->l0[i] = h[(size_t)8 * N + i];
      if (b->diameter) b->diameter[i] = h[(size_t)9 * N + i];
      if (b->ldot) b->ldot[i] = h[(size_t)10 * N + i];
      for (int k = 0; k < 3; ++k) {
        if (b->force) b->force[3 * i + k] = F[6 * i + k];
        if (b->torque) b->torque[3 * i + k] = F[6 * i + 3 + k];
        if (b->velocity) b->velocity[3 * i + k] = U[6 * i + k];
        if (b->omega) b->omega[3 * i + k] = U[6 * i + 3 + k];
      }
      if (b->impedance) b->impedance[i] = im[i];
      if (b->stress) b->stress[i] = sg[i];
      if (b->num_constraints) b->num_constraints[i] = nc[i];
    }
    return PCC_OK;
  });
}

pcc_status pcc_snapshot(pcc_ctx* ctx) {
  if (!ctx) return PCC_E_INVALID_ARGUMENT;
  if (!ctx->has_state) return fail(ctx, PCC_E_NO_STATE, "no state");
  return guarded(ctx, [&]() {
    PCC_HIP(hipSetDevice(ctx->device));
    ctx->snap.ensure(kBodyFields * ctx->N + 1);
    PCC_HIP(hipMemcpyAsync(ctx->snap.p, ctx->body.p, sizeof(double) * kBodyFields * ctx->N,
                           hipMemcpyDeviceToDevice, ctx->stream));
    PCC_HIP(hipStreamSynchronize(ctx->stream));
    ctx->has_snapshot = true;
    return PCC_OK;
  });
}

pcc_status pcc_restore(pcc_ctx* ctx) {
  if (!ctx) return PCC_E_INVALID_ARGUMENT;
  if (!ctx->has_snapshot) return fail(ctx, PCC_E_NO_STATE, "no snapshot");
  return guarded(ctx, [&]() {
    PCC_HIP(hipSetDevice(ctx->device));
    PCC_HIP(hipMemcpyAsync(ctx->body.p, ctx->snap.p, sizeof(double) * kBodyFields * ctx->N,
                           hipMemcpyDeviceToDevice, ctx->stream));
    PCC_HIP(hipStreamSynchronize(ctx->stream));
    return PCC_OK;
  });
}

pcc_status pcc_step(pcc_ctx* ctx, int32_t iter, pcc_step_result* out) {
  (void)iter;  // the reference uses it only to gate BBPGD tracing (Hard.cpp:109)
  if (!ctx || !out) return PCC_E_INVALID_ARGUMENT;
  memset(out, 0, sizeof(*out));
  if (!ctx->has_state) return fail(ctx, PCC_E_NO_STATE, "no state");
  std::string why;
  pcc_status ps = check_params(&ctx->params, &why);
  if (ps != PCC_OK) return fail(ctx, ps, why);
  pcc_status s = guarded(ctx, [&]() {
    PCC_HIP(hipSetDevice(ctx->device));
    if (ctx->params.mode == PCC_MODE_HARD) return hard_step(ctx, out);
    return soft_step(ctx, out);
  });
  pcc_status t = guarded(ctx, [&]() {
    resolve_timing(ctx);
    return PCC_OK;
  });
  return s != PCC_OK ? s : t;
}

pcc_status pcc_get_constraints(pcc_ctx* ctx, pcc_constraints* out, int64_t* count) {
  if (!ctx || !out) return PCC_E_INVALID_ARGUMENT;
  if (count) *count = ctx->S.C;
  return guarded(ctx, [&]() {
    PCC_HIP(hipSetDevice(ctx->device));
    copy_set_to_host(ctx, out);
    return PCC_OK;
  });
}

pcc_status pcc_detect(pcc_ctx* ctx, double cutoff, pcc_constraints* out, int64_t* count,
                      uint32_t* bucket_out, int32_t* perm_out) {
  if (!ctx) return PCC_E_INVALID_ARGUMENT;
  if (!ctx->has_state) return fail(ctx, PCC_E_NO_STATE, "no state");
  return guarded(ctx, [&]() {
    PCC_HIP(hipSetDevice(ctx->device));
    detect_batch(ctx, cutoff, 0);
    if (count) *count = ctx->batch_count;
    batch_to_host(ctx, out, 0);
    const int64_t N = ctx->N;
    if (N > 0 && bucket_out)
      PCC_HIP(hipMemcpyAsync(bucket_out, ctx->bucket.p, sizeof(uint32_t) * N,
                             hipMemcpyDeviceToHost, ctx->stream));
    if (N > 0 && perm_out)
      PCC_HIP(hipMemcpyAsync(perm_out, ctx->perm.p, sizeof(int32_t) * N, hipMemcpyDeviceToHost,
                             ctx->stream));
    PCC_HIP(hipStreamSynchronize(ctx->stream));
    resolve_timing(ctx);
    return PCC_OK;
  });
}

pcc_status pcc_get_phase_stats(pcc_ctx* ctx, int64_t* bbpgd_iterations, int64_t* n_constraints,
                               int64_t max, int64_t* count) {
  if (!ctx) return PCC_E_INVALID_ARGUMENT;
  const int64_t n = (int64_t)ctx->phase_iters.size();
  if (count) *count = n;
  for (int64_t i = 0; i < n && i < max; ++i) {
    if (bbpgd_iterations) bbpgd_iterations[i] = ctx->phase_iters[i];
    if (n_constraints) n_constraints[i] = ctx->phase_C[i];
  }
  return PCC_OK;
}

pcc_status pcc_set_timing(pcc_ctx* ctx, int enable) {
  if (!ctx) return PCC_E_INVALID_ARGUMENT;
  return guarded(ctx, [&]() {
    resolve_timing(ctx);
    Timing& T = ctx->timing;
    T.enabled = enable != 0;
    for (int k = 0; k < PCC_K_COUNT; ++k) {
      T.launches[k] = 0;
      T.total_ms[k] = 0;
      T.bytes[k] = 0;
    }
    return PCC_OK;
  });
}

pcc_status pcc_get_timing(pcc_ctx* ctx, int kernel, int64_t* launches, double* total_ms,
                          double* bytes_per_launch) {
  if (!ctx || kernel < 0 || kernel >= PCC_K_COUNT) return PCC_E_INVALID_ARGUMENT;
  return guarded(ctx, [&]() {
    resolve_timing(ctx);
    const Timing& T = ctx->timing;
    if (launches) *launches = T.launches[kernel];
    if (total_ms) *total_ms = T.total_ms[kernel];
    if (bytes_per_launch)
      *bytes_per_launch = T.launches[kernel] ? T.bytes[kernel] / (double)T.launches[kernel] : 0.0;
    return PCC_OK;
  });
}

const char* pcc_kernel_name(int kernel) {
  static const char* names[PCC_K_COUNT] = {"body_apply", "constraint_apply", "detect",
                                           "assemble",   "integrate",        "hertz"};
  return (kernel >= 0 && kernel < PCC_K_COUNT) ? names[kernel] : "?";
}

}  // extern "C"
