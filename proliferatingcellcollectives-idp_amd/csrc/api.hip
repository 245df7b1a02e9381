// api.hip — the C ABI of libpcc (include/pcc.h) and the per-step orchestration.
//
// pcc_step() follows ParticleManager::step (ParticleManager.cpp:150-166) and
// dispatches to the hard-model ReLCP loop (solveHardModel, Hard.cpp:91-193) or
// the soft model (solveSoftPotential, Soft.cpp:21-114).  All device work runs
// on the context's stream; the host only reads the handful of scalars that
// steer the ReLCP loop (constraint counts, max overlaps, BBPGD flags).
//
// A context holds one colony (pcc_set_state) or a batch of independent
// colonies (pcc_batch_*).  A batched step runs every colony's ReLCP loop in
// lockstep: phase k detects, assembles and solves all colonies still in their
// loop with the same launches; a colony leaves the loop exactly where it
// would alone, and its results are the ones it would get alone.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "common.hpp"
#include "internal.hpp"

namespace pcc {

void batch_to_host(pcc_ctx* ctx, pcc_constraints* out, int32_t iteration);

std::recursive_mutex& device_mutex() {
  static std::recursive_mutex m;
  return m;
}

std::string& thread_error() {
  static thread_local std::string e;
  return e;
}

hipStream_t& alloc_stream() {
  static thread_local hipStream_t s = nullptr;
  return s;
}

void* dev_alloc(size_t bytes, bool* pooled) {
  void* q = nullptr;
  const hipStream_t s = alloc_stream();
  const hipError_t e = s ? hipMallocAsync(&q, bytes, s) : hipMalloc(&q, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw StatusError{PCC_E_OUT_OF_MEMORY, "device allocation failed"};
  }
  *pooled = s != nullptr;
  return q;
}

void dev_free(void* p, bool pooled) {
  const hipStream_t s = alloc_stream();
  if (pooled && s)
    (void)hipFreeAsync(p, s);
  else
    (void)hipFree(p);
}

// Keeps freed pool memory in the process, so that a buffer freed on one
// context's stream serves the next allocation without a trip to the driver.
static void keep_pool_memory(int device) {
  static std::once_flag once[64];
  if (device < 0 || device >= 64) return;
  std::call_once(once[device], [device]() {
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool) {
      uint64_t thr = UINT64_MAX;
      (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
    }
  });
}

BodyArrays body_arrays(double* base, int64_t N) {
  BodyArrays b;
  double** f[kBodyFields] = {&b.px, &b.py, &b.pz,   &b.qw,   &b.qx,  &b.qy, &b.qz,
                             &b.len, &b.l0, &b.diam, &b.ldot, &b.age, &b.imp};
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
      DeviceLock lock(device_mutex());
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

// Per-body outputs of the last step cleared after a division (F, U: 6 per
// body; sigma, ldot_new; constraint counts), in one launch.
__global__ void k_clear_outputs(int64_t N, double* __restrict__ F, double* __restrict__ U,
                                double* __restrict__ sig, double* __restrict__ ldn,
                                int32_t* __restrict__ nc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > N) return;  // the +1 entries too, as the arrays' full length
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    F[6 * i + k] = 0.0;
    U[6 * i + k] = 0.0;
  }
  sig[i] = 0.0;
  ldn[i] = 0.0;
  nc[i] = 0;
}

// Colony of every body from the colony records on the device (colonies are
// stored one after another in record order).
__global__ void k_binst(int64_t N, int64_t B, const InstDev* __restrict__ idev,
                        int32_t* __restrict__ binst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int64_t lo = 0, hi = B - 1;  // last colony with boff <= i
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) / 2;
    if (idev[mid].boff <= i)
      lo = mid;
    else
      hi = mid - 1;
  }
  binst[i] = (int32_t)lo;
}

void reset_step(pcc_ctx* ctx) {
  const int64_t N = ctx->N;
  BodyArrays b = body_arrays(ctx->body.p, N);
  if (N > 0) k_reset<<<(unsigned)((N + 255) / 256), 256, 0, ctx->stream>>>(N, b.ldot, ctx->nc.p);
  ctx->S.C = 0;
  ctx->append_log.clear();
  ctx->set_unmerged = false;
  for (auto& I : ctx->inst) {
    I.coff = 0;
    I.C = 0;
    I.m = 0;
    I.bstart = 0;
  }
}

// Instance records to the device, through a pinned staging slot (no stream
// synchronisation; the host copy may change right after).
void upload_instances(pcc_ctx* ctx) {
  const int64_t B = (int64_t)ctx->inst.size();
  ctx->inst_host.resize(B);
  for (int64_t q = 0; q < B; ++q) {
    const Instance& I = ctx->inst[q];
    const pcc_params& p = I.p;
    InstDev& d = ctx->inst_host[q];
    d.boff = I.boff;
    d.N = I.N;
    d.coff = I.coff;
    d.C = I.C;
    d.G = I.G;
    d.K = I.K;
    d.lb = I.lb;
    d.bb = I.bb;
    d.bstart = I.bstart;
    d.maxit = p.max_bbpgd_iterations;
    d.dt = p.dt;
    d.xi = p.xi;
    d.itau = 1 / p.tau;
    d.mlam = -p.lambda;
    d.kcc = p.kcc;
    d.tol = p.bbpgd_tol;
    d.cutoff = I.cutoff;
    d.warm = p.warm_start_factor;
    d.origin = p.grid_origin;
    d.h = p.linked_cell_size;
    d.lam = p.lambda;
    d.tau = p.tau;
    d.temp = p.temperature;
    d.bseed = p.brownian_seed;
    d.bstep = I.nstep;
    d.mono = p.monolayer ? 1 : 0;
    d.pad1 = 0;
    d.phase = I.k;
    d.flags = I.flags;
    d.check = ctx->trace_on ? 1 : 10;
  }
  if (B == 0) return;
  ctx->idev.ensure(B);
  constexpr int kSlots = 4;
  if (ctx->inst_pin_cap < B) {
    PCC_HIP(hipStreamSynchronize(ctx->stream));  // earlier copies from the old slots have run
    DeviceLock lock(device_mutex());
    if (ctx->inst_pin) PCC_HIP(hipHostFree(ctx->inst_pin));
    ctx->inst_pin = nullptr;
    ctx->inst_pin_cap = 0;
    const int64_t cap = std::max<int64_t>(B, 16);
    PCC_HIP(hipHostMalloc((void**)&ctx->inst_pin, sizeof(InstDev) * cap * kSlots,
                          hipHostMallocDefault));
    ctx->inst_pin_cap = cap;
    for (auto& e : ctx->inst_ev)
      if (!e) PCC_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  const int slot = ctx->inst_slot;
  ctx->inst_slot = (slot + 1) % kSlots;
  PCC_HIP(hipEventSynchronize(ctx->inst_ev[slot]));  // the slot's previous copy has run
  InstDev* stage = ctx->inst_pin + (size_t)slot * ctx->inst_pin_cap;
  memcpy(stage, ctx->inst_host.data(), sizeof(InstDev) * B);
  PCC_HIP(hipMemcpyAsync(ctx->idev.p, stage, sizeof(InstDev) * B, hipMemcpyHostToDevice,
                         ctx->stream));
  PCC_HIP(hipEventRecord(ctx->inst_ev[slot], ctx->stream));
}

int32_t* readback_pin(pcc_ctx* ctx, int64_t n) {
  if (ctx->rb_pin_cap < n) {
    DeviceLock lock(device_mutex());
    if (ctx->rb_pin) PCC_HIP(hipHostFree(ctx->rb_pin));
    ctx->rb_pin = nullptr;
    ctx->rb_pin_cap = 0;
    const int64_t cap = std::max<int64_t>(n, 64);
    PCC_HIP(hipHostMalloc((void**)&ctx->rb_pin, sizeof(int32_t) * cap, hipHostMallocDefault));
    ctx->rb_pin_cap = cap;
  }
  return ctx->rb_pin;
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
  StreamScope scope(ctx ? ctx->stream : nullptr);
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
  if (!(p->temperature >= 0.0) || !std::isfinite(p->temperature)) {
    *why = "invalid temperature";
    return PCC_E_INVALID_ARGUMENT;
  }
  if (!(p->linked_cell_size > 0) || !(p->dt > 0) || !(p->xi > 0) || !(p->tau != 0) ||
      p->max_bbpgd_iterations < 1 || p->max_recursive_iterations < 1) {
    *why = "invalid parameter value";
    return PCC_E_INVALID_ARGUMENT;
  }
  return PCC_OK;
}

// Copies colony q's constraints (its range of the colony-major set) to host
// arrays, with body indices local to the colony.
void copy_set_to_host(pcc_ctx* ctx, int64_t q, pcc_constraints* out) {
  const ConstraintSet& S = ctx->S;
  const Instance& I = ctx->inst[q];
  const int64_t m = std::min<int64_t>(I.C, out->capacity);
  if (m <= 0) return;
  const int64_t o = I.coff;
  hipStream_t s = ctx->stream;
  auto cp = [&](void* dst, const void* src, size_t bytes) {
    if (dst) PCC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
  };
  cp(out->gidI, S.I.p + o, sizeof(int32_t) * m);
  cp(out->gidJ, S.J.p + o, sizeof(int32_t) * m);
  cp(out->signed_distance, S.phi.p + o, sizeof(double) * m);
  cp(out->normal, S.n.p + 3 * o, sizeof(double) * 3 * m);
  cp(out->rPosI, S.rI.p + 3 * o, sizeof(double) * 3 * m);
  cp(out->rPosJ, S.rJ.p + 3 * o, sizeof(double) * 3 * m);
  cp(out->contact_point, S.p.p + 3 * o, sizeof(double) * 3 * m);
  cp(out->stressI, S.sI.p + o, sizeof(double) * m);
  cp(out->stressJ, S.sJ.p + o, sizeof(double) * m);
  cp(out->gamma, S.gamma.p + o, sizeof(double) * m);
  cp(out->gid, S.gid.p + o, sizeof(int32_t) * m);
  cp(out->iteration, S.iter.p + o, sizeof(int32_t) * m);
  PCC_HIP(hipStreamSynchronize(s));
  const int32_t bo = (int32_t)I.boff;
  if (bo != 0) {
    for (int64_t c = 0; c < m; ++c) {
      if (out->gidI) out->gidI[c] -= bo;
      if (out->gidJ) out->gidJ[c] -= bo;
    }
  }
}

void fill_result(const Instance& I, pcc_step_result* r) {
  r->n_constraints = I.C;
  r->constraint_iterations = I.k;
  r->reserved = 0;
  r->bbpgd_iterations = I.its;
  r->residual = I.res;
  r->max_overlap = I.maxov;
}

void begin_step(pcc_ctx* ctx) {
  reset_step(ctx);
  ctx->brownian = false;
  const int64_t B = (int64_t)ctx->inst.size();
  ctx->trace_B = 0;
  if (ctx->trace_on && B > 0) {
    ctx->trace.ensure(B * std::max<int64_t>(ctx->trace_cap, 1));
    ctx->trace_n.ensure(B);
    PCC_HIP(hipMemsetAsync(ctx->trace_n.p, 0, sizeof(int64_t) * B, ctx->stream));
    ctx->trace_B = B;
  }
  for (auto& I : ctx->inst) {
    I.k = 0;
    I.its = 0;
    I.res = 0.0;
    I.maxov = 0.0;
    I.status = PCC_OK;
    I.error.clear();
    I.phase_iters.clear();
    I.phase_C.clear();
    I.active = true;
    I.flags = kInstActive;
  }
}

// Colonies whose ReLCP loops run together.  A colony of at least
// ctx->solo_cells cells fills the device alone and runs its loop by itself;
// the others share one loop in lockstep, which pays for small colonies (one
// launch serves them all) but, in every phase, runs as many BBPGD
// iterations as the group's slowest colony needs there.  Colonies never
// interact, so the grouping changes no result.
std::vector<std::vector<int64_t>> relcp_groups(const pcc_ctx* ctx) {
  std::vector<std::vector<int64_t>> g(1);
  for (int64_t q = 0; q < (int64_t)ctx->inst.size(); ++q) {
    if (ctx->solo_cells > 0 && ctx->inst[q].N >= ctx->solo_cells)
      g.push_back({q});
    else
      g[0].push_back(q);
  }
  if (g[0].empty()) g.erase(g.begin());
  return g;
}

// The ReLCP loop of solveHardModel (Hard.cpp:113-183) for the colonies
// marked active.
void relcp_loop(pcc_ctx* ctx) {
  std::vector<double> mo;
  for (int32_t k = 0;; ++k) {
    bool any = false;
    for (auto& I : ctx->inst) {
      if (I.active && k >= I.p.max_recursive_iterations) I.active = false;  // Hard.cpp:113
      I.flags = I.active ? kInstActive : 0;
      I.cutoff = k == 0 ? I.p.cutoff_phase0 : I.p.cutoff_phaseK;  // Hard.cpp:121
      any = any || I.active;
    }
    if (!any) break;
    upload_instances(ctx);
    detect_batch(ctx);
    append_batch(ctx, k);  // Hard.cpp:123
    max_overlap_inst(ctx, mo);  // Hard.cpp:125-130
    any = false;
    for (size_t q = 0; q < ctx->inst.size(); ++q) {
      Instance& I = ctx->inst[q];
      if (!I.active) continue;
      I.maxov = mo[q];
      if (I.maxov <= I.p.relcp_tol && k > 0) I.active = false;  // Hard.cpp:132-134
      I.flags = I.active ? kInstActive : 0;
      any = any || I.active;
    }
    if (!any) break;
    build_incidence(ctx, true);  // createMatrices, Hard.cpp:136
    if (k == 0) {
      upload_instances(ctx);
      brownian_inputs(ctx);  // U_ext of phase 0 (Hard.cpp:140-144)
    }
    bbpgd_solve(ctx);            // Hard.cpp:153
    for (auto& I : ctx->inst) {
      if (!I.active) continue;
      I.its += I.solve_its;
      I.res = I.solve_res;
      I.phase_iters.push_back(I.solve_its);
      I.phase_C.push_back(I.C);
      if (I.solve_failed) {
        I.status = PCC_E_BBPGD_NOT_CONVERGED;
        I.error = "BBPGD did not converge";
        I.active = false;
        I.flags = 0;
      } else {
        I.flags = kInstActive | kInstInteg;
      }
    }
    upload_instances(ctx);
    integrate_hard(ctx);  // Hard.cpp:160-170
    ctx->brownian = false;  // U_ext = 0 from phase 1 on
    for (auto& I : ctx->inst)
      if (I.active) I.k = k + 1;
  }
}

// solveHardModel (Hard.cpp:91-193) for every colony of the context.
void hard_step(pcc_ctx* ctx) {
  begin_step(ctx);
  // planar kernels while every body lies in z = 0 with an in-plane
  // orientation and no Brownian input turns it out of the plane
  ctx->planar = ctx->allow_planar && !brownian_offplane(ctx) && state_planar(ctx);
  ctx->detect_flat = ctx->planar;  // z = 0 stays exact through a planar step
  ctx->flat_known = ctx->planar;   // and after it
  for (const auto& grp : relcp_groups(ctx)) {
    for (auto& I : ctx->inst) I.active = false;
    for (int64_t q : grp) ctx->inst[q].active = true;
    relcp_loop(ctx);
  }
  for (auto& I : ctx->inst) I.flags = I.status == PCC_OK ? kInstGrow : 0;
  upload_instances(ctx);
  grow(ctx);  // Hard.cpp:185
  merge_set(ctx);
  for (auto& I : ctx->inst) {
    if (I.status == PCC_OK && I.k == I.p.max_recursive_iterations) {  // Hard.cpp:187-190
      I.status = PCC_E_RELCP_MAX_ITERS;
      I.error = "Maximum number of constraint iterations reached";
    }
  }
  PCC_HIP(hipStreamSynchronize(ctx->stream));
}

// solveSoftPotential (Soft.cpp:21-114) for every colony of the context.
void soft_step(pcc_ctx* ctx) {
  begin_step(ctx);
  const bool flat = state_planar(ctx);
  ctx->detect_flat = ctx->allow_planar && flat;
  // a planar state stays planar through the step unless a Brownian turn
  // leaves the plane (Hertz forces between planar bodies have no z part)
  ctx->flat_known = flat && !brownian_offplane(ctx);
  for (auto& I : ctx->inst) I.cutoff = 0.0;  // Soft.cpp:35
  upload_instances(ctx);
  detect_batch(ctx);
  append_batch(ctx, 0);
  max_overlap_launch(ctx);  // Soft.cpp:37-41, read at the end of the step
  build_incidence(ctx, false);
  brownian_inputs(ctx);  // U_ext of the step (Soft.cpp:94)
  soft_forces_and_integrate(ctx);
  ctx->brownian = false;
  merge_set(ctx);
  PCC_HIP(hipStreamSynchronize(ctx->stream));
  std::vector<double> mo;
  max_overlap_read(ctx, mo);
  for (size_t q = 0; q < ctx->inst.size(); ++q) {
    ctx->inst[q].maxov = mo[q];
    ctx->inst[q].res = mo[q];
  }
}

pcc_status run_step(pcc_ctx* ctx) {
  if (ctx->params.mode == PCC_MODE_HARD)
    hard_step(ctx);
  else
    soft_step(ctx);
  return PCC_OK;
}

pcc_status check_colony_params(pcc_ctx* ctx, const pcc_params* p, std::string* why) {
  pcc_status s = check_params(p, why);
  if (s != PCC_OK) return s;
  if (p->mode != ctx->params.mode || p->precision != ctx->params.precision) {
    *why = "a colony's mode and precision must be the context's";
    return PCC_E_INVALID_ARGUMENT;
  }
  return PCC_OK;
}

pcc_status check_bodies(pcc_ctx* ctx, const pcc_bodies* b) {
  if (!b || b->n < 0) return fail(ctx, PCC_E_INVALID_ARGUMENT, "invalid bodies");
  if (b->n > 0 && (!b->pos || !b->quat || !b->length || !b->l0 || !b->diameter || !b->ldot))
    return fail(ctx, PCC_E_INVALID_ARGUMENT, "missing body field");
  return PCC_OK;
}

// Device arrays sized for N bodies in total (body state, outputs, colony map).
void ensure_body_arrays(pcc_ctx* ctx, int64_t N) {
  ctx->body.ensure(kBodyFields * N + 1);
  ctx->F.ensure(6 * N + 6);
  ctx->U.ensure(6 * N + 6);
  ctx->sig.ensure(N + 1);
  ctx->ldot_new.ensure(N + 1);
  ctx->nc.ensure(N + 1);
  ctx->binst.ensure(N + 1);
}

// Colony of every body, from the colony ranges.
void upload_binst(pcc_ctx* ctx) {
  std::vector<int32_t> h(ctx->N);
  for (size_t q = 0; q < ctx->inst.size(); ++q)
    for (int64_t i = 0; i < ctx->inst[q].N; ++i) h[ctx->inst[q].boff + i] = (int32_t)q;
  if (ctx->N > 0)
    PCC_HIP(hipMemcpyAsync(ctx->binst.p, h.data(), sizeof(int32_t) * ctx->N, hipMemcpyHostToDevice,
                           ctx->stream));
  PCC_HIP(hipStreamSynchronize(ctx->stream));
}

// Packs host bodies (interleaved vectors) into the field-major layout of a
// context holding Ntot bodies, at body offset boff.
void pack_bodies(const pcc_bodies* b, int64_t Ntot, int64_t boff, double* h) {
  for (int64_t i = 0; i < b->n; ++i) {
    const int64_t j = boff + i;
    for (int k = 0; k < 3; ++k) h[(size_t)k * Ntot + j] = b->pos[3 * i + k];
    for (int k = 0; k < 4; ++k) h[(size_t)(3 + k) * Ntot + j] = b->quat[4 * i + k];
    h[(size_t)7 * Ntot + j] = b->length[i];
    h[(size_t)8 * Ntot + j] = b->l0[i];
    h[(size_t)9 * Ntot + j] = b->diameter[i];
    h[(size_t)10 * Ntot + j] = b->ldot[i];
    h[(size_t)11 * Ntot + j] = b->age ? (double)b->age[i] : 1.0;
    h[(size_t)12 * Ntot + j] = b->impedance ? b->impedance[i] : 0.0;  // Particle.cpp:25
  }
}

// Replaces the context's colonies (state uploaded, outputs zeroed).
pcc_status set_colonies(pcc_ctx* ctx, int64_t B, const pcc_bodies* bodies,
                        const pcc_params* params) {
  if (B < 1) return fail(ctx, PCC_E_INVALID_ARGUMENT, "no colony");
  int64_t Ntot = 0;
  for (int64_t q = 0; q < B; ++q) {
    pcc_status st = check_bodies(ctx, &bodies[q]);
    if (st != PCC_OK) return st;
    std::string why;
    if (params) {
      st = check_colony_params(ctx, &params[q], &why);
      if (st != PCC_OK) return fail(ctx, st, why);
    }
    Ntot += bodies[q].n;
  }
  if (Ntot >= (int64_t)1 << 30) return fail(ctx, PCC_E_INVALID_ARGUMENT, "too many bodies");
  return guarded(ctx, [&]() {
    PCC_HIP(hipSetDevice(ctx->device));
    std::vector<double> h((size_t)kBodyFields * Ntot);
    ctx->inst.assign(B, Instance());
    int64_t off = 0;
    for (int64_t q = 0; q < B; ++q) {
      Instance& I = ctx->inst[q];
      I.p = params ? params[q] : ctx->params;
      I.boff = off;
      I.N = bodies[q].n;
      pack_bodies(&bodies[q], Ntot, off, h.data());
      off += I.N;
    }
    ctx->N = Ntot;
    ensure_body_arrays(ctx, Ntot);
    hipStream_t s = ctx->stream;
    if (Ntot > 0)
      PCC_HIP(hipMemcpyAsync(ctx->body.p, h.data(), sizeof(double) * h.size(),
                             hipMemcpyHostToDevice, s));
    PCC_HIP(hipMemsetAsync(ctx->F.p, 0, sizeof(double) * (6 * Ntot + 6), s));
    PCC_HIP(hipMemsetAsync(ctx->U.p, 0, sizeof(double) * (6 * Ntot + 6), s));
    PCC_HIP(hipMemsetAsync(ctx->sig.p, 0, sizeof(double) * (Ntot + 1), s));
    PCC_HIP(hipMemsetAsync(ctx->nc.p, 0, sizeof(int32_t) * (Ntot + 1), s));
    upload_binst(ctx);
    upload_instances(ctx);
    ctx->S.C = 0;
    ctx->flat_known = false;
    ctx->has_state = true;
    ctx->has_snapshot = false;
    return PCC_OK;
  });
}

// Downloads colony q's state into host bodies (gid order within the colony).
pcc_status get_colony_state(pcc_ctx* ctx, int64_t q, pcc_bodies* b) {
  const Instance& I = ctx->inst[q];
  if (b->n != I.N) return fail(ctx, PCC_E_INVALID_ARGUMENT, "body count mismatch");
  return guarded(ctx, [&]() {
    PCC_HIP(hipSetDevice(ctx->device));
    const int64_t N = I.N, Nt = ctx->N, o = I.boff;
    std::vector<double> h((size_t)kBodyFields * N), F(6 * N), U(6 * N), sg(N);
    std::vector<int32_t> nc(N);
    hipStream_t s = ctx->stream;
    if (N > 0) {
      for (int k = 0; k < kBodyFields; ++k)
        PCC_HIP(hipMemcpyAsync(h.data() + (size_t)k * N, ctx->body.p + (size_t)k * Nt + o,
                               sizeof(double) * N, hipMemcpyDeviceToHost, s));
      PCC_HIP(hipMemcpyAsync(F.data(), ctx->F.p + 6 * o, sizeof(double) * 6 * N, hipMemcpyDeviceToHost, s));
      PCC_HIP(hipMemcpyAsync(U.data(), ctx->U.p + 6 * o, sizeof(double) * 6 * N, hipMemcpyDeviceToHost, s));
      PCC_HIP(hipMemcpyAsync(sg.data(), ctx->sig.p + o, sizeof(double) * N, hipMemcpyDeviceToHost, s));
      PCC_HIP(hipMemcpyAsync(nc.data(), ctx->nc.p + o, sizeof(int32_t) * N, hipMemcpyDeviceToHost, s));
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
      if (b->age) b->age[i] = (int32_t)h[(size_t)11 * N + i];
      for (int k = 0; k < 3; ++k) {
        if (b->force) b->force[3 * i + k] = F[6 * i + k];
        if (b->torque) b->torque[3 * i + k] = F[6 * i + 3 + k];
        if (b->velocity) b->velocity[3 * i + k] = U[6 * i + k];
        if (b->omega) b->omega[3 * i + k] = U[6 * i + 3 + k];
      }
      if (b->impedance) b->impedance[i] = h[(size_t)12 * N + i];
      if (b->stress) b->stress[i] = sg[i];
      if (b->num_constraints) b->num_constraints[i] = nc[i];
    }
    return PCC_OK;
  });
}

// Moves body ranges (src, dst, count) of the current arrays into fresh arrays
// for newN bodies; the rest of the fresh arrays is zero.
struct Move {
  int64_t src, dst, count;
};

template <class T>
void move_rows(pcc_ctx* ctx, DArr<T>& a, int64_t oldN, int64_t newN, int width, bool field_major,
               const std::vector<Move>& mv) {
  DArr<T> fresh;
  const int64_t rows = field_major ? width : 1;
  const int64_t per = field_major ? 1 : width;
  fresh.ensure(rows * per * newN + per + 1);
  hipStream_t s = ctx->stream;
  PCC_HIP(hipMemsetAsync(fresh.p, 0, sizeof(T) * (size_t)(rows * per * newN + per + 1), s));
  for (int64_t r = 0; r < rows; ++r)
    for (const Move& m : mv)
      if (m.count > 0)
        PCC_HIP(hipMemcpyAsync(fresh.p + r * newN * per + m.dst * per,
                               a.p + r * oldN * per + m.src * per, sizeof(T) * m.count * per,
                               hipMemcpyDeviceToDevice, s));
  PCC_HIP(hipStreamSynchronize(s));
  a.swap(fresh);
}

void repack(pcc_ctx* ctx, int64_t newN, const std::vector<Move>& mv) {
  const int64_t oldN = ctx->N;
  move_rows(ctx, ctx->body, oldN, newN, kBodyFields, true, mv);
  move_rows(ctx, ctx->F, oldN, newN, 6, false, mv);
  move_rows(ctx, ctx->U, oldN, newN, 6, false, mv);
  move_rows(ctx, ctx->sig, oldN, newN, 1, false, mv);
  move_rows(ctx, ctx->ldot_new, oldN, newN, 1, false, mv);
  move_rows(ctx, ctx->nc, oldN, newN, 1, false, mv);
  ctx->binst.release();
  ctx->binst.ensure(newN + 1);
  ctx->N = newN;
  ctx->S.C = 0;
  ctx->has_snapshot = false;
  int64_t off = 0;
  for (auto& I : ctx->inst) {
    I.boff = off;
    off += I.N;
    I.C = 0;
    I.coff = 0;
  }
  upload_binst(ctx);
  upload_instances(ctx);
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
  p->temperature = 1e-28;  // CLIParser.h:181
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
  p->monolayer = 1;  // CLIParser.h:183
  p->brownian_seed = 0;
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
  const char* graph_env = getenv("PCC_GRAPH");
  ctx->use_graphs = !(graph_env && graph_env[0] == '0');
  const char* tune_env = getenv("PCC_AUTOTUNE");
  ctx->autotune = !(tune_env && tune_env[0] == '0');
  const char* persist_env = getenv("PCC_PERSIST");
  ctx->use_persist = !(persist_env && persist_env[0] == '0');
  if (const char* v = getenv("PCC_PERSIST_MAX_C")) ctx->persist_max_c = atoll(v);
  if (const char* v = getenv("PCC_PERSIST_MAX_N")) ctx->persist_max_n = atoll(v);
  if (const char* v = getenv("PCC_SOLO_CELLS")) ctx->solo_cells = atoll(v);
  const char* compact_env = getenv("PCC_COMPACT");
  ctx->compact = !(compact_env && compact_env[0] == '0');
  if (const char* v = getenv("PCC_COMPACT_MIN_BLOCKS")) ctx->compact_min_blocks = atoll(v);
  if (const char* v = getenv("PCC_CHUNK_ITERS")) {
    const int c = atoi(v);
    if (c >= 1 && c <= 1000) ctx->chunk_iters = c;
  }
  if (const char* v = getenv("PCC_COMPACT_RATIO")) {
    const double r = atof(v);
    if (r > 0.0 && r < 1.0) ctx->compact_ratio = r;
  }
  // PCC_BODY_VARIANT=v (0..PCC_BODY_VARIANTS-1) fixes the body-kernel variant
  if (const char* v = getenv("PCC_BODY_VARIANT")) {
    const int x = atoi(v);
    if (v[0] >= '0' && v[0] <= '9' && x >= 0 && x < kBodyVariants) {
      ctx->body_variant = x;
      ctx->autotune = false;
    }
  }
  pcc_status s = guarded(ctx, [&]() {
    PCC_HIP(hipSetDevice(device));
    keep_pool_memory(device);
    {
      DeviceLock lock(device_mutex());
      PCC_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
      PCC_HIP(hipHostMalloc((void**)&ctx->state_host, 4 * sizeof(SolverState),
                            hipHostMallocDefault));
    }
    ctx->state_host_cap = 4;
    StreamScope scope(ctx->stream);
    ctx->state.ensure(1);
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
  const hipStream_t s = ctx->stream;
  if (s) (void)hipStreamSynchronize(s);
  {
    DeviceLock lock(device_mutex());
    for (auto& p : ctx->timing.pending) {
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    for (auto e : ctx->timing.free_events) (void)hipEventDestroy(e);
    if (ctx->state_host) (void)hipHostFree(ctx->state_host);
    if (ctx->inst_pin) (void)hipHostFree(ctx->inst_pin);
    if (ctx->ov_pin) (void)hipHostFree(ctx->ov_pin);
    if (ctx->rb_pin) (void)hipHostFree(ctx->rb_pin);
    for (auto e : ctx->inst_ev)
      if (e) (void)hipEventDestroy(e);
  }
  {
    StreamScope scope(s);  // pooled buffers go back to the pool in the stream's order
    delete ctx;            // DArr members free their device memory
  }
  if (s) {
    DeviceLock lock(device_mutex());
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
  }
}

pcc_status pcc_set_params(pcc_ctx* ctx, const pcc_params* params) {
  if (!ctx || !params) return PCC_E_INVALID_ARGUMENT;
  std::string why;
  pcc_status s = check_params(params, &why);
  if (s != PCC_OK) return fail(ctx, s, why);
  ctx->params = *params;
  for (auto& I : ctx->inst) I.p = *params;
  return PCC_OK;
}

const char* pcc_last_error(const pcc_ctx* ctx) {
  return ctx ? ctx->last_error.c_str() : pcc::thread_error().c_str();
}

pcc_status pcc_set_state(pcc_ctx* ctx, const pcc_bodies* b) {
  if (!ctx || !b) return PCC_E_INVALID_ARGUMENT;
  return set_colonies(ctx, 1, b, nullptr);
}

pcc_status pcc_get_state(pcc_ctx* ctx, pcc_bodies* b) {
  if (!ctx || !b) return PCC_E_INVALID_ARGUMENT;
  if (!ctx->has_state) return fail(ctx, PCC_E_NO_STATE, "no state");
  if (ctx->inst.size() != 1) return fail(ctx, PCC_E_INVALID_ARGUMENT, "batched context: use pcc_batch_get_state");
  return get_colony_state(ctx, 0, b);
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
    ctx->snap_nstep.clear();
    for (const auto& I : ctx->inst) ctx->snap_nstep.push_back(I.nstep);
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
    ctx->flat_known = false;
    for (size_t q = 0; q < ctx->inst.size() && q < ctx->snap_nstep.size(); ++q)
      ctx->inst[q].nstep = ctx->snap_nstep[q];
    return PCC_OK;
  });
}

pcc_status pcc_step(pcc_ctx* ctx, int32_t iter, pcc_step_result* out) {
  if (!ctx || !out) return PCC_E_INVALID_ARGUMENT;
  memset(out, 0, sizeof(*out));
  if (!ctx->has_state) return fail(ctx, PCC_E_NO_STATE, "no state");
  std::string why;
  pcc_status ps = check_params(&ctx->params, &why);
  if (ps != PCC_OK) return fail(ctx, ps, why);
  if (ctx->inst.size() != 1) return fail(ctx, PCC_E_INVALID_ARGUMENT, "batched context: use pcc_batch_step");
  (void)iter;  // the reference uses it only to gate BBPGD tracing (Hard.cpp:109)
  pcc_status s = guarded(ctx, [&]() {
    PCC_HIP(hipSetDevice(ctx->device));
    return run_step(ctx);
  });
  pcc_status t = guarded(ctx, [&]() {
    resolve_timing(ctx);
    return PCC_OK;
  });
  if (s != PCC_OK) return s;
  ++ctx->inst[0].nstep;
  const Instance& I = ctx->inst[0];
  fill_result(I, out);
  if (I.status != PCC_OK) return fail(ctx, I.status, I.error);
  return t;
}

pcc_status pcc_get_constraints(pcc_ctx* ctx, pcc_constraints* out, int64_t* count) {
  if (!ctx || !out) return PCC_E_INVALID_ARGUMENT;
  if (ctx->inst.size() != 1) return fail(ctx, PCC_E_INVALID_ARGUMENT, "batched context: use pcc_batch_get_constraints");
  if (count) *count = ctx->inst[0].C;
  return guarded(ctx, [&]() {
    PCC_HIP(hipSetDevice(ctx->device));
    copy_set_to_host(ctx, 0, out);
    return PCC_OK;
  });
}

pcc_status pcc_detect(pcc_ctx* ctx, double cutoff, pcc_constraints* out, int64_t* count,
                      uint32_t* bucket_out, int32_t* perm_out) {
  if (!ctx) return PCC_E_INVALID_ARGUMENT;
  if (!ctx->has_state) return fail(ctx, PCC_E_NO_STATE, "no state");
  if (ctx->inst.size() != 1) return fail(ctx, PCC_E_INVALID_ARGUMENT, "batched context");
  return guarded(ctx, [&]() {
    PCC_HIP(hipSetDevice(ctx->device));
    Instance& I = ctx->inst[0];
    I.flags = kInstActive;
    I.cutoff = cutoff;
    ctx->detect_flat = ctx->allow_planar && bodies_planar(ctx);
    upload_instances(ctx);
    detect_batch(ctx);
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

pcc_status pcc_batch_get_phase_stats(pcc_ctx* ctx, int64_t colony, int64_t* bbpgd_iterations,
                                     int64_t* n_constraints, int64_t max, int64_t* count) {
  if (!ctx) return PCC_E_INVALID_ARGUMENT;
  if (colony < 0 || colony >= (int64_t)ctx->inst.size())
    return fail(ctx, PCC_E_INVALID_ARGUMENT, "no such colony");
  const Instance& I = ctx->inst[colony];
  const int64_t n = (int64_t)I.phase_iters.size();
  if (count) *count = n;
  for (int64_t i = 0; i < n && i < max; ++i) {
    if (bbpgd_iterations) bbpgd_iterations[i] = I.phase_iters[i];
    if (n_constraints) n_constraints[i] = I.phase_C[i];
  }
  return PCC_OK;
}

pcc_status pcc_get_phase_stats(pcc_ctx* ctx, int64_t* bbpgd_iterations, int64_t* n_constraints,
                               int64_t max, int64_t* count) {
  if (!ctx) return PCC_E_INVALID_ARGUMENT;
  if (ctx->inst.empty()) {
    if (count) *count = 0;
    return PCC_OK;
  }
  return pcc_batch_get_phase_stats(ctx, 0, bbpgd_iterations, n_constraints, max, count);
}

pcc_status pcc_batch_set_state(pcc_ctx* ctx, int64_t n_colonies, const pcc_bodies* bodies,
                               const pcc_params* params) {
  if (!ctx || !bodies || n_colonies < 1) return PCC_E_INVALID_ARGUMENT;
  return set_colonies(ctx, n_colonies, bodies, params);
}

pcc_status pcc_batch_size(pcc_ctx* ctx, int64_t* n_colonies, int64_t* n_bodies) {
  if (!ctx) return PCC_E_INVALID_ARGUMENT;
  if (n_colonies) *n_colonies = ctx->has_state ? (int64_t)ctx->inst.size() : 0;
  if (n_bodies) *n_bodies = ctx->has_state ? ctx->N : 0;
  return PCC_OK;
}

pcc_status pcc_batch_step(pcc_ctx* ctx, int32_t iter, pcc_step_result* results,
                          pcc_status* statuses) {
  if (!ctx || !results) return PCC_E_INVALID_ARGUMENT;
  if (!ctx->has_state) return fail(ctx, PCC_E_NO_STATE, "no state");
  std::string why;
  pcc_status ps = check_params(&ctx->params, &why);
  if (ps != PCC_OK) return fail(ctx, ps, why);
  (void)iter;
  pcc_status s = guarded(ctx, [&]() {
    PCC_HIP(hipSetDevice(ctx->device));
    return run_step(ctx);
  });
  pcc_status t = guarded(ctx, [&]() {
    resolve_timing(ctx);
    return PCC_OK;
  });
  if (s != PCC_OK) return s;
  for (auto& I : ctx->inst) ++I.nstep;
  for (size_t q = 0; q < ctx->inst.size(); ++q) {
    fill_result(ctx->inst[q], &results[q]);
    if (statuses) statuses[q] = ctx->inst[q].status;
  }
  return t;
}

pcc_status pcc_batch_get_state(pcc_ctx* ctx, int64_t colony, pcc_bodies* bodies) {
  if (!ctx || !bodies) return PCC_E_INVALID_ARGUMENT;
  if (!ctx->has_state) return fail(ctx, PCC_E_NO_STATE, "no state");
  if (colony < 0 || colony >= (int64_t)ctx->inst.size())
    return fail(ctx, PCC_E_INVALID_ARGUMENT, "no such colony");
  return get_colony_state(ctx, colony, bodies);
}

pcc_status pcc_batch_get_constraints(pcc_ctx* ctx, int64_t colony, pcc_constraints* out,
                                     int64_t* count) {
  if (!ctx || !out) return PCC_E_INVALID_ARGUMENT;
  if (colony < 0 || colony >= (int64_t)ctx->inst.size())
    return fail(ctx, PCC_E_INVALID_ARGUMENT, "no such colony");
  if (count) *count = ctx->inst[colony].C;
  return guarded(ctx, [&]() {
    PCC_HIP(hipSetDevice(ctx->device));
    copy_set_to_host(ctx, colony, out);
    return PCC_OK;
  });
}

static pcc_status divide_impl(pcc_ctx* ctx, const uint64_t* seeds, int32_t step,
                              double length_noise, double angle_noise) {
  if (!ctx->has_state) return fail(ctx, PCC_E_NO_STATE, "no state");
  if (!(length_noise >= 0.0 && length_noise < 1.0) || !(angle_noise >= 0.0 && angle_noise <= 0.2))
    return fail(ctx, PCC_E_INVALID_ARGUMENT, "division noise out of range");
  return guarded(ctx, [&]() {
    PCC_HIP(hipSetDevice(ctx->device));
    divide_colonies(ctx, seeds, step, length_noise, angle_noise);
    {
      const int64_t N = ctx->N, B = (int64_t)ctx->inst.size();
      ensure_body_arrays(ctx, N);
      hipStream_t s = ctx->stream;
      k_clear_outputs<<<(unsigned)((N + 1 + 255) / 256), 256, 0, s>>>(N, ctx->F.p, ctx->U.p,
                                                                      ctx->sig.p, ctx->ldot_new.p,
                                                                      ctx->nc.p);
      upload_instances(ctx);
      if (N > 0 && B > 0)
        k_binst<<<(unsigned)((N + 255) / 256), 256, 0, s>>>(N, B, ctx->idev.p, ctx->binst.p);
    }
    return PCC_OK;
  });
}

pcc_status pcc_divide(pcc_ctx* ctx, uint64_t seed, int32_t step, double length_noise,
                      double angle_noise, int64_t* n_bodies) {
  if (!ctx) return PCC_E_INVALID_ARGUMENT;
  if (ctx->has_state && ctx->inst.size() != 1)
    return fail(ctx, PCC_E_INVALID_ARGUMENT, "batched context: use pcc_batch_divide");
  pcc_status st = divide_impl(ctx, &seed, step, length_noise, angle_noise);
  if (st == PCC_OK && n_bodies) *n_bodies = ctx->N;
  return st;
}

pcc_status pcc_batch_divide(pcc_ctx* ctx, const uint64_t* seeds, int32_t step,
                            double length_noise, double angle_noise, int64_t* sizes) {
  if (!ctx || !seeds) return PCC_E_INVALID_ARGUMENT;
  pcc_status st = divide_impl(ctx, seeds, step, length_noise, angle_noise);
  if (st == PCC_OK && sizes)
    for (size_t q = 0; q < ctx->inst.size(); ++q) sizes[q] = ctx->inst[q].N;
  return st;
}

pcc_status pcc_adapt_dt(pcc_ctx* ctx, double cfl_factor, double alpha, double* dt,
                        double* median) {
  if (!ctx) return PCC_E_INVALID_ARGUMENT;
  if (!ctx->has_state) return fail(ctx, PCC_E_NO_STATE, "no state");
  if (ctx->inst.size() != 1)
    return fail(ctx, PCC_E_INVALID_ARGUMENT, "batched context: use pcc_batch_adapt_dt");
  if (!(alpha >= 0.0 && alpha <= 1.0) || !(cfl_factor > 0.0))
    return fail(ctx, PCC_E_INVALID_ARGUMENT, "invalid time-step adaptation parameter");
  return guarded(ctx, [&]() {
    PCC_HIP(hipSetDevice(ctx->device));
    std::vector<double> m;
    colony_median_speeds(ctx, m);
    // the colony's relcp_tol: the reference's one tolerance (Domain.cpp:106)
    const double nd =
        adapted_dt(ctx->params.dt, m[0], cfl_factor, ctx->inst[0].p.relcp_tol, alpha);
    ctx->params.dt = nd;
    ctx->inst[0].p.dt = nd;
    if (dt) *dt = nd;
    if (median) *median = m[0];
    return PCC_OK;
  });
}

pcc_status pcc_batch_adapt_dt(pcc_ctx* ctx, double cfl_factor, double alpha, double* dts,
                              double* medians) {
  if (!ctx) return PCC_E_INVALID_ARGUMENT;
  if (!ctx->has_state) return fail(ctx, PCC_E_NO_STATE, "no state");
  if (!(alpha >= 0.0 && alpha <= 1.0) || !(cfl_factor > 0.0))
    return fail(ctx, PCC_E_INVALID_ARGUMENT, "invalid time-step adaptation parameter");
  return guarded(ctx, [&]() {
    PCC_HIP(hipSetDevice(ctx->device));
    std::vector<double> m;
    colony_median_speeds(ctx, m);
    for (size_t q = 0; q < ctx->inst.size(); ++q) {
      pcc_params& p = ctx->inst[q].p;
      p.dt = adapted_dt(p.dt, m[q], cfl_factor, p.relcp_tol, alpha);
      if (dts) dts[q] = p.dt;
      if (medians) medians[q] = m[q];
    }
    return PCC_OK;
  });
}

pcc_status pcc_batch_get_step_count(pcc_ctx* ctx, int64_t colony, int64_t* count) {
  if (!ctx || !count) return PCC_E_INVALID_ARGUMENT;
  if (colony < 0 || colony >= (int64_t)ctx->inst.size())
    return fail(ctx, PCC_E_INVALID_ARGUMENT, "no such colony");
  *count = ctx->inst[colony].nstep;
  return PCC_OK;
}

pcc_status pcc_batch_set_step_count(pcc_ctx* ctx, int64_t colony, int64_t count) {
  if (!ctx || count < 0) return PCC_E_INVALID_ARGUMENT;
  if (colony < 0 || colony >= (int64_t)ctx->inst.size())
    return fail(ctx, PCC_E_INVALID_ARGUMENT, "no such colony");
  ctx->inst[colony].nstep = count;
  return PCC_OK;
}

pcc_status pcc_batch_set_params(pcc_ctx* ctx, int64_t colony, const pcc_params* params) {
  if (!ctx || !params) return PCC_E_INVALID_ARGUMENT;
  if (colony < 0 || colony >= (int64_t)ctx->inst.size())
    return fail(ctx, PCC_E_INVALID_ARGUMENT, "no such colony");
  std::string why;
  pcc_status s = check_colony_params(ctx, params, &why);
  if (s != PCC_OK) return fail(ctx, s, why);
  ctx->inst[colony].p = *params;
  return PCC_OK;
}

pcc_status pcc_batch_add(pcc_ctx* ctx, const pcc_bodies* bodies, const pcc_params* params,
                         int64_t* colony) {
  if (!ctx || !bodies) return PCC_E_INVALID_ARGUMENT;
  if (!ctx->has_state) {
    pcc_status s = set_colonies(ctx, 1, bodies, params);
    if (s == PCC_OK && colony) *colony = 0;
    return s;
  }
  pcc_status st = check_bodies(ctx, bodies);
  if (st != PCC_OK) return st;
  std::string why;
  if (params) {
    st = check_colony_params(ctx, params, &why);
    if (st != PCC_OK) return fail(ctx, st, why);
  }
  const int64_t oldN = ctx->N, n = bodies->n, newN = oldN + n;
  if (newN >= (int64_t)1 << 30) return fail(ctx, PCC_E_INVALID_ARGUMENT, "too many bodies");
  return guarded(ctx, [&]() {
    PCC_HIP(hipSetDevice(ctx->device));
    Instance I;
    I.p = params ? *params : ctx->params;
    I.N = n;
    ctx->inst.push_back(I);
    repack(ctx, newN, {{0, 0, oldN}});
    // the new colony's fields
    std::vector<double> h((size_t)kBodyFields * n);
    pack_bodies(bodies, n, 0, h.data());
    for (int k = 0; k < kBodyFields && n > 0; ++k)
      PCC_HIP(hipMemcpyAsync(ctx->body.p + (size_t)k * newN + oldN, h.data() + (size_t)k * n,
                             sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
    PCC_HIP(hipStreamSynchronize(ctx->stream));
    ctx->flat_known = false;
    if (colony) *colony = (int64_t)ctx->inst.size() - 1;
    return PCC_OK;
  });
}

pcc_status pcc_batch_remove(pcc_ctx* ctx, int64_t colony) {
  if (!ctx) return PCC_E_INVALID_ARGUMENT;
  if (colony < 0 || colony >= (int64_t)ctx->inst.size())
    return fail(ctx, PCC_E_INVALID_ARGUMENT, "no such colony");
  if (ctx->inst.size() == 1) {
    ctx->inst.clear();
    ctx->N = 0;
    ctx->S.C = 0;
    ctx->has_state = false;
    ctx->has_snapshot = false;
    return PCC_OK;
  }
  return guarded(ctx, [&]() {
    PCC_HIP(hipSetDevice(ctx->device));
    const Instance& I = ctx->inst[colony];
    const int64_t oldN = ctx->N, newN = oldN - I.N;
    const std::vector<Move> mv = {{0, 0, I.boff}, {I.boff + I.N, I.boff, oldN - I.boff - I.N}};
    ctx->inst.erase(ctx->inst.begin() + colony);
    repack(ctx, newN, mv);
    return PCC_OK;
  });
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

pcc_status pcc_get_tuning(pcc_ctx* ctx, int32_t* variant, int32_t* body_chunk,
                          int32_t* body_threads, int32_t* tuned, double* variant_us) {
  if (!ctx) return PCC_E_INVALID_ARGUMENT;
  if (variant) *variant = ctx->body_variant;
  if (body_chunk) *body_chunk = body_chunk_of(ctx->body_variant, ctx->planar);
  if (body_threads) *body_threads = body_threads_of(ctx->body_variant);
  if (tuned) *tuned = ctx->body_tuned ? (ctx->body_from_cache ? 2 : 1) : 0;
  if (variant_us)
    for (int v = 0; v < kBodyVariants; ++v) variant_us[v] = ctx->body_variant_ms[v] * 1e3;
  return PCC_OK;
}

pcc_status pcc_get_solve_stats(pcc_ctx* ctx, int64_t* compactions) {
  if (!ctx) return PCC_E_INVALID_ARGUMENT;
  if (compactions) *compactions = ctx->compactions;
  return PCC_OK;
}

pcc_status pcc_set_trace(pcc_ctx* ctx, int enable, int64_t capacity) {
  if (!ctx) return PCC_E_INVALID_ARGUMENT;
  if (enable && capacity < 0)
    return fail(ctx, PCC_E_INVALID_ARGUMENT, "trace capacity must be >= 0");
  ctx->trace_on = enable != 0;
  ctx->trace_cap = enable ? capacity : 0;
  return PCC_OK;
}

pcc_status pcc_get_trace(pcc_ctx* ctx, int64_t colony, pcc_bbpgd_record* out, int64_t capacity,
                         int64_t* count) {
  if (!ctx || !count || capacity < 0 || (capacity > 0 && !out))
    return PCC_E_INVALID_ARGUMENT;
  *count = 0;
  if (ctx->trace_B == 0) return PCC_OK;  // the last step was not traced
  if (colony < 0 || colony >= ctx->trace_B)
    return fail(ctx, PCC_E_INVALID_ARGUMENT, "colony out of range of the traced step");
  return guarded(ctx, [&]() {
    int64_t n = 0;
    PCC_HIP(hipMemcpyAsync(&n, ctx->trace_n.p + colony, sizeof(int64_t), hipMemcpyDeviceToHost,
                           ctx->stream));
    PCC_HIP(hipStreamSynchronize(ctx->stream));
    const int64_t m = std::min(std::min(n, capacity), ctx->trace_cap);
    if (m > 0) {
      PCC_HIP(hipMemcpyAsync(out, ctx->trace.p + colony * std::max<int64_t>(ctx->trace_cap, 1),
                             sizeof(pcc_bbpgd_record) * m, hipMemcpyDeviceToHost, ctx->stream));
      PCC_HIP(hipStreamSynchronize(ctx->stream));
    }
    *count = n;
    return PCC_OK;
  });
}

pcc_status pcc_device_bytes(pcc_ctx* ctx, int64_t* bytes) {
  if (!ctx || !bytes) return PCC_E_INVALID_ARGUMENT;
  int64_t t = 0;
  auto add = [&](const auto& a) { t += a.cap * (int64_t)sizeof(*a.p); };
  auto add_set = [&](const ConstraintSet& S) {
    add(S.I); add(S.J); add(S.gid); add(S.iter); add(S.inst); add(S.phi); add(S.gamma);
    add(S.n); add(S.rI); add(S.rJ); add(S.p); add(S.sI); add(S.sJ); add(S.coef64);
    add(S.coef32);
  };
  add(ctx->body); add(ctx->snap); add(ctx->F); add(ctx->U); add(ctx->sig); add(ctx->uext);
  add(ctx->dv_flag); add(ctx->dv_off); add(ctx->dv_bounds); add(ctx->dv_perm);
  add(ctx->dv_sinst); add(ctx->dv_stage); add(ctx->dv_body); add(ctx->dv_key); add(ctx->dv_col);
  add(ctx->ldot_new); add(ctx->nc); add(ctx->cell); add(ctx->bucket); add(ctx->perm);
  add(ctx->bstart); add(ctx->bcursor); add(ctx->dcount); add(ctx->dslot); add(ctx->dslot_off); add(ctx->dcstart); add(ctx->dcand); add(ctx->daccept); add(ctx->dbcount); add(ctx->dgeo); add(ctx->scell); add(ctx->sgeo); add(ctx->sfrec); add(ctx->rI_); add(ctx->rJ_);
  add(ctx->rrec); add(ctx->ridx); add(ctx->ridx2); add(ctx->rkey); add(ctx->rkey2);
  add(ctx->xs_cnt); add(ctx->xs_off); add(ctx->xs_bounds);
  add(ctx->sort_tmp); add_set(ctx->S); add_set(ctx->S2); add(ctx->rowptr); add(ctx->inccur);
  add(ctx->inc); add(ctx->increc); add(ctx->increc32); add(ctx->incidx); add(ctx->gold_e); add(ctx->fold); add(ctx->plist);
  add(ctx->sord); add(ctx->spos); add(ctx->skey); add(ctx->skey2); add(ctx->sval); add(ctx->sIJ);
  add(ctx->sphi); add(ctx->prow); add(ctx->prow32); add(ctx->flag);
  add(ctx->iter); add(ctx->gprev); add(ctx->ubuf); add(ctx->partials); add(ctx->merge_offs);
  add(ctx->state); add(ctx->scratch_d); add(ctx->hertz); add(ctx->hertz32); add(ctx->idev);
  add(ctx->binst); add(ctx->bmap); add(ctx->cmap); add(ctx->bbounds); add(ctx->maxbits);
  add(ctx->regroup_dst); add(ctx->trace); add(ctx->trace_n); add(ctx->epart);
  *bytes = t;
  return PCC_OK;
}

const char* pcc_kernel_name(int kernel) {
  static const char* names[PCC_K_COUNT] = {"body_apply", "constraint_apply", "detect",
                                           "assemble",   "integrate",        "hertz"};
  return (kernel >= 0 && kernel < PCC_K_COUNT) ? names[kernel] : "?";
}

}  // extern "C"
