// solve.hip — hard-model ReLCP phase on the GPU: operators, BBPGD, integration.
//
// Replaces createMatrices (Hard.cpp:10-47, Physics.cpp:118-265),
// HardModelGradient::gradient/residual (HardModelGradient.cpp:89-158),
// BBPGD (BBPGD.cpp:32-127) and the write-back of Hard.cpp:160-170,
// ParticleManager.cpp:29-137, Particle.cpp:37-65,110-120.
//
// One BBPGD iteration is two kernels:
//   k_body    one thread per body: projection step gamma = max(gp - a*g, 0)
//             evaluated on the fly for each incident constraint, then the
//             segmented D^T (gamma - gamma_old) and L^T gamma sums in ascending
//             constraint order (PETSc's MatMultTranspose row order), mobility,
//             impedance = exp(-lambda sigma), growth rate; writes U and
//             ldot - ldot_prev per body.  No float atomics.
//   k_constr  one thread per constraint: the same projection, D U and L dldot
//             gathers, g = Phi + dt*phidot - dt*phidot_growth, BB partial dot
//             products and the projected-gradient residual with a fixed
//             reduction tree; the last block (integer ticket) reduces the
//             partials in a fixed order and updates alpha / iteration /
//             convergence on the device.
// Both kernels exit immediately once the device flags say the solve is done,
// so the host can enqueue iterations speculatively.
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "common.hpp"
#include "internal.hpp"

namespace pcc {

void exclusive_scan_i32(pcc_ctx* ctx, const int32_t* in, int32_t* out, int64_t n);
void seg_sort_i32(pcc_ctx* ctx, int64_t nseg, const int32_t* start, int32_t* v);

namespace {

inline unsigned nblk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

__device__ __forceinline__ double proj_step(double yp, double malpha, double gp) {
  double v = yp + malpha * gp;  // VecWAXPY(gamma_i, -alpha, g_prev, gamma_prev)
  return v < 0.0 ? 0.0 : v;     // VecPointwiseMax(gamma_i, gamma_i, 0)
}

__device__ __forceinline__ float proj_step_f(float yp, double malpha, float gp) {
  double v = (double)yp + malpha * (double)gp;
  return v < 0.0 ? 0.0f : (float)v;
}

__global__ void k_inc_count(int64_t C, const int32_t* __restrict__ I, const int32_t* __restrict__ J,
                            int32_t* __restrict__ deg) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  atomicAdd(&deg[I[c]], 1);
  atomicAdd(&deg[J[c]], 1);
}

__global__ void k_inc_scatter(int64_t C, const int32_t* __restrict__ I,
                              const int32_t* __restrict__ J, const int32_t* __restrict__ rowptr,
                              int32_t* __restrict__ cur, int32_t* __restrict__ inc) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  int32_t a = I[c], b = J[c];
  inc[rowptr[a] + atomicAdd(&cur[a], 1)] = (int32_t)(2 * c);
  inc[rowptr[b] + atomicAdd(&cur[b], 1)] = (int32_t)(2 * c + 1);
}

// Incidence records, body-major: coefficients of the D row on this body's
// six DOFs ([n, rI x n] for side I, [-n, rJ x -n] for side J), the stress
// coefficient, and the constraint index.
__global__ void k_inc_records(int64_t E, const int32_t* __restrict__ inc,
                              const double* __restrict__ n, const double* __restrict__ tI,
                              const double* __restrict__ tJ, const double* __restrict__ sI,
                              const double* __restrict__ sJ, double* __restrict__ rec,
                              float* __restrict__ rec32) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const int32_t v = inc[e];
  const int64_t c = v >> 1;
  const int side = v & 1;
  double co[8];
  if (side == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      co[k] = n[3 * c + k];
      co[3 + k] = tI[3 * c + k];
    }
    co[6] = sI[c];
  } else {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      co[k] = -n[3 * c + k];
      co[3 + k] = tJ[3 * c + k];
    }
    co[6] = sJ[c];
  }
  co[7] = __longlong_as_double((long long)c);
  double* r = rec + 8 * e;
#pragma unroll
  for (int k = 0; k < 8; ++k) r[k] = co[k];
  float* r32 = rec32 + 8 * e;
#pragma unroll
  for (int k = 0; k < 7; ++k) r32[k] = (float)co[k];
  r32[7] = __int_as_float((int)c);
}

__global__ void k_phase_init(int64_t C, const double* __restrict__ Sgam,
                             const double* __restrict__ Sphi, double* __restrict__ gam,
                             double* __restrict__ gold, double* __restrict__ phi,
                             float* __restrict__ gam32, float* __restrict__ gold32,
                             float* __restrict__ phi32) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  double g = Sgam[c], p = Sphi[c];
  gam[c] = g;
  gold[c] = g;
  phi[c] = p;
  gam32[c] = (float)g;
  gold32[c] = (float)g;
  phi32[c] = (float)p;
}

__global__ void k_mobility(int64_t N, const double* __restrict__ len, double xi,
                           double* __restrict__ mob) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double l = len[i];
  mob[2 * i] = 1.0 / (l * xi);               // Physics.cpp:203
  mob[2 * i + 1] = 12.0 / (l * l * l * xi);  // Physics.cpp:204
}

// ---------------------------------------------------------------------------
// FP64 BBPGD kernels
// ---------------------------------------------------------------------------
struct BodyArgs {
  int64_t N, C;
  const int32_t* rowptr;
  const double* rec;
  const float* rec32;
  const double* gam;   // 2C
  const double* grd;   // 2C
  const double* gold;  // C
  const float* gam32;
  const float* grd32;
  const float* gold32;
  const double* mob;
  const double* len;
  const double* ldot_prev;
  double mlam, itau;
  const SolverState* st;
  double* ubuf;  // 8 per body
  float* ubuf32;
};

__global__ void __launch_bounds__(256) k_body64(BodyArgs a) {
  const SolverState* st = a.st;
  if (st->converged | st->failed) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.N) return;
  const int64_t k = st->iter;
  const double malpha = -st->alpha;
  const double* yp = a.gam + ((k - 1) & 1) * a.C;
  const double* gp = a.grd + ((k - 1) & 1) * a.C;
  double F[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double sig = 0.0;
  const int32_t e1 = a.rowptr[i + 1];
  for (int32_t e = a.rowptr[i]; e < e1; ++e) {
    const double* r = a.rec + 8 * (int64_t)e;
    const int64_t c = (int64_t)__double_as_longlong(r[7]);
    const double gc = (k == 0) ? a.gam[c] : proj_step(yp[c], malpha, gp[c]);
    const double dg = gc - a.gold[c];
#pragma unroll
    for (int q = 0; q < 6; ++q) F[q] += dg * r[q];
    sig += gc * r[6];
  }
  const double mt = a.mob[2 * i], mr = a.mob[2 * i + 1];
  double* u = a.ubuf + 8 * i;
#pragma unroll
  for (int q = 0; q < 3; ++q) u[q] = (0.0 + mt * F[q]) + 0.0;
#pragma unroll
  for (int q = 3; q < 6; ++q) u[q] = (0.0 + mr * F[q]) + 0.0;
  const double imp = det_exp(sig * a.mlam);
  const double ld = (a.len[i] * imp) * a.itau;
  u[6] = ld - a.ldot_prev[i];
}

struct ConstrArgs {
  int64_t C, nb;
  const int32_t* I;
  const int32_t* J;
  const double* n;
  const double* tI;
  const double* tJ;
  const double* sI;
  const double* sJ;
  const double* phi;
  const float* fn;
  const float* ftI;
  const float* ftJ;
  const float* fsI;
  const float* fsJ;
  const float* phi32;
  double* gam;
  double* grd;
  float* gam32;
  float* grd32;
  const double* ubuf;
  const float* ubuf32;
  double dt, tol;
  int64_t maxit;
  SolverState* st;
  double* partials;
  uint32_t* ticket;
};

// Final reduction of the per-block partials by the last block, and the BB
// step-size / convergence update (BBPGD.cpp:49-116).
__device__ void finish_iteration(const ConstrArgs& a, int64_t k, double s0, double s1, double s2,
                                 double r, double* lds) {
  __shared__ int is_last;
  const int64_t nb = a.nb;
  if (threadIdx.x == 0) {
    a.partials[blockIdx.x] = s0;
    a.partials[nb + blockIdx.x] = s1;
    a.partials[2 * nb + blockIdx.x] = s2;
    a.partials[3 * nb + blockIdx.x] = r;
    __threadfence();
    uint32_t t = atomicAdd(a.ticket, 1u);
    is_last = (t == (uint32_t)(nb - 1));
  }
  __syncthreads();
  if (!is_last) return;
  __threadfence();
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, am = 0.0;
  for (int64_t i = threadIdx.x; i < nb; i += PCC_BLOCK) {
    a0 += __hip_atomic_load(a.partials + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    a1 += __hip_atomic_load(a.partials + nb + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    a2 += __hip_atomic_load(a.partials + 2 * nb + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    double m = __hip_atomic_load(a.partials + 3 * nb + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    am = am < m ? m : am;
  }
  const double S0 = block_sum(a0, lds);
  const double S1 = block_sum(a1, lds);
  const double S2 = block_sum(a2, lds);
  const double R = block_max(am, lds);
  if (threadIdx.x != 0) return;
  SolverState* st = a.st;
  *a.ticket = 0u;
  if (k == 0) {  // initial gradient: res = residual(g, Gamma), alpha = 1/res
    st->res = R;
    st->alpha = 1.0 / R;
    st->iter = 1;
    if (1 == a.maxit) st->failed = 1;
    return;
  }
  if (k % 10 == 0) {
    st->res = R;
    if (R <= a.tol) {
      st->converged = 1;
      return;
    }
  }
  double num, den;
  if (k % 2 == 0) {  // BB1: (dg.dg)/(dg.dphi)
    num = S0;
    den = S1;
  } else {  // BB2: (dphi.dg)/(dphi.dphi)
    den = S2;
    num = S1;
  }
  const double eps10 = 10 * 2.220446049250313080847e-16;
  if (fabs(den) < eps10) den += eps10;
  st->alpha = num / den;
  st->iter = k + 1;
  if (k + 1 == a.maxit) st->failed = 1;
}

__global__ void __launch_bounds__(256) k_constr64(ConstrArgs a) {
  __shared__ double lds[4];
  const SolverState* st = a.st;
  if (st->converged | st->failed) return;
  const int64_t k = st->iter;
  const double malpha = -st->alpha;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v0 = 0.0, v1 = 0.0, v2 = 0.0, vr = 0.0;
  if (c < a.C) {
    const double* yp = a.gam + ((k - 1) & 1) * a.C;
    const double* gp = a.grd + ((k - 1) & 1) * a.C;
    const double gc = (k == 0) ? a.gam[c] : proj_step(yp[c], malpha, gp[c]);
    const double* UI = a.ubuf + 8 * (int64_t)a.I[c];
    const double* UJ = a.ubuf + 8 * (int64_t)a.J[c];
    const double n0 = a.n[3 * c], n1 = a.n[3 * c + 1], n2 = a.n[3 * c + 2];
    double pd = 0.0;  // D U_total, columns of body I (lower gid) first
    pd += n0 * UI[0];
    pd += n1 * UI[1];
    pd += n2 * UI[2];
    pd += a.tI[3 * c] * UI[3];
    pd += a.tI[3 * c + 1] * UI[4];
    pd += a.tI[3 * c + 2] * UI[5];
    pd += (-n0) * UJ[0];
    pd += (-n1) * UJ[1];
    pd += (-n2) * UJ[2];
    pd += a.tJ[3 * c] * UJ[3];
    pd += a.tJ[3 * c + 1] * UJ[4];
    pd += a.tJ[3 * c + 2] * UJ[5];
    double pg = 0.0;  // L ldot_diff
    pg += a.sI[c] * UI[6];
    pg += a.sJ[c] * UJ[6];
    const double g = a.phi[c] + (a.dt * pd + (-a.dt) * pg);
    if (k > 0) a.gam[(k & 1) * a.C + c] = gc;
    a.grd[(k & 1) * a.C + c] = g;
    if (k > 0) {
      const double dgam = gc - yp[c];
      const double dphi = g - gp[c];
      v0 = dgam * dgam;
      v1 = dgam * dphi;
      v2 = dphi * dphi;
    }
    const double pj = gc > 0 ? g : (g < 0.0 ? g : 0.0);
    vr = fabs(pj);
  }
  const double s0 = block_sum(v0, lds);
  const double s1 = block_sum(v1, lds);
  const double s2 = block_sum(v2, lds);
  const double r = block_max(vr, lds);
  finish_iteration(a, k, s0, s1, s2, r, lds);
}

// ---------------------------------------------------------------------------
// Mixed precision: FP32 storage and applies of D, D^T, L, L^T, gamma, g and
// the body velocities; FP64 for the projection step, the BB dot products and
// the residual.  The exp of the growth law runs in FP32 (HIP expf).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_body32(BodyArgs a) {
  const SolverState* st = a.st;
  if (st->converged | st->failed) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.N) return;
  const int64_t k = st->iter;
  const double malpha = -st->alpha;
  const float* yp = a.gam32 + ((k - 1) & 1) * a.C;
  const float* gp = a.grd32 + ((k - 1) & 1) * a.C;
  float F[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  float sig = 0.f;
  const int32_t e1 = a.rowptr[i + 1];
  for (int32_t e = a.rowptr[i]; e < e1; ++e) {
    const float4* r = reinterpret_cast<const float4*>(a.rec32 + 8 * (int64_t)e);
    const float4 r0 = r[0], r1 = r[1];
    const int64_t c = (int64_t)__float_as_int(r1.w);
    const float gc = (k == 0) ? a.gam32[c] : proj_step_f(yp[c], malpha, gp[c]);
    const float dg = gc - a.gold32[c];
    F[0] += dg * r0.x;
    F[1] += dg * r0.y;
    F[2] += dg * r0.z;
    F[3] += dg * r0.w;
    F[4] += dg * r1.x;
    F[5] += dg * r1.y;
    sig += gc * r1.z;
  }
  const float mt = (float)a.mob[2 * i], mr = (float)a.mob[2 * i + 1];
  const float imp = expf(sig * (float)a.mlam);
  const float ld = ((float)a.len[i] * imp) * (float)a.itau;
  float4* u = reinterpret_cast<float4*>(a.ubuf32 + 8 * i);
  u[0] = make_float4(mt * F[0], mt * F[1], mt * F[2], mr * F[3]);
  u[1] = make_float4(mr * F[4], mr * F[5], ld - (float)a.ldot_prev[i], 0.f);
}

__global__ void __launch_bounds__(256) k_constr32(ConstrArgs a) {
  __shared__ double lds[4];
  const SolverState* st = a.st;
  if (st->converged | st->failed) return;
  const int64_t k = st->iter;
  const double malpha = -st->alpha;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v0 = 0.0, v1 = 0.0, v2 = 0.0, vr = 0.0;
  if (c < a.C) {
    const float* yp = a.gam32 + ((k - 1) & 1) * a.C;
    const float* gp = a.grd32 + ((k - 1) & 1) * a.C;
    const float gc = (k == 0) ? a.gam32[c] : proj_step_f(yp[c], malpha, gp[c]);
    const float4* UI = reinterpret_cast<const float4*>(a.ubuf32 + 8 * (int64_t)a.I[c]);
    const float4* UJ = reinterpret_cast<const float4*>(a.ubuf32 + 8 * (int64_t)a.J[c]);
    const float4 ui0 = UI[0], ui1 = UI[1], uj0 = UJ[0], uj1 = UJ[1];
    const float n0 = a.fn[3 * c], n1 = a.fn[3 * c + 1], n2 = a.fn[3 * c + 2];
    float pd = n0 * (ui0.x - uj0.x) + n1 * (ui0.y - uj0.y) + n2 * (ui0.z - uj0.z);
    pd += a.ftI[3 * c] * ui0.w + a.ftI[3 * c + 1] * ui1.x + a.ftI[3 * c + 2] * ui1.y;
    pd += a.ftJ[3 * c] * uj0.w + a.ftJ[3 * c + 1] * uj1.x + a.ftJ[3 * c + 2] * uj1.y;
    const float pg = a.fsI[c] * ui1.z + a.fsJ[c] * uj1.z;
    const float g = a.phi32[c] + ((float)a.dt * pd - (float)a.dt * pg);
    if (k > 0) a.gam32[(k & 1) * a.C + c] = gc;
    a.grd32[(k & 1) * a.C + c] = g;
    if (k > 0) {
      const double dgam = (double)gc - (double)yp[c];
      const double dphi = (double)g - (double)gp[c];
      v0 = dgam * dgam;
      v1 = dgam * dphi;
      v2 = dphi * dphi;
    }
    const float pj = gc > 0 ? g : (g < 0.0f ? g : 0.0f);
    vr = fabs((double)pj);
  }
  const double s0 = block_sum(v0, lds);
  const double s1 = block_sum(v1, lds);
  const double s2 = block_sum(v2, lds);
  const double r = block_max(vr, lds);
  finish_iteration(a, k, s0, s1, s2, r, lds);
}

// ---------------------------------------------------------------------------
// Finalisation: the workspace of the last gradient evaluation, gamma_i =
// gam[kf & 1] (HardModelGradient.cpp:89-121), in FP64.  In mixed precision
// the FP32 iterate is widened and the body sums are redone in FP64.
// ---------------------------------------------------------------------------
__global__ void k_body_final(BodyArgs a, int64_t kf, int mixed, double* __restrict__ F_out,
                             double* __restrict__ U_out, double* __restrict__ sig_out,
                             double* __restrict__ imp_out, double* __restrict__ ld_out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.N) return;
  const double* gi = a.gam + (kf & 1) * a.C;
  const float* gi32 = a.gam32 + (kf & 1) * a.C;
  double F[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double sig = 0.0;
  const int32_t e1 = a.rowptr[i + 1];
  for (int32_t e = a.rowptr[i]; e < e1; ++e) {
    const double* r = a.rec + 8 * (int64_t)e;
    const int64_t c = (int64_t)__double_as_longlong(r[7]);
    const double gc = mixed ? (double)gi32[c] : gi[c];
    const double dg = gc - (mixed ? (double)a.gold32[c] : a.gold[c]);
#pragma unroll
    for (int q = 0; q < 6; ++q) F[q] += dg * r[q];
    sig += gc * r[6];
  }
  const double mt = a.mob[2 * i], mr = a.mob[2 * i + 1];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    F_out[6 * i + q] = F[q];
    U_out[6 * i + q] = (0.0 + mt * F[q]) + 0.0;
  }
#pragma unroll
  for (int q = 3; q < 6; ++q) {
    F_out[6 * i + q] = F[q];
    U_out[6 * i + q] = (0.0 + mr * F[q]) + 0.0;
  }
  const double imp = det_exp(sig * a.mlam);
  sig_out[i] = sig;
  imp_out[i] = imp;
  ld_out[i] = (a.len[i] * imp) * a.itau;
}

// dC = G(q) U (Physics.cpp:220-265), Euler step and renormalisation
// (Particle.cpp:37-57,110-120), write-back of ldot/impedance/stress
// (ParticleManager.cpp:89-130).
__global__ void k_integrate(int64_t N, BodyArrays b, double dt, const double* __restrict__ U,
                            const double* __restrict__ ld, double* __restrict__ ldot_field) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double* u = U + 6 * i;
  double q[4] = {b.qw[i], b.qx[i], b.qy[i], b.qz[i]};
  const double s = q[0], wx = q[1], wy = q[2], wz = q[3];
  const double xi[4][3] = {{-0.5 * wx, -0.5 * wy, -0.5 * wz},
                           {0.5 * s, -0.5 * wz, 0.5 * wy},
                           {0.5 * wz, 0.5 * s, -0.5 * wx},
                           {-0.5 * wy, 0.5 * wx, 0.5 * s}};
  double dC[7];
#pragma unroll
  for (int k = 0; k < 3; ++k) dC[k] = 0.0 + 1.0 * u[k];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    double sum = 0.0;
    sum += xi[r][0] * u[3];
    sum += xi[r][1] * u[4];
    sum += xi[r][2] * u[5];
    dC[3 + r] = sum;
  }
  b.px[i] += dt * dC[0];
  b.py[i] += dt * dC[1];
  b.pz[i] += dt * dC[2];
#pragma unroll
  for (int k = 0; k < 4; ++k) q[k] += dt * dC[3 + k];
  const double norm = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  b.qw[i] = q[0] / norm;
  b.qx[i] = q[1] / norm;
  b.qy[i] = q[2] / norm;
  b.qz[i] = q[3] / norm;
  if (ld) ldot_field[i] = ld[i];
}

// updateConstraintsFromSolution (Hard.cpp:49-89): gamma <- lagging iterate,
// phi <- linearised gap of the last evaluated iterate.
__global__ void k_update_constraints(int64_t C, int mixed, const double* __restrict__ gam_prev,
                                     const double* __restrict__ g_last,
                                     const float* __restrict__ gam_prev32,
                                     const float* __restrict__ g_last32,
                                     double* __restrict__ Sgam, double* __restrict__ Sphi) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  if (mixed) {
    Sgam[c] = (double)gam_prev32[c];
    Sphi[c] = (double)g_last32[c];
  } else {
    Sgam[c] = gam_prev[c];
    Sphi[c] = g_last[c];
  }
}

// max over the set of max(0, -phi) (Hard.cpp:125-130); max is order-free.
__global__ void k_maxov(int64_t C, const double* __restrict__ phi, double* __restrict__ part) {
  __shared__ double lds[4];
  double m = 0.0;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < C;
       c += (int64_t)gridDim.x * blockDim.x) {
    const double p = phi[c];
    const double v = p < 0 ? -p : 0;
    m = m < v ? v : m;
  }
  const double r = block_max(m, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// grow (ParticleManager.cpp:132-137 -> Particle.cpp:59-65)
__global__ void k_grow(int64_t N, double dt, double* __restrict__ len,
                       const double* __restrict__ l0, const double* __restrict__ ldot) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double l = len[i] + dt * ldot[i];
  if (l > 2 * l0[i]) l = 2 * l0[i];
  len[i] = l;
}

}  // namespace

double max_overlap(pcc_ctx* ctx) {
  const int64_t C = ctx->S.C;
  if (C == 0) return 0.0;
  const int nb = 256;
  ctx->scratch_d.ensure(nb);
  k_maxov<<<nb, 256, 0, ctx->stream>>>(C, ctx->S.phi.p, ctx->scratch_d.p);
  double h[256];
  PCC_HIP(hipMemcpyAsync(h, ctx->scratch_d.p, sizeof(double) * nb, hipMemcpyDeviceToHost,
                         ctx->stream));
  PCC_HIP(hipStreamSynchronize(ctx->stream));
  double m = 0.0;
  for (int i = 0; i < nb; ++i) m = m < h[i] ? h[i] : m;
  return m;
}

// Body-incidence CSR over the whole constraint set, each body's incidences in
// ascending constraint order, plus the per-incidence coefficient records.
void build_incidence(pcc_ctx* ctx, bool records) {
  const int64_t N = ctx->N, C = ctx->S.C;
  hipStream_t s = ctx->stream;
  KTimer tm(ctx, PCC_K_ASSEMBLE, 0.0);
  ctx->rowptr.ensure(N + 1);
  ctx->inccur.ensure(N + 1);
  ctx->inc.ensure(2 * C + 1);
  ctx->increc.ensure(8 * (2 * C + 1));
  ctx->increc32.ensure(8 * (2 * C + 1));
  PCC_HIP(hipMemsetAsync(ctx->inccur.p, 0, sizeof(int32_t) * (size_t)(N + 1), s));
  if (C > 0) k_inc_count<<<nblk(C), 256, 0, s>>>(C, ctx->S.I.p, ctx->S.J.p, ctx->inccur.p);
  exclusive_scan_i32(ctx, ctx->inccur.p, ctx->rowptr.p, N + 1);
  PCC_HIP(hipMemsetAsync(ctx->inccur.p, 0, sizeof(int32_t) * (size_t)(N + 1), s));
  if (C > 0) {
    k_inc_scatter<<<nblk(C), 256, 0, s>>>(C, ctx->S.I.p, ctx->S.J.p, ctx->rowptr.p, ctx->inccur.p,
                                          ctx->inc.p);
    seg_sort_i32(ctx, N, ctx->rowptr.p, ctx->inc.p);
    if (records)
      k_inc_records<<<nblk(2 * C), 256, 0, s>>>(2 * C, ctx->inc.p, ctx->S.n.p, ctx->S.tI.p,
                                              ctx->S.tJ.p, ctx->S.sI.p, ctx->S.sJ.p,
                                              ctx->increc.p, ctx->increc32.p);
  }
}

static BodyArgs body_args(pcc_ctx* ctx) {
  BodyArgs a;
  const int64_t N = ctx->N;
  BodyArrays b = body_arrays(ctx->body.p, N);
  a.N = N;
  a.C = ctx->S.C;
  a.rowptr = ctx->rowptr.p;
  a.rec = ctx->increc.p;
  a.rec32 = ctx->increc32.p;
  a.gam = ctx->gam.p;
  a.grd = ctx->grd.p;
  a.gold = ctx->gold.p;
  a.gam32 = ctx->gam32.p;
  a.grd32 = ctx->grd32.p;
  a.gold32 = ctx->gold32.p;
  a.mob = ctx->mob.p;
  a.len = b.len;
  a.ldot_prev = b.ldot;
  a.mlam = -ctx->params.lambda;
  a.itau = 1 / ctx->params.tau;
  a.st = ctx->state.p;
  a.ubuf = ctx->ubuf.p;
  a.ubuf32 = ctx->ubuf32.p;
  return a;
}

static ConstrArgs constr_args(pcc_ctx* ctx, int64_t nb) {
  ConstrArgs a;
  ConstraintSet& S = ctx->S;
  a.C = S.C;
  a.nb = nb;
  a.I = S.I.p;
  a.J = S.J.p;
  a.n = S.n.p;
  a.tI = S.tI.p;
  a.tJ = S.tJ.p;
  a.sI = S.sI.p;
  a.sJ = S.sJ.p;
  a.phi = ctx->phi.p;
  a.fn = S.f_n.p;
  a.ftI = S.f_tI.p;
  a.ftJ = S.f_tJ.p;
  a.fsI = S.f_sI.p;
  a.fsJ = S.f_sJ.p;
  a.phi32 = ctx->phi32.p;
  a.gam = ctx->gam.p;
  a.grd = ctx->grd.p;
  a.gam32 = ctx->gam32.p;
  a.grd32 = ctx->grd32.p;
  a.ubuf = ctx->ubuf.p;
  a.ubuf32 = ctx->ubuf32.p;
  a.dt = ctx->params.dt;
  a.tol = ctx->params.bbpgd_tol;
  a.maxit = ctx->params.max_bbpgd_iterations;
  a.st = ctx->state.p;
  a.partials = ctx->partials.p;
  a.ticket = ctx->ticket.p;
  return a;
}

// Bytes one launch of each BBPGD kernel must move at minimum (each array
// touched once per launch), used for the achieved-bandwidth figure.
static double body_bytes(int64_t N, int64_t C, bool mixed) {
  const double sv = mixed ? 4 : 8;
  // per incidence: record (8 values), gamma_p, g_p, gamma_old gathers
  // per body: rowptr, mobility (2), length, ldot_prev, U+dldot out (8)
  return 2.0 * C * (8 * sv + 3 * sv) + N * (4 + 16 + 16 + 8 * sv);
}
static double constr_bytes(int64_t N, int64_t C, bool mixed) {
  const double sv = mixed ? 4 : 8;
  // per constraint: I,J; n,tI,tJ (9); sI,sJ; phi; gamma_p, g_p in; gamma, g out
  // per body: the U+dldot record read once (8 values)
  return C * (8 + 9 * sv + 2 * sv + sv + 4 * sv) + N * 8 * sv;
}

// Runs BBPGD on the current phase operators.  Returns its status; *iters and
// *res as BBPGDResult (BBPGD.h:13-16).
pcc_status bbpgd_solve(pcc_ctx* ctx, int64_t* iters, double* res) {
  const int64_t N = ctx->N, C = ctx->S.C;
  const bool mixed = ctx->params.precision == PCC_PREC_MIXED;
  hipStream_t s = ctx->stream;
  const int64_t nb = C > 0 ? (C + 255) / 256 : 1;
  ctx->gam.ensure(2 * C + 2);
  ctx->grd.ensure(2 * C + 2);
  ctx->gold.ensure(C + 1);
  ctx->phi.ensure(C + 1);
  ctx->gam32.ensure(2 * C + 2);
  ctx->grd32.ensure(2 * C + 2);
  ctx->gold32.ensure(C + 1);
  ctx->phi32.ensure(C + 1);
  ctx->ubuf.ensure(8 * N + 8);
  ctx->ubuf32.ensure(8 * N + 8);
  ctx->mob.ensure(2 * N + 2);
  ctx->partials.ensure(4 * nb + 4);
  BodyArrays b = body_arrays(ctx->body.p, N);
  if (C > 0)
    k_phase_init<<<nblk(C), 256, 0, s>>>(C, ctx->S.gamma.p, ctx->S.phi.p, ctx->gam.p, ctx->gold.p,
                                         ctx->phi.p, ctx->gam32.p, ctx->gold32.p, ctx->phi32.p);
  if (N > 0) k_mobility<<<nblk(N), 256, 0, s>>>(N, b.len, ctx->params.xi, ctx->mob.p);
  SolverState init = {};
  PCC_HIP(hipMemcpyAsync(ctx->state.p, &init, sizeof(SolverState), hipMemcpyHostToDevice, s));
  BodyArgs ba = body_args(ctx);
  ConstrArgs ca = constr_args(ctx, nb);
  const unsigned gb = nblk(N > 0 ? N : 1), gc = (unsigned)nb;
  const double bb = body_bytes(N, C, mixed), cb = constr_bytes(N, C, mixed);
  auto one_iteration = [&]() {
    {
      KTimer t(ctx, PCC_K_BODY_APPLY, bb);
      if (mixed)
        k_body32<<<gb, 256, 0, s>>>(ba);
      else
        k_body64<<<gb, 256, 0, s>>>(ba);
    }
    {
      KTimer t(ctx, PCC_K_CONSTR_APPLY, cb);
      if (mixed)
        k_constr32<<<gc, 256, 0, s>>>(ca);
      else
        k_constr64<<<gc, 256, 0, s>>>(ca);
    }
  };
  one_iteration();  // k = 0
  const int chunk = 10;
  SolverState* hs = ctx->state_host;
  hipEvent_t ev[2];
  PCC_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  PCC_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  int64_t j = 0;
  bool done = false;
  while (!done) {
    for (int q = 0; q < chunk; ++q) one_iteration();
    PCC_HIP(hipMemcpyAsync(&hs[j & 1], ctx->state.p, sizeof(SolverState), hipMemcpyDeviceToHost, s));
    PCC_HIP(hipEventRecord(ev[j & 1], s));
    if (j >= 1) {
      PCC_HIP(hipEventSynchronize(ev[(j - 1) & 1]));
      const SolverState& h = hs[(j - 1) & 1];
      if (h.converged || h.failed) done = true;
    }
    ++j;
  }
  PCC_HIP(hipStreamSynchronize(s));
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  SolverState fin;
  PCC_HIP(hipMemcpy(&fin, ctx->state.p, sizeof(SolverState), hipMemcpyDeviceToHost));
  *iters = fin.iter;
  *res = fin.res;
  return fin.failed ? PCC_E_BBPGD_NOT_CONVERGED : PCC_OK;
}

// Workspace of the converged iterate, move, write-back, constraint update.
void integrate_hard(pcc_ctx* ctx, int64_t kf) {
  const int64_t N = ctx->N, C = ctx->S.C;
  const bool mixed = ctx->params.precision == PCC_PREC_MIXED;
  hipStream_t s = ctx->stream;
  BodyArrays b = body_arrays(ctx->body.p, N);
  KTimer t(ctx, PCC_K_INTEGRATE, 0.0);
  if (N > 0) {
    BodyArgs ba = body_args(ctx);
    k_body_final<<<nblk(N), 256, 0, s>>>(ba, kf, mixed ? 1 : 0, ctx->F.p, ctx->U.p, ctx->sig.p,
                                         ctx->imp.p, ctx->ldot_new.p);
    k_integrate<<<nblk(N), 256, 0, s>>>(N, b, ctx->params.dt, ctx->U.p, ctx->ldot_new.p, b.ldot);
  }
  if (C > 0)
    k_update_constraints<<<nblk(C), 256, 0, s>>>(
        C, mixed ? 1 : 0, ctx->gam.p + ((kf - 1) & 1) * C, ctx->grd.p + (kf & 1) * C,
        ctx->gam32.p + ((kf - 1) & 1) * C, ctx->grd32.p + (kf & 1) * C, ctx->S.gamma.p,
        ctx->S.phi.p);
}

void grow(pcc_ctx* ctx) {
  const int64_t N = ctx->N;
  if (N == 0) return;
  BodyArrays b = body_arrays(ctx->body.p, N);
  k_grow<<<nblk(N), 256, 0, ctx->stream>>>(N, ctx->params.dt, b.len, b.l0, b.ldot);
}

}  // namespace pcc
