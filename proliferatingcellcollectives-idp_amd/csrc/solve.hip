// solve.hip — hard-model ReLCP phase on the GPU: operators, BBPGD, integration.
//
// Replaces createMatrices (Hard.cpp:10-47, Physics.cpp:118-265),
// HardModelGradient::gradient/residual (HardModelGradient.cpp:89-158),
// BBPGD (BBPGD.cpp:32-127) and the write-back of Hard.cpp:160-170,
// ParticleManager.cpp:29-137, Particle.cpp:37-65,110-120.
//
// One BBPGD iteration is three launches:
//   k_body_coop  one block per 256 consecutive bodies: the block walks their
//             incidence range in LDS chunks; all lanes evaluate the projection
//             step gamma = max(gamma_p - alpha g_p, 0) of one incidence each
//             (independent gathers) and store the products with the D / L
//             coefficients; then each thread adds its own body's products in
//             ascending constraint order (PETSc's MatMultTranspose row order),
//             applies mobility, impedance exp(-lambda sigma) and the growth
//             rate, and writes U and ldot - ldot_prev.  No float atomics.
//             The same block walk writes the converged iterate's F, U,
//             sigma, impedance and ldot (MODE 1) and, once per phase in
//             mixed precision, D^T gamma_old per body (MODE 2).
//   k_constr  the same projection per constraint, D U and L dldot gathers,
//             g = Phi + dt*phidot - dt*phidot_growth, BB dot-product and
//             residual partials per block (fixed tile->block assignment and
//             reduction tree, no atomics);
//   k_finish  one block per colony: reduces its partials in a fixed order and
//             updates alpha / iteration / convergence on the device.
// A context may hold many colonies (batched ensembles): every launch covers
// all colonies of the current phase, each colony has its own solver state,
// and its blocks exit immediately once its flags say its solve is done, so the
// host can enqueue iterations speculatively until every colony has finished.
//
// Blocks are mapped XCD-aware: the hardware dispatches block h to XCD h % 8,
// so logical block (h % 8) * (G/8) + h / 8 makes each XCD sweep one
// contiguous range of bodies / constraints and keeps its gathers in its own L2.
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "common.hpp"
#include "internal.hpp"

namespace pcc {

void exclusive_scan_i32(pcc_ctx* ctx, const int32_t* in, int32_t* out, int64_t n);
void seg_sort_i32(pcc_ctx* ctx, int64_t nseg, const int32_t* start, int32_t* v);

// Solver layout.  Inside a BBPGD solve the per-constraint arrays are kept in
// "solver order": the constraint set stably sorted by the lower body index I
// (sord maps solver position -> constraint, spos the inverse).  A block of
// consecutive bodies then finds the constraints it gathers from in a few short
// runs of memory instead of one run per ReLCP phase.  Solver-order arrays:
// I, J, the operator row, Phi, the iterate record (gamma, g) of the last
// evaluated iterate (16 bytes, overwritten in place by the constraint kernel)
// and the lagging iterate gprev that the reference returns (BBPGD.cpp:124),
// saved on iterations % 10 == 0.  Each body still sums its incidences in
// ascending constraint order; gamma_old (fixed within a phase) travels with
// the incidence records.
static constexpr int kBbGridMax = 2048;
static constexpr int64_t kTuneMinConstraints = 100000;
// chunks after which a BBPGD solve's chunks double in length
static constexpr int64_t kChunkGrowAfter = 10;


namespace {

inline unsigned nblk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

__device__ __forceinline__ double proj_step(double yp, double malpha, double gp) {
  double v = yp + malpha * gp;  // VecWAXPY(gamma_i, -alpha, g_prev, gamma_prev)
  return v < 0.0 ? 0.0 : v;     // VecPointwiseMax(gamma_i, gamma_i, 0)
}

// Iterate record of constraint c (solver order): the impulse gamma and the
// gradient g of the last evaluated iterate.  FP64: one double2 (gamma, g) per
// constraint in `it`.  Mixed: a packed 12-byte record, gamma in FP64 (as its
// two 32-bit halves) and g in FP32, loaded and stored as three dwords, so that
// a gather touches one record; g is rounded once where it is evaluated, and
// every later use (projection, residual, BB dot products, the linearised gap)
// reads the rounded value.  Phi likewise: FP64, or FP32 in mixed precision.
struct Rec12 {
  uint32_t lo, hi;  // gamma's bits
  float g;
};
static_assert(sizeof(Rec12) == 12, "packed iterate record");
template <class CT>
__device__ __forceinline__ double2 load_it(const double* __restrict__ it,
                                           const float* __restrict__ itg, int64_t c) {
  (void)itg;
  if constexpr (sizeof(CT) == 8) {
    return reinterpret_cast<const double2*>(it)[c];
  } else {
    const Rec12 r = reinterpret_cast<const Rec12*>(it)[c];
    const double gam = __hiloint2double((int)r.hi, (int)r.lo);
    return make_double2(gam, (double)r.g);
  }
}
template <class CT>
__device__ __forceinline__ double load_gamma(const double* __restrict__ it, int64_t c) {
  return load_it<CT>(it, nullptr, c).x;
}
template <class CT>
__device__ __forceinline__ void store_it(double* __restrict__ it, float* __restrict__ itg,
                                         int64_t c, double gam, double g) {
  (void)itg;
  if constexpr (sizeof(CT) == 8) {
    reinterpret_cast<double2*>(it)[c] = make_double2(gam, g);
  } else {
    Rec12 r;
    r.lo = (uint32_t)__double2loint(gam);
    r.hi = (uint32_t)__double2hiint(gam);
    r.g = (float)g;
    reinterpret_cast<Rec12*>(it)[c] = r;
  }
}
template <class CT>
__device__ __forceinline__ double load_phi(const void* __restrict__ phi, int64_t c) {
  if constexpr (sizeof(CT) == 8)
    return static_cast<const double*>(phi)[c];
  else
    return (double)static_cast<const float*>(phi)[c];
}
// g as stored: FP64, or rounded to FP32 in mixed precision.
template <class CT>
__device__ __forceinline__ double stored_g(double g) {
  if constexpr (sizeof(CT) == 8)
    return g;
  else
    return (double)(float)g;
}


__global__ void k_inc_count(int64_t C, const int32_t* __restrict__ I, const int32_t* __restrict__ J,
                            int32_t* __restrict__ deg) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  atomicAdd(&deg[I[c]], 1);
  atomicAdd(&deg[J[c]], 1);
}

// Incidences into their bodies' ranges; cnt holds each body's count (from
// k_inc_count) and ends at zero.  The order within a range is the order of
// arrival; k_seg_sort then orders it.
__global__ void k_inc_scatter(int64_t C, const int32_t* __restrict__ I,
                              const int32_t* __restrict__ J, const int32_t* __restrict__ rowptr,
                              int32_t* __restrict__ cnt, int32_t* __restrict__ inc) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  int32_t a = I[c], b = J[c];
  inc[rowptr[a] + atomicSub(&cnt[a], 1) - 1] = (int32_t)(2 * c);
  inc[rowptr[b] + atomicSub(&cnt[b], 1) - 1] = (int32_t)(2 * c + 1);
}

// Solver order: positions of the constraints stably sorted by I.
__global__ void k_iota(int64_t n, int32_t* __restrict__ v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (int32_t)i;
}

__global__ void k_invert(int64_t n, const int32_t* __restrict__ ord, int32_t* __restrict__ pos) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) pos[ord[p]] = (int32_t)p;
}

// Solver-order copies of the phase's constraint data, and the initial iterate
// record (gamma_0, 0).  Row layout: planar (n_x, n_y, tauI_z, tauJ_z, sI, sJ)
// (the 2D monolayer operator: every body has z = 0 and q = (w, 0, 0, z), so
// n_z, r_z, the x/y torque components and U_z, omega_x, omega_y are exact
// zeros); 3D the packed row (n, rI x n, rJ x -n, sI, sJ, pad).
template <class CT, bool PLANAR>
__global__ void k_solver_rows(int64_t C, const int32_t* __restrict__ ord,
                              const int32_t* __restrict__ I, const int32_t* __restrict__ J,
                              const double* __restrict__ phi, const double* __restrict__ gam,
                              const double* __restrict__ coef, int2* __restrict__ sIJ,
                              void* __restrict__ sphi,
                              double* __restrict__ it, float* __restrict__ itg,
                              double* __restrict__ row, float* __restrict__ row32) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= C) return;
  const int64_t c = ord[p];
  sIJ[p] = make_int2(I[c], J[c]);
  if constexpr (sizeof(CT) == 8)
    static_cast<double*>(sphi)[p] = phi[c];
  else
    static_cast<float*>(sphi)[p] = (float)phi[c];
  store_it<CT>(it, itg, p, gam[c], 0.0);
  const double* r = coef + 12 * c;
  // only the row of the solve's precision is read
  if (PLANAR) {
    const double v[6] = {r[0], r[1], r[5], r[8], r[9], r[10]};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      if constexpr (sizeof(CT) == 8)
        row[6 * p + k] = v[k];
      else
        row32[6 * p + k] = (float)v[k];
    }
  } else {
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      if constexpr (sizeof(CT) == 8)
        row[12 * p + k] = r[k];
      else
        row32[12 * p + k] = (float)r[k];
    }
  }
}

// Incidence records, body-major: the coefficients of the D row on this body's
// DOFs ([n, rI x n] for side I, [-n, rJ x -n] for side J; planar (c_x, c_y,
// tau_z)), the stress coefficient, the constraint's solver position (3D: in
// the last slot; planar: in rc) and gamma_old, in the solve's precision.
template <bool PLANAR, bool MIXED>
__global__ void k_inc_records(int64_t E, const int32_t* __restrict__ inc,
                              const double* __restrict__ coef, const int32_t* __restrict__ spos,
                              const double* __restrict__ gam, double* __restrict__ rec,
                              float* __restrict__ rec32, int32_t* __restrict__ rc,
                              double* __restrict__ gold) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const int32_t v = inc[e];
  const int64_t c = v >> 1;
  const int side = v & 1;
  const double* row = coef + 12 * c;
  const int32_t pos = spos[c];
  gold[e] = gam[c];
  if (PLANAR) {
    double co[4];
    if (side == 0) {
      co[0] = row[0];
      co[1] = row[1];
      co[2] = row[5];
      co[3] = row[9];
    } else {
      co[0] = -row[0];
      co[1] = -row[1];
      co[2] = row[8];
      co[3] = row[10];
    }
    if (MIXED) {
      reinterpret_cast<float4*>(rec32)[e] =
          make_float4((float)co[0], (float)co[1], (float)co[2], (float)co[3]);
    } else {
      reinterpret_cast<double2*>(rec)[2 * e] = make_double2(co[0], co[1]);
      reinterpret_cast<double2*>(rec)[2 * e + 1] = make_double2(co[2], co[3]);
    }
    rc[e] = pos;
  } else {
    double co[8];
    if (side == 0) {
#pragma unroll
      for (int k = 0; k < 6; ++k) co[k] = row[k];
      co[6] = row[9];
    } else {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        co[k] = -row[k];
        co[3 + k] = row[6 + k];
      }
      co[6] = row[10];
    }
    co[7] = __longlong_as_double((long long)pos);
    if (MIXED) {
      float* r32 = rec32 + 8 * e;
#pragma unroll
      for (int k = 0; k < 7; ++k) r32[k] = (float)co[k];
      r32[7] = __int_as_float(pos);
    } else {
      double* r = rec + 8 * e;
#pragma unroll
      for (int k = 0; k < 8; ++k) r[k] = co[k];
    }
  }
}

// Translational and rotational mobility of a body of length l (Physics.cpp:
// 203-204), evaluated where a kernel needs it from the length it reads anyway
// (the same two correctly rounded divisions in every kernel).
__device__ __forceinline__ void mobility(double l, double xi, double& mt, double& mr) {
  mt = 1.0 / (l * xi);
  mr = 12.0 / (l * l * l * xi);
}

// Brownian angular velocities of the step (calculate_external_velocities,
// Physics.cpp:59-81): draws for the bodies of colonies with T > 0, zeros for
// the others.
__global__ void k_brownian(int64_t N, const double* __restrict__ imp,
                           const int32_t* __restrict__ binst, const InstDev* __restrict__ idev,
                           double* __restrict__ uext) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const InstDev& D = idev[binst[i]];
  double w[3] = {0.0, 0.0, 0.0};
  if (D.temp > 0.0 && (D.flags & kInstActive))
    brownian_omega(D.temp, D.mono, D.bseed, D.bstep, i - D.boff, imp[i], w);
  uext[3 * i] = w[0];
  uext[3 * i + 1] = w[1];
  uext[3 * i + 2] = w[2];
}

// 1 if every body is in the z = 0 plane with an in-plane orientation.
__global__ void k_planar_check(int64_t N, BodyArrays b, int32_t* __restrict__ nonplanar) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  if (b.pz[i] != 0.0 || b.qx[i] != 0.0 || b.qy[i] != 0.0) atomicOr(nonplanar, 1);
}

// ---------------------------------------------------------------------------
// BBPGD kernels, templated on the storage type CT of the operator
// coefficients (D rows, L entries, incidence records):
//   CT = double  plain FP64: bit-identical to the CPU restatement;
//   CT = float   mixed: the operator, the bulk of the bytes, is stored in
//                FP32; impulses, gaps, body velocities, every accumulation,
//                the BB dot products and the residual stay FP64.  (FP32
//                accumulation cannot reach a 1e-6 residual once impulses
//                reach ~1e5, as they do on random dense packings.)
// and on PLANAR (2D monolayer records, the exact-zero terms skipped; the
// results are bit-identical to the 3D kernels).
// ---------------------------------------------------------------------------
struct BodyArgs {
  int64_t N;           // bodies of the context
  int64_t nbb, gp;     // logical body blocks (active colonies); padded grid (multiple of 8)
  const int32_t* bmap; // logical body block -> colony
  const int32_t* binst;
  const InstDev* idev;
  const int32_t* rowptr;
  const void* rec;     // 3D: 8 CT per incidence (coef[6], s, index bits); 2D: 4 CT
  const int32_t* rc;   // 2D: constraint index per incidence
  const double* gold;  // gamma_old per incidence (FP64 kernels)
  const double* fold;  // mixed kernels: D^T gamma_old per body (3D: 6, planar: x, y, z-torque)
  double* it;          // iterate records (gamma, g), solver order (mixed: gamma)
  float* itg;          // mixed: g, solver order
  const double* len;
  const double* ldot_prev;
  const SolverState* st;  // one per colony
  double* ubuf;  // 3D: 8 per body (U[6], dldot, pad); 2D: 4 (ux, uy, wz, dldot)
  const double* uext;  // Brownian angular velocity (3 per body) of phase 0, or null
};

// U_total = U_ext + M F (HardModelGradient.cpp:15-21, VecWAXPY with
// U_ext = 0 except the Brownian angular part on phase 0).
__device__ __forceinline__ void body_velocity(const double* __restrict__ uext, int64_t i,
                                              double mt, double mr, const double F[6],
                                              double u[6]) {
  double ue[3] = {0.0, 0.0, 0.0};
  if (uext) {
    ue[0] = uext[3 * i];
    ue[1] = uext[3 * i + 1];
    ue[2] = uext[3 * i + 2];
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) u[q] = (0.0 + mt * F[q]) + 0.0;
#pragma unroll
  for (int q = 3; q < 6; ++q) u[q] = (0.0 + mr * F[q]) + ue[q - 3];
}

template <class CT>
__device__ __forceinline__ void load_rec(const void* base, int64_t e, double co[7], int64_t& c);

template <>
__device__ __forceinline__ void load_rec<double>(const void* base, int64_t e, double co[7],
                                                 int64_t& c) {
  const double2* r = reinterpret_cast<const double2*>(base) + 4 * e;
  const double2 a = r[0], b = r[1], d = r[2], f = r[3];
  co[0] = a.x;
  co[1] = a.y;
  co[2] = b.x;
  co[3] = b.y;
  co[4] = d.x;
  co[5] = d.y;
  co[6] = f.x;
  c = (int64_t)__double_as_longlong(f.y);
}

template <>
__device__ __forceinline__ void load_rec<float>(const void* base, int64_t e, double co[7],
                                                int64_t& c) {
  const float4* r = reinterpret_cast<const float4*>(base) + 2 * e;
  const float4 a = r[0], b = r[1];
  co[0] = a.x;
  co[1] = a.y;
  co[2] = a.z;
  co[3] = a.w;
  co[4] = b.x;
  co[5] = b.y;
  co[6] = b.z;
  c = (int64_t)__float_as_int(b.w);
}

template <class CT>
__device__ __forceinline__ void load_prec(const void* base, int64_t e, double co[4]);

template <>
__device__ __forceinline__ void load_prec<double>(const void* base, int64_t e, double co[4]) {
  const double2* r = reinterpret_cast<const double2*>(base) + 2 * e;
  const double2 a = r[0], b = r[1];
  co[0] = a.x;
  co[1] = a.y;
  co[2] = b.x;
  co[3] = b.y;
}

template <>
__device__ __forceinline__ void load_prec<float>(const void* base, int64_t e, double co[4]) {
  const float4 a = reinterpret_cast<const float4*>(base)[e];
  co[0] = a.x;
  co[1] = a.y;
  co[2] = a.z;
  co[3] = a.w;
}

// One incidence record in the form it is loaded (the values of load_rec /
// load_prec, converted only where they are used), so that a kernel can keep
// the next records in flight while it finishes the current ones.
template <class CT, bool PLANAR>
struct IncRaw;

template <>
struct IncRaw<float, true> {
  float4 r;
  int32_t c;
  __device__ __forceinline__ void load(const void* rec, const int32_t* rc, int64_t e) {
    r = reinterpret_cast<const float4*>(rec)[e];
    c = rc[e];
  }
  __device__ __forceinline__ int64_t index() const { return c; }
  __device__ __forceinline__ void coef(double co[7]) const {
    co[0] = r.x;
    co[1] = r.y;
    co[5] = r.z;
    co[6] = r.w;
  }
};

template <>
struct IncRaw<double, true> {
  double2 r0, r1;
  int32_t c;
  __device__ __forceinline__ void load(const void* rec, const int32_t* rc, int64_t e) {
    const double2* p = reinterpret_cast<const double2*>(rec) + 2 * e;
    r0 = p[0];
    r1 = p[1];
    c = rc[e];
  }
  __device__ __forceinline__ int64_t index() const { return c; }
  __device__ __forceinline__ void coef(double co[7]) const {
    co[0] = r0.x;
    co[1] = r0.y;
    co[5] = r1.x;
    co[6] = r1.y;
  }
};

template <>
struct IncRaw<float, false> {
  float4 r0, r1;
  __device__ __forceinline__ void load(const void* rec, const int32_t*, int64_t e) {
    const float4* p = reinterpret_cast<const float4*>(rec) + 2 * e;
    r0 = p[0];
    r1 = p[1];
  }
  __device__ __forceinline__ int64_t index() const { return (int64_t)__float_as_int(r1.w); }
  __device__ __forceinline__ void coef(double co[7]) const {
    co[0] = r0.x;
    co[1] = r0.y;
    co[2] = r0.z;
    co[3] = r0.w;
    co[4] = r1.x;
    co[5] = r1.y;
    co[6] = r1.z;
  }
};

template <>
struct IncRaw<double, false> {
  double2 r0, r1, r2, r3;
  __device__ __forceinline__ void load(const void* rec, const int32_t*, int64_t e) {
    const double2* p = reinterpret_cast<const double2*>(rec) + 4 * e;
    r0 = p[0];
    r1 = p[1];
    r2 = p[2];
    r3 = p[3];
  }
  __device__ __forceinline__ int64_t index() const { return (int64_t)__double_as_longlong(r3.y); }
  __device__ __forceinline__ void coef(double co[7]) const {
    co[0] = r0.x;
    co[1] = r0.y;
    co[2] = r1.x;
    co[3] = r1.y;
    co[4] = r2.x;
    co[5] = r2.y;
    co[6] = r3.x;
  }
};

// Mixed precision: F = D^T (gamma - gamma_old) is evaluated as D^T gamma
// minus the per-body sum D^T gamma_old, which is fixed within a phase, so the
// iteration reads 3 or 6 doubles per body instead of gamma_old per incidence.
// Both sums use the same FP32 coefficients and FP64 accumulation.
__device__ __forceinline__ void unfold(const double* __restrict__ fold, int64_t i, double F[6],
                                       bool planar) {
  if (planar) {
    const double* f = fold + 3 * i;
    F[0] -= f[0];
    F[1] -= f[1];
    F[5] -= f[2];
  } else {
    const double* f = fold + 6 * i;
#pragma unroll
    for (int q = 0; q < 6; ++q) F[q] -= f[q];
  }
}

// Outputs of the converged iterate's workspace (the final body walk): F and
// U (6 per body), sigma, impedance, ldot; the fold walk writes D^T gamma_old
// through F.
struct FinalOut {
  double *F, *U, *sig, *imp, *ld;
};

// Element idx of an array of T at `base`, addressed by a 32-bit byte offset.
template <class T>
__device__ __forceinline__ T* at(const void* base, uint32_t idx) {
  return reinterpret_cast<T*>(const_cast<char*>(static_cast<const char*>(base)) +
                              (size_t)(idx * (uint32_t)sizeof(T)));
}

// Iterate record as stored (FP64 pair, or the packed 12-byte record),
// converted where it is used so that a prefetched tile holds fewer registers.
template <class CT>
struct ItRaw {
  double2 v;
  __device__ __forceinline__ void load(const double* __restrict__ it, int64_t c) {
    v = reinterpret_cast<const double2*>(it)[c];
  }
  __device__ __forceinline__ double2 value() const { return v; }
};
template <>
struct ItRaw<float> {
  Rec12 v;
  __device__ __forceinline__ void load(const double* __restrict__ it, int64_t c) {
    v = reinterpret_cast<const Rec12*>(it)[c];
  }
  __device__ __forceinline__ double2 value() const {
    return make_double2(__hiloint2double((int)v.hi, (int)v.lo), (double)v.g);
  }
};

// Body apply of one BBPGD iteration, block-cooperative: a block owns 256
// consecutive bodies and walks their incidence range in chunks.  Phase 1: every
// lane takes incidences of the chunk in turn, gathers the constraint's
// iterate record (independent loads, full memory-level parallelism) and
// writes the products dg*coef and gamma*s to LDS.  Phase 2: each thread adds
// the products of its own bodies (256 / TPB of them) in ascending incidence
// order, so every sum has exactly the order of the thread-per-body loop
// above.  Blocks never straddle two colonies (each colony starts a new
// logical block).  CHUNK (incidences staged per pass) and TPB (threads per
// block: 256, or 128 so that a 600k-body grid fits the device in one round)
// do not change any sum and are picked by the autotuner.  With PIPE, each
// lane issues the next chunk's incidence records as soon as the current
// chunk's iterate gathers are in flight, so that their latency overlaps the
// products, the LDS sums and the two barriers of the current chunk.
// MODE 0: the BBPGD body apply (velocity records of the iterate about to
// run).  MODE 1: the workspace of the converged iterate for the colonies
// being integrated (F = D^T (gamma - gamma_old), U = M F, sigma, impedance,
// ldot into fo; HardModelGradient.cpp:89-121).  MODE 2 (mixed precision
// only): D^T gamma_old per body with the FP32 coefficients, once per phase.
// All three sum each body's incidences in ascending order.
template <class CT, bool PLANAR, int CHUNK, int TPB, bool PIPE, int MODE>
__device__ __forceinline__ void body_block(const BodyArgs& a, int64_t l,
                                           double (*prod)[CHUNK], const FinalOut& fo) {
  constexpr bool FOLD = sizeof(CT) == 4;  // mixed: gamma_old enters as the per-body sum
  constexpr bool GOLD = !FOLD || MODE == 2;  // gamma_old per incidence is read
  static_assert(MODE != 2 || FOLD, "the fold is mixed precision");
  constexpr int R = PCC_BLOCK / TPB;      // bodies per thread
  static_assert(CHUNK % TPB == 0 && CHUNK >= TPB && PCC_BLOCK % TPB == 0, "staging shape");
  if (l >= a.nbb) return;
  const int32_t inst = a.bmap[l];
  const SolverState* st = a.st + inst;
  const InstDev& D = a.idev[inst];
  if (MODE == 0 && (st->skip | st->converged | st->failed)) return;
  if (MODE == 1 && !(D.flags & kInstInteg)) return;
  if (MODE == 2 && st->skip) return;
  const int64_t bend = D.boff + D.N;
  const int64_t b0 = D.boff + (l - D.bb) * PCC_BLOCK;
  const int64_t b1 = b0 + PCC_BLOCK < bend ? b0 + PCC_BLOCK : bend;
  const int64_t k = MODE == 0 ? st->iter : 0;  // MODE 1: gamma as stored
  const double malpha = MODE == 0 ? -st->alpha : 0.0;
  const int32_t E0 = a.rowptr[b0], E1 = a.rowptr[b1];
  int32_t e_my[R], e_end[R];
  double F[R][6], sig[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int64_t i = b0 + threadIdx.x + r * TPB;
    const bool act = i < b1;
    e_my[r] = act ? a.rowptr[i] : 0;
    e_end[r] = act ? a.rowptr[i + 1] : 0;
#pragma unroll
    for (int q = 0; q < 6; ++q) F[r][q] = 0.0;
    sig[r] = 0.0;
  }
  constexpr int PER = CHUNK / TPB;  // incidences per lane and chunk
  // Phase 1 in three sweeps so that every lane has all of its loads in flight
  // at once: records, then the iterate-record gathers, then products.  Every
  // lane loads and writes all PER slots of a chunk, also past the end of the
  // range (loads clamped to the block's last incidence; the sums never read
  // those slots), and scheduling barriers keep the sweeps apart: with no
  // loads under branches and no loads interleaved, the compiler waits for
  // each load exactly, so the gathers of a sweep overlap and, with PIPE, the
  // next chunk's records stay in flight through the products and the sums.
  IncRaw<CT, PLANAR> raw[PER];
  bool ok[PER];
  auto fetch = [&](int32_t from, int32_t lim) {  // requires E0 < E1
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int32_t e = from + (int32_t)threadIdx.x + j * TPB;
      ok[j] = e < lim;
      raw[j].load(a.rec, a.rc, ok[j] ? e : E1 - 1);
    }
  };
  if (PIPE && E0 < E1) fetch(E0, E0 + CHUNK < E1 ? E0 + CHUNK : E1);
  for (int32_t base = E0; base < E1; base += CHUNK) {
    const int32_t lim = base + CHUNK < E1 ? base + CHUNK : E1;
    if (!PIPE) fetch(base, lim);
    __builtin_amdgcn_sched_barrier(0);
    double2 yg[PER];
    double gold[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      if (MODE != 2) yg[j] = load_it<CT>(a.it, a.itg, raw[j].index());
      if (GOLD) gold[j] = a.gold[ok[j] ? base + (int32_t)threadIdx.x + j * TPB : E1 - 1];
    }
    IncRaw<CT, PLANAR> cur[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) cur[j] = raw[j];
    __builtin_amdgcn_sched_barrier(0);
    // the next chunk's records (past E1 on the last chunk: unused)
    if (PIPE) fetch(lim, lim + CHUNK < E1 ? lim + CHUNK : E1);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      double co[7];
      cur[j].coef(co);
      const double gc = MODE == 2 ? gold[j] : (k == 0) ? yg[j].x : proj_step(yg[j].x, malpha, yg[j].y);
      const double dg = (FOLD || MODE == 2) ? gc : gc - gold[j];
      const int q = (int)threadIdx.x + j * TPB;
      if (PLANAR) {
        prod[0][q] = dg * co[0];
        prod[1][q] = dg * co[1];
        prod[2][q] = dg * co[5];
        prod[3][q] = gc * co[6];
      } else {
#pragma unroll
        for (int m = 0; m < 6; ++m) prod[m][q] = dg * co[m];
        prod[6][q] = gc * co[6];
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) {
      for (; e_my[r] < e_end[r] && e_my[r] < lim; ++e_my[r]) {
        const int q = e_my[r] - base;
        if (PLANAR) {
          F[r][0] += prod[0][q];
          F[r][1] += prod[1][q];
          F[r][5] += prod[2][q];
          sig[r] += prod[3][q];
        } else {
#pragma unroll
          for (int m = 0; m < 6; ++m) F[r][m] += prod[m][q];
          sig[r] += prod[6][q];
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int64_t i = b0 + threadIdx.x + r * TPB;
    if (i >= b1) continue;
    if (MODE == 2) {
      if (PLANAR) {
        double* f = fo.F + 3 * i;
        f[0] = F[r][0];
        f[1] = F[r][1];
        f[2] = F[r][5];
      } else {
#pragma unroll
        for (int q = 0; q < 6; ++q) fo.F[6 * i + q] = F[r][q];
      }
      continue;
    }
    if (FOLD) unfold(a.fold, i, F[r], PLANAR);
    const double li = a.len[i];
    double mt, mr;
    mobility(li, D.xi, mt, mr);
    double u[6];
    body_velocity(a.uext, i, mt, mr, F[r], u);
    const double imp = det_exp(sig[r] * D.mlam);
    const double ld = (li * imp) * D.itau;
    if (MODE == 1) {
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        fo.F[6 * i + q] = F[r][q];
        fo.U[6 * i + q] = u[q];
      }
      fo.sig[i] = sig[r];
      fo.imp[i] = imp;
      fo.ld[i] = ld;
      continue;
    }
    if (PLANAR) {
      double2* o = reinterpret_cast<double2*>(a.ubuf + 4 * i);
      o[0] = make_double2(u[0], u[1]);
      o[1] = make_double2(u[5], ld - a.ldot_prev[i]);
    } else {
      double2* o = reinterpret_cast<double2*>(a.ubuf + 8 * i);
      o[0] = make_double2(u[0], u[1]);
      o[1] = make_double2(u[2], u[3]);
      o[2] = make_double2(u[4], u[5]);
      o[3] = make_double2(ld - a.ldot_prev[i], 0.0);
    }
  }
}

template <class CT, bool PLANAR, int CHUNK, int TPB, bool PIPE, int MODE = 0>
__global__ void __launch_bounds__(TPB) k_body_coop(BodyArgs a, FinalOut fo) {
  constexpr int NC = PLANAR ? 4 : 7;  // products per incidence
  __shared__ double prod[NC][CHUNK];
  body_block<CT, PLANAR, CHUNK, TPB, PIPE, MODE>(a, xcd_block(blockIdx.x, a.gp), prod, fo);
}

struct ConstrArgs {
  int64_t ncb, gp;      // logical constraint blocks (active colonies); padded grid
  int64_t Gtot;         // stride of the partials arrays
  const int32_t* cmap;  // logical constraint block -> colony
  const InstDev* idev;
  const int2* IJ;       // (I, J) per constraint, solver order
  const void* coef;     // 3D: 12 CT per constraint (n, rIxn, rJx-n, sI, sJ, pad); 2D: 6 CT
  const void* phi;      // Phi, solver order (mixed: FP32)
  double* it;
  float* itg;           // mixed: g
  double* gprev;        // lagging iterate saved on iterations % 10 == 0
  const double* ubuf;
  SolverState* st;      // one per colony
  double* partials;     // 4 * Gtot (5 * Gtot while tracing: gamma.Phi)
  // BBPGD trace (all null / 0 when off)
  int32_t trace;
  int64_t trace_cap;
  pcc_bbpgd_record* trace_buf;  // trace_cap records per colony
  int64_t* trace_n;             // records per colony this step
  const double* epart;          // energy partials: [nbb] F.MF, [nbb] l exp(-lambda sigma)
  int64_t nbb;
};

// Operator row of one constraint as stored (CT), kept unconverted in
// registers: planar (nx, ny, tauI_z, tauJ_z, sI, sJ); 3D (n, rI x n,
// rJ x -n, sI, sJ).
template <class CT, bool PLANAR>
struct alignas(PLANAR && sizeof(CT) == 4 ? 8 : 16) Row {
  CT v[PLANAR ? 6 : 12];
};

template <class CT, bool PLANAR>
__device__ __forceinline__ Row<CT, PLANAR> load_row(const void* base, int64_t c) {
  Row<CT, PLANAR> r;
  constexpr int W = PLANAR ? 6 : 12;
  if constexpr (sizeof(CT) == 8) {
    const double2* p = reinterpret_cast<const double2*>(base) + (W / 2) * c;
#pragma unroll
    for (int q = 0; q < W / 2; ++q) {
      const double2 v = p[q];
      r.v[2 * q] = v.x;
      r.v[2 * q + 1] = v.y;
    }
  } else if constexpr (PLANAR) {
    const float2* p = reinterpret_cast<const float2*>(base) + 3 * c;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const float2 v = p[q];
      r.v[2 * q] = v.x;
      r.v[2 * q + 1] = v.y;
    }
  } else {
    const float4* p = reinterpret_cast<const float4*>(base) + 3 * c;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const float4 v = p[q];
      r.v[4 * q] = v.x;
      r.v[4 * q + 1] = v.y;
      r.v[4 * q + 2] = v.z;
      r.v[4 * q + 3] = v.w;
    }
  }
  return r;
}

// Inputs of one constraint of a tile, as stored.
template <class CT, bool PLANAR>
struct TileIn {
  int32_t I, J;
  ItRaw<CT> yg;  // (gamma_p, g_p)
  CT phi;
  Row<CT, PLANAR> row;
};

// A colony with Q tiles of 256 constraints gets G = min(Q, 2048) logical
// blocks of K = ceil(Q / G) consecutive tiles; every thread accumulates its
// lanes of those tiles in order, the block tree combines the threads, and
// k_finish combines the G block partials in a fixed order.  The tile loop is
// software-pipelined: tile j+1's indices, row and iterate record are loaded
// while tile j's body-velocity gathers are in flight.
// The scalar update of one colony's BBPGD iteration (BBPGD.cpp:49-116):
// its block partials reduced in a fixed order, then residual, convergence,
// BB step size and iteration count on the device.
__device__ __forceinline__ void finish_colony(const ConstrArgs& a, int32_t q, double* lds,
                                              double (*l4)[4]) {
  SolverState* st = a.st + q;
  if (st->skip | st->converged | st->failed) return;
  const InstDev& D = a.idev[q];
  const int64_t k = st->iter;
  const int64_t G = D.G, o = D.lb, T = a.Gtot;
  // Thread t adds partials t, t + 256, ... (G <= kBbGridMax, so at most KP of
  // them) in ascending order.  All of its loads are issued first, at an
  // in-range index for the slots past G, and every sum is formed and then
  // selected (a slot past G leaves it unchanged), so that no load waits
  // behind a branch and they all overlap.
  constexpr int KP = kBbGridMax / PCC_BLOCK;
  double p0[KP], p1[KP], p2[KP], pm[KP];
#pragma unroll
  for (int j = 0; j < KP; ++j) {
    const int64_t i = threadIdx.x + j * PCC_BLOCK;
    const int64_t c = o + (i < G ? i : 0);
    p0[j] = a.partials[c];
    p1[j] = a.partials[T + c];
    p2[j] = a.partials[2 * T + c];
    pm[j] = a.partials[3 * T + c];
  }
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, am = 0.0;
#pragma unroll
  for (int j = 0; j < KP; ++j) {
    const bool in = threadIdx.x + j * PCC_BLOCK < G;
    const double s0 = a0 + p0[j], s1 = a1 + p1[j], s2 = a2 + p2[j];
    const double sm = am < pm[j] ? pm[j] : am;
    a0 = in ? s0 : a0;
    a1 = in ? s1 : a1;
    a2 = in ? s2 : a2;
    am = in ? sm : am;
  }
  // block_sum x 3 and block_max with one LDS exchange: the same wave
  // butterflies and the same ((w0 + w1) + w2) + w3 / max chain over waves
  {
    const double w0 = wave_sum(a0), w1 = wave_sum(a1), w2 = wave_sum(a2), wm = wave_max(am);
    if ((threadIdx.x & (PCC_WAVE - 1)) == 0) {
      const int w = threadIdx.x >> 6;
      l4[0][w] = w0;
      l4[1][w] = w1;
      l4[2][w] = w2;
      l4[3][w] = wm;
    }
    __syncthreads();
  }
  double S0 = 0.0, S1 = 0.0, S2 = 0.0, R = 0.0;
  if (threadIdx.x == 0) {
    S0 = ((l4[0][0] + l4[0][1]) + l4[0][2]) + l4[0][3];
    S1 = ((l4[1][0] + l4[1][1]) + l4[1][2]) + l4[1][3];
    S2 = ((l4[2][0] + l4[2][1]) + l4[2][2]) + l4[2][3];
    R = l4[3][0];
    for (int i = 1; i < 4; ++i) R = R < l4[3][i] ? l4[3][i] : R;
  }
  // BBPGD trace record of this iterate (BBPGD.cpp:57-60,76-80): energy terms
  // from the gamma.Phi block partials and the colony's body-block partials,
  // each combined in the fixed strided order above
  double E1 = 0.0, E2 = 0.0, E3 = 0.0;
  if (a.trace) {
    double a4 = 0.0, a5 = 0.0, a6 = 0.0;
    for (int64_t i = threadIdx.x; i < G; i += PCC_BLOCK) a4 += a.partials[4 * T + o + i];
    const int64_t nb = (D.N + PCC_BLOCK - 1) / PCC_BLOCK;
    for (int64_t i = threadIdx.x; i < nb; i += PCC_BLOCK) {
      a5 += a.epart[D.bb + i];
      a6 += a.epart[a.nbb + D.bb + i];
    }
    E1 = block_sum(a4, lds);
    E2 = block_sum(a5, lds);
    E3 = block_sum(a6, lds);
  }
  if (threadIdx.x != 0) return;
  auto record = [&](double res, double alpha) {
    const int64_t n = a.trace_n[q];
    if (n < a.trace_cap) {
      pcc_bbpgd_record* r = a.trace_buf + (int64_t)q * a.trace_cap + n;
      const double t2 = E2 * (D.dt / 2.0);
      const double t3 = E3 * (D.dt / (D.lam * D.tau));
      r->step = k;
      r->recursive_iteration = D.phase;
      r->total_constraints = D.C;
      r->residual = res;
      r->step_size = alpha;
      r->linear = E1;
      r->quadratic = t2;
      r->growth = t3;
      r->total = E1 + t2 + t3;
    }
    a.trace_n[q] = n + 1;
  };
  if (k == 0) {  // initial gradient: res = residual(g, Gamma), alpha = 1/res
    st->res = R;
    st->alpha = 1.0 / R;
    st->iter = 1;
    if (a.trace) record(R, st->alpha);
    if (1 == D.maxit) st->failed = 1;
    return;
  }
  if (k % D.check == 0) {
    st->res = R;
    if (a.trace) record(R, st->alpha);
    if (R <= D.tol) {
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
  if (k + 1 == D.maxit) st->failed = 1;
}

template <class CT, bool PLANAR>
__device__ __forceinline__ void constr_block(const ConstrArgs& a, int64_t l, double* lds) {
  if (l >= a.ncb) return;
  const int32_t inst = a.cmap[l];
  const SolverState* st = a.st + inst;
  if (st->skip | st->converged | st->failed) return;
  const InstDev& D = a.idev[inst];
  const int64_t b = l - D.lb;
  const int64_t K = D.K, cend = D.coff + D.C;
  const double dt = D.dt;
  const int64_t k = st->iter;
  const double malpha = -st->alpha;
  const bool save_prev = k > 0 && k % D.check == 0;
  constexpr int UW = PLANAR ? 4 : 8;  // doubles per body velocity record
  // Element addresses are the arrays' (uniform) bases plus 32-bit byte
  // offsets (the host keeps every array of the set under 4 GiB), so the loads
  // take the scalar-base form and the loop carries one index instead of a
  // 64-bit pointer per array.
  auto fetch = [&](uint32_t c, TileIn<CT, PLANAR>& t) {
    const int2 ij = *at<int2>(a.IJ, c);
    t.I = ij.x;
    t.J = ij.y;
    t.yg.v = *at<decltype(t.yg.v)>(a.it, c);
    t.phi = *at<CT>(a.phi, c);
    t.row = *at<Row<CT, PLANAR>>(a.coef, c);
  };
  double v0 = 0.0, v1 = 0.0, v2 = 0.0, vr = 0.0, ve = 0.0;
  auto gatherU = [&](const TileIn<CT, PLANAR>& t, double2* ui, double2* uj) {
    const double2* UI = at<double2>(a.ubuf, (uint32_t)t.I * (UW / 2));
    const double2* UJ = at<double2>(a.ubuf, (uint32_t)t.J * (UW / 2));
#pragma unroll
    for (int q = 0; q < UW / 2; ++q) {
      ui[q] = UI[q];
      uj[q] = UJ[q];
    }
  };
  auto apply = [&](const TileIn<CT, PLANAR>& cur, const double2* ui, const double2* uj,
                   uint32_t c) {
    const double2 yg = cur.yg.value();
    const double phi = (double)cur.phi;
    const double gc = (k == 0) ? yg.x : proj_step(yg.x, malpha, yg.y);
    double pd = 0.0;  // D U_total, columns of body I (lower gid) first
    double pg = 0.0;  // L ldot_diff
    const CT* co = cur.row.v;
    if (PLANAR) {
      pd += (double)co[0] * ui[0].x;
      pd += (double)co[1] * ui[0].y;
      pd += (double)co[2] * ui[1].x;
      pd += (-(double)co[0]) * uj[0].x;
      pd += (-(double)co[1]) * uj[0].y;
      pd += (double)co[3] * uj[1].x;
      pg += (double)co[4] * ui[1].y;
      pg += (double)co[5] * uj[1].y;
    } else {
      pd += (double)co[0] * ui[0].x;
      pd += (double)co[1] * ui[0].y;
      pd += (double)co[2] * ui[1].x;
      pd += (double)co[3] * ui[1].y;
      pd += (double)co[4] * ui[2].x;
      pd += (double)co[5] * ui[2].y;
      pd += (-(double)co[0]) * uj[0].x;
      pd += (-(double)co[1]) * uj[0].y;
      pd += (-(double)co[2]) * uj[1].x;
      pd += (double)co[6] * uj[1].y;
      pd += (double)co[7] * uj[2].x;
      pd += (double)co[8] * uj[2].y;
      pg += (double)co[9] * ui[3].x;
      pg += (double)co[10] * uj[3].x;
    }
    const double g = stored_g<CT>(phi + (dt * pd + (-dt) * pg));
    if constexpr (sizeof(CT) == 8) {
      *at<double2>(a.it, c) = make_double2(gc, g);
    } else {
      Rec12 r;
      r.lo = (uint32_t)__double2loint(gc);
      r.hi = (uint32_t)__double2hiint(gc);
      r.g = (float)g;
      *at<Rec12>(a.it, c) = r;
    }
    if (save_prev) *at<double>(a.gprev, c) = yg.x;
    if (k > 0) {
      const double dgam = gc - yg.x;
      const double dphi = g - yg.y;
      v0 += dgam * dgam;
      v1 += dgam * dphi;
      v2 += dphi * dphi;
    }
    const double pj = gc > 0 ? g : (g < 0.0 ? g : 0.0);
    const double ar = fabs(pj);
    vr = vr < ar ? ar : vr;
    if (a.trace) ve += gc * phi;  // energy: gamma . Phi
  };
  const int64_t c0 = D.coff + b * K * PCC_BLOCK + threadIdx.x;
  bool ok = K > 0 && c0 < cend;
  uint32_t c = (uint32_t)c0;
  const uint32_t cend32 = (uint32_t)cend;
  const int32_t K32 = (int32_t)K;
  TileIn<CT, PLANAR> cur;
  if (ok) fetch(c, cur);
  for (int32_t j = 0; j < K32 && ok; ++j) {
    double2 ui[UW / 2], uj[UW / 2];
    gatherU(cur, ui, uj);
    const uint32_t cn = c + PCC_BLOCK;
    const bool okn = j + 1 < K32 && cn < cend32;
    // (measured: loading the next tile in every lane, clamped, instead of
    // under okn, or no prefetch at all, was no faster)
    TileIn<CT, PLANAR> nxt;
    if (okn) fetch(cn, nxt);
    apply(cur, ui, uj, c);
    c = cn;
    ok = okn;
    if (okn) cur = nxt;
  }
  // block_sum x 3 (4 while tracing) and block_max with one LDS exchange: the
  // same wave butterflies and the same ((w0 + w1) + w2) + w3 combination
  const double w0 = wave_sum(v0), w1 = wave_sum(v1), w2 = wave_sum(v2), wm = wave_max(vr);
  const double we = a.trace ? wave_sum(ve) : 0.0;
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & (PCC_WAVE - 1)) == 0) {
    lds[wv] = w0;
    lds[4 + wv] = w1;
    lds[8 + wv] = w2;
    lds[12 + wv] = wm;
    lds[16 + wv] = we;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double rm = lds[12];
    for (int i = 1; i < 4; ++i) rm = rm < lds[12 + i] ? lds[12 + i] : rm;
    a.partials[l] = ((lds[0] + lds[1]) + lds[2]) + lds[3];
    a.partials[a.Gtot + l] = ((lds[4] + lds[5]) + lds[6]) + lds[7];
    a.partials[2 * a.Gtot + l] = ((lds[8] + lds[9]) + lds[10]) + lds[11];
    a.partials[3 * a.Gtot + l] = rm;
    if (a.trace) a.partials[4 * a.Gtot + l] = ((lds[16] + lds[17]) + lds[18]) + lds[19];
  }
}

template <class CT, bool PLANAR>
__global__ void __launch_bounds__(256) k_constr(ConstrArgs a) {
  __shared__ double lds[20];
  constr_block<CT, PLANAR>(a, xcd_block(blockIdx.x, a.gp), lds);
}

// Energy of the iterate just evaluated (HardModelGradient.cpp:160-203), body
// part, while tracing: per body F = D^T gamma and sigma = L^T gamma in
// ascending incidence order, U = M F, F.U (six products in order) and
// l exp(-lambda sigma); per logical body block their block sums.
template <class CT, bool PLANAR>
__global__ void __launch_bounds__(256) k_energy_body(BodyArgs a, double* __restrict__ epart) {
  __shared__ double lds[4];
  const int64_t l = xcd_block(blockIdx.x, a.gp);
  if (l >= a.nbb) return;
  const int32_t inst = a.bmap[l];
  const SolverState* st = a.st + inst;
  if (st->skip | st->converged | st->failed) return;
  const InstDev& D = a.idev[inst];
  const int64_t i = D.boff + (l - D.bb) * PCC_BLOCK + threadIdx.x;
  double q2 = 0.0, q3 = 0.0;
  if (i < D.boff + D.N) {
    double F[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double sig = 0.0;
    const int32_t e1 = a.rowptr[i + 1];
    for (int32_t e = a.rowptr[i]; e < e1; ++e) {
      double co[7];
      int64_t c;
      if (PLANAR) {
        double p[4];
        load_prec<CT>(a.rec, e, p);
        co[0] = p[0];
        co[1] = p[1];
        co[5] = p[2];
        co[6] = p[3];
        c = a.rc[e];
      } else {
        load_rec<CT>(a.rec, e, co, c);
      }
      const double gc = load_gamma<CT>(a.it, c);
      if (PLANAR) {
        F[0] += gc * co[0];
        F[1] += gc * co[1];
        F[5] += gc * co[5];
      } else {
#pragma unroll
        for (int q = 0; q < 6; ++q) F[q] += gc * co[q];
      }
      sig += gc * co[6];
    }
    const double li = a.len[i];
    double mt, mr;
    mobility(li, D.xi, mt, mr);
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const double u = (0.0 + (q < 3 ? mt : mr) * F[q]) + 0.0;
      s += F[q] * u;
    }
    q2 = s;
    q3 = li * det_exp(sig * D.mlam);
  }
  const double s2 = block_sum(q2, lds);
  const double s3 = block_sum(q3, lds);
  if (threadIdx.x == 0) {
    epart[l] = s2;
    epart[a.nbb + l] = s3;
  }
}

// Reduction of a colony's block partials in a fixed order and its BB
// step-size / convergence update (BBPGD.cpp:49-116).  One block of
// PCC_BLOCK threads per colony.
__global__ void __launch_bounds__(256) k_finish(ConstrArgs a) {
  __shared__ double lds[4];
  __shared__ double l4[4][4];
  finish_colony(a, (int32_t)blockIdx.x, lds, l4);
}

// Persistent BBPGD for colonies small enough for one workgroup: a block of
// kPersistThreads threads runs one colony's whole solve (BBPGD.cpp:32-127)
// in a single launch, with no launch and no host round trip per iteration.
// It evaluates exactly the sums of k_body_coop, k_constr and k_finish, in
// their order:
//   * body apply as k_body_coop, over groups of kPersistThreads bodies: the
//     group's incidences staged through LDS in chunks (all loads of a chunk
//     in flight), each thread adding its own body's products in ascending
//     incidence order;
//   * constraint apply as k_constr: the lanes of each logical block of 256
//     (taken by one group of four waves) reduced by a 64-lane butterfly per
//     wave, kept in LDS, and combined as ((w0 + w1) + w2) + w3;
//   * the block partials strided over 256 threads and reduced as k_finish
//     does.
// So it is bit-identical to the multi-kernel path.  The loop ends at
// convergence or at max_bbpgd_iterations.  Requires G <= kPersistMaxTiles.
constexpr int kPersistThreads = 1024;
constexpr int kPersistGroups = kPersistThreads / PCC_BLOCK;
constexpr int kPersistMaxTiles = 32;
constexpr int kPersistChunk = 1024;  // incidences staged per pass

template <class CT, bool PLANAR>
__global__ void __launch_bounds__(kPersistThreads)
    k_bbpgd_persist(BodyArgs ba, ConstrArgs ca, const int32_t* __restrict__ plist) {
  constexpr bool FOLD = sizeof(CT) == 4;
  constexpr int UW = PLANAR ? 4 : 8;
  constexpr int NC = PLANAR ? 4 : 7;  // products per incidence
  constexpr int CHUNK = kPersistChunk;
  constexpr int PER = CHUNK / kPersistThreads;
  __shared__ double prod[NC][CHUNK];
  __shared__ double red[kPersistMaxTiles][4][4];  // logical block, quantity, wave
  __shared__ double fin[4][4];                    // quantity, wave
  __shared__ double s_alpha;
  __shared__ int64_t s_iter;
  __shared__ int32_t s_done;
  const int32_t q = plist[blockIdx.x];
  SolverState* st = ca.st + q;
  const InstDev& D = ca.idev[q];
  const int tid = threadIdx.x;
  const int grp = tid / PCC_BLOCK, lt = tid % PCC_BLOCK;
  const int wv = lt / PCC_WAVE, lane = lt % PCC_WAVE;
  if (tid == 0) {
    s_alpha = st->alpha;
    s_iter = st->iter;
    s_done = (st->skip | st->converged | st->failed) ? 1 : 0;
  }
  __syncthreads();
  const int64_t bend = D.boff + D.N, cend = D.coff + D.C;
  const int64_t G = D.G, K = D.K;
  const double dt = D.dt;
  while (!s_done) {
    const int64_t k = s_iter;
    const double malpha = -s_alpha;
    // ---- body apply (k_body_coop), a group of kPersistThreads bodies at a time
    for (int64_t g0 = D.boff; g0 < bend; g0 += kPersistThreads) {
      const int64_t g1 = g0 + kPersistThreads < bend ? g0 + kPersistThreads : bend;
      const int64_t i = g0 + tid;
      const bool active = i < g1;
      const int32_t E0 = ba.rowptr[g0], E1 = ba.rowptr[g1];
      int32_t e_my = active ? ba.rowptr[i] : 0;
      const int32_t e_end = active ? ba.rowptr[i + 1] : 0;
      double F[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      double sig = 0.0;
      for (int32_t base = E0; base < E1; base += CHUNK) {
        const int32_t lim = base + CHUNK < E1 ? base + CHUNK : E1;
        double co[PER][7];
        int64_t cc[PER];
        bool ok[PER];
#pragma unroll
        for (int j = 0; j < PER; ++j) {
          const int32_t e = base + tid + j * kPersistThreads;
          ok[j] = e < lim;
          cc[j] = 0;
          if (ok[j]) {
            if (PLANAR) {
              double p[4];
              load_prec<CT>(ba.rec, e, p);
              co[j][0] = p[0];
              co[j][1] = p[1];
              co[j][5] = p[2];
              co[j][6] = p[3];
              cc[j] = ba.rc[e];
            } else {
              load_rec<CT>(ba.rec, e, co[j], cc[j]);
            }
          }
        }
        double2 yg[PER];
        double gold[PER];
#pragma unroll
        for (int j = 0; j < PER; ++j) {
          if (ok[j]) {
            yg[j] = load_it<CT>(ba.it, ba.itg, cc[j]);
            if (!FOLD) gold[j] = ba.gold[base + tid + j * kPersistThreads];
          }
        }
#pragma unroll
        for (int j = 0; j < PER; ++j) {
          if (!ok[j]) continue;
          const double gc = (k == 0) ? yg[j].x : proj_step(yg[j].x, malpha, yg[j].y);
          const double dg = FOLD ? gc : gc - gold[j];
          const int qq = tid + j * kPersistThreads;
          if (PLANAR) {
            prod[0][qq] = dg * co[j][0];
            prod[1][qq] = dg * co[j][1];
            prod[2][qq] = dg * co[j][5];
            prod[3][qq] = gc * co[j][6];
          } else {
#pragma unroll
            for (int m = 0; m < 6; ++m) prod[m][qq] = dg * co[j][m];
            prod[6][qq] = gc * co[j][6];
          }
        }
        __syncthreads();
        for (; e_my < e_end && e_my < lim; ++e_my) {
          const int qq = e_my - base;
          if (PLANAR) {
            F[0] += prod[0][qq];
            F[1] += prod[1][qq];
            F[5] += prod[2][qq];
            sig += prod[3][qq];
          } else {
#pragma unroll
            for (int m = 0; m < 6; ++m) F[m] += prod[m][qq];
            sig += prod[6][qq];
          }
        }
        __syncthreads();
      }
      if (active) {
        if (FOLD) unfold(ba.fold, i, F, PLANAR);
        const double li = ba.len[i];
        double mt, mr;
        mobility(li, D.xi, mt, mr);
        double u[6];
        body_velocity(ba.uext, i, mt, mr, F, u);
        const double imp = det_exp(sig * D.mlam);
        const double ld = (li * imp) * D.itau;
        if (PLANAR) {
          double2* w = reinterpret_cast<double2*>(ba.ubuf + 4 * i);
          w[0] = make_double2(u[0], u[1]);
          w[1] = make_double2(u[5], ld - ba.ldot_prev[i]);
        } else {
          double2* w = reinterpret_cast<double2*>(ba.ubuf + 8 * i);
          w[0] = make_double2(u[0], u[1]);
          w[1] = make_double2(u[2], u[3]);
          w[2] = make_double2(u[4], u[5]);
          w[3] = make_double2(ld - ba.ldot_prev[i], 0.0);
        }
      }
    }
    __syncthreads();
    // ---- constraint apply (k_constr): logical block b is taken by group b % 4
    const bool save_prev = k > 0 && k % D.check == 0;
    for (int64_t b = grp; b < G; b += kPersistGroups) {
      double v0 = 0.0, v1 = 0.0, v2 = 0.0, vr = 0.0;
      for (int64_t j = 0; j < K; ++j) {
        const int64_t c = D.coff + (b * K + j) * PCC_BLOCK + lt;
        if (c >= cend) break;
        const int2 ij = ca.IJ[c];
        const int32_t I = ij.x, J = ij.y;
        const double2 yg = load_it<CT>(ca.it, ca.itg, c);
        const double phi = load_phi<CT>(ca.phi, c);
        const Row<CT, PLANAR> row = load_row<CT, PLANAR>(ca.coef, c);
        const double2* UI = reinterpret_cast<const double2*>(ca.ubuf + UW * (int64_t)I);
        const double2* UJ = reinterpret_cast<const double2*>(ca.ubuf + UW * (int64_t)J);
        double2 ui[UW / 2], uj[UW / 2];
#pragma unroll
        for (int m = 0; m < UW / 2; ++m) {
          ui[m] = UI[m];
          uj[m] = UJ[m];
        }
        const double gc = (k == 0) ? yg.x : proj_step(yg.x, malpha, yg.y);
        double pd = 0.0;
        double pg = 0.0;
        const CT* co = row.v;
        if (PLANAR) {
          pd += (double)co[0] * ui[0].x;
          pd += (double)co[1] * ui[0].y;
          pd += (double)co[2] * ui[1].x;
          pd += (-(double)co[0]) * uj[0].x;
          pd += (-(double)co[1]) * uj[0].y;
          pd += (double)co[3] * uj[1].x;
          pg += (double)co[4] * ui[1].y;
          pg += (double)co[5] * uj[1].y;
        } else {
          pd += (double)co[0] * ui[0].x;
          pd += (double)co[1] * ui[0].y;
          pd += (double)co[2] * ui[1].x;
          pd += (double)co[3] * ui[1].y;
          pd += (double)co[4] * ui[2].x;
          pd += (double)co[5] * ui[2].y;
          pd += (-(double)co[0]) * uj[0].x;
          pd += (-(double)co[1]) * uj[0].y;
          pd += (-(double)co[2]) * uj[1].x;
          pd += (double)co[6] * uj[1].y;
          pd += (double)co[7] * uj[2].x;
          pd += (double)co[8] * uj[2].y;
          pg += (double)co[9] * ui[3].x;
          pg += (double)co[10] * uj[3].x;
        }
        const double g = stored_g<CT>(phi + (dt * pd + (-dt) * pg));
        if (save_prev) ca.gprev[c] = yg.x;
        store_it<CT>(ca.it, ca.itg, c, gc, g);
        if (k > 0) {
          const double dgam = gc - yg.x;
          const double dphi = g - yg.y;
          v0 += dgam * dgam;
          v1 += dgam * dphi;
          v2 += dphi * dphi;
        }
        const double pj = gc > 0 ? g : (g < 0.0 ? g : 0.0);
        const double ar = fabs(pj);
        vr = vr < ar ? ar : vr;
      }
      const double w0 = wave_sum(v0), w1 = wave_sum(v1), w2 = wave_sum(v2);
      const double wm = wave_max(vr);
      if (lane == 0) {
        red[b][0][wv] = w0;
        red[b][1][wv] = w1;
        red[b][2][wv] = w2;
        red[b][3][wv] = wm;
      }
    }
    __syncthreads();
    // ---- finish (k_finish): block partials, then the strided reduction
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, am = 0.0;
    if (grp == 0) {
      for (int64_t i = lt; i < G; i += PCC_BLOCK) {
        const double(*r)[4] = red[i];
        const double p0 = ((r[0][0] + r[0][1]) + r[0][2]) + r[0][3];
        const double p1 = ((r[1][0] + r[1][1]) + r[1][2]) + r[1][3];
        const double p2 = ((r[2][0] + r[2][1]) + r[2][2]) + r[2][3];
        double pm = r[3][0];
        for (int m = 1; m < 4; ++m) pm = pm < r[3][m] ? r[3][m] : pm;
        a0 += p0;
        a1 += p1;
        a2 += p2;
        am = am < pm ? pm : am;
      }
    }
    const double f0 = wave_sum(a0), f1 = wave_sum(a1), f2 = wave_sum(a2);
    const double fm = wave_max(am);
    if (grp == 0 && lane == 0) {
      fin[0][wv] = f0;
      fin[1][wv] = f1;
      fin[2][wv] = f2;
      fin[3][wv] = fm;
    }
    __syncthreads();
    if (tid == 0) {
      const double S0 = ((fin[0][0] + fin[0][1]) + fin[0][2]) + fin[0][3];
      const double S1 = ((fin[1][0] + fin[1][1]) + fin[1][2]) + fin[1][3];
      const double S2 = ((fin[2][0] + fin[2][1]) + fin[2][2]) + fin[2][3];
      double R = fin[3][0];
      for (int m = 1; m < 4; ++m) R = R < fin[3][m] ? fin[3][m] : R;
      if (k == 0) {  // initial gradient: res = residual(g, Gamma), alpha = 1/res
        st->res = R;
        st->alpha = 1.0 / R;
        st->iter = 1;
        if (1 == D.maxit) st->failed = 1;
      } else {
        bool stop = false;
        if (k % D.check == 0) {
          st->res = R;
          if (R <= D.tol) {
            st->converged = 1;
            stop = true;
          }
        }
        if (!stop) {
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
          if (k + 1 == D.maxit) st->failed = 1;
        }
      }
      s_alpha = st->alpha;
      s_iter = st->iter;
      s_done = (st->converged | st->failed) ? 1 : 0;
    }
    __syncthreads();
  }
}

// dC = G(q) U (Physics.cpp:220-265), Euler step and renormalisation
// (Particle.cpp:37-57,110-120), write-back of ldot/impedance/stress
// (ParticleManager.cpp:89-130).
__global__ void k_integrate(int64_t N, BodyArrays b, const int32_t* __restrict__ binst,
                            const InstDev* __restrict__ idev, const double* __restrict__ U,
                            const double* __restrict__ ld, double* __restrict__ ldot_field) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const InstDev& D = idev[binst[i]];
  if (!(D.flags & kInstInteg)) return;
  const double dt = D.dt;
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

// updateConstraintsFromSolution (Hard.cpp:49-89): gamma <- lagging iterate
// (saved at the converging iteration), phi <- linearised gap of the last
// evaluated iterate; read at the constraint's solver position.
template <class CT>
__global__ void k_update_constraints(int64_t C, const int32_t* __restrict__ Sinst,
                                     const InstDev* __restrict__ idev,
                                     const int32_t* __restrict__ spos,
                                     const double* __restrict__ it, const float* __restrict__ itg,
                                     const double* __restrict__ gprev,
                                     double* __restrict__ Sgam, double* __restrict__ Sphi) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  if (!(idev[Sinst[c]].flags & kInstInteg)) return;
  const int32_t p = spos[c];
  Sgam[c] = gprev[p];
  Sphi[c] = load_it<CT>(it, itg, p).y;
}

// Per colony, max over its constraints of max(0, -phi) (Hard.cpp:125-130).
// Max is order-free; non-negative doubles order like their bit patterns, so
// an integer atomic max per colony is exact and deterministic.  The set is
// colony-major: a block whose constraints share one colony reduces first.
__global__ void k_maxov(int64_t C, const int32_t* __restrict__ Sinst,
                        const double* __restrict__ phi, unsigned long long* __restrict__ bits) {
  __shared__ double lds[4];
  const int64_t c0 = (int64_t)blockIdx.x * blockDim.x;
  const int64_t c = c0 + threadIdx.x;
  double v = 0.0;
  if (c < C) {
    const double p = phi[c];
    v = p < 0 ? -p : 0.0;
  }
  // one block maximum when every constraint of the block belongs to one
  // colony (the set need not be in colony-major order during a step)
  const int32_t first = Sinst[c0];
  const int32_t mine = c < C ? Sinst[c] : first;
  if (__syncthreads_and(mine == first)) {
    const double r = block_max(v, lds);
    if (threadIdx.x == 0) atomicMax(&bits[first], (unsigned long long)__double_as_longlong(r));
  } else if (c < C) {
    atomicMax(&bits[mine], (unsigned long long)__double_as_longlong(v));
  }
}

// grow (ParticleManager.cpp:132-137 -> Particle.cpp:59-65)
__global__ void k_grow(int64_t N, const int32_t* __restrict__ binst,
                       const InstDev* __restrict__ idev, double* __restrict__ len,
                       const double* __restrict__ l0, const double* __restrict__ ldot) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const InstDev& D = idev[binst[i]];
  if (!(D.flags & kInstGrow)) return;
  const double dt = D.dt;
  double l = len[i] + dt * ldot[i];
  if (l > 2 * l0[i]) l = 2 * l0[i];
  len[i] = l;
}

}  // namespace

// Max overlap of every colony's current constraint set (Hard.cpp:125-130):
// the reduction and its copy into pinned host memory are queued on the
// stream; max_overlap_read takes the values after a stream synchronisation.
void max_overlap_launch(pcc_ctx* ctx) {
  const int64_t B = (int64_t)ctx->inst.size(), C = ctx->S.C;
  if (ctx->ov_pin_cap < B) {
    PCC_HIP(hipStreamSynchronize(ctx->stream));
    DeviceLock lock(device_mutex());
    if (ctx->ov_pin) PCC_HIP(hipHostFree(ctx->ov_pin));
    ctx->ov_pin = nullptr;
    ctx->ov_pin_cap = 0;
    const int64_t cap = std::max<int64_t>(B, 16);
    PCC_HIP(hipHostMalloc((void**)&ctx->ov_pin, sizeof(unsigned long long) * cap,
                          hipHostMallocDefault));
    ctx->ov_pin_cap = cap;
  }
  if (B == 0) return;
  if (C == 0) {
    for (int64_t b = 0; b < B; ++b) ctx->ov_pin[b] = 0ull;  // +0.0
    return;
  }
  hipStream_t s = ctx->stream;
  ctx->maxbits.ensure(B);
  PCC_HIP(hipMemsetAsync(ctx->maxbits.p, 0, sizeof(unsigned long long) * B, s));
  k_maxov<<<nblk(C), 256, 0, s>>>(C, ctx->S.inst.p, ctx->S.phi.p, ctx->maxbits.p);
  PCC_HIP(hipMemcpyAsync(ctx->ov_pin, ctx->maxbits.p, sizeof(unsigned long long) * B,
                         hipMemcpyDeviceToHost, s));
}

void max_overlap_read(pcc_ctx* ctx, std::vector<double>& out) {
  const int64_t B = (int64_t)ctx->inst.size();
  out.assign(B, 0.0);
  for (int64_t b = 0; b < B && b < ctx->ov_pin_cap; ++b) memcpy(&out[b], &ctx->ov_pin[b], sizeof(double));
}

void max_overlap_inst(pcc_ctx* ctx, std::vector<double>& out) {
  max_overlap_launch(ctx);
  PCC_HIP(hipStreamSynchronize(ctx->stream));
  max_overlap_read(ctx, out);
}

// The planarity of the body state, from the last step when it kept the
// state in the plane (planar kernels, no out-of-plane Brownian turn, device
// division), otherwise checked on the device.
bool state_planar(pcc_ctx* ctx) {
  if (!ctx->flat_known) {
    ctx->flat_value = bodies_planar(ctx);
    ctx->flat_known = true;
  }
  return ctx->flat_value;
}

// 1 if every body lies in the z = 0 plane with an in-plane orientation.
bool bodies_planar(pcc_ctx* ctx) {
  const int64_t N = ctx->N;
  if (N == 0) return true;
  ctx->flag.ensure(1);
  PCC_HIP(hipMemsetAsync(ctx->flag.p, 0, sizeof(int32_t), ctx->stream));
  k_planar_check<<<nblk(N), 256, 0, ctx->stream>>>(N, body_arrays(ctx->body.p, N), ctx->flag.p);
  int32_t h = 0;
  PCC_HIP(hipMemcpyAsync(&h, ctx->flag.p, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  PCC_HIP(hipStreamSynchronize(ctx->stream));
  return h == 0;
}

// The iterate records hold g in place (FP64 pairs or packed mixed records):
// no separate array.
static float* iter_g(pcc_ctx*) { return nullptr; }

// The mixed body kernels read 3 or 6 doubles of D^T gamma_old per body.
static void check_fold(pcc_ctx* ctx) {
  if (ctx->params.precision != PCC_PREC_MIXED) return;
  if (ctx->N > 0 && (ctx->fold.p == nullptr || ctx->fold.cap < 6 * ctx->N))
    throw StatusError{PCC_E_DEVICE, "internal: gamma_old sums missing for the mixed kernels"};
}

// Body-incidence CSR over the whole constraint set, each body's incidences in
// ascending constraint order, plus the per-incidence coefficient records (3D
// or planar, FP64 and FP32), gamma_old per incidence and, planar, the compact
// constraint rows.
void build_incidence(pcc_ctx* ctx, bool records) {
  const int64_t N = ctx->N, C = ctx->S.C;
  hipStream_t s = ctx->stream;
  KTimer tm(ctx, PCC_K_ASSEMBLE, 0.0);
  ctx->rowptr.ensure(N + 1);
  ctx->inccur.ensure(N + 1);
  ctx->inc.ensure(2 * C + 1);
  if (!ctx->inccur_zero.known(ctx->inccur.p, N + 1))
    PCC_HIP(hipMemsetAsync(ctx->inccur.p, 0, sizeof(int32_t) * (size_t)(N + 1), s));
  ctx->inccur_zero.clear();
  if (C > 0) k_inc_count<<<nblk(C), 256, 0, s>>>(C, ctx->S.I.p, ctx->S.J.p, ctx->inccur.p);
  exclusive_scan_i32(ctx, ctx->inccur.p, ctx->rowptr.p, N + 1);
  if (C == 0) {
    ctx->inccur_zero.mark(ctx->inccur.p, N + 1);
    return;
  }
  k_inc_scatter<<<nblk(C), 256, 0, s>>>(C, ctx->S.I.p, ctx->S.J.p, ctx->rowptr.p, ctx->inccur.p,
                                        ctx->inc.p);
  ctx->inccur_zero.mark(ctx->inccur.p, N + 1);  // the scatter consumed the counts
  seg_sort_i32(ctx, N, ctx->rowptr.p, ctx->inc.p);
  if (!records) return;
  // solver order: stable radix sort of the constraints by I
  ctx->sord.ensure(C + 1);
  ctx->spos.ensure(C + 1);
  ctx->skey.ensure(C + 1);
  ctx->sval.ensure(C + 1);
  PCC_HIP(hipMemcpyAsync(ctx->skey.p, ctx->S.I.p, sizeof(int32_t) * C, hipMemcpyDeviceToDevice, s));
  k_iota<<<nblk(C), 256, 0, s>>>(C, ctx->sval.p);
  int bits = 1;
  while (((int64_t)1 << bits) < N) ++bits;
  ctx->skey2.ensure(C + 1);
  size_t bytes = 0;
  PCC_HIP(rocprim::radix_sort_pairs(nullptr, bytes, ctx->skey.p, ctx->skey2.p, ctx->sval.p,
                                    ctx->sord.p, (size_t)C, 0, bits, s));
  ctx->sort_tmp.ensure((int64_t)bytes + 16);
  PCC_HIP(rocprim::radix_sort_pairs(ctx->sort_tmp.p, bytes, ctx->skey.p, ctx->skey2.p,
                                    ctx->sval.p, ctx->sord.p, (size_t)C, 0, bits, s));
  k_invert<<<nblk(C), 256, 0, s>>>(C, ctx->sord.p, ctx->spos.p);
  const int W = ctx->planar ? 6 : 12;
  ctx->sIJ.ensure(C + 1);
  ctx->sphi.ensure(C + 1);
  ctx->iter.ensure(2 * (C + 1));  // (gamma, g): 16-byte FP64 pairs or 12-byte mixed records
  ctx->gprev.ensure(C + 1);
  ConstraintSet& S = ctx->S;
  const bool mixed = ctx->params.precision == PCC_PREC_MIXED;
  // rows and incidence records exist in the solve's precision only
  if (mixed)
    ctx->prow32.ensure(W * (C + 1));
  else
    ctx->prow.ensure(W * (C + 1));
  float* itg = iter_g(ctx);
#define PCC_SOLVER_ROWS(CT, PL)                                                              \
  k_solver_rows<CT, PL><<<nblk(C), 256, 0, s>>>(C, ctx->sord.p, S.I.p, S.J.p, S.phi.p,      \
                                                S.gamma.p, S.coef64.p, ctx->sIJ.p, \
                                                ctx->sphi.p, ctx->iter.p, itg, ctx->prow.p,  \
                                                ctx->prow32.p)
  if (ctx->planar) {
    if (mixed)
      PCC_SOLVER_ROWS(float, true);
    else
      PCC_SOLVER_ROWS(double, true);
  } else {
    if (mixed)
      PCC_SOLVER_ROWS(float, false);
    else
      PCC_SOLVER_ROWS(double, false);
  }
#undef PCC_SOLVER_ROWS
  const int RW = ctx->planar ? 4 : 8;
  if (mixed)
    ctx->increc32.ensure(RW * (2 * C + 1));
  else
    ctx->increc.ensure(RW * (2 * C + 1));
  ctx->incidx.ensure(2 * C + 1);
  ctx->gold_e.ensure(2 * C + 1);
#define PCC_INC_RECORDS(PL, MX)                                                             \
  k_inc_records<PL, MX><<<nblk(2 * C), 256, 0, s>>>(2 * C, ctx->inc.p, S.coef64.p, ctx->spos.p, \
                                                    S.gamma.p, ctx->increc.p, ctx->increc32.p,  \
                                                    ctx->incidx.p, ctx->gold_e.p)
  if (ctx->planar) {
    if (mixed)
      PCC_INC_RECORDS(true, true);
    else
      PCC_INC_RECORDS(true, false);
  } else {
    if (mixed)
      PCC_INC_RECORDS(false, true);
    else
      PCC_INC_RECORDS(false, false);
  }
#undef PCC_INC_RECORDS
}

// Logical block layout of the BBPGD kernels over the active colonies (with
// `keep`, only those it marks): colony b gets ceil(N_b / 256) body blocks
// and G_b = min(Q_b, 2048) blocks of K_b = ceil(Q_b / G_b) constraint tiles
// (Q_b = ceil(C_b / 256); a colony with no constraints still gets one
// block, as a lone colony does).
static void solver_layout(pcc_ctx* ctx, const std::vector<char>* keep = nullptr) {
  std::vector<int32_t> bm, cm;
  int64_t lb = 0, bb = 0;
  for (size_t q = 0; q < ctx->inst.size(); ++q) {
    Instance& I = ctx->inst[q];
    if (!(I.flags & kInstActive) || (keep && !(*keep)[q])) {
      I.G = 1;
      I.K = 0;
      I.lb = 0;
      I.bb = 0;
      I.nbb = 0;
      continue;
    }
    const int64_t Q = (I.C + PCC_BLOCK - 1) / PCC_BLOCK;
    I.G = Q < kBbGridMax ? (Q > 0 ? Q : 1) : kBbGridMax;
    I.K = (Q + I.G - 1) / I.G;
    I.lb = lb;
    lb += I.G;
    cm.insert(cm.end(), (size_t)I.G, (int32_t)q);
    I.nbb = (I.N + PCC_BLOCK - 1) / PCC_BLOCK;
    I.bb = bb;
    bb += I.nbb;
    bm.insert(bm.end(), (size_t)I.nbb, (int32_t)q);
  }
  ctx->ncb_total = lb;
  ctx->nbb_total = bb;
  hipStream_t s = ctx->stream;
  ctx->cmap.ensure(lb + 1);
  ctx->bmap.ensure(bb + 1);
  if (lb > 0)
    PCC_HIP(hipMemcpyAsync(ctx->cmap.p, cm.data(), sizeof(int32_t) * lb, hipMemcpyHostToDevice, s));
  if (bb > 0)
    PCC_HIP(hipMemcpyAsync(ctx->bmap.p, bm.data(), sizeof(int32_t) * bb, hipMemcpyHostToDevice, s));
  upload_instances(ctx);  // synchronises: the host vectors may go
}

static BodyArgs body_args(pcc_ctx* ctx) {
  BodyArgs a;
  const int64_t N = ctx->N;
  const bool mixed = ctx->params.precision == PCC_PREC_MIXED;
  BodyArrays b = body_arrays(ctx->body.p, N);
  a.N = N;
  a.nbb = ctx->nbb_total;
  a.gp = std::max<int64_t>(pad8(ctx->nbb_total), kXcd);
  a.bmap = ctx->bmap.p;
  a.binst = ctx->binst.p;
  a.idev = ctx->idev.p;
  a.rowptr = ctx->rowptr.p;
  a.rec = mixed ? (const void*)ctx->increc32.p : (const void*)ctx->increc.p;
  a.rc = ctx->incidx.p;
  a.gold = ctx->gold_e.p;
  a.fold = ctx->fold.p;
  a.it = ctx->iter.p;
  a.itg = iter_g(ctx);
  a.len = b.len;
  a.ldot_prev = b.ldot;
  a.st = ctx->state.p;
  a.ubuf = ctx->ubuf.p;
  a.uext = ctx->brownian ? ctx->uext.p : nullptr;
  return a;
}

static ConstrArgs constr_args(pcc_ctx* ctx) {
  ConstrArgs a;
  ConstraintSet& S = ctx->S;
  const bool mixed = ctx->params.precision == PCC_PREC_MIXED;
  a.ncb = ctx->ncb_total;
  a.gp = std::max<int64_t>(pad8(ctx->ncb_total), kXcd);
  a.Gtot = std::max<int64_t>(ctx->ncb_total, 1);
  a.cmap = ctx->cmap.p;
  a.idev = ctx->idev.p;
  (void)S;
  a.IJ = ctx->sIJ.p;
  a.coef = mixed ? (const void*)ctx->prow32.p : (const void*)ctx->prow.p;
  a.phi = ctx->sphi.p;
  a.it = ctx->iter.p;
  a.itg = iter_g(ctx);
  a.gprev = ctx->gprev.p;
  a.ubuf = ctx->ubuf.p;
  a.st = ctx->state.p;
  a.partials = ctx->partials.p;
  a.trace = ctx->trace_on ? 1 : 0;
  a.trace_cap = ctx->trace_on ? ctx->trace_cap : 0;
  a.trace_buf = ctx->trace_on ? ctx->trace.p : nullptr;
  a.trace_n = ctx->trace_on ? ctx->trace_n.p : nullptr;
  a.epart = ctx->trace_on ? ctx->epart.p : nullptr;
  a.nbb = ctx->nbb_total;
  return a;
}

// Algorithmic bytes of one launch of each BBPGD kernel: every array the
// kernel must touch, counted once (a constraint's iterate record counts once
// per incidence in the body kernel, a body's velocity record once).
double body_bytes(int64_t N, int64_t C, bool mixed, bool planar) {
  const double sc = mixed ? 4 : 8;
  // per incidence: record (8 or 4 coefficient slots [+ 4-byte index]) +
  // the iterate record's (gamma, g) (FP64 16 bytes; mixed 8 + 4) +
  // gamma_old (FP64 only)
  // per body: rowptr, length, ldot_prev, velocity record out, and in mixed
  // precision D^T gamma_old (6 or 3 doubles); the mobility is evaluated from
  // the length
  const double rec = planar ? 4 * sc + 4 : 8 * sc;
  const double ub = planar ? 32 : 64;
  const double yg = mixed ? 12 : 16;
  const double gold_e = mixed ? 0 : 8;
  const double fold = mixed ? (planar ? 24 : 48) : 0;
  return 2.0 * C * (rec + yg + gold_e) + N * (4 + 16 + ub + fold);
}
double constr_bytes(int64_t N, int64_t C, bool mixed, bool planar) {
  const double sc = mixed ? 4 : 8;
  // per constraint: I, J; operator row (12 or 6 slots); iterate record in
  // and out ((gamma, g): FP64 16 bytes, mixed 8 + 4) and Phi (8, mixed 4).
  // Per body: velocity record read once.
  const double row = planar ? 6 * sc : 12 * sc;
  const double ub = planar ? 32 : 64;
  const double yg = mixed ? 12 : 16;
  return C * (8 + row + sc + 2 * yg) + N * ub;
}

static void launch_constr(bool planar, bool mixed, unsigned grid, hipStream_t s,
                          const ConstrArgs& ca) {
  if (planar) {
    if (mixed)
      k_constr<float, true><<<grid, 256, 0, s>>>(ca);
    else
      k_constr<double, true><<<grid, 256, 0, s>>>(ca);
  } else {
    if (mixed)
      k_constr<float, false><<<grid, 256, 0, s>>>(ca);
    else
      k_constr<double, false><<<grid, 256, 0, s>>>(ca);
  }
}

// Body-kernel variants the autotuner chooses from: staging size (incidences
// per pass: planar 512 / 1024 / 2048, 3D 256 / 512 / 1024) times threads per
// block (variants 0-2: 256, one body each; 3-5: 128, two bodies each), and
// variants 6-8: 256 threads with the next chunk's records prefetched, staging
// 256 / 512 / 1024.
template <class CT, bool PL, int C0>
static void launch_body_t(int variant, unsigned grid, hipStream_t s, const BodyArgs& ba) {
  switch (variant) {
    case 0: k_body_coop<CT, PL, C0, 256, false><<<grid, 256, 0, s>>>(ba, FinalOut{}); break;
    case 1: k_body_coop<CT, PL, 2 * C0, 256, false><<<grid, 256, 0, s>>>(ba, FinalOut{}); break;
    case 2: k_body_coop<CT, PL, 4 * C0, 256, false><<<grid, 256, 0, s>>>(ba, FinalOut{}); break;
    case 3: k_body_coop<CT, PL, C0, 128, false><<<grid, 128, 0, s>>>(ba, FinalOut{}); break;
    case 4: k_body_coop<CT, PL, 2 * C0, 128, false><<<grid, 128, 0, s>>>(ba, FinalOut{}); break;
    case 5: k_body_coop<CT, PL, 4 * C0, 128, false><<<grid, 128, 0, s>>>(ba, FinalOut{}); break;
    case 6: k_body_coop<CT, PL, 256, 256, true><<<grid, 256, 0, s>>>(ba, FinalOut{}); break;
    case 7: k_body_coop<CT, PL, 512, 256, true><<<grid, 256, 0, s>>>(ba, FinalOut{}); break;
    default: k_body_coop<CT, PL, 1024, 256, true><<<grid, 256, 0, s>>>(ba, FinalOut{}); break;
  }
}

static void launch_body(int variant, bool planar, bool mixed, unsigned grid, hipStream_t s,
                        const BodyArgs& ba) {
  if (planar) {
    if (mixed)
      launch_body_t<float, true, 512>(variant, grid, s, ba);
    else
      launch_body_t<double, true, 512>(variant, grid, s, ba);
  } else {
    if (mixed)
      launch_body_t<float, false, 256>(variant, grid, s, ba);
    else
      launch_body_t<double, false, 256>(variant, grid, s, ba);
  }
}

int body_chunk_of(int variant, bool planar) {
  return variant >= 6 ? 256 << (variant - 6) : (planar ? 512 : 256) << (variant % 3);
}
int body_threads_of(int variant) { return variant < 3 || variant >= 6 ? 256 : 128; }

// Process-wide memory of the autotuner: the body-kernel variant chosen for a
// (device architecture, planar, precision, N bucket, C/N bucket) class of
// solves, so a new context (serving creates one per batch) reuses it instead
// of timing every variant inside its first step.  N buckets by powers of
// two of the bodies of the solve, C/N by halves of the mean incidences per
// body.
using TuneKey = std::tuple<std::string, int, int, int, int>;
static std::mutex& tune_mutex() {
  static std::mutex m;
  return m;
}
struct TuneEntry {
  int variant;
  float ms[PCC_BODY_VARIANTS];  // the timings it was chosen from
};
static std::map<TuneKey, TuneEntry>& tune_cache() {
  static std::map<TuneKey, TuneEntry> m;
  return m;
}
static TuneKey tune_key(pcc_ctx* ctx, int64_t N, int64_t C) {
  static std::mutex arch_m;
  static std::map<int, std::string> arch;
  std::string a;
  {
    std::lock_guard<std::mutex> g(arch_m);
    auto it = arch.find(ctx->device);
    if (it == arch.end()) {
      hipDeviceProp_t prop;
      std::string name = "?";
      if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess) name = prop.gcnArchName;
      it = arch.emplace(ctx->device, name).first;
    }
    a = it->second;
  }
  int nb = 0;
  while (((int64_t)1 << (nb + 1)) <= std::max<int64_t>(N, 1)) ++nb;
  const int cn = (int)std::min<int64_t>(4 * C / std::max<int64_t>(N, 1), 64);
  return TuneKey(a, ctx->planar ? 1 : 0, ctx->params.precision, nb, cn);
}

static void ensure_state_host(pcc_ctx* ctx, int64_t slots) {
  if (ctx->state_host_cap >= slots) return;
  DeviceLock lock(device_mutex());
  if (ctx->state_host) PCC_HIP(hipHostFree(ctx->state_host));
  ctx->state_host = nullptr;
  ctx->state_host_cap = 0;
  PCC_HIP(hipHostMalloc((void**)&ctx->state_host, sizeof(SolverState) * slots,
                        hipHostMallocDefault));
  ctx->state_host_cap = slots;
}

// Final solver state of every active colony (BBPGDResult, BBPGD.h:13-16);
// with `account`, the algorithmic bytes of the launches in which each
// colony's blocks worked: iterations 0..iter (the converging one included),
// or 0..iter-1 on failure.
static void collect_solve_results(pcc_ctx* ctx, bool account) {
  const int64_t B = (int64_t)ctx->inst.size();
  const bool mixed = ctx->params.precision == PCC_PREC_MIXED;
  hipStream_t s = ctx->stream;
  std::vector<SolverState> fin(B);
  PCC_HIP(hipMemcpyAsync(fin.data(), ctx->state.p, sizeof(SolverState) * B, hipMemcpyDeviceToHost, s));
  PCC_HIP(hipStreamSynchronize(s));
  for (int64_t q = 0; q < B; ++q) {
    Instance& I = ctx->inst[q];
    if (!(I.flags & kInstActive)) continue;
    I.solve_its = fin[q].iter;
    I.solve_res = fin[q].res;
    I.solve_failed = fin[q].failed != 0;
    if (!account) continue;
    const double worked = (double)(fin[q].iter + (I.solve_failed ? 0 : 1));
    add_timing_bytes(ctx, PCC_K_BODY_APPLY, worked * body_bytes(I.N, I.C, mixed, ctx->planar));
    add_timing_bytes(ctx, PCC_K_CONSTR_APPLY,
                     worked * constr_bytes(I.N, I.C, mixed, ctx->planar));
  }
}

// Runs BBPGD for every active colony at once on the current phase operators
// (BBPGD.cpp:32-127).  Each colony has its own step size, iteration count and
// convergence flag on the device; a converged colony's blocks exit at once
// while the others go on.  On return every active colony's solve_its /
// solve_res / solve_failed are set (BBPGDResult, BBPGD.h:13-16).
void bbpgd_solve(pcc_ctx* ctx) {
  const int64_t N = ctx->N, C = ctx->S.C;
  const int64_t B = (int64_t)ctx->inst.size();
  const bool mixed = ctx->params.precision == PCC_PREC_MIXED;
  const bool planar = ctx->planar;
  // k_constr addresses the solver-order arrays by 32-bit byte offsets
  const int64_t row_bytes = (planar ? 6 : 12) * (mixed ? 4 : 8);
  if (C * std::max<int64_t>(row_bytes, 16) >= ((int64_t)1 << 32) || N * 64 >= ((int64_t)1 << 32))
    throw StatusError{PCC_E_INVALID_ARGUMENT,
                      "constraint set too large for one context (operator rows beyond 4 GiB)"};
  hipStream_t s = ctx->stream;
  solver_layout(ctx);
  ctx->ubuf.ensure(8 * N + 8);
  ctx->partials.ensure((ctx->trace_on ? 5 : 4) * std::max<int64_t>(ctx->ncb_total, 1) + 4);
  if (ctx->trace_on) ctx->epart.ensure(2 * ctx->nbb_total + 2);
  ctx->state.ensure(B);
  ensure_state_host(ctx, 2 * B);
  {
    std::vector<SolverState> init(B);
    memset(init.data(), 0, sizeof(SolverState) * B);
    for (int64_t q = 0; q < B; ++q)
      if (!(ctx->inst[q].flags & kInstActive)) init[q].skip = 1;
    PCC_HIP(hipMemcpyAsync(ctx->state.p, init.data(), sizeof(SolverState) * B,
                           hipMemcpyHostToDevice, s));
    PCC_HIP(hipStreamSynchronize(s));
  }
  if (mixed && N > 0) ctx->fold.ensure(6 * N + 6);
  BodyArgs ba = body_args(ctx);
  ConstrArgs ca = constr_args(ctx);
  unsigned gb = (unsigned)ba.gp, gc = (unsigned)ca.gp;
  if (mixed && N > 0) {
    // D^T gamma_old of every body of the solve, once per phase
    FinalOut fo{};
    fo.F = ctx->fold.p;
    if (planar)
      k_body_coop<float, true, 512, 256, true, 2><<<gb, 256, 0, s>>>(ba, fo);
    else
      k_body_coop<float, false, 256, 256, true, 2><<<gb, 256, 0, s>>>(ba, fo);
  }
  check_fold(ctx);
  // Persistent path: when every active colony fits one workgroup, one launch
  // runs all their solves (a block per colony) to the end.
  {
    std::vector<int32_t> plist;
    bool persist = ctx->use_persist && !ctx->trace_on && !ctx->timing.enabled;
    for (int64_t q = 0; q < B && persist; ++q) {
      const Instance& I = ctx->inst[q];
      if (!(I.flags & kInstActive)) continue;
      if (I.C > ctx->persist_max_c || I.N > ctx->persist_max_n || I.G > kPersistMaxTiles)
        persist = false;
      else
        plist.push_back((int32_t)q);
    }
    if (persist && !plist.empty()) {
      const int64_t np = (int64_t)plist.size();
      ctx->plist.ensure(np);
      PCC_HIP(hipMemcpyAsync(ctx->plist.p, plist.data(), sizeof(int32_t) * np,
                             hipMemcpyHostToDevice, s));
      const unsigned g = (unsigned)np;
      if (planar) {
        if (mixed)
          k_bbpgd_persist<float, true><<<g, kPersistThreads, 0, s>>>(ba, ca, ctx->plist.p);
        else
          k_bbpgd_persist<double, true><<<g, kPersistThreads, 0, s>>>(ba, ca, ctx->plist.p);
      } else {
        if (mixed)
          k_bbpgd_persist<float, false><<<g, kPersistThreads, 0, s>>>(ba, ca, ctx->plist.p);
        else
          k_bbpgd_persist<double, false><<<g, kPersistThreads, 0, s>>>(ba, ca, ctx->plist.p);
      }
      collect_solve_results(ctx, false);
      return;
    }
  }
  auto body = [&](int variant) { launch_body(variant, planar, mixed, gb, s, ba); };
  auto one_iteration = [&]() {
    {
      KTimer t(ctx, PCC_K_BODY_APPLY, 0.0);
      body(ctx->body_variant);
    }
    {
      KTimer t(ctx, PCC_K_CONSTR_APPLY, 0.0);
      launch_constr(planar, mixed, gc, s, ca);
    }
    if (ctx->trace_on) {
      double* ep = ctx->epart.p;
      if (planar) {
        if (mixed)
          k_energy_body<float, true><<<gb, 256, 0, s>>>(ba, ep);
        else
          k_energy_body<double, true><<<gb, 256, 0, s>>>(ba, ep);
      } else {
        if (mixed)
          k_energy_body<float, false><<<gb, 256, 0, s>>>(ba, ep);
        else
          k_energy_body<double, false><<<gb, 256, 0, s>>>(ba, ep);
      }
    }
    k_finish<<<(unsigned)B, 256, 0, s>>>(ca);
  };
  one_iteration();  // k = 0
  // Autotuner: the body-kernel variant for this solve's class (N bucket,
  // C/N bucket), from the process-wide memory or timed here on the solve's
  // own operators.  A context re-tunes when a solve falls in another class
  // (the phases of a ReLCP step grow C/N, and larger staging pays off as it
  // does).  The kernel only rewrites the velocity records of the iteration
  // about to run, identically in every variant, so timing it changes no
  // state.
  if (ctx->autotune && C >= kTuneMinConstraints) {
    int64_t na = 0;
    for (const auto& I : ctx->inst)
      if (I.flags & kInstActive) na += I.N;
    const TuneKey key = tune_key(ctx, na, C);
    if (!ctx->body_tuned || key != ctx->tuned_key) {
      bool found = false;
      {
        std::lock_guard<std::mutex> g(tune_mutex());
        auto it = tune_cache().find(key);
        if (it != tune_cache().end()) {
          ctx->body_variant = it->second.variant;
          memcpy(ctx->body_variant_ms, it->second.ms, sizeof(ctx->body_variant_ms));
          found = true;
        }
      }
      if (found) {
        ctx->body_tuned = true;
        ctx->body_from_cache = true;
        ctx->tuned_key = key;
      } else if (!ctx->timing.enabled) {
        float best = 0.f;
        hipEvent_t e0, e1;
        {
          DeviceLock lock(device_mutex());
          PCC_HIP(hipEventCreate(&e0));
          PCC_HIP(hipEventCreate(&e1));
        }
        constexpr int kReps = 8;
        for (int v = 0; v < kBodyVariants; ++v) {
          body(v);  // warm
          PCC_HIP(hipEventRecord(e0, s));
          for (int r = 0; r < kReps; ++r) body(v);
          PCC_HIP(hipEventRecord(e1, s));
          PCC_HIP(hipEventSynchronize(e1));
          float ms = 0.f;
          PCC_HIP(hipEventElapsedTime(&ms, e0, e1));
          ctx->body_variant_ms[v] = ms / kReps;
          if (v == 0 || ms < best) {
            best = ms;
            ctx->body_variant = v;
          }
        }
        {
          TuneEntry en;
          en.variant = ctx->body_variant;
          memcpy(en.ms, ctx->body_variant_ms, sizeof(en.ms));
          std::lock_guard<std::mutex> g(tune_mutex());
          tune_cache()[key] = en;
        }
        {
          DeviceLock lock(device_mutex());
          (void)hipEventDestroy(e0);
          (void)hipEventDestroy(e1);
        }
        ctx->body_tuned = true;
        ctx->body_from_cache = false;
        ctx->tuned_key = key;
      }
    }
  }
  // The chunk of `chunk` iterations is launched as one captured HIP graph
  // (kernel nodes only; the state read-back below stays outside it), unless
  // per-kernel event timing is on.  A solve still running after
  // kChunkGrowAfter chunks switches to chunks twice as long: fewer graph
  // launches and flag read-backs for long solves (2 x 300k colonies: 738 ->
  // 731 ms per step), the short chunks' early stop for short ones.
  int chunk = ctx->chunk_iters;
  hipGraphExec_t exec = nullptr;
  auto capture = [&]() {
    DeviceLock lock(device_mutex());
    hipGraph_t graph = nullptr;
    PCC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int q = 0; q < chunk; ++q) one_iteration();
    PCC_HIP(hipStreamEndCapture(s, &graph));
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    PCC_HIP(e);
  };
  if (ctx->use_graphs && !ctx->timing.enabled) capture();
  // graphs replaced after a compaction (possibly still in flight): destroyed
  // after the final synchronisation
  std::vector<hipGraphExec_t> retired;
  struct ExecGuard {
    hipGraphExec_t& e;
    std::vector<hipGraphExec_t>& r;
    ~ExecGuard() {
      if (e || !r.empty()) {
        DeviceLock lock(device_mutex());
        if (e) (void)hipGraphExecDestroy(e);
        for (auto x : r) (void)hipGraphExecDestroy(x);
      }
    }
  } guard{exec, retired};
  bool compacted = false;
  SolverState* hs = ctx->state_host;
  hipEvent_t ev[2];
  {
    DeviceLock lock(device_mutex());
    PCC_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    PCC_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  }
  int64_t j = 0;
  bool done = false;
  while (!done) {
    if (j == kChunkGrowAfter) {
      chunk *= 2;
      if (exec) {
        retired.push_back(exec);
        exec = nullptr;
        capture();
      }
    }
    if (exec)
      PCC_HIP(hipGraphLaunch(exec, s));
    else
      for (int q = 0; q < chunk; ++q) one_iteration();
    PCC_HIP(hipMemcpyAsync(hs + (j & 1) * B, ctx->state.p, sizeof(SolverState) * B,
                           hipMemcpyDeviceToHost, s));
    PCC_HIP(hipEventRecord(ev[j & 1], s));
    if (j >= 1) {
      PCC_HIP(hipEventSynchronize(ev[(j - 1) & 1]));
      const SolverState* h = hs + ((j - 1) & 1) * B;
      done = true;
      int64_t live_bb = 0;
      for (int64_t q = 0; q < B; ++q)
        if (!(h[q].skip | h[q].converged | h[q].failed)) {
          done = false;
          live_bb += ctx->inst[q].nbb;
        }
      // Active-set compaction: once the colonies still iterating hold at most
      // compact_ratio (default one half) of the launches' body blocks, the
      // grids are rebuilt over them alone (the finished colonies' blocks
      // would only exit).  A colony's
      // logical blocks and sums are the same in any layout, so no result
      // changes; the maps are rewritten in stream order, after the chunk in
      // flight, which still uses the old ones.
      if (!done && ctx->compact && ctx->nbb_total >= ctx->compact_min_blocks &&
          (double)live_bb <= ctx->compact_ratio * (double)ctx->nbb_total) {
        std::vector<char> keep(B, 0);
        for (int64_t q = 0; q < B; ++q)
          keep[q] = !(h[q].skip | h[q].converged | h[q].failed);
        solver_layout(ctx, &keep);
        ba = body_args(ctx);
        ca = constr_args(ctx);
        gb = (unsigned)ba.gp;
        gc = (unsigned)ca.gp;
        if (exec) {
          retired.push_back(exec);
          exec = nullptr;
          capture();
        }
        compacted = true;
        ++ctx->compactions;
      }
    }
    ++j;
  }
  PCC_HIP(hipStreamSynchronize(s));
  if (compacted) solver_layout(ctx);  // every active colony's blocks, for the integration
  {
    DeviceLock lock(device_mutex());
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
  }
  collect_solve_results(ctx, true);
}

// Workspace of the converged iterate, move, write-back, constraint update for
// the colonies flagged kInstInteg.
void integrate_hard(pcc_ctx* ctx) {
  const int64_t N = ctx->N, C = ctx->S.C;
  const bool mixed = ctx->params.precision == PCC_PREC_MIXED;
  hipStream_t s = ctx->stream;
  BodyArrays b = body_arrays(ctx->body.p, N);
  KTimer t(ctx, PCC_K_INTEGRATE, 0.0);
  if (N > 0) {
    check_fold(ctx);
    BodyArgs ba = body_args(ctx);  // the solve's body blocks (solver_layout)
    const FinalOut fo = {ctx->F.p, ctx->U.p, ctx->sig.p, b.imp, ctx->ldot_new.p};
    const unsigned gb = (unsigned)ba.gp, g = nblk(N);
    if (ctx->planar) {
      if (mixed)
        k_body_coop<float, true, 512, 256, true, 1><<<gb, 256, 0, s>>>(ba, fo);
      else
        k_body_coop<double, true, 512, 256, true, 1><<<gb, 256, 0, s>>>(ba, fo);
    } else {
      if (mixed)
        k_body_coop<float, false, 256, 256, true, 1><<<gb, 256, 0, s>>>(ba, fo);
      else
        k_body_coop<double, false, 256, 256, true, 1><<<gb, 256, 0, s>>>(ba, fo);
    }
    k_integrate<<<g, 256, 0, s>>>(N, b, ctx->binst.p, ctx->idev.p, ctx->U.p, ctx->ldot_new.p,
                                  b.ldot);
  }
  if (C > 0) {
    const float* itg = iter_g(ctx);
    if (mixed)
      k_update_constraints<float><<<nblk(C), 256, 0, s>>>(C, ctx->S.inst.p, ctx->idev.p,
                                                          ctx->spos.p, ctx->iter.p, itg,
                                                          ctx->gprev.p, ctx->S.gamma.p,
                                                          ctx->S.phi.p);
    else
      k_update_constraints<double><<<nblk(C), 256, 0, s>>>(C, ctx->S.inst.p, ctx->idev.p,
                                                           ctx->spos.p, ctx->iter.p, itg,
                                                           ctx->gprev.p, ctx->S.gamma.p,
                                                           ctx->S.phi.p);
  }
}

// Brownian inputs of the current phase-0 solve or soft step: ctx->uext holds
// every body's draws (zeros outside colonies with T > 0) and ctx->brownian
// is set when some active colony has T > 0.  The instance records must be
// current (flags, colony ranges).
void brownian_inputs(pcc_ctx* ctx) {
  ctx->brownian = false;
  bool any = false;
  for (const auto& I : ctx->inst)
    if (I.p.temperature > 0.0 && (I.flags & kInstActive)) any = true;
  const int64_t N = ctx->N;
  if (!any || N == 0) return;
  ctx->uext.ensure(3 * N + 3);
  BodyArrays b = body_arrays(ctx->body.p, N);
  k_brownian<<<nblk(N), 256, 0, ctx->stream>>>(N, b.imp, ctx->binst.p, ctx->idev.p,
                                                ctx->uext.p);
  ctx->brownian = true;
}

// True if some colony's Brownian input turns bodies out of the plane.
bool brownian_offplane(const pcc_ctx* ctx) {
  for (const auto& I : ctx->inst)
    if (I.p.temperature > 0.0 && !I.p.monolayer) return true;
  return false;
}

// Growth of the colonies flagged kInstGrow.
void grow(pcc_ctx* ctx) {
  const int64_t N = ctx->N;
  if (N == 0) return;
  BodyArrays b = body_arrays(ctx->body.p, N);
  k_grow<<<nblk(N), 256, 0, ctx->stream>>>(N, ctx->binst.p, ctx->idev.p, b.len, b.l0, b.ldot);
}

}  // namespace pcc
