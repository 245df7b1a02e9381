// soft.hip — soft (Hertz) model step on the GPU.
//
// Replaces solveSoftPotential (Soft.cpp:21-114): per overlapping contact
// F = sqrt(R1 R2/(R1+R2)) kcc delta^1.5 along n (Soft.cpp:53-67), accumulated
// per body in ascending contact order (the order of the reference's
// VecSetValues(ADD_VALUES) loop, Soft.cpp:46-83) through the body-incidence
// CSR, i.e. a deterministic segmented sum with no float atomics; then
// U = M F, the same integration as the hard model and the growth law
// (Soft.cpp:97-111).  In mixed precision the pair forces and their sums run
// in FP32 and the integration in FP64.  Every colony of the context takes its
// own kcc, xi, dt, lambda and tau from its instance record.
#include "common.hpp"
#include "internal.hpp"

namespace pcc {

namespace {

inline unsigned nblk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// record (16 doubles): fi[3], ti[3], fj[3], tj[3], sI*F, sJ*F, active, pad
__global__ void k_hertz(int64_t C, int mixed, const int32_t* __restrict__ I,
                        const int32_t* __restrict__ J, const double* __restrict__ phi,
                        const double* __restrict__ n, const double* __restrict__ rI,
                        const double* __restrict__ rJ, const double* __restrict__ sI,
                        const double* __restrict__ sJ, const double* __restrict__ diam,
                        const int32_t* __restrict__ Sinst, const InstDev* __restrict__ idev,
                        double* __restrict__ rec, float* __restrict__ rec32) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  const double kcc = idev[Sinst[c]].kcc;
  const double overlap = -phi[c];
  double* r = rec + 16 * c;
  float* r32 = rec32 + 16 * c;
  if (overlap <= 0.00) {
    r[14] = 0.0;
    r32[14] = 0.f;
    return;
  }
  const double R1 = diam[I[c]] / 2.0;
  const double R2 = diam[J[c]] / 2.0;
  const double Reff = sqrt(R1 * R2 / (R1 + R2));
  if (!mixed) {
    const double Fel = Reff * kcc * pow(overlap, 1.5);
    double fi[3], fj[3], ti[3], tj[3], ri[3], rj[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      fi[k] = n[3 * c + k] * Fel;
      fj[k] = -fi[k];
      ri[k] = rI[3 * c + k];
      rj[k] = rJ[3 * c + k];
    }
    cross3(ri, fi, ti);
    cross3(rj, fj, tj);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      r[k] = fi[k];
      r[3 + k] = ti[k];
      r[6 + k] = fj[k];
      r[9 + k] = tj[k];
    }
    r[12] = sI[c] * Fel;
    r[13] = sJ[c] * Fel;
    r[14] = 1.0;
  } else {
    const float ov = (float)overlap;
    const float Fel = (float)Reff * (float)kcc * (ov * sqrtf(ov));
    float fi[3], ri[3], rj[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      fi[k] = (float)n[3 * c + k] * Fel;
      ri[k] = (float)rI[3 * c + k];
      rj[k] = (float)rJ[3 * c + k];
    }
    r32[0] = fi[0];
    r32[1] = fi[1];
    r32[2] = fi[2];
    r32[3] = ri[1] * fi[2] - ri[2] * fi[1];
    r32[4] = ri[2] * fi[0] - ri[0] * fi[2];
    r32[5] = ri[0] * fi[1] - ri[1] * fi[0];
    r32[6] = -fi[0];
    r32[7] = -fi[1];
    r32[8] = -fi[2];
    r32[9] = -(rj[1] * fi[2] - rj[2] * fi[1]);
    r32[10] = -(rj[2] * fi[0] - rj[0] * fi[2]);
    r32[11] = -(rj[0] * fi[1] - rj[1] * fi[0]);
    r32[12] = (float)sI[c] * Fel;
    r32[13] = (float)sJ[c] * Fel;
    r32[14] = 1.f;
  }
}

__global__ void k_soft_body(int64_t N, int mixed, BodyArrays b, const int32_t* __restrict__ rowptr,
                            const int32_t* __restrict__ inc, const double* __restrict__ rec,
                            const float* __restrict__ rec32, const int32_t* __restrict__ binst,
                            const InstDev* __restrict__ idev, const double* __restrict__ uext,
                            double* __restrict__ F_out, double* __restrict__ U_out,
                            double* __restrict__ sig_out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const InstDev& D = idev[binst[i]];
  const double xi = D.xi, dt = D.dt, mlam = D.mlam, itau = D.itau;
  double F[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double sig = 0.0;
  const int32_t e1 = rowptr[i + 1];
  if (!mixed) {
    for (int32_t e = rowptr[i]; e < e1; ++e) {
      const int32_t v = inc[e];
      const int64_t c = v >> 1;
      const int side = v & 1;
      const double* r = rec + 16 * c;
      if (r[14] == 0.0) continue;
#pragma unroll
      for (int k = 0; k < 6; ++k) F[k] += r[6 * side + k];
      sig += r[12 + side];
    }
  } else {
    float Ff[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float sf = 0.f;
    for (int32_t e = rowptr[i]; e < e1; ++e) {
      const int32_t v = inc[e];
      const int64_t c = v >> 1;
      const int side = v & 1;
      const float* r = rec32 + 16 * c;
      if (r[14] == 0.f) continue;
#pragma unroll
      for (int k = 0; k < 6; ++k) Ff[k] += r[6 * side + k];
      sf += r[12 + side];
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) F[k] = (double)Ff[k];
    sig = (double)sf;
  }
  // U = M F + U_ext (Soft.cpp:8-19), then the same move as the hard model
  const double l = b.len[i];
  const double mt = 1.0 / (l * xi);
  const double mr = 12.0 / (l * l * l * xi);
  double ue[3] = {0.0, 0.0, 0.0};  // U_ext: Brownian angular velocity (T > 0)
  if (uext) {
    ue[0] = uext[3 * i];
    ue[1] = uext[3 * i + 1];
    ue[2] = uext[3 * i + 2];
  }
  double u[6];
#pragma unroll
  for (int k = 0; k < 3; ++k) u[k] = (0.0 + mt * F[k]) + 0.0;
#pragma unroll
  for (int k = 3; k < 6; ++k) u[k] = (0.0 + mr * F[k]) + ue[k - 3];
  double q[4] = {b.qw[i], b.qx[i], b.qy[i], b.qz[i]};
  const double s = q[0], wx = q[1], wy = q[2], wz = q[3];
  const double xq[4][3] = {{-0.5 * wx, -0.5 * wy, -0.5 * wz},
                           {0.5 * s, -0.5 * wz, 0.5 * wy},
                           {0.5 * wz, 0.5 * s, -0.5 * wx},
                           {-0.5 * wy, 0.5 * wx, 0.5 * s}};
  double dC[7];
#pragma unroll
  for (int k = 0; k < 3; ++k) dC[k] = 0.0 + 1.0 * u[k];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    double sum = 0.0;
    sum += xq[r][0] * u[3];
    sum += xq[r][1] * u[4];
    sum += xq[r][2] * u[5];
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
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    F_out[6 * i + k] = F[k];
    U_out[6 * i + k] = u[k];
  }
  // impedance, growth rate and growth (Soft.cpp:101-111)
  const double imp = det_exp(sig * mlam);
  const double ld = (l * imp) * itau;
  sig_out[i] = sig;
  b.imp[i] = imp;
  b.ldot[i] = ld;
  double ln = l + dt * ld;
  if (ln > 2 * b.l0[i]) ln = 2 * b.l0[i];
  b.len[i] = ln;
}

}  // namespace

// Algorithmic bytes per launch.  k_hertz, per contact: I, J, phi, n, rI, rJ,
// sI, sJ, the two diameters and the colony index read (124 B); 15 values of
// the pair record written (FP32 in mixed precision).  k_soft_body: per
// incidence the index and the 8 record values it reads (6 force/torque
// entries, the stress term, the active flag); per body rowptr and colony
// index, 9 state doubles read (length, l0, position, quaternion) and 23
// written (position, quaternion, ldot, length, F, U, stress, impedance).
double hertz_bytes(int64_t C, bool mixed) { return (double)C * (124.0 + 15.0 * (mixed ? 4 : 8)); }
double soft_body_bytes(int64_t N, int64_t C, bool mixed) {
  return 2.0 * C * (4.0 + 8.0 * (mixed ? 4 : 8)) + (double)N * (8.0 + 9 * 8 + 23 * 8);
}

void soft_forces_and_integrate(pcc_ctx* ctx) {
  const int64_t N = ctx->N, C = ctx->S.C;
  const bool mixed = ctx->params.precision == PCC_PREC_MIXED;
  hipStream_t s = ctx->stream;
  BodyArrays b = body_arrays(ctx->body.p, N);
  ctx->hertz.ensure(16 * (C + 1));
  ctx->hertz32.ensure(16 * (C + 1));
  {
    KTimer t(ctx, PCC_K_HERTZ, hertz_bytes(C, mixed));
    ConstraintSet& S = ctx->S;
    if (C > 0)
      k_hertz<<<nblk(C), 256, 0, s>>>(C, mixed ? 1 : 0, S.I.p, S.J.p, S.phi.p, S.n.p, S.rI.p,
                                      S.rJ.p, S.sI.p, S.sJ.p, b.diam, S.inst.p, ctx->idev.p,
                                      ctx->hertz.p, ctx->hertz32.p);
  }
  KTimer t(ctx, PCC_K_INTEGRATE, soft_body_bytes(N, C, mixed));
  if (N > 0)
    k_soft_body<<<nblk(N), 256, 0, s>>>(N, mixed ? 1 : 0, b, ctx->rowptr.p, ctx->inc.p,
                                        ctx->hertz.p, ctx->hertz32.p, ctx->binst.p, ctx->idev.p,
                                        ctx->brownian ? ctx->uext.p : nullptr, ctx->F.p,
                                        ctx->U.p, ctx->sig.p);
}

}  // namespace pcc
