// divide.hip — cell division and gid reassignment on the device, between steps.
//
// Replaces ParticleManager::divideParticles / Particle::divide
// (ParticleManager.cpp:139-148, Particle.cpp:122-150) and the gid
// reassignment by ascending x of Domain::assignGlobalIDs (Domain.cpp:492-508),
// for every colony of a context, so that a growing colony never leaves HBM.
//
//  1. k_div_flags   1 for every cell with length >= 2 l0
//  2. scan          rank of each dividing cell; per-colony counts
//  3. k_div_write   staging layout: colony q's cells in place (a dividing cell
//                   becomes its left child), then its right children in
//                   parent order; l1 = U[1-a, 1+a] L/2, l2 = L - l1, centres
//                   x -/+ (l/2) t, each child turned in-plane by an angle drawn
//                   from N(0, angle_noise) (Irwin-Hall, twelve uniforms), both
//                   children one generation older than the parent
//  4. xsort_run     gids by ascending x per colony, ties in staging order
//  5. k_div_gather  body fields into the new order
// The draws are counter-based on (seed of the colony, step, index of the parent
// within its colony), so a colony divides the same way alone and in a batch.
// Only +, -, *, / and sqrt are used, in the order of the CPU restatement.
#include <algorithm>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "internal.hpp"

namespace pcc {

void exclusive_scan_i32(pcc_ctx* ctx, const int32_t* in, int32_t* out, int64_t n);

namespace {

inline unsigned nblk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// splitmix64 finaliser (Steele, Lea, Flood 2014)
__device__ __forceinline__ uint64_t split_mix(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__device__ __forceinline__ double div_uniform(uint64_t seed, int32_t step, int64_t p, int k) {
  const uint64_t key = split_mix(split_mix(seed) ^ (uint64_t)(uint32_t)step);
  const uint64_t z = split_mix(key ^ ((uint64_t)p * 32u + (uint64_t)k));
  return (double)(z >> 11) * 0x1.0p-53;
}

__device__ __forceinline__ double div_normal(uint64_t seed, int32_t step, int64_t p, int k0) {
  double s = 0.0;
  for (int k = 0; k < 12; ++k) s += div_uniform(seed, step, p, k0 + k);
  return s - 6.0;
}

// cos and sin of a half-angle |x| <= 0.6: Taylor series to x^21, Horner in
// x^2 with correctly rounded 1/n!.
__device__ __forceinline__ void half_angle_cs(double x, double& c, double& s) {
  const double y = x * x;
  double pc = -0x1.0ce396db7f853p-70;
  pc = pc * y + 0x1.e542ba4020225p-62;
  pc = pc * y - 0x1.6827863b97d97p-53;
  pc = pc * y + 0x1.ae7f3e733b81fp-45;
  pc = pc * y - 0x1.93974a8c07c9dp-37;
  pc = pc * y + 0x1.1eed8eff8d898p-29;
  pc = pc * y - 0x1.27e4fb7789f5cp-22;
  pc = pc * y + 0x1.a01a01a01a01ap-16;
  pc = pc * y - 0x1.6c16c16c16c17p-10;
  pc = pc * y + 0x1.5555555555555p-5;
  pc = pc * y - 0.5;
  c = pc * y + 1.0;
  double ps = 0x1.71b8ef6dcf572p-66;
  ps = ps * y - 0x1.2f49b46814157p-57;
  ps = ps * y + 0x1.952c77030ad4ap-49;
  ps = ps * y - 0x1.ae7f3e733b81fp-41;
  ps = ps * y + 0x1.6124613a86d09p-33;
  ps = ps * y - 0x1.ae64567f544e4p-26;
  ps = ps * y + 0x1.71de3a556c734p-19;
  ps = ps * y - 0x1.a01a01a01a01ap-13;
  ps = ps * y + 0x1.1111111111111p-7;
  ps = ps * y - 0x1.5555555555555p-3;
  ps = ps * y + 1.0;
  s = ps * x;
}

// q <- normalize((cos(d/2), 0, 0, sin(d/2)) * q), renormalised as
// Particle.cpp:110-120.
__device__ __forceinline__ void turn_z(double q[4], double d) {
  double c, s;
  half_angle_cs(0.5 * d, c, s);
  double r[4];
  r[0] = c * q[0] - s * q[3];
  r[1] = c * q[1] - s * q[2];
  r[2] = c * q[2] + s * q[1];
  r[3] = c * q[3] + s * q[0];
  const double norm = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + r[3] * r[3]);
#pragma unroll
  for (int k = 0; k < 4; ++k) q[k] = r[k] / norm;
}

// flag[N] (the scan's closing entry) is 0.
__global__ void k_div_flags(int64_t N, const double* __restrict__ len,
                            const double* __restrict__ l0, int32_t* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > N) return;
  flag[i] = i < N && len[i] >= 2 * l0[i] ? 1 : 0;
}

struct DivColony {
  int64_t boff, N;    // before
  int64_t noff;       // first body after
  int64_t rank0;      // dividing cells before this colony (scan value at boff)
  uint64_t seed;
};

// Dividing cells before each colony; also clears the per-colony x-key
// bounds of the x sort.
__global__ void k_div_bounds(int64_t B, int64_t N, const InstDev* __restrict__ idev,
                             const int32_t* __restrict__ off, int32_t* __restrict__ out,
                             unsigned long long* __restrict__ kmin,
                             unsigned long long* __restrict__ kmax) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b > B) return;
  out[b] = off[b < B ? idev[b].boff : N];
  if (b < B) {
    kmin[b] = ~0ull;
    kmax[b] = 0ull;
  }
}

// Staging: one thread per old body.  stage is field-major with stride M (new
// total), key (the order key of x) feeds the x sort, sinst the colony of every
// staged body.
__global__ void k_div_write(int64_t N, int64_t M, BodyArrays b, const int32_t* __restrict__ binst,
                            const DivColony* __restrict__ col, const int32_t* __restrict__ off,
                            int32_t step, double length_noise, double angle_noise,
                            double* __restrict__ stage, uint64_t* __restrict__ key,
                            int32_t* __restrict__ sinst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int32_t q = binst[i];
  const DivColony c = col[q];
  const int64_t local = i - c.boff;
  const int64_t j = c.noff + local;
  double x[3] = {b.px[i], b.py[i], b.pz[i]};
  double qq[4] = {b.qw[i], b.qx[i], b.qy[i], b.qz[i]};
  const double L = b.len[i], l0 = b.l0[i], dm = b.diam[i], ld = b.ldot[i], age = b.age[i];
  const double imp = b.imp[i];
  auto put = [&](int64_t s, const double p[3], const double qv[4], double len, double ag,
                 double im) {
    stage[0 * M + s] = p[0];
    stage[1 * M + s] = p[1];
    stage[2 * M + s] = p[2];
    stage[3 * M + s] = qv[0];
    stage[4 * M + s] = qv[1];
    stage[5 * M + s] = qv[2];
    stage[6 * M + s] = qv[3];
    stage[7 * M + s] = len;
    stage[8 * M + s] = l0;
    stage[9 * M + s] = dm;
    stage[10 * M + s] = ld;
    stage[11 * M + s] = ag;
    stage[12 * M + s] = im;
    key[s] = order_key(p[0]);
    sinst[s] = q;
  };
  if (!(L >= 2 * l0)) {
    put(j, x, qq, L, age, imp);
    return;
  }
  const double half = L / 2;
  const double lo = half * (1.0 - length_noise), hi = half * (1.0 + length_noise);
  const double l1 = lo + div_uniform(c.seed, step, local, 0) * (hi - lo);
  const double l2 = L - l1;
  double t[3];
  direction(qq[0], qq[1], qq[2], qq[3], t);
  double ql[4], qr[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) ql[k] = qr[k] = qq[k];
  turn_z(ql, angle_noise * div_normal(c.seed, step, local, 1));
  turn_z(qr, angle_noise * div_normal(c.seed, step, local, 13));
  double pl[3], pr[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    pr[k] = x[k] + (l2 / 2) * t[k];
    pl[k] = x[k] - (l1 / 2) * t[k];
  }
  // Particle::divide: the parent's age grows by one and the new cell copies
  // it; the parent keeps its impedance, the new cell starts at 0 as a newly
  // constructed Particle does (Particle.cpp:25)
  put(j, pl, ql, l1, age + 1.0, imp);
  const int64_t jr = c.noff + c.N + (off[i] - c.rank0);
  put(jr, pr, qr, l2, age + 1.0, 0.0);
}

__global__ void k_div_gather(int64_t M, const double* __restrict__ stage,
                             const int32_t* __restrict__ perm, double* __restrict__ body) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M) return;
  const int64_t s = perm[j];
#pragma unroll
  for (int k = 0; k < kBodyFields; ++k) body[(int64_t)k * M + j] = stage[(int64_t)k * M + s];
}

}  // namespace

// Divides every cell with length >= 2 l0 in every colony and reassigns gids by
// ascending x per colony, as the reference does before every step whether or
// not a cell divided.  seeds[q] is colony q's seed.  The constraint set is
// emptied; the caller resets the per-body outputs of the last step.
void divide_colonies(pcc_ctx* ctx, const uint64_t* seeds, int32_t step, double length_noise,
                     double angle_noise) {
  const int64_t N = ctx->N;
  const int64_t B = (int64_t)ctx->inst.size();
  hipStream_t s = ctx->stream;
  if (N == 0) return;
  BodyArrays b = body_arrays(ctx->body.p, N);
  DArr<int32_t>& flag = ctx->dv_flag;
  DArr<int32_t>& off = ctx->dv_off;
  DArr<int32_t>& bounds = ctx->dv_bounds;
  flag.ensure(N + 1);
  off.ensure(N + 1);
  bounds.ensure(B + 1);
  k_div_flags<<<nblk(N + 1), 256, 0, s>>>(N, b.len, b.l0, flag.p);
  exclusive_scan_i32(ctx, flag.p, off.p, N + 1);
  ctx->xs_bounds.ensure(2 * B);
  k_div_bounds<<<nblk(B + 1), 256, 0, s>>>(B, N, ctx->idev.p, off.p, bounds.p,
                                           ctx->xs_bounds.p, ctx->xs_bounds.p + B);
  int32_t* pin = readback_pin(ctx, B + 1);
  PCC_HIP(hipMemcpyAsync(pin, bounds.p, sizeof(int32_t) * (B + 1), hipMemcpyDeviceToHost, s));
  PCC_HIP(hipStreamSynchronize(s));
  const std::vector<int32_t> hb(pin, pin + B + 1);
  const int64_t D = hb[B] - hb[0];  // with no divider the cells are still renumbered by x
  const int64_t M = N + D;
  if (M >= (int64_t)1 << 30) throw StatusError{PCC_E_INVALID_ARGUMENT, "too many bodies"};
  std::vector<DivColony> hc(B);
  int64_t noff = 0;
  for (int64_t q = 0; q < B; ++q) {
    const Instance& I = ctx->inst[q];
    hc[q] = {I.boff, I.N, noff, (int64_t)hb[q], seeds ? seeds[q] : 0};
    noff += I.N + (hb[q + 1] - hb[q]);
  }
  // one upload: the colony records, then each colony's first new gid (B + 1
  // entries, the groups of the x order)
  const size_t col_bytes = (sizeof(DivColony) * B + 15) / 16 * 16;
  std::vector<uint8_t> blob(col_bytes + sizeof(int32_t) * (B + 1));
  memcpy(blob.data(), hc.data(), sizeof(DivColony) * B);
  {
    int32_t* gs = reinterpret_cast<int32_t*>(blob.data() + col_bytes);
    for (int64_t q = 0; q < B; ++q) gs[q] = (int32_t)hc[q].noff;
    gs[B] = (int32_t)M;
  }
  ctx->dv_col.ensure((int64_t)blob.size());
  DivColony* dcol = reinterpret_cast<DivColony*>(ctx->dv_col.p);
  const int32_t* gstart = reinterpret_cast<const int32_t*>(ctx->dv_col.p + col_bytes);
  PCC_HIP(hipMemcpyAsync(ctx->dv_col.p, blob.data(), blob.size(), hipMemcpyHostToDevice, s));
  DArr<double>& stage = ctx->dv_stage;
  DArr<uint64_t>& key = ctx->dv_key;
  DArr<int32_t>& perm = ctx->dv_perm;
  DArr<int32_t>& sinst = ctx->dv_sinst;
  stage.ensure(kBodyFields * M + 1);
  key.ensure(M + 1);
  perm.ensure(M + 1);
  sinst.ensure(M + 1);
  k_div_write<<<nblk(N), 256, 0, s>>>(N, M, b, ctx->binst.p, dcol, off.p, step, length_noise,
                                      angle_noise, stage.p, key.p, sinst.p);
  // gids by ascending x within each colony, ties in staging order
  ctx->xs_cnt.ensure(M + 1);
  ctx->xs_off.ensure(M + 1);
  ctx->sort_tmp.ensure((int64_t)xsort_tmp_bytes(M) + 16);
  XSort xs;
  xs.n = M;
  xs.G = B;
  xs.gstart = gstart;
  xs.key = key.p;
  xs.tie = nullptr;  // the staging index
  xs.grp = sinst.p;
  xs.cnt = ctx->xs_cnt.p;
  xs.off = ctx->xs_off.p;
  xs.kmin = ctx->xs_bounds.p;
  xs.kmax = ctx->xs_bounds.p + B;
  xs.bounds_clear = true;  // by k_div_bounds
  xs.perm = perm.p;
  xs.tmp = ctx->sort_tmp.p;
  xs.tmp_bytes = (size_t)ctx->sort_tmp.cap;
  xs.s = s;
  xsort_run(xs);
  const int32_t* order = perm.p;
  DArr<double>& body = ctx->dv_body;
  body.ensure(kBodyFields * M + 1);
  k_div_gather<<<nblk(M), 256, 0, s>>>(M, stage.p, order, body.p);
  ctx->body.swap(body);  // stream order keeps every later use after the gather
  // new colony ranges; per-body outputs of the last step no longer apply
  for (int64_t q = 0; q < B; ++q) {
    Instance& I = ctx->inst[q];
    I.boff = hc[q].noff;
    I.N += hb[q + 1] - hb[q];
    I.C = 0;
    I.coff = 0;
  }
  ctx->N = M;
  ctx->S.C = 0;
  ctx->has_snapshot = false;
}

}  // namespace pcc
