// adapt.hip — CFL time-step adaptation on the device, between steps.
//
// Replaces Domain::adaptDt (Domain.cpp:69-118) for every colony of a context:
// the median over the colony's cells of |v| + ldot after the last step (for an
// even count the mean of the two middle values, as nth_element plus the
// max of the lower half gives), then dt' = dt (1 - alpha) + alpha cfl / median
// clamped to [0.9 dt, 1.1 dt], with cfl = cfl_factor * tolerance; dt is kept
// when the median is 0.  The median is exact: a stable radix sort of the
// speeds, then by colony.
#include <algorithm>
#include <cmath>
#include <vector>
#include <rocprim/device/device_radix_sort.hpp>

#include "common.hpp"
#include "internal.hpp"

namespace pcc {

namespace {

inline unsigned nblk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// |v| = sqrt(dot(v, v)) (ArrayMath magnitude, dot from 0) plus ldot.
__global__ void k_speed(int64_t N, const double* __restrict__ U, const double* __restrict__ ldot,
                        double* __restrict__ v, uint64_t* __restrict__ key,
                        int32_t* __restrict__ val) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double* u = U + 6 * i;
  double s = 0.0;
  s += u[0] * u[0];
  s += u[1] * u[1];
  s += u[2] * u[2];
  const double x = sqrt(s) + ldot[i];
  v[i] = x;
  key[i] = order_key(x);
  val[i] = (int32_t)i;
}

__global__ void k_speed_colony_keys(int64_t N, const int32_t* __restrict__ perm,
                                    const int32_t* __restrict__ binst, uint32_t* __restrict__ k) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < N) k[p] = (uint32_t)binst[perm[p]];
}

// The two middle speeds of every colony (the lower one equals the upper one
// for odd counts).
__global__ void k_pick_middle(int64_t B, const InstDev* __restrict__ idev,
                              const int32_t* __restrict__ order, const double* __restrict__ v,
                              double* __restrict__ out) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= B) return;
  const int64_t n = idev[q].N, o = idev[q].boff;
  if (n == 0) {
    out[2 * q] = out[2 * q + 1] = 0.0;
    return;
  }
  const int64_t hi = o + n / 2;
  const int64_t lo = (n % 2 == 0) ? hi - 1 : hi;
  out[2 * q] = v[order[lo]];
  out[2 * q + 1] = v[order[hi]];
}

}  // namespace

// Median speed of every colony.
void colony_median_speeds(pcc_ctx* ctx, std::vector<double>& median) {
  const int64_t N = ctx->N;
  const int64_t B = (int64_t)ctx->inst.size();
  hipStream_t s = ctx->stream;
  median.assign(B, 0.0);
  if (N == 0) return;
  BodyArrays b = body_arrays(ctx->body.p, N);
  DArr<double> v, mid;
  DArr<uint64_t> key, key2;
  DArr<int32_t> val, perm, perm2;
  DArr<uint32_t> ck, ck2;
  v.ensure(N + 1);
  key.ensure(N + 1);
  key2.ensure(N + 1);
  val.ensure(N + 1);
  perm.ensure(N + 1);
  mid.ensure(2 * B + 2);
  k_speed<<<nblk(N), 256, 0, s>>>(N, ctx->U.p, b.ldot, v.p, key.p, val.p);
  size_t bytes = 0;
  PCC_HIP(rocprim::radix_sort_pairs(nullptr, bytes, key.p, key2.p, val.p, perm.p, (size_t)N, 0, 64,
                                    s));
  ctx->sort_tmp.ensure((int64_t)bytes + 16);
  PCC_HIP(rocprim::radix_sort_pairs(ctx->sort_tmp.p, bytes, key.p, key2.p, val.p, perm.p,
                                    (size_t)N, 0, 64, s));
  const int32_t* order = perm.p;
  if (B > 1) {
    perm2.ensure(N + 1);
    ck.ensure(N + 1);
    ck2.ensure(N + 1);
    k_speed_colony_keys<<<nblk(N), 256, 0, s>>>(N, perm.p, ctx->binst.p, ck.p);
    int bits = 1;
    while (((int64_t)1 << bits) < B) ++bits;
    bytes = 0;
    PCC_HIP(rocprim::radix_sort_pairs(nullptr, bytes, ck.p, ck2.p, perm.p, perm2.p, (size_t)N, 0,
                                      bits, s));
    ctx->sort_tmp.ensure((int64_t)bytes + 16);
    PCC_HIP(rocprim::radix_sort_pairs(ctx->sort_tmp.p, bytes, ck.p, ck2.p, perm.p, perm2.p,
                                      (size_t)N, 0, bits, s));
    order = perm2.p;
  }
  k_pick_middle<<<nblk(B), 256, 0, s>>>(B, ctx->idev.p, order, v.p, mid.p);
  std::vector<double> h(2 * B);
  PCC_HIP(hipMemcpyAsync(h.data(), mid.p, sizeof(double) * 2 * B, hipMemcpyDeviceToHost, s));
  PCC_HIP(hipStreamSynchronize(s));
  for (int64_t q = 0; q < B; ++q) {
    const int64_t n = ctx->inst[q].N;
    median[q] = (n % 2 == 0 && n > 0) ? 0.5 * (h[2 * q] + h[2 * q + 1]) : h[2 * q + 1];
  }
}

// Domain.cpp:103-117 for one colony.
double adapted_dt(double dt, double median, double cfl_factor, double tolerance, double alpha) {
  if (median == 0.0) return dt;
  const double cfl = cfl_factor * tolerance;
  const double cap = cfl / median;
  const double nd = dt * (1.0 - alpha) + alpha * cap;
  return std::clamp(nd, 0.9 * dt, 1.1 * dt);
}

}  // namespace pcc
