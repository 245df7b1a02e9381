// common.hpp — device helpers shared by the libpcc kernels (gfx950).
//
// All FP64 helpers evaluate the reference's expressions in the reference's
// order; the library is compiled with -ffp-contract=off so no multiply-add is
// fused behind our back.  Together with the fixed reduction trees below this
// makes the FP64 path bit-identical to the CPU restatement of the reference.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PCC_BLOCK 256
#define PCC_WAVE 64

namespace pcc {

// ---------------------------------------------------------------------------
// ArrayMath.h (dot from 0, cross, normalize) and Quaternion.cpp:7-38
// ---------------------------------------------------------------------------
__device__ __forceinline__ double dot3(const double a[3], const double b[3]) {
  double r = 0.0;
  r += a[0] * b[0];
  r += a[1] * b[1];
  r += a[2] * b[2];
  return r;
}

__device__ __forceinline__ void cross3(const double a[3], const double b[3], double o[3]) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

// ArrayMath.h:218-231
__device__ __forceinline__ void normalize3(const double a[3], double o[3]) {
  double mag = sqrt(dot3(a, a));
  if (mag > 2.220446049250313080847e-16) {
    double inv = 1.0 / mag;
    o[0] = a[0] * inv;
    o[1] = a[1] * inv;
    o[2] = a[2] * inv;
  } else {
    o[0] = 1.0;
    o[1] = 0.0;
    o[2] = 0.0;
  }
}

// Quaternion::rotateVectorOfPositions(q, e_x) (Quaternion.cpp:7-38)
__device__ __forceinline__ void direction(double w, double x, double y, double z, double t[3]) {
  const double vx = 1.0, vy = 0.0, vz = 0.0;
  double c1x = y * vz - z * vy;
  double c1y = z * vx - x * vz;
  double c1z = x * vy - y * vx;
  double c2x = y * c1z - z * c1y;
  double c2y = z * c1x - x * c1z;
  double c2z = x * c1y - y * c1x;
  t[0] = vx + 2.0 * w * c1x + 2.0 * c2x;
  t[1] = vy + 2.0 * w * c1y + 2.0 * c2y;
  t[2] = vz + 2.0 * w * c1z + 2.0 * c2z;
}

// ---------------------------------------------------------------------------
// Deterministic exp: Cody-Waite reduction with ln2 split 33+53 bits, degree-13
// Taylor polynomial in Horner form, exact 2^k scaling.  Coefficients are the
// correctly rounded 1/n!.  Identical expression order to the CPU restatement.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double det_exp(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return __builtin_inf();
  if (x < -745.1332191019412) return 0.0;
  const double inv_ln2 = 0x1.71547652b82fep+0;
  const double ln2_hi = 0x1.62e42fef00000p-1;
  const double ln2_lo = 0x1.473de6af278edp-34;
  double k = rint(x * inv_ln2);
  double r = (x - k * ln2_hi) - k * ln2_lo;
  double p = 0x1.6124613a86d09p-33;
  p = p * r + 0x1.1eed8eff8d898p-29;
  p = p * r + 0x1.ae64567f544e4p-26;
  p = p * r + 0x1.27e4fb7789f5cp-22;
  p = p * r + 0x1.71de3a556c734p-19;
  p = p * r + 0x1.a01a01a01a01ap-16;
  p = p * r + 0x1.a01a01a01a01ap-13;
  p = p * r + 0x1.6c16c16c16c17p-10;
  p = p * r + 0x1.1111111111111p-7;
  p = p * r + 0x1.5555555555555p-5;
  p = p * r + 0x1.5555555555555p-3;
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  return ldexp(p, (int)k);
}

// ---------------------------------------------------------------------------
// Segment-segment closest points: DCPQuery<double,3>::operator()
// (SpherocylinderCell.h:64-261), region analysis in the same branch order.
// out: s, t, c0[3], c1[3], dist
// ---------------------------------------------------------------------------
struct DcpOut {
  double c0[3], c1[3];
  double dist;
};

__device__ __forceinline__ void dcp(const double P0[3], const double P1[3], const double Q0[3],
                                    const double Q1[3], DcpOut& o) {
  double P1mP0[3], Q1mQ0[3], P0mQ0[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    P1mP0[k] = P1[k] - P0[k];
    Q1mQ0[k] = Q1[k] - Q0[k];
    P0mQ0[k] = P0[k] - Q0[k];
  }
  const double a = dot3(P1mP0, P1mP0);
  const double b = dot3(P1mP0, Q1mQ0);
  const double c = dot3(Q1mQ0, Q1mQ0);
  const double d = dot3(P1mP0, P0mQ0);
  const double e = dot3(Q1mQ0, P0mQ0);
  const double det = a * c - b * b;
  double s, t, nd, bmd, bte, ctd, bpe, ate, btd;
  const double zero = 0.0, one = 1.0;
  if (det > zero) {
    bte = b * e;
    ctd = c * d;
    if (bte <= ctd) {
      s = zero;
      if (e <= zero) {
        t = zero;
        nd = -d;
        if (nd >= a) {
          s = one;
        } else if (nd > zero) {
          s = nd / a;
        }
      } else if (e < c) {
        t = e / c;
      } else {
        t = one;
        bmd = b - d;
        if (bmd >= a) {
          s = one;
        } else if (bmd > zero) {
          s = bmd / a;
        }
      }
    } else {
      s = bte - ctd;
      if (s >= det) {
        s = one;
        bpe = b + e;
        if (bpe <= zero) {
          t = zero;
          nd = -d;
          if (nd <= zero) {
            s = zero;
          } else if (nd < a) {
            s = nd / a;
          }
        } else if (bpe < c) {
          t = bpe / c;
        } else {
          t = one;
          bmd = b - d;
          if (bmd <= zero) {
            s = zero;
          } else if (bmd < a) {
            s = bmd / a;
          }
        }
      } else {
        ate = a * e;
        btd = b * d;
        if (ate <= btd) {
          t = zero;
          nd = -d;
          if (nd <= zero) {
            s = zero;
          } else if (nd >= a) {
            s = one;
          } else {
            s = nd / a;
          }
        } else {
          t = ate - btd;
          if (t >= det) {
            t = one;
            bmd = b - d;
            if (bmd <= zero) {
              s = zero;
            } else if (bmd >= a) {
              s = one;
            } else {
              s = bmd / a;
            }
          } else {
            s /= det;
            t /= det;
          }
        }
      }
    }
  } else {
    if (e <= zero) {
      t = zero;
      nd = -d;
      if (nd <= zero) {
        s = zero;
      } else if (nd >= a) {
        s = one;
      } else {
        s = nd / a;
      }
    } else if (e >= c) {
      t = one;
      bmd = b - d;
      if (bmd <= zero) {
        s = zero;
      } else if (bmd >= a) {
        s = one;
      } else {
        s = bmd / a;
      }
    } else {
      s = zero;
      t = e / c;
    }
  }
  double diff[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    o.c0[k] = P0[k] + P1mP0[k] * s;
    o.c1[k] = Q0[k] + Q1mQ0[k] * t;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) diff[k] = o.c0[k] - o.c1[k];
  o.dist = sqrt(dot3(diff, diff));
}

// ---------------------------------------------------------------------------
// Hashed cell list: floor((x - origin)/h) per axis (SpatialGrid.cpp:29-36),
// then a 32-bit mix (odd multipliers + MurmurHash3 fmix32).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int32_t cell_coord(double x, double origin, double h) {
  return (int32_t)floor((x - origin) / h);
}

__host__ __device__ __forceinline__ uint32_t cell_hash(int32_t cx, int32_t cy, int32_t cz) {
  uint32_t h = (uint32_t)cx * 0x9E3779B1u + (uint32_t)cy * 0x85EBCA77u + (uint32_t)cz * 0xC2B2AE3Du;
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return h;
}

// Per-colony salt of the cell hash: colonies of one context share the hashed
// table without sharing cells.  Colony 0 gets 0, so a single colony hashes
// exactly as above.
__host__ __device__ __forceinline__ uint32_t inst_mix(int32_t b) {
  return b == 0 ? 0u : cell_hash(b, 0x5bd1e995, 0x27d4eb2f);
}

// XCD-aware block mapping: the hardware dispatches block h to XCD h % 8, so
// logical block (h % 8) * (G/8) + h / 8 makes each XCD sweep one contiguous
// range and keeps its gathers in its own L2 (gp, the grid, is a multiple of 8).
static constexpr int kXcd = 8;
__device__ __forceinline__ int64_t xcd_block(int64_t h, int64_t gp) {
  return (h % kXcd) * (gp / kXcd) + h / kXcd;
}
inline int64_t pad8(int64_t g) { return (g + kXcd - 1) / kXcd * kXcd; }

// Table slot of a cell (before the mask).  A single z-layer (flat) maps cells
// row by row, 1024 to a row, so that neighbouring cells sit in nearby slots
// and a contiguous range of the table covers a band of the colony (the grid
// walk then stays in one L2); other layouts use the hash.  Colonies are
// offset by their salt.  Cells that share a slot are told apart by their
// coordinates wherever the table is walked.
__host__ __device__ __forceinline__ uint32_t cell_slot(int32_t cx, int32_t cy, int32_t cz,
                                                       int32_t inst, bool flat) {
  const uint32_t k = flat ? (uint32_t)cx + ((uint32_t)cy << 10) : cell_hash(cx, cy, cz);
  return flat ? k + inst_mix(inst) : k ^ inst_mix(inst);
}

// ---------------------------------------------------------------------------
// Deterministic reductions.  A 64-lane butterfly (xor 32,16,...,1) leaves the
// same value in every lane; a 256-thread block combines its four waves as
// ((w0 + w1) + w2) + w3.  The CPU restatement evaluates the same tree.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, PCC_WAVE);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    double o = __shfl_xor(v, off, PCC_WAVE);
    v = v < o ? o : v;
  }
  return v;
}

// Block sum over PCC_BLOCK threads; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* lds4) {
  v = wave_sum(v);
  const int lane = threadIdx.x & (PCC_WAVE - 1);
  const int w = threadIdx.x >> 6;
  if (lane == 0) lds4[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = ((lds4[0] + lds4[1]) + lds4[2]) + lds4[3];
  __syncthreads();
  return r;
}

__device__ __forceinline__ double block_max(double v, double* lds4) {
  v = wave_max(v);
  const int lane = threadIdx.x & (PCC_WAVE - 1);
  const int w = threadIdx.x >> 6;
  if (lane == 0) lds4[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    r = lds4[0];
    for (int i = 1; i < 4; ++i) r = r < lds4[i] ? lds4[i] : r;
  }
  __syncthreads();
  return r;
}

// splitmix64 finaliser (Steele, Lea, Flood 2014)
__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// Brownian draw k of body i (index within its colony) at the colony's step
// count `iter`:
// uniform on [-1, 1) with 53 random bits, exact in double arithmetic
// (std::uniform_real_distribution(-1, 1), Particle.cpp:177).  The stream
// constant keeps these draws apart from the division draws of the same seed.
__device__ __forceinline__ double brownian_uniform(uint64_t seed, int64_t iter, int64_t i, int k) {
  const uint64_t key = mix64(mix64(seed ^ 0xB7E151628AED2A6Bull) ^ (uint64_t)(uint32_t)iter);
  const uint64_t z = mix64(key ^ ((uint64_t)i * 4u + (uint64_t)k));
  return (double)(z >> 11) * 0x1.0p-52 - 1.0;
}

// Particle::calculateBrownianVelocity (Particle.cpp:174-190) added into a
// zeroed U_ext (Physics.cpp:60,75): angular velocity (wx, wy, wz) of a body
// with impedance imp; every product in the reference's left-to-right order.
__device__ __forceinline__ void brownian_omega(double T, int mono, uint64_t seed, int64_t iter,
                                               int64_t i, double imp, double w[3]) {
  const double rot = T * brownian_uniform(seed, iter, i, 0);
  if (mono) {
    w[0] = 0.0 + 0.0;
    w[1] = 0.0 + 0.0;
    w[2] = 0.0 + (((rot * brownian_uniform(seed, iter, i, 1)) * imp) * imp) * imp;
  } else {
    w[0] = 0.0 + (((rot * brownian_uniform(seed, iter, i, 1)) * imp) * imp) * imp;
    w[1] = 0.0 + (((rot * brownian_uniform(seed, iter, i, 2)) * imp) * imp) * imp;
    w[2] = 0.0 + (((rot * brownian_uniform(seed, iter, i, 3)) * imp) * imp) * imp;
  }
}

// Order-preserving map of a double to uint64 for radix sorting (-0.0 is
// folded onto +0.0 first, matching operator< semantics).
__device__ __forceinline__ uint64_t order_key(double x) {
  x = x + 0.0;
  uint64_t u = (uint64_t)__double_as_longlong(x);
  return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}

}  // namespace pcc
