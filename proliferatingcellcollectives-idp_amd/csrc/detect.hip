// detect.hip — contact detection on the GPU.
//
// Replaces SpatialGrid::findPotentialCollisions (SpatialGrid.cpp:58-121),
// CollisionDetector::checkParticlePairs / tryCreateConstraint
// (CollisionDetector.cpp:144-183, 201-311) and the contact-x sort of
// detectCollisions (CollisionDetector.cpp:84-111).
//
//  1. k_cells      cell = floor((x+1)/h) per axis, bucket = hash(cell) & (T-1),
//                  histogram of buckets (integer atomics)
//  2. scan         bucket offsets
//  3. k_scatter    bodies into buckets, then k_seg_sort orders each bucket
//                  by body index: a stable counting sort, identical to the
//                  CPU restatement's; k_stage copies every body's cell and
//                  detection record into bucket order, so that a bucket's
//                  members are contiguous (one bucket walk reads consecutive
//                  records, with no index indirection before the loads)
//  4. k_filter     one thread per (body I, neighbour cell), threads in bucket
//                  order, tests every J > I of that cell against the
//                  reference's centre-distance filter (after a conservative
//                  FP32 pre-filter) and emits the passing pairs as one
//                  contiguous run of candidates; 27 neighbour cells, or the 9
//                  of the body's own z-layer when every body has z = 0 (the
//                  other cells are then empty)
//     k_pair_test  one thread per candidate: the exact FP64 segment-segment
//                  DCP on dense wavefronts (a thread per (body, cell) would
//                  run it with a few lanes of each wavefront busy); counts
//                  the contacts of each (body, neighbour cell) slot
//     scan         slot offsets, so each body's records keep the cell order
//                  dz, dy, dx and the bucket order
//     k_pair_write recomputes each accepted pair and writes its record
//                  with the keys of the order (colony, contact x, I, J)
//  5. xsort_run:   that order by a bucket sort (xsort.hip)
//
// Colonies of one context share the hashed table: the colony index salts the
// bucket hash and neighbour tests require the same colony, so each colony's
// contact list is the one it would have alone.
#include <algorithm>
#include <cstring>
#include <vector>
#include <rocprim/device/device_scan.hpp>

#include "common.hpp"
#include "internal.hpp"

namespace pcc {

namespace {

// Per body: its cell and colony (cell4), and the detection geometry record
// geo = (x, y, z, t_x | t_y, t_z, length, diameter) with t = rotate(q, e_x), so
// that the detection kernel reads a candidate from two cache lines.
__global__ void k_cells(int64_t N, BodyArrays b, const int32_t* __restrict__ binst,
                        const InstDev* __restrict__ idev, uint32_t mask, bool flat,
                        int4* __restrict__ cell4, double2* __restrict__ geo,
                        uint32_t* __restrict__ bucket, int32_t* __restrict__ count) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int32_t inst = binst[i];
  const double origin = idev[inst].origin, h = idev[inst].h;
  const double x = b.px[i], y = b.py[i], z = b.pz[i];
  int32_t cx = cell_coord(x, origin, h);
  int32_t cy = cell_coord(y, origin, h);
  int32_t cz = cell_coord(z, origin, h);
  cell4[i] = make_int4(cx, cy, cz, inst);
  double t[3];
  direction(b.qw[i], b.qx[i], b.qy[i], b.qz[i], t);
  geo[4 * i] = make_double2(x, y);
  geo[4 * i + 1] = make_double2(z, t[0]);
  geo[4 * i + 2] = make_double2(t[1], t[2]);
  geo[4 * i + 3] = make_double2(b.len[i], b.diam[i]);
  uint32_t bk = cell_slot(cx, cy, cz, inst, flat) & mask;
  bucket[i] = bk;
  atomicAdd(&count[bk], 1);
}

// Bodies into their buckets; cnt holds each bucket's count (from k_cells)
// and ends at zero.  The order within a bucket is the order of arrival;
// k_seg_sort then orders it by body index.
__global__ void k_scatter(int64_t N, const uint32_t* __restrict__ bucket,
                          const int32_t* __restrict__ start, int32_t* __restrict__ cnt,
                          int32_t* __restrict__ perm) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  uint32_t bk = bucket[i];
  int32_t pos = start[bk] + atomicSub(&cnt[bk], 1) - 1;
  perm[pos] = (int32_t)i;
}

// Every body's cell and detection record in bucket order (perm[pos] is the
// body at position pos), and its filter record: FP32 centre relative to its
// cell's corner (origin + cell * h per axis) and FP32 length.  Also clears
// the body's S per-slot contact counts (and the closing entry and the
// candidate-region flag) for the filter that follows.
__global__ void k_stage(int64_t N, int S, const int32_t* __restrict__ perm,
                        const int4* __restrict__ cell4, const double2* __restrict__ geo,
                        const InstDev* __restrict__ idev, int4* __restrict__ scell,
                        double2* __restrict__ sgeo, float4* __restrict__ frec,
                        int32_t* __restrict__ acnt, int32_t* __restrict__ need) {
  const int64_t pos = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (pos == 0) {
    acnt[N * S] = 0;
    *need = 0;
  }
  if (pos >= N) return;
  const int64_t J = perm[pos];
  for (int sl = 0; sl < S; ++sl) acnt[J * S + sl] = 0;
  const int4 c = cell4[J];
  scell[pos] = c;
  double2 g[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    g[q] = geo[4 * J + q];
    sgeo[4 * pos + q] = g[q];
  }
  const double origin = idev[c.w].origin, h = idev[c.w].h;
  frec[pos] = make_float4((float)(g[0].x - (origin + c.x * h)), (float)(g[0].y - (origin + c.y * h)),
                          (float)(g[1].x - (origin + c.z * h)), (float)g[3].x);
}

// Orders the members of each segment ascending (insertion sort; segments are
// short: bodies per hash bucket, or incidences per body).
__global__ void k_seg_sort(int64_t nseg, const int32_t* __restrict__ start,
                           int32_t* __restrict__ v) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  int32_t a = start[s], e = start[s + 1];
  for (int32_t i = a + 1; i < e; ++i) {
    int32_t x = v[i];
    int32_t j = i - 1;
    while (j >= a && v[j] > x) {
      v[j + 1] = v[j];
      --j;
    }
    v[j + 1] = x;
  }
}

// Raw contact record layout (16 doubles): phi, n[3], rI[3], rJ[3], p[3], sI, sJ, gamma0
enum { R_PHI = 0, R_N = 1, R_RI = 4, R_RJ = 7, R_P = 10, R_SI = 13, R_SJ = 14, R_G = 15, R_LEN = 16 };

// A body's segment as the contact tests use it: centre x, direction t,
// length, diameter and the end points P0 = x - t h, P1 = x + t h with
// h = length/2 - diameter/2 (CollisionDetector.cpp:215-217), from its
// detection record g[0..3].
struct Seg {
  double x[3], t[3], len, d, P0[3], P1[3];
};

__device__ __forceinline__ void seg_of(const double2* __restrict__ g, Seg& a) {
  const double2 g0 = g[0], g1 = g[1], g2 = g[2], g3 = g[3];
  a.x[0] = g0.x;
  a.x[1] = g0.y;
  a.x[2] = g1.x;
  a.t[0] = g1.y;
  a.t[1] = g2.x;
  a.t[2] = g2.y;
  a.len = g3.x;
  a.d = g3.y;
  const double half = a.len / 2 - a.d / 2;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double hv = a.t[k] * half;
    a.P0[k] = a.x[k] - hv;
    a.P1[k] = a.x[k] + hv;
  }
}

// Signed distance of the pair (closest points in o): segment-segment DCP
// minus the sum of the radii (CollisionDetector.cpp:219-237).
__device__ __forceinline__ double pair_phi(const Seg& a, const Seg& b, DcpOut& o) {
  dcp(a.P0, a.P1, b.P0, b.P1, o);
  const double sum_radii = (a.d + b.d) / 2.0;
  return o.dist - sum_radii;
}

// The raw record of contact (I = a, J = b) (CollisionDetector.cpp:239-258, 306).
__device__ __forceinline__ void pair_record(const Seg& a, const Seg& b, const DcpOut& o,
                                            double phi, double warm, double* __restrict__ r) {
  double p[3], d01[3], nrm[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    p[k] = (o.c0[k] + o.c1[k]) * 0.5;
    d01[k] = o.c0[k] - o.c1[k];
  }
  normalize3(d01, nrm);
  const double mn[3] = {-nrm[0], -nrm[1], -nrm[2]};
  r[R_PHI] = phi;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    r[R_N + k] = nrm[k];
    r[R_RI + k] = p[k] - a.x[k];
    r[R_RJ + k] = p[k] - b.x[k];
    r[R_P + k] = p[k];
  }
  r[R_SI] = 0.5 * fabs(dot3(nrm, a.t));
  r[R_SJ] = 0.5 * fabs(dot3(mn, b.t));
  r[R_G] = phi >= 0.0 ? 0.0 : -phi * warm;
}

constexpr int kFilterThreads = 256;

// Candidate pairs.  One thread per (body I, neighbour cell), threads in
// bucket order (the bodies of one cell are neighbours in the grid and their
// threads walk the same buckets; blocks are mapped XCD-aware, so each XCD
// walks one band of the table), tests the J > I of that cell against the
// reference's centre-distance filter (CollisionDetector.cpp:208-213).  A
// bucket lists its bodies in ascending index, so the walk runs from its end
// and stops at the first J <= I.  A conservative FP32 pre-filter on the
// filter records comes first; the cell and the exact FP64 test only for the
// pairs it keeps.  The thread's passing positions form one run of candidates
// (slot u = I * S + cell, bucket position of J), in the workgroup's own
// region of the candidate list: logical block l owns [l * cb, (l + 1) * cb),
// its runs placed by a scan in the workgroup (no atomics: atomics on one
// address serialise across the whole device).  cstart[u] = the run's first
// candidate, bcount[l] = the region's fill.  A region that would overflow
// writes nothing and raises *need to its total (the caller repeats the batch
// with larger regions).
template <int S>  // neighbour cells per body: 9 (one z-layer) or 27
__global__ void __launch_bounds__(kFilterThreads)
k_filter(int64_t N, int64_t gp, const int4* __restrict__ scell, const double2* __restrict__ sgeo,
         const float4* __restrict__ frec, const int32_t* __restrict__ bstart,
         const int32_t* __restrict__ perm, uint32_t mask, const InstDev* __restrict__ idev,
         int32_t cb, int32_t* __restrict__ bcount, int32_t* __restrict__ need,
         int32_t* __restrict__ cstart, int2* __restrict__ cand) {
  __shared__ int32_t wsum[kFilterThreads / 64];
  __shared__ int32_t btot;
  const int64_t l = xcd_block(blockIdx.x, gp);
  const int64_t t = l * kFilterThreads + threadIdx.x;
  bool live = t < N * S;
  int64_t I = 0, u = 0;
  int32_t nx = 0, ny = 0, nz = 0, inst = 0, e0 = 0, e1 = 0;
  double xI[3] = {0.0, 0.0, 0.0}, lenI = 0.0, cutoff = 0.0;
  float fx = 0.f, fy = 0.f, fz = 0.f, flen = 0.f, fcut = 0.f, fh = 0.f;
  if (live) {
    const int64_t h = t / S;  // the home body's position in bucket order
    const int sl = (int)(t - h * S);
    I = perm[h];
    u = I * S + sl;
    const int4 cI = scell[h];
    inst = cI.w;
    live = (idev[inst].flags & kInstActive) != 0;
    if (live) {
      cutoff = idev[inst].cutoff;
      const double2 g0 = sgeo[4 * h], g1 = sgeo[4 * h + 1], g3 = sgeo[4 * h + 3];
      xI[0] = g0.x;
      xI[1] = g0.y;
      xI[2] = g1.x;
      lenI = g3.x;
      const int dx = sl % 3 - 1, dy = (sl / 3) % 3 - 1, dz = S == 27 ? sl / 9 - 1 : 0;
      nx = cI.x + dx;
      ny = cI.y + dy;
      nz = cI.z + dz;
      // home centre relative to the neighbour cell's corner
      const float4 f = frec[h];
      fh = (float)idev[inst].h;
      fx = f.x - (float)dx * fh;
      fy = f.y - (float)dy * fh;
      fz = f.z - (float)dz * fh;
      flen = f.w;
      fcut = (float)cutoff;
      const uint32_t bk = cell_slot(nx, ny, nz, inst, S == 9) & mask;
      e0 = bstart[bk];
      e1 = bstart[bk + 1];
    }
  }
  // Filter records hold centres relative to their cell's corner: the FP32
  // centre difference is within about 4 * 2^-24 (h + |d|) of the FP64 one,
  // and the pre-filter's margin, 1e-5 (reach + h), is far above that, so
  // every pair that the FP64 test keeps passes it.
  auto passes = [&](int32_t e) -> bool {
    const float4 f = frec[e];
    const float ex = fx - f.x, ey = fy - f.y, ez = fz - f.z;
    const float fr = (flen + f.w) * 0.5f + fcut;
    const float lim = fr + 1e-5f * (fr + fh);
    if (ex * ex + ey * ey + ez * ez > lim * lim) return false;
    const int4 cJ = scell[e];
    if (cJ.x != nx || cJ.y != ny || cJ.z != nz || cJ.w != inst) return false;
    const double2 g0 = sgeo[4 * (int64_t)e], g1 = sgeo[4 * (int64_t)e + 1];
    const double lenJ = sgeo[4 * (int64_t)e + 3].x;
    const double dd[3] = {xI[0] - g0.x, xI[1] - g0.y, xI[2] - g1.x};
    const double rough = sqrt(dot3(dd, dd));
    return !(rough > (lenI + lenJ) / 2.0 + cutoff);
  };
  int32_t n = 0;
  uint64_t bits = 0;
  for (int32_t e = e1 - 1; e >= e0; --e) {
    const int32_t J = perm[e];
    if (J <= I) break;  // the rest of the bucket has smaller indices
    if (passes(e)) {
      if (e - e0 < 64) bits |= 1ull << (e - e0);
      ++n;
    }
  }
  // the workgroup's runs: inclusive scan in each wavefront, then across them
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int32_t x = n;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wv] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t tot = 0;
#pragma unroll
    for (int q = 0; q < kFilterThreads / 64; ++q) {
      const int32_t v = wsum[q];
      wsum[q] = tot;
      tot += v;
    }
    btot = tot;
    bcount[l] = tot <= cb ? tot : 0;
    if (tot > cb) atomicMax(need, tot);  // rare: the batch is repeated
  }
  __syncthreads();
  if (n == 0 || btot > cb) return;
  int64_t k = l * cb + wsum[wv] + (x - n);
  cstart[u] = (int32_t)k;
  if (e1 - e0 <= 64) {
    for (uint64_t m = bits; m; m &= m - 1, ++k)
      cand[k] = make_int2((int32_t)u, e0 + (int32_t)__builtin_ctzll(m));
  } else {
    for (int32_t e = e1 - 1; e >= e0; --e) {
      const int32_t J = perm[e];
      if (J <= I) break;
      if (passes(e)) cand[k++] = make_int2((int32_t)u, e);
    }
  }
}

// Every candidate through the reference's segment-segment test
// (CollisionDetector.cpp:215-237): accept[k] = 1 for a contact, and
// acnt[u] counts the contacts of slot u.  Candidates fill every lane, so
// the FP64 DCP runs on dense wavefronts.  Workgroup l takes region l of the
// list (the same XCD as the workgroup that filled it).
__global__ void __launch_bounds__(256)
k_pair_test(int S, int64_t gp, int32_t cb, const int32_t* __restrict__ bcount,
            const int2* __restrict__ cand, const double2* __restrict__ geo,
            const double2* __restrict__ sgeo, const int32_t* __restrict__ binst,
            const InstDev* __restrict__ idev, uint8_t* __restrict__ accept,
            int32_t* __restrict__ acnt) {
  const int64_t l = xcd_block(blockIdx.x, gp);
  const int32_t nb = bcount[l];
  for (int32_t j = threadIdx.x; j < nb; j += blockDim.x) {
    const int64_t k = l * cb + j;
    const int2 c = cand[k];
    const int64_t I = c.x / S;
    Seg a, b;
    seg_of(geo + 4 * I, a);
    seg_of(sgeo + 4 * (int64_t)c.y, b);
    DcpOut o;
    const double phi = pair_phi(a, b, o);
    const bool contact = !(phi > idev[binst[I]].cutoff);
    accept[k] = contact ? 1 : 0;
    if (contact) atomicAdd(&acnt[c.x], 1);
  }
}

// The record of every accepted candidate at its place: slot u's contacts
// start at aoff[u] and keep the run's (bucket) order.
__global__ void __launch_bounds__(256)
k_pair_write(int S, int64_t gp, int32_t cb, const int32_t* __restrict__ bcount,
             const int2* __restrict__ cand, const uint8_t* __restrict__ accept,
             const int32_t* __restrict__ cstart, const int32_t* __restrict__ aoff,
             const double2* __restrict__ geo, const double2* __restrict__ sgeo,
             const int32_t* __restrict__ perm, const int32_t* __restrict__ binst,
             const InstDev* __restrict__ idev, int32_t* __restrict__ outI,
             int32_t* __restrict__ outJ, double* __restrict__ rec, uint64_t* __restrict__ xkey,
             uint64_t* __restrict__ xtie, int32_t* __restrict__ xgrp) {
  const int64_t l = xcd_block(blockIdx.x, gp);
  const int32_t nb = bcount[l];
  for (int32_t j = threadIdx.x; j < nb; j += blockDim.x) {
    const int64_t k = l * cb + j;
    if (!accept[k]) continue;
    const int2 c = cand[k];
    int32_t rank = 0;
    for (int64_t q = cstart[c.x]; q < k; ++q) rank += accept[q];
    const int64_t slot = (int64_t)aoff[c.x] + rank;
    const int64_t I = c.x / S;
    Seg a, b;
    seg_of(geo + 4 * I, a);
    seg_of(sgeo + 4 * (int64_t)c.y, b);
    DcpOut o;
    const double phi = pair_phi(a, b, o);
    const int32_t J = perm[c.y];
    const int32_t inst = binst[I];
    outI[slot] = (int32_t)I;
    outJ[slot] = J;
    double* r = rec + R_LEN * slot;
    pair_record(a, b, o, phi, idev[inst].warm, r);
    // keys of the (colony, contact x, I, J) order
    xkey[slot] = order_key(r[R_P]);
    xtie[slot] = ((uint64_t)(uint32_t)I << 32) | (uint32_t)J;
    xgrp[slot] = inst;
  }
}

// Per-body record offsets from the per-(body, neighbour cell) offsets.
// Per-body record offsets from the per-(body, neighbour cell) offsets, and
// the batch offsets at the colony boundaries: colony b's records occupy
// [off[boff_b], off[boff_b + N_b]) of the batch.  Also clears the per-colony
// x-key bounds of the x sort that follows the writing pass.
__global__ void k_body_offsets(int64_t N, int S, int64_t B, const InstDev* __restrict__ idev,
                               const int32_t* __restrict__ slot_off, int32_t* __restrict__ off,
                               int32_t* __restrict__ bounds, unsigned long long* __restrict__ kmin,
                               unsigned long long* __restrict__ kmax) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= N) off[i] = slot_off[i * S];
  if (i <= B) bounds[i] = slot_off[(i < B ? idev[i].boff : N) * S];
  if (i < B) {
    kmin[i] = ~0ull;
    kmax[i] = 0ull;
  }
}

// Appends the sorted batch to the constraint set at offset C0 and builds the
// Jacobian blocks of the new rows (Physics.cpp:118-154).
__global__ void k_append(int64_t m, int64_t C0, int32_t iteration, const int32_t* __restrict__ idx,
                         const int32_t* __restrict__ rI_, const int32_t* __restrict__ rJ_,
                         const double* __restrict__ rec, const int32_t* __restrict__ binst,
                         const InstDev* __restrict__ idev, int32_t* __restrict__ Sinst,
                         int32_t* __restrict__ SI,
                         int32_t* __restrict__ SJ, int32_t* __restrict__ Sgid,
                         int32_t* __restrict__ Siter, double* __restrict__ Sphi,
                         double* __restrict__ Sgam, double* __restrict__ Sn,
                         double* __restrict__ SrI, double* __restrict__ SrJ,
                         double* __restrict__ Sp, double* __restrict__ SsI,
                         double* __restrict__ SsJ, double* __restrict__ coef64,
                         float* __restrict__ coef32, int32_t* __restrict__ nc) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int32_t src = idx[i];
  const double* r = rec + R_LEN * (int64_t)src;
  const int64_t c = C0 + i;
  const int32_t I = rI_[src], J = rJ_[src];
  const int32_t inst = binst[I];
  SI[c] = I;
  SJ[c] = J;
  Sinst[c] = inst;
  Sgid[c] = (int32_t)(i - idev[inst].bstart);  // index within the colony's batch
  Siter[c] = iteration;
  Sphi[c] = r[R_PHI];
  Sgam[c] = r[R_G];
  double n[3], ri[3], rj[3], nj[3], ti[3], tj[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    n[k] = r[R_N + k];
    ri[k] = r[R_RI + k];
    rj[k] = r[R_RJ + k];
    nj[k] = -n[k];
    Sn[3 * c + k] = n[k];
    SrI[3 * c + k] = ri[k];
    SrJ[3 * c + k] = rj[k];
    Sp[3 * c + k] = r[R_P + k];
  }
  cross3(ri, n, ti);
  cross3(rj, nj, tj);
  SsI[c] = r[R_SI];
  SsJ[c] = r[R_SJ];
  // packed operator row: D at I = [n, rI x n], D at J = [-n, rJ x -n], L = (sI, sJ)
  const double row[12] = {n[0],  n[1],  n[2],     ti[0],    ti[1], ti[2],
                          tj[0], tj[1], tj[2], r[R_SI], r[R_SJ], 0.0};
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    coef64[12 * c + k] = row[k];
    coef32[12 * c + k] = (float)row[k];
  }
  atomicAdd(&nc[I], 1);
  atomicAdd(&nc[J], 1);
}

// Colony-major merge at the end of a step: the set holds one block per append
// (a ReLCP phase of the colonies that ran it), each in colony order, so a
// constraint's destination is its position plus the offset of its (colony,
// phase) block.
__global__ void k_regroup_phases(int64_t total, const int32_t* __restrict__ inst,
                                 const int32_t* __restrict__ iter,
                                 const int64_t* __restrict__ offs, int32_t kmax,
                                 int64_t* __restrict__ dst) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  dst[g] = g + offs[(int64_t)inst[g] * kmax + iter[g]];
}

template <class T, int W>
__global__ void k_scatter_rows(int64_t total, const int64_t* __restrict__ dst,
                               const T* __restrict__ in, T* __restrict__ out) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const int64_t d = dst[g];
#pragma unroll
  for (int k = 0; k < W; ++k) out[W * d + k] = in[W * g + k];
}

inline unsigned nblk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// Records [0, min(m, cap)) of the sorted batch into SoA arrays in device
// memory (the pcc_constraints layout; a NULL array is skipped).
__global__ void k_records_out(int64_t m, const int32_t* __restrict__ idx,
                              const int32_t* __restrict__ rI_, const int32_t* __restrict__ rJ_,
                              const double* __restrict__ rec, pcc_constraints o) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m || i >= o.capacity) return;
  const int32_t src = idx[i];
  const double* r = rec + R_LEN * (int64_t)src;
  if (o.gidI) o.gidI[i] = rI_[src];
  if (o.gidJ) o.gidJ[i] = rJ_[src];
  if (o.signed_distance) o.signed_distance[i] = r[R_PHI];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (o.normal) o.normal[3 * i + k] = r[R_N + k];
    if (o.rPosI) o.rPosI[3 * i + k] = r[R_RI + k];
    if (o.rPosJ) o.rPosJ[3 * i + k] = r[R_RJ + k];
    if (o.contact_point) o.contact_point[3 * i + k] = r[R_P + k];
  }
  if (o.stressI) o.stressI[i] = r[R_SI];
  if (o.stressJ) o.stressJ[i] = r[R_SJ];
  if (o.gamma) o.gamma[i] = r[R_G];
  if (o.gid) o.gid[i] = (int32_t)i;
  if (o.iteration) o.iteration[i] = 0;
}

}  // namespace

void detect_records_out(const DetectBufs& d, int64_t m, const pcc_constraints& out_dev) {
  if (m > 0 && out_dev.capacity > 0)
    k_records_out<<<nblk(m), 256, 0, d.s>>>(m, d.ridx, d.rI, d.rJ, d.rec, out_dev);
}

template <class T>
static void excl_scan(pcc_ctx* ctx, const T* in, T* out, int64_t n) {
  size_t bytes = 0;
  PCC_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, T(0), (size_t)n, rocprim::plus<T>(),
                                  ctx->stream));
  ctx->sort_tmp.ensure((int64_t)bytes + 16);
  PCC_HIP(rocprim::exclusive_scan(ctx->sort_tmp.p, bytes, in, out, T(0), (size_t)n,
                                  rocprim::plus<T>(), ctx->stream));
}

void exclusive_scan_i32(pcc_ctx* ctx, const int32_t* in, int32_t* out, int64_t n) {
  excl_scan<int32_t>(ctx, in, out, n);
}

void seg_sort_i32(pcc_ctx* ctx, int64_t nseg, const int32_t* start, int32_t* v) {
  if (nseg > 0) k_seg_sort<<<nblk(nseg), 256, 0, ctx->stream>>>(nseg, start, v);
}

int64_t table_size_for(int64_t n) {
  int64_t t = 1024;
  while (t < 2 * n) t <<= 1;
  return t;
}

namespace {

__global__ void k_fill_i32(int64_t n, int32_t v, int32_t* __restrict__ a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// rocPRIM calls on caller-provided temporary storage: the required size is
// queried first and must fit.
void scan_on(DetectBufs& d, const int32_t* in, int32_t* out, int64_t n) {
  size_t bytes = 0;
  PCC_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, int32_t(0), (size_t)n,
                                  rocprim::plus<int32_t>(), d.s));
  if (bytes > d.tmp_bytes) throw StatusError{PCC_E_DEVICE, "internal: scan storage too small"};
  PCC_HIP(rocprim::exclusive_scan(d.tmp, bytes, in, out, int32_t(0), (size_t)n,
                                  rocprim::plus<int32_t>(), d.s));
}

}  // namespace

size_t detect_scan_bytes(int64_t N, int S, int64_t T) {
  size_t a = 0, b = 0;
  int32_t* p = nullptr;
  PCC_HIP(rocprim::exclusive_scan(nullptr, a, p, p, int32_t(0), (size_t)(T + 1),
                                  rocprim::plus<int32_t>(), (hipStream_t)0));
  PCC_HIP(rocprim::exclusive_scan(nullptr, b, p, p, int32_t(0), (size_t)(N * S + 1),
                                  rocprim::plus<int32_t>(), (hipStream_t)0));
  return std::max(a, b);
}

int64_t filter_grid(int64_t N, int S) {
  return pad8((N * S + kFilterThreads - 1) / kFilterThreads);
}

size_t detect_sort_bytes(int64_t m, int64_t B) {
  (void)B;
  return xsort_tmp_bytes(m);
}

// Stage 1 of a detection batch: the hashed cell list (a stable counting sort
// of salted cell hashes), the candidate pairs, their contact test and the
// offsets.  bounds_host[b] (B + 2 entries) receives the first record of
// colony b in the batch; bounds_host[B] is the record count and
// bounds_host[B + 1] is 0, or the candidate region size that some workgroup
// needed beyond d.cand_block: then the function returns false and nothing
// else is valid.
bool detect_count(DetectBufs& d, int32_t* bounds_host) {
  const int64_t N = d.N, T = d.T, NS = N * d.S;
  const hipStream_t s = d.s;
  const uint32_t mask = (uint32_t)(T - 1);
  if (!d.cursor_zero) k_fill_i32<<<nblk(T + 1), 256, 0, s>>>(T + 1, 0, d.bcursor);
  if (N > 0)
    k_cells<<<nblk(N), 256, 0, s>>>(N, d.body, d.binst, d.idev, mask, d.S == 9, d.cell4, d.geo,
                                    d.bucket,
                                    d.bcursor);
  scan_on(d, d.bcursor, d.bstart, T + 1);
  // candidates, then their test; dslot counts each slot's contacts
  int32_t* need = d.bbounds + d.B + 1;
  if (N > 0) {
    k_scatter<<<nblk(N), 256, 0, s>>>(N, d.bucket, d.bstart, d.bcursor, d.perm);
    k_seg_sort<<<nblk(T), 256, 0, s>>>(T, d.bstart, d.perm);
    k_stage<<<nblk(N), 256, 0, s>>>(N, d.S, d.perm, d.cell4, d.geo, d.idev, d.scell, d.sgeo,
                                    d.frec, d.dslot, need);
  } else {
    k_fill_i32<<<1, 64, 0, s>>>(1, 0, need);
    k_fill_i32<<<1, 64, 0, s>>>(1, 0, d.dslot);
  }
  if (N > 0) {
    const int64_t gp = filter_grid(N, d.S);
    if (d.S == 9)
      k_filter<9><<<(unsigned)gp, kFilterThreads, 0, s>>>(
          N, gp, d.scell, d.sgeo, d.frec, d.bstart, d.perm, mask, d.idev, d.cand_block,
          d.bcount, need, d.cstart, d.cand);
    else
      k_filter<27><<<(unsigned)gp, kFilterThreads, 0, s>>>(
          N, gp, d.scell, d.sgeo, d.frec, d.bstart, d.perm, mask, d.idev, d.cand_block,
          d.bcount, need, d.cstart, d.cand);
    k_pair_test<<<(unsigned)gp, 256, 0, s>>>(d.S, gp, d.cand_block, d.bcount, d.cand, d.geo,
                                             d.sgeo, d.binst, d.idev, d.accept, d.dslot);
  }
  // per-slot offsets, then per-body offsets into dcount, then the colony
  // boundaries of the batch (bodies are stored colony after colony, and a
  // stable sort by colony keeps them); bounds[B] is the batch's record count
  // and bounds[B + 1] the candidate count, so one read-back gives all
  scan_on(d, d.dslot, d.dslot_off, NS + 1);
  k_body_offsets<<<nblk(std::max<int64_t>(N, d.B) + 1), 256, 0, s>>>(
      N, d.S, d.B, d.idev, d.dslot_off, d.dcount, d.bbounds, d.xs_kmin, d.xs_kmax);
  PCC_HIP(hipMemcpyAsync(bounds_host, d.bbounds, sizeof(int32_t) * (d.B + 2),
                         hipMemcpyDeviceToHost, s));
  PCC_HIP(hipStreamSynchronize(s));
  return bounds_host[d.B + 1] == 0;
}

// Stage 2: the writing pass of the m records found by detect_count, then
// their order (colony, contact x, I, J); on return d.ridx lists them in that
// order.
void detect_write(DetectBufs& d, int64_t m) {
  if (m == 0) return;
  const int64_t N = d.N;
  const hipStream_t s = d.s;
  const int64_t gp = filter_grid(N, d.S);
  // ridx2: each record's colony (then its bucket); rkey / rkey2: its x key
  // and (I, J) tie-break
  k_pair_write<<<(unsigned)gp, 256, 0, s>>>(d.S, gp, d.cand_block, d.bcount, d.cand, d.accept,
                                            d.cstart, d.dslot_off, d.geo, d.sgeo, d.perm,
                                            d.binst, d.idev, d.rI, d.rJ, d.rec, d.rkey, d.rkey2,
                                            d.ridx2);
  XSort x;
  x.n = m;
  x.G = d.B;
  x.gstart = d.bbounds;  // records are colony-major (bodies are, and records are I-major)
  x.key = d.rkey;
  x.tie = d.rkey2;
  x.grp = d.ridx2;
  x.cnt = d.xs_cnt;
  x.off = d.xs_off;
  x.kmin = d.xs_kmin;
  x.kmax = d.xs_kmax;
  x.bounds_clear = true;  // by k_body_offsets
  x.perm = d.ridx;
  x.tmp = d.tmp;
  x.tmp_bytes = d.tmp_bytes;
  x.s = s;
  xsort_run(x);
}

// Builds the hashed cell list of the current body state (stable counting sort).
static DetectBufs ctx_detect_bufs(pcc_ctx* ctx) {
  const int64_t N = ctx->N;
  const int64_t T = table_size_for(N);
  const int S = ctx->detect_flat ? 9 : 27;
  const int64_t NS = N * S;
  const int64_t B = (int64_t)ctx->inst.size();
  ctx->table_size = T;
  ctx->cell.ensure(4 * N + 4);
  ctx->dgeo.ensure(8 * N + 8);
  ctx->scell.ensure(4 * N + 4);
  ctx->sgeo.ensure(8 * N + 8);
  ctx->sfrec.ensure(4 * N + 4);
  ctx->bucket.ensure(N + 1);
  ctx->perm.ensure(N + 1);
  ctx->bstart.ensure(T + 1);
  ctx->bcursor.ensure(T + 1);
  ctx->dslot.ensure(NS + 1);
  ctx->dslot_off.ensure(NS + 1);
  ctx->dcstart.ensure(NS + 1);
  // candidate regions: a workgroup of 256 (body, cell) slots finds ~100
  // candidates in dense 2D packings; regions grow when one overflows
  const int64_t gp = filter_grid(N, S);
  if (ctx->cand_block == 0) ctx->cand_block = kCandBlock;
  ctx->dcand.ensure(2 * gp * ctx->cand_block + 2);
  ctx->daccept.ensure(gp * ctx->cand_block + 1);
  ctx->dbcount.ensure(gp + 1);
  ctx->dcount.ensure(N + 1);
  ctx->bbounds.ensure(B + 2);
  ctx->xs_bounds.ensure(2 * B);
  ctx->sort_tmp.ensure((int64_t)detect_scan_bytes(N, S, T) + 16);
  DetectBufs d;
  d.N = N;
  d.body = body_arrays(ctx->body.p, N);
  d.binst = ctx->binst.p;
  d.idev = ctx->idev.p;
  d.B = B;
  d.S = S;
  d.T = T;
  d.cell4 = reinterpret_cast<int4*>(ctx->cell.p);
  d.geo = reinterpret_cast<double2*>(ctx->dgeo.p);
  d.scell = reinterpret_cast<int4*>(ctx->scell.p);
  d.sgeo = reinterpret_cast<double2*>(ctx->sgeo.p);
  d.frec = reinterpret_cast<float4*>(ctx->sfrec.p);
  d.bucket = ctx->bucket.p;
  d.perm = ctx->perm.p;
  d.bstart = ctx->bstart.p;
  d.bcursor = ctx->bcursor.p;
  d.dslot = ctx->dslot.p;
  d.dslot_off = ctx->dslot_off.p;
  d.cstart = ctx->dcstart.p;
  d.cand = reinterpret_cast<int2*>(ctx->dcand.p);
  d.accept = ctx->daccept.p;
  d.bcount = ctx->dbcount.p;
  d.cand_block = ctx->cand_block;
  d.dcount = ctx->dcount.p;
  d.bbounds = ctx->bbounds.p;
  d.xs_kmin = ctx->xs_bounds.p;
  d.xs_kmax = ctx->xs_bounds.p + B;
  d.cursor_zero = ctx->bcursor_zero.known(ctx->bcursor.p, T + 1);
  d.tmp = ctx->sort_tmp.p;
  d.tmp_bytes = (size_t)ctx->sort_tmp.cap;
  d.s = ctx->stream;
  return d;
}

// One detection batch over the active colonies, each at the cutoff of its
// instance record; on return ctx->ridx[0..batch_count) lists the raw records
// in (colony, contact x, I, J) order and every colony's m / bstart is set.
void detect_batch(pcc_ctx* ctx) {
  const int64_t B = (int64_t)ctx->inst.size();
  KTimer tm(ctx, PCC_K_DETECT, 0.0);
  DetectBufs d = ctx_detect_bufs(ctx);
  int32_t* bounds = readback_pin(ctx, B + 2);
  ctx->bcursor_zero.clear();
  while (!detect_count(d, bounds)) {
    // a workgroup found more candidates than its region holds: larger
    // regions, and the batch again
    int32_t cb = ctx->cand_block;
    while (cb < bounds[B + 1]) cb *= 2;
    ctx->cand_block = cb;
    d = ctx_detect_bufs(ctx);
  }
  // k_scatter consumed the bucket counts back to zero
  ctx->bcursor_zero.mark(ctx->bcursor.p, d.T + 1);
  const int64_t m = bounds[B];
  ctx->batch_count = m;
  ctx->rI_.ensure(m + 1);
  ctx->rJ_.ensure(m + 1);
  ctx->rrec.ensure(R_LEN * (m + 1));
  ctx->ridx.ensure(m + 1);
  ctx->ridx2.ensure(m + 1);
  ctx->rkey.ensure(m + 1);
  ctx->rkey2.ensure(m + 1);
  ctx->xs_cnt.ensure(m + 1);
  ctx->xs_off.ensure(m + 1);
  ctx->sort_tmp.ensure((int64_t)detect_sort_bytes(m, B) + 16);
  for (int64_t q = 0; q < B; ++q) {
    ctx->inst[q].bstart = bounds[q];
    ctx->inst[q].m = bounds[q + 1] - bounds[q];
  }
  if (m == 0) return;
  d.rI = ctx->rI_.p;
  d.rJ = ctx->rJ_.p;
  d.rec = ctx->rrec.p;
  d.ridx = ctx->ridx.p;
  d.ridx2 = ctx->ridx2.p;
  d.rkey = ctx->rkey.p;
  d.rkey2 = ctx->rkey2.p;
  d.xs_cnt = ctx->xs_cnt.p;
  d.xs_off = ctx->xs_off.p;
  d.tmp = ctx->sort_tmp.p;
  d.tmp_bytes = (size_t)ctx->sort_tmp.cap;
  detect_write(d, m);
}

static void ensure_set_capacity(ConstraintSet& S, int64_t need, hipStream_t s) {
  if (need <= S.cap) return;
  int64_t nc = S.cap > 0 ? S.cap : 4096;
  while (nc < need) nc *= 2;
  const int64_t keep = S.C;
  S.inst.grow_keep(nc, keep, s);
  S.I.grow_keep(nc, keep, s);
  S.J.grow_keep(nc, keep, s);
  S.gid.grow_keep(nc, keep, s);
  S.iter.grow_keep(nc, keep, s);
  S.phi.grow_keep(nc, keep, s);
  S.gamma.grow_keep(nc, keep, s);
  S.sI.grow_keep(nc, keep, s);
  S.sJ.grow_keep(nc, keep, s);
  S.n.grow_keep(3 * nc, 3 * keep, s);
  S.rI.grow_keep(3 * nc, 3 * keep, s);
  S.rJ.grow_keep(3 * nc, 3 * keep, s);
  S.p.grow_keep(3 * nc, 3 * keep, s);
  S.coef64.grow_keep(12 * nc, 12 * keep, s);
  S.coef32.grow_keep(12 * nc, 12 * keep, s);
  S.cap = nc;
}

template <class T>
static void swap_arr(DArr<T>& a, DArr<T>& b) {
  a.swap(b);
}

// Appends the sorted batch at the end of the set (merge_set puts the set in
// colony-major order at the end of the step).  The instance records must
// hold the colonies' current counts, m and bstart; on return their ranges
// (coff, C) are the colony-major ones, which the solver order follows.
void append_batch(pcc_ctx* ctx, int32_t iteration) {
  const int64_t m = ctx->batch_count;
  const int64_t B = (int64_t)ctx->inst.size();
  ConstraintSet& S = ctx->S;
  hipStream_t s = ctx->stream;
  ensure_set_capacity(S, S.C + m + 1, s);
  int64_t off = 0;
  for (auto& q : ctx->inst) {
    q.coff_new_ = off;
    off += q.C + q.m;
  }
  upload_instances(ctx);
  const int64_t C0 = S.C;
  if (m > 0) {
    KTimer tm(ctx, PCC_K_ASSEMBLE, 0.0);
    k_append<<<nblk(m), 256, 0, s>>>(
        m, C0, iteration, ctx->ridx.p, ctx->rI_.p, ctx->rJ_.p, ctx->rrec.p, ctx->binst.p,
        ctx->idev.p, S.inst.p, S.I.p, S.J.p, S.gid.p, S.iter.p, S.phi.p, S.gamma.p, S.n.p,
        S.rI.p, S.rJ.p, S.p.p, S.sI.p, S.sJ.p, S.coef64.p, S.coef32.p, ctx->nc.p);
  }
  S.C = C0 + m;
  // The set is left in append order: one block per append, each in colony
  // order.  Within a colony that is the colony-major order (phase-major, then
  // batch order), and nothing during the step depends on where the other
  // colonies' constraints sit (the solver order is sorted by body index and
  // a body's incidences by set index, and both keep a colony's relative
  // order), so the merge into colony-major ranges is done once, at the end of
  // the step (merge_set).  It moves constraints only if some colony after the
  // first one with new constraints already holds constraints (not in phase
  // 0, a soft step, or a colony stepped alone in a batch).
  {
    bool appended = false;
    for (const auto& q : ctx->inst) {
      if (appended && q.C > 0) ctx->set_unmerged = true;
      if (q.m > 0) appended = true;
    }
  }
  for (size_t b = 0; b < ctx->inst.size(); ++b) {
    const auto& q = ctx->inst[b];
    if (q.m > 0) ctx->append_log.push_back({(int32_t)b, iteration, C0 + q.bstart, q.m});
  }
  for (auto& q : ctx->inst) {
    q.coff = q.coff_new_;
    q.C += q.m;
  }
  upload_instances(ctx);
}

// The colony-major merge of the step's constraint set (see append_batch):
// colony b's phase-k block moves to coff_b + (its earlier phases' counts).
void merge_set(pcc_ctx* ctx) {
  const int64_t B = (int64_t)ctx->inst.size();
  ConstraintSet& S = ctx->S;
  if (!ctx->set_unmerged || B < 2 || S.C == 0) {
    ctx->set_unmerged = false;
    return;
  }
  hipStream_t s = ctx->stream;
  int32_t kmax = 1;
  for (const auto& e : ctx->append_log) kmax = std::max(kmax, e.phase + 1);
  auto& offs = ctx->merge_offs_host;
  offs.assign((size_t)B * kmax, 0);
  std::vector<int64_t> next(B);
  for (int64_t b = 0; b < B; ++b) next[b] = ctx->inst[b].coff;
  // the log is in append order, and a colony's phases are appended in order
  for (const auto& e : ctx->append_log) {
    offs[(size_t)e.colony * kmax + e.phase] = next[e.colony] - e.start;
    next[e.colony] += e.count;
  }
  const int64_t T = S.C;
  ctx->merge_offs.ensure((int64_t)offs.size());
  PCC_HIP(hipMemcpyAsync(ctx->merge_offs.p, offs.data(), sizeof(int64_t) * offs.size(),
                         hipMemcpyHostToDevice, s));
  ctx->regroup_dst.ensure(T + 1);
  k_regroup_phases<<<nblk(T), 256, 0, s>>>(T, S.inst.p, S.iter.p, ctx->merge_offs.p, kmax,
                                           ctx->regroup_dst.p);
  ConstraintSet& D = ctx->S2;
  ensure_set_capacity(D, T + 1, s);
  const int64_t* dst = ctx->regroup_dst.p;
  const unsigned g = nblk(T);
  k_scatter_rows<int32_t, 1><<<g, 256, 0, s>>>(T, dst, S.inst.p, D.inst.p);
  k_scatter_rows<int32_t, 1><<<g, 256, 0, s>>>(T, dst, S.I.p, D.I.p);
  k_scatter_rows<int32_t, 1><<<g, 256, 0, s>>>(T, dst, S.J.p, D.J.p);
  k_scatter_rows<int32_t, 1><<<g, 256, 0, s>>>(T, dst, S.gid.p, D.gid.p);
  k_scatter_rows<int32_t, 1><<<g, 256, 0, s>>>(T, dst, S.iter.p, D.iter.p);
  k_scatter_rows<double, 1><<<g, 256, 0, s>>>(T, dst, S.phi.p, D.phi.p);
  k_scatter_rows<double, 1><<<g, 256, 0, s>>>(T, dst, S.gamma.p, D.gamma.p);
  k_scatter_rows<double, 1><<<g, 256, 0, s>>>(T, dst, S.sI.p, D.sI.p);
  k_scatter_rows<double, 1><<<g, 256, 0, s>>>(T, dst, S.sJ.p, D.sJ.p);
  k_scatter_rows<double, 3><<<g, 256, 0, s>>>(T, dst, S.n.p, D.n.p);
  k_scatter_rows<double, 3><<<g, 256, 0, s>>>(T, dst, S.rI.p, D.rI.p);
  k_scatter_rows<double, 3><<<g, 256, 0, s>>>(T, dst, S.rJ.p, D.rJ.p);
  k_scatter_rows<double, 3><<<g, 256, 0, s>>>(T, dst, S.p.p, D.p.p);
  k_scatter_rows<double, 12><<<g, 256, 0, s>>>(T, dst, S.coef64.p, D.coef64.p);
  k_scatter_rows<float, 12><<<g, 256, 0, s>>>(T, dst, S.coef32.p, D.coef32.p);
  swap_arr(S.inst, D.inst);
  swap_arr(S.I, D.I);
  swap_arr(S.J, D.J);
  swap_arr(S.gid, D.gid);
  swap_arr(S.iter, D.iter);
  swap_arr(S.phi, D.phi);
  swap_arr(S.gamma, D.gamma);
  swap_arr(S.sI, D.sI);
  swap_arr(S.sJ, D.sJ);
  swap_arr(S.n, D.n);
  swap_arr(S.rI, D.rI);
  swap_arr(S.rJ, D.rJ);
  swap_arr(S.p, D.p);
  swap_arr(S.coef64, D.coef64);
  swap_arr(S.coef32, D.coef32);
  std::swap(S.cap, D.cap);
  D.C = 0;
  ctx->set_unmerged = false;
}

// Copies the current (sorted) detection batch to host arrays.
void batch_to_host(pcc_ctx* ctx, pcc_constraints* out, int32_t iteration) {
  const int64_t m = ctx->batch_count;
  if (!out || m == 0) return;
  std::vector<int32_t> idx(m), I(m), J(m);
  std::vector<double> rec(R_LEN * m);
  hipStream_t s = ctx->stream;
  PCC_HIP(hipMemcpyAsync(idx.data(), ctx->ridx.p, sizeof(int32_t) * m, hipMemcpyDeviceToHost, s));
  PCC_HIP(hipMemcpyAsync(I.data(), ctx->rI_.p, sizeof(int32_t) * m, hipMemcpyDeviceToHost, s));
  PCC_HIP(hipMemcpyAsync(J.data(), ctx->rJ_.p, sizeof(int32_t) * m, hipMemcpyDeviceToHost, s));
  PCC_HIP(hipMemcpyAsync(rec.data(), ctx->rrec.p, sizeof(double) * R_LEN * m,
                         hipMemcpyDeviceToHost, s));
  PCC_HIP(hipStreamSynchronize(s));
  for (int64_t i = 0; i < m && i < out->capacity; ++i) {
    const int32_t src = idx[i];
    const double* r = rec.data() + R_LEN * (int64_t)src;
    if (out->gidI) out->gidI[i] = I[src];
    if (out->gidJ) out->gidJ[i] = J[src];
    if (out->signed_distance) out->signed_distance[i] = r[R_PHI];
    for (int k = 0; k < 3; ++k) {
      if (out->normal) out->normal[3 * i + k] = r[R_N + k];
      if (out->rPosI) out->rPosI[3 * i + k] = r[R_RI + k];
      if (out->rPosJ) out->rPosJ[3 * i + k] = r[R_RJ + k];
      if (out->contact_point) out->contact_point[3 * i + k] = r[R_P + k];
    }
    if (out->stressI) out->stressI[i] = r[R_SI];
    if (out->stressJ) out->stressJ[i] = r[R_SJ];
    if (out->gamma) out->gamma[i] = r[R_G];
    if (out->gid) out->gid[i] = (int32_t)i;
    if (out->iteration) out->iteration[i] = iteration;
  }
}

}  // namespace pcc
