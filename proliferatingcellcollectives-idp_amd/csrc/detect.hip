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
//  3. k_scatter    bodies into buckets, then k_bucket_sort orders each bucket
//                  by body index: a stable counting sort, identical to the
//                  CPU restatement's
//  4. k_detect     one thread per (body I, neighbour cell) tests every J > I
//                  of that cell with the reference's filters and exact FP64
//                  segment-segment DCP; 27 neighbour cells, or the 9 of the
//                  body's own z-layer when every body has z = 0 (the other
//                  cells are then empty); pass 0 counts and keeps, per thread,
//                  a bit mask of the accepted positions of its bucket; pass 1
//                  re-evaluates only those (buckets of at most 64 bodies;
//                  longer ones are walked again) and writes at the scanned
//                  offset, so each body's records keep the cell order dz, dy,
//                  dx and the bucket order
//  5. k_segsort + radix sort: order = (contact x, I, J); with several colonies
//                  in one context a second stable radix sort by colony makes
//                  the order (colony, contact x, I, J)
//
// Colonies of one context share the hashed table: the colony index salts the
// bucket hash and neighbour tests require the same colony, so each colony's
// contact list is the one it would have alone.
#include <algorithm>
#include <cstring>
#include <vector>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "common.hpp"
#include "internal.hpp"

namespace pcc {

namespace {

// Per body: its cell and colony (cell4), and the detection geometry record
// geo = (x, y, z, t_x | t_y, t_z, length, diameter) with t = rotate(q, e_x), so
// that the detection kernel reads a candidate from two cache lines.
__global__ void k_cells(int64_t N, BodyArrays b, const int32_t* __restrict__ binst,
                        const InstDev* __restrict__ idev, uint32_t mask,
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
  uint32_t bk = (cell_hash(cx, cy, cz) ^ inst_mix(inst)) & mask;
  bucket[i] = bk;
  atomicAdd(&count[bk], 1);
}

__global__ void k_scatter(int64_t N, const uint32_t* __restrict__ bucket,
                          const int32_t* __restrict__ start, int32_t* __restrict__ cursor,
                          int32_t* __restrict__ perm) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  uint32_t bk = bucket[i];
  int32_t pos = start[bk] + atomicAdd(&cursor[bk], 1);
  perm[pos] = (int32_t)i;
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

template <int S>  // neighbour cells per body: 9 (one z-layer) or 27
__global__ void __launch_bounds__(256)
k_detect(int pass, int64_t N, const int4* __restrict__ cell4, const double2* __restrict__ geo,
         const int32_t* __restrict__ bstart, const int32_t* __restrict__ perm, uint32_t mask,
         const InstDev* __restrict__ idev, int32_t* __restrict__ cnt_off,
         uint64_t* __restrict__ acc, int32_t* __restrict__ outI, int32_t* __restrict__ outJ,
         double* __restrict__ rec) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N * S) return;
  const int64_t I = t / S;
  const int sl = (int)(t - I * S);
  const int dx = sl % 3 - 1, dy = (sl / 3) % 3 - 1, dz = S == 27 ? sl / 9 - 1 : 0;
  const int4 cI = cell4[I];
  const int32_t inst = cI.w;
  if (!(idev[inst].flags & kInstActive)) return;  // count stays 0
  const double cutoff = idev[inst].cutoff, warm = idev[inst].warm;
  const uint32_t salt = inst_mix(inst);
  const double2 gI0 = geo[4 * I], gI1 = geo[4 * I + 1], gI2 = geo[4 * I + 2], gI3 = geo[4 * I + 3];
  const double xI[3] = {gI0.x, gI0.y, gI1.x};
  const double tI[3] = {gI1.y, gI2.x, gI2.y};
  const double lenI = gI3.x, dI = gI3.y;
  const double halfI = lenI / 2 - dI / 2;
  double P0[3], P1[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double hv = tI[k] * halfI;
    P0[k] = xI[k] - hv;
    P1[k] = xI[k] + hv;
  }
  const int32_t cx = cI.x, cy = cI.y, cz = cI.z;
  const int32_t w = pass ? cnt_off[t] : 0;
  int32_t c = 0;
  const int32_t nx = cx + dx, ny = cy + dy, nz = cz + dz;
  const uint32_t bk = (cell_hash(nx, ny, nz) ^ salt) & mask;
  const int32_t e0 = bstart[bk], e1 = bstart[bk + 1];
  // The reference's filters and the record of candidate perm[e]; true if it
  // is a contact (then, in pass 1, written at slot w + c).
  auto candidate = [&](int32_t e) -> bool {
          const int32_t J = perm[e];
          if (J <= I) return false;
          const int4 cJ = cell4[J];
          if (cJ.x != nx || cJ.y != ny || cJ.z != nz || cJ.w != inst) return false;
          const double2 g0 = geo[4 * (int64_t)J], g1 = geo[4 * (int64_t)J + 1];
          const double2 g2 = geo[4 * (int64_t)J + 2], g3 = geo[4 * (int64_t)J + 3];
          const double xJ[3] = {g0.x, g0.y, g1.x};
          const double lenJ = g3.x, dJ = g3.y;
          double dd[3] = {xI[0] - xJ[0], xI[1] - xJ[1], xI[2] - xJ[2]};
          double rough = sqrt(dot3(dd, dd));
          if (rough > (lenI + lenJ) / 2.0 + cutoff) return false;
          const double tJ[3] = {g1.y, g2.x, g2.y};
          const double halfJ = lenJ / 2 - dJ / 2;
          double Q0[3], Q1[3];
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            double hv = tJ[k] * halfJ;
            Q0[k] = xJ[k] - hv;
            Q1[k] = xJ[k] + hv;
          }
          DcpOut o;
          dcp(P0, P1, Q0, Q1, o);
          const double sum_radii = (dI + dJ) / 2.0;
          const double phi = o.dist - sum_radii;
          if (phi > cutoff) return false;
          if (pass) {
            const int64_t slot = (int64_t)w + c;
            outI[slot] = (int32_t)I;
            outJ[slot] = J;
            double* r = rec + R_LEN * slot;
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
              r[R_RI + k] = p[k] - xI[k];
              r[R_RJ + k] = p[k] - xJ[k];
              r[R_P + k] = p[k];
            }
            r[R_SI] = 0.5 * fabs(dot3(nrm, tI));
            r[R_SJ] = 0.5 * fabs(dot3(mn, tJ));
            r[R_G] = phi >= 0.0 ? 0.0 : -phi * warm;
          }
          return true;
  };
  if (pass && e1 - e0 <= 64) {
    // the contacts pass 0 found, in ascending bucket position
    for (uint64_t m = acc[t]; m; m &= m - 1)
      if (candidate(e0 + (int32_t)__builtin_ctzll(m))) ++c;
  } else {
    uint64_t bits = 0;
    for (int32_t e = e0; e < e1; ++e)
      if (candidate(e)) {
        if (e - e0 < 64) bits |= 1ull << (e - e0);
        ++c;
      }
    if (!pass) {
      cnt_off[t] = c;
      acc[t] = bits;
    }
  }
}

// Per-body record offsets from the per-(body, neighbour cell) offsets.
__global__ void k_body_offsets(int64_t N, int S, const int32_t* __restrict__ slot_off,
                               int32_t* __restrict__ off) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= N) off[i] = slot_off[i * S];
}

// Per body segment of the batch: order its records by J (insertion sort of
// record indices).
__global__ void k_segsort_by_J(int64_t N, const int32_t* __restrict__ off,
                               const int32_t* __restrict__ J, int32_t* __restrict__ idx) {
  int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= N) return;
  int32_t a = off[I], e = off[I + 1];
  for (int32_t i = a; i < e; ++i) idx[i] = i;
  for (int32_t i = a + 1; i < e; ++i) {
    int32_t x = idx[i];
    int32_t jx = J[x];
    int32_t j = i - 1;
    while (j >= a && J[idx[j]] > jx) {
      idx[j + 1] = idx[j];
      --j;
    }
    idx[j + 1] = x;
  }
}

__global__ void k_make_keys(int64_t m, const int32_t* __restrict__ idx,
                           const double* __restrict__ rec, uint64_t* __restrict__ key) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  key[i] = order_key(rec[R_LEN * (int64_t)idx[i] + R_P]);
}

// Colony of each sorted batch record (key of the second, stable sort).
__global__ void k_inst_keys(int64_t m, const int32_t* __restrict__ idx,
                            const int32_t* __restrict__ rI_, const int32_t* __restrict__ binst,
                            uint32_t* __restrict__ key) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  key[i] = (uint32_t)binst[rI_[idx[i]]];
}

// Batch offsets at the colony boundaries: colony b's records occupy
// [off[boff_b], off[boff_b + N_b]) of the sorted batch.
__global__ void k_batch_bounds(int64_t B, int64_t N, const InstDev* __restrict__ idev,
                               const int32_t* __restrict__ off, int32_t* __restrict__ out) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b > B) return;
  out[b] = off[b < B ? idev[b].boff : N];
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

// Instance-major merge: constraint g of the set (old entries first, then the
// batch appended at C0) moves to its colony's new range, old constraints first.
__global__ void k_regroup(int64_t total, int64_t C0, const InstDev* __restrict__ idev,
                          const int32_t* __restrict__ inst_in, int64_t* __restrict__ dst) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const InstDev& D = idev[inst_in[g]];
  dst[g] = g < C0 ? g - D.coff + D.coff_new : D.coff_new + D.C + (g - C0 - D.bstart);
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

template <class K>
void sort_pairs_on(DetectBufs& d, K* k0, K* k1, int32_t* v0, int32_t* v1, int64_t m, int bits) {
  size_t bytes = 0;
  PCC_HIP(rocprim::radix_sort_pairs(nullptr, bytes, k0, k1, v0, v1, (size_t)m, 0, bits, d.s));
  if (bytes > d.tmp_bytes) throw StatusError{PCC_E_DEVICE, "internal: sort storage too small"};
  PCC_HIP(rocprim::radix_sort_pairs(d.tmp, bytes, k0, k1, v0, v1, (size_t)m, 0, bits, d.s));
}

int colony_bits(int64_t B) {
  int bits = 1;
  while (((int64_t)1 << bits) < B) ++bits;
  return bits;
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

size_t detect_sort_bytes(int64_t m, int64_t B) {
  size_t a = 0, b = 0;
  uint64_t* k = nullptr;
  uint32_t* k32 = nullptr;
  int32_t* v = nullptr;
  PCC_HIP(rocprim::radix_sort_pairs(nullptr, a, k, k, v, v, (size_t)m, 0, 64, (hipStream_t)0));
  if (B > 1)
    PCC_HIP(rocprim::radix_sort_pairs(nullptr, b, k32, k32, v, v, (size_t)m, 0, colony_bits(B),
                                      (hipStream_t)0));
  return std::max(a, b);
}

// Stage 1 of a detection batch: the hashed cell list (a stable counting sort
// of salted cell hashes), the counting pass over every (body, neighbour cell)
// slot, and the offsets.  bounds_host[b] (B + 1 entries) receives the first
// record of colony b in the batch; bounds_host[B] is the record count.
void detect_count(DetectBufs& d, int32_t* bounds_host) {
  const int64_t N = d.N, T = d.T, NS = N * d.S;
  const hipStream_t s = d.s;
  const uint32_t mask = (uint32_t)(T - 1);
  k_fill_i32<<<nblk(T + 1), 256, 0, s>>>(T + 1, 0, d.bcursor);
  if (N > 0)
    k_cells<<<nblk(N), 256, 0, s>>>(N, d.body, d.binst, d.idev, mask, d.cell4, d.geo, d.bucket,
                                    d.bcursor);
  scan_on(d, d.bcursor, d.bstart, T + 1);
  k_fill_i32<<<nblk(T + 1), 256, 0, s>>>(T + 1, 0, d.bcursor);
  if (N > 0) {
    k_scatter<<<nblk(N), 256, 0, s>>>(N, d.bucket, d.bstart, d.bcursor, d.perm);
    k_seg_sort<<<nblk(T), 256, 0, s>>>(T, d.bstart, d.perm);
  }
  k_fill_i32<<<nblk(NS + 1), 256, 0, s>>>(NS + 1, 0, d.dslot);
  if (N > 0) {
    if (d.S == 9)
      k_detect<9><<<nblk(NS), 256, 0, s>>>(0, N, d.cell4, d.geo, d.bstart, d.perm, mask, d.idev,
                                           d.dslot, d.dacc, nullptr, nullptr, nullptr);
    else
      k_detect<27><<<nblk(NS), 256, 0, s>>>(0, N, d.cell4, d.geo, d.bstart, d.perm, mask, d.idev,
                                            d.dslot, d.dacc, nullptr, nullptr, nullptr);
  }
  // per-slot offsets, then per-body offsets into dcount, then the colony
  // boundaries of the batch (bodies are stored colony after colony, and a
  // stable sort by colony keeps them); bounds[B] is the batch's record count,
  // so one read-back gives both
  scan_on(d, d.dslot, d.dslot_off, NS + 1);
  k_body_offsets<<<nblk(N + 1), 256, 0, s>>>(N, d.S, d.dslot_off, d.dcount);
  k_batch_bounds<<<nblk(d.B + 1), 256, 0, s>>>(d.B, N, d.idev, d.dcount, d.bbounds);
  PCC_HIP(hipMemcpyAsync(bounds_host, d.bbounds, sizeof(int32_t) * (d.B + 1),
                         hipMemcpyDeviceToHost, s));
  PCC_HIP(hipStreamSynchronize(s));
}

// Stage 2: the writing pass of the m records found by detect_count, then
// their order (colony, contact x, I, J); on return d.ridx lists them in that
// order (d.ridx and d.ridx2 may have been exchanged).
void detect_write(DetectBufs& d, int64_t m) {
  if (m == 0) return;
  const int64_t N = d.N, NS = N * d.S;
  const hipStream_t s = d.s;
  const uint32_t mask = (uint32_t)(d.T - 1);
  if (d.S == 9)
    k_detect<9><<<nblk(NS), 256, 0, s>>>(1, N, d.cell4, d.geo, d.bstart, d.perm, mask, d.idev,
                                         d.dslot_off, d.dacc, d.rI, d.rJ, d.rec);
  else
    k_detect<27><<<nblk(NS), 256, 0, s>>>(1, N, d.cell4, d.geo, d.bstart, d.perm, mask, d.idev,
                                          d.dslot_off, d.dacc, d.rI, d.rJ, d.rec);
  k_segsort_by_J<<<nblk(N), 256, 0, s>>>(N, d.dcount, d.rJ, d.ridx);
  k_make_keys<<<nblk(m), 256, 0, s>>>(m, d.ridx, d.rec, d.rkey);
  sort_pairs_on<uint64_t>(d, d.rkey, d.rkey2, d.ridx, d.ridx2, m, 64);
  std::swap(d.ridx, d.ridx2);  // sorted order now in ridx
  if (d.B > 1) {
    // (colony, contact x, I, J): stable LSD sort by colony on top
    uint32_t* k0 = reinterpret_cast<uint32_t*>(d.rkey);
    uint32_t* k1 = reinterpret_cast<uint32_t*>(d.rkey2);
    k_inst_keys<<<nblk(m), 256, 0, s>>>(m, d.ridx, d.rI, d.binst, k0);
    sort_pairs_on<uint32_t>(d, k0, k1, d.ridx, d.ridx2, m, colony_bits(d.B));
    std::swap(d.ridx, d.ridx2);
  }
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
  ctx->bucket.ensure(N + 1);
  ctx->perm.ensure(N + 1);
  ctx->bstart.ensure(T + 1);
  ctx->bcursor.ensure(T + 1);
  ctx->dslot.ensure(NS + 1);
  ctx->dslot_off.ensure(NS + 1);
  ctx->dacc.ensure(NS + 1);
  ctx->dcount.ensure(N + 1);
  ctx->bbounds.ensure(B + 1);
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
  d.bucket = ctx->bucket.p;
  d.perm = ctx->perm.p;
  d.bstart = ctx->bstart.p;
  d.bcursor = ctx->bcursor.p;
  d.dslot = ctx->dslot.p;
  d.dslot_off = ctx->dslot_off.p;
  d.dacc = ctx->dacc.p;
  d.dcount = ctx->dcount.p;
  d.bbounds = ctx->bbounds.p;
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
  std::vector<int32_t> bounds(B + 1);
  detect_count(d, bounds.data());
  const int64_t m = bounds[B];
  ctx->batch_count = m;
  ctx->rI_.ensure(m + 1);
  ctx->rJ_.ensure(m + 1);
  ctx->rrec.ensure(R_LEN * (m + 1));
  ctx->ridx.ensure(m + 1);
  ctx->ridx2.ensure(m + 1);
  ctx->rkey.ensure(m + 1);
  ctx->rkey2.ensure(m + 1);
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
  d.tmp = ctx->sort_tmp.p;
  d.tmp_bytes = (size_t)ctx->sort_tmp.cap;
  detect_write(d, m);
  if (d.ridx != ctx->ridx.p) ctx->ridx.swap(ctx->ridx2);
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

// Appends the sorted batch; with several colonies the set is then regrouped
// instance-major (each colony's old constraints, then its new ones).  The
// instance records must hold the colonies' current ranges, m and bstart; on
// return they hold the new ranges.
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
  if (B > 1 && m > 0) {
    const int64_t T = S.C;
    ConstraintSet& D = ctx->S2;
    ensure_set_capacity(D, T + 1, s);
    ctx->regroup_dst.ensure(T + 1);
    int64_t* dst = ctx->regroup_dst.p;
    const unsigned g = nblk(T);
    k_regroup<<<g, 256, 0, s>>>(T, C0, ctx->idev.p, S.inst.p, dst);
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
  }
  for (auto& q : ctx->inst) {
    q.coff = q.coff_new_;
    q.C += q.m;
  }
  upload_instances(ctx);
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
