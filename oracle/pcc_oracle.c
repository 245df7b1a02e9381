/*
 * pcc_oracle.c — FP64 CPU restatement of the reference's per-timestep collision
 * pipeline.  TEST INFRASTRUCTURE ONLY (checker for libpcc and the CPU baseline
 * of bench.py); the product never links or calls it.
 *
 * Built twice by oracle/Makefile:
 *   liboracle.so        strict IEEE (-O2 -ffp-contract=off): the parity oracle;
 *   liboracle_fast.so   the reference's Release flags (-O3 -ffast-math
 *                       -funroll-loops, code/cpp/CMakeLists.txt:17) + OpenMP:
 *                       the timed CPU baseline.
 *
 * Operation order follows the reference line by line (cited per function) so
 * that the FP64 GPU path, which evaluates the same expressions, reproduces it
 * bit for bit.  The deliberate deviations are listed in pcc_oracle.h.
 */
#include "pcc_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#ifdef _OPENMP
#include <omp.h>
#endif

int oracle_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

static int64_t g_iteration_cap = 0;
void oracle_set_iteration_cap(int64_t cap) { g_iteration_cap = cap; }

/* wall time of each ReLCP phase of the last hard step (timing diagnostics) */
#define ORACLE_MAX_PHASES 64
static double g_phase_time[ORACLE_MAX_PHASES];
static int64_t g_phase_count = 0;
static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}
int64_t oracle_phase_times(double* out, int64_t max) {
  for (int64_t i = 0; i < g_phase_count && i < max; ++i) out[i] = g_phase_time[i];
  return g_phase_count;
}

/* ------------------------------------------------------------------------ */
/* Small vector math: ArrayMath.h                                            */
/* ------------------------------------------------------------------------ */

/* dot: `result = 0; result += a[d]*b[d]` (ArrayMath.h:152-159). */
static inline double dot3(const double a[3], const double b[3]) {
  double r = 0.0;
  r += a[0] * b[0];
  r += a[1] * b[1];
  r += a[2] * b[2];
  return r;
}

/* cross_product (ArrayMath.h:139-142). */
static inline void cross3(const double a[3], const double b[3], double o[3]) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

/* normalize (ArrayMath.h:218-231): a * (1/|a|), or (1,0,0) if |a| <= eps. */
void oracle_normalize(const double a[3], double o[3]) {
  double mag = sqrt(dot3(a, a));
  if (mag > DBL_EPSILON) {
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

/* Quaternion::rotateVectorOfPositions with v = e_x (Quaternion.cpp:7-38). */
void oracle_direction(const double q[4], double t[3]) {
  const double w = q[0], x = q[1], y = q[2], z = q[3];
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

/* ------------------------------------------------------------------------ */
/* Deterministic exp: Cody & Waite (1980) argument reduction by ln2 split into */
/* a 33-bit head and a tail, then the degree-13 Taylor polynomial of e^r on   */
/* |r| <= ln2/2 in Horner form, then exact scaling by 2^k.  Coefficients are */
/* 1/n! correctly rounded.  Same expression order on CPU and GPU.            */
/* ------------------------------------------------------------------------ */
double oracle_exp(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return INFINITY;
  if (x < -745.1332191019412) return 0.0;
  const double inv_ln2 = 0x1.71547652b82fep+0;
  const double ln2_hi = 0x1.62e42fef00000p-1;
  const double ln2_lo = 0x1.473de6af278edp-34;
  double k = rint(x * inv_ln2);
  double r = (x - k * ln2_hi) - k * ln2_lo;
  double p = 0x1.6124613a86d09p-33;      /* 1/13! */
  p = p * r + 0x1.1eed8eff8d898p-29;     /* 1/12! */
  p = p * r + 0x1.ae64567f544e4p-26;     /* 1/11! */
  p = p * r + 0x1.27e4fb7789f5cp-22;     /* 1/10! */
  p = p * r + 0x1.71de3a556c734p-19;     /* 1/9!  */
  p = p * r + 0x1.a01a01a01a01ap-16;     /* 1/8!  */
  p = p * r + 0x1.a01a01a01a01ap-13;     /* 1/7!  */
  p = p * r + 0x1.6c16c16c16c17p-10;     /* 1/6!  */
  p = p * r + 0x1.1111111111111p-7;      /* 1/5!  */
  p = p * r + 0x1.5555555555555p-5;      /* 1/4!  */
  p = p * r + 0x1.5555555555555p-3;      /* 1/3!  */
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  return ldexp(p, (int)k);
}

/* ------------------------------------------------------------------------ */
/* DCP: DCPQuery<double,3>::operator() (SpherocylinderCell.h:64-261)         */
/* ------------------------------------------------------------------------ */
void oracle_dcp(const double P0[3], const double P1[3], const double Q0[3], const double Q1[3],
                double out[10]) {
  double P1mP0[3], Q1mQ0[3], P0mQ0[3];
  for (int k = 0; k < 3; ++k) {
    P1mP0[k] = P1[k] - P0[k];
    Q1mQ0[k] = Q1[k] - Q0[k];
    P0mQ0[k] = P0[k] - Q0[k];
  }
  double a = dot3(P1mP0, P1mP0);
  double b = dot3(P1mP0, Q1mQ0);
  double c = dot3(Q1mQ0, Q1mQ0);
  double d = dot3(P1mP0, P0mQ0);
  double e = dot3(Q1mQ0, P0mQ0);
  double det = a * c - b * b;
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
  double c0[3], c1[3], diff[3];
  for (int k = 0; k < 3; ++k) {
    c0[k] = P0[k] + P1mP0[k] * s;
    c1[k] = Q0[k] + Q1mQ0[k] * t;
  }
  for (int k = 0; k < 3; ++k) diff[k] = c0[k] - c1[k];
  double sqr = dot3(diff, diff);
  out[0] = s;
  out[1] = t;
  out[2] = c0[0];
  out[3] = c0[1];
  out[4] = c0[2];
  out[5] = c1[0];
  out[6] = c1[1];
  out[7] = c1[2];
  out[8] = sqr;
  out[9] = sqrt(sqr);
}

/* ------------------------------------------------------------------------ */
/* Broad phase: hashed cell list with a stable counting sort                 */
/* ------------------------------------------------------------------------ */

/* Cell coordinate: static_cast<int>(floor((x - min)/h)) (SpatialGrid.cpp:29-36). */
static inline int32_t cell_coord(double x, double origin, double h) {
  return (int32_t)floor((x - origin) / h);
}

/* 32-bit mix of the three cell coordinates: odd-multiplier combination then
 * the MurmurHash3 fmix32 finaliser (Appleby). */
static inline uint32_t cell_hash(int32_t cx, int32_t cy, int32_t cz) {
  uint32_t h = (uint32_t)cx * 0x9E3779B1u + (uint32_t)cy * 0x85EBCA77u + (uint32_t)cz * 0xC2B2AE3Du;
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return h;
}

/* Table slot of a cell (before the mask): bodies in one z-layer with their
 * axes in the plane (z = 0, quaternion x = y = 0) map cells row by row, 1024
 * to a row; others use the hash.  The slots only place cells in the table:
 * every walk compares cell coordinates, so the contacts do not depend on it. */
static inline uint32_t cell_slot(int32_t cx, int32_t cy, int32_t cz, int flat) {
  return flat ? (uint32_t)cx + ((uint32_t)cy << 10) : cell_hash(cx, cy, cz);
}

int oracle_grid_flat(int64_t n, const double* pos, const double* quat) {
  for (int64_t i = 0; i < n; ++i) {
    if (pos[3 * i + 2] != 0.0) return 0;
    if (quat && (quat[4 * i + 1] != 0.0 || quat[4 * i + 2] != 0.0)) return 0;
  }
  return 1;
}

int64_t oracle_table_size(int64_t n) {
  int64_t t = 1024;
  while (t < 2 * n) t <<= 1;
  return t;
}

int64_t oracle_grid(int64_t n, const double* pos, const double* quat, double h, double origin,
                    int32_t* cell, uint32_t* bucket, int32_t* perm, int32_t* bucket_start,
                    int64_t bucket_start_cap) {
  int64_t T = oracle_table_size(n);
  const int flat = oracle_grid_flat(n, pos, quat);
  uint32_t mask = (uint32_t)(T - 1);
  int32_t* cl = cell ? cell : (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)(n > 0 ? n : 1));
  uint32_t* bk = bucket ? bucket : (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
  for (int64_t i = 0; i < n; ++i) {
    for (int k = 0; k < 3; ++k) cl[3 * i + k] = cell_coord(pos[3 * i + k], origin, h);
    bk[i] = cell_slot(cl[3 * i], cl[3 * i + 1], cl[3 * i + 2], flat) & mask;
  }
  if (perm && bucket_start && bucket_start_cap >= T + 1) {
    memset(bucket_start, 0, sizeof(int32_t) * (size_t)(T + 1));
    for (int64_t i = 0; i < n; ++i) bucket_start[bk[i] + 1]++;
    for (int64_t b = 0; b < T; ++b) bucket_start[b + 1] += bucket_start[b];
    int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * (size_t)T);
    memcpy(cur, bucket_start, sizeof(int32_t) * (size_t)T);
    for (int64_t i = 0; i < n; ++i) perm[cur[bk[i]]++] = (int32_t)i;
    free(cur);
  }
  if (!cell) free(cl);
  if (!bucket) free(bk);
  return T;
}

/* ------------------------------------------------------------------------ */
/* Narrow phase: tryCreateConstraint (CollisionDetector.cpp:201-311)         */
/* ------------------------------------------------------------------------ */
typedef struct rec_t {
  int32_t I, J;
  double phi;
  double n[3], rI[3], rJ[3], p[3];
  double sI, sJ, gamma;
} rec_t;

typedef struct body_view {
  const double* pos;
  const double* quat;
  const double* len;
  const double* diam;
} body_view;

/* getParticleEndpoints (CollisionDetector.cpp:61-75). */
static inline void endpoints(const body_view* bv, int64_t i, double P0[3], double P1[3],
                             double t[3]) {
  oracle_direction(bv->quat + 4 * i, t);
  double half = bv->len[i] / 2 - bv->diam[i] / 2;
  for (int k = 0; k < 3; ++k) {
    double hv = t[k] * half;
    P0[k] = bv->pos[3 * i + k] - hv;
    P1[k] = bv->pos[3 * i + k] + hv;
  }
}

/* Returns 1 and fills *r if (I, J) yields a constraint at this cutoff. */
static int try_pair(const body_view* bv, int32_t I, int32_t J, double cutoff, double warm,
                    rec_t* r) {
  const double* xi = bv->pos + 3 * (int64_t)I;
  const double* xj = bv->pos + 3 * (int64_t)J;
  double dx[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
  double rough = sqrt(dot3(dx, dx));
  if (rough > (bv->len[I] + bv->len[J]) / 2.0 + cutoff) return 0;
  double P0[3], P1[3], Q0[3], Q1[3], tI[3], tJ[3], dcp[10];
  endpoints(bv, I, P0, P1, tI);
  endpoints(bv, J, Q0, Q1, tJ);
  oracle_dcp(P0, P1, Q0, Q1, dcp);
  const double* c0 = dcp + 2;
  const double* c1 = dcp + 5;
  double sum_radii = (bv->diam[I] + bv->diam[J]) / 2.0;
  double phi = dcp[9] - sum_radii;
  if (phi > cutoff) return 0;
  r->I = I;
  r->J = J;
  r->phi = phi;
  double d01[3];
  for (int k = 0; k < 3; ++k) {
    r->p[k] = (c0[k] + c1[k]) * 0.5;
    d01[k] = c0[k] - c1[k];
  }
  for (int k = 0; k < 3; ++k) {
    r->rI[k] = r->p[k] - xi[k];
    r->rJ[k] = r->p[k] - xj[k];
  }
  oracle_normalize(d01, r->n);
  double mn[3] = {-r->n[0], -r->n[1], -r->n[2]};
  r->sI = 0.5 * fabs(dot3(r->n, tI));
  r->sJ = 0.5 * fabs(dot3(mn, tJ));
  r->gamma = phi >= 0.0 ? 0.0 : -phi * warm;
  return 1;
}

static int rec_cmp(const void* pa, const void* pb) {
  const rec_t* a = (const rec_t*)pa;
  const rec_t* b = (const rec_t*)pb;
  if (a->p[0] < b->p[0]) return -1;
  if (b->p[0] < a->p[0]) return 1;
  if (a->I != b->I) return a->I < b->I ? -1 : 1;
  if (a->J != b->J) return a->J < b->J ? -1 : 1;
  return 0;
}

/* Detection into a malloc'd record array sorted by (p.x, I, J). */
static int64_t detect_records(const body_view* bv, int64_t n, double cutoff, double h,
                              double origin, double warm, int brute, rec_t** out_recs) {
  int64_t T = oracle_table_size(n);
  int32_t* cell = (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)(n > 0 ? n : 1));
  uint32_t* bkt = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
  int32_t* perm = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  int32_t* start = (int32_t*)malloc(sizeof(int32_t) * (size_t)(T + 1));
  oracle_grid(n, bv->pos, bv->quat, h, origin, cell, bkt, perm, start, T + 1);
  const int flat = oracle_grid_flat(n, bv->pos, bv->quat);
  uint32_t mask = (uint32_t)(T - 1);

  int64_t* cnt = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
  /* pass 1 counts, pass 2 writes (both visit the same candidates) */
  rec_t* recs = NULL;
  for (int pass = 0; pass < 2; ++pass) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t I = 0; I < n; ++I) {
      int64_t w = pass ? cnt[I] : 0;
      int64_t c = 0;
      rec_t r;
      if (brute) {
        for (int64_t J = I + 1; J < n; ++J) {
          if (try_pair(bv, (int32_t)I, (int32_t)J, cutoff, warm, &r)) {
            if (pass) recs[w + c] = r;
            ++c;
          }
        }
      } else {
        const int32_t cx = cell[3 * I], cy = cell[3 * I + 1], cz = cell[3 * I + 2];
        for (int dz = -1; dz <= 1; ++dz)
          for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
              int32_t nx = cx + dx, ny = cy + dy, nz = cz + dz;
              uint32_t b = cell_slot(nx, ny, nz, flat) & mask;
              for (int32_t e = start[b]; e < start[b + 1]; ++e) {
                int32_t J = perm[e];
                if (J <= I) continue;
                if (cell[3 * (int64_t)J] != nx || cell[3 * (int64_t)J + 1] != ny ||
                    cell[3 * (int64_t)J + 2] != nz)
                  continue;
                if (try_pair(bv, (int32_t)I, J, cutoff, warm, &r)) {
                  if (pass) recs[w + c] = r;
                  ++c;
                }
              }
            }
      }
      if (!pass) cnt[I] = c;
    }
    if (!pass) {
      int64_t acc = 0;
      for (int64_t I = 0; I < n; ++I) {
        int64_t c = cnt[I];
        cnt[I] = acc;
        acc += c;
      }
      cnt[n] = acc;
      recs = (rec_t*)malloc(sizeof(rec_t) * (size_t)(acc > 0 ? acc : 1));
    }
  }
  int64_t total = cnt[n];
  qsort(recs, (size_t)total, sizeof(rec_t), rec_cmp);
  free(cnt);
  free(cell);
  free(bkt);
  free(perm);
  free(start);
  *out_recs = recs;
  return total;
}

static void write_record(pcc_constraints* out, int64_t i, const rec_t* r, int32_t gid,
                         int32_t iteration) {
  if (!out || i >= out->capacity) return;
  if (out->gidI) out->gidI[i] = r->I;
  if (out->gidJ) out->gidJ[i] = r->J;
  if (out->signed_distance) out->signed_distance[i] = r->phi;
  for (int k = 0; k < 3; ++k) {
    if (out->normal) out->normal[3 * i + k] = r->n[k];
    if (out->rPosI) out->rPosI[3 * i + k] = r->rI[k];
    if (out->rPosJ) out->rPosJ[3 * i + k] = r->rJ[k];
    if (out->contact_point) out->contact_point[3 * i + k] = r->p[k];
  }
  if (out->stressI) out->stressI[i] = r->sI;
  if (out->stressJ) out->stressJ[i] = r->sJ;
  if (out->gamma) out->gamma[i] = r->gamma;
  if (out->gid) out->gid[i] = gid;
  if (out->iteration) out->iteration[i] = iteration;
}

int64_t oracle_detect(const pcc_bodies* b, double cutoff, double h, double origin, double warm,
                      int32_t iteration, int brute, pcc_constraints* out,
                      int32_t* num_constraints) {
  body_view bv = {b->pos, b->quat, b->length, b->diameter};
  rec_t* recs = NULL;
  int64_t C = detect_records(&bv, b->n, cutoff, h, origin, warm, brute, &recs);
  for (int64_t i = 0; i < C; ++i) {
    write_record(out, i, &recs[i], (int32_t)i, iteration);
    if (num_constraints) {
      num_constraints[recs[i].I]++;
      num_constraints[recs[i].J]++;
    }
  }
  free(recs);
  return C;
}

/* ------------------------------------------------------------------------ */
/* Deterministic blocked sum (mirrors the GPU's block/wave reduction tree)   */
/* ------------------------------------------------------------------------ */
static double butterfly64(const double* x) {
  double v[64], w[64];
  memcpy(v, x, sizeof(v));
  for (int off = 32; off >= 1; off >>= 1) {
    for (int l = 0; l < 64; ++l) w[l] = v[l] + v[l ^ off];
    memcpy(v, w, sizeof(v));
  }
  return v[0];
}

static double block256(const double* x) {
  double w0 = butterfly64(x), w1 = butterfly64(x + 64), w2 = butterfly64(x + 128),
         w3 = butterfly64(x + 192);
  return ((w0 + w1) + w2) + w3;
}

/* Sum of per-tile partials: thread t accumulates partials t, t+256, ...
 * sequentially from 0.0, then the same block tree. */
static double reduce_partials(const double* part, int64_t nb) {
  double acc[256];
  for (int t = 0; t < 256; ++t) {
    double a = 0.0;
    for (int64_t i = t; i < nb; i += 256) a += part[i];
    acc[t] = a;
  }
  return block256(acc);
}

double oracle_block_sum(const double* v, int64_t n) {
  int64_t nb = (n + 255) / 256;
  double* part = (double*)malloc(sizeof(double) * (size_t)(nb > 0 ? nb : 1));
  for (int64_t b = 0; b < nb; ++b) {
    double x[256];
    for (int t = 0; t < 256; ++t) {
      int64_t i = b * 256 + t;
      x[t] = i < n ? v[i] : 0.0;
    }
    part[b] = block256(x);
  }
  double r = reduce_partials(part, nb);
  free(part);
  return r;
}

/* Three BB sums at once: sum(dg*dg), sum(dg*dp), sum(dp*dp), in the GPU's
 * order: the constraints, taken in solver order ord[] (stably sorted by the
 * lower body index I, the layout the GPU solver keeps them in), form
 * Q = ceil(C/256) tiles of 256; G = min(Q, 2048) logical blocks each own
 * K = ceil(Q/G) consecutive tiles; thread t of a block accumulates element t
 * of its tiles in tile order starting from 0.0; then the block tree per block
 * and the partial reduction over blocks. */
#define ORACLE_BB_GMAX 2048
static void bb_sums(const double* gi, const double* gp, const double* ri, const double* rp,
                    const int64_t* ord, int64_t C, double* Sgg, double* Sgp, double* Spp) {
  const int64_t Q = (C + 255) / 256;
  const int64_t G = Q < ORACLE_BB_GMAX ? (Q > 0 ? Q : 1) : ORACLE_BB_GMAX;
  const int64_t K = (Q + G - 1) / G;
  double* part = (double*)malloc(sizeof(double) * 3 * (size_t)G);
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < G; ++b) {
    double x0[256], x1[256], x2[256];
    for (int t = 0; t < 256; ++t) {
      double a0 = 0.0, a1 = 0.0, a2 = 0.0;
      for (int64_t j = 0; j < K; ++j) {
        int64_t pos = (b * K + j) * 256 + t;
        if (pos >= C) break;
        const int64_t i = ord[pos];
        double dgam = gi[i] - gp[i];
        double dphi = ri[i] - rp[i];
        a0 += dgam * dgam;
        a1 += dgam * dphi;
        a2 += dphi * dphi;
      }
      x0[t] = a0;
      x1[t] = a1;
      x2[t] = a2;
    }
    part[b] = block256(x0);
    part[G + b] = block256(x1);
    part[2 * G + b] = block256(x2);
  }
  *Sgg = reduce_partials(part, G);
  *Sgp = reduce_partials(part + G, G);
  *Spp = reduce_partials(part + 2 * G, G);
  free(part);
}

/* ------------------------------------------------------------------------ */
/* Brownian input                                                            */
/* ------------------------------------------------------------------------ */
/* splitmix64 finaliser (Steele, Lea, Flood 2014) */
static inline uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

/* Draw k of body i at the colony's step count iter, uniform on [-1, 1) (the
 * std::uniform_real_distribution(-1, 1) of Particle.cpp:177), from a
 * counter-based generator in place of the reference's unseeded global
 * std::mt19937 (Physics.cpp:19). */
static inline double brownian_uniform(uint64_t seed, int64_t iter, int64_t i, int k) {
  const uint64_t key = mix64(mix64(seed ^ 0xB7E151628AED2A6Bull) ^ (uint64_t)(uint32_t)iter);
  const uint64_t z = mix64(key ^ ((uint64_t)i * 4u + (uint64_t)k));
  return (double)(z >> 11) * 0x1.0p-52 - 1.0;
}

/* calculate_external_velocities (Physics.cpp:59-81) with
 * Particle::calculateBrownianVelocity (Particle.cpp:174-190): U_ext zeroed,
 * then per body the angular part added, each product in the reference's
 * left-to-right order and the draws in its order (the monolayer ternaries
 * skip the x/y draws).  imp: each body's impedance from its last step (NULL:
 * 0, a new Particle's value, Particle.cpp:25).  Returns NULL when T <= 0. */
static double* brownian_uext(const pcc_params* p, const pcc_bodies* b, int32_t iter) {
  if (!(p->temperature > 0)) return NULL;
  const int64_t N = b->n;
  double* w = (double*)malloc(sizeof(double) * (size_t)(3 * N > 0 ? 3 * N : 1));
  for (int64_t i = 0; i < N; ++i) {
    const double imp = b->impedance ? b->impedance[i] : 0.0;
    const double rot = p->temperature * brownian_uniform(p->brownian_seed, iter, i, 0);
    double v[3];
    if (p->monolayer) {
      v[0] = 0.0;
      v[1] = 0.0;
      v[2] = (((rot * brownian_uniform(p->brownian_seed, iter, i, 1)) * imp) * imp) * imp;
    } else {
      for (int k = 0; k < 3; ++k)
        v[k] = (((rot * brownian_uniform(p->brownian_seed, iter, i, 1 + k)) * imp) * imp) * imp;
    }
    for (int k = 0; k < 3; ++k) w[3 * i + k] = 0.0 + v[k]; /* VecSetValues ADD_VALUES on 0 */
  }
  return w;
}

/* ------------------------------------------------------------------------ */
/* Hard model                                                                */
/* ------------------------------------------------------------------------ */
typedef struct cset_t { /* all_constraints_set (Hard.cpp:101) as SoA */
  int64_t C, cap;
  rec_t* r;
  int32_t* gid;
  int32_t* iter;
} cset_t;

static void cset_append(cset_t* S, const rec_t* recs, int64_t m, int32_t iteration) {
  if (S->C + m > S->cap) {
    int64_t nc = S->cap ? S->cap : 1024;
    while (nc < S->C + m) nc *= 2;
    S->r = (rec_t*)realloc(S->r, sizeof(rec_t) * (size_t)nc);
    S->gid = (int32_t*)realloc(S->gid, sizeof(int32_t) * (size_t)nc);
    S->iter = (int32_t*)realloc(S->iter, sizeof(int32_t) * (size_t)nc);
    S->cap = nc;
  }
  for (int64_t i = 0; i < m; ++i) {
    S->r[S->C + i] = recs[i];
    S->gid[S->C + i] = (int32_t)i;
    S->iter[S->C + i] = iteration;
  }
  S->C += m;
}

/* Operators of one ReLCP phase (createMatrices, Hard.cpp:10-47; Physics.cpp). */
typedef struct phase_t {
  int64_t N, C;
  double dt, lambda, tau;
  double* D;        /* C x 12: body I [n, rI x n], body J [-n, rJ x -n] (Physics.cpp:118-154) */
  double* sI;       /* L row: (I, sI), (J, sJ) (Physics.cpp:174-190) */
  double* sJ;
  int32_t* I;
  int32_t* J;
  double* phi;      /* PHI */
  double* gold;     /* gamma_old (Hard.cpp:146-147) */
  double* mt;       /* mobility 1/(xi l) (Physics.cpp:203) */
  double* mr;       /* 12/(xi l^3) (Physics.cpp:204) */
  double* len;
  double* ldot_prev;
  int64_t* rowptr;  /* body -> incidences (ascending constraint index) */
  int64_t* inc;     /* 2*c + side */
  int64_t* ord;     /* solver order: constraints stably sorted by I */
  /* workspace (Workspace.h:5-20) */
  double* dgam;
  double* F;        /* 6N */
  double* U;        /* 6N */
  double* sig;      /* N */
  double* imp;      /* N */
  double* ldot;     /* N */
  double* dldot;    /* N */
  const double* uext; /* 3N Brownian angular velocity of phase 0 (U_ext), or NULL */
} phase_t;

static void phase_free(phase_t* ph) {
  free(ph->D);
  free(ph->sI);
  free(ph->sJ);
  free(ph->I);
  free(ph->J);
  free(ph->phi);
  free(ph->gold);
  free(ph->mt);
  free(ph->mr);
  free(ph->len);
  free(ph->ldot_prev);
  free(ph->rowptr);
  free(ph->inc);
  free(ph->ord);
  free(ph->dgam);
  free(ph->F);
  free(ph->U);
  free(ph->sig);
  free(ph->imp);
  free(ph->ldot);
  free(ph->dldot);
}

#define ALLOC(T, n) ((T*)malloc(sizeof(T) * (size_t)((n) > 0 ? (n) : 1)))

static void phase_build(phase_t* ph, const cset_t* S, const pcc_bodies* b, const pcc_params* p) {
  const int64_t N = b->n, C = S->C;
  ph->N = N;
  ph->C = C;
  ph->dt = p->dt;
  ph->lambda = p->lambda;
  ph->tau = p->tau;
  ph->D = ALLOC(double, 12 * C);
  ph->sI = ALLOC(double, C);
  ph->sJ = ALLOC(double, C);
  ph->I = ALLOC(int32_t, C);
  ph->J = ALLOC(int32_t, C);
  ph->phi = ALLOC(double, C);
  ph->gold = ALLOC(double, C);
  ph->dgam = ALLOC(double, C);
  for (int64_t c = 0; c < C; ++c) {
    const rec_t* r = &S->r[c];
    double nj[3] = {-r->n[0], -r->n[1], -r->n[2]};
    double ti[3], tj[3];
    cross3(r->rI, r->n, ti);
    cross3(r->rJ, nj, tj);
    double* Dc = ph->D + 12 * c;
    Dc[0] = r->n[0];
    Dc[1] = r->n[1];
    Dc[2] = r->n[2];
    Dc[3] = ti[0];
    Dc[4] = ti[1];
    Dc[5] = ti[2];
    Dc[6] = nj[0];
    Dc[7] = nj[1];
    Dc[8] = nj[2];
    Dc[9] = tj[0];
    Dc[10] = tj[1];
    Dc[11] = tj[2];
    ph->sI[c] = r->sI;
    ph->sJ[c] = r->sJ;
    ph->I[c] = r->I;
    ph->J[c] = r->J;
    ph->phi[c] = r->phi;
    ph->gold[c] = r->gamma;
  }
  ph->mt = ALLOC(double, N);
  ph->mr = ALLOC(double, N);
  ph->len = ALLOC(double, N);
  ph->ldot_prev = ALLOC(double, N);
  for (int64_t i = 0; i < N; ++i) {
    double l = b->length[i];
    ph->mt[i] = 1.0 / (l * p->xi);
    ph->mr[i] = 12.0 / (l * l * l * p->xi);
    ph->len[i] = l;
    ph->ldot_prev[i] = b->ldot[i];
  }
  /* incidence CSR, ascending constraint index per body (the row order in which
   * PETSc's MatMultTranspose accumulates). */
  ph->rowptr = (int64_t*)calloc((size_t)N + 1, sizeof(int64_t));
  ph->inc = ALLOC(int64_t, 2 * C);
  for (int64_t c = 0; c < C; ++c) {
    ph->rowptr[ph->I[c] + 1]++;
    ph->rowptr[ph->J[c] + 1]++;
  }
  for (int64_t i = 0; i < N; ++i) ph->rowptr[i + 1] += ph->rowptr[i];
  int64_t* cur = ALLOC(int64_t, N);
  memcpy(cur, ph->rowptr, sizeof(int64_t) * (size_t)N);
  for (int64_t c = 0; c < C; ++c) {
    ph->inc[cur[ph->I[c]]++] = 2 * c;
    ph->inc[cur[ph->J[c]]++] = 2 * c + 1;
  }
  /* solver order: counting sort of the constraints by I, stable (ties keep
   * ascending constraint index) */
  ph->ord = ALLOC(int64_t, C);
  int64_t* cnt = (int64_t*)calloc((size_t)N + 1, sizeof(int64_t));
  for (int64_t c = 0; c < C; ++c) cnt[ph->I[c] + 1]++;
  for (int64_t i = 0; i < N; ++i) cnt[i + 1] += cnt[i];
  for (int64_t c = 0; c < C; ++c) ph->ord[cnt[ph->I[c]]++] = c;
  free(cnt);
  free(cur);
  ph->F = ALLOC(double, 6 * N);
  ph->U = ALLOC(double, 6 * N);
  ph->sig = ALLOC(double, N);
  ph->imp = ALLOC(double, N);
  ph->ldot = ALLOC(double, N);
  ph->dldot = ALLOC(double, N);
}

/* HardModelGradient::gradient (HardModelGradient.cpp:89-121). */
static void gradient(phase_t* ph, const double* gam, double* g) {
  const int64_t N = ph->N, C = ph->C;
  const double dt = ph->dt;
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < C; ++c) ph->dgam[c] = gam[c] - ph->gold[c]; /* :92 */
  const double mlam = -ph->lambda;
  const double itau = 1 / ph->tau;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    /* F = D^T dgam (:15), sigma = L^T gamma (:56) */
    double F[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double sig = 0.0;
    for (int64_t e = ph->rowptr[i]; e < ph->rowptr[i + 1]; ++e) {
      int64_t c = ph->inc[e] >> 1;
      int side = (int)(ph->inc[e] & 1);
      const double* Dv = ph->D + 12 * c + 6 * side;
      double a = ph->dgam[c];
      for (int k = 0; k < 6; ++k) F[k] += a * Dv[k];
      sig += gam[c] * (side ? ph->sJ[c] : ph->sI[c]);
    }
    /* U_c = M F (:18), U_total = 1.0 U_ext + U_c (:21); U_ext is zero
     * except the Brownian angular part on phase 0 */
    for (int k = 0; k < 3; ++k) {
      ph->F[6 * i + k] = F[k];
      ph->U[6 * i + k] = (0.0 + ph->mt[i] * F[k]) + 0.0;
    }
    for (int k = 3; k < 6; ++k) {
      ph->F[6 * i + k] = F[k];
      ph->U[6 * i + k] = (0.0 + ph->mr[i] * F[k]) + (ph->uext ? ph->uext[3 * i + k - 3] : 0.0);
    }
    /* impedance = exp(-lambda sigma), ldot = (l .* impedance) * (1/tau) (:34-52) */
    ph->sig[i] = sig;
    double imp = oracle_exp(sig * mlam);
    ph->imp[i] = imp;
    double ld = (ph->len[i] * imp) * itau;
    ph->ldot[i] = ld;
    ph->dldot[i] = ld - ph->ldot_prev[i]; /* :108 */
  }
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < C; ++c) {
    /* phi_dot = D U_total (:24): 12 products in column order (I < J). */
    const double* Dc = ph->D + 12 * c;
    const double* UI = ph->U + 6 * (int64_t)ph->I[c];
    const double* UJ = ph->U + 6 * (int64_t)ph->J[c];
    double pd = 0.0;
    for (int k = 0; k < 6; ++k) pd += Dc[k] * UI[k];
    for (int k = 0; k < 6; ++k) pd += Dc[6 + k] * UJ[k];
    /* phi_dot_growth = L ldot_diff (:111) */
    double pg = 0.0;
    pg += ph->sI[c] * ph->dldot[ph->I[c]];
    pg += ph->sJ[c] * ph->dldot[ph->J[c]];
    /* phi_next = PHI + dt*phi_dot + (-dt)*phi_dot_growth (VecMAXPY, :114-118) */
    g[c] = ph->phi[c] + (dt * pd + (-dt) * pg);
  }
}

/* HardModelGradient::residual (HardModelGradient.cpp:123-158). */
static double residual(const double* g, const double* gam, int64_t C) {
  double res = 0.0;
  for (int64_t c = 0; c < C; ++c) {
    double v;
    if (gam[c] > 0) {
      v = g[c];
    } else {
      v = g[c] < 0.0 ? g[c] : 0.0;
    }
    double a = fabs(v);
    res = res < a ? a : res;
  }
  return res;
}

/* HardModelGradient::energy (HardModelGradient.cpp:160-203) at gam:
 * out = {gamma.Phi, dt/2 (D gamma).M(D gamma), dt/(lambda tau) sum l exp(-lambda
 * L gamma), their sum}.  Summed in the GPU's order: gamma.Phi over the
 * constraints in solver order in the blocked tree of bb_sums; per body the six
 * products F_k U_k in k order and l exp(-lambda sigma), then blocks of 256
 * bodies through block256 and the block partials through reduce_partials.
 * F = D^T gamma and sigma = L^T gamma in ascending incidence order, as in
 * gradient(). */
static void energy(const phase_t* ph, const double* gam, double out[4]) {
  const int64_t N = ph->N, C = ph->C;
  const int64_t Q = (C + 255) / 256;
  const int64_t G = Q < ORACLE_BB_GMAX ? (Q > 0 ? Q : 1) : ORACLE_BB_GMAX;
  const int64_t K = (Q + G - 1) / G;
  double* part = ALLOC(double, G);
  for (int64_t b = 0; b < G; ++b) {
    double x[256];
    for (int t = 0; t < 256; ++t) {
      double a = 0.0;
      for (int64_t j = 0; j < K; ++j) {
        int64_t pos = (b * K + j) * 256 + t;
        if (pos >= C) break;
        const int64_t i = ph->ord[pos];
        a += gam[i] * ph->phi[i];
      }
      x[t] = a;
    }
    part[b] = block256(x);
  }
  const double term1 = reduce_partials(part, G);
  free(part);
  const int64_t nb = (N + 255) / 256;
  double* p2 = ALLOC(double, nb);
  double* p3 = ALLOC(double, nb);
  const double mlam = -ph->lambda;
  for (int64_t b = 0; b < nb; ++b) {
    double x2[256], x3[256];
    for (int t = 0; t < 256; ++t) {
      const int64_t i = b * 256 + t;
      x2[t] = 0.0;
      x3[t] = 0.0;
      if (i >= N) continue;
      double F[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      double sig = 0.0;
      for (int64_t e = ph->rowptr[i]; e < ph->rowptr[i + 1]; ++e) {
        const int64_t c = ph->inc[e] >> 1;
        const int side = (int)(ph->inc[e] & 1);
        const double* Dv = ph->D + 12 * c + 6 * side;
        for (int k = 0; k < 6; ++k) F[k] += gam[c] * Dv[k];
        sig += gam[c] * (side ? ph->sJ[c] : ph->sI[c]);
      }
      double q = 0.0;
      for (int k = 0; k < 6; ++k) {
        const double m = k < 3 ? ph->mt[i] : ph->mr[i];
        const double u = (0.0 + m * F[k]) + 0.0;
        q += F[k] * u;
      }
      x2[t] = q;
      x3[t] = ph->len[i] * oracle_exp(sig * mlam);
    }
    p2[b] = block256(x2);
    p3[b] = block256(x3);
  }
  double term2 = reduce_partials(p2, nb);
  double term3 = reduce_partials(p3, nb);
  free(p2);
  free(p3);
  term2 *= (ph->dt / 2.0);
  term3 *= (ph->dt / (ph->lambda * ph->tau));
  out[0] = term1;
  out[1] = term2;
  out[2] = term3;
  out[3] = term1 + term2 + term3;
}

/* BBPGD trace (BBPGDLogger; collected at BBPGD.cpp:57-60,73-80). */
static pcc_bbpgd_record* g_trace = NULL;
static int64_t g_trace_cap = 0, g_trace_n = 0, g_trace_phase = 0;

void oracle_set_trace(pcc_bbpgd_record* buf, int64_t capacity) {
  g_trace = buf;
  g_trace_cap = buf ? capacity : 0;
  g_trace_n = 0;
}
int64_t oracle_trace_count(void) { return g_trace_n; }

static void trace_collect(const phase_t* ph, const double* gam, int64_t step, double res,
                          double alpha) {
  if (g_trace_n < g_trace_cap) {
    double E[4];
    energy(ph, gam, E);
    pcc_bbpgd_record* r = g_trace + g_trace_n;
    r->step = step;
    r->recursive_iteration = g_trace_phase;
    r->total_constraints = ph->C;
    r->residual = res;
    r->step_size = alpha;
    r->linear = E[0];
    r->quadratic = E[1];
    r->growth = E[2];
    r->total = E[3];
  }
  ++g_trace_n;
}

/* BBPGD (BBPGD.cpp:32-127).  On return Gamma holds the lagging iterate (Q3),
 * g_last the gradient of the last evaluated iterate (the workspace).  While
 * a trace buffer is set, the residual is checked at every iteration
 * (BBPGD.cpp:73, `|| bbpgd_logger`) and every iteration is recorded. */
static pcc_status bbpgd(phase_t* ph, double* Gamma, double tol, int64_t maxit, double* g_last,
                        int64_t* iters, double* res_out) {
  const int64_t C = ph->C;
  double* gp = ALLOC(double, C);
  double* gi = ALLOC(double, C);
  double* yp = ALLOC(double, C);
  double* yi = ALLOC(double, C);
  memcpy(yp, Gamma, sizeof(double) * (size_t)C);
  gradient(ph, yp, gp);             /* :49 */
  double res = residual(gp, yp, C); /* :51 */
  double alpha = 1.0 / res;         /* :52 */
  const int tracing = g_trace != NULL;
  if (tracing) trace_collect(ph, yp, 0, res, alpha); /* :57-60 */
  int64_t it = 1;
  pcc_status st = PCC_OK;
  int64_t cap = g_iteration_cap;
  while (it < maxit) {
    const double malpha = -alpha;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < C; ++c) { /* :66-67 */
      double v = yp[c] + malpha * gp[c];
      yi[c] = v < 0.0 ? 0.0 : v;
    }
    gradient(ph, yi, gi); /* :70 */
    if (it % 10 == 0 || tracing) { /* :73-88 */
      res = residual(gi, yi, C);
      if (tracing) trace_collect(ph, yi, it, res, alpha);
      if (res <= tol) break;
    }
    if (cap > 0 && it >= cap) break;
    double Sgg, Sgp, Spp; /* :91-104 */
    bb_sums(yi, yp, gi, gp, ph->ord, C, &Sgg, &Sgp, &Spp);
    double num, den;
    if (it % 2 == 0) {
      num = Sgg;
      den = Sgp;
    } else {
      den = Spp;
      num = Sgp;
    }
    if (fabs(den) < 10 * DBL_EPSILON) den += 10 * DBL_EPSILON; /* :106-108 */
    alpha = num / den;
    double* t = yp;
    yp = yi;
    yi = t;
    t = gp;
    gp = gi;
    gi = t;
    ++it;
  }
  if (it == maxit) st = PCC_E_BBPGD_NOT_CONVERGED; /* :118-121 */
  memcpy(Gamma, yp, sizeof(double) * (size_t)C); /* :124 (self-move, Q3) */
  memcpy(g_last, gi, sizeof(double) * (size_t)C);
  *iters = it;
  *res_out = res;
  free(gp);
  free(gi);
  free(yp);
  free(yi);
  return st;
}

/* G U (Physics.cpp:220-265 via MatMult, Hard.cpp:161) then
 * Particle::eulerStepPosition (Particle.cpp:37-57,91-94,110-120). */
static void integrate_body(pcc_bodies* b, int64_t i, const double* U, double dt) {
  double* q = b->quat + 4 * i;
  const double s = q[0], wx = q[1], wy = q[2], wz = q[3];
  const double xi[4][3] = {{-0.5 * wx, -0.5 * wy, -0.5 * wz},
                           {0.5 * s, -0.5 * wz, 0.5 * wy},
                           {0.5 * wz, 0.5 * s, -0.5 * wx},
                           {-0.5 * wy, 0.5 * wx, 0.5 * s}};
  double dC[7];
  for (int k = 0; k < 3; ++k) dC[k] = 0.0 + 1.0 * U[k];
  for (int r = 0; r < 4; ++r) {
    double sum = 0.0;
    sum += xi[r][0] * U[3];
    sum += xi[r][1] * U[4];
    sum += xi[r][2] * U[5];
    dC[3 + r] = sum;
  }
  for (int k = 0; k < 3; ++k) b->pos[3 * i + k] += dt * dC[k];
  for (int k = 0; k < 4; ++k) q[k] += dt * dC[3 + k];
  double norm = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  for (int k = 0; k < 4; ++k) q[k] /= norm;
}

static void set_err(char* err, int64_t len, const char* msg) {
  if (err && len > 0) {
    strncpy(err, msg, (size_t)len - 1);
    err[len - 1] = 0;
  }
}

static void write_set(pcc_constraints* out, const cset_t* S) {
  if (!out) return;
  for (int64_t c = 0; c < S->C; ++c) write_record(out, c, &S->r[c], S->gid[c], S->iter[c]);
}

/* solveHardModel (Hard.cpp:91-193). */
static pcc_status hard_step(const pcc_params* p, pcc_bodies* b, int32_t iter,
                            pcc_step_result* result, pcc_constraints* cons,
                            int64_t* phase_iters, int64_t max_phase, char* err,
                            int64_t err_len) {
  const int64_t N = b->n;
  const double dt = p->dt;
  body_view bv = {b->pos, b->quat, b->length, b->diameter};
  cset_t S = {0, 0, NULL, NULL, NULL};
  int32_t k = 0;
  int64_t total_it = 0;
  double res = 0.0, maxov = 0.0;
  pcc_status st = PCC_OK;
  for (int64_t i = 0; i < N; ++i) { /* Particle::reset (Particle.h:146-149) */
    b->ldot[i] = 0.0;
    if (b->num_constraints) b->num_constraints[i] = 0;
  }
  g_phase_count = 0;
  while (k < p->max_recursive_iterations) { /* :113 */
    const double t_phase = now_s();
    double cutoff = k == 0 ? p->cutoff_phase0 : p->cutoff_phaseK; /* :121 */
    rec_t* recs = NULL;
    int64_t m = detect_records(&bv, N, cutoff, p->linked_cell_size, p->grid_origin,
                               p->warm_start_factor, 0, &recs);
    if (b->num_constraints)
      for (int64_t i = 0; i < m; ++i) {
        b->num_constraints[recs[i].I]++;
        b->num_constraints[recs[i].J]++;
      }
    cset_append(&S, recs, m, k); /* :123 */
    free(recs);
    double acc = 0.0; /* :125-130 */
    for (int64_t c = 0; c < S.C; ++c) {
      double ph = S.r[c].phi;
      double v = ph < 0 ? -ph : 0;
      acc = acc < v ? v : acc;
    }
    maxov = acc;
    if (maxov <= p->relcp_tol && k > 0) break; /* :132-134 */

    phase_t ph;
    memset(&ph, 0, sizeof(ph));
    phase_build(&ph, &S, b, p); /* :136 */
    double* uext = k == 0 ? brownian_uext(p, b, iter) : NULL; /* :140-144 */
    ph.uext = uext;
    double* Gamma = ALLOC(double, S.C);
    double* g_last = ALLOC(double, S.C);
    for (int64_t c = 0; c < S.C; ++c) Gamma[c] = S.r[c].gamma;
    int64_t its = 0;
    g_trace_phase = k; /* set_recursive_iteration, Hard.cpp:114-116 */
    st = bbpgd(&ph, Gamma, p->bbpgd_tol, p->max_bbpgd_iterations, g_last, &its, &res); /* :153 */
    if (phase_iters && k < max_phase) phase_iters[k] = its;
    if (st != PCC_OK) {
      set_err(err, err_len, "BBPGD did not converge");
      free(Gamma);
      free(g_last);
      free(uext);
      phase_free(&ph);
      total_it += its;
      break;
    }
    total_it += its;
    /* move (Hard.cpp:160-164, ParticleManager.cpp:29-87) */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
      integrate_body(b, i, ph.U + 6 * i, dt);
      for (int kk = 0; kk < 3; ++kk) {
        if (b->force) b->force[3 * i + kk] = ph.F[6 * i + kk];
        if (b->torque) b->torque[3 * i + kk] = ph.F[6 * i + 3 + kk];
        if (b->velocity) b->velocity[3 * i + kk] = ph.U[6 * i + kk];
        if (b->omega) b->omega[3 * i + kk] = ph.U[6 * i + 3 + kk];
      }
      /* setGrowParamsFromSolution (ParticleManager.cpp:89-130) */
      b->ldot[i] = ph.ldot[i];
      if (b->impedance) b->impedance[i] = ph.imp[i];
      if (b->stress) b->stress[i] = ph.sig[i];
    }
    /* updateConstraintsFromSolution (Hard.cpp:49-89) */
    for (int64_t c = 0; c < S.C; ++c) {
      S.r[c].gamma = Gamma[c];
      S.r[c].phi = g_last[c];
    }
    free(Gamma);
    free(g_last);
    free(uext);
    phase_free(&ph);
    if (k < ORACLE_MAX_PHASES) {
      g_phase_time[k] = now_s() - t_phase;
      g_phase_count = k + 1;
    }
    ++k;
  }
  if (st == PCC_OK) {
    for (int64_t i = 0; i < N; ++i) { /* grow (ParticleManager.cpp:132-137, Particle.cpp:59-65) */
      b->length[i] += dt * b->ldot[i];
      if (b->length[i] > 2 * b->l0[i]) b->length[i] = 2 * b->l0[i];
    }
    if (k == p->max_recursive_iterations) { /* :187-190 */
      st = PCC_E_RELCP_MAX_ITERS;
      set_err(err, err_len, "Maximum number of constraint iterations reached");
    }
  }
  result->n_constraints = S.C;
  result->constraint_iterations = k;
  result->reserved = 0;
  result->bbpgd_iterations = total_it;
  result->residual = res;
  result->max_overlap = maxov;
  write_set(cons, &S);
  free(S.r);
  free(S.gid);
  free(S.iter);
  return st;
}

/* solveSoftPotential (Soft.cpp:21-114). */
static pcc_status soft_step(const pcc_params* p, pcc_bodies* b, int32_t iter,
                            pcc_step_result* result, pcc_constraints* cons) {
  const int64_t N = b->n;
  const double dt = p->dt;
  body_view bv = {b->pos, b->quat, b->length, b->diameter};
  for (int64_t i = 0; i < N; ++i) {
    b->ldot[i] = 0.0;
    if (b->num_constraints) b->num_constraints[i] = 0;
  }
  rec_t* recs = NULL;
  int64_t C = detect_records(&bv, N, 0.0, p->linked_cell_size, p->grid_origin,
                             p->warm_start_factor, 0, &recs); /* :35 */
  if (b->num_constraints)
    for (int64_t i = 0; i < C; ++i) {
      b->num_constraints[recs[i].I]++;
      b->num_constraints[recs[i].J]++;
    }
  double maxov = 0.0; /* :37-41 */
  for (int64_t c = 0; c < C; ++c) {
    double ph = recs[c].phi;
    double v = ph < 0 ? -ph : 0;
    maxov = maxov < v ? v : maxov;
  }
  double* F = (double*)calloc((size_t)(6 * N > 0 ? 6 * N : 1), sizeof(double));
  double* sig = (double*)calloc((size_t)(N > 0 ? N : 1), sizeof(double));
  for (int64_t c = 0; c < C; ++c) { /* :46-83 */
    const rec_t* r = &recs[c];
    double overlap = -r->phi;
    if (overlap <= 0.00) continue;
    double R1 = b->diameter[r->I] / 2.0;
    double R2 = b->diameter[r->J] / 2.0;
    double Reff = sqrt(R1 * R2 / (R1 + R2));
    double Fel = Reff * p->kcc * pow(overlap, 1.5);
    double fi[3], fj[3], ti[3], tj[3];
    for (int k = 0; k < 3; ++k) fi[k] = r->n[k] * Fel;
    for (int k = 0; k < 3; ++k) fj[k] = -fi[k];
    cross3(r->rI, fi, ti);
    cross3(r->rJ, fj, tj);
    double* FI = F + 6 * (int64_t)r->I;
    double* FJ = F + 6 * (int64_t)r->J;
    for (int k = 0; k < 3; ++k) {
      FI[k] += fi[k];
      FI[3 + k] += ti[k];
    }
    for (int k = 0; k < 3; ++k) {
      FJ[k] += fj[k];
      FJ[3 + k] += tj[k];
    }
    sig[r->I] += r->sI * Fel;
    sig[r->J] += r->sJ * Fel;
  }
  const double mlam = -p->lambda;
  const double itau = 1 / p->tau;
  double* uext = brownian_uext(p, b, iter); /* :94, before the impedance is updated */
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    double l = b->length[i];
    double mt = 1.0 / (l * p->xi);
    double mr = 12.0 / (l * l * l * p->xi);
    double U[6];
    for (int k = 0; k < 3; ++k) U[k] = (0.0 + mt * F[6 * i + k]) + 0.0; /* calculate_forces :8-19 */
    for (int k = 3; k < 6; ++k) U[k] = (0.0 + mr * F[6 * i + k]) + (uext ? uext[3 * i + k - 3] : 0.0);
    integrate_body(b, i, U, dt); /* :99 */
    for (int k = 0; k < 3; ++k) {
      if (b->force) b->force[3 * i + k] = F[6 * i + k];
      if (b->torque) b->torque[3 * i + k] = F[6 * i + 3 + k];
      if (b->velocity) b->velocity[3 * i + k] = U[k];
      if (b->omega) b->omega[3 * i + k] = U[3 + k];
    }
    double imp = oracle_exp(sig[i] * mlam); /* :102-107 */
    double ld = (l * imp) * itau;
    if (b->impedance) b->impedance[i] = imp;
    if (b->stress) b->stress[i] = sig[i];
    b->ldot[i] = ld;
    b->length[i] += dt * ld; /* grow :111 */
    if (b->length[i] > 2 * b->l0[i]) b->length[i] = 2 * b->l0[i];
  }
  for (int64_t c = 0; c < C; ++c) write_record(cons, c, &recs[c], (int32_t)c, 0);
  result->n_constraints = C;
  result->constraint_iterations = 0;
  result->reserved = 0;
  result->bbpgd_iterations = 0;
  result->residual = maxov;
  result->max_overlap = maxov;
  free(F);
  free(sig);
  free(recs);
  free(uext);
  return PCC_OK;
}

int64_t oracle_phase0_energy(const pcc_params* p, const pcc_bodies* b, const double* gamma,
                             double E[4], double* g) {
  body_view bv = {b->pos, b->quat, b->length, b->diameter};
  cset_t S = {0, 0, NULL, NULL, NULL};
  rec_t* recs = NULL;
  int64_t m = detect_records(&bv, b->n, p->cutoff_phase0, p->linked_cell_size, p->grid_origin,
                             p->warm_start_factor, 0, &recs);
  cset_append(&S, recs, m, 0);
  free(recs);
  if (gamma) {
    phase_t ph;
    memset(&ph, 0, sizeof(ph));
    phase_build(&ph, &S, b, p);
    if (E) energy(&ph, gamma, E);
    if (g) gradient(&ph, gamma, g);
    phase_free(&ph);
  }
  free(S.r);
  free(S.gid);
  free(S.iter);
  return m;
}

pcc_status oracle_step(const pcc_params* p, pcc_bodies* b, int32_t iter, pcc_step_result* result,
                       pcc_constraints* cons, int64_t* phase_iters, int64_t max_phase,
                       char* err, int64_t err_len) {
  memset(result, 0, sizeof(*result));
  if (!(p->temperature >= 0.0)) {
    set_err(err, err_len, "invalid temperature");
    return PCC_E_INVALID_ARGUMENT;
  }
  if (p->mode == PCC_MODE_HARD)
    return hard_step(p, b, iter, result, cons, phase_iters, max_phase, err, err_len);
  if (p->mode == PCC_MODE_SOFT) return soft_step(p, b, iter, result, cons);
  set_err(err, err_len, "Invalid mode");
  return PCC_E_INVALID_MODE;
}

/* ------------------------------------------------------------------------ */
/* Division and gid reassignment (outside the step)                          */
/* ------------------------------------------------------------------------ */
/* splitmix64 finaliser (Steele, Lea, Flood 2014) */
static inline uint64_t split_mix(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

/* Draw k in [0, 1) of parent p (index within its colony) at (seed, step). */
static inline double div_uniform(uint64_t seed, int32_t step, int64_t p, int k) {
  const uint64_t key = split_mix(split_mix(seed) ^ (uint64_t)(uint32_t)step);
  const uint64_t z = split_mix(key ^ ((uint64_t)p * 32u + (uint64_t)k));
  return (double)(z >> 11) * 0x1.0p-53;
}

/* Approximately N(0, 1): Irwin-Hall sum of twelve uniforms minus 6. */
static inline double div_normal(uint64_t seed, int32_t step, int64_t p, int k0) {
  double s = 0.0;
  for (int k = 0; k < 12; ++k) s += div_uniform(seed, step, p, k0 + k);
  return s - 6.0;
}

/* cos and sin of a half-angle |x| <= 0.6 by their Taylor series to x^21
 * (truncation below 1e-25), Horner in x^2 with correctly rounded 1/n!. */
static inline void half_angle_cs(double x, double* c, double* s) {
  const double y = x * x;
  double pc = -0x1.0ce396db7f853p-70;        /* -1/22! */
  pc = pc * y + 0x1.e542ba4020225p-62;       /* 1/20! */
  pc = pc * y - 0x1.6827863b97d97p-53;       /* 1/18! */
  pc = pc * y + 0x1.ae7f3e733b81fp-45;       /* 1/16! */
  pc = pc * y - 0x1.93974a8c07c9dp-37;       /* 1/14! */
  pc = pc * y + 0x1.1eed8eff8d898p-29;       /* 1/12! */
  pc = pc * y - 0x1.27e4fb7789f5cp-22;       /* 1/10! */
  pc = pc * y + 0x1.a01a01a01a01ap-16;       /* 1/8! */
  pc = pc * y - 0x1.6c16c16c16c17p-10;       /* 1/6! */
  pc = pc * y + 0x1.5555555555555p-5;        /* 1/4! */
  pc = pc * y - 0.5;
  *c = pc * y + 1.0;
  double ps = 0x1.71b8ef6dcf572p-66;         /* 1/21! */
  ps = ps * y - 0x1.2f49b46814157p-57;       /* 1/19! */
  ps = ps * y + 0x1.952c77030ad4ap-49;       /* 1/17! */
  ps = ps * y - 0x1.ae7f3e733b81fp-41;       /* 1/15! */
  ps = ps * y + 0x1.6124613a86d09p-33;       /* 1/13! */
  ps = ps * y - 0x1.ae64567f544e4p-26;       /* 1/11! */
  ps = ps * y + 0x1.71de3a556c734p-19;       /* 1/9! */
  ps = ps * y - 0x1.a01a01a01a01ap-13;       /* 1/7! */
  ps = ps * y + 0x1.1111111111111p-7;        /* 1/5! */
  ps = ps * y - 0x1.5555555555555p-3;        /* 1/3! */
  ps = ps * y + 1.0;
  *s = ps * x;
}

/* q <- normalize((cos(d/2), 0, 0, sin(d/2)) * q): an in-plane turn by d
 * about the z axis, renormalised as Particle.cpp:110-120 does. */
static void turn_z(double q[4], double d) {
  double c, s;
  half_angle_cs(0.5 * d, &c, &s);
  double r[4];
  r[0] = c * q[0] - s * q[3];
  r[1] = c * q[1] - s * q[2];
  r[2] = c * q[2] + s * q[1];
  r[3] = c * q[3] + s * q[0];
  const double norm = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + r[3] * r[3]);
  for (int k = 0; k < 4; ++k) q[k] = r[k] / norm;
}

typedef struct {
  uint64_t key;
  int64_t idx;
} xkey_t;

static uint64_t order_key_host(double x) {
  x = x + 0.0;
  uint64_t u;
  memcpy(&u, &x, sizeof(u));
  return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}

static int xkey_cmp(const void* a, const void* b) {
  const xkey_t* x = (const xkey_t*)a;
  const xkey_t* y = (const xkey_t*)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

/* Particle::divide (Particle.cpp:122-150) for every cell with length >= 2 l0,
 * then gid reassignment by ascending x (Domain::assignGlobalIDs,
 * Domain.cpp:492-508; ties keep the staging order).  l1 = U[1-a, 1+a] L/2,
 * l2 = L - l1, centres x -/+ (l/2) t; each child then turns in-plane by
 * angle_noise * N(0,1).  Staging order: every parent's left child in its
 * place, right children appended in parent order.  age (may be NULL): the
 * parent's age grows by one and the new cell copies it (Particle.cpp:144,148). */
int64_t oracle_divide(int64_t n, int64_t cap, double* pos, double* quat, double* length,
                      double* l0, double* diameter, double* ldot, double* age,
                      double* impedance, uint64_t seed, int32_t step, double length_noise,
                      double angle_noise) {
  int64_t d = 0;
  for (int64_t i = 0; i < n; ++i) d += length[i] >= 2 * l0[i];
  const int64_t m = n + d;
  if (m > cap) return -1;
  int64_t r = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (!(length[i] >= 2 * l0[i])) continue;
    const double L = length[i];
    const double half = L / 2;
    const double lo = half * (1.0 - length_noise), hi = half * (1.0 + length_noise);
    const double l1 = lo + div_uniform(seed, step, i, 0) * (hi - lo);
    const double l2 = L - l1;
    double t[3];
    oracle_direction(quat + 4 * i, t);
    const int64_t j = n + r++;
    double ql[4], qr[4];
    for (int k = 0; k < 4; ++k) ql[k] = qr[k] = quat[4 * i + k];
    turn_z(ql, angle_noise * div_normal(seed, step, i, 1));
    turn_z(qr, angle_noise * div_normal(seed, step, i, 13));
    for (int k = 0; k < 3; ++k) {
      const double x = pos[3 * i + k];
      pos[3 * j + k] = x + (l2 / 2) * t[k];
      pos[3 * i + k] = x - (l1 / 2) * t[k];
    }
    for (int k = 0; k < 4; ++k) {
      quat[4 * i + k] = ql[k];
      quat[4 * j + k] = qr[k];
    }
    length[i] = l1;
    length[j] = l2;
    l0[j] = l0[i];
    diameter[j] = diameter[i];
    ldot[j] = ldot[i];
    if (age) {
      age[i] = age[i] + 1;
      age[j] = age[i];
    }
    if (impedance) impedance[j] = 0.0; /* a new Particle's (Particle.cpp:25) */
  }
  /* gids by ascending x */
  xkey_t* ord = (xkey_t*)malloc(sizeof(xkey_t) * (size_t)(m > 0 ? m : 1));
  for (int64_t i = 0; i < m; ++i) {
    ord[i].key = order_key_host(pos[3 * i]);
    ord[i].idx = i;
  }
  qsort(ord, (size_t)m, sizeof(xkey_t), xkey_cmp);
  double* tmp = (double*)malloc(sizeof(double) * 4 * (size_t)(m > 0 ? m : 1));
#define PERMUTE(arr, w)                                                           \
  do {                                                                            \
    for (int64_t i = 0; i < m; ++i)                                               \
      for (int k = 0; k < (w); ++k) tmp[(w) * i + k] = (arr)[(w) * ord[i].idx + k]; \
    memcpy((arr), tmp, sizeof(double) * (w) * (size_t)m);                        \
  } while (0)
  PERMUTE(pos, 3);
  PERMUTE(quat, 4);
  PERMUTE(length, 1);
  PERMUTE(l0, 1);
  PERMUTE(diameter, 1);
  PERMUTE(ldot, 1);
  if (age) PERMUTE(age, 1);
  if (impedance) PERMUTE(impedance, 1);
#undef PERMUTE
  free(tmp);
  free(ord);
  return m;
}
