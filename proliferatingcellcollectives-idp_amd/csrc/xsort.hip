// xsort.hip — orders items by (group, x, tie) with a bucket sort.
//
// Two orders of the hot path are orders by an x coordinate: the contact list
// sorted by contact-point x with ties by (I, J) (CollisionDetector.cpp:94-96),
// and the gids reassigned by ascending body x after division
// (Domain.cpp:492-508).  Both are per colony, and in both the items of a
// colony form one contiguous, ascending range of the item order, which is also
// the colony's range of output positions.  The full key (x, tie) is unique, so
// any exact sort gives the one order; x ties compare equal as doubles (-0 and
// +0 included: the order keys map both to one value).
//
// Instead of a 64-bit radix sort (about a dozen launches of rocPRIM's block
// and merge passes at the sizes of a step), a bucket sort in six small
// launches:
//   1. k_xs_minmax: per colony the smallest and largest order key (block
//      reductions, two atomics per block and colony);
//   2. k_xs_hist: colony g's n_g items get n_g buckets by the linear map
//      floor((x - lo) / (hi - lo) * n_g), which is monotone in x (rounded
//      subtraction, division by a positive number and truncation of a
//      non-negative value are monotone), so bucket order never contradicts x
//      order; items count into their bucket;
//   3. an exclusive scan of the counts gives each bucket's output range;
//   4. k_xs_scatter places the items in their buckets (arrival order);
//   5. k_xs_sort orders each bucket by (x, tie) with one thread per bucket
//      (~1.3 items per bucket in dense colonies: up to 8 sorted in
//      registers, more by insertion sort, O(k^2) for k items, so inputs with
//      many equal x values are slower, never wrong).
#include <rocprim/device/device_scan.hpp>

#include "common.hpp"
#include "internal.hpp"

namespace pcc {

namespace {

inline unsigned nblk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// Inverse of order_key: the double an order key was made from.
__device__ __forceinline__ double key_value(uint64_t k) {
  const uint64_t u = (k & 0x8000000000000000ull) ? (k ^ 0x8000000000000000ull) : ~k;
  return __longlong_as_double((long long)u);
}

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int off) {
  const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, off, PCC_WAVE);
  const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), off, PCC_WAVE);
  return ((uint64_t)hi << 32) | lo;
}

__global__ void k_xs_init(int64_t G, unsigned long long* __restrict__ kmin,
                          unsigned long long* __restrict__ kmax) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < G) {
    kmin[t] = ~0ull;
    kmax[t] = 0ull;
  }
}

// Block b covers items [b * 256 * ipt, (b + 1) * 256 * ipt) in rounds of
// 256 * 8, thread t taking base + t + j * 256 of a round; the eight loads of
// a round are issued together (clamped to the last item; duplicates do not
// change a min or max).  Groups ascend with the item index, so the block
// visits the groups from its first item's to its last item's, one block
// reduction each.  Also clears the bucket counts (n + 1 entries).
__global__ void __launch_bounds__(256) k_xs_minmax(int64_t n, int ipt, const uint64_t* __restrict__ key,
                                                   const int32_t* __restrict__ grp,
                                                   unsigned long long* __restrict__ kmin,
                                                   unsigned long long* __restrict__ kmax,
                                                   int32_t* __restrict__ cnt) {
  constexpr int U = 8;
  __shared__ uint64_t lmin[4], lmax[4];
  const int64_t base = (int64_t)blockIdx.x * 256 * ipt;
  if (base >= n) return;
  const int64_t last = base + 256 * (int64_t)ipt - 1 < n - 1 ? base + 256 * (int64_t)ipt - 1 : n - 1;
  for (int j = 0; j < ipt; ++j) {
    const int64_t r = base + threadIdx.x + (int64_t)j * 256;
    if (r <= last) cnt[r] = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) cnt[n] = 0;
  const int32_t g0 = grp[base], g1 = grp[last];
  const int lane = threadIdx.x & (PCC_WAVE - 1), w = threadIdx.x >> 6;
  for (int32_t g = g0; g <= g1; ++g) {
    uint64_t mn = ~0ull, mx = 0ull;
    for (int j0 = 0; j0 < ipt; j0 += U) {
      uint64_t k[U];
      int32_t q[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        int64_t r = base + threadIdx.x + (int64_t)(j0 + u) * 256;
        r = r <= last ? r : last;
        k[u] = key[r];
        q[u] = grp[r];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const bool in = q[u] == g;
        mn = in && k[u] < mn ? k[u] : mn;
        mx = in && k[u] > mx ? k[u] : mx;
      }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      const uint64_t a = shfl_xor_u64(mn, off), b = shfl_xor_u64(mx, off);
      mn = a < mn ? a : mn;
      mx = b > mx ? b : mx;
    }
    if (lane == 0) {
      lmin[w] = mn;
      lmax[w] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int q = 1; q < 4; ++q) {
        mn = lmin[q] < mn ? lmin[q] : mn;
        mx = lmax[q] > mx ? lmax[q] : mx;
      }
      if (mn <= mx) {  // the group has items in this block
        atomicMin(&kmin[g], (unsigned long long)mn);
        atomicMax(&kmax[g], (unsigned long long)mx);
      }
    }
    __syncthreads();
  }
}

// grp_bkt: the group of each item in, its bucket out.
__global__ void k_xs_hist(int64_t n, const uint64_t* __restrict__ key, int32_t* __restrict__ grp_bkt,
                          const int32_t* __restrict__ gstart,
                          const unsigned long long* __restrict__ kmin,
                          const unsigned long long* __restrict__ kmax, int32_t* __restrict__ cnt) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int32_t g = grp_bkt[r];
  const int64_t s = gstart[g], ng = (int64_t)gstart[g + 1] - s;
  const double lo = key_value(kmin[g]), hi = key_value(kmax[g]);
  int64_t b = 0;
  if (hi > lo) {
    const double t = ((key_value(key[r]) - lo) / (hi - lo)) * (double)ng;  // in [0, ng]
    b = t < (double)ng ? (int64_t)t : ng - 1;
  }
  const int64_t bk = s + b;
  grp_bkt[r] = (int32_t)bk;
  atomicAdd(&cnt[bk], 1);
}

// cnt holds each bucket's count and ends at zero.
__global__ void k_xs_scatter(int64_t n, const int32_t* __restrict__ bkt,
                             const int32_t* __restrict__ off, int32_t* __restrict__ cnt,
                             int32_t* __restrict__ perm) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int32_t b = bkt[r];
  perm[off[b] + atomicSub(&cnt[b], 1) - 1] = (int32_t)r;
}

__device__ __forceinline__ bool xs_less(uint64_t ka, uint64_t ta, uint64_t kb, uint64_t tb) {
  return ka < kb || (ka == kb && ta < tb);
}

// One thread per bucket.  A bucket of up to kSmall items is loaded into
// registers with independent loads and sorted there (a sorting network of
// compare-exchanges over the valid prefix); larger buckets use an insertion
// sort in place.
__global__ void k_xs_sort(int64_t n, const int32_t* __restrict__ off,
                          const uint64_t* __restrict__ key, const uint64_t* __restrict__ tie,
                          int32_t* __restrict__ perm) {
  constexpr int kSmall = 8;
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  const int32_t lo = off[b], hi = off[b + 1], m = hi - lo;
  if (m <= 1) return;
  if (m <= kSmall) {
    int32_t x[kSmall];
    uint64_t kx[kSmall], tx[kSmall];
#pragma unroll
    for (int i = 0; i < kSmall; ++i) x[i] = perm[lo + (i < m ? i : m - 1)];
#pragma unroll
    for (int i = 0; i < kSmall; ++i) {
      kx[i] = key[x[i]];
      tx[i] = tie ? tie[x[i]] : (uint64_t)x[i];
    }
    // odd-even transposition over the first m slots (m rounds)
#pragma unroll
    for (int r = 0; r < kSmall; ++r) {
#pragma unroll
      for (int i = r & 1; i + 1 < kSmall; i += 2) {
        const bool sw = i + 1 < m && xs_less(kx[i + 1], tx[i + 1], kx[i], tx[i]);
        const uint64_t k0 = kx[i], t0 = tx[i];
        const int32_t x0 = x[i];
        kx[i] = sw ? kx[i + 1] : k0;
        tx[i] = sw ? tx[i + 1] : t0;
        x[i] = sw ? x[i + 1] : x0;
        kx[i + 1] = sw ? k0 : kx[i + 1];
        tx[i + 1] = sw ? t0 : tx[i + 1];
        x[i + 1] = sw ? x0 : x[i + 1];
      }
    }
#pragma unroll
    for (int i = 0; i < kSmall; ++i)
      if (i < m) perm[lo + i] = x[i];
    return;
  }
  for (int32_t i = lo + 1; i < hi; ++i) {
    const int32_t x = perm[i];
    const uint64_t kx = key[x], tx = tie ? tie[x] : (uint64_t)x;
    int32_t j = i - 1;
    while (j >= lo) {
      const int32_t y = perm[j];
      if (!xs_less(kx, tx, key[y], tie ? tie[y] : (uint64_t)y)) break;
      perm[j + 1] = y;
      --j;
    }
    perm[j + 1] = x;
  }
}

}  // namespace

size_t xsort_tmp_bytes(int64_t n) {
  size_t bytes = 0;
  int32_t* p = nullptr;
  PCC_HIP(rocprim::exclusive_scan(nullptr, bytes, p, p, int32_t(0), (size_t)(n + 1),
                                  rocprim::plus<int32_t>(), (hipStream_t)0));
  return bytes;
}

void xsort_run(const XSort& a) {
  if (a.n <= 0) return;
  const hipStream_t s = a.s;
  if (a.G > 0 && !a.bounds_clear) k_xs_init<<<nblk(a.G), 256, 0, s>>>(a.G, a.kmin, a.kmax);
  // at most ~512 reduction blocks (two atomics each per colony they cover)
  int ipt = 8;
  while ((a.n + 256 * (int64_t)ipt - 1) / (256 * (int64_t)ipt) > 512) ipt *= 2;
  const int64_t span = 256 * (int64_t)ipt;
  k_xs_minmax<<<(unsigned)((a.n + span - 1) / span), 256, 0, s>>>(a.n, ipt, a.key, a.grp, a.kmin,
                                                                   a.kmax, a.cnt);
  k_xs_hist<<<nblk(a.n), 256, 0, s>>>(a.n, a.key, a.grp, a.gstart, a.kmin, a.kmax, a.cnt);
  size_t bytes = 0;
  PCC_HIP(rocprim::exclusive_scan(nullptr, bytes, a.cnt, a.off, int32_t(0), (size_t)(a.n + 1),
                                  rocprim::plus<int32_t>(), s));
  if (bytes > a.tmp_bytes) throw StatusError{PCC_E_DEVICE, "internal: x-sort storage too small"};
  PCC_HIP(rocprim::exclusive_scan(a.tmp, bytes, a.cnt, a.off, int32_t(0), (size_t)(a.n + 1),
                                  rocprim::plus<int32_t>(), s));
  k_xs_scatter<<<nblk(a.n), 256, 0, s>>>(a.n, a.grp, a.off, a.cnt, a.perm);
  k_xs_sort<<<nblk(a.n), 256, 0, s>>>(a.n, a.off, a.key, a.tie, a.perm);
}

}  // namespace pcc
