// noalloc.hip — pcc_detect_noalloc: contact detection that allocates nothing.
//
// The detection batch of pcc_detect (CollisionDetector::detectCollisions,
// CollisionDetector.cpp:84-111, through the same kernels and the same
// detect_count / detect_write pipeline) for one colony whose state the caller
// already holds in device memory, written into caller-owned device arrays.
// Every buffer the pipeline needs beyond the inputs and outputs lives in one
// file-scope device array, `scratch`, of fixed size: a grid and records for
// at most PCC_NOALLOC_MAX_BODIES bodies, PCC_NOALLOC_MAX_BLOCK_CANDIDATES
// candidate pairs per filter workgroup and PCC_NOALLOC_MAX_CONTACTS contacts,
// and the rocPRIM temporary storage.  Its device address is looked
// up once per device with hipGetSymbolAddress and kept; host code never
// passes the array's name itself to a kernel or a copy.
//
// Each GPU holds its own copy of `scratch` (reserved when this code object is
// loaded on that GPU, kept until the process ends).  Calls on one GPU share
// it: overlapping calls race, and every call re-initialises what must start
// at zero (the grid counters and slot counts are cleared by fill kernels in
// the pipeline, the colony map and the planarity flag here).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <mutex>
#include <string>

#include "internal.hpp"

namespace pcc {
namespace {

constexpr int64_t kN = PCC_NOALLOC_MAX_BODIES;
constexpr int64_t kM = PCC_NOALLOC_MAX_CONTACTS;
constexpr int64_t kS = 27;  // neighbour cells per body, worst case (3D)
constexpr int64_t kT = [] {  // hash table size of kN bodies (table_size_for)
  int64_t t = 1024;
  while (t < 2 * kN) t <<= 1;
  return t;
}();
constexpr int64_t kGp = (kN * kS + 255) / 256 / 8 * 8 + 8;  // filter workgroups, at most
constexpr int64_t kTmp = 8ll << 20;  // rocPRIM temporary storage (checked at set-up)
constexpr int kBodyFieldsRead = 9;   // x y z qw qx qy qz length diameter

constexpr int64_t pad(int64_t b) { return (b + 255) / 256 * 256; }

// Byte layout of the scratch array.
struct Layout {
  int64_t body, binst, idev, flag, cell4, geo, scell, sgeo, frec, bucket, perm, bstart, bcursor, dslot, dslot_off,
      cstart, cand, accept, bcount, dcount, bbounds, rI, rJ, rec, ridx, ridx2, rkey, rkey2, xs_cnt,
      xs_off, xs_bounds, tmp, total;
};
constexpr Layout make_layout() {
  Layout L{};
  int64_t o = 0;
  auto take = [&o](int64_t bytes) {
    const int64_t at = o;
    o += pad(bytes);
    return at;
  };
  L.body = take(8 * kBodyFieldsRead * kN);
  L.binst = take(4 * kN);
  L.idev = take((int64_t)sizeof(InstDev));
  L.flag = take(4);
  L.cell4 = take(16 * kN);
  L.geo = take(64 * kN);
  L.scell = take(16 * kN);
  L.sgeo = take(64 * kN);
  L.frec = take(16 * kN);
  L.bucket = take(4 * (kN + 1));
  L.perm = take(4 * (kN + 1));
  L.bstart = take(4 * (kT + 1));
  L.bcursor = take(4 * (kT + 1));
  L.dslot = take(4 * (kN * kS + 1));
  L.dslot_off = take(4 * (kN * kS + 1));
  L.cstart = take(4 * (kN * kS + 1));
  L.cand = take(8 * kGp * kCandBlock);
  L.accept = take(kGp * kCandBlock);
  L.bcount = take(4 * kGp);
  L.dcount = take(4 * (kN + 1));
  L.bbounds = take(4 * 3);
  L.rI = take(4 * (kM + 1));
  L.rJ = take(4 * (kM + 1));
  L.rec = take(8 * 16 * (kM + 1));
  L.ridx = take(4 * (kM + 1));
  L.ridx2 = take(4 * (kM + 1));
  L.rkey = take(8 * (kM + 1));
  L.rkey2 = take(8 * (kM + 1));
  L.xs_cnt = take(4 * (kM + 1));
  L.xs_off = take(4 * (kM + 1));
  L.xs_bounds = take(8 * 2);
  L.tmp = take(kTmp);
  L.total = o;
  return L;
}
constexpr Layout kL = make_layout();
static_assert(kCandBlock == PCC_NOALLOC_MAX_BLOCK_CANDIDATES, "candidate region");
static_assert(kL.total < (int64_t)96 << 20, "scratch of the no-allocation detection");

}  // namespace
}  // namespace pcc

// The scratch: one per device, device memory for the life of the process.
static __device__ __align__(256) unsigned char scratch[pcc::kL.total];

namespace pcc {
namespace {

// Device address of `scratch` per device, looked up once with the device
// current (hipGetSymbolAddress); host code uses only these.
std::mutex& scratch_mutex() {
  static std::mutex m;
  return m;
}
std::array<unsigned char*, 64>& scratch_ptrs() {
  static std::array<unsigned char*, 64> p{};
  return p;
}

unsigned char* scratch_on(int device) {
  if (device < 0 || device >= 64)
    throw StatusError{PCC_E_INVALID_ARGUMENT, "device index out of range"};
  std::lock_guard<std::mutex> g(scratch_mutex());
  unsigned char*& p = scratch_ptrs()[device];
  if (p) return p;
  void* q = nullptr;
  const hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !q) {
    (void)hipGetLastError();
    throw StatusError{PCC_E_DEVICE, std::string("hipGetSymbolAddress(scratch) failed: ") +
                                        hipGetErrorString(e)};
  }
  if (detect_scan_bytes(kN, (int)kS, kT) > (size_t)kTmp || detect_sort_bytes(kM, 1) > (size_t)kTmp)
    throw StatusError{PCC_E_DEVICE, "internal: the scratch's sort storage is too small"};
  p = static_cast<unsigned char*>(q);
  return p;
}

inline unsigned nblk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// The colony record and the cleared planarity flag.
__global__ void k_na_setup(InstDev v, InstDev* __restrict__ idev, int32_t* __restrict__ flag) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    *idev = v;
    *flag = 0;
  }
}

// Interleaved caller state -> the SoA fields the detection kernels read;
// every body in colony 0; flag = 1 if some body has z != 0.
__global__ void k_na_load(int64_t n, const double* __restrict__ pos,
                          const double* __restrict__ quat, const double* __restrict__ len,
                          const double* __restrict__ diam, BodyArrays b,
                          int32_t* __restrict__ binst, int32_t* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double z = pos[3 * i + 2];
  b.px[i] = pos[3 * i];
  b.py[i] = pos[3 * i + 1];
  b.pz[i] = z;
  b.qw[i] = quat[4 * i];
  b.qx[i] = quat[4 * i + 1];
  b.qy[i] = quat[4 * i + 2];
  b.qz[i] = quat[4 * i + 3];
  b.len[i] = len[i];
  b.diam[i] = diam[i];
  binst[i] = 0;
  if (z != 0.0) atomicOr(flag, 1);
}

struct DeviceScope {
  int prev = -1;
  explicit DeviceScope(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    PCC_HIP(hipSetDevice(dev));
  }
  ~DeviceScope() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

pcc_status detect_noalloc(int device, hipStream_t s, const pcc_params& p, double cutoff,
                          const pcc_device_bodies& in, const pcc_constraints* out,
                          int64_t* count) {
  const int64_t n = in.n;
  DeviceScope dscope(device);
  unsigned char* base = scratch_on(device);
  auto at = [base](int64_t off) { return static_cast<void*>(base + off); };
  double* f = static_cast<double*>(at(kL.body));
  BodyArrays b{};
  b.px = f;
  b.py = f + kN;
  b.pz = f + 2 * kN;
  b.qw = f + 3 * kN;
  b.qx = f + 4 * kN;
  b.qy = f + 5 * kN;
  b.qz = f + 6 * kN;
  b.len = f + 7 * kN;
  b.diam = f + 8 * kN;  // l0, ldot, age, imp: not read by detection
  InstDev v{};
  v.boff = 0;
  v.N = n;
  v.cutoff = cutoff;
  v.warm = p.warm_start_factor;
  v.origin = p.grid_origin;
  v.h = p.linked_cell_size;
  v.flags = kInstActive;
  InstDev* idev = static_cast<InstDev*>(at(kL.idev));
  int32_t* flag = static_cast<int32_t*>(at(kL.flag));
  int32_t* binst = static_cast<int32_t*>(at(kL.binst));
  k_na_setup<<<1, 64, 0, s>>>(v, idev, flag);
  if (n > 0)
    k_na_load<<<nblk(n), 256, 0, s>>>(n, in.pos, in.quat, in.length, in.diameter, b, binst, flag);
  int32_t nonflat = 0;
  PCC_HIP(hipMemcpyAsync(&nonflat, flag, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  PCC_HIP(hipStreamSynchronize(s));
  DetectBufs d;
  d.N = n;
  d.body = b;
  d.binst = binst;
  d.idev = idev;
  d.B = 1;
  d.S = nonflat ? 27 : 9;  // every body in z = 0: only that layer's 9 cells hold neighbours
  d.T = table_size_for(n);
  d.cell4 = static_cast<int4*>(at(kL.cell4));
  d.geo = static_cast<double2*>(at(kL.geo));
  d.scell = static_cast<int4*>(at(kL.scell));
  d.sgeo = static_cast<double2*>(at(kL.sgeo));
  d.frec = static_cast<float4*>(at(kL.frec));
  d.bucket = static_cast<uint32_t*>(at(kL.bucket));
  d.perm = static_cast<int32_t*>(at(kL.perm));
  d.bstart = static_cast<int32_t*>(at(kL.bstart));
  d.bcursor = static_cast<int32_t*>(at(kL.bcursor));
  d.dslot = static_cast<int32_t*>(at(kL.dslot));
  d.dslot_off = static_cast<int32_t*>(at(kL.dslot_off));
  d.cstart = static_cast<int32_t*>(at(kL.cstart));
  d.cand = static_cast<int2*>(at(kL.cand));
  d.accept = static_cast<uint8_t*>(at(kL.accept));
  d.bcount = static_cast<int32_t*>(at(kL.bcount));
  d.cand_block = kCandBlock;
  d.dcount = static_cast<int32_t*>(at(kL.dcount));
  d.bbounds = static_cast<int32_t*>(at(kL.bbounds));
  d.rI = static_cast<int32_t*>(at(kL.rI));
  d.rJ = static_cast<int32_t*>(at(kL.rJ));
  d.rec = static_cast<double*>(at(kL.rec));
  d.ridx = static_cast<int32_t*>(at(kL.ridx));
  d.ridx2 = static_cast<int32_t*>(at(kL.ridx2));
  d.rkey = static_cast<uint64_t*>(at(kL.rkey));
  d.rkey2 = static_cast<uint64_t*>(at(kL.rkey2));
  d.xs_cnt = static_cast<int32_t*>(at(kL.xs_cnt));
  d.xs_off = static_cast<int32_t*>(at(kL.xs_off));
  d.xs_kmin = static_cast<unsigned long long*>(at(kL.xs_bounds));
  d.xs_kmax = d.xs_kmin + 1;
  d.tmp = at(kL.tmp);
  d.tmp_bytes = (size_t)kTmp;
  d.s = s;
  if (d.T > kT) throw StatusError{PCC_E_DEVICE, "internal: hash table beyond the scratch"};
  int32_t bounds[3] = {0, 0, 0};
  if (filter_grid(n, d.S) > kGp) throw StatusError{PCC_E_DEVICE, "internal: filter grid beyond the scratch"};
  if (!detect_count(d, bounds)) {
    thread_error() = "pcc_detect_noalloc: " + std::to_string(bounds[2]) +
                     " candidate pairs in one block of 256 (body, neighbour cell) slots exceed "
                     "PCC_NOALLOC_MAX_BLOCK_CANDIDATES (" + std::to_string(kCandBlock) + ")";
    return PCC_E_INVALID_ARGUMENT;
  }
  const int64_t m = bounds[1];
  if (count) *count = m;
  if (m > kM) {
    thread_error() = "pcc_detect_noalloc: " + std::to_string(m) +
                     " contacts exceed PCC_NOALLOC_MAX_CONTACTS (" + std::to_string(kM) + ")";
    return PCC_E_INVALID_ARGUMENT;
  }
  detect_write(d, m);
  if (out) detect_records_out(d, m, *out);
  PCC_HIP(hipStreamSynchronize(s));
  return PCC_OK;
}

}  // namespace
}  // namespace pcc

extern "C" pcc_status pcc_detect_noalloc(int device, void* stream, const pcc_params* params,
                                         double cutoff, const pcc_device_bodies* bodies,
                                         const pcc_constraints* out, int64_t* count) {
  using namespace pcc;
  auto bad = [](const char* why) {
    thread_error() = std::string("pcc_detect_noalloc: ") + why;
    return PCC_E_INVALID_ARGUMENT;
  };
  if (!params || !bodies) return bad("params and bodies are required");
  if (bodies->n < 0) return bad("negative body count");
  if (bodies->n > PCC_NOALLOC_MAX_BODIES)
    return bad("more bodies than PCC_NOALLOC_MAX_BODIES");
  if (bodies->n > 0 && (!bodies->pos || !bodies->quat || !bodies->length || !bodies->diameter))
    return bad("missing body array");
  if (out && out->capacity < 0) return bad("negative capacity");
  if (!(cutoff >= 0.0) || !std::isfinite(cutoff)) return bad("invalid cutoff");
  if (!(params->linked_cell_size > 0.0) || !std::isfinite(params->grid_origin))
    return bad("invalid grid parameters");
  try {
    return detect_noalloc(device, static_cast<hipStream_t>(stream), *params, cutoff, *bodies, out,
                          count);
  } catch (const HipError& e) {
    thread_error() = std::string(e.what) + ": " + hipGetErrorString(e.err);
    return PCC_E_DEVICE;
  } catch (const StatusError& e) {
    thread_error() = e.msg;
    return e.status;
  } catch (const std::exception& e) {
    thread_error() = e.what();
    return PCC_E_DEVICE;
  }
}
