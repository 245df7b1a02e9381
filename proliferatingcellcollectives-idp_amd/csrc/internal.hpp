// internal.hpp — libpcc context, device arrays and kernel launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <tuple>
#include <utility>
#include <string>
#include <vector>

#include "pcc.h"

namespace pcc {

// One lock for the whole library.  HIP refuses device-wide runtime calls
// (allocation, freeing, device synchronisation, pinned host memory, stream
// and event creation) on every thread while any stream of the process is being
// captured into a graph, so graph captures and those calls take turns.  A
// capture holds the lock only while it records one chunk of kernel launches.
// Recursive: a holder may call into code that takes it again.
std::recursive_mutex& device_mutex();
using DeviceLock = std::lock_guard<std::recursive_mutex>;

struct HipError {
  hipError_t err;
  const char* what;
};

#define PCC_HIP(call)                                                   \
  do {                                                                  \
    hipError_t e_ = (call);                                             \
    if (e_ != hipSuccess) throw ::pcc::HipError{e_, #call};             \
  } while (0)

struct StatusError {
  pcc_status status;
  std::string msg;
};

// Device buffers are allocated from the device's stream-ordered pool on the
// stream of the context whose API call runs on this thread (set by
// StreamScope): a buffer that grows is freed in that stream's order, after the
// work already queued on it, so growing never waits for other contexts'
// streams (a plain hipFree synchronises the whole device).  Outside an API
// call (no stream set) the plain calls are used.
// Message of the calling thread's last failed context-free call.
std::string& thread_error();

hipStream_t& alloc_stream();
struct StreamScope {
  hipStream_t prev;
  explicit StreamScope(hipStream_t s) : prev(alloc_stream()) { alloc_stream() = s; }
  ~StreamScope() { alloc_stream() = prev; }
};
// *pooled receives whether the buffer came from the stream-ordered pool;
// dev_free must be given the same flag.
void* dev_alloc(size_t bytes, bool* pooled);
void dev_free(void* p, bool pooled);

// Leading entries of a count array known to be zero.  The scatters that
// consume counts (atomicSub down to zero) leave their arrays clear, so a fill
// is needed only for entries never cleared since the array was allocated.
// Keyed by the array's address: a grown array is a new allocation made while
// the old one is still held, so it never matches the address last marked.
struct ZeroTrack {
  const void* p = nullptr;
  int64_t n = 0;
  bool known(const void* q, int64_t k) const { return q == p && k <= n; }
  void mark(const void* q, int64_t k) {
    if (q != p) {
      p = q;
      n = 0;
    }
    n = k > n ? k : n;
  }
  void clear() { p = nullptr; n = 0; }
};

// Growable device array (never shrinks; contents are not preserved on growth
// unless grow_keep is used).  Capacity doubles, so regrowth is rare.
template <class T>
struct DArr {
  T* p = nullptr;
  int64_t cap = 0;
  bool pooled = false;
  DArr() = default;
  DArr(const DArr&) = delete;
  DArr& operator=(const DArr&) = delete;
  ~DArr() { release(); }
  void ensure(int64_t n) {
    if (n <= cap) return;
    int64_t nc = cap > 0 ? cap : 1024;
    while (nc < n) nc *= 2;
    DeviceLock lock(device_mutex());
    bool pl = false;
    T* q = static_cast<T*>(dev_alloc(sizeof(T) * (size_t)nc, &pl));
    if (p) dev_free(p, pooled);  // ordered after the queued work that may read it
    p = q;
    cap = nc;
    pooled = pl;
  }
  // Grows keeping the first `keep` elements; s must be the stream the array
  // is used on.
  void grow_keep(int64_t n, int64_t keep, hipStream_t s) {
    if (n <= cap) return;
    int64_t nc = cap > 0 ? cap : 1024;
    while (nc < n) nc *= 2;
    DeviceLock lock(device_mutex());
    bool pl = false;
    T* q = static_cast<T*>(dev_alloc(sizeof(T) * (size_t)nc, &pl));
    if (p && keep > 0) PCC_HIP(hipMemcpyAsync(q, p, sizeof(T) * (size_t)keep, hipMemcpyDeviceToDevice, s));
    if (p) dev_free(p, pooled);
    p = q;
    cap = nc;
    pooled = pl;
  }
  void release() {
    if (p) {
      DeviceLock lock(device_mutex());
      dev_free(p, pooled);
    }
    p = nullptr;
    cap = 0;
    pooled = false;
  }
  // Exchanges buffers with another array (ping-pong without copies).
  void swap(DArr& o) {
    std::swap(p, o.p);
    std::swap(cap, o.cap);
    std::swap(pooled, o.pooled);
  }
};

// Body state, FP64 SoA, gid order.
struct BodyArrays {
  double *px, *py, *pz, *qw, *qx, *qy, *qz, *len, *l0, *diam, *ldot;
  double* age;  // generation, an integer held exactly
  double* imp;  // impedance of the last step (input of the Brownian velocity)
};
static constexpr int kBodyFields = 13;
static constexpr int kBodyVariants = PCC_BODY_VARIANTS;
int body_chunk_of(int variant, bool planar);
int body_threads_of(int variant);

// Device-side solver scalars of one instance's BBPGD solve.
struct SolverState {
  double alpha;
  double res;
  double maxov;
  int32_t skip;       // instance takes no part in this solve
  int32_t pad0;
  int64_t iter;       // BBPGD iteration about to run (0 = initial gradient)
  int32_t converged;  // residual <= tol reached at iteration `iter`
  int32_t failed;     // iteration == max
};

// Instance flags for the current ReLCP phase.
enum { kInstActive = 1, kInstInteg = 2, kInstGrow = 4 };

// One colony of a context as the kernels see it.  A context holds B >= 1
// independent colonies; their bodies are concatenated (instance b owns bodies
// [boff, boff + N)) and the solver order is instance-major (instance b owns
// solver positions [coff, coff + C)), as is the constraint set at the end of
// a step, so every per-instance sum and reduction has exactly the order of a
// context holding that colony alone.
struct InstDev {
  int64_t boff, N;
  int64_t coff, C;
  int64_t G, K;      // BB reduction: G logical blocks of K tiles of 256 constraints
  int64_t lb;        // first logical constraint block (= first partial slot)
  int64_t bb;        // first logical body block
  int64_t bstart;    // first record of the instance in the sorted detection batch
  int64_t maxit;
  double dt, xi, itau, mlam, kcc, tol, cutoff, warm, origin, h;
  double lam, tau;   // for the energy's growth term dt / (lambda tau)
  double temp;       // Brownian temperature
  uint64_t bseed;    // Brownian draws: seed
  int64_t bstep;     // Brownian draws: the colony's step count
  int32_t mono;      // Brownian draws: monolayer (omega_z only)
  int32_t pad1;
  int64_t phase;     // ReLCP phase of the current solve
  int32_t flags;
  int32_t check;     // residual checked every `check` iterations (10; 1 while tracing)
};

// Constraint set of the current step (all ReLCP phases), FP64 SoA.
struct ConstraintSet {
  DArr<int32_t> I, J, gid, iter, inst;
  DArr<double> phi, gamma, n, rI, rJ, p, sI, sJ;  // n, rI, rJ, p hold 3*C
  // packed operator row per constraint: n[3], rI x n, rJ x -n, sI, sJ, pad
  DArr<double> coef64;                           // 12*C
  DArr<float> coef32;                            // 12*C (mixed precision)
  int64_t C = 0;
  int64_t cap = 0;
};

// Per-kernel HIP-event accounting on the context stream (enabled on demand).
struct Timing {
  struct Pending {
    hipEvent_t a, b;
    int kernel;
    double bytes;
  };
  bool enabled = false;
  int64_t launches[PCC_K_COUNT] = {0};
  double total_ms[PCC_K_COUNT] = {0};
  double bytes[PCC_K_COUNT] = {0};
  std::vector<hipEvent_t> free_events;
  std::vector<Pending> pending;
};

// Host-side record of one colony.
struct Instance {
  pcc_params p;
  int64_t boff = 0, N = 0;
  int64_t coff = 0, C = 0;   // constraint range of the current step
  int64_t m = 0, bstart = 0; // current detection batch
  int64_t coff_new_ = 0;     // range offset once the batch is appended
  double cutoff = 0.0;       // contact cutoff of the current detection
  int64_t G = 1, K = 0, lb = 0, bb = 0, nbb = 0;
  int32_t flags = 0;
  int64_t nstep = 0;         // steps taken since the colony was set or added (Brownian key)
  // per-step outcome (SolverSolution, Hard.cpp:192)
  bool active = false;
  int32_t k = 0;
  int64_t its = 0;
  double res = 0.0, maxov = 0.0;
  pcc_status status = PCC_OK;
  std::string error;
  std::vector<int64_t> phase_iters, phase_C;
  // last BBPGD solve
  int64_t solve_its = 0;
  double solve_res = 0.0;
  bool solve_failed = false;
};

}  // namespace pcc

struct pcc_ctx {
  pcc_params params;
  int device = 0;
  hipStream_t stream = nullptr;
  std::string last_error;
  int64_t N = 0;
  bool has_state = false;
  bool planar = false;        // current step uses the 2D monolayer kernels
  bool detect_flat = false;   // every body has z = 0: detection visits 9 neighbour cells
  bool allow_planar = true;   // PCC_PLANAR=0 in the environment forces the 3D kernels
  bool use_graphs = true;     // BBPGD chunks as captured HIP graphs (PCC_GRAPH=0: plain launches)
  // one-launch BBPGD when every active colony has at most persist_max_c
  // constraints and persist_max_n bodies (PCC_PERSIST=0: off); measured
  // faster than the multi-kernel path up to ~2000 constraints, slower from
  // ~7000 (one workgroup per colony)
  bool use_persist = true;
  int64_t persist_max_c = 4096;
  int64_t persist_max_n = 1024;
  // hard step of a batch: colonies of at least solo_cells cells run their
  // ReLCP loop alone, one after another; the others share one lockstep loop
  // (PCC_SOLO_CELLS=n sets the threshold, 0 keeps every colony in lockstep)
  int64_t solo_cells = 150000;
  // BBPGD grids rebuilt over the colonies still iterating once they hold at
  // most compact_ratio of the body blocks, for solves of at least
  // compact_min_blocks body blocks (PCC_COMPACT=0: off;
  // PCC_COMPACT_MIN_BLOCKS=n; PCC_COMPACT_RATIO=r, measured flat from 0.5 to
  // 0.9 on a 128-colony ensemble)
  bool compact = true;
  int64_t compact_min_blocks = 64;
  double compact_ratio = 0.5;  // PCC_COMPACT_RATIO
  // BBPGD iterations per speculative chunk (one captured graph; PCC_CHUNK_ITERS)
  int chunk_iters = 10;
  int64_t compactions = 0;
  // autotuner (PCC_AUTOTUNE=0: off): body-kernel staging size
  bool autotune = true;
  bool body_tuned = false;
  bool body_from_cache = false;  // the variant came from the process-wide tuning memory
  std::tuple<std::string, int, int, int, int> tuned_key;  // class of the last tuned solve
  int body_variant = 1;
  float body_variant_ms[PCC_BODY_VARIANTS] = {};

  // body state + snapshot + outputs
  pcc::DArr<double> body;      // kBodyFields * N
  pcc::DArr<double> snap;      // kBodyFields * N
  bool has_snapshot = false;
  pcc::DArr<double> F, U, sig, ldot_new;  // 6N, 6N, N, N
  pcc::DArr<double> uext;  // 3N: Brownian angular velocity of the step (T > 0)
  bool brownian = false;   // uext holds this step's draws
  std::vector<int64_t> snap_nstep;  // step counts at pcc_snapshot
  pcc::DArr<int32_t> nc;                        // N

  // grid
  pcc::DArr<int32_t> cell;     // 4N: cell x, y, z and colony per body
  pcc::DArr<uint32_t> bucket;  // N
  pcc::DArr<int32_t> perm;     // N
  pcc::DArr<int32_t> bstart;   // T+1
  pcc::DArr<int32_t> bcursor;  // T
  pcc::ZeroTrack bcursor_zero;  // bucket counts left at zero by the last detection
  int64_t table_size = 0;

  // detection batch
  pcc::DArr<int32_t> dcount;   // N+1 (per-body record offsets)
  pcc::DArr<int32_t> dslot, dslot_off;  // per (body, neighbour cell) counts and offsets
  pcc::DArr<int32_t> dcstart;  // per (body, neighbour cell): first candidate of its run
  pcc::DArr<int32_t> dcand;    // candidate pairs (slot, bucket position of J)
  pcc::DArr<uint8_t> daccept;  // per candidate: 1 if a contact
  pcc::DArr<int32_t> dbcount;  // per filter workgroup: candidates in its region
  int32_t cand_block = 0;      // candidate region per filter workgroup (grows on overflow)
  // division scratch (kept between calls)
  pcc::DArr<int32_t> dv_flag, dv_off, dv_bounds, dv_perm, dv_sinst;
  pcc::DArr<double> dv_stage, dv_body;
  pcc::DArr<uint64_t> dv_key;
  pcc::DArr<uint8_t> dv_col;
  pcc::DArr<double> dgeo;    // per body: x, y, z, t (3), length, diameter (detection)
  pcc::DArr<int32_t> scell;  // cell records in bucket order
  pcc::DArr<double> sgeo;    // detection records in bucket order
  pcc::DArr<float> sfrec;    // FP32 filter records in bucket order
  pcc::DArr<int32_t> rI_, rJ_; // raw pair ids
  pcc::DArr<double> rrec;      // raw geometry: 16 doubles per contact
  pcc::DArr<int32_t> ridx, ridx2;
  pcc::DArr<uint64_t> rkey, rkey2;
  pcc::DArr<int32_t> xs_cnt, xs_off;          // bucket counts and offsets of the x orders
  pcc::DArr<unsigned long long> xs_bounds;    // 2 per colony: smallest and largest x key
  pcc::DArr<uint8_t> sort_tmp;
  int64_t batch_count = 0;

  // constraint set and phase operators
  pcc::ConstraintSet S;
  pcc::DArr<int32_t> rowptr, inccur, inc;  // N+1, N, 2C
  pcc::ZeroTrack inccur_zero;              // incidence counts left at zero by the last build
  pcc::DArr<double> increc;                // per incidence: 8 (3D) or 4 (planar) doubles
  pcc::DArr<float> increc32;               // per incidence: 8 or 4 floats
  pcc::DArr<int32_t> incidx;               // planar: solver position per incidence
  pcc::DArr<double> gold_e;                // gamma_old per incidence
  pcc::DArr<double> fold;                  // mixed: D^T gamma_old per body (6, planar 3)
  pcc::DArr<int32_t> plist;                // colonies of a persistent BBPGD launch
  pcc::DArr<int32_t> sord, spos;           // solver order and its inverse
  pcc::DArr<int32_t> skey, skey2, sval;    // sort scratch
  pcc::DArr<int2> sIJ;                     // solver order: (I, J) per constraint
  pcc::DArr<double> sphi;                  // solver order
  pcc::DArr<double> prow;                  // operator rows, solver order (6 planar, 12 3D)
  pcc::DArr<float> prow32;
  pcc::DArr<int32_t> flag;                 // small device flags
  pcc::DArr<double> iter;                  // 2 per constraint: (gamma, g), solver order
  pcc::DArr<double> gprev;                 // lagging iterate at convergence checks
  pcc::DArr<double> ubuf;                  // 8 per body: U[6], dldot, pad
  pcc::DArr<double> partials;              // 4 * nblocks (5 while tracing)
  // BBPGD trace: per colony `trace_cap` records and a count per step
  bool trace_on = false;
  int64_t trace_cap = 0;
  int64_t trace_B = 0;  // colonies traced in the last step (0: last step untraced)
  pcc::DArr<pcc_bbpgd_record> trace;
  pcc::DArr<int64_t> trace_n;
  pcc::DArr<double> epart;                 // energy partials per body block (2 * nbb)
  pcc::DArr<pcc::SolverState> state;       // B
  pcc::SolverState* state_host = nullptr;  // pinned, 2*B slots
  int64_t state_host_cap = 0;
  pcc::DArr<double> scratch_d;             // small reductions

  // soft
  pcc::DArr<double> hertz;                 // 16 per constraint
  pcc::DArr<float> hertz32;                // 16 per constraint

  pcc::Timing timing;

  // colonies of this context (one for the single-colony entry points)
  std::vector<pcc::Instance> inst;
  std::vector<pcc::InstDev> inst_host;
  // pinned staging of the colony records: kInstSlots slots of inst_pin_cap
  // records, each reused once the copy recorded on its event has run
  pcc::InstDev* inst_pin = nullptr;
  int64_t inst_pin_cap = 0;
  int inst_slot = 0;
  hipEvent_t inst_ev[4] = {nullptr, nullptr, nullptr, nullptr};
  // pinned copy of the per-colony max overlap (max_overlap_launch)
  unsigned long long* ov_pin = nullptr;
  int64_t ov_pin_cap = 0;
  // pinned target of the small synchronous read-backs (detection and
  // division counts): a copy into pageable memory takes a staging detour
  int32_t* rb_pin = nullptr;
  int64_t rb_pin_cap = 0;
  // planarity of the body state as of the end of the last step (cleared by
  // every host write of body state; kept by device division)
  bool flat_known = false;
  bool flat_value = false;
  pcc::DArr<pcc::InstDev> idev;            // B
  pcc::DArr<int32_t> binst;                // N: instance of every body
  pcc::DArr<int32_t> bmap;                 // logical body block -> instance
  pcc::DArr<int32_t> cmap;                 // logical constraint block -> instance
  pcc::DArr<int32_t> bbounds;              // B+1 detection-batch offsets, then the candidate count
  pcc::DArr<unsigned long long> maxbits;   // B: max overlap bits
  pcc::ConstraintSet S2;                   // merge target (B > 1)
  // appends of the current step, in order (colony, its phase, first set
  // index, count); set_unmerged: the set is not yet in colony-major order
  struct AppendRec {
    int32_t colony, phase;
    int64_t start, count;
  };
  std::vector<AppendRec> append_log;
  bool set_unmerged = false;
  std::vector<int64_t> merge_offs_host;
  pcc::DArr<int64_t> merge_offs;
  pcc::DArr<int64_t> regroup_dst;
  int64_t nbb_total = 0, ncb_total = 0;
};

namespace pcc {

// Device buffers of one detection batch as raw pointers (the context's
// arrays, or a caller-independent static scratch).  Sizes: N bodies, S
// neighbour cells per body (9 when every body has z = 0, else 27), hash table
// of T (a power of two >= 2N) buckets, m records.
// Candidate region of one filter workgroup (256 (body, neighbour cell)
// slots), in candidate pairs: dense 2D packings fill ~100.
constexpr int32_t kCandBlock = 512;

struct DetectBufs {
  int64_t N = 0;
  BodyArrays body{};                 // body state (position, orientation, length, diameter read)
  const int32_t* binst = nullptr;    // colony of every body
  const InstDev* idev = nullptr;     // colony records (grid, cutoff, warm start, flags)
  int64_t B = 1;
  int S = 27;
  int64_t T = 1024;
  int4* cell4 = nullptr;             // N
  double2* geo = nullptr;            // 4N
  int4* scell = nullptr;             // N: cell4 in bucket order
  double2* sgeo = nullptr;           // 4N: geo in bucket order
  float4* frec = nullptr;            // N: FP32 filter records in bucket order
  uint32_t* bucket = nullptr;        // N
  int32_t *perm = nullptr, *bstart = nullptr, *bcursor = nullptr;  // N, T+1, T+1
  int32_t *dslot = nullptr, *dslot_off = nullptr;                  // N*S+1 each
  int32_t* cstart = nullptr;         // N*S+1: first candidate of each slot's run
  int2* cand = nullptr;              // filter_grid * cand_block: (slot, bucket position of J)
  uint8_t* accept = nullptr;         // filter_grid * cand_block
  int32_t* bcount = nullptr;         // filter_grid: candidates in each workgroup's region
  int32_t cand_block = kCandBlock;   // candidate region per workgroup
  int32_t *dcount = nullptr, *bbounds = nullptr;                   // N+1, B+2
  int32_t *rI = nullptr, *rJ = nullptr;                            // m
  double* rec = nullptr;             // 16 m
  int32_t *ridx = nullptr, *ridx2 = nullptr;                       // m
  uint64_t *rkey = nullptr, *rkey2 = nullptr;                      // m
  int32_t *xs_cnt = nullptr, *xs_off = nullptr;                    // m+1 each (x order)
  unsigned long long *xs_kmin = nullptr, *xs_kmax = nullptr;       // B each
  bool cursor_zero = false;          // bcursor[0..T] known to be zero (no fill)
  void* tmp = nullptr;               // rocPRIM temporary storage
  size_t tmp_bytes = 0;
  hipStream_t s = nullptr;
};
int64_t table_size_for(int64_t n);
size_t detect_scan_bytes(int64_t N, int S, int64_t T);
size_t detect_sort_bytes(int64_t m, int64_t B);
bool detect_count(DetectBufs& d, int32_t* bounds_host);
int64_t filter_grid(int64_t N, int S);  // workgroups of the candidate filter
void detect_write(DetectBufs& d, int64_t m);
void detect_records_out(const DetectBufs& d, int64_t m, const pcc_constraints& out_dev);

// Order of n items by (group, x, tie) (xsort.hip).  The items of group g are
// the contiguous range [gstart[g], gstart[g + 1]) of the item order (groups
// ascend with the item index) and keep that range of output positions.
struct XSort {
  int64_t n = 0, G = 0;
  const int32_t* gstart = nullptr;   // G + 1 entries, on the device
  const uint64_t* key = nullptr;     // n: order_key of each item's x
  const uint64_t* tie = nullptr;     // n: tie-break, or null for the item index
  int32_t* grp = nullptr;            // n: each item's group (overwritten)
  int32_t *cnt = nullptr, *off = nullptr;            // n + 1 each
  unsigned long long *kmin = nullptr, *kmax = nullptr;  // G each
  bool bounds_clear = false;         // kmin / kmax already hold ~0 / 0 (else cleared here)
  int32_t* perm = nullptr;           // n: the items in order (out)
  void* tmp = nullptr;               // rocPRIM scan storage
  size_t tmp_bytes = 0;
  hipStream_t s = nullptr;
};
size_t xsort_tmp_bytes(int64_t n);
void xsort_run(const XSort& a);

// launchers (kernels in detect.hip / solve.hip / soft.hip)
void detect_batch(pcc_ctx* ctx);  // active colonies at their cutoffs; fills ctx->ridx sorted
void append_batch(pcc_ctx* ctx, int32_t iteration);               // appends sorted batch to S
void merge_set(pcc_ctx* ctx);  // the step's set into colony-major order (end of a step)
void build_incidence(pcc_ctx* ctx, bool records);
bool bodies_planar(pcc_ctx* ctx);
bool state_planar(pcc_ctx* ctx);  // bodies_planar, or the value known from the last step
void bbpgd_solve(pcc_ctx* ctx);
void integrate_hard(pcc_ctx* ctx);
void grow(pcc_ctx* ctx);
void brownian_inputs(pcc_ctx* ctx);
bool brownian_offplane(const pcc_ctx* ctx);
void soft_forces_and_integrate(pcc_ctx* ctx);
void reset_step(pcc_ctx* ctx);
void divide_colonies(pcc_ctx* ctx, const uint64_t* seeds, int32_t step, double length_noise,
                     double angle_noise);
void colony_median_speeds(pcc_ctx* ctx, std::vector<double>& median);
double adapted_dt(double dt, double median, double cfl_factor, double tolerance, double alpha);
void upload_instances(pcc_ctx* ctx);
int32_t* readback_pin(pcc_ctx* ctx, int64_t n);  // pinned int32[n]; no copy may be pending
void max_overlap_inst(pcc_ctx* ctx, std::vector<double>& out);
void max_overlap_launch(pcc_ctx* ctx);  // reduction and copy to pinned memory, no wait
void max_overlap_read(pcc_ctx* ctx, std::vector<double>& out);  // after a stream sync
void merge_batch(pcc_ctx* ctx);

BodyArrays body_arrays(double* base, int64_t N);

// timing helpers
inline void add_timing_bytes(pcc_ctx* ctx, int kernel, double bytes) {
  if (ctx->timing.enabled) ctx->timing.bytes[kernel] += bytes;
}

struct KTimer {
  pcc_ctx* ctx;
  int k;
  hipEvent_t a = nullptr, b = nullptr;
  KTimer(pcc_ctx* c, int kernel, double bytes);
  ~KTimer();
};

}  // namespace pcc
