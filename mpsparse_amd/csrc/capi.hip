// C ABI of the library: handle management, host staging, solver drivers.
// Everything below the ABI runs as HIP kernels on the handle's GPU; there is no
// CPU compute path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mpsparse_amd.h"
#include "bandqr.hpp"
#include "batched.hpp"
#include "bcsr.hpp"
#include "common.hpp"
#include "dia.hpp"
#include "fsai.hpp"
#include "krylov.hpp"
#include "mlp.hpp"
#include "refine.hpp"
#include "sort.hpp"
#include "spmv.hpp"

using namespace msp;

// ---------------------------------------------------------------------------
// errors
static thread_local std::string g_err;

static msp_status fail(msp_status s, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return s;
}

#define HIPCHK(expr)                                                                                  \
  do {                                                                                                \
    hipError_t e_ = (expr);                                                                           \
    if (e_ != hipSuccess)                                                                             \
      return fail(MSP_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

#define MSPCHK(expr)                    \
  do {                                  \
    msp_status s_ = (expr);             \
    if (s_ != MSP_OK) return s_;        \
  } while (0)

// Waits for a stream's queued work (or an event) with the calling thread
// polling the runtime, for the solvers' short waits between launches (the
// launches queued after such a wait start on an idle GPU, so the host should
// resume as soon as the work is done).
static hipError_t spin_sync(hipStream_t s) {
  for (;;) {
    const hipError_t e = hipStreamQuery(s);
    if (e != hipErrorNotReady) return e;
  }
}
static hipError_t spin_event(hipEvent_t ev) {
  for (;;) {
    const hipError_t e = hipEventQuery(ev);
    if (e != hipErrorNotReady) return e;
  }
}

// ---------------------------------------------------------------------------
// device facts
static int cu_count() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (!cached[dev]) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
    cached[dev] = v;
  }
  return cached[dev];
}

static uint32_t spmv_grid(uint32_t nblk) {
  uint32_t cap = (uint32_t)cu_count() * 8u;
  return std::max(1u, std::min(nblk, cap));
}

static uint32_t vec_grid(uint64_t n) {
  uint64_t chunks = (n + 3) / 4;
  uint64_t g = (chunks + kBlock - 1) / kBlock;
  uint64_t cap = (uint64_t)cu_count() * 4u;
  return (uint32_t)std::max<uint64_t>(1, std::min(g, cap));
}

static uint32_t grid_for(uint64_t work) {
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((work + kBlock - 1) / kBlock, 65536));
}

static uint32_t bitlen(uint32_t v) { return v ? 32u - (uint32_t)__builtin_clz(v) : 0u; }

// ---------------------------------------------------------------------------
// memory helpers
// Freed device blocks of the long-lived buffers (matrices, preconditioners,
// solver workspaces) are kept per device for the next allocation of about the
// same size: a rebuilt preconditioner or matrix asks for the same sizes again,
// and hipMalloc / hipFree cost 60-180 us per call (1.3 + 1.4 ms of a 2.3 ms
// learned-FSAI set-up). A block enters the cache only after the device has
// finished all its work (the wait hipFree itself performs), so no kernel still
// uses it when it is handed out again. At most kBlockCacheCap bytes per device
// are held (beyond that, frees go back to the system); a failed hipMalloc
// empties the cache and retries; msp_release_workspace returns everything.
// Process teardown. Destructors of handles may run inside exit() (static
// objects of a host program, handles still owned by a scripting runtime that
// is finalised late). The HIP runtime tears itself down from its own exit
// handlers, and a HIP call after that point can crash. The library registers
// an exit handler right after its first successful HIP call, i.e. after the
// runtime has registered its own, so it runs first (exit handlers run in
// reverse order of registration): from then on destructors make no HIP call
// and leave device memory, pinned memory, events, streams and graphs to the
// end of the process. Library state with destructors is allocated once and
// never destroyed (no static destructor of this library calls HIP).
static std::atomic<bool> g_exiting{false};
static void on_process_exit() { g_exiting.store(true, std::memory_order_relaxed); }
static void register_exit_guard() {
  static std::once_flag once;
  std::call_once(once, [] { std::atexit(on_process_exit); });
}
static bool in_teardown() { return g_exiting.load(std::memory_order_relaxed); }

constexpr size_t kBlockCacheCap = (size_t)8 << 30;
constexpr int kMaxDevices = 64;
struct BlockCache {
  std::mutex mu;
  std::multimap<size_t, void*> free[kMaxDevices];  // capacity -> block
  size_t bytes[kMaxDevices] = {};
};
// never destroyed: buffers released during process teardown still find it
static BlockCache& g_blocks = *new BlockCache;

// Capacity of a block serving `bytes`: 64 KiB steps up to 16 MiB, 2 MiB above.
static size_t block_capacity(size_t bytes) {
  const size_t g = bytes <= ((size_t)16 << 20) ? ((size_t)64 << 10) : ((size_t)2 << 20);
  return (bytes + g - 1) / g * g;
}

static void cache_flush(int dev) {
  std::lock_guard<std::mutex> lk(g_blocks.mu);
  for (auto& kv : g_blocks.free[dev]) (void)hipFree(kv.second);
  g_blocks.free[dev].clear();
  g_blocks.bytes[dev] = 0;
}

// No-allocation calls: while set (for the calling thread), every temporary
// comes from this fixed arena (a scratch array) and a buffer allocation fails
// instead of calling hipMalloc.
struct Arena;
static thread_local Arena* t_fixed_arena = nullptr;

struct DevMem {
  void* p = nullptr;
  size_t cap = 0;
  int dev = -1;
  DevMem() = default;
  DevMem(const DevMem&) = delete;
  DevMem& operator=(const DevMem&) = delete;
  ~DevMem() { reset(); }
  msp_status alloc(size_t bytes) {
    reset();
    if (t_fixed_arena) return fail(MSP_ERR_UNSUPPORTED, "a buffer allocation was requested inside a no-allocation call");
    if (bytes == 0) bytes = 256;
    int d = 0;
    HIPCHK(hipGetDevice(&d));
    register_exit_guard();
    const size_t want = block_capacity(bytes);
    if (d >= 0 && d < kMaxDevices) {
      std::lock_guard<std::mutex> lk(g_blocks.mu);
      auto& fl = g_blocks.free[d];
      auto it = fl.lower_bound(want);
      if (it != fl.end() && it->first <= want + want / 4) {
        p = it->second;
        cap = it->first;
        dev = d;
        g_blocks.bytes[d] -= cap;
        fl.erase(it);
        return MSP_OK;
      }
    }
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess && d >= 0 && d < kMaxDevices) {
      (void)hipGetLastError();
      (void)hipDeviceSynchronize();
      cache_flush(d);
      e = hipMalloc(&p, want);
    }
    if (e != hipSuccess) {
      p = nullptr;
      return fail(MSP_ERR_HIP, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    }
    cap = want;
    dev = d;
    return MSP_OK;
  }
  void reset() {
    if (!p) return;
    if (in_teardown()) {  // the process is ending: no HIP call
      p = nullptr;
      cap = 0;
      dev = -1;
      return;
    }
    int cur = -1;
    bool cached = false;
    if (dev >= 0 && dev < kMaxDevices && hipGetDevice(&cur) == hipSuccess && cur == dev &&
        hipDeviceSynchronize() == hipSuccess) {
      std::lock_guard<std::mutex> lk(g_blocks.mu);
      if (g_blocks.bytes[dev] + cap <= kBlockCacheCap) {
        g_blocks.free[dev].emplace(cap, p);
        g_blocks.bytes[dev] += cap;
        cached = true;
      }
    }
    if (!cached) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    dev = -1;
  }
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
};

// Temporaries (sort buffers, probe counters, ...) come from a scratch arena
// per (device, stream) that the library keeps between calls: a stack of
// blocks carved from chunks that are never returned on their own, so repeated
// builds pay neither hipMalloc nor hipFree (whose implicit synchronisation
// stalled the host for ~1 ms per large buffer). Reuse is safe because every
// user of a block is enqueued on the arena's stream: work enqueued after a
// block is released runs after the work that used it. Blocks are released in
// reverse order of allocation (a block released early is popped once those
// above it are). msp_release_workspace() frees the chunks of idle arenas.
struct ArenaChunk {
  void* p;
  size_t size;
};
struct ArenaRec {
  size_t chunk, offset, end;
  bool freed;
};
struct Arena {
  std::mutex mu;
  std::vector<ArenaChunk> chunks;
  std::vector<ArenaRec> stack;
  bool fixed = false;  // one caller-provided chunk, never grown (no-allocation calls)
};

// never destroyed, like g_blocks
static std::mutex& g_arena_mu = *new std::mutex;
static std::map<std::pair<int, hipStream_t>, Arena*>& g_arenas = *new std::map<std::pair<int, hipStream_t>, Arena*>;

static msp_status arena_for(hipStream_t s, Arena** out) {
  if (t_fixed_arena) {
    *out = t_fixed_arena;
    return MSP_OK;
  }
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  register_exit_guard();
  std::lock_guard<std::mutex> lk(g_arena_mu);
  Arena*& a = g_arenas[std::make_pair(dev, s)];
  if (!a) a = new Arena;
  *out = a;
  return MSP_OK;
}

static msp_status arena_alloc(Arena* a, size_t bytes, void** p, size_t* rec) {
  bytes = (bytes + 255) & ~(size_t)255;
  std::lock_guard<std::mutex> lk(a->mu);
  size_t c = 0, off = 0;
  if (!a->stack.empty()) {
    c = a->stack.back().chunk;
    off = a->stack.back().end;
  }
  while (c < a->chunks.size() && off + bytes > a->chunks[c].size) {
    ++c;
    off = 0;
  }
  if (c == a->chunks.size()) {
    if (a->fixed)
      return fail(MSP_ERR_UNSUPPORTED, "the fixed scratch of a no-allocation call is exhausted (%zu more bytes)",
                  bytes);
    size_t sz = std::max<size_t>(bytes, (size_t)64 << 20);
    if (!a->chunks.empty()) sz = std::max(sz, a->chunks.back().size);
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, sz);
    if (e != hipSuccess) return fail(MSP_ERR_HIP, "hipMalloc(%zu bytes) for scratch: %s", sz, hipGetErrorString(e));
    a->chunks.push_back(ArenaChunk{q, sz});
  }
  *p = static_cast<char*>(a->chunks[c].p) + off;
  a->stack.push_back(ArenaRec{c, off, off + bytes, false});
  *rec = a->stack.size() - 1;
  return MSP_OK;
}

static void arena_free(Arena* a, size_t rec) {
  std::lock_guard<std::mutex> lk(a->mu);
  if (rec < a->stack.size()) a->stack[rec].freed = true;
  while (!a->stack.empty() && a->stack.back().freed) a->stack.pop_back();
}

struct TmpMem {
  void* p = nullptr;
  Arena* a = nullptr;
  size_t rec = 0;
  TmpMem() = default;
  TmpMem(const TmpMem&) = delete;
  TmpMem& operator=(const TmpMem&) = delete;
  ~TmpMem() { reset(); }
  // a block of the arena of stream st; every user must be enqueued on st
  msp_status alloc(size_t bytes, hipStream_t st) {
    reset();
    if (bytes == 0) bytes = 256;
    MSPCHK(arena_for(st, &a));
    return arena_alloc(a, bytes, &p, &rec);
  }
  void reset() {
    if (p) arena_free(a, rec);
    p = nullptr;
  }
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
};

static bool is_device_ptr(const void* p) {
  if (!p) return false;
  hipPointerAttribute_t a;
  hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.isManaged;
}

// A read-only array on the device: the caller's pointer, or a staged copy.
struct InArr {
  const void* d = nullptr;
  DevMem own;
  msp_status set(const void* p, size_t bytes, hipStream_t s) {
    if (!p && bytes) return fail(MSP_ERR_INVALID, "null input array");
    if (bytes == 0 || is_device_ptr(p)) {
      d = p;
      return MSP_OK;
    }
    MSPCHK(own.alloc(bytes));
    HIPCHK(hipMemcpyAsync(own.p, p, bytes, hipMemcpyHostToDevice, s));
    d = own.p;
    return MSP_OK;
  }
};

// Copy `bytes` from device memory `src` to `dst` (host or device).
static msp_status copy_out(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (!dst || bytes == 0) return MSP_OK;
  if (is_device_ptr(dst)) {
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
  } else {
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  return MSP_OK;
}

static msp_status copy_in(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes == 0 || dst == src) return MSP_OK;
  if (!src) return fail(MSP_ERR_INVALID, "null input array");
  HIPCHK(hipMemcpyAsync(dst, src, bytes, is_device_ptr(src) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
  return MSP_OK;
}

static msp_status launch_check(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(MSP_ERR_HIP, "kernel launch (%s): %s", what, hipGetErrorString(e));
  return MSP_OK;
}

// ---------------------------------------------------------------------------
// handle
struct Workspace {
  size_t npad = 0;
  DevMem vecs;            // kNumVec vectors of npad floats
  DevMem scal;            // Scalars
  DevMem red;             // grid-reduction partials
  Scalars* hS = nullptr;  // pinned host mirror
  Scalars* hsnap = nullptr;  // pinned: the state after each of the last two batches
  // pinned: the refinement's residual sums (a copy into pageable memory is
  // staged through a runtime buffer)
  double* hsum = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipEvent_t t0 = nullptr, t1 = nullptr;
  static constexpr int kNumVec = 12;
  // X2 is reserved for the refinement's correction (no inner solver touches
  // it); the refinement writes each correction's right-hand side into B
  enum { X = 0, R, P, Q, R0, S, T, Z, W, B, X2, XL };
  float* v(int i) const { return vecs.as<float>() + (size_t)i * npad; }
  Scalars* S_dev() const { return scal.as<Scalars>(); }
  // staging copy of the state for decisions taken inside a step kernel
  Scalars* S_stage() const { return scal.as<Scalars>() + 1; }
  // two halves: SpMV partials, and those of the kernel that consumes them
  GridRed gr() const { return GridRed{red.as<CSum>()}; }
  GridRed gr2() const { return GridRed{red.as<CSum>() + (size_t)16 * cu_count()}; }
  ~Workspace() {
    if (in_teardown()) return;
    if (hS) (void)hipHostFree(hS);
    if (hsnap) (void)hipHostFree(hsnap);
    if (hsum) (void)hipHostFree(hsum);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
  }
};

// State shared by every matrix kind: device, shape, solver workspace, Jacobi.
struct MatBase {
  int device = 0;
  uint32_t n = 0, m = 0;
  // CG step form per preconditioner (none, Jacobi): -1 not measured yet,
  // 1 direction formed in the gather, 0 separate update
  int cg_fused[2] = {-1, -1};
  Workspace ws;
  DevMem dinv;
  int dinv_state = 0;  // 0 not built, 1 ok, 2 singular
  DevMem w64;          // refinement: x64, r64, b64 (npad doubles each) + 2 sums
};

struct msp_csr_s : MatBase {
  uint64_t nnz = 0;
  DevMem row_ptr, col, val, val16, dlo, blk;  // val16 + dlo: 16-bit values, FP32 diagonal remainder
  int val16_type = -1;
  uint32_t nblk = 0, T = 0, max_row = 0;
  // diagonal storage (stencil-like matrices): state 0 not probed, 1 in use, 2 not applicable
  int format = MSP_FORMAT_AUTO;
  int dia_state = 0, dia_dd = -1, dia16_type = -1;
  DiaOffsets dia_o{};
  // symmetric values: the lower diagonals may be read from the upper planes
  // (dia_sym_o); used only when the planes exceed kDiaMirrorMinBytes
  bool dia_sym = false;
  DiaOffsets dia_sym_o{};
  uint64_t dia_ld = 0;
  DevMem dia, dia16, dia_dlo;
  // a transpose kept as a view of another matrix's diagonal storage: its
  // diagonal +k is that matrix's -k plane read k rows on (dia_alias_o holds
  // the plane / shift of each diagonal); no value planes or CSR arrays of its
  // own
  const msp_csr_s* dia_alias = nullptr;
  DiaOffsets dia_alias_o{};
};

struct msp_bcsr_s : MatBase {
  uint32_t B = 0, nbr = 0, nbc = 0;
  uint64_t nblk = 0;
  DevMem brow_ptr, bcol, blocks;
};

static msp_status check_handle(const MatBase* A) {
  if (!A) return fail(MSP_ERR_INVALID, "null matrix handle");
  HIPCHK(hipSetDevice(A->device));
  return MSP_OK;
}

static msp_status ensure_workspace(MatBase* A, hipStream_t s) {
  Workspace& w = A->ws;
  size_t need = ((size_t)std::max(A->n, A->m) + 255) / 256 * 256 + 256;
  if (w.npad >= need && w.scal.p) return MSP_OK;
  w.npad = need;
  MSPCHK(w.vecs.alloc(sizeof(float) * need * Workspace::kNumVec));
  HIPCHK(hipMemsetAsync(w.vecs.p, 0, sizeof(float) * need * Workspace::kNumVec, s));
  if (!w.scal.p) {
    MSPCHK(w.scal.alloc(2 * sizeof(Scalars)));
    HIPCHK(hipMemsetAsync(w.scal.p, 0, 2 * sizeof(Scalars), s));
    // partials of up to 4 values for grids of up to 8 workgroups per CU
    size_t red_bytes = sizeof(CSum) * 6 * (size_t)cu_count() * 8;  // two regions of 16 CSums per CU in use (gr, gr2), a third spare
    MSPCHK(w.red.alloc(red_bytes));
    HIPCHK(hipMemsetAsync(w.red.p, 0, red_bytes, s));
    HIPCHK(hipHostMalloc((void**)&w.hS, sizeof(Scalars), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void**)&w.hsnap, 2 * sizeof(Scalars), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void**)&w.hsum, 2 * sizeof(double), hipHostMallocDefault));
    for (auto& e : w.ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIPCHK(hipEventCreate(&w.t0));
    HIPCHK(hipEventCreate(&w.t1));
  }
  return MSP_OK;
}

// Row blocks for the SpMV (and the max row length).
// Diagonal storage when every entry lies on one of at most kMaxDiag
// diagonals that are at least 60% full: candidate offsets come from a few
// interior sample rows, then one pass over all entries confirms that none lies
// elsewhere. Probed once per matrix (its pattern never changes).
// Symmetric storage (see DiaOffsets): when every lower diagonal has an upper
// partner and the filled values agree bit for bit, the lower diagonals are
// read from their partners' planes.
static msp_status dia_try_mirror(msp_csr A, hipStream_t s) {
  A->dia_sym = false;
  DiaOffsets o = A->dia_o;
  if (A->n != A->m) return MSP_OK;
  int32_t partner[kMaxDiag];
  bool any = false;
  for (int d = 0; d < o.ndiag; ++d) {
    partner[d] = d;
    if (o.off[d] >= 0) continue;
    int p = -1;
    for (int q = 0; q < o.ndiag; ++q)
      if (o.off[q] == -o.off[d]) p = q;
    if (p <= d) return MSP_OK;  // (offsets ascend: the +k plane follows the -k plane)
    partner[d] = p;
    any = true;
  }
  if (!any) return MSP_OK;
  TmpMem pt, miss;
  MSPCHK(pt.alloc(sizeof(partner), s));
  MSPCHK(miss.alloc(sizeof(uint32_t), s));
  HIPCHK(hipMemcpyAsync(pt.p, partner, sizeof(partner), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemsetAsync(miss.p, 0, sizeof(uint32_t), s));
  k_dia_symcheck<<<grid_for(A->n), kBlock, 0, s>>>(A->dia.as<float>(), A->n, o, pt.as<int32_t>(), A->dia_ld,
                                                   miss.as<uint32_t>());
  MSPCHK(launch_check("k_dia_symcheck"));
  uint32_t hm = 0;
  HIPCHK(hipMemcpyAsync(&hm, miss.p, sizeof hm, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (hm) return MSP_OK;
  for (int d = 0; d < o.ndiag; ++d) {
    if (o.off[d] >= 0) continue;
    o.plane[d] = partner[d];
    o.shift[d] = o.off[d];
  }
  A->dia_sym_o = o;
  A->dia_sym = true;
  return MSP_OK;
}

// fill = false: the values are set later (refresh_dia fills the storage then),
// so only the pattern probe and the allocation happen here.
static msp_status ensure_dia(msp_csr A, hipStream_t s, bool fill = true) {
  if (A->dia_state != 0) return MSP_OK;
  A->dia_state = 2;
  // the diagonal kernel indexes rows, columns and byte offsets in 32 bits
  if (A->format != MSP_FORMAT_AUTO || A->n == 0 || A->m == 0 || A->nnz == 0 || A->max_row > (uint32_t)kMaxDiag ||
      A->n >= kDiaMaxDim || A->m >= kDiaMaxDim)
    return MSP_OK;
  const uint32_t n = A->n;
  std::vector<int64_t> offs;
  // kDiaSamples sample rows from a fixed-seed LCG (rows on a grid face lack
  // some offsets; a sample that misses one only sends the matrix back to CSR),
  // gathered by one kernel and read back with one copy
  DiaSample smp;
  uint64_t st = 0x9E3779B97F4A7C15ull;
  for (int k = 0; k < kDiaSamples; ++k) {
    st = st * 6364136223846793005ull + 1442695040888963407ull;
    smp.row[k] = (uint32_t)((st >> 33) % n);
  }
  constexpr size_t kSampleWords = (size_t)kDiaSamples * (kMaxDiag + 1);
  TmpMem sbuf;
  MSPCHK(sbuf.alloc(sizeof(uint32_t) * kSampleWords, s));
  k_dia_sample<<<1, kDiaSamples, 0, s>>>(A->row_ptr.as<uint32_t>(), A->col.as<uint32_t>(), smp,
                                         sbuf.as<uint32_t>());
  MSPCHK(launch_check("k_dia_sample"));
  std::vector<uint32_t> hs(kSampleWords);
  HIPCHK(hipMemcpyAsync(hs.data(), sbuf.p, sizeof(uint32_t) * kSampleWords, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (int k = 0; k < kDiaSamples; ++k) {
    const uint32_t* rec = hs.data() + (size_t)k * (kMaxDiag + 1);
    const uint32_t len = std::min<uint32_t>(rec[0], (uint32_t)kMaxDiag);
    for (uint32_t e = 0; e < len; ++e) {
      const int64_t o = (int64_t)rec[1 + e] - (int64_t)smp.row[k];
      if (std::find(offs.begin(), offs.end(), o) == offs.end()) offs.push_back(o);
    }
    if (offs.size() > (size_t)kMaxDiag) return MSP_OK;
  }
  if (offs.empty()) return MSP_OK;
  std::sort(offs.begin(), offs.end());
  for (int64_t o : offs)
    if (o < INT32_MIN || o > INT32_MAX) return MSP_OK;
  if ((double)A->nnz < 0.6 * (double)n * (double)offs.size()) return MSP_OK;
  DiaOffsets o{};
  dia_plain(o);
  o.ndiag = (int32_t)offs.size();
  int dd = -1;
  for (size_t d = 0; d < offs.size(); ++d) {
    o.off[d] = (int32_t)offs[d];
    if (offs[d] == 0) dd = (int)d;
  }
  TmpMem miss;
  MSPCHK(miss.alloc(sizeof(uint32_t), s));
  HIPCHK(hipMemsetAsync(miss.p, 0, sizeof(uint32_t), s));
  k_dia_check<<<grid_for(n), kBlock, 0, s>>>(A->row_ptr.as<uint32_t>(), A->col.as<uint32_t>(), n, o,
                                             miss.as<uint32_t>());
  MSPCHK(launch_check("k_dia_check"));
  uint32_t hm = 0;
  HIPCHK(hipMemcpyAsync(&hm, miss.p, sizeof hm, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (hm) return MSP_OK;
  A->dia_ld = ((uint64_t)n + 63) / 64 * 64;
  const size_t bytes = sizeof(float) * A->dia_ld * o.ndiag;
  MSPCHK(A->dia.alloc(bytes));
  if (fill) {
    HIPCHK(hipMemsetAsync(A->dia.p, 0, bytes, s));
    k_dia_fill<<<grid_for(n), kBlock, 0, s>>>(A->row_ptr.as<uint32_t>(), A->col.as<uint32_t>(), A->val.as<float>(),
                                              n, o, A->dia_ld, A->dia.as<float>());
    MSPCHK(launch_check("k_dia_fill"));
  }
  A->dia_o = o;
  A->dia_dd = dd;
  A->dia_state = 1;
  A->dia16_type = -1;
  if (fill) MSPCHK(dia_try_mirror(A, s));
  return MSP_OK;
}

// T = G^T as a view of G's diagonal storage: T's diagonal d (offset +k,
// ascending) reads G's plane of offset -k, k rows on (G(i + k, i) = T(i, i + k)).
// G's main diagonal is its last plane, so every such read stays inside G's
// planes (row i + k < n + k <= the next plane's end). T keeps no values or CSR
// arrays of its own.
static void dia_alias_transpose(msp_csr T, const msp_csr_s* g) {
  DiaOffsets o{}, a{};
  dia_plain(o);
  const int nd = g->dia_o.ndiag;
  o.ndiag = a.ndiag = nd;
  int dd = -1;
  for (int d = 0; d < nd; ++d) {
    o.off[d] = -g->dia_o.off[nd - 1 - d];  // ascending
    a.off[d] = o.off[d];
    a.plane[d] = nd - 1 - d;
    a.shift[d] = o.off[d];
    if (o.off[d] == 0) dd = d;
  }
  T->dia_o = o;
  T->dia_alias_o = a;
  T->dia_dd = dd;
  T->dia_ld = g->dia_ld;
  T->dia_alias = g;
  T->dia16_type = -1;
  T->dia_state = 1;
}

// Refill the FP32 diagonal storage after A's values changed (pattern fixed).
static msp_status refresh_dia(msp_csr A, hipStream_t s) {
  if (A->dia_state != 1 || A->dia_alias) return MSP_OK;
  A->dia_sym = false;  // new values: symmetry no longer known
  HIPCHK(hipMemsetAsync(A->dia.p, 0, sizeof(float) * A->dia_ld * A->dia_o.ndiag, s));
  k_dia_fill<<<grid_for(A->n), kBlock, 0, s>>>(A->row_ptr.as<uint32_t>(), A->col.as<uint32_t>(), A->val.as<float>(),
                                               A->n, A->dia_o, A->dia_ld, A->dia.as<float>());
  A->dia16_type = -1;
  return launch_check("k_dia_fill");
}

static msp_status ensure_dia16(msp_csr A, int32_t vt, hipStream_t s) {
  if (A->dia_alias) return ensure_dia16(const_cast<msp_csr>(A->dia_alias), vt, s);
  if (A->dia_state != 1 || vt == MSP_VALUES_F32 || A->dia16_type == vt) return MSP_OK;
  const size_t cnt = A->dia_ld * A->dia_o.ndiag;
  MSPCHK(A->dia16.alloc(sizeof(uint16_t) * cnt));
  HIPCHK(hipMemsetAsync(A->dia16.p, 0, sizeof(uint16_t) * cnt, s));
  MSPCHK(A->dia_dlo.alloc(sizeof(float) * ((size_t)A->n + 1)));
  if (vt == kValuesBF16Plain)
    k_dia_cvt16<true, false><<<grid_for(A->n), kBlock, 0, s>>>(A->dia.as<float>(), A->n, A->m, A->dia_o, A->dia_dd,
                                                               A->dia_ld, A->dia16.as<uint16_t>(),
                                                               A->dia_dlo.as<float>());
  else if (vt == MSP_VALUES_BF16)
    k_dia_cvt16<true><<<grid_for(A->n), kBlock, 0, s>>>(A->dia.as<float>(), A->n, A->m, A->dia_o, A->dia_dd,
                                                        A->dia_ld, A->dia16.as<uint16_t>(), A->dia_dlo.as<float>());
  else
    k_dia_cvt16<false><<<grid_for(A->n), kBlock, 0, s>>>(A->dia.as<float>(), A->n, A->m, A->dia_o, A->dia_dd,
                                                         A->dia_ld, A->dia16.as<uint16_t>(), A->dia_dlo.as<float>());
  MSPCHK(launch_check("k_dia_cvt16"));
  A->dia16_type = vt;
  return MSP_OK;
}

// Everything an SpMV with values of type vt needs: the diagonal-storage probe,
// and the 16-bit copies (CSR and, when in use, diagonal).
static msp_status ensure_format(msp_csr A, int32_t vt, hipStream_t s) {
  MSPCHK(ensure_dia(A, s));
  if (vt != MSP_VALUES_F32) MSPCHK(ensure_dia16(A, vt, s));
  return MSP_OK;
}

// max_row_bound >= 0: a known upper bound on the row length stands in for the
// measured maximum (no device round trip; a pattern derived from another).
static msp_status finalize_matrix(msp_csr A, hipStream_t s, int64_t max_row_bound = -1) {
  const uint32_t n = A->n;
  if (max_row_bound >= 0) {
    A->max_row = (uint32_t)std::min<int64_t>(max_row_bound, UINT32_MAX);
  } else {
    TmpMem mx;
    MSPCHK(mx.alloc(sizeof(uint32_t), s));
    HIPCHK(hipMemsetAsync(mx.p, 0, sizeof(uint32_t), s));
    if (n > 0) {
      uint32_t g = std::min<uint32_t>((n + kBlock - 1) / kBlock, 1024);
      k_max_row<<<g, kBlock, 0, s>>>(A->row_ptr.as<uint32_t>(), n, mx.as<uint32_t>());
      MSPCHK(launch_check("k_max_row"));
    }
    HIPCHK(hipMemcpyAsync(&A->max_row, mx.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  // nonzeros per row block: as large as the LDS window allows with room for
  // the longest row (measured best near 3584 for 7- and 144-entry rows)
  const uint32_t room = A->max_row + 4 < kCap ? kCap - 4 - A->max_row : 0;
  const uint32_t T = std::min<uint32_t>(3584u, std::max<uint32_t>(512u, room));
  A->T = T;
  A->nblk = (uint32_t)std::max<uint64_t>(1, (A->nnz + T - 1) / T);
  MSPCHK(A->blk.alloc(sizeof(uint32_t) * (A->nblk + 1)));
  uint32_t g = (A->nblk + 1 + kBlock - 1) / kBlock;
  k_row_blocks<<<g, kBlock, 0, s>>>(A->row_ptr.as<uint32_t>(), n, T, A->nblk, A->blk.as<uint32_t>());
  MSPCHK(launch_check("k_row_blocks"));
  return MSP_OK;
}

static msp_status alloc_csr(msp_csr A) {
  MSPCHK(A->row_ptr.alloc(sizeof(uint32_t) * ((size_t)A->n + 1)));
  size_t padded = (size_t)A->nnz + 64;
  MSPCHK(A->col.alloc(sizeof(uint32_t) * padded));
  MSPCHK(A->val.alloc(sizeof(float) * padded));
  return MSP_OK;
}

static msp_status zero_pad(msp_csr A, hipStream_t s) {
  HIPCHK(hipMemsetAsync(A->col.as<uint32_t>() + A->nnz, 0, 64 * sizeof(uint32_t), s));
  HIPCHK(hipMemsetAsync(A->val.as<float>() + A->nnz, 0, 64 * sizeof(float), s));
  return MSP_OK;
}

// ---------------------------------------------------------------------------
// COO -> CSR driver
template <typename K>
static msp_status sort_passes(const uint64_t* keys, const float* vals, uint64_t nnz, sortk::KeyMap km, int passes,
                              int shift0, uint32_t* col_out, float* val_out, uint32_t* rows_out, int* err,
                              hipStream_t s) {
  using namespace sortk;
  const uint32_t num_tiles = (uint32_t)((nnz + kTile - 1) / kTile);
  const uint64_t N = (uint64_t)kRadix * num_tiles;
  const uint32_t nchunks = (num_tiles + kColChunk - 1) / kColChunk;
  TmpMem counts, sums, dst, ka, kb, va, vb;
  MSPCHK(counts.alloc(sizeof(uint32_t) * N, s));
  MSPCHK(sums.alloc(sizeof(uint32_t) * kRadix * nchunks, s));
  MSPCHK(dst.alloc(sizeof(uint32_t) * (kRadix + 1), s));
  if (passes > 1) {
    MSPCHK(ka.alloc(sizeof(K) * nnz, s));
    MSPCHK(va.alloc(sizeof(float) * nnz, s));
  }
  if (passes > 2) {
    MSPCHK(kb.alloc(sizeof(K) * nnz, s));
    MSPCHK(vb.alloc(sizeof(float) * nnz, s));
  }
  K* kbuf[2] = {ka.as<K>(), kb.as<K>()};
  float* vbuf[2] = {va.as<float>(), vb.as<float>()};
  const K* kin = nullptr;
  const float* vin = vals;
  for (int p = 0; p < passes; ++p) {
    const int shift = shift0 + p * kRadixBits;
    const bool first = p == 0, last = p == passes - 1;
    if (first) k_upsweep<uint64_t, true><<<num_tiles, kBlock, 0, s>>>(keys, nnz, km, shift, num_tiles, counts.as<uint32_t>(), err);
    else k_upsweep<K, false><<<num_tiles, kBlock, 0, s>>>(kin, nnz, km, shift, num_tiles, counts.as<uint32_t>(), err);
    MSPCHK(launch_check("k_upsweep"));
    k_colscan_reduce<<<nchunks, kRadix, 0, s>>>(counts.as<uint32_t>(), num_tiles, sums.as<uint32_t>());
    k_colscan_top<<<1, kRadix, 0, s>>>(sums.as<uint32_t>(), nchunks, dst.as<uint32_t>(), nnz);
    k_colscan_down<<<nchunks, kRadix, 0, s>>>(counts.as<uint32_t>(), num_tiles, sums.as<uint32_t>(),
                                              dst.as<uint32_t>());
    MSPCHK(launch_check("k_colscan"));
    K* kout = kbuf[p & 1];
    float* vout = vbuf[p & 1];
    const uint32_t* off = counts.as<uint32_t>();
    if (first && last)
      k_downsweep<uint64_t, K, true, true><<<num_tiles, kBlock, 0, s>>>(keys, vin, nullptr, val_out, col_out, rows_out, nnz, km, shift, num_tiles, off);
    else if (first)
      k_downsweep<uint64_t, K, true, false><<<num_tiles, kBlock, 0, s>>>(keys, vin, kout, vout, nullptr, nullptr, nnz, km, shift, num_tiles, off);
    else if (last)
      k_downsweep<K, K, false, true><<<num_tiles, kBlock, 0, s>>>(kin, vin, nullptr, val_out, col_out, rows_out, nnz, km, shift, num_tiles, off);
    else
      k_downsweep<K, K, false, false><<<num_tiles, kBlock, 0, s>>>(kin, vin, kout, vout, nullptr, nullptr, nnz, km, shift, num_tiles, off);
    MSPCHK(launch_check("k_downsweep"));
    kin = kout;
    vin = vout;
  }
  // temporaries go back to the pool on return, ordered after the passes on s
  return MSP_OK;
}

// Two-level sort (sortk::k_segment_sort): phase-1 LSD passes over the segment
// bits of the row only, segment ranges by binary search, then one LDS sort per
// segment that writes col / values / row_ptr (and rows when asked) directly.
// *overflow is raised on the device when a segment exceeds sortk::kSegMax keys
// (the caller then runs sort_passes instead). K: key type of the phase-1 output.
template <typename K>
static msp_status segment_sort(const uint64_t* keys, const float* vals, uint64_t nnz, uint32_t n,
                               const sortk::KeyMap& km, int sb, int passes, uint32_t* row_ptr, uint32_t* col_out,
                               float* val_out, uint32_t* rows_out, int* err, int* overflow, hipStream_t s) {
  using namespace sortk;
  const uint32_t G = (uint32_t)(((uint64_t)n + (1ull << sb) - 1) >> sb);
  const int shift0 = (int)km.cbits + sb;
  TmpMem bounds, counts, sums, dst, ka, kb, va, vb;
  const K* kres = nullptr;
  const float* vres = vals;
  if (passes > 0) {
    const uint32_t num_tiles = (uint32_t)((nnz + kTile - 1) / kTile);
    const uint32_t nchunks = (num_tiles + kColChunk - 1) / kColChunk;
    MSPCHK(counts.alloc(sizeof(uint32_t) * (uint64_t)kRadix * num_tiles, s));
    MSPCHK(sums.alloc(sizeof(uint32_t) * kRadix * nchunks, s));
    MSPCHK(dst.alloc(sizeof(uint32_t) * (kRadix + 1), s));
    MSPCHK(ka.alloc(sizeof(K) * nnz, s));
    MSPCHK(va.alloc(sizeof(float) * nnz, s));
    if (passes > 1) {
      MSPCHK(kb.alloc(sizeof(K) * nnz, s));
      MSPCHK(vb.alloc(sizeof(float) * nnz, s));
    }
    K* kbuf[2] = {ka.as<K>(), kb.as<K>()};
    float* vbuf[2] = {va.as<float>(), vb.as<float>()};
    for (int p = 0; p < passes; ++p) {
      const int shift = shift0 + p * kRadixBits;
      if (p == 0)
        k_upsweep<uint64_t, true><<<num_tiles, kBlock, 0, s>>>(keys, nnz, km, shift, num_tiles, counts.as<uint32_t>(), err);
      else
        k_upsweep<K, false><<<num_tiles, kBlock, 0, s>>>(kres, nnz, km, shift, num_tiles, counts.as<uint32_t>(), err);
      k_colscan_reduce<<<nchunks, kRadix, 0, s>>>(counts.as<uint32_t>(), num_tiles, sums.as<uint32_t>());
      k_colscan_top<<<1, kRadix, 0, s>>>(sums.as<uint32_t>(), nchunks, dst.as<uint32_t>(), nnz);
      k_colscan_down<<<nchunks, kRadix, 0, s>>>(counts.as<uint32_t>(), num_tiles, sums.as<uint32_t>(),
                                                dst.as<uint32_t>());
      if (p == 0)
        k_downsweep<uint64_t, K, true, false><<<num_tiles, kBlock, 0, s>>>(keys, vres, kbuf[p & 1], vbuf[p & 1], nullptr,
                                                                           nullptr, nnz, km, shift, num_tiles,
                                                                           counts.as<uint32_t>());
      else
        k_downsweep<K, K, false, false><<<num_tiles, kBlock, 0, s>>>(kres, vres, kbuf[p & 1], vbuf[p & 1], nullptr,
                                                                     nullptr, nnz, km, shift, num_tiles,
                                                                     counts.as<uint32_t>());
      MSPCHK(launch_check("segment phase 1"));
      kres = kbuf[p & 1];
      vres = vbuf[p & 1];
    }
  }
  MSPCHK(bounds.alloc(sizeof(uint32_t) * ((size_t)G + 1), s));
  k_seg_bounds<K><<<grid_for((uint64_t)G + 1), kBlock, 0, s>>>(kres, nnz, shift0, G, passes == 0,
                                                              bounds.as<uint32_t>());
  const uint32_t grid = std::max(1u, std::min<uint32_t>(G, 1u << 20));
  if (passes == 0)
    k_segment_sort<uint64_t, true><<<grid, kSegThreads, 0, s>>>(keys, vals, bounds.as<uint32_t>(), G, km, sb, n, nnz,
                                                           col_out, val_out, rows_out, row_ptr, overflow, err);
  else
    k_segment_sort<K, false><<<grid, kSegThreads, 0, s>>>(kres, vres, bounds.as<uint32_t>(), G, km, sb, n, nnz, col_out,
                                                     val_out, rows_out, row_ptr, overflow, err);
  return launch_check("segment sort");
}

// The two-level sort with two phase-1 passes over keys of 33..40 bits, the
// keys kept in 32 bits throughout: the first pass writes each key without its
// own digit (recoverable from the position through that pass's digit starts),
// the second ranks on the segment id's high digit (now at the same shift) and
// writes the key with that digit replaced by the first pass's, recovered from
// the input position. The segment bounds come from the second pass's digit
// starts and the low digit; the segment sort reads only the local bits below
// the digit. Against 64-bit keys: 4 bytes less per key in the first pass's
// writes, the second pass's reads and writes, and the segment sort's reads.
static msp_status segment_sort_split(const uint64_t* keys, const float* vals, uint64_t nnz, uint32_t n,
                                     const sortk::KeyMap& km, int sb, uint32_t* row_ptr, uint32_t* col_out,
                                     float* val_out, uint32_t* rows_out, int* err, int* overflow, hipStream_t s) {
  using namespace sortk;
  const uint32_t G = (uint32_t)(((uint64_t)n + (1ull << sb) - 1) >> sb);
  const int shift = (int)km.cbits + sb;
  const uint32_t num_tiles = (uint32_t)((nnz + kTile - 1) / kTile);
  const uint32_t nchunks = (num_tiles + kColChunk - 1) / kColChunk;
  TmpMem bounds, counts, sums, ka, kb, va, vb, st1, st2;
  MSPCHK(counts.alloc(sizeof(uint32_t) * (uint64_t)kRadix * num_tiles, s));
  MSPCHK(sums.alloc(sizeof(uint32_t) * kRadix * nchunks, s));
  MSPCHK(ka.alloc(sizeof(uint32_t) * nnz, s));
  MSPCHK(va.alloc(sizeof(float) * nnz, s));
  MSPCHK(kb.alloc(sizeof(uint32_t) * nnz, s));
  MSPCHK(vb.alloc(sizeof(float) * nnz, s));
  MSPCHK(st1.alloc(sizeof(uint32_t) * (kRadix + 1), s));
  MSPCHK(st2.alloc(sizeof(uint32_t) * (kRadix + 1), s));
  uint32_t* cnt = counts.as<uint32_t>();
  // the scanned offsets, and the pass's digit starts into `st`
  auto colscan = [&](uint32_t* st) {
    k_colscan_reduce<<<nchunks, kRadix, 0, s>>>(cnt, num_tiles, sums.as<uint32_t>());
    k_colscan_top<<<1, kRadix, 0, s>>>(sums.as<uint32_t>(), nchunks, st, nnz);
    k_colscan_down<<<nchunks, kRadix, 0, s>>>(cnt, num_tiles, sums.as<uint32_t>(), st);
  };
  // pass 1: the segment id's low digit (compacted keys, bits [shift, shift + 8))
  k_upsweep<uint64_t, true><<<num_tiles, kBlock, 0, s>>>(keys, nnz, km, shift, num_tiles, cnt, err);
  colscan(st1.as<uint32_t>());
  k_downsweep<uint64_t, uint32_t, true, false, kXfDrop><<<num_tiles, kBlock, 0, s>>>(
      keys, vals, ka.as<uint32_t>(), va.as<float>(), nullptr, nullptr, nnz, km, shift, num_tiles, cnt);
  // pass 2: the high digit, now at the same shift
  k_upsweep<uint32_t, false><<<num_tiles, kBlock, 0, s>>>(ka.as<uint32_t>(), nnz, km, shift, num_tiles, cnt, err);
  colscan(st2.as<uint32_t>());
  k_downsweep<uint32_t, uint32_t, false, false, kXfSwap><<<num_tiles, kBlock, 0, s>>>(
      ka.as<uint32_t>(), va.as<float>(), kb.as<uint32_t>(), vb.as<float>(), nullptr, nullptr, nnz, km, shift,
      num_tiles, cnt, st1.as<uint32_t>());
  MSPCHK(launch_check("segment phase 1 (32-bit keys)"));
  MSPCHK(bounds.alloc(sizeof(uint32_t) * ((size_t)G + 1), s));
  k_seg_bounds_split<<<grid_for((uint64_t)G + 1), kBlock, 0, s>>>(kb.as<uint32_t>(), nnz, shift, G,
                                                                 st2.as<uint32_t>(), bounds.as<uint32_t>());
  const uint32_t grid = std::max(1u, std::min<uint32_t>(G, 1u << 20));
  k_segment_sort<uint32_t, false><<<grid, kSegThreads, 0, s>>>(kb.as<uint32_t>(), vb.as<float>(),
                                                               bounds.as<uint32_t>(), G, km, sb, n, nnz, col_out,
                                                               val_out, rows_out, row_ptr, overflow, err);
  return launch_check("segment sort");
}

// Segment size for the two-level sort: the largest s (rows per segment 2^s)
// whose average segment holds at most 60% of sortk::kSegMax keys; -1 when the
// two-level sort does not apply (local bits plus the position index exceed 64,
// or no phase-1 pass would be saved).
static int segment_bits(uint64_t nnz, uint32_t n, uint32_t rbits, uint32_t cbits) {
  const double avg = (double)nnz / std::max<uint32_t>(1, n);
  int sb = 0;
  while (sb < (int)rbits && sb < 31 && avg * (double)(1ull << (sb + 1)) <= 0.6 * sortk::kSegMax) ++sb;
  if (cbits + (uint32_t)sb + sortk::kSegIdxBits > 64) return -1;
  return sb;
}

// Sort (keys, vals) on the device into CSR arrays. `row_ptr` has n+1 entries,
// `col_out`/`val_out` nnz entries. cols_in_order: the input already lists each
// row's entries in ascending column order (a transposed CSR pattern, say), so
// the stable passes over the row bits alone give the full order (half the
// passes for 2M x 2M).
// Flags start at zero through a kernel (not a memset), so the sort can run
// under graph capture.
__global__ void k_zero_flags(int* __restrict__ f, int count) {
  if ((int)threadIdx.x < count) f[threadIdx.x] = 0;
}

// max_row: when not null and the two-level sort completes, receives the
// longest row (read back with the sort's flags, in the same round trip); it
// is set to -1 otherwise.
static msp_status coo_to_csr_device(const uint64_t* keys, const float* vals, uint64_t nnz, uint32_t n, uint32_t m,
                                    uint32_t* row_ptr, uint32_t* col_out, float* val_out, hipStream_t s,
                                    uint32_t* rows_out = nullptr, bool cols_in_order = false,
                                    int64_t* max_row = nullptr) {
  TmpMem err, rows_own;
  if (max_row) *max_row = -1;
  // [0] key out of range, [1] segment overflow, [2] longest row
  MSPCHK(err.alloc(3 * sizeof(int), s));
  k_zero_flags<<<1, 64, 0, s>>>(err.as<int>(), 3);
  if (nnz > 0) {
    sortk::KeyMap km;
    km.cbits = m > 1 ? bitlen(m - 1) : 0;
    km.rows = n;
    km.cols = m;
    km.cmask = km.cbits ? ((1ull << km.cbits) - 1ull) : 0ull;
    const uint32_t rbits = n > 1 ? bitlen(n - 1) : 0;
    const uint32_t tb = km.cbits + rbits;
    const uint32_t sorted_bits = cols_in_order ? rbits : tb;
    const int shift0 = cols_in_order ? (int)km.cbits : 0;
    const int passes = std::max(1, (int)((sorted_bits + sortk::kRadixBits - 1) / sortk::kRadixBits));
    const int sb = cols_in_order ? -1 : segment_bits(nnz, n, rbits, km.cbits);
    if (sb >= 0) {
      const int passes1 = (int)((rbits - (uint32_t)sb + sortk::kRadixBits - 1) / sortk::kRadixBits);
      int* ovf = err.as<int>() + 1;
      if (tb <= 32)
        MSPCHK(segment_sort<uint32_t>(keys, vals, nnz, n, km, sb, passes1, row_ptr, col_out, val_out, rows_out,
                                      err.as<int>(), ovf, s));
      else if (passes1 == 2 && tb <= 40)
        MSPCHK(segment_sort_split(keys, vals, nnz, n, km, sb, row_ptr, col_out, val_out, rows_out, err.as<int>(),
                                  ovf, s));
      else
        MSPCHK(segment_sort<uint64_t>(keys, vals, nnz, n, km, sb, passes1, row_ptr, col_out, val_out, rows_out,
                                      err.as<int>(), ovf, s));
      if (max_row && n > 0) {
        const uint32_t g = std::min<uint32_t>((n + kBlock - 1) / kBlock, 1024);
        k_max_row<<<g, kBlock, 0, s>>>(row_ptr, n, reinterpret_cast<uint32_t*>(err.as<int>() + 2));
      }
      int hf[3] = {0, 0, 0};
      HIPCHK(hipMemcpyAsync(hf, err.p, 3 * sizeof(int), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      if (hf[0]) return fail(MSP_ERR_KEY_RANGE, "a COO key lies outside [0,%u) x [0,%u)", n, m);
      if (!hf[1]) {
        if (max_row) *max_row = (int64_t)(uint32_t)hf[2];
        return MSP_OK;
      }
      // a segment too large for one workgroup: the single-level sort below
    }
    if (!rows_out) {
      MSPCHK(rows_own.alloc(sizeof(uint32_t) * nnz, s));
      rows_out = rows_own.as<uint32_t>();
    }
    // 32-bit keys while the compacted key and every pass's shift fit them
    if (tb <= 32 && shift0 + sortk::kRadixBits * (passes - 1) < 32)
      MSPCHK(sort_passes<uint32_t>(keys, vals, nnz, km, passes, shift0, col_out, val_out, rows_out, err.as<int>(),
                                   s));
    else
      MSPCHK(sort_passes<uint64_t>(keys, vals, nnz, km, passes, shift0, col_out, val_out, rows_out, err.as<int>(),
                                   s));
  }
  uint64_t work = std::max<uint64_t>(nnz, (uint64_t)n + 1);
  uint32_t g = (uint32_t)std::min<uint64_t>((work + kBlock - 1) / kBlock, 65536);
  sortk::k_row_ptr<<<g, kBlock, 0, s>>>(rows_out, nnz, n, row_ptr, err.as<int>());
  MSPCHK(launch_check("k_row_ptr"));
  int herr = 0;
  HIPCHK(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (herr) return fail(MSP_ERR_KEY_RANGE, "a COO key lies outside [0,%u) x [0,%u)", n, m);
  return MSP_OK;
}

// ---------------------------------------------------------------------------
// ABI: library
extern "C" int32_t msp_abi_version(void) { return MSP_ABI_VERSION; }

extern "C" const char* msp_last_error(void) { return g_err.c_str(); }

extern "C" msp_status msp_release_workspace(void) {
  int cur = 0;
  HIPCHK(hipGetDevice(&cur));
  for (int d = 0; d < kMaxDevices; ++d) {
    bool any = false;
    {
      std::lock_guard<std::mutex> lk(g_blocks.mu);
      any = !g_blocks.free[d].empty();
    }
    if (!any) continue;
    HIPCHK(hipSetDevice(d));
    cache_flush(d);
  }
  HIPCHK(hipSetDevice(cur));
  std::lock_guard<std::mutex> lk(g_arena_mu);
  if (g_arenas.empty()) return MSP_OK;
  for (auto& kv : g_arenas) {
    Arena* a = kv.second;
    std::lock_guard<std::mutex> la(a->mu);
    if (!a->stack.empty()) continue;  // in use by a call in progress
    HIPCHK(hipSetDevice(kv.first.first));
    HIPCHK(hipDeviceSynchronize());
    for (auto& c : a->chunks) HIPCHK(hipFree(c.p));
    a->chunks.clear();
  }
  HIPCHK(hipSetDevice(cur));
  return MSP_OK;
}

extern "C" void msp_default_opts(msp_solve_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof *o);
  o->tol = 1e-6;
  o->maxiter = 10000;
  o->check_every = 16;
  o->criterion = MSP_CRIT_RELRES;
  o->use_x0 = 0;
  o->precond = MSP_PRECOND_NONE;
  o->values = MSP_VALUES_F32;
  o->verify = 1;
}

extern "C" int32_t msp_device_count(void) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return c;
}

static msp_status need_device() {
  if (msp_device_count() <= 0) return fail(MSP_ERR_NO_DEVICE, "no HIP device visible");
  return MSP_OK;
}

extern "C" msp_status msp_coo_to_csr(const uint64_t* keys, const float* vals, uint64_t nnz, uint32_t num_rows,
                                     uint32_t num_cols, uint32_t* row_ptr, uint32_t* col_ind, float* out_vals,
                                     void* stream) {
  MSPCHK(need_device());
  if (!row_ptr || (nnz && (!keys || !vals || !col_ind || !out_vals)))
    return fail(MSP_ERR_INVALID, "msp_coo_to_csr: null array");
  if (nnz > 0xFFFFFFFFull - 64) return fail(MSP_ERR_INVALID, "nnz %llu exceeds the u32 index range", (unsigned long long)nnz);
  hipStream_t s = (hipStream_t)stream;
  InArr k, v;
  MSPCHK(k.set(keys, sizeof(uint64_t) * nnz, s));
  MSPCHK(v.set(vals, sizeof(float) * nnz, s));
  DevMem rp, ci, ov;
  const bool rp_dev = is_device_ptr(row_ptr), ci_dev = nnz == 0 || is_device_ptr(col_ind),
             ov_dev = nnz == 0 || is_device_ptr(out_vals);
  uint32_t* d_rp = row_ptr;
  uint32_t* d_ci = col_ind;
  float* d_ov = out_vals;
  if (!rp_dev) {
    MSPCHK(rp.alloc(sizeof(uint32_t) * ((size_t)num_rows + 1)));
    d_rp = rp.as<uint32_t>();
  }
  if (!ci_dev) {
    MSPCHK(ci.alloc(sizeof(uint32_t) * nnz));
    d_ci = ci.as<uint32_t>();
  }
  if (!ov_dev) {
    MSPCHK(ov.alloc(sizeof(float) * nnz));
    d_ov = ov.as<float>();
  }
  MSPCHK(coo_to_csr_device(static_cast<const uint64_t*>(k.d), static_cast<const float*>(v.d), nnz, num_rows, num_cols,
                           d_rp, d_ci, d_ov, s));
  if (!rp_dev) MSPCHK(copy_out(row_ptr, d_rp, sizeof(uint32_t) * ((size_t)num_rows + 1), s));
  if (!ci_dev) MSPCHK(copy_out(col_ind, d_ci, sizeof(uint32_t) * nnz, s));
  if (!ov_dev) MSPCHK(copy_out(out_vals, d_ov, sizeof(float) * nnz, s));
  HIPCHK(hipStreamSynchronize(s));
  return MSP_OK;
}

// ---------------------------------------------------------------------------
// No-allocation COO -> CSR. Its temporaries (sort buffers, count tables,
// segment bounds, flags) come from `scratch`, one fixed array per device,
// reached through the device address hipGetSymbolAddress returns for that
// device (the array's name in host code is only a stand-in). Every call on a
// device uses that device's array: calls that overlap on one device race.
constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20;
constexpr uint64_t kNoAllocMaxKeys = (uint64_t)1 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

struct NoAllocScratch {
  std::mutex mu;
  Arena* arena[kMaxDevices] = {};  // fixed arenas over each device's scratch (never destroyed)
};
static NoAllocScratch& g_noalloc = *new NoAllocScratch;

// The fixed arena of the current device, set up on the first call there.
static msp_status noalloc_arena(Arena** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(MSP_ERR_UNSUPPORTED, "device %d beyond the library's limit", dev);
  std::lock_guard<std::mutex> lk(g_noalloc.mu);
  if (!g_noalloc.arena[dev]) {
    void* ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(MSP_ERR_HIP, "hipGetSymbolAddress(scratch) returned no address");
    Arena* a = new Arena;
    a->fixed = true;
    a->chunks.push_back(ArenaChunk{ptr, kNoAllocScratchBytes});
    g_noalloc.arena[dev] = a;
  }
  *out = g_noalloc.arena[dev];
  return MSP_OK;
}

// Routes the calling thread's temporaries to a fixed arena for one call.
struct FixedArenaScope {
  explicit FixedArenaScope(Arena* a) { t_fixed_arena = a; }
  ~FixedArenaScope() { t_fixed_arena = nullptr; }
};

extern "C" msp_status msp_coo_to_csr_noalloc(const uint64_t* keys, const float* vals, uint64_t nnz,
                                             uint32_t num_rows, uint32_t num_cols, uint32_t* row_ptr,
                                             uint32_t* col_ind, float* out_vals, void* stream) {
  MSPCHK(need_device());
  if (nnz > kNoAllocMaxKeys)
    return fail(MSP_ERR_INVALID, "msp_coo_to_csr_noalloc: %llu keys exceed its limit of %llu",
                (unsigned long long)nnz, (unsigned long long)kNoAllocMaxKeys);
  if (!row_ptr || !is_device_ptr(row_ptr))
    return fail(MSP_ERR_INVALID, "msp_coo_to_csr_noalloc: row_ptr must be device memory");
  if (nnz && (!is_device_ptr(keys) || !is_device_ptr(vals) || !is_device_ptr(col_ind) || !is_device_ptr(out_vals)))
    return fail(MSP_ERR_INVALID, "msp_coo_to_csr_noalloc: keys, vals, col_ind and out_vals must be device memory");
  Arena* a = nullptr;
  MSPCHK(noalloc_arena(&a));
  FixedArenaScope scope(a);
  return coo_to_csr_device(keys, vals, nnz, num_rows, num_cols, row_ptr, col_ind, out_vals, (hipStream_t)stream);
}

// ---------------------------------------------------------------------------
// ABI: handles
extern "C" msp_status msp_csr_create_coo(msp_csr* out, const uint64_t* keys, const float* vals, uint64_t nnz,
                                         uint32_t num_rows, uint32_t num_cols, void* stream) {
  if (!out) return fail(MSP_ERR_INVALID, "null output handle");
  *out = nullptr;
  MSPCHK(need_device());
  if (nnz && (!keys || !vals)) return fail(MSP_ERR_INVALID, "msp_csr_create_coo: null keys/values");
  if (nnz > 0xFFFFFFFFull - 64) return fail(MSP_ERR_INVALID, "nnz exceeds the u32 index range");
  hipStream_t s = (hipStream_t)stream;
  msp_csr A = new msp_csr_s;
  HIPCHK(hipGetDevice(&A->device));
  A->n = num_rows;
  A->m = num_cols;
  A->nnz = nnz;
  msp_status st = alloc_csr(A);
  InArr k, v;
  if (st == MSP_OK) st = k.set(keys, sizeof(uint64_t) * nnz, s);
  if (st == MSP_OK) st = v.set(vals, sizeof(float) * nnz, s);
  int64_t max_row = -1;
  if (st == MSP_OK)
    st = coo_to_csr_device(static_cast<const uint64_t*>(k.d), static_cast<const float*>(v.d), nnz, num_rows, num_cols,
                           A->row_ptr.as<uint32_t>(), A->col.as<uint32_t>(), A->val.as<float>(), s, nullptr, false,
                           &max_row);
  if (st == MSP_OK) st = zero_pad(A, s);
  if (st == MSP_OK) st = finalize_matrix(A, s, max_row);
  if (st == MSP_OK && hipStreamSynchronize(s) != hipSuccess) st = fail(MSP_ERR_HIP, "stream sync failed");
  if (st != MSP_OK) {
    delete A;
    return st;
  }
  *out = A;
  return MSP_OK;
}

extern "C" msp_status msp_csr_create_csr(msp_csr* out, const uint32_t* row_ptr, const uint32_t* col_ind,
                                         const float* vals, uint32_t num_rows, uint32_t num_cols, void* stream) {
  if (!out) return fail(MSP_ERR_INVALID, "null output handle");
  *out = nullptr;
  MSPCHK(need_device());
  if (!row_ptr) return fail(MSP_ERR_INVALID, "null row_ptr");
  hipStream_t s = (hipStream_t)stream;
  // nnz = row_ptr[num_rows]
  uint32_t nnz32 = 0;
  if (is_device_ptr(row_ptr)) {
    HIPCHK(hipMemcpyAsync(&nnz32, row_ptr + num_rows, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  } else {
    nnz32 = row_ptr[num_rows];
  }
  if (nnz32 && (!col_ind || !vals)) return fail(MSP_ERR_INVALID, "null col_ind/values");
  msp_csr A = new msp_csr_s;
  HIPCHK(hipGetDevice(&A->device));
  A->n = num_rows;
  A->m = num_cols;
  A->nnz = nnz32;
  msp_status st = alloc_csr(A);
  if (st == MSP_OK) st = copy_in(A->row_ptr.p, row_ptr, sizeof(uint32_t) * ((size_t)num_rows + 1), s);
  if (st == MSP_OK) st = copy_in(A->col.p, col_ind, sizeof(uint32_t) * A->nnz, s);
  if (st == MSP_OK) st = copy_in(A->val.p, vals, sizeof(float) * A->nnz, s);
  if (st == MSP_OK) st = zero_pad(A, s);
  if (st == MSP_OK) st = finalize_matrix(A, s);
  if (st == MSP_OK && hipStreamSynchronize(s) != hipSuccess) st = fail(MSP_ERR_HIP, "stream sync failed");
  if (st != MSP_OK) {
    delete A;
    return st;
  }
  *out = A;
  return MSP_OK;
}

extern "C" msp_status msp_csr_export(msp_csr A, uint32_t* row_ptr, uint32_t* col_ind, float* vals, void* stream) {
  MSPCHK(check_handle(A));
  hipStream_t s = (hipStream_t)stream;
  MSPCHK(copy_out(row_ptr, A->row_ptr.p, sizeof(uint32_t) * ((size_t)A->n + 1), s));
  MSPCHK(copy_out(col_ind, A->col.p, sizeof(uint32_t) * A->nnz, s));
  MSPCHK(copy_out(vals, A->val.p, sizeof(float) * A->nnz, s));
  HIPCHK(hipStreamSynchronize(s));
  return MSP_OK;
}

extern "C" msp_status msp_csr_shape(msp_csr A, uint32_t* num_rows, uint32_t* num_cols, uint64_t* nnz) {
  if (!A) return fail(MSP_ERR_INVALID, "null matrix handle");
  if (num_rows) *num_rows = A->n;
  if (num_cols) *num_cols = A->m;
  if (nnz) *nnz = A->nnz;
  return MSP_OK;
}

extern "C" msp_status msp_csr_prepare_values(msp_csr A, int32_t value_type, void* stream) {
  MSPCHK(check_handle(A));
  hipStream_t s = (hipStream_t)stream;
  if (value_type == MSP_VALUES_F32) return MSP_OK;
  if (value_type != MSP_VALUES_BF16 && value_type != MSP_VALUES_F16)
    return fail(MSP_ERR_INVALID, "unknown value type %d", value_type);
  if (A->val16_type == value_type && A->val16.p) return MSP_OK;
  size_t padded = (size_t)A->nnz + 64;
  MSPCHK(A->val16.alloc(sizeof(uint16_t) * padded));
  MSPCHK(A->dlo.alloc(sizeof(float) * ((size_t)A->n + 1)));
  const uint32_t g = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(((uint64_t)A->n + kBlock - 1) / kBlock, 65536));
  if (value_type == MSP_VALUES_BF16)
    k_cvt16_rowsum<true><<<g, kBlock, 0, s>>>(A->row_ptr.as<uint32_t>(), A->col.as<uint32_t>(), A->val.as<float>(),
                                              A->n, A->m, A->nnz, A->val16.as<uint16_t>(), A->dlo.as<float>());
  else
    k_cvt16_rowsum<false><<<g, kBlock, 0, s>>>(A->row_ptr.as<uint32_t>(), A->col.as<uint32_t>(), A->val.as<float>(),
                                               A->n, A->m, A->nnz, A->val16.as<uint16_t>(), A->dlo.as<float>());
  MSPCHK(launch_check("k_cvt16_rowsum"));
  A->val16_type = value_type;
  HIPCHK(hipStreamSynchronize(s));
  return MSP_OK;
}

extern "C" msp_status msp_csr_set_format(msp_csr A, int32_t format) {
  MSPCHK(check_handle(A));
  if (format != MSP_FORMAT_AUTO && format != MSP_FORMAT_CSR) return fail(MSP_ERR_INVALID, "unknown format %d", format);
  if (format == A->format) return MSP_OK;
  A->format = format;
  A->dia_state = 0;  // re-probe on the next use
  A->cg_fused[0] = A->cg_fused[1] = -1;
  A->dia16_type = -1;
  if (format == MSP_FORMAT_CSR) {
    A->dia_state = 2;
    A->dia.reset();
    A->dia16.reset();
    A->dia_dlo.reset();
  }
  return MSP_OK;
}

extern "C" msp_status msp_csr_get_format(msp_csr A, int32_t* format_in_use) {
  MSPCHK(check_handle(A));
  if (format_in_use) *format_in_use = A->dia_state == 1 ? (A->dia_sym ? 3 : 2) : MSP_FORMAT_CSR;
  return MSP_OK;
}

extern "C" msp_status msp_csr_destroy(msp_csr A) {
  if (!A) return MSP_OK;
  if (!in_teardown()) {
    (void)hipSetDevice(A->device);
    (void)hipDeviceSynchronize();
  }
  delete A;
  return MSP_OK;
}

// ---------------------------------------------------------------------------
// Operators: what a solver needs from a matrix kind (SpMV with fused epilogue,
// the true-residual check, the inverse diagonal).
template <typename V>
static CsrView<V> view_of(msp_csr A) {
  CsrView<V> v;
  v.row_ptr = A->row_ptr.as<uint32_t>();
  v.col = A->col.as<uint32_t>();
  if constexpr (std::is_same<V, ValF32>::value) {
    v.val = A->val.as<float>();
    v.dlo = nullptr;
  } else {
    v.val = A->val16.as<uint16_t>();
    v.dlo = A->dlo.as<float>();
  }
  v.blk = A->blk.as<uint32_t>();
  v.nblk = A->nblk;
  v.n = A->n;
  v.m = A->m;
  return v;
}

static GridRed red_of(const MatBase* A) { return A->ws.scal.p ? A->ws.gr() : GridRed{nullptr}; }

// Second half of a grid reduction: one workgroup folds the nb block partials
// of the kernel just launched into the finisher (nothing for FinNone).
template <int ND, typename Fin>
static void finish(const Fin& fin, GridRed rb, uint32_t nb, hipStream_t s) {
  if constexpr (ND > 0 && !std::is_same<Fin, FinNone>::value) k_finish<ND, Fin><<<1, kBlock, 0, s>>>(rb, nb, fin);
}

// Symmetric storage (the lower diagonals read from the upper planes) for
// value planes above this size: 256^3 FP32 product, 470 MB: 121 -> 94 us; at
// 128^3 (BF16 planes 29 MB, FP32 59 MB) the learned-FSAI solve 13.1 -> 12.9
// ms and plain CG 0.494 -> 0.450 s. (With the kernels' earlier 64-bit
// addressing and fewer rows per thread, the same learned-FSAI solve was
// slower with it, 17.8 -> 19.7 ms, and the threshold was 128 MB.)
constexpr size_t kDiaMirrorMinBytes = (size_t)16 << 20;

template <typename V>
static DiaView<V> dia_view_of(msp_csr A) {
  DiaView<V> v;
  if (const msp_csr_s* g = A->dia_alias) {  // a transpose read from the planes of g
    if constexpr (std::is_same<V, ValF32>::value) {
      v.val = g->dia.as<float>();
      v.dlo = nullptr;
    } else {
      v.val = g->dia16.as<uint16_t>();
      v.dlo = g->dia_dlo.as<float>();
    }
    v.ld = g->dia_ld;
    v.n = A->n;
    v.m = A->m;
    v.mirrored = 1;
    v.o = A->dia_alias_o;
    return v;
  }
  if constexpr (std::is_same<V, ValF32>::value) {
    v.val = A->dia.as<float>();
    v.dlo = nullptr;
  } else {
    v.val = A->dia16.as<uint16_t>();
    v.dlo = A->dia_dlo.as<float>();
  }
  v.ld = A->dia_ld;
  v.n = A->n;
  v.m = A->m;
  const size_t plane_bytes = sizeof(typename V::storage) * A->dia_ld * (size_t)A->dia_o.ndiag;
  v.mirrored = A->dia_sym && plane_bytes >= kDiaMirrorMinBytes ? 1 : 0;
  v.o = v.mirrored ? A->dia_sym_o : A->dia_o;
  return v;
}

// Workgroups of a kernel resident on one CU at once (1..8): the CSR kernel's
// persistent grid must not exceed residency, or the surplus workgroups run as
// a second wave after the first has finished.
template <typename K>
static uint32_t resident_per_cu(K kernel) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, kBlock, 0) != hipSuccess || nb < 1) nb = 1;
  return (uint32_t)std::min(nb, 8);
}

template <typename V, int ND, typename XL, typename Epi>
static uint32_t csr_grid(uint32_t nblk) {
  static const uint32_t per_cu = resident_per_cu(k_spmv<V, ND, XL, Epi>);
  return std::max(1u, std::min(nblk, per_cu * (uint32_t)cu_count()));
}

// Rows per thread of the diagonal kernel's load batch: 4 while a row needs at
// most 14 loads (factors, 2D stencils, the 7-point operator with a one-vector
// gather), 2 above (the 7-point operator forming its direction from two
// vectors: 4 there leaves 2-3 waves per SIMD), 1 for the generic diagonal
// count. Against one row per thread on the 128^3 solves: FP32 FSAI-PCG steps
// 66.3 -> 60.5 us, the other solvers 0.3-0.7% faster; 4 instead of 2 rows for
// 11-14 loads: the BF16 FSAI-PCG step's x / r update with w = G r 14.0 ->
// 13.3 us (with the 32-bit form for products with one sum, plain CG at 128^3
// 0.498 -> 0.472 s).
template <typename XL, int NDIAG>
constexpr int dia_rows_per_thread() {
  constexpr int loads = NDIAG * (1 + (int)(sizeof(typename XL::Raw) / sizeof(float)));
  return NDIAG == 0 ? 1 : (loads <= 14 ? 4 : 2);
}

// The diagonal kernel is grid-stride over rows: a grid beyond residency would
// leave its last workgroups running alone at the end.
// The marched walk (DiaMarch) for symmetric storage whose far offset c is
// large enough that c rows of the streamed planes outgrow an XCD's 4 MB L2:
// every work item must get its own workgroup within the residency (and the
// partial-sum buffer's 8 per CU), and a work item should cover several levels
// (one extra level of carried loads per item).
constexpr uint64_t kDiaMarchMinBytes = (uint64_t)1 << 20;
constexpr uint32_t kDiaMarchMinLevels = 8;

template <typename V, int NDIAG>
static bool dia_march_plan(const DiaView<V>& v, uint32_t cap, DiaMarch& mk) {
  if constexpr (NDIAG < 3 || (NDIAG & 1) == 0) {
    return false;
  } else {
    if (!v.mirrored) return false;
    const int32_t c = v.o.off[NDIAG - 1];
    if ((NDIAG & 1) == 0 || v.o.off[NDIAG / 2] != 0 || c <= 0 || v.o.off[0] != -c || (uint32_t)c >= v.n ||
        v.n >= kDiaMaxDim)
      return false;
    const uint64_t row_bytes = (uint64_t)(NDIAG / 2 + 1) * sizeof(typename V::storage) + 2 * sizeof(float);
    if ((uint64_t)c * row_bytes < kDiaMarchMinBytes) return false;
    const uint32_t ntile = ((uint32_t)c + kBlock - 1) / kBlock;
    const uint32_t nlev = (uint32_t)(((uint64_t)v.n + (uint32_t)c - 1) / (uint32_t)c);
    if (nlev < kDiaMarchMinLevels || ntile > cap) return false;
    uint32_t nchunk = std::max(1u, std::min(nlev / kDiaMarchMinLevels, cap / ntile));
    const uint32_t lpc = (nlev + nchunk - 1) / nchunk;
    nchunk = (nlev + lpc - 1) / lpc;
    mk.c = (uint32_t)c;
    mk.ntile = ntile;
    mk.nlev = nlev;
    mk.lpc = lpc;
    mk.nwork = ntile * nchunk;
    return true;
  }
}

template <typename V, int ND, typename XL, typename Epi, int NDIAG>
static uint32_t dia_launch(msp_csr A, XL xl, Epi epi, GridRed rb, const int* done, hipStream_t s) {
  constexpr int U = dia_rows_per_thread<XL, NDIAG>();
  if constexpr (NDIAG >= 3 && (NDIAG & 1)) {
    // levels per memory round trip: 4 with one-load gathers, 2 with
    // two-vector gathers (direction formed in the gather)
    constexpr int UZ = sizeof(typename RawOf<XL>::type) == sizeof(float) ? 4 : 2;
    const auto mkern = [] {
      if constexpr (ND > 0) return &k_dia_march_sum<V, ND, XL, Epi, NDIAG, UZ>;
      else return &k_dia_march<V, ND, XL, Epi, NDIAG, UZ>;
    }();
    static const uint32_t mper_cu = resident_per_cu(mkern);
    const DiaView<V> v = dia_view_of<V>(A);
    DiaMarch mk;
    if (dia_march_plan<V, NDIAG>(v, mper_cu * (uint32_t)cu_count(), mk)) {
      mkern<<<mk.nwork, kBlock, 0, s>>>(v, mk, xl, epi, rb, done);
      return mk.nwork;
    }
  }
  const auto kern = [] {
    if constexpr (ND > 0) return &k_dia_sum<V, ND, XL, Epi, NDIAG, U>;
    else return &k_dia<V, ND, XL, Epi, NDIAG, U>;
  }();
  static const uint32_t per_cu = resident_per_cu(kern);
  // as many grid-stride passes as residency requires, all equally full: a
  // resident-sized grid left the last pass mostly empty (2.1M rows at U = 4
  // over 1792 workgroups: 1.83M rows, then 0.26M) at the cost of a full round
  // trip
  const uint64_t per_wg = (uint64_t)kBlock * U;
  const uint64_t need = ((uint64_t)A->n + per_wg - 1) / per_wg;
  const uint64_t cap = (uint64_t)per_cu * (uint32_t)cu_count();
  const uint64_t passes = (need + cap - 1) / cap;
  const uint32_t g = (uint32_t)std::max<uint64_t>(1, (need + passes - 1) / passes);
  kern<<<g, kBlock, 0, s>>>(dia_view_of<V>(A), xl, epi, rb, done);
  return g;
}

// y = A x with a fused epilogue: diagonal storage when the matrix has it (and
// its values of type V are prepared), the CSR row-block kernel otherwise.
template <typename V, int ND, typename XL, typename Epi, typename Fin>
static uint32_t spmv_any(msp_csr A, XL xl, Epi epi, Fin fin, GridRed rb, const int* done, hipStream_t s) {
  const msp_csr_s* planes = A->dia_alias ? A->dia_alias : A;
  const bool dia = A->dia_state == 1 && (V::kType == MSP_VALUES_F32 || planes->dia16_type == V::kType);
  uint32_t g;
  if (dia) {
    // the diagonal kernel is grid-stride over rows: a grid beyond residency
    // would leave its last workgroups running alone at the end
    // (3, 5, 7: 1D/2D/3D stencils; 3 and 4 also their FSAI factors' lower triangles)
    const int nd = A->dia_o.ndiag;
    if (nd == 7) g = dia_launch<V, ND, XL, Epi, 7>(A, xl, epi, rb, done, s);
    else if (nd == 5) g = dia_launch<V, ND, XL, Epi, 5>(A, xl, epi, rb, done, s);
    else if (nd == 4) g = dia_launch<V, ND, XL, Epi, 4>(A, xl, epi, rb, done, s);
    else if (nd == 3) g = dia_launch<V, ND, XL, Epi, 3>(A, xl, epi, rb, done, s);
    else g = dia_launch<V, ND, XL, Epi, 0>(A, xl, epi, rb, done, s);
  } else {
    g = csr_grid<V, ND, XL, Epi>(A->nblk);
    k_spmv<V, ND, XL, Epi><<<g, kBlock, 0, s>>>(view_of<V>(A), xl, epi, rb, done);
  }
  finish<ND>(fin, rb, g, s);
  return g;
}

template <typename V>
struct OpCsr {
  msp_csr A;
  MatBase* base() const { return A; }
  template <int ND, typename Epi, typename Fin>
  uint32_t spmv(const float* x, Epi epi, Fin fin, const int* done, hipStream_t s) const {
    return spmv_any<V, ND>(A, XPlain{x}, epi, fin, red_of(A), done, s);
  }
  // y = A (r - alpha v)
  template <int ND, typename Epi, typename Fin>
  void spmv_axpy(const float* r, const float* v, const float* alpha, Epi epi, Fin fin, const int* done,
                 hipStream_t s) const {
    spmv_any<V, ND>(A, XAxpy{r, v, alpha}, epi, fin, red_of(A), done, s);
  }
  // CG step head: p = z + beta p_old formed in the gather, q = A p, <p,q>
  // partials (first reduction half); returns the number of partials.
  template <int PC>
  uint32_t cg_dir(const float* r, const float* pold, const float* dinv, CgDecide dec, float* q, float* pnew,
                  const int* done, hipStream_t s) const {
    const XPNew<PC> xl{r, pold, dinv, dec, 0.0f};
    return spmv_any<V, 1>(A, xl, EpiCgQ<XPNew<PC>>{q, pnew, xl}, FinNone{}, red_of(A), done, s);
  }
  RowsCsr rows() const { return RowsCsr{A->row_ptr.as<uint32_t>(), A->col.as<uint32_t>(), A->val.as<float>()}; }
  void inv_diag(float* dinv, int* bad, hipStream_t s) const {
    uint32_t g = std::max<uint32_t>(1, std::min<uint32_t>((A->n + kBlock - 1) / kBlock, 8192));
    k_inv_diag<<<g, kBlock, 0, s>>>(A->row_ptr.as<uint32_t>(), A->col.as<uint32_t>(), A->val.as<float>(), A->n, dinv,
                                    bad);
  }
};

static BcsrView bview_of(msp_bcsr A) {
  BcsrView v;
  v.brow_ptr = A->brow_ptr.as<uint32_t>();
  v.bcol = A->bcol.as<uint32_t>();
  v.blocks = A->blocks.as<float>();
  v.nbr = A->nbr;
  v.n = A->n;
  return v;
}

static uint32_t bspmv_grid(const msp_bcsr A) {
  const uint32_t rpb = (uint32_t)(kWave / (A->B * A->B / 4)) * kWavesPerBlock;
  return spmv_grid((A->nbr + rpb - 1) / rpb);
}

struct OpBcsr {
  msp_bcsr A;
  MatBase* base() const { return A; }
  template <int ND, typename XL, typename Epi, typename Fin>
  uint32_t launch(XL xl, Epi epi, Fin fin, const int* done, hipStream_t s) const {
    const uint32_t g = bspmv_grid(A);
    if (A->B == 4) k_bspmv<4, ND, XL, Epi><<<g, kBlock, 0, s>>>(bview_of(A), xl, epi, red_of(A), done);
    else if (A->B == 8) k_bspmv<8, ND, XL, Epi><<<g, kBlock, 0, s>>>(bview_of(A), xl, epi, red_of(A), done);
    else k_bspmv<16, ND, XL, Epi><<<g, kBlock, 0, s>>>(bview_of(A), xl, epi, red_of(A), done);
    finish<ND>(fin, red_of(A), g, s);
    return g;
  }
  template <int PC>
  uint32_t cg_dir(const float* r, const float* pold, const float* dinv, CgDecide dec, float* q, float* pnew,
                  const int* done, hipStream_t s) const {
    const X4PNew<PC> xl{r, pold, dinv, dec, 0.0f};
    return launch<1>(xl, EpiCgQ<X4PNew<PC>>{q, pnew, xl}, FinNone{}, done, s);
  }
  template <int ND, typename Epi, typename Fin>
  uint32_t spmv(const float* x, Epi epi, Fin fin, const int* done, hipStream_t s) const {
    return launch<ND>(X4Plain{x}, epi, fin, done, s);
  }
  template <int ND, typename Epi, typename Fin>
  void spmv_axpy(const float* r, const float* v, const float* alpha, Epi epi, Fin fin, const int* done,
                 hipStream_t s) const {
    launch<ND>(X4Axpy{r, v, alpha}, epi, fin, done, s);
  }
  RowsBcsr rows() const { return RowsBcsr{bview_of(A), A->B, A->m}; }
  void inv_diag(float* dinv, int* bad, hipStream_t s) const {
    uint32_t g = std::max<uint32_t>(1, std::min<uint32_t>((A->n + kBlock - 1) / kBlock, 8192));
    k_inv_diag_bcsr<<<g, kBlock, 0, s>>>(bview_of(A), A->B, A->m, dinv, bad);
  }
};

static msp_status check_values(msp_csr A, int32_t vt) {
  if (vt == MSP_VALUES_F32) return MSP_OK;
  if (vt != MSP_VALUES_BF16 && vt != MSP_VALUES_F16) return fail(MSP_ERR_INVALID, "unknown value type %d", vt);
  if (A->val16_type != vt) return fail(MSP_ERR_INVALID, "values of type %d not prepared (msp_csr_prepare_values)", vt);
  return MSP_OK;
}

// Run f(OpCsr<V>{A}) for the requested value type.
template <typename F>
static msp_status with_csr_op(msp_csr A, int32_t vt, hipStream_t s, F&& f) {
  MSPCHK(check_values(A, vt));
  MSPCHK(ensure_format(A, vt, s));
  if (vt == MSP_VALUES_BF16) return f(OpCsr<ValBF16>{A});
  if (vt == MSP_VALUES_F16) return f(OpCsr<ValF16>{A});
  return f(OpCsr<ValF32>{A});
}

// ---------------------------------------------------------------------------
// solver drivers (generic over the operator)
static void fill_opts(msp_solve_opts& o, const msp_solve_opts* in) {
  msp_default_opts(&o);
  if (in) o = *in;
  if (o.check_every <= 0) o.check_every = 16;
}

template <class Op>
static msp_status ensure_dinv(const Op& op, hipStream_t s) {
  MatBase* A = op.base();
  if (A->dinv_state == 1) return MSP_OK;
  if (A->dinv_state == 2) return fail(MSP_ERR_SINGULAR, "zero or non-finite diagonal entry (Jacobi)");
  MSPCHK(A->dinv.alloc(sizeof(float) * A->ws.npad));
  HIPCHK(hipMemsetAsync(A->dinv.p, 0, sizeof(float) * A->ws.npad, s));
  DevMem bad;
  MSPCHK(bad.alloc(sizeof(int)));
  HIPCHK(hipMemsetAsync(bad.p, 0, sizeof(int), s));
  op.inv_diag(A->dinv.as<float>(), bad.as<int>(), s);
  MSPCHK(launch_check("inv_diag"));
  int hb = 0;
  HIPCHK(hipMemcpyAsync(&hb, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  A->dinv_state = hb ? 2 : 1;
  if (hb) return fail(MSP_ERR_SINGULAR, "zero or non-finite diagonal entry (Jacobi)");
  return MSP_OK;
}

// Upload scalar settings, run `init`, then batches of `step` with one host
// poll of the done flag per batch (the poll trails by one batch, so the GPU
// never idles waiting for the host). A recurrence that reports convergence
// halts the steps until the true-residual check `verify` at the end of the
// batch (launched only when o.verify).
// Solver state on the host before a solve (tolerance, limits, criterion).
// (written by a one-thread kernel: a host-to-device copy cost a copy launch
// and its gap at the start of every correction of a refined solve)
static msp_status start_solve(Workspace& w, const msp_solve_opts& o, hipStream_t s) {
  k_start_solve<<<1, 1, 0, s>>>(w.S_dev(), (float)o.tol, o.maxiter, o.criterion, o.verify ? 1 : 0);
  MSPCHK(launch_check("k_start_solve"));
  HIPCHK(hipEventRecord(w.t0, s));
  return MSP_OK;
}

// Final state back to the host and into the info struct.
struct NoTail {
  msp_status operator()() const { return MSP_OK; }
};

// The solver's final state to the host. `tail` enqueues the work that follows
// the last step and needs no host decision (the held x update, the collapse
// of the compensated iterate, the copy out) before the one wait, so the GPU
// runs it straight after the last batch instead of after a host round trip.
template <typename Tail = NoTail>
static msp_status finish_solve(Workspace& w, msp_solve_info* info, hipStream_t s, Tail tail = Tail{}) {
  HIPCHK(hipEventRecord(w.t1, s));
  HIPCHK(hipMemcpyAsync(w.hS, w.S_dev(), sizeof(Scalars), hipMemcpyDeviceToHost, s));
  MSPCHK(tail());
  HIPCHK(spin_sync(s));
  if (info) {
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, w.t0, w.t1);
    info->iterations = w.hS->iter;
    info->status = w.hS->done == kRunning ? MSP_SOLVE_MAXITER : w.hS->done;
    info->verifications = w.hS->nverify;
    info->rr = w.hS->rr;
    info->bb = w.hS->bb;
    info->relres = w.hS->bb > 0 ? std::sqrt((double)w.hS->rr / (double)w.hS->bb) : std::sqrt((double)w.hS->rr);
    info->seconds = ms * 1e-3;
  }
  return MSP_OK;
}

struct NoFlush {
  void operator()() const {}
};

// `flush` runs after each batch of steps, before the verification and the
// host's read of the state (a solver whose steps leave a decision pending
// takes it there).
template <typename Init, typename Step, typename Verify, typename Flush = NoFlush, typename Tail = NoTail>
static msp_status run_solver(MatBase* A, const msp_solve_opts& o, msp_solve_info* info, hipStream_t s, Init init,
                             Step step, Verify verify, Flush flush = Flush{}, Tail tail = Tail{}) {
  Workspace& w = A->ws;
  MSPCHK(start_solve(w, o, s));
  MSPCHK(init());
  int enqueued = 0, slot = 0;
  bool have_prev = false;
  // a few extra steps beyond maxiter leave room for the final verification
  const int limit = o.maxiter + (o.verify ? 2 * o.check_every : 0);
  const int full = std::max(1, o.check_every);
  // batches of 4, 4, 8, 16, ... up to check_every: the second batch is queued
  // before the host has read the first batch's state, so a solve that
  // converges within the first batch pays 4 skipped steps, not 8 (BiCGSTAB on
  // the block-sparse system: 4 steps)
  int grow = std::min(4, full);
  int nbatch = 0;
  // Near convergence the batches shrink to the steps the recurrence still
  // needs: the host reads the state after each batch (one batch behind the
  // GPU), takes the residual's rate of decrease over the last two snapshots
  // and sizes the next batch to the estimated remainder past the batch in
  // flight, plus 2 (a halted step still costs a launch per kernel).
  int cap = full;
  int32_t it_a = -1, nv_a = 0;
  float rr_a = 0.0f;
  while (true) {
    const int batch = std::min(std::min(grow, cap), std::max(1, limit - enqueued));
    if (++nbatch > 1) grow = std::min(2 * grow, full);
    for (int k = 0; k < batch; ++k) step();
    flush();
    if (o.verify) verify();
    MSPCHK(launch_check("solver step"));
    enqueued += batch;
    HIPCHK(hipMemcpyAsync(&w.hsnap[slot], w.S_dev(), sizeof(Scalars), hipMemcpyDeviceToHost, s));
    HIPCHK(hipEventRecord(w.ev[slot], s));
    if (have_prev) {
      HIPCHK(spin_event(w.ev[slot ^ 1]));
      const Scalars& p = w.hsnap[slot ^ 1];
      if (p.done != 0) break;
      cap = full;
      if (o.criterion == kCritRelres && !p.verify && it_a >= 0 && p.iter > it_a && p.rr > 0.0f &&
          rr_a > p.rr && p.thresh_rec > 0.0f && std::isfinite(p.rr)) {
        const double rate = std::log((double)p.rr / rr_a) / (double)(p.iter - it_a);  // < 0
        const double rem = std::log((double)p.thresh_rec / p.rr) / rate;               // steps from p
        const double after = rem - batch;  // past the batch now in flight
        cap = (int)std::min<double>(full, std::max(2.0, std::ceil(after) + 2.0));
      }
      // a verification replaces the recurrence residual: start a new estimate
      it_a = p.nverify == nv_a ? p.iter : -1;
      nv_a = p.nverify;
      rr_a = p.rr;
    }
    have_prev = true;
    if (enqueued >= limit) break;
    slot ^= 1;
  }
  return finish_solve(w, info, s, tail);
}

static msp_status solver_prologue(MatBase* A, const float* b, const float* x, const msp_solve_opts& o,
                                  hipStream_t s) {
  if (A->n != A->m) return fail(MSP_ERR_INVALID, "solvers need a square matrix (%u x %u)", A->n, A->m);
  if ((!b || !x) && A->n) return fail(MSP_ERR_INVALID, "null b / x");
  if (o.tol < 0 || o.maxiter < 0) return fail(MSP_ERR_INVALID, "negative tol / maxiter");
  MSPCHK(ensure_workspace(A, s));
  MSPCHK(copy_in(A->ws.v(Workspace::B), b, sizeof(float) * A->n, s));
  if (o.use_x0) MSPCHK(copy_in(A->ws.v(Workspace::X), x, sizeof(float) * A->n, s));
  return MSP_OK;
}

// True-residual check (k_verify): lane groups per row for CSR matrices with
// long rows and for block rows (a lane per block column), one thread per row
// otherwise. Returns the workgroups launched (the finisher's partial count):
// the lane-group forms take up to 8 workgroups per CU, since a group walks
// its rows one after another and more groups in flight hide that latency.
template <int MODE, typename Op>
static uint32_t verify_launch(const Op& op, uint32_t g, const float* X, const float* XLo, const float* B, float* R,
                              float* P, float* R0, const float* D, uint32_t n, Scalars* S, GridRed rb,
                              hipStream_t s) {
  const auto lanes_grid = [&](uint32_t L) {
    const uint64_t groups = ((uint64_t)n + kBlock / L - 1) / (kBlock / L);
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(groups, (uint64_t)cu_count() * 8));
  };
  if constexpr (std::is_same<decltype(op.rows()), RowsCsr>::value) {
    const uint64_t avg = op.A->n ? op.A->nnz / op.A->n : 0;
    if (avg >= 64) {
      const uint32_t gl = lanes_grid(64);
      k_verify_lanes<MODE, 64><<<gl, kBlock, 0, s>>>(op.rows(), X, XLo, B, R, P, R0, D, n, S, rb);
      return gl;
    }
    if (avg >= 16) {
      const uint32_t gl = lanes_grid(16);
      k_verify_lanes<MODE, 16><<<gl, kBlock, 0, s>>>(op.rows(), X, XLo, B, R, P, R0, D, n, S, rb);
      return gl;
    }
  }
  if constexpr (std::is_same<decltype(op.rows()), RowsBcsr>::value) {
    const RowsBcsr rows = op.rows();
    const uint32_t gl = lanes_grid(rows.B);
    if (rows.B == 16) {
      k_verify_bcsr<MODE, 16><<<gl, kBlock, 0, s>>>(rows.A, rows.m, X, XLo, B, R, P, R0, D, n, S, rb);
      return gl;
    }
    if (rows.B == 8) {
      k_verify_bcsr<MODE, 8><<<gl, kBlock, 0, s>>>(rows.A, rows.m, X, XLo, B, R, P, R0, D, n, S, rb);
      return gl;
    }
    if (rows.B == 4) {
      k_verify_bcsr<MODE, 4><<<gl, kBlock, 0, s>>>(rows.A, rows.m, X, XLo, B, R, P, R0, D, n, S, rb);
      return gl;
    }
  }
  k_verify<MODE><<<g, kBlock, 0, s>>>(op.rows(), X, XLo, B, R, P, R0, D, n, S, rb);
  return g;
}

// The CG iterate carries its rounding error in a second vector (x + xl) only
// when the target is near FP32's attainable accuracy; looser solves (the
// corrections of iterative refinement, 1e-3 by default) update x alone and
// move two vectors less per step.
static bool compensate(const msp_solve_opts& o) { return o.tol < 1e-4; }

template <class Op>
static msp_status cg_impl(const Op& op, const float* b, float* x, const msp_solve_opts& o, msp_solve_info* info,
                          hipStream_t s,
                                bool sync_out = true) {
  MatBase* A = op.base();
  MSPCHK(solver_prologue(A, b, x, o, s));
  if (o.precond != MSP_PRECOND_NONE && o.precond != MSP_PRECOND_JACOBI)
    return fail(MSP_ERR_UNSUPPORTED, "unknown preconditioner %d", o.precond);
  if (o.precond == MSP_PRECOND_JACOBI) MSPCHK(ensure_dinv(op, s));
  Workspace& w = A->ws;
  float *X = w.v(Workspace::X), *R = w.v(Workspace::R), *P = w.v(Workspace::P), *Q = w.v(Workspace::Q);
  float* XLo = w.v(Workspace::XL);
  const float* B = w.v(Workspace::B);
  const float* D = A->dinv.as<float>();
  Scalars* S = w.S_dev();
  Scalars* S2 = w.S_stage();
  const bool comp = compensate(o);
  if constexpr (std::is_same<Op, OpCsr<ValF32>>::value) {
    // small systems: the whole solve in one workgroup (no per-step launches)
    if (o.cg_form == 0 && A->n > 0 && A->n <= kSmallN) {
      MSPCHK(start_solve(w, o, s));
      const RowsCsr rows = op.rows();
      const float* Bv = w.v(Workspace::B);
      const bool j = o.precond == MSP_PRECOND_JACOBI;
      if (j && comp) k_cg_small<kPcJacobi, true><<<1, kSmallThreads, 0, s>>>(rows, Bv, X, D, A->n, o.use_x0, S);
      else if (j) k_cg_small<kPcJacobi, false><<<1, kSmallThreads, 0, s>>>(rows, Bv, X, D, A->n, o.use_x0, S);
      else if (comp) k_cg_small<kPcNone, true><<<1, kSmallThreads, 0, s>>>(rows, Bv, X, D, A->n, o.use_x0, S);
      else k_cg_small<kPcNone, false><<<1, kSmallThreads, 0, s>>>(rows, Bv, X, D, A->n, o.use_x0, S);
      MSPCHK(launch_check("k_cg_small"));
      MSPCHK(finish_solve(w, info, s, [&]() { return copy_out(x, X, sizeof(float) * A->n, s); }));
      HIPCHK(hipStreamSynchronize(s));
      return MSP_OK;
    }
  }
  GridRed rb = w.gr();
  const uint32_t n = A->n, gv = vec_grid(n);
  const bool jac = o.precond == MSP_PRECOND_JACOBI;
  auto init = [&]() -> msp_status {
    if (o.use_x0) op.template spmv<1>(X, EpiResidual{R, B}, FinNone{}, nullptr, s);
    k_cg_init<<<gv, kBlock, 0, s>>>(X, XLo, R, P, B, D, n, o.use_x0, jac ? kPcJacobi : kPcNone, S, rb);
    finish<3>(FinCgInit{S}, rb, gv, s);
    return launch_check("cg init");
  };
  // ping-pong directions: step k reads P[k & 1], writes P[(k + 1) & 1] (P, Z).
  // Two forms of the step compute the same recurrence: the direction formed in
  // the SpMV's gather (2 launches; best while the gathered neighbourhood of r
  // and p_old stays in cache), or a separate p update feeding a plain SpMV
  // (3 launches). The first solve on a matrix times
  // steps 0-7 in one form and 8-15 in the other and keeps the faster. The
  // first kernel of a step takes the previous step's decision from the
  // update's partials (rb2), stages it in S2 and the update commits it; the
  // last step's decision is taken by a finisher at the end of each batch.
  float* Pb[2] = {P, w.v(Workspace::Z)};
  int k = 0;
  const GridRed rb2 = w.gr2();
  hipEvent_t tev[3] = {nullptr, nullptr, nullptr};
  int& form = A->cg_fused[jac ? 1 : 0];
  if (o.cg_form < 0 || o.cg_form > 2) return fail(MSP_ERR_INVALID, "cg_form must be 0, 1 or 2");
  const bool tune = o.cg_form == 0 && form < 0;
  if (tune)
    for (auto& e : tev) HIPCHK(hipEventCreate(&e));
  auto step = [&]() {
    const float* pold = Pb[k & 1];
    float* pnew = Pb[(k + 1) & 1];
    const bool fused = o.cg_form == 1 || (o.cg_form == 0 && (tune ? k < 8 : form == 1));
    if (tune && (k == 0 || k == 8 || k == 16)) (void)hipEventRecord(tev[k / 8], s);
    if (tune && k == 16) {
      // decide only if all 16 steps did work (none halted by convergence): the
      // state holds the decisions of steps 0-14 (the last one pending), and
      // step 15 ran if none of them halted
      float ms_f = 0.f, ms_u = 0.f;
      Scalars hs;
      if (hipEventSynchronize(tev[2]) == hipSuccess && hipEventElapsedTime(&ms_f, tev[0], tev[1]) == hipSuccess &&
          hipEventElapsedTime(&ms_u, tev[1], tev[2]) == hipSuccess &&
          hipMemcpy(&hs, S, sizeof hs, hipMemcpyDeviceToHost) == hipSuccess && hs.iter >= 15 && hs.halt == 0)
        form = ms_f <= ms_u ? 1 : 0;
    }
    // the update's <r,r>, <r,z> partials of the previous step (none before the first)
    const CgDecide dec{S, S2, rb2, k ? gv : 0u};
    uint32_t g;
    if (fused) {
      g = jac ? op.template cg_dir<kPcJacobi>(R, pold, D, dec, Q, pnew, &S->halt, s)
              : op.template cg_dir<kPcNone>(R, pold, D, dec, Q, pnew, &S->halt, s);
    } else {
      if (jac) k_cg_update_p<kPcJacobi><<<gv, kBlock, 0, s>>>(pnew, pold, R, D, n, dec);
      else k_cg_update_p<kPcNone><<<gv, kBlock, 0, s>>>(pnew, pold, R, D, n, dec);
      g = op.template spmv<1>(pnew, EpiStoreDotX{Q, pnew}, FinNone{}, &S2->halt, s);
    }
    // (pold holds the previous step's direction for a held update of x; pnew
    // is buffer (k + 1) & 1, recorded when this step holds its own)
    const int32_t pbuf = (k + 1) & 1;
    if (jac) {
      if (comp)
        k_cg_update_xr<kPcJacobi, true><<<gv, kBlock, 0, s>>>(X, XLo, R, pnew, pold, Q, D, n, S, S2, rb2, rb, g, pbuf);
      else
        k_cg_update_xr<kPcJacobi, false><<<gv, kBlock, 0, s>>>(X, XLo, R, pnew, pold, Q, D, n, S, S2, rb2, rb, g, pbuf);
    } else {
      if (comp)
        k_cg_update_xr<kPcNone, true><<<gv, kBlock, 0, s>>>(X, XLo, R, pnew, pold, Q, D, n, S, S2, rb2, rb, g, pbuf);
      else
        k_cg_update_xr<kPcNone, false><<<gv, kBlock, 0, s>>>(X, XLo, R, pnew, pold, Q, D, n, S, S2, rb2, rb, g, pbuf);
    }
    ++k;
  };
  // a held update of x applied before a true-residual check runs (VER) or
  // before the iterate is returned
  auto xflush = [&](bool ver) {
    if (ver) {
      if (comp) k_cg_xflush<true, true><<<gv, kBlock, 0, s>>>(X, XLo, Pb[0], Pb[1], n, S);
      else k_cg_xflush<false, true><<<gv, kBlock, 0, s>>>(X, XLo, Pb[0], Pb[1], n, S);
      k_cg_xflush_clear<true><<<1, 1, 0, s>>>(S);
    } else {
      if (comp) k_cg_xflush<true, false><<<gv, kBlock, 0, s>>>(X, XLo, Pb[0], Pb[1], n, S);
      else k_cg_xflush<false, false><<<gv, kBlock, 0, s>>>(X, XLo, Pb[0], Pb[1], n, S);
      k_cg_xflush_clear<false><<<1, 1, 0, s>>>(S);
    }
  };
  // the last step's decision, before the host reads the state
  auto flush = [&]() {
    if (!k) return;
    if (jac) finish<2>(FinCgStepPending<kPcJacobi>{S}, rb2, gv, s);
    else finish<2>(FinCgStepPending<kPcNone>{S}, rb2, gv, s);
  };
  // a restart (beta = 0) makes the next direction z whatever P[k & 1] holds
  auto verify = [&]() {
    xflush(true);
    if (jac) {
      const uint32_t gvf = verify_launch<1>(op, gv, X, XLo, B, R, Pb[k & 1], nullptr, D, n, S, rb, s);
      finish<2>(FinVerify<1>{S}, rb, gvf, s);
    } else {
      const uint32_t gvf = verify_launch<0>(op, gv, X, XLo, B, R, Pb[k & 1], nullptr, D, n, S, rb, s);
      finish<2>(FinVerify<0>{S}, rb, gvf, s);
    }
  };
  auto tail = [&]() -> msp_status {
    xflush(false);
    k_collapse<<<gv, kBlock, 0, s>>>(X, XLo, n);
    return copy_out(x, X, sizeof(float) * n, s);
  };
  const msp_status st = run_solver(A, o, info, s, init, step, verify, flush, tail);
  for (auto& e : tev)
    if (e) (void)hipEventDestroy(e);
  MSPCHK(st);
  // (a caller inside the library that keeps working on s skips the wait)
  if (sync_out) HIPCHK(hipStreamSynchronize(s));
  return MSP_OK;
}

template <class Op>
static msp_status bicgstab_impl(const Op& op, const float* b, float* x, const msp_solve_opts& o,
                                msp_solve_info* info, hipStream_t s,
                                bool sync_out = true) {
  MatBase* A = op.base();
  MSPCHK(solver_prologue(A, b, x, o, s));
  if (o.precond != MSP_PRECOND_NONE) return fail(MSP_ERR_UNSUPPORTED, "BiCGSTAB takes no preconditioner");
  Workspace& w = A->ws;
  float *X = w.v(Workspace::X), *R = w.v(Workspace::R), *P = w.v(Workspace::P), *Vv = w.v(Workspace::Q);
  float *R0 = w.v(Workspace::R0), *Sv = w.v(Workspace::S), *T = w.v(Workspace::T);
  float* XLo = w.v(Workspace::XL);
  const float* B = w.v(Workspace::B);
  Scalars* S = w.S_dev();
  GridRed rb = w.gr();
  const uint32_t n = A->n, gv = vec_grid(n);
  auto init = [&]() -> msp_status {
    if (o.use_x0) op.template spmv<1>(X, EpiResidual{R, B}, FinNone{}, nullptr, s);
    k_bicg_init<<<gv, kBlock, 0, s>>>(X, XLo, R, R0, P, B, n, o.use_x0, S, rb);
    finish<2>(FinBicgInit{S}, rb, gv, s);
    return launch_check("bicgstab init");
  };
  auto step = [&]() {
    op.template spmv<1>(P, EpiStoreDot{Vv, R0}, FinBicgAlpha{S}, &S->halt, s);
    op.template spmv_axpy<2>(R, Vv, &S->alpha, EpiBicgT{T, Sv, R, Vv, &S->alpha}, FinBicgOmega{S}, &S->halt, s);
    k_bicg_update<<<gv, kBlock, 0, s>>>(X, XLo, R, P, Sv, T, R0, n, S, rb);
    finish<2>(FinBicgStep{S}, rb, gv, s);
    k_bicg_p<<<gv, kBlock, 0, s>>>(P, R, Vv, n, S);
  };
  auto verify = [&]() {
    const uint32_t gvf = verify_launch<2>(op, gv, X, XLo, B, R, P, R0, nullptr, n, S, rb, s);
    finish<2>(FinVerify<2>{S}, rb, gvf, s);
  };
  auto tail = [&]() -> msp_status {
    k_collapse<<<gv, kBlock, 0, s>>>(X, XLo, n);
    return copy_out(x, X, sizeof(float) * n, s);
  };
  MSPCHK(run_solver(A, o, info, s, init, step, verify, NoFlush{}, tail));
  // (a caller inside the library that keeps working on s skips the wait)
  if (sync_out) HIPCHK(hipStreamSynchronize(s));
  return MSP_OK;
}

template <class Op>
static msp_status richardson_impl(const Op& op, const float* b, float* x, float omega, int32_t iters,
                                  int32_t use_x0, hipStream_t s) {
  MatBase* A = op.base();
  if (iters < 0) return fail(MSP_ERR_INVALID, "negative iteration count");
  msp_solve_opts o;
  fill_opts(o, nullptr);
  o.use_x0 = use_x0;
  MSPCHK(solver_prologue(A, b, x, o, s));
  Workspace& w = A->ws;
  float* xa = w.v(Workspace::X);
  float* xb = w.v(Workspace::X2);
  const float* B = w.v(Workspace::B);
  if (!use_x0) HIPCHK(hipMemsetAsync(xa, 0, sizeof(float) * w.npad, s));
  for (int32_t it = 0; it < iters; ++it) {
    op.template spmv<0>(xa, EpiRichardson{xb, xa, B, omega}, FinNone{}, nullptr, s);
    std::swap(xa, xb);
  }
  MSPCHK(launch_check("richardson"));
  MSPCHK(copy_out(x, xa, sizeof(float) * A->n, s));
  HIPCHK(hipStreamSynchronize(s));
  return MSP_OK;
}

// y = A x for caller arrays: used in place when both are device pointers with
// the alignment / padding the kernel needs, else staged through the workspace.
template <class Op>
static msp_status spmv_impl(const Op& op, const float* x, float* y, bool need_padded_x, hipStream_t s) {
  MatBase* A = op.base();
  if ((!x && A->m) || (!y && A->n)) return fail(MSP_ERR_INVALID, "null x / y");
  const bool xin = is_device_ptr(x) && !need_padded_x, yin = is_device_ptr(y);
  if (xin && yin) {
    op.template spmv<0>(x, EpiStore{y}, FinNone{}, nullptr, s);
    return launch_check("spmv");
  }
  MSPCHK(ensure_workspace(A, s));
  float* xw = A->ws.v(Workspace::X2);
  float* yw = A->ws.v(Workspace::W);
  const float* xd = x;
  if (!xin) {
    MSPCHK(copy_in(xw, x, sizeof(float) * A->m, s));
    xd = xw;
  }
  op.template spmv<0>(xd, EpiStore{yin ? y : yw}, FinNone{}, nullptr, s);
  MSPCHK(launch_check("spmv"));
  if (!yin) MSPCHK(copy_out(y, yw, sizeof(float) * A->n, s));
  HIPCHK(hipStreamSynchronize(s));
  return MSP_OK;
}

// ---------------------------------------------------------------------------
// ABI: CSR compute
static bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

extern "C" msp_status msp_csr_spmv(msp_csr A, const float* x, float* y, int32_t value_type, void* stream) {
  MSPCHK(check_handle(A));
  hipStream_t s = (hipStream_t)stream;
  // the kernel may read x in aligned 16-byte quads: stage a misaligned x
  return with_csr_op(A, value_type, s, [&](auto op) { return spmv_impl(op, x, y, !aligned16(x), s); });
}

extern "C" msp_status msp_csr_cg(msp_csr A, const float* b, float* x, const msp_solve_opts* opts,
                                 msp_solve_info* info, void* stream) {
  MSPCHK(check_handle(A));
  msp_solve_opts o;
  fill_opts(o, opts);
  hipStream_t s = (hipStream_t)stream;
  return with_csr_op(A, o.values, s, [&](auto op) { return cg_impl(op, b, x, o, info, s); });
}

extern "C" msp_status msp_csr_bicgstab(msp_csr A, const float* b, float* x, const msp_solve_opts* opts,
                                       msp_solve_info* info, void* stream) {
  MSPCHK(check_handle(A));
  msp_solve_opts o;
  fill_opts(o, opts);
  hipStream_t s = (hipStream_t)stream;
  return with_csr_op(A, o.values, s, [&](auto op) { return bicgstab_impl(op, b, x, o, info, s); });
}

extern "C" msp_status msp_csr_richardson(msp_csr A, const float* b, float* x, float omega, int32_t iters,
                                         int32_t use_x0, int32_t value_type, void* stream) {
  MSPCHK(check_handle(A));
  hipStream_t s = (hipStream_t)stream;
  return with_csr_op(A, value_type, s,
                     [&](auto op) { return richardson_impl(op, b, x, omega, iters, use_x0, s); });
}

// ---------------------------------------------------------------------------
// ABI: BCSR
static bool block_ok(uint32_t B) { return B == 4 || B == 8 || B == 16; }

static msp_status bcsr_alloc(msp_bcsr A, uint32_t num_rows, uint32_t num_cols, uint32_t B, uint64_t nblk) {
  A->n = num_rows;
  A->m = num_cols;
  A->B = B;
  A->nbr = (num_rows + B - 1) / B;
  A->nbc = (num_cols + B - 1) / B;
  A->nblk = nblk;
  MSPCHK(A->brow_ptr.alloc(sizeof(uint32_t) * ((size_t)A->nbr + 1)));
  MSPCHK(A->bcol.alloc(sizeof(uint32_t) * (nblk + 1)));
  MSPCHK(A->blocks.alloc(sizeof(float) * (nblk * B * B + 16)));
  return MSP_OK;
}

extern "C" msp_status msp_bcsr_create(msp_bcsr* out, const uint32_t* brow_ptr, const uint32_t* bcol,
                                      const float* blocks, uint32_t num_rows, uint32_t num_cols, uint32_t block,
                                      void* stream) {
  if (!out) return fail(MSP_ERR_INVALID, "null output handle");
  *out = nullptr;
  MSPCHK(need_device());
  if (!block_ok(block)) return fail(MSP_ERR_UNSUPPORTED, "block size %u (4, 8 or 16)", block);
  if (!brow_ptr) return fail(MSP_ERR_INVALID, "null brow_ptr");
  hipStream_t s = (hipStream_t)stream;
  const uint32_t nbr = (num_rows + block - 1) / block;
  uint32_t nblk = 0;
  if (is_device_ptr(brow_ptr)) {
    HIPCHK(hipMemcpyAsync(&nblk, brow_ptr + nbr, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  } else {
    nblk = brow_ptr[nbr];
  }
  if (nblk && (!bcol || !blocks)) return fail(MSP_ERR_INVALID, "null bcol / blocks");
  msp_bcsr A = new msp_bcsr_s;
  HIPCHK(hipGetDevice(&A->device));
  msp_status st = bcsr_alloc(A, num_rows, num_cols, block, nblk);
  if (st == MSP_OK) st = copy_in(A->brow_ptr.p, brow_ptr, sizeof(uint32_t) * ((size_t)nbr + 1), s);
  if (st == MSP_OK) st = copy_in(A->bcol.p, bcol, sizeof(uint32_t) * nblk, s);
  if (st == MSP_OK) st = copy_in(A->blocks.p, blocks, sizeof(float) * (size_t)nblk * block * block, s);
  if (st == MSP_OK && hipStreamSynchronize(s) != hipSuccess) st = fail(MSP_ERR_HIP, "stream sync failed");
  if (st != MSP_OK) {
    delete A;
    return st;
  }
  *out = A;
  return MSP_OK;
}

// CSR (device arrays) -> BCSR on the GPU.
static msp_status csr_to_bcsr_device(msp_bcsr A, const uint32_t* row_ptr, const uint32_t* col, const float* val,
                                     uint64_t nnz, hipStream_t s) {
  const uint32_t B = A->B, n = A->n, nbr = A->nbr, nbc = A->nbc;
  const uint32_t g = std::max<uint32_t>(1, std::min<uint32_t>((n + kBlock - 1) / kBlock, 65536));
  const uint32_t gz = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((nnz + kBlock - 1) / kBlock, 65536));
  TmpMem bkeys, payload, brp_nnz, sbcol, sbrow, spayload, rows, flag, excl, sums;
  MSPCHK(bkeys.alloc(sizeof(uint64_t) * nnz, s));
  MSPCHK(payload.alloc(sizeof(float) * nnz, s));
  MSPCHK(brp_nnz.alloc(sizeof(uint32_t) * ((size_t)nbr + 1), s));
  MSPCHK(sbcol.alloc(sizeof(uint32_t) * nnz, s));
  MSPCHK(sbrow.alloc(sizeof(uint32_t) * nnz, s));
  MSPCHK(spayload.alloc(sizeof(float) * nnz, s));
  MSPCHK(rows.alloc(sizeof(uint32_t) * nnz, s));
  MSPCHK(flag.alloc(sizeof(uint32_t) * nnz, s));
  MSPCHK(excl.alloc(sizeof(uint32_t) * nnz, s));
  uint32_t L = 1;
  while (L < 64 && L * n < nnz) L <<= 1;  // lanes per row ~ mean row length
  k_block_keys<<<g, kBlock, 0, s>>>(row_ptr, col, n, B, L, bkeys.as<uint64_t>(), payload.as<float>(),
                                    rows.as<uint32_t>());
  MSPCHK(launch_check("block keys"));
  MSPCHK(coo_to_csr_device(bkeys.as<uint64_t>(), payload.as<float>(), nnz, nbr, nbc, brp_nnz.as<uint32_t>(),
                           sbcol.as<uint32_t>(), spayload.as<float>(), s, sbrow.as<uint32_t>()));
  k_block_heads<<<gz, kBlock, 0, s>>>(sbrow.as<uint32_t>(), sbcol.as<uint32_t>(), nnz, flag.as<uint32_t>());
  // exclusive scan of the head flags
  const uint32_t nscan = (uint32_t)((nnz + sortk::kScanTile - 1) / sortk::kScanTile);
  MSPCHK(sums.alloc(sizeof(uint32_t) * nscan, s));
  HIPCHK(hipMemcpyAsync(excl.p, flag.p, sizeof(uint32_t) * nnz, hipMemcpyDeviceToDevice, s));
  sortk::k_scan_reduce<<<nscan, kBlock, 0, s>>>(excl.as<uint32_t>(), nnz, sums.as<uint32_t>());
  sortk::k_scan_top<<<1, kBlock, 0, s>>>(sums.as<uint32_t>(), nscan);
  sortk::k_scan_down<<<nscan, kBlock, 0, s>>>(excl.as<uint32_t>(), nnz, sums.as<uint32_t>());
  MSPCHK(launch_check("block scan"));
  uint32_t last[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(&last[0], excl.as<uint32_t>() + nnz - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(&last[1], flag.as<uint32_t>() + nnz - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  const uint64_t nblk = (uint64_t)last[0] + last[1];
  A->nblk = nblk;
  MSPCHK(A->bcol.alloc(sizeof(uint32_t) * (nblk + 1)));
  MSPCHK(A->blocks.alloc(sizeof(float) * (nblk * B * B + 16)));
  HIPCHK(hipMemsetAsync(A->blocks.p, 0, sizeof(float) * (nblk * B * B + 16), s));
  k_block_fill<<<gz, kBlock, 0, s>>>(sbcol.as<uint32_t>(), flag.as<uint32_t>(), excl.as<uint32_t>(),
                                     spayload.as<float>(), rows.as<uint32_t>(), col, val, nnz, B,
                                     A->bcol.as<uint32_t>(), A->blocks.as<float>());
  const uint32_t gb = (nbr + 1 + kBlock - 1) / kBlock;
  k_brow_ptr<<<gb, kBlock, 0, s>>>(brp_nnz.as<uint32_t>(), excl.as<uint32_t>(), flag.as<uint32_t>(), nbr, nnz,
                                   A->brow_ptr.as<uint32_t>());
  MSPCHK(launch_check("block fill"));
  HIPCHK(hipStreamSynchronize(s));
  return MSP_OK;
}

extern "C" msp_status msp_bcsr_create_coo(msp_bcsr* out, const uint64_t* keys, const float* vals, uint64_t nnz,
                                          uint32_t num_rows, uint32_t num_cols, uint32_t block, void* stream) {
  if (!out) return fail(MSP_ERR_INVALID, "null output handle");
  *out = nullptr;
  MSPCHK(need_device());
  if (!block_ok(block)) return fail(MSP_ERR_UNSUPPORTED, "block size %u (4, 8 or 16)", block);
  if (nnz && (!keys || !vals)) return fail(MSP_ERR_INVALID, "null keys / values");
  if (nnz > 0xFFFFFFFFull - 64) return fail(MSP_ERR_INVALID, "nnz exceeds the u32 index range");
  hipStream_t s = (hipStream_t)stream;
  InArr k, v;
  MSPCHK(k.set(keys, sizeof(uint64_t) * nnz, s));
  MSPCHK(v.set(vals, sizeof(float) * nnz, s));
  DevMem rp, ci, cv;
  MSPCHK(rp.alloc(sizeof(uint32_t) * ((size_t)num_rows + 1)));
  MSPCHK(ci.alloc(sizeof(uint32_t) * (nnz + 1)));
  MSPCHK(cv.alloc(sizeof(float) * (nnz + 1)));
  MSPCHK(coo_to_csr_device(static_cast<const uint64_t*>(k.d), static_cast<const float*>(v.d), nnz, num_rows, num_cols,
                           rp.as<uint32_t>(), ci.as<uint32_t>(), cv.as<float>(), s));
  msp_bcsr A = new msp_bcsr_s;
  HIPCHK(hipGetDevice(&A->device));
  msp_status st = bcsr_alloc(A, num_rows, num_cols, block, 0);
  if (st == MSP_OK && nnz > 0)
    st = csr_to_bcsr_device(A, rp.as<uint32_t>(), ci.as<uint32_t>(), cv.as<float>(), nnz, s);
  if (st == MSP_OK && nnz == 0) {
    if (hipMemsetAsync(A->brow_ptr.p, 0, sizeof(uint32_t) * ((size_t)A->nbr + 1), s) != hipSuccess)
      st = fail(MSP_ERR_HIP, "memset failed");
  }
  if (st == MSP_OK && hipStreamSynchronize(s) != hipSuccess) st = fail(MSP_ERR_HIP, "stream sync failed");
  if (st != MSP_OK) {
    delete A;
    return st;
  }
  *out = A;
  return MSP_OK;
}

extern "C" msp_status msp_bcsr_shape(msp_bcsr A, uint32_t* num_rows, uint32_t* num_cols, uint32_t* block,
                                     uint64_t* num_blocks) {
  if (!A) return fail(MSP_ERR_INVALID, "null matrix handle");
  if (num_rows) *num_rows = A->n;
  if (num_cols) *num_cols = A->m;
  if (block) *block = A->B;
  if (num_blocks) *num_blocks = A->nblk;
  return MSP_OK;
}

extern "C" msp_status msp_bcsr_export(msp_bcsr A, uint32_t* brow_ptr, uint32_t* bcol, float* blocks, void* stream) {
  MSPCHK(check_handle(A));
  hipStream_t s = (hipStream_t)stream;
  MSPCHK(copy_out(brow_ptr, A->brow_ptr.p, sizeof(uint32_t) * ((size_t)A->nbr + 1), s));
  MSPCHK(copy_out(bcol, A->bcol.p, sizeof(uint32_t) * A->nblk, s));
  MSPCHK(copy_out(blocks, A->blocks.p, sizeof(float) * A->nblk * A->B * A->B, s));
  HIPCHK(hipStreamSynchronize(s));
  return MSP_OK;
}

extern "C" msp_status msp_bcsr_destroy(msp_bcsr A) {
  if (!A) return MSP_OK;
  if (!in_teardown()) {
    (void)hipSetDevice(A->device);
    (void)hipDeviceSynchronize();
  }
  delete A;
  return MSP_OK;
}

extern "C" msp_status msp_bcsr_spmv(msp_bcsr A, const float* x, float* y, void* stream) {
  MSPCHK(check_handle(A));
  hipStream_t s = (hipStream_t)stream;
  // the kernel reads x in 16-byte quads up to nbc * B: stage unless that is safe
  const bool padded_ok = aligned16(x) && (A->m % A->B == 0);
  return spmv_impl(OpBcsr{A}, x, y, !padded_ok, s);
}

extern "C" msp_status msp_bcsr_cg(msp_bcsr A, const float* b, float* x, const msp_solve_opts* opts,
                                  msp_solve_info* info, void* stream) {
  MSPCHK(check_handle(A));
  msp_solve_opts o;
  fill_opts(o, opts);
  if (o.values != MSP_VALUES_F32) return fail(MSP_ERR_UNSUPPORTED, "BCSR stores FP32 values");
  return cg_impl(OpBcsr{A}, b, x, o, info, (hipStream_t)stream);
}

extern "C" msp_status msp_bcsr_bicgstab(msp_bcsr A, const float* b, float* x, const msp_solve_opts* opts,
                                        msp_solve_info* info, void* stream) {
  MSPCHK(check_handle(A));
  msp_solve_opts o;
  fill_opts(o, opts);
  if (o.values != MSP_VALUES_F32) return fail(MSP_ERR_UNSUPPORTED, "BCSR stores FP32 values");
  return bicgstab_impl(OpBcsr{A}, b, x, o, info, (hipStream_t)stream);
}

extern "C" msp_status msp_bcsr_richardson(msp_bcsr A, const float* b, float* x, float omega, int32_t iters,
                                          int32_t use_x0, void* stream) {
  MSPCHK(check_handle(A));
  return richardson_impl(OpBcsr{A}, b, x, omega, iters, use_x0, (hipStream_t)stream);
}

// ---------------------------------------------------------------------------
// ABI: mixed-precision iterative refinement
extern "C" void msp_default_refine_opts(msp_refine_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof *o);
  o->tol = 1e-10;
  o->inner_tol = 0.0;  // automatic (resolved_inner_tol)
  o->max_outer = 30;
  o->inner_maxiter = 10000;
  o->inner_solver = 0;
  o->values = MSP_VALUES_F32;
  o->precond = MSP_PRECOND_NONE;
  o->use_x0 = 0;
}

static bool is_device_ptr_d(const double* p) { return is_device_ptr(static_cast<const void*>(p)); }

// FP64 outer loop: r = b - A x in FP64 (FP32 matrix values), correction
// A d = r / ||r|| by `inner(rhs, d, inner_tol, &info, &stop)` in FP32,
// x += ||r|| d in FP64. A correction is solved only as far as the target
// needs: to max(inner_tol, 0.3 tol / relres) (at most 0.5), so the last one
// does not over-solve. Stops at relres <= tol, after two outer steps that fail
// to halve the residual (stagnated), or after max_outer corrections.
template <class Inner>
static msp_status refine_impl(msp_csr A, const double* b, double* x, const msp_refine_opts& ro, msp_refine_info* info,
                              hipStream_t s, Inner&& inner) {
  auto t_start = std::chrono::steady_clock::now();
  MSPCHK(ensure_workspace(A, s));
  const size_t np = A->ws.npad;
  if (!A->w64.p) MSPCHK(A->w64.alloc(sizeof(double) * (3 * np + 2)));
  double* X64 = A->w64.as<double>();
  double* R64 = X64 + np;
  double* B64 = X64 + 2 * np;
  double* SUMS = X64 + 3 * np;
  const uint32_t n = A->n, gv = vec_grid(n);
  HIPCHK(hipMemcpyAsync(B64, b, sizeof(double) * n, is_device_ptr_d(b) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                        s));
  if (ro.use_x0)
    HIPCHK(hipMemcpyAsync(X64, x, sizeof(double) * n,
                          is_device_ptr_d(x) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
  else
    k_fill64<<<gv, kBlock, 0, s>>>(X64, 0.0, n);
  // the scaled residual goes straight into the solvers' b vector (their
  // copy of b onto itself is skipped)
  float* RHS = A->ws.v(Workspace::B);
  float* D = A->ws.v(Workspace::X2);
  GridRed rb = A->ws.gr();
  const RowsCsr rows{A->row_ptr.as<uint32_t>(), A->col.as<uint32_t>(), A->val.as<float>()};
  msp_refine_info inf{};
  inf.status = MSP_SOLVE_MAXITER;
  double prev = 0.0;
  int slow = 0;
  for (int k = 0;; ++k) {
    double* part = reinterpret_cast<double*>(rb.partials);
    const uint64_t avg = n ? A->nnz / n : 0;
    if (avg >= 64) k_residual64_lanes<64><<<gv, kBlock, 0, s>>>(rows, X64, B64, R64, n, part);
    else if (avg >= 16) k_residual64_lanes<16><<<gv, kBlock, 0, s>>>(rows, X64, B64, R64, n, part);
    else k_residual64<<<gv, kBlock, 0, s>>>(rows, X64, B64, R64, n, part);
    k_finish_d<2><<<1, kBlock, 0, s>>>(reinterpret_cast<const double*>(rb.partials), gv, SUMS);
    MSPCHK(launch_check("k_residual64"));
    double* sums = A->ws.hsum;
    HIPCHK(hipMemcpyAsync(sums, SUMS, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(spin_sync(s));
    const double rnorm = std::sqrt(sums[0]), bnorm = std::sqrt(sums[1]);
    inf.relres = bnorm > 0 ? rnorm / bnorm : rnorm;
    inf.outer_iterations = k;
    if (!std::isfinite(inf.relres)) {
      inf.status = MSP_SOLVE_BREAKDOWN;
      break;
    }
    if (inf.relres <= ro.tol || rnorm == 0.0) {
      inf.status = MSP_SOLVE_CONVERGED;
      break;
    }
    if (k > 0 && inf.relres > 0.5 * prev) {
      if (++slow >= 2) {
        inf.status = MSP_SOLVE_STAGNATED;
        break;
      }
    } else {
      slow = 0;
    }
    if (k >= ro.max_outer) break;
    prev = inf.relres;
    k_scale_to_f32<<<gv, kBlock, 0, s>>>(R64, RHS, 1.0 / rnorm, n);
    bool stop = false;
    const double itol = std::max(ro.inner_tol, std::min(0.5, 0.3 * ro.tol / inf.relres));
    MSPCHK(inner(RHS, D, itol, &inf, &stop));
    if (stop) {
      inf.status = MSP_SOLVE_BREAKDOWN;
      break;
    }
    k_axpy64<<<gv, kBlock, 0, s>>>(X64, D, rnorm, n);
    MSPCHK(launch_check("k_axpy64"));
  }
  HIPCHK(hipMemcpyAsync(x, X64, sizeof(double) * n, is_device_ptr_d(x) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                        s));
  HIPCHK(spin_sync(s));
  inf.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  if (info) *info = inf;
  return MSP_OK;
}

// inner_tol <= 0: 1e-3, or the unit roundoff of the matrix values inside the
// correction if larger (BF16 2^-8; FP16 2^-11 and FP32 fall back to 1e-3): a
// correction cannot be more accurate than its operator's rounding, so solving
// further is wasted steps (128^3 learned-FSAI refinement with BF16 values: 355
// inner steps at 1e-3, 336 at 4e-3, same outer count and final accuracy).
static double resolved_inner_tol(const msp_refine_opts& ro) {
  if (ro.inner_tol > 0.0) return ro.inner_tol;
  const double u = ro.values == MSP_VALUES_BF16 ? std::ldexp(1.0, -8)
                   : ro.values == MSP_VALUES_F16 ? std::ldexp(1.0, -11)
                                                 : std::ldexp(1.0, -24);
  return std::max(1e-3, u);
}

static msp_solve_opts inner_opts(const msp_refine_opts& ro) {
  msp_solve_opts io;
  msp_default_opts(&io);
  io.tol = ro.inner_tol;
  io.maxiter = ro.inner_maxiter;
  io.precond = ro.precond;
  io.verify = 0;
  io.values = ro.values;
  return io;
}

extern "C" msp_status msp_csr_solve_refined(msp_csr A, const double* b, double* x, const msp_refine_opts* opts,
                                            msp_refine_info* info, void* stream) {
  MSPCHK(check_handle(A));
  msp_refine_opts ro;
  msp_default_refine_opts(&ro);
  if (opts) ro = *opts;
  if (ro.inner_tol < 0.0) return fail(MSP_ERR_INVALID, "negative inner_tol");
  ro.inner_tol = resolved_inner_tol(ro);
  if (A->n != A->m) return fail(MSP_ERR_INVALID, "refinement needs a square matrix");
  if ((!b || !x) && A->n) return fail(MSP_ERR_INVALID, "null b / x");
  if (ro.inner_solver != 0 && ro.inner_solver != 1) return fail(MSP_ERR_INVALID, "inner_solver must be 0 or 1");
  hipStream_t s = (hipStream_t)stream;
  if (ro.values != MSP_VALUES_F32) MSPCHK(msp_csr_prepare_values(A, ro.values, stream));
  msp_solve_opts io = inner_opts(ro);
  auto inner = [&](const float* rhs, float* d, double itol, msp_refine_info* inf, bool* stop) -> msp_status {
    msp_solve_info ii{};
    io.tol = itol;
    auto run = [&](int32_t vt) -> msp_status {
      io.values = vt;
      return with_csr_op(A, vt, s, [&](auto op) {
        return ro.inner_solver == 0 ? cg_impl(op, rhs, d, io, &ii, s, false)
                                    : bicgstab_impl(op, rhs, d, io, &ii, s, false);
      });
    };
    MSPCHK(run(ro.values));
    inf->inner_iterations += ii.iterations;
    if (ii.status == MSP_SOLVE_BREAKDOWN && ro.values != MSP_VALUES_F32) {
      inf->fallbacks += 1;
      MSPCHK(run(MSP_VALUES_F32));
      inf->inner_iterations += ii.iterations;
    }
    *stop = ii.status == MSP_SOLVE_BREAKDOWN && ii.iterations == 0;
    return MSP_OK;
  };
  return refine_impl(A, b, x, ro, info, s, inner);
}

// ---------------------------------------------------------------------------
// ABI: FSAI preconditioner M = G^T G (G on the lower-triangle pattern of A)
// A pinned host word for a small device-to-host read that the host waits
// for (a copy into pageable memory is staged through a runtime buffer).
struct PinnedU32 {
  uint32_t* p = nullptr;
  PinnedU32() = default;
  PinnedU32(const PinnedU32&) = delete;
  PinnedU32& operator=(const PinnedU32&) = delete;
  ~PinnedU32() {
    if (p && !in_teardown()) (void)hipHostFree(p);
  }
  // the device word's value, after the work queued on s
  msp_status read(const void* dev, hipStream_t s, uint32_t& out) {
    if (!p) HIPCHK(hipHostMalloc((void**)&p, sizeof(uint32_t), hipHostMallocDefault));
    HIPCHK(hipMemcpyAsync(p, dev, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(spin_sync(s));
    out = *p;
    return MSP_OK;
  }
};

struct msp_fsai_s {
  int device = 0;
  msp_csr A = nullptr;  // borrowed
  msp_csr_s G, GT;      // G and its transpose as CSR matrices
  uint64_t ne = 0;
  DevMem Lrow, Lsrc, GTmap, feat, colscale, diag, vec, sums;
  DevMem rowscale;  // stated-5 network: the row normalisation s_i of G
  int features = kFeatCoupled8;  // feature set of the edge network (FsaiFeatures)
  int ready = 0;    // values set
  int feat_for = -1;  // the feature set whose features (and column scaling) feat holds
  int nonpos = 0;   // diagonal entries of G that are not positive
  PinnedU32 hword;  // host side of the set-up's counts
};

static msp_status fsai_compute_features(msp_fsai F, hipStream_t s) {
  msp_csr A = F->A;
  const uint32_t n = A->n;
  if (F->features == kFeatStated5)
    k_fsai_features<kFeatStated5><<<grid_for(n), kBlock, 0, s>>>(
        A->row_ptr.as<uint32_t>(), A->col.as<uint32_t>(), A->val.as<float>(), n, F->G.row_ptr.as<uint32_t>(),
        F->G.col.as<uint32_t>(), F->Lsrc.as<uint32_t>(), F->diag.as<float>(), F->feat.as<float>(),
        F->colscale.as<float>());
  else
    k_fsai_features<kFeatCoupled8><<<grid_for(n), kBlock, 0, s>>>(
        A->row_ptr.as<uint32_t>(), A->col.as<uint32_t>(), A->val.as<float>(), n, F->G.row_ptr.as<uint32_t>(),
        F->G.col.as<uint32_t>(), F->Lsrc.as<uint32_t>(), F->diag.as<float>(), F->feat.as<float>(),
        F->colscale.as<float>());
  return launch_check("features");
}

// The selected set's features and column scaling, computed on first use (a
// handle that selects another set before using any costs one pass, not two).
static msp_status fsai_features_ready(msp_fsai F, hipStream_t s) {
  if (F->feat_for == F->features) return MSP_OK;
  MSPCHK(fsai_compute_features(F, s));
  F->feat_for = F->features;
  return MSP_OK;
}

static msp_status exclusive_scan_u32(uint32_t* data, uint64_t N, hipStream_t s) {
  const uint32_t nscan = (uint32_t)((N + sortk::kScanTile - 1) / sortk::kScanTile);
  DevMem sums;
  MSPCHK(sums.alloc(sizeof(uint32_t) * std::max<uint32_t>(1, nscan)));
  sortk::k_scan_reduce<<<nscan, kBlock, 0, s>>>(data, N, sums.as<uint32_t>());
  sortk::k_scan_top<<<1, kBlock, 0, s>>>(sums.as<uint32_t>(), nscan);
  sortk::k_scan_down<<<nscan, kBlock, 0, s>>>(data, N, sums.as<uint32_t>());
  MSPCHK(launch_check("scan"));
  HIPCHK(hipStreamSynchronize(s));
  return MSP_OK;
}


extern "C" msp_status msp_fsai_create(msp_fsai* out, msp_csr A, void* stream) {
  if (!out) return fail(MSP_ERR_INVALID, "null output handle");
  *out = nullptr;
  MSPCHK(check_handle(A));
  if (A->n != A->m) return fail(MSP_ERR_INVALID, "FSAI needs a square matrix");
  hipStream_t s = (hipStream_t)stream;
  const uint32_t n = A->n;
  msp_fsai F = new msp_fsai_s;
  F->device = A->device;
  F->A = A;
  auto body = [&]() -> msp_status {
    // lower pattern of A -> G's CSR
    F->G.device = F->GT.device = A->device;
    F->G.n = F->G.m = F->GT.n = F->GT.m = n;
    MSPCHK(F->G.row_ptr.alloc(sizeof(uint32_t) * ((size_t)n + 1)));
    k_lower_count<<<grid_for((uint64_t)n + 1), kBlock, 0, s>>>(A->row_ptr.as<uint32_t>(), A->col.as<uint32_t>(), n,
                                                               F->G.row_ptr.as<uint32_t>());
    MSPCHK(launch_check("lower count"));
    MSPCHK(exclusive_scan_u32(F->G.row_ptr.as<uint32_t>(), (uint64_t)n + 1, s));
    uint32_t ne32 = 0;
    HIPCHK(hipMemcpy(&ne32, F->G.row_ptr.as<uint32_t>() + n, sizeof(uint32_t), hipMemcpyDeviceToHost));
    F->ne = F->G.nnz = F->GT.nnz = ne32;
    const uint64_t ne = F->ne;
    MSPCHK(F->G.col.alloc(sizeof(uint32_t) * (ne + 64)));
    MSPCHK(F->G.val.alloc(sizeof(float) * (ne + 64)));
    MSPCHK(F->Lrow.alloc(sizeof(uint32_t) * (ne + 1)));
    MSPCHK(F->Lsrc.alloc(sizeof(uint32_t) * (ne + 1)));
    k_lower_fill<<<grid_for(n), kBlock, 0, s>>>(A->row_ptr.as<uint32_t>(), A->col.as<uint32_t>(), n,
                                                F->G.row_ptr.as<uint32_t>(), F->G.col.as<uint32_t>(),
                                                F->Lrow.as<uint32_t>(), F->Lsrc.as<uint32_t>());
    MSPCHK(launch_check("lower fill"));
    HIPCHK(hipMemsetAsync(F->G.val.p, 0, sizeof(float) * (ne + 64), s));
    MSPCHK(zero_pad(&F->G, s));
    // G's rows are the lower parts of A's: A's longest row bounds them
    MSPCHK(finalize_matrix(&F->G, s, A->max_row));
    // G's values are set later (fsai_commit refills the diagonal storage)
    MSPCHK(ensure_dia(&F->G, s, false));
    if (F->G.dia_state == 1 && F->G.dia_o.off[F->G.dia_o.ndiag - 1] == 0) {
      // G^T is read from G's planes (G's last plane is its main diagonal, which
      // keeps every shifted read inside them): no transposed pattern, no copy
      dia_alias_transpose(&F->GT, &F->G);
      F->GT.max_row = (uint32_t)F->G.dia_o.ndiag;
    } else {
    // transpose pattern by sorting (col << 32 | row) with the edge index as payload
    TmpMem keys, payload;
    MSPCHK(keys.alloc(sizeof(uint64_t) * (ne + 1), s));
    MSPCHK(payload.alloc(sizeof(float) * (ne + 1), s));
    k_transpose_keys<<<grid_for(ne), kBlock, 0, s>>>(F->Lrow.as<uint32_t>(), F->G.col.as<uint32_t>(), ne,
                                                     keys.as<uint64_t>(), payload.as<float>());
    MSPCHK(launch_check("transpose keys"));
    MSPCHK(alloc_csr(&F->GT));
    MSPCHK(F->GTmap.alloc(sizeof(float) * (ne + 64)));
    // (G's edges are listed row by row with ascending columns, so each row of
    // G^T receives its entries in ascending column order: sort by row only)
    MSPCHK(coo_to_csr_device(keys.as<uint64_t>(), payload.as<float>(), ne, n, n, F->GT.row_ptr.as<uint32_t>(),
                             F->GT.col.as<uint32_t>(), F->GTmap.as<float>(), s, nullptr, true));
    HIPCHK(hipMemsetAsync(F->GT.val.p, 0, sizeof(float) * (ne + 64), s));
    MSPCHK(zero_pad(&F->GT, s));
    // with G in diagonal storage, G^T's rows hold at most its diagonal count
    // and its storage follows from G's
    MSPCHK(finalize_matrix(&F->GT, s));
    MSPCHK(ensure_dia(&F->GT, s, false));
    }
    // features and column scaling
    MSPCHK(F->diag.alloc(sizeof(float) * ((size_t)n + 1)));
    MSPCHK(F->feat.alloc(sizeof(float) * 8 * (ne + 64)));
    MSPCHK(F->colscale.alloc(sizeof(float) * (ne + 64)));
    // k_fsai_features writes all 8 features of every edge: zero the padding only
    HIPCHK(hipMemsetAsync(F->feat.as<float>() + 8 * ne, 0, sizeof(float) * 8 * 64, s));
    k_diag<<<grid_for(n), kBlock, 0, s>>>(A->row_ptr.as<uint32_t>(), A->col.as<uint32_t>(), A->val.as<float>(), n,
                                          F->diag.as<float>());
    MSPCHK(ensure_workspace(A, s));
    MSPCHK(F->vec.alloc(sizeof(float) * 6 * A->ws.npad));
    HIPCHK(hipMemsetAsync(F->vec.p, 0, sizeof(float) * 6 * A->ws.npad, s));
    MSPCHK(F->sums.alloc(sizeof(float) * 4));
    HIPCHK(hipStreamSynchronize(s));
    return MSP_OK;
  };
  msp_status st = body();
  if (st != MSP_OK) {
    delete F;
    return st;
  }
  *out = F;
  return MSP_OK;
}

extern "C" msp_status msp_fsai_destroy(msp_fsai F) {
  if (!F) return MSP_OK;
  if (!in_teardown()) {
    (void)hipSetDevice(F->device);
    (void)hipDeviceSynchronize();
  }
  delete F;
  return MSP_OK;
}

extern "C" msp_status msp_fsai_num_edges(msp_fsai F, uint64_t* ne) {
  if (!F || !ne) return fail(MSP_ERR_INVALID, "null argument");
  *ne = F->ne;
  return MSP_OK;
}

extern "C" msp_status msp_fsai_edges(msp_fsai F, uint32_t* rows, uint32_t* cols, void* stream) {
  if (!F) return fail(MSP_ERR_INVALID, "null FSAI handle");
  HIPCHK(hipSetDevice(F->device));
  hipStream_t s = (hipStream_t)stream;
  MSPCHK(copy_out(rows, F->Lrow.p, sizeof(uint32_t) * F->ne, s));
  MSPCHK(copy_out(cols, F->G.col.p, sizeof(uint32_t) * F->ne, s));
  HIPCHK(hipStreamSynchronize(s));
  return MSP_OK;
}

extern "C" msp_status msp_fsai_features(msp_fsai F, float* feat, float* colscale, void* stream) {
  if (!F) return fail(MSP_ERR_INVALID, "null FSAI handle");
  HIPCHK(hipSetDevice(F->device));
  hipStream_t s = (hipStream_t)stream;
  MSPCHK(fsai_features_ready(F, s));
  MSPCHK(copy_out(feat, F->feat.p, sizeof(float) * 8 * F->ne, s));
  MSPCHK(copy_out(colscale, F->colscale.p, sizeof(float) * F->ne, s));
  HIPCHK(hipStreamSynchronize(s));
  return MSP_OK;
}

extern "C" msp_status msp_fsai_select_features(msp_fsai F, int32_t feature_set, void* stream) {
  if (!F) return fail(MSP_ERR_INVALID, "null FSAI handle");
  if (feature_set != MSP_FSAI_FEATURES_COUPLED8 && feature_set != MSP_FSAI_FEATURES_STATED5)
    return fail(MSP_ERR_INVALID, "unknown FSAI feature set %d", (int)feature_set);
  HIPCHK(hipSetDevice(F->device));
  hipStream_t s = (hipStream_t)stream;
  if (feature_set == F->features) return MSP_OK;
  F->features = feature_set;
  if (feature_set == MSP_FSAI_FEATURES_STATED5 && !F->rowscale.p)
    MSPCHK(F->rowscale.alloc(sizeof(float) * ((size_t)F->A->n + 1)));
  (void)s;  // the features follow on first use
  return MSP_OK;
}

extern "C" msp_status msp_fsai_get_features(msp_fsai F, int32_t* feature_set) {
  if (!F || !feature_set) return fail(MSP_ERR_INVALID, "null argument");
  *feature_set = F->features;
  return MSP_OK;
}

// Refresh G^T from G and count non-positive diagonal entries of G.
static msp_status fsai_commit(msp_fsai F, hipStream_t s) {
  if (!F->GT.dia_alias) {
    k_gather_by_payload<<<grid_for(F->ne), kBlock, 0, s>>>(F->GT.val.as<float>(), F->G.val.as<float>(),
                                                           F->GTmap.as<float>(), F->ne);
    MSPCHK(launch_check("gather G^T"));
  }
  MSPCHK(refresh_dia(&F->G, s));
  MSPCHK(refresh_dia(&F->GT, s));
  F->ready = 1;
  return MSP_OK;
}

extern "C" msp_status msp_fsai_set_values(msp_fsai F, const float* g, int32_t* nonpositive_diag, void* stream) {
  if (!F) return fail(MSP_ERR_INVALID, "null FSAI handle");
  HIPCHK(hipSetDevice(F->device));
  hipStream_t s = (hipStream_t)stream;
  MSPCHK(copy_in(F->G.val.p, g, sizeof(float) * F->ne, s));
  MSPCHK(fsai_features_ready(F, s));  // the column scaling
  // count non-positive diagonal entries with the scaling kernel (its scaled
  // output goes to a throw-away buffer; colscale > 0 keeps the signs)
  TmpMem cnt, scaled;
  MSPCHK(cnt.alloc(sizeof(uint32_t), s));
  HIPCHK(hipMemsetAsync(cnt.p, 0, sizeof(uint32_t), s));
  MSPCHK(scaled.alloc(sizeof(float) * (F->ne + 1), s));
  k_fsai_scale<<<grid_for(F->ne), kBlock, 0, s>>>(F->G.val.as<float>(), F->colscale.as<float>(),
                                                  F->Lrow.as<uint32_t>(), F->G.col.as<uint32_t>(), F->ne,
                                                  scaled.as<float>(), cnt.as<uint32_t>());
  MSPCHK(fsai_commit(F, s));
  uint32_t h = 0;
  MSPCHK(F->hword.read(cnt.p, s, h));
  F->nonpos = (int)h;
  if (nonpositive_diag) *nonpositive_diag = (int32_t)h;
  return MSP_OK;
}

extern "C" msp_status msp_fsai_set_network_output(msp_fsai F, const float* out, int32_t* nonpositive_diag,
                                                  void* stream) {
  if (!F) return fail(MSP_ERR_INVALID, "null FSAI handle");
  HIPCHK(hipSetDevice(F->device));
  hipStream_t s = (hipStream_t)stream;
  MSPCHK(fsai_features_ready(F, s));  // the column (and row) scaling
  InArr o;
  MSPCHK(o.set(out, sizeof(float) * F->ne, s));
  TmpMem cnt;
  MSPCHK(cnt.alloc(sizeof(uint32_t), s));
  HIPCHK(hipMemsetAsync(cnt.p, 0, sizeof(uint32_t), s));
  if (F->features == kFeatStated5) {
    msp_csr A = F->A;
    // A's diagonal storage (probed and filled here, else at its first product)
    // turns each A(J_p, J_q) of the normalisation into one load
    MSPCHK(ensure_dia(A, s));
    if (A->dia_state == 1)
      k_fsai_rownorm<true><<<grid_for(A->n), kBlock, 0, s>>>(
          A->row_ptr.as<uint32_t>(), A->col.as<uint32_t>(), A->val.as<float>(), A->n, F->G.row_ptr.as<uint32_t>(),
          F->G.col.as<uint32_t>(), static_cast<const float*>(o.d), F->colscale.as<float>(), F->G.val.as<float>(),
          F->rowscale.as<float>(), cnt.as<uint32_t>(), dia_view_of<ValF32>(A));
    else
      k_fsai_rownorm<false><<<grid_for(A->n), kBlock, 0, s>>>(
          A->row_ptr.as<uint32_t>(), A->col.as<uint32_t>(), A->val.as<float>(), A->n, F->G.row_ptr.as<uint32_t>(),
          F->G.col.as<uint32_t>(), static_cast<const float*>(o.d), F->colscale.as<float>(), F->G.val.as<float>(),
          F->rowscale.as<float>(), cnt.as<uint32_t>());
    MSPCHK(launch_check("fsai rownorm"));
  } else {
    k_fsai_scale<<<grid_for(F->ne), kBlock, 0, s>>>(static_cast<const float*>(o.d), F->colscale.as<float>(),
                                                    F->Lrow.as<uint32_t>(), F->G.col.as<uint32_t>(), F->ne,
                                                    F->G.val.as<float>(), cnt.as<uint32_t>());
    MSPCHK(launch_check("fsai scale"));
  }
  MSPCHK(fsai_commit(F, s));
  uint32_t h = 0;
  MSPCHK(F->hword.read(cnt.p, s, h));
  F->nonpos = (int)h;
  if (nonpositive_diag) *nonpositive_diag = (int32_t)h;
  return MSP_OK;
}

extern "C" msp_status msp_fsai_classical(msp_fsai F, int32_t* fallback_rows, void* stream) {
  if (!F) return fail(MSP_ERR_INVALID, "null FSAI handle");
  HIPCHK(hipSetDevice(F->device));
  hipStream_t s = (hipStream_t)stream;
  msp_csr A = F->A;
  TmpMem cnt;
  MSPCHK(cnt.alloc(sizeof(uint32_t), s));
  HIPCHK(hipMemsetAsync(cnt.p, 0, sizeof(uint32_t), s));
  k_fsai_classical<<<grid_for(A->n), kBlock, 0, s>>>(A->row_ptr.as<uint32_t>(), A->col.as<uint32_t>(),
                                                     A->val.as<float>(), A->n, F->G.row_ptr.as<uint32_t>(),
                                                     F->G.col.as<uint32_t>(), F->G.val.as<float>(),
                                                     cnt.as<uint32_t>());
  MSPCHK(launch_check("fsai classical"));
  MSPCHK(fsai_commit(F, s));
  uint32_t h = 0;
  MSPCHK(F->hword.read(cnt.p, s, h));
  F->nonpos = 0;
  if (fallback_rows) *fallback_rows = (int32_t)h;
  return MSP_OK;
}

extern "C" msp_status msp_fsai_export_values(msp_fsai F, float* g, void* stream) {
  if (!F) return fail(MSP_ERR_INVALID, "null FSAI handle");
  HIPCHK(hipSetDevice(F->device));
  MSPCHK(copy_out(g, F->G.val.p, sizeof(float) * F->ne, (hipStream_t)stream));
  HIPCHK(hipStreamSynchronize((hipStream_t)stream));
  return MSP_OK;
}

template <int ND, typename XL, typename Epi, typename Fin>
static uint32_t spmv_with(msp_csr M, XL xl, Epi epi, Fin fin, GridRed rb, const int* done, hipStream_t s) {
  return spmv_any<ValF32, ND>(M, xl, epi, fin, rb, done, s);
}

// A product with an FSAI factor (G or G^T), FP32 values or their one-by-one
// BF16 rounding (g16; diagonal storage only, prepared by ensure_dia16).
template <int ND, typename XL, typename Epi>
static uint32_t factor_spmv(bool g16, msp_csr M, XL xl, Epi epi, GridRed rb, const int* done, hipStream_t s) {
  if (g16) return spmv_any<ValBF16Plain, ND>(M, xl, epi, FinNone{}, rb, done, s);
  return spmv_any<ValF32, ND>(M, xl, epi, FinNone{}, rb, done, s);
}

extern "C" msp_status msp_fsai_apply(msp_fsai F, const float* r, float* z, void* stream) {
  if (!F || !F->ready) return fail(MSP_ERR_INVALID, "FSAI values not set");
  HIPCHK(hipSetDevice(F->device));
  hipStream_t s = (hipStream_t)stream;
  const size_t np = F->A->ws.npad;
  float* rin = F->vec.as<float>();
  float* w = rin + np;
  float* zo = rin + 2 * np;
  MSPCHK(copy_in(rin, r, sizeof(float) * F->A->n, s));
  spmv_with<0>(&F->G, XPlain{rin}, EpiStore{w}, FinNone{}, GridRed{nullptr}, nullptr, s);
  spmv_with<0>(&F->GT, XPlain{w}, EpiStore{zo}, FinNone{}, GridRed{nullptr}, nullptr, s);
  MSPCHK(launch_check("fsai apply"));
  MSPCHK(copy_out(z, zo, sizeof(float) * F->A->n, s));
  HIPCHK(hipStreamSynchronize(s));
  return MSP_OK;
}

// L = sum_k ||G A G^T w_k - w_k||^2 over k probes (rows of W, n floats each,
// device or host); adds dL/dG per edge into grad (ne floats, device).
extern "C" msp_status msp_fsai_probe_loss_grad(msp_fsai F, const float* W, int32_t k, float* grad, double* loss,
                                               void* stream) {
  if (!F || !F->ready) return fail(MSP_ERR_INVALID, "FSAI values not set");
  if (k < 0 || (k > 0 && (!W || !grad))) return fail(MSP_ERR_INVALID, "bad probes / gradient");
  HIPCHK(hipSetDevice(F->device));
  hipStream_t s = (hipStream_t)stream;
  msp_csr A = F->A;
  const uint32_t n = A->n;
  const size_t np = A->ws.npad;
  float* w = F->vec.as<float>();
  float* u = w + np;
  float* y = w + 2 * np;
  float* z = w + 3 * np;
  float* e = w + 4 * np;
  float* h = w + 5 * np;
  GridRed rb = A->ws.gr();
  double total = 0.0;
  const bool wdev = is_device_ptr(W);
  for (int32_t q = 0; q < k; ++q) {
    HIPCHK(hipMemcpyAsync(w, W + (size_t)q * n, sizeof(float) * n,
                          wdev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    spmv_with<0>(&F->GT, XPlain{w}, EpiStore{u}, FinNone{}, rb, nullptr, s);
    spmv_with<0>(A, XPlain{u}, EpiStore{y}, FinNone{}, rb, nullptr, s);
    spmv_with<0>(&F->G, XPlain{y}, EpiStore{z}, FinNone{}, rb, nullptr, s);
    k_probe_resid<<<vec_grid(n), kBlock, 0, s>>>(z, w, n, e, rb);
    finish<2>(FinStore{F->sums.as<float>()}, rb, vec_grid(n), s);
    spmv_with<0>(&F->GT, XPlain{e}, EpiStore{u}, FinNone{}, rb, nullptr, s);
    spmv_with<0>(A, XPlain{u}, EpiStore{h}, FinNone{}, rb, nullptr, s);
    k_fsai_edge_grad<<<grid_for(F->ne), kBlock, 0, s>>>(F->Lrow.as<uint32_t>(), F->G.col.as<uint32_t>(), F->ne, e, y,
                                                        w, h, 1.0f, grad);
    MSPCHK(launch_check("probe grad"));
    float sums[2];
    HIPCHK(hipMemcpyAsync(sums, F->sums.p, sizeof sums, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    total += sums[0];
  }
  if (loss) *loss = total;
  return MSP_OK;
}

// FSAI-preconditioned CG. Step order puts the preconditioner first, so that a
// restart after a failed true-residual check is just beta = 0:
//   z = G^T w, <r,z> = ||w||^2 -> beta; p = z + beta p; q = A p, <p,q> -> alpha;
//   x += alpha p, r -= alpha q, w = G r, ||r||^2 -> convergence, ||w||^2;
//   [true residual check].
template <typename V>
static msp_status pcg_fsai_impl(const OpCsr<V>& op, msp_fsai F, const float* b, float* x, const msp_solve_opts& o,
                                msp_solve_info* info, hipStream_t s,
                                bool sync_out = true) {
  MatBase* A = op.base();
  MSPCHK(solver_prologue(A, b, x, o, s));
  // With 16-bit values of A, G's values are BF16 too (each rounded on its
  // own, so the stored G^T is the exact transpose of the stored G and
  // M = G^T G stays symmetric positive definite); diagonal storage only.
  const bool g16 = o.values != MSP_VALUES_F32 && F->G.dia_state == 1 && F->GT.dia_state == 1;
  if (g16) {
    MSPCHK(ensure_dia16(&F->G, kValuesBF16Plain, s));
    MSPCHK(ensure_dia16(&F->GT, kValuesBF16Plain, s));
  }
  Workspace& w = A->ws;
  float *X = w.v(Workspace::X), *R = w.v(Workspace::R), *P = w.v(Workspace::P), *Q = w.v(Workspace::Q);
  float *Wv = w.v(Workspace::W), *XLo = w.v(Workspace::XL);
  const float* B = w.v(Workspace::B);
  Scalars* S = w.S_dev();
  Scalars* S2 = w.S_stage();
  GridRed rb = w.gr();
  const uint32_t n = A->n, gv = vec_grid(n);
  const bool comp = compensate(o);
  // ping-pong directions and residuals: step k reads Pb[k & 1], Rb[k & 1]
  // and writes Pb[(k + 1) & 1], Rb[(k + 1) & 1] (the residual is formed in a
  // gather, so the vector being written cannot be the one other workgroups
  // gather; the direction is written while the held update of x may still
  // need the previous one)
  float* Pb[2] = {P, w.v(Workspace::T)};
  float* Rb[2] = {R, w.v(Workspace::R0)};
  int k = 0;
  uint32_t gu = 0;  // workgroups of the latest x / r update (its decision pending on rb2)
  // workgroups whose <r,r>, <w,w> partials in rb2 describe the current w
  uint32_t nw = gv;
  GridRed rb2 = w.gr2();
  // w = G r of the current residual and its partials, both gated (skipped
  // when *gate != 0); the partials take the update's block count once one ran
  auto g_of_r = [&](const int* gate) {
    factor_spmv<0>(g16, &F->G, XPlain{Rb[k & 1]}, EpiStore{Wv}, rb, gate, s);
    k_pcg_wnorm<<<nw, kBlock, 0, s>>>(Rb[k & 1], Wv, n, rb2, gate);
  };
  auto init = [&]() -> msp_status {
    if (o.use_x0) op.template spmv<1>(X, EpiResidual{R, B}, FinNone{}, nullptr, s);
    k_cg_init<<<gv, kBlock, 0, s>>>(X, XLo, R, P, B, nullptr, n, o.use_x0, kPcExplicit, S, rb);
    finish<3>(FinCgInit{S}, rb, gv, s);
    k_set_restart<<<1, 1, 0, s>>>(S);
    g_of_r(&S->halt);
    return launch_check("pcg init");
  };
  // one step, three launches: z = G^T w with p = z + beta p_old formed in the
  // epilogue, the previous step's decision (pending on rb2) taken by every
  // workgroup and staged in S2 with <r,z> = ||w||^2; q = A p with <p,q>
  // partials (rb); the decision committed to S and the x / r update with
  // alpha = <r,z> / <p,q> fused with the next step's w = G r (r formed in the
  // gather, ||r||^2 and ||w||^2 partials on rb2, decision pending).
  auto step = [&]() {
    const float* pold = Pb[k & 1];
    float* pnew = Pb[(k + 1) & 1];
    const float* rcur = Rb[k & 1];
    float* rnext = Rb[(k + 1) & 1];
    factor_spmv<0>(g16, &F->GT, XPlain{Wv}, EpiPcgDir{pnew, pold, S, S2, rb2, nw, 0.0f}, rb, &S->halt, s);
    const uint32_t gq = op.template spmv<1>(pnew, EpiStoreDot{Q, pnew}, FinNone{}, &S2->halt, s);
    // (pold: the previous direction of a held update of x; pnew is buffer
    // (k + 1) & 1, recorded when this step holds its own)
    const int32_t pbuf = (k + 1) & 1;
    if (comp)
      gu = factor_spmv<2>(g16, &F->G, XRNew{rcur, Q, 0.0f},
                          EpiCgXrW<true>{X, XLo, rcur, rnext, pnew, Q, Wv, S, S2, rb, gq, 0.0f, 0, pold, pbuf, 0.0f,
                                         0, 0},
                          rb2, &S2->halt, s);
    else
      gu = factor_spmv<2>(g16, &F->G, XRNew{rcur, Q, 0.0f},
                          EpiCgXrW<false>{X, XLo, rcur, rnext, pnew, Q, Wv, S, S2, rb, gq, 0.0f, 0, pold, pbuf, 0.0f,
                                          0, 0},
                          rb2, &S2->halt, s);
    nw = gu;
    ++k;
  };
  // a held update of x applied before a true-residual check runs (ver) or
  // before the iterate is returned
  auto xflush = [&](bool ver) {
    if (ver) {
      if (comp) k_cg_xflush<true, true><<<gv, kBlock, 0, s>>>(X, XLo, Pb[0], Pb[1], n, S);
      else k_cg_xflush<false, true><<<gv, kBlock, 0, s>>>(X, XLo, Pb[0], Pb[1], n, S);
      k_cg_xflush_clear<true><<<1, 1, 0, s>>>(S);
    } else {
      if (comp) k_cg_xflush<true, false><<<gv, kBlock, 0, s>>>(X, XLo, Pb[0], Pb[1], n, S);
      else k_cg_xflush<false, false><<<gv, kBlock, 0, s>>>(X, XLo, Pb[0], Pb[1], n, S);
      k_cg_xflush_clear<false><<<1, 1, 0, s>>>(S);
    }
  };
  // the last step's decision, before the host reads the state
  auto flush = [&]() {
    if (gu) finish<2>(FinCgStepPending<kPcExplicit>{S}, rb2, gu, s);
  };
  // a verification rewrites the current residual; after a restart w = G r and
  // its partials are recomputed (gated by skip_g, which FinVerify clears only
  // then)
  auto verify = [&]() {
    xflush(true);
    const uint32_t gvf = verify_launch<3>(op, gv, X, XLo, B, Rb[k & 1], P, nullptr, nullptr, n, S, rb, s);
    finish<2>(FinVerify<3>{S}, rb, gvf, s);
    g_of_r(&S->skip_g);
  };
  auto tail = [&]() -> msp_status {
    xflush(false);
    k_collapse<<<gv, kBlock, 0, s>>>(X, XLo, n);
    return copy_out(x, X, sizeof(float) * n, s);
  };
  MSPCHK(run_solver(A, o, info, s, init, step, verify, flush, tail));
  // (a caller inside the library that keeps working on s skips the wait)
  if (sync_out) HIPCHK(hipStreamSynchronize(s));
  return MSP_OK;
}

extern "C" msp_status msp_csr_pcg_fsai(msp_csr A, msp_fsai F, const float* b, float* x, const msp_solve_opts* opts,
                                       msp_solve_info* info, void* stream) {
  MSPCHK(check_handle(A));
  if (!F || F->A != A) return fail(MSP_ERR_INVALID, "FSAI handle does not belong to this matrix");
  if (!F->ready) return fail(MSP_ERR_INVALID, "FSAI values not set");
  msp_solve_opts o;
  fill_opts(o, opts);
  hipStream_t s = (hipStream_t)stream;
  if (F->nonpos > 0) {
    // G has a non-positive diagonal: M = G^T G may be singular -> plain CG
    o.precond = MSP_PRECOND_NONE;
    msp_status st = with_csr_op(A, o.values, s, [&](auto op) { return cg_impl(op, b, x, o, info, s); });
    if (info) info->reserved = 1;
    return st;
  }
  msp_status st = with_csr_op(A, o.values, s, [&](auto op) { return pcg_fsai_impl(op, F, b, x, o, info, s); });
  if (info) info->reserved = 0;
  return st;
}

// Refinement whose corrections are FSAI-preconditioned CG solves. A correction
// that breaks down, or that ends without reaching inner_tol, is repeated with
// plain CG (counted in fallbacks); a G with a non-positive diagonal uses plain
// CG throughout.
extern "C" msp_status msp_csr_solve_refined_fsai(msp_csr A, msp_fsai F, const double* b, double* x,
                                                 const msp_refine_opts* opts, msp_refine_info* info, void* stream) {
  MSPCHK(check_handle(A));
  if (!F || F->A != A) return fail(MSP_ERR_INVALID, "FSAI handle does not belong to this matrix");
  if (!F->ready) return fail(MSP_ERR_INVALID, "FSAI values not set");
  msp_refine_opts ro;
  msp_default_refine_opts(&ro);
  if (opts) ro = *opts;
  if (ro.inner_tol < 0.0) return fail(MSP_ERR_INVALID, "negative inner_tol");
  ro.inner_tol = resolved_inner_tol(ro);
  if ((!b || !x) && A->n) return fail(MSP_ERR_INVALID, "null b / x");
  hipStream_t s = (hipStream_t)stream;
  if (ro.values != MSP_VALUES_F32) MSPCHK(msp_csr_prepare_values(A, ro.values, stream));
  msp_solve_opts io = inner_opts(ro);
  io.precond = MSP_PRECOND_NONE;
  auto inner = [&](const float* rhs, float* d, double itol, msp_refine_info* inf, bool* stop) -> msp_status {
    msp_solve_info ii{};
    io.tol = itol;
    const bool use_fsai = F->nonpos == 0;
    if (use_fsai)
      MSPCHK(with_csr_op(A, io.values, s, [&](auto op) { return pcg_fsai_impl(op, F, rhs, d, io, &ii, s, false); }));
    inf->inner_iterations += ii.iterations;
    const bool reached = ii.status == MSP_SOLVE_CONVERGED;
    if (!use_fsai || !reached) {
      if (use_fsai) inf->fallbacks += 1;
      MSPCHK(with_csr_op(A, io.values, s, [&](auto op) { return cg_impl(op, rhs, d, io, &ii, s, false); }));
      inf->inner_iterations += ii.iterations;
    }
    *stop = ii.status == MSP_SOLVE_BREAKDOWN && ii.iterations == 0;
    return MSP_OK;
  };
  return refine_impl(A, b, x, ro, info, s, inner);
}

// ---------------------------------------------------------------------------
// ABI: edge MLP of the learned FSAI (BF16 MFMA, FP32 accumulate)
extern "C" int32_t msp_mlp_num_params(void) { return mlp::kParams; }

static uint32_t mlp_grid(uint64_t ne, uint32_t per_cu = 2) {
  const uint64_t tiles = (ne + 31) / 32;
  const uint64_t g = (tiles + kWavesPerBlock - 1) / kWavesPerBlock;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(g, (uint64_t)cu_count() * per_cu));
}

static msp_status need_dev(const void* p, const char* what) {
  if (!p || !is_device_ptr(p)) return fail(MSP_ERR_INVALID, "%s must be device memory", what);
  return MSP_OK;
}

extern "C" msp_status msp_mlp_forward(const float* params, const float* feat, uint64_t ne, float* out, void* stream) {
  MSPCHK(need_device());
  MSPCHK(need_dev(params, "params"));
  if (ne == 0) return MSP_OK;
  MSPCHK(need_dev(feat, "feat"));
  MSPCHK(need_dev(out, "out"));
  if ((reinterpret_cast<uintptr_t>(feat) & 15u) != 0) return fail(MSP_ERR_INVALID, "feat must be 16-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  static const uint32_t fwd_per_cu = resident_per_cu(mlp::k_mlp_forward);
  mlp::k_mlp_forward<<<mlp_grid(ne, fwd_per_cu), kBlock, 0, s>>>(params, feat, ne, out);
  return launch_check("mlp forward");
}

extern "C" msp_status msp_mlp_backward(const float* params, const float* feat, const float* grad_out, uint64_t ne,
                                       float* grad_params, void* stream) {
  MSPCHK(need_device());
  MSPCHK(need_dev(params, "params"));
  MSPCHK(need_dev(grad_params, "grad_params"));
  hipStream_t s = (hipStream_t)stream;
  if (ne == 0) {
    HIPCHK(hipMemsetAsync(grad_params, 0, sizeof(float) * mlp::kParams, s));
    return MSP_OK;
  }
  MSPCHK(need_dev(feat, "feat"));
  MSPCHK(need_dev(grad_out, "grad_out"));
  if ((reinterpret_cast<uintptr_t>(feat) & 15u) != 0) return fail(MSP_ERR_INVALID, "feat must be 16-byte aligned");
  const uint32_t g = mlp_grid(ne);
  DevMem partial;
  MSPCHK(partial.alloc(sizeof(float) * (size_t)g * mlp::kParams));
  mlp::k_mlp_backward<<<g, kBlock, 0, s>>>(params, feat, grad_out, ne, partial.as<float>());
  mlp::k_mlp_reduce<<<(mlp::kParams + kBlock - 1) / kBlock, kBlock, 0, s>>>(partial.as<float>(), g, grad_params);
  MSPCHK(launch_check("mlp backward"));
  HIPCHK(hipStreamSynchronize(s));
  return MSP_OK;
}

extern "C" msp_status msp_mlp_adam(float* params, const float* grad, float* m, float* v, float lr, float beta1,
                                   float beta2, float eps, int32_t step, float weight_decay, void* stream) {
  MSPCHK(need_device());
  MSPCHK(need_dev(params, "params"));
  MSPCHK(need_dev(grad, "grad"));
  MSPCHK(need_dev(m, "m"));
  MSPCHK(need_dev(v, "v"));
  if (step < 1) return fail(MSP_ERR_INVALID, "adam step counts from 1");
  const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
  mlp::k_adam<<<(mlp::kParams + kBlock - 1) / kBlock, kBlock, 0, (hipStream_t)stream>>>(params, grad, m, v, lr, beta1,
                                                                                          beta2, eps, c1, c2,
                                                                                          weight_decay);
  return launch_check("adam");
}

// G = network(features of F) scaled by 1/sqrt(a_jj); counts non-positive diagonal.
extern "C" msp_status msp_fsai_apply_network(msp_fsai F, const float* params, int32_t* nonpositive_diag,
                                             void* stream) {
  if (!F) return fail(MSP_ERR_INVALID, "null FSAI handle");
  HIPCHK(hipSetDevice(F->device));
  TmpMem out;
  MSPCHK(fsai_features_ready(F, (hipStream_t)stream));
  MSPCHK(out.alloc(sizeof(float) * (F->ne + 1), (hipStream_t)stream));
  MSPCHK(msp_mlp_forward(params, F->feat.as<float>(), F->ne, out.as<float>(), stream));
  return msp_fsai_set_network_output(F, out.as<float>(), nonpositive_diag, stream);
}

// One training evaluation on one matrix: forward, G from the output, probe loss
// and dL/dG, chain rule through the column scaling, backward to the parameters.
extern "C" msp_status msp_fsai_train_step(msp_fsai F, const float* params, const float* W, int32_t k,
                                          float* grad_params, double* loss, void* stream) {
  if (!F) return fail(MSP_ERR_INVALID, "null FSAI handle");
  HIPCHK(hipSetDevice(F->device));
  hipStream_t s = (hipStream_t)stream;
  int32_t nonpos = 0;
  MSPCHK(msp_fsai_apply_network(F, params, &nonpos, stream));
  DevMem gg, go;
  MSPCHK(gg.alloc(sizeof(float) * (F->ne + 1)));
  MSPCHK(go.alloc(sizeof(float) * (F->ne + 1)));
  HIPCHK(hipMemsetAsync(gg.p, 0, sizeof(float) * (F->ne + 1), s));
  MSPCHK(msp_fsai_probe_loss_grad(F, W, k, gg.as<float>(), loss, stream));
  if (F->features == kFeatStated5) {
    msp_csr A = F->A;
    k_fsai_rownorm_grad<<<grid_for(A->n), kBlock, 0, s>>>(A->row_ptr.as<uint32_t>(), A->col.as<uint32_t>(),
                                                          A->val.as<float>(), A->n, F->G.row_ptr.as<uint32_t>(),
                                                          F->G.col.as<uint32_t>(), F->G.val.as<float>(),
                                                          F->rowscale.as<float>(), F->colscale.as<float>(),
                                                          gg.as<float>(), go.as<float>());
  } else {
    mlp::k_scale_grad<<<grid_for(F->ne), kBlock, 0, s>>>(gg.as<float>(), F->colscale.as<float>(), F->ne,
                                                          go.as<float>());
  }
  MSPCHK(launch_check("scale grad"));
  return msp_mlp_backward(params, F->feat.as<float>(), go.as<float>(), F->ne, grad_params, stream);
}

// ---------------------------------------------------------------------------
// ABI: batched Jacobi-PCG over the diagonal blocks of one CSR matrix
using batch::PairState;

struct msp_batch_s {
  int device = 0;
  msp_csr A = nullptr;  // borrowed
  uint32_t nsys = 0, n = 0;
  std::vector<uint32_t> h_sys_off;
  DevMem sys_off, dinv;
  // tiling (chosen by the tuner)
  uint32_t T = 0, ntiles = 0;
  int rpt = 1;
  DevMem tile_r0, tile_sys, sys_tile, active, n_active;
  DevMem partials, state, sys_active;
  DevMem vec;  // X, XL, R, P0, P1, Q, B: n * 4 floats each
  // batched diagonal storage (every system on at most batch::kBD diagonals)
  DevMem bd_planes, bd_offs, bd_nds;
  int bd_nd = 0;  // 0: CSR product; else the planes stored (5 when every system has exactly 5)
  bool bd_mid = false;  // 5 planes and every system's middle offset is 0 (the main diagonal)
  bool bd_mirror = false;  // bd_mid and every system exactly symmetric: lower diagonals read the upper planes
  int32_t* h_active = nullptr;  // pinned, 2 * nsys
  hipStream_t cs = nullptr;     // capture stream
  hipEvent_t ev[2] = {nullptr, nullptr}, join = nullptr, t0 = nullptr, t1 = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_steps = 0, graph_rpt = 0, graph_maxiter = -1;
  bool tuned = false;
  int tuned_for[2] = {-1, -1};  // (tile_rows, rows_per_thread) requested when tuned
  ~msp_batch_s() {
    if (in_teardown()) return;
    if (exec) (void)hipGraphExecDestroy(exec);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    if (join) (void)hipEventDestroy(join);
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    if (cs) (void)hipStreamDestroy(cs);
    if (h_active) (void)hipHostFree(h_active);
  }
  float* v(int i) const { return vec.as<float>() + (size_t)i * 4 * n; }
  batch::View view() const {
    batch::View V;
    V.row_ptr = A->row_ptr.as<uint32_t>();
    V.col = A->col.as<uint32_t>();
    V.val = A->val.as<float>();
    V.dinv = dinv.as<float>();
    V.tile_r0 = tile_r0.as<uint32_t>();
    V.tile_sys = tile_sys.as<uint32_t>();
    V.active = active.as<uint32_t>();
    V.n_active = n_active.as<uint32_t>();
    V.sys_active = sys_active.as<int32_t>();
    V.planes = bd_nd ? bd_planes.as<float>() : nullptr;
    V.ld = n;
    V.offs = bd_offs.as<int32_t>();
    V.sys_off = sys_off.as<uint32_t>();
    return V;
  }
};

enum { kBX = 0, kBXL, kBR, kBP0, kBP1, kBQ, kBB, kBVecs };

// Tiles of at most T rows inside each system.
static msp_status batch_tile(msp_batch B, uint32_t T, hipStream_t s) {
  std::vector<uint32_t> r0, sy, st(B->nsys + 1, 0);
  for (uint32_t q = 0; q < B->nsys; ++q) {
    st[q] = (uint32_t)r0.size();
    for (uint32_t i = B->h_sys_off[q]; i < B->h_sys_off[q + 1]; i += T) {
      r0.push_back(i);
      sy.push_back(q);
    }
  }
  st[B->nsys] = (uint32_t)r0.size();
  const uint32_t nt = (uint32_t)r0.size();
  r0.push_back(B->n);
  sy.push_back(0);
  MSPCHK(B->tile_r0.alloc(sizeof(uint32_t) * r0.size()));
  MSPCHK(B->tile_sys.alloc(sizeof(uint32_t) * sy.size()));
  MSPCHK(B->sys_tile.alloc(sizeof(uint32_t) * st.size()));
  MSPCHK(B->active.alloc(sizeof(uint32_t) * (nt + 1)));
  MSPCHK(B->partials.alloc(sizeof(CSum) * 12 * ((size_t)nt + 1)));
  HIPCHK(hipMemcpyAsync(B->tile_r0.p, r0.data(), sizeof(uint32_t) * r0.size(), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(B->tile_sys.p, sy.data(), sizeof(uint32_t) * sy.size(), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(B->sys_tile.p, st.data(), sizeof(uint32_t) * st.size(), hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));  // the host vectors go out of scope
  B->T = T;
  B->ntiles = nt;
  if (B->exec) {
    (void)hipGraphExecDestroy(B->exec);
    B->exec = nullptr;
  }
  return MSP_OK;
}

static uint32_t batch_grid(const msp_batch B) {
  return std::max<uint32_t>(1, std::min<uint32_t>(B->ntiles, (uint32_t)cu_count() * 8));
}
static uint32_t sys_grid(const msp_batch B) {
  return std::max<uint32_t>(1, (B->nsys + kWavesPerBlock - 1) / kWavesPerBlock);
}

static void launch_spmm(msp_batch B, int rpt, int parity, hipStream_t s) {
  const batch::View V = B->view();
  const uint32_t g = batch_grid(B);
  const float* pold = B->v(parity ? kBP1 : kBP0);
  float* pnew = B->v(parity ? kBP0 : kBP1);
  PairState* ps = B->state.as<PairState>();
  CSum* pt = B->partials.as<CSum>();
  if (B->bd_nd == 5 && B->bd_mirror) {  // symmetric 2D 5-point systems
    if (rpt == 2)
      batch::k_bt_spmm_dia<2, 5, true, true><<<g, kBlock, 0, s>>>(V, B->v(kBR), pold, pnew, B->v(kBQ), ps, pt);
    else
      batch::k_bt_spmm_dia<1, 5, true, true><<<g, kBlock, 0, s>>>(V, B->v(kBR), pold, pnew, B->v(kBQ), ps, pt);
    return;
  }
  if (B->bd_nd == 5 && B->bd_mid) {  // 2D 5-point systems
    if (rpt == 2) batch::k_bt_spmm_dia<2, 5, true><<<g, kBlock, 0, s>>>(V, B->v(kBR), pold, pnew, B->v(kBQ), ps, pt);
    else batch::k_bt_spmm_dia<1, 5, true><<<g, kBlock, 0, s>>>(V, B->v(kBR), pold, pnew, B->v(kBQ), ps, pt);
    return;
  }
  if (B->bd_nd == 5) {
    if (rpt == 2) batch::k_bt_spmm_dia<2, 5><<<g, kBlock, 0, s>>>(V, B->v(kBR), pold, pnew, B->v(kBQ), ps, pt);
    else batch::k_bt_spmm_dia<1, 5><<<g, kBlock, 0, s>>>(V, B->v(kBR), pold, pnew, B->v(kBQ), ps, pt);
    return;
  }
  if (B->bd_nd > 0) {
    if (rpt == 2) batch::k_bt_spmm_dia<2, 0><<<g, kBlock, 0, s>>>(V, B->v(kBR), pold, pnew, B->v(kBQ), ps, pt);
    else batch::k_bt_spmm_dia<1, 0><<<g, kBlock, 0, s>>>(V, B->v(kBR), pold, pnew, B->v(kBQ), ps, pt);
    return;
  }
  if (rpt == 2)
    batch::k_bt_spmm<2><<<g, kBlock, 0, s>>>(V, B->v(kBR), pold, pnew, B->v(kBQ), B->state.as<PairState>(),
                                             B->partials.as<CSum>());
  else
    batch::k_bt_spmm<1><<<g, kBlock, 0, s>>>(V, B->v(kBR), pold, pnew, B->v(kBQ), B->state.as<PairState>(),
                                             B->partials.as<CSum>());
}

// One PCG iteration (4 kernels); `parity` selects the direction buffers, and
// odd iterations also update the iterate (graphs hold an even number).
static void batch_iteration(msp_batch B, int rpt, int parity, int32_t maxiter, hipStream_t s) {
  const batch::View V = B->view();
  const uint32_t g = batch_grid(B), gs = sys_grid(B);
  launch_spmm(B, rpt, parity, s);
  batch::k_bt_fin_alpha<<<gs, kBlock, 0, s>>>(B->partials.as<CSum>(), B->sys_tile.as<uint32_t>(), B->nsys,
                                               B->state.as<PairState>(), B->sys_active.as<int32_t>(), parity ? 0 : 1);
  // the even iteration's direction is P1, the odd one's P0
  if (parity)
    batch::k_bt_update<true><<<g, kBlock, 0, s>>>(V, B->v(kBX), B->v(kBXL), B->v(kBR), B->v(kBP1), B->v(kBP0),
                                                  B->v(kBQ), B->state.as<PairState>(), B->partials.as<CSum>());
  else
    batch::k_bt_update<false><<<g, kBlock, 0, s>>>(V, B->v(kBX), B->v(kBXL), B->v(kBR), B->v(kBP0), B->v(kBP1),
                                                   B->v(kBQ), B->state.as<PairState>(), B->partials.as<CSum>());
  batch::k_bt_fin_beta<<<gs, kBlock, 0, s>>>(B->partials.as<CSum>(), B->sys_tile.as<uint32_t>(), B->nsys, maxiter,
                                              B->state.as<PairState>(), B->sys_active.as<int32_t>(), parity);
}

static void batch_active_list(msp_batch B, hipStream_t s) {
  batch::k_bt_active_list<<<1, kBlock, 0, s>>>(B->sys_tile.as<uint32_t>(), B->nsys, B->sys_active.as<int32_t>(),
                                               B->active.as<uint32_t>(), B->n_active.as<uint32_t>());
}

// hipGraph of (active list + `steps` iterations), captured once per
// configuration on the handle's own stream.
static msp_status batch_graph(msp_batch B, int steps, int32_t maxiter) {
  if (B->exec && B->graph_steps == steps && B->graph_rpt == B->rpt && B->graph_maxiter == maxiter) return MSP_OK;
  if (B->exec) {
    (void)hipGraphExecDestroy(B->exec);
    B->exec = nullptr;
  }
  hipGraph_t graph = nullptr;
  HIPCHK(hipStreamBeginCapture(B->cs, hipStreamCaptureModeThreadLocal));
  batch_active_list(B, B->cs);
  for (int k = 0; k < steps; ++k) batch_iteration(B, B->rpt, k & 1, maxiter, B->cs);
  hipError_t e = hipStreamEndCapture(B->cs, &graph);
  if (e != hipSuccess) return fail(MSP_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(&B->exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    B->exec = nullptr;
    return fail(MSP_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
  }
  B->graph_steps = steps;
  B->graph_rpt = B->rpt;
  B->graph_maxiter = maxiter;
  return MSP_OK;
}

extern "C" void msp_default_batch_opts(msp_batch_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof *o);
  o->tol = 1e-6;
  o->maxiter = 10000;
  o->nrhs = 4;
  o->steps_per_graph = 16;
  o->tile_rows = 0;
  o->rows_per_thread = 0;
  o->max_verify = 3;
}

// Batched diagonal storage: every system's distinct offsets (at most
// batch::kBD), then the value planes. Left off (CSR product) when a system
// has more offsets or the batch opts out (B->bd_nd stays 0).
static msp_status batch_diagonal_storage(msp_batch B, hipStream_t s) {
  msp_csr A = B->A;
  const uint32_t nsys = B->nsys;
  if (nsys == 0 || B->n == 0) return MSP_OK;
  MSPCHK(B->bd_offs.alloc(sizeof(int32_t) * batch::kBD * (size_t)nsys));
  MSPCHK(B->bd_nds.alloc(sizeof(int32_t) * (size_t)nsys));
  TmpMem bad;
  MSPCHK(bad.alloc(sizeof(uint32_t), s));
  HIPCHK(hipMemsetAsync(bad.p, 0, sizeof(uint32_t), s));
  const uint32_t g = std::min<uint32_t>(nsys, (uint32_t)cu_count() * 8);
  batch::k_bd_offsets<<<g, kBlock, 0, s>>>(A->row_ptr.as<uint32_t>(), A->col.as<uint32_t>(),
                                           B->sys_off.as<uint32_t>(), nsys, B->bd_offs.as<int32_t>(),
                                           B->bd_nds.as<int32_t>(), bad.as<uint32_t>());
  MSPCHK(launch_check("k_bd_offsets"));
  std::vector<int32_t> nds(nsys);
  uint32_t hb = 0;
  HIPCHK(hipMemcpyAsync(&hb, bad.p, sizeof hb, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(nds.data(), B->bd_nds.p, sizeof(int32_t) * nsys, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (hb) return MSP_OK;  // some system has more than kBD diagonals: CSR product
  bool all5 = true;
  for (int32_t c : nds) all5 = all5 && c == 5;
  const int planes = all5 ? 5 : batch::kBD;
  if (all5) {
    std::vector<int32_t> offs((size_t)batch::kBD * nsys);
    HIPCHK(hipMemcpy(offs.data(), B->bd_offs.p, sizeof(int32_t) * offs.size(), hipMemcpyDeviceToHost));
    bool mid = true;
    for (uint32_t q = 0; q < nsys && mid; ++q) mid = offs[(size_t)q * batch::kBD + 2] == 0;
    B->bd_mid = mid;
  }
  const size_t bytes = sizeof(float) * (size_t)planes * B->n;
  MSPCHK(B->bd_planes.alloc(bytes));
  HIPCHK(hipMemsetAsync(B->bd_planes.p, 0, bytes, s));
  batch::k_bd_fill<<<g, kBlock, 0, s>>>(A->row_ptr.as<uint32_t>(), A->col.as<uint32_t>(), A->val.as<float>(),
                                        B->sys_off.as<uint32_t>(), nsys, B->bd_offs.as<int32_t>(), B->n,
                                        B->bd_planes.as<float>());
  MSPCHK(launch_check("k_bd_fill"));
  B->bd_nd = planes;
  if (B->bd_mid) {  // symmetric pattern (-b, -a, 0, a, b) in every system, then the values
    std::vector<int32_t> offs((size_t)batch::kBD * nsys);
    HIPCHK(hipMemcpy(offs.data(), B->bd_offs.p, sizeof(int32_t) * offs.size(), hipMemcpyDeviceToHost));
    bool pat = true;
    for (uint32_t q = 0; q < nsys && pat; ++q) {
      const int32_t* o = &offs[(size_t)q * batch::kBD];
      pat = o[0] == -o[4] && o[1] == -o[3] && o[3] > 0;
    }
    if (pat) {
      HIPCHK(hipMemsetAsync(bad.p, 0, sizeof(uint32_t), s));
      batch::k_bd_symcheck<<<g, kBlock, 0, s>>>(B->bd_planes.as<float>(), B->n, B->sys_off.as<uint32_t>(), nsys,
                                                B->bd_offs.as<int32_t>(), bad.as<uint32_t>());
      MSPCHK(launch_check("k_bd_symcheck"));
      HIPCHK(hipMemcpyAsync(&hb, bad.p, sizeof hb, hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      B->bd_mirror = hb == 0;
    }
  }
  return MSP_OK;
}

extern "C" msp_status msp_batch_create(msp_batch* out, msp_csr A, const uint32_t* sys_off, uint32_t num_systems,
                                       void* stream) {
  if (!out) return fail(MSP_ERR_INVALID, "null output handle");
  *out = nullptr;
  MSPCHK(check_handle(A));
  if (A->n != A->m) return fail(MSP_ERR_INVALID, "batched systems need a square matrix");
  if (!sys_off && num_systems) return fail(MSP_ERR_INVALID, "null sys_off");
  hipStream_t s = (hipStream_t)stream;
  msp_batch B = new msp_batch_s;
  B->device = A->device;
  B->A = A;
  B->nsys = num_systems;
  B->n = A->n;
  auto body = [&]() -> msp_status {
    B->h_sys_off.resize((size_t)num_systems + 1);
    if (is_device_ptr(sys_off)) {
      HIPCHK(hipMemcpy(B->h_sys_off.data(), sys_off, sizeof(uint32_t) * (num_systems + 1), hipMemcpyDeviceToHost));
    } else {
      std::memcpy(B->h_sys_off.data(), sys_off, sizeof(uint32_t) * (num_systems + 1));
    }
    if (B->h_sys_off[0] != 0 || B->h_sys_off[num_systems] != A->n)
      return fail(MSP_ERR_INVALID, "sys_off must run from 0 to num_rows (%u)", A->n);
    for (uint32_t q = 0; q < num_systems; ++q)
      if (B->h_sys_off[q + 1] < B->h_sys_off[q]) return fail(MSP_ERR_INVALID, "sys_off must not decrease");
    MSPCHK(B->sys_off.alloc(sizeof(uint32_t) * (num_systems + 1)));
    HIPCHK(hipMemcpyAsync(B->sys_off.p, B->h_sys_off.data(), sizeof(uint32_t) * (num_systems + 1),
                          hipMemcpyHostToDevice, s));
    MSPCHK(batch_tile(B, 1024, s));
    DevMem bad;
    MSPCHK(bad.alloc(sizeof(uint32_t) * 2));
    HIPCHK(hipMemsetAsync(bad.p, 0, sizeof(uint32_t) * 2, s));
    batch::k_bt_check<<<batch_grid(B), kBlock, 0, s>>>(A->row_ptr.as<uint32_t>(), A->col.as<uint32_t>(),
                                                       B->tile_r0.as<uint32_t>(), B->tile_sys.as<uint32_t>(),
                                                       B->sys_off.as<uint32_t>(), B->ntiles, bad.as<uint32_t>());
    MSPCHK(B->dinv.alloc(sizeof(float) * ((size_t)A->n + 1)));
    batch::k_bt_dinv<<<grid_for(A->n), kBlock, 0, s>>>(A->row_ptr.as<uint32_t>(), A->col.as<uint32_t>(),
                                                       A->val.as<float>(), A->n, B->dinv.as<float>(),
                                                       bad.as<uint32_t>() + 1);
    MSPCHK(launch_check("batch check"));
    uint32_t hb[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(hb, bad.p, sizeof hb, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (hb[0]) return fail(MSP_ERR_INVALID, "%u entries couple different systems (not block diagonal)", hb[0]);
    if (hb[1]) return fail(MSP_ERR_SINGULAR, "%u rows without a positive diagonal (Jacobi needs one)", hb[1]);
    MSPCHK(batch_diagonal_storage(B, s));
    MSPCHK(B->state.alloc(sizeof(PairState) * batch::kRhs * ((size_t)num_systems + 1)));
    MSPCHK(B->sys_active.alloc(sizeof(int32_t) * ((size_t)num_systems + 1)));
    MSPCHK(B->n_active.alloc(sizeof(uint32_t)));
    MSPCHK(B->vec.alloc(sizeof(float) * 4 * (size_t)A->n * kBVecs + 64));
    HIPCHK(hipMemsetAsync(B->vec.p, 0, sizeof(float) * 4 * (size_t)A->n * kBVecs, s));
    HIPCHK(hipMemsetAsync(B->state.p, 0, sizeof(PairState) * batch::kRhs * ((size_t)num_systems + 1), s));
    HIPCHK(hipHostMalloc((void**)&B->h_active, sizeof(int32_t) * 2 * ((size_t)num_systems + 1),
                         hipHostMallocDefault));
    HIPCHK(hipStreamCreateWithFlags(&B->cs, hipStreamNonBlocking));
    for (auto& e : B->ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&B->join, hipEventDisableTiming));
    HIPCHK(hipEventCreate(&B->t0));
    HIPCHK(hipEventCreate(&B->t1));
    HIPCHK(hipStreamSynchronize(s));
    return MSP_OK;
  };
  msp_status st = body();
  if (st != MSP_OK) {
    delete B;
    return st;
  }
  *out = B;
  return MSP_OK;
}

extern "C" msp_status msp_batch_destroy(msp_batch B) {
  if (!B) return MSP_OK;
  if (!in_teardown()) {
    (void)hipSetDevice(B->device);
    (void)hipDeviceSynchronize();
  }
  delete B;
  return MSP_OK;
}

// Tile size and rows per thread for k_bt_spmm: each candidate is timed (3
// launches after one warm-up, all tiles active) and the fastest kept.
static msp_status batch_tune(msp_batch B, const msp_batch_opts& o, hipStream_t s) {
  std::vector<uint32_t> Ts = {256, 512, 1024, 2048};
  std::vector<int> rpts = {1, 2};
  if (o.tile_rows > 0) Ts = {(uint32_t)o.tile_rows};
  if (o.rows_per_thread > 0) rpts = {o.rows_per_thread};
  float best = 1e30f;
  uint32_t bestT = Ts[0];
  int bestR = rpts[0];
  const bool single = Ts.size() == 1 && rpts.size() == 1;
  for (uint32_t T : Ts) {
    if (B->T != T) MSPCHK(batch_tile(B, T, s));
    if (single) break;
    // every system active, every beta 0 (the state is reset by the solve's init)
    std::vector<int32_t> ones(B->nsys, 1);
    HIPCHK(hipMemcpyAsync(B->sys_active.p, ones.data(), sizeof(int32_t) * B->nsys, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemsetAsync(B->state.p, 0, sizeof(PairState) * batch::kRhs * B->nsys, s));
    batch_active_list(B, s);
    for (int rpt : rpts) {
      launch_spmm(B, rpt, 0, s);
      HIPCHK(hipEventRecord(B->t0, s));
      for (int k = 0; k < 3; ++k) launch_spmm(B, rpt, 0, s);
      HIPCHK(hipEventRecord(B->t1, s));
      HIPCHK(hipEventSynchronize(B->t1));
      float ms = 0.0f;
      HIPCHK(hipEventElapsedTime(&ms, B->t0, B->t1));
      if (ms < best) {
        best = ms;
        bestT = T;
        bestR = rpt;
      }
    }
    HIPCHK(hipStreamSynchronize(s));
  }
  if (B->T != bestT) MSPCHK(batch_tile(B, bestT, s));
  B->rpt = bestR;
  return MSP_OK;
}

static msp_status batch_pcg_impl(msp_batch B, const float* b, void* x, bool x64, const msp_batch_opts* opts,
                                 msp_batch_info* info, int32_t* pair_status, int32_t* pair_iterations,
                                 double* pair_relres, void* stream) {
  if (!B) return fail(MSP_ERR_INVALID, "null batch handle");
  HIPCHK(hipSetDevice(B->device));
  msp_batch_opts o;
  msp_default_batch_opts(&o);
  if (opts) o = *opts;
  if (o.nrhs < 1 || o.nrhs > batch::kRhs) return fail(MSP_ERR_INVALID, "nrhs must be 1..4");
  if (o.tol < 0 || o.maxiter < 0) return fail(MSP_ERR_INVALID, "negative tol / maxiter");
  if (o.steps_per_graph < 2 || (o.steps_per_graph & 1)) return fail(MSP_ERR_INVALID, "steps_per_graph: even, >= 2");
  if (o.tile_rows < 0 || (o.tile_rows > 0 && o.tile_rows % 64)) return fail(MSP_ERR_INVALID, "tile_rows: multiple of 64");
  if (o.rows_per_thread < 0 || o.rows_per_thread > 2) return fail(MSP_ERR_INVALID, "rows_per_thread: 0 (tune), 1, 2");
  if ((!b || !x) && B->n) return fail(MSP_ERR_INVALID, "null b / x");
  hipStream_t s = (hipStream_t)stream;
  const auto wall0 = std::chrono::steady_clock::now();
  const uint32_t n = B->n, gv = vec_grid(n);
  // right-hand sides into the [n][4] layout
  const size_t bytes = sizeof(float) * (size_t)n * o.nrhs;
  InArr bin;
  MSPCHK(bin.set(b, bytes, s));
  if (o.nrhs == batch::kRhs)
    HIPCHK(hipMemcpyAsync(B->v(kBB), bin.d, bytes, hipMemcpyDeviceToDevice, s));
  else
    batch::k_bt_pack<<<gv, kBlock, 0, s>>>(static_cast<const float*>(bin.d), n, o.nrhs, B->v(kBB));
  if (!B->tuned || B->tuned_for[0] != o.tile_rows || B->tuned_for[1] != o.rows_per_thread) {
    MSPCHK(batch_tune(B, o, s));
    B->tuned = true;
    B->tuned_for[0] = o.tile_rows;
    B->tuned_for[1] = o.rows_per_thread;
  }
  const auto wall1 = std::chrono::steady_clock::now();
  // the solve runs on the handle's capture stream, ordered after the caller's work
  HIPCHK(hipEventRecord(B->join, s));
  HIPCHK(hipStreamWaitEvent(B->cs, B->join, 0));
  MSPCHK(batch_graph(B, o.steps_per_graph, o.maxiter));
  hipStream_t cs = B->cs;
  HIPCHK(hipEventRecord(B->t0, cs));
  const batch::View V = B->view();
  batch::k_bt_init<<<batch_grid(B), kBlock, 0, cs>>>(V, B->v(kBB), B->v(kBX), B->v(kBXL), B->v(kBR), B->v(kBP0),
                                                     B->ntiles, B->partials.as<CSum>());
  batch::k_bt_fin_init<<<sys_grid(B), kBlock, 0, cs>>>(B->partials.as<CSum>(), B->sys_tile.as<uint32_t>(), B->nsys,
                                                       o.nrhs, (float)o.tol, o.maxiter, B->state.as<PairState>(),
                                                       B->sys_active.as<int32_t>());
  MSPCHK(launch_check("batch init"));
  const size_t act_bytes = sizeof(int32_t) * B->nsys;
  auto any_active = [&](int slot) {
    const int32_t* a = B->h_active + (size_t)slot * (B->nsys + 1);
    for (uint32_t q = 0; q < B->nsys; ++q)
      if (a[q]) return true;
    return false;
  };
  int launches = 0, verifications = 0;
  const int max_launches = (o.maxiter / o.steps_per_graph + 3) * (o.max_verify + 2);
  while (true) {
    // graphs until the trailing poll sees no active system
    int slot = 0;
    bool have_prev = false;
    while (launches < max_launches) {
      HIPCHK(hipGraphLaunch(B->exec, cs));
      ++launches;
      HIPCHK(hipMemcpyAsync(B->h_active + (size_t)slot * (B->nsys + 1), B->sys_active.p, act_bytes,
                            hipMemcpyDeviceToHost, cs));
      HIPCHK(hipEventRecord(B->ev[slot], cs));
      if (have_prev) {
        HIPCHK(hipEventSynchronize(B->ev[slot ^ 1]));
        if (!any_active(slot ^ 1)) break;
      }
      have_prev = true;
      slot ^= 1;
    }
    // true residuals; converged pairs above tol restart
    if (x64)
      batch::k_bt_residual<true><<<batch_grid(B), kBlock, 0, cs>>>(V, B->v(kBB), B->v(kBX), B->v(kBXL), B->v(kBR),
                                                                   B->ntiles, B->partials.as<CSum>());
    else
      batch::k_bt_residual<false><<<batch_grid(B), kBlock, 0, cs>>>(V, B->v(kBB), B->v(kBX), B->v(kBXL), B->v(kBR),
                                                                    B->ntiles, B->partials.as<CSum>());
    batch::k_bt_fin_verify<<<sys_grid(B), kBlock, 0, cs>>>(B->partials.as<CSum>(), B->sys_tile.as<uint32_t>(),
                                                           B->nsys, o.maxiter, o.max_verify, x64 ? 1 : 0,
                                                           B->state.as<PairState>(), B->sys_active.as<int32_t>());
    MSPCHK(launch_check("batch verify"));
    ++verifications;
    HIPCHK(hipMemcpyAsync(B->h_active, B->sys_active.p, act_bytes, hipMemcpyDeviceToHost, cs));
    HIPCHK(hipStreamSynchronize(cs));
    if (!any_active(0) || launches >= max_launches) break;
  }
  HIPCHK(hipEventRecord(B->t1, cs));
  // solution out (x + xl), ordered before the caller's later work
  {
    const size_t xbytes = (x64 ? sizeof(double) : sizeof(float)) * (size_t)n * o.nrhs;
    DevMem tmp;
    void* dst = x;
    if (!is_device_ptr(x)) {
      MSPCHK(tmp.alloc(xbytes + 16));
      dst = tmp.p;
    }
    if (x64)
      batch::k_bt_unpack64<<<gv, kBlock, 0, cs>>>(B->v(kBX), B->v(kBXL), n, o.nrhs, static_cast<double*>(dst));
    else
      batch::k_bt_unpack<<<gv, kBlock, 0, cs>>>(B->v(kBX), B->v(kBXL), n, o.nrhs, static_cast<float*>(dst));
    MSPCHK(launch_check("batch unpack"));
    if (dst != x) MSPCHK(copy_out(x, dst, xbytes, cs));
  }
  std::vector<PairState> ps((size_t)B->nsys * batch::kRhs);
  HIPCHK(hipMemcpyAsync(ps.data(), B->state.p, sizeof(PairState) * ps.size(), hipMemcpyDeviceToHost, cs));
  HIPCHK(hipStreamSynchronize(cs));
  HIPCHK(hipEventRecord(B->join, cs));
  HIPCHK(hipStreamWaitEvent(s, B->join, 0));
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, B->t0, B->t1);
  msp_batch_info inf{};
  inf.total_pairs = (int32_t)(B->nsys * o.nrhs);
  for (uint32_t q = 0; q < B->nsys; ++q)
    for (int k = 0; k < o.nrhs; ++k) {
      const PairState& p = ps[(size_t)q * batch::kRhs + k];
      const size_t idx = (size_t)q * o.nrhs + k;
      const int32_t st = p.status == batch::kPairRunning ? MSP_SOLVE_MAXITER
                         : p.status == batch::kPairConverged ? MSP_SOLVE_CONVERGED
                         : p.status == batch::kPairBreakdown ? MSP_SOLVE_BREAKDOWN
                         : p.status == batch::kPairStagnated ? MSP_SOLVE_STAGNATED
                                                             : MSP_SOLVE_MAXITER;
      if (st == MSP_SOLVE_CONVERGED) inf.converged_pairs += 1;
      inf.max_iterations = std::max(inf.max_iterations, p.iter);
      if (pair_status) pair_status[idx] = st;
      if (pair_iterations) pair_iterations[idx] = p.iter;
      if (pair_relres) pair_relres[idx] = p.bb > 0 ? std::sqrt((double)p.rr_true / (double)p.bb) : 0.0;
    }
  inf.graph_launches = launches;
  inf.verifications = verifications;
  inf.tile_rows = (int32_t)B->T;
  inf.diagonal_planes = B->bd_nd;
  inf.symmetric_storage = B->bd_mirror ? 1 : 0;
  inf.rows_per_thread = B->rpt;
  inf.seconds = ms * 1e-3;
  inf.setup_seconds = std::chrono::duration<double>(wall1 - wall0).count();
  if (info) *info = inf;
  return MSP_OK;
}

extern "C" msp_status msp_batch_pcg(msp_batch B, const float* b, float* x, const msp_batch_opts* opts,
                                    msp_batch_info* info, int32_t* pair_status, int32_t* pair_iterations,
                                    double* pair_relres, void* stream) {
  return batch_pcg_impl(B, b, x, false, opts, info, pair_status, pair_iterations, pair_relres, stream);
}

extern "C" msp_status msp_batch_pcg_x64(msp_batch B, const float* b, double* x, const msp_batch_opts* opts,
                                        msp_batch_info* info, int32_t* pair_status, int32_t* pair_iterations,
                                        double* pair_relres, void* stream) {
  return batch_pcg_impl(B, b, x, true, opts, info, pair_status, pair_iterations, pair_relres, stream);
}

// ---------------------------------------------------------------------------
// Direct sparse solve: banded Householder QR (bandqr.hpp). Replaces
// sparse_bridge/bridge.cpp:14-191 (SparseQRSolver: symbolic + numeric QR at
// construction, solve(B) for one or many right-hand sides).
struct msp_qr_s {
  int device = 0;
  int64_t m = 0, n = 0;
  int kl = 0, ku = 0, wc = 1;
  DevMem Rb, Vb, tau;   // R rows [n][wc], reflectors [n][kl], tau [n]
  DevMem src_row;       // ordered row -> caller's row (int32 [m])
  DevMem pcol;          // caller's column -> ordered column (int32 [n])
};

// Reverse Cuthill-McKee order of the symmetric pattern adj (CSR, no self
// loops): each component from a pseudo-peripheral start (the lowest-degree
// node of the last BFS level, repeated while the eccentricity grows),
// neighbours visited by increasing degree, the whole order reversed.
static std::vector<int32_t> rcm_order(int64_t n, const std::vector<int64_t>& ptr, const std::vector<int32_t>& adj) {
  std::vector<int32_t> order;
  order.reserve((size_t)n);
  std::vector<int32_t> level((size_t)n, -1);
  std::vector<char> done((size_t)n, 0);
  auto deg = [&](int32_t v) { return ptr[v + 1] - ptr[v]; };
  // BFS from s over unfinished nodes; returns the last level's nodes and the depth
  auto bfs = [&](int32_t s, std::vector<int32_t>& last) {
    std::vector<int32_t> cur{s}, seen{s};
    level[s] = 0;
    int depth = 0;
    while (true) {
      std::vector<int32_t> nxt;
      for (int32_t v : cur)
        for (int64_t e = ptr[v]; e < ptr[v + 1]; ++e) {
          const int32_t w = adj[e];
          if (done[w] || level[w] >= 0) continue;
          level[w] = depth + 1;
          nxt.push_back(w);
          seen.push_back(w);
        }
      if (nxt.empty()) break;
      cur.swap(nxt);
      ++depth;
    }
    last = cur;
    for (int32_t v : seen) level[v] = -1;
    return depth;
  };
  std::vector<int32_t> byDeg((size_t)n);
  for (int64_t v = 0; v < n; ++v) byDeg[v] = (int32_t)v;
  std::stable_sort(byDeg.begin(), byDeg.end(), [&](int32_t a, int32_t b) { return deg(a) < deg(b); });
  for (int32_t s0 : byDeg) {
    if (done[s0]) continue;
    int32_t s = s0;
    std::vector<int32_t> last;
    int ecc = bfs(s, last);
    for (int it = 0; it < 8; ++it) {
      int32_t cand = last[0];
      for (int32_t v : last)
        if (deg(v) < deg(cand)) cand = v;
      std::vector<int32_t> last2;
      const int e2 = bfs(cand, last2);
      if (e2 <= ecc) break;
      s = cand;
      ecc = e2;
      last.swap(last2);
    }
    // Cuthill-McKee from s
    size_t head = order.size();
    order.push_back(s);
    done[s] = 1;
    std::vector<int32_t> nb;
    while (head < order.size()) {
      const int32_t v = order[head++];
      nb.clear();
      for (int64_t e = ptr[v]; e < ptr[v + 1]; ++e)
        if (!done[adj[e]]) {
          done[adj[e]] = 1;
          nb.push_back(adj[e]);
        }
      std::stable_sort(nb.begin(), nb.end(), [&](int32_t a, int32_t b) { return deg(a) < deg(b); });
      order.insert(order.end(), nb.begin(), nb.end());
    }
  }
  std::reverse(order.begin(), order.end());
  return order;
}

// Band of the pattern under row map prow (caller row -> ordered row) and
// column map pcol.
static void band_of(int64_t n, const std::vector<int64_t>& cs, const std::vector<int64_t>& ri,
                    const std::vector<int32_t>& prow, const std::vector<int32_t>& pcol, int64_t* kl, int64_t* ku) {
  int64_t l = 0, u = 0;
  for (int64_t c = 0; c < n; ++c)
    for (int64_t e = cs[c]; e < cs[c + 1]; ++e) {
      const int64_t d = (int64_t)pcol[c] - (int64_t)prow[ri[e]];
      if (d > u) u = d;
      if (-d > l) l = -d;
    }
  *kl = l;
  *ku = u;
}

static size_t bqr_factor_lds(int kl, int wc) {
  return sizeof(double) * ((size_t)(kl + 1) * wc + (size_t)bqr::kStage * wc + (size_t)(kl + 1));
}
static size_t bqr_qt_lds(int kl) {
  return sizeof(double) * ((size_t)(kl + 1) * bqr::kRhs + (size_t)bqr::kChunk * kl + bqr::kChunk);
}
static size_t bqr_rs_lds(int wc) { return sizeof(double) * ((size_t)wc * bqr::kRhs + (size_t)bqr::kChunk * wc); }

static msp_status lds_allow(const void* kernel, size_t bytes) {
  int dev = 0, maxb = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&maxb, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  if (bytes > (size_t)maxb)
    return fail(MSP_ERR_UNSUPPORTED, "banded QR: the band window needs %zu bytes of LDS per workgroup (device allows %d)",
                bytes, maxb);
  if (bytes > 65536) HIPCHK(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  return MSP_OK;
}

extern "C" msp_status msp_qr_create_csc(msp_qr* out, const int64_t* col_starts, const int64_t* row_indices,
                                        const double* values, int64_t num_rows, int64_t num_cols, void* stream) {
  if (!out) return fail(MSP_ERR_INVALID, "null output handle");
  *out = nullptr;
  MSPCHK(need_device());
  if (!col_starts || num_rows <= 0 || num_cols <= 0) return fail(MSP_ERR_INVALID, "msp_qr_create_csc: empty matrix");
  if (num_rows < num_cols)
    return fail(MSP_ERR_UNSUPPORTED, "banded QR needs num_rows >= num_cols (got %lld x %lld)", (long long)num_rows,
                (long long)num_cols);
  if (num_rows > INT32_MAX) return fail(MSP_ERR_INVALID, "num_rows exceeds the int32 index range");
  hipStream_t s = (hipStream_t)stream;
  const int64_t m = num_rows, n = num_cols;
  // symbolic phase on the host: the pattern, an ordering, the band
  std::vector<int64_t> cs((size_t)n + 1);
  HIPCHK(hipMemcpy(cs.data(), col_starts, sizeof(int64_t) * (n + 1), hipMemcpyDefault));
  if (cs[0] != 0) return fail(MSP_ERR_INVALID, "col_starts[0] must be 0");
  for (int64_t c = 0; c < n; ++c)
    if (cs[c + 1] < cs[c]) return fail(MSP_ERR_INVALID, "col_starts must be non-decreasing");
  const int64_t nnz = cs[n];
  if (nnz > 0 && (!row_indices || !values)) return fail(MSP_ERR_INVALID, "null row_indices / values");
  std::vector<int64_t> ri((size_t)nnz);
  if (nnz) HIPCHK(hipMemcpy(ri.data(), row_indices, sizeof(int64_t) * nnz, hipMemcpyDefault));
  for (int64_t e = 0; e < nnz; ++e)
    if (ri[e] < 0 || ri[e] >= m) return fail(MSP_ERR_INVALID, "row index %lld out of range", (long long)ri[e]);
  std::vector<int32_t> prow((size_t)m), pcol((size_t)n);
  int64_t kl = 0, ku = 0;
  if (m == n) {
    // identity or reverse Cuthill-McKee on A + A^T, whichever gives the smaller window
    std::vector<int64_t> ptr((size_t)n + 1, 0);
    for (int64_t c = 0; c < n; ++c)
      for (int64_t e = cs[c]; e < cs[c + 1]; ++e)
        if (ri[e] != c) {
          ++ptr[ri[e] + 1];
          ++ptr[c + 1];
        }
    for (int64_t v = 0; v < n; ++v) ptr[v + 1] += ptr[v];
    std::vector<int32_t> adj((size_t)ptr[n]);
    std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
    for (int64_t c = 0; c < n; ++c)
      for (int64_t e = cs[c]; e < cs[c + 1]; ++e)
        if (ri[e] != c) {
          adj[fill[ri[e]]++] = (int32_t)c;
          adj[fill[c]++] = (int32_t)ri[e];
        }
    std::vector<int64_t> uptr((size_t)n + 1, 0);  // sorted, unique neighbour lists
    std::vector<int32_t> uadj;
    uadj.reserve(adj.size());
    for (int64_t v = 0; v < n; ++v) {
      std::sort(adj.begin() + ptr[v], adj.begin() + ptr[v + 1]);
      for (int64_t e = ptr[v]; e < ptr[v + 1]; ++e)
        if (e == ptr[v] || adj[e] != adj[e - 1]) uadj.push_back(adj[e]);
      uptr[v + 1] = (int64_t)uadj.size();
    }
    for (int64_t v = 0; v < n; ++v) prow[v] = pcol[v] = (int32_t)v;
    band_of(n, cs, ri, prow, pcol, &kl, &ku);
    const std::vector<int32_t> order = rcm_order(n, uptr, uadj);
    std::vector<int32_t> inv((size_t)n);
    for (int64_t k = 0; k < n; ++k) inv[order[k]] = (int32_t)k;
    int64_t kl2 = 0, ku2 = 0;
    band_of(n, cs, ri, inv, inv, &kl2, &ku2);
    if ((kl2 + 1) * (kl2 + ku2 + 1) < (kl + 1) * (kl + ku + 1)) {
      prow = inv;
      pcol = inv;
      kl = kl2;
      ku = ku2;
    }
  } else {
    // tall: columns as given, rows stably sorted by their leftmost column
    std::vector<int64_t> lead((size_t)m, n);
    for (int64_t c = 0; c < n; ++c)
      for (int64_t e = cs[c]; e < cs[c + 1]; ++e) lead[ri[e]] = std::min(lead[ri[e]], c);
    std::vector<int32_t> rows((size_t)m);
    for (int64_t r = 0; r < m; ++r) rows[r] = (int32_t)r;
    std::stable_sort(rows.begin(), rows.end(), [&](int32_t a, int32_t b) { return lead[a] < lead[b]; });
    for (int64_t k = 0; k < m; ++k) prow[rows[k]] = (int32_t)k;
    for (int64_t c = 0; c < n; ++c) pcol[c] = (int32_t)c;
    band_of(n, cs, ri, prow, pcol, &kl, &ku);
  }
  if (kl + ku + 1 > bqr::kMaxWC || kl > bqr::kMaxKl)
    return fail(MSP_ERR_UNSUPPORTED,
                "banded QR: the ordered matrix has bandwidths kl=%lld ku=%lld; kl+ku+1 must be at most %d and "
                "kl at most %d",
                (long long)kl, (long long)ku, bqr::kMaxWC, bqr::kMaxKl);
  const int ikl = (int)kl, iku = (int)ku, wc = ikl + iku + 1;
  const bool narrow = wc <= kWave;
  const void* fk = narrow ? (const void*)bqr::k_bqr_factor<kWave> : (const void*)bqr::k_bqr_factor<kBlock>;
  MSPCHK(lds_allow(fk, bqr_factor_lds(ikl, wc)));
  MSPCHK(lds_allow((const void*)bqr::k_bqr_apply_qt, bqr_qt_lds(ikl)));
  MSPCHK(lds_allow((const void*)bqr::k_bqr_rsolve, bqr_rs_lds(wc)));
  msp_qr F = new msp_qr_s;
  std::unique_ptr<msp_qr_s> guard(F);
  HIPCHK(hipGetDevice(&F->device));
  F->m = m;
  F->n = n;
  F->kl = ikl;
  F->ku = iku;
  F->wc = wc;
  std::vector<int32_t> src((size_t)m);
  for (int64_t r = 0; r < m; ++r) src[prow[r]] = (int32_t)r;
  MSPCHK(F->src_row.alloc(sizeof(int32_t) * m));
  MSPCHK(F->pcol.alloc(sizeof(int32_t) * n));
  HIPCHK(hipMemcpyAsync(F->src_row.p, src.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(F->pcol.p, pcol.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
  TmpMem d_prow, d_cs, d_ri, RB, bad;
  MSPCHK(d_prow.alloc(sizeof(int32_t) * m, s));
  HIPCHK(hipMemcpyAsync(d_prow.p, prow.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice, s));
  MSPCHK(d_cs.alloc(sizeof(int64_t) * (n + 1), s));
  HIPCHK(hipMemcpyAsync(d_cs.p, cs.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, s));
  MSPCHK(d_ri.alloc(sizeof(int64_t) * std::max<int64_t>(nnz, 1), s));
  if (nnz) HIPCHK(hipMemcpyAsync(d_ri.p, ri.data(), sizeof(int64_t) * nnz, hipMemcpyHostToDevice, s));
  InArr v;
  MSPCHK(v.set(values, sizeof(double) * nnz, s));
  MSPCHK(RB.alloc(sizeof(double) * m * wc, s));
  HIPCHK(hipMemsetAsync(RB.p, 0, sizeof(double) * m * wc, s));
  MSPCHK(bad.alloc(sizeof(int), s));
  HIPCHK(hipMemsetAsync(bad.p, 0, sizeof(int), s));
  bqr::k_bqr_fill<<<grid_for((uint64_t)n), kBlock, 0, s>>>(d_cs.as<int64_t>(), d_ri.as<int64_t>(),
                                                          static_cast<const double*>(v.d), n, d_prow.as<int32_t>(),
                                                          F->pcol.as<int32_t>(), ikl, wc, RB.as<double>(),
                                                          bad.as<int>());
  MSPCHK(launch_check("k_bqr_fill"));
  MSPCHK(F->Rb.alloc(sizeof(double) * n * wc));
  MSPCHK(F->Vb.alloc(sizeof(double) * std::max<int64_t>(n * ikl, 1)));
  MSPCHK(F->tau.alloc(sizeof(double) * n));
  const size_t lds = bqr_factor_lds(ikl, wc);
  if (ikl == 1 && iku == 1)
    bqr::k_bqr_factor_small<1, 1><<<1, kWave, 0, s>>>(RB.as<double>(), m, n, ikl, iku, F->Rb.as<double>(),
                                                      F->Vb.as<double>(), F->tau.as<double>());
  else if (ikl + 1 <= bqr::kSmallWR && wc <= bqr::kSmallWC)
    bqr::k_bqr_factor_small<-1, -1><<<1, kWave, 0, s>>>(RB.as<double>(), m, n, ikl, iku, F->Rb.as<double>(),
                                                        F->Vb.as<double>(), F->tau.as<double>());
  else if (narrow)
    bqr::k_bqr_factor<kWave><<<1, kWave, lds, s>>>(RB.as<double>(), m, n, ikl, iku, F->Rb.as<double>(),
                                                   F->Vb.as<double>(), F->tau.as<double>());
  else
    bqr::k_bqr_factor<kBlock><<<1, kBlock, lds, s>>>(RB.as<double>(), m, n, ikl, iku, F->Rb.as<double>(),
                                                     F->Vb.as<double>(), F->tau.as<double>());
  MSPCHK(launch_check("k_bqr_factor"));
  bqr::k_bqr_check<<<grid_for((uint64_t)n), kBlock, 0, s>>>(F->Rb.as<double>(), n, wc, bad.as<int>() + 0);
  MSPCHK(launch_check("k_bqr_check"));
  int hb = 0;
  HIPCHK(hipMemcpyAsync(&hb, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (hb) return fail(MSP_ERR_SINGULAR, "banded QR: R has a zero or non-finite diagonal entry (rank-deficient matrix)");
  *out = guard.release();
  return MSP_OK;
}

extern "C" msp_status msp_qr_bandwidth(msp_qr F, int32_t* lower, int32_t* upper) {
  if (!F) return fail(MSP_ERR_INVALID, "null QR handle");
  if (lower) *lower = F->kl;
  if (upper) *upper = F->ku;
  return MSP_OK;
}

extern "C" msp_status msp_qr_solve(msp_qr F, const double* B, int64_t ldb, int64_t nrhs, double* X, int64_t ldx,
                                   void* stream) {
  if (!F) return fail(MSP_ERR_INVALID, "null QR handle");
  if (nrhs <= 0) return MSP_OK;
  if (!B || !X) return fail(MSP_ERR_INVALID, "null B / X");
  if (ldb < F->m || ldx < F->n) return fail(MSP_ERR_INVALID, "ldb must be >= num_rows and ldx >= num_cols");
  HIPCHK(hipSetDevice(F->device));
  hipStream_t s = (hipStream_t)stream;
  const int64_t m = F->m, n = F->n;
  InArr bin;
  MSPCHK(bin.set(B, sizeof(double) * (size_t)(ldb * (nrhs - 1) + m), s));
  TmpMem Y, Xp, xo;
  MSPCHK(Y.alloc(sizeof(double) * m * nrhs, s));
  MSPCHK(Xp.alloc(sizeof(double) * n * nrhs, s));
  const uint64_t gm = (uint64_t)(m * nrhs);
  bqr::k_bqr_gather_rhs<<<grid_for(gm), kBlock, 0, s>>>(static_cast<const double*>(bin.d), ldb, m, nrhs,
                                                        F->src_row.as<int32_t>(), Y.as<double>());
  const uint32_t g = (uint32_t)((nrhs + bqr::kRhs - 1) / bqr::kRhs);
  if (F->kl + 1 <= bqr::kSmallWR)
    bqr::k_bqr_apply_qt_small<<<g, bqr::kRhs, 0, s>>>(F->Vb.as<double>(), F->tau.as<double>(), m, n, F->kl,
                                                      Y.as<double>(), nrhs);
  else
    bqr::k_bqr_apply_qt<<<g, bqr::kRhs, bqr_qt_lds(F->kl), s>>>(F->Vb.as<double>(), F->tau.as<double>(), m, n,
                                                                F->kl, Y.as<double>(), nrhs);
  if (F->wc <= bqr::kSmallWC)
    bqr::k_bqr_rsolve_small<<<g, bqr::kRhs, 0, s>>>(F->Rb.as<double>(), n, F->wc, Y.as<double>(), Xp.as<double>(),
                                                    nrhs);
  else
    bqr::k_bqr_rsolve<<<g, bqr::kRhs, bqr_rs_lds(F->wc), s>>>(F->Rb.as<double>(), n, F->wc, Y.as<double>(),
                                                              Xp.as<double>(), nrhs);
  MSPCHK(launch_check("banded QR solve"));
  const bool xdev = is_device_ptr(X);
  double* xd = X;
  if (!xdev) {
    MSPCHK(xo.alloc(sizeof(double) * (size_t)(ldx * (nrhs - 1) + n), s));
    xd = xo.as<double>();
  }
  bqr::k_bqr_scatter_x<<<grid_for((uint64_t)(n * nrhs)), kBlock, 0, s>>>(Xp.as<double>(), n, nrhs,
                                                                        F->pcol.as<int32_t>(), xd, ldx);
  MSPCHK(launch_check("k_bqr_scatter_x"));
  if (!xdev)  // column by column: the caller's rows n .. ldx-1 stay untouched
    HIPCHK(hipMemcpy2DAsync(X, sizeof(double) * ldx, xd, sizeof(double) * ldx, sizeof(double) * n, nrhs,
                            hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return MSP_OK;
}

extern "C" msp_status msp_qr_destroy(msp_qr F) {
  if (!F) return MSP_OK;
  if (!in_teardown()) {
    (void)hipSetDevice(F->device);
    (void)hipDeviceSynchronize();
  }
  delete F;
  return MSP_OK;
}
