// C-ABI of libodba_hip: host orchestration of the DTW / DBA kernels.
// Declarations and contracts: include/odba/odba.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/odba/odba.h"
#include "odba_dtw.hpp"

using namespace odba;

struct odba_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int cus = 0;
  int clock_khz = 0;
  uint64_t mem = 0;
  uint64_t mem_cap = 0;   // device memory this context may plan for (serving slots get a share)
  std::string name;
  std::string err;
  // grow-only device scratch
  void* esc = nullptr;
  size_t esc_bytes = 0;
  void* small = nullptr;  // counters, flags (256 bytes)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::map<std::string, odba_schedule> tuned;  // all-vs-all schedules by workload shape
  // Pinned host staging for graph replays, grow-only: pinned allocations
  // can wait for the whole device, which would stall other contexts' work.
  void* pinned = nullptr;
  size_t pinned_bytes = 0;
  bool graphs = true;  // replay DBA rounds as hipGraphs (off for contexts that share a busy device)
  // Grow-only buffers of the calls that would otherwise allocate and free
  // per call (all-vs-all distances and flags, z-norm statistics, the pair
  // order and small-pair records of a job's all-vs-all): freeing device
  // memory waits for the device's queued work, so a serving slot that
  // freed at the end of each preparation waited for every other slot's and
  // engine's kernels.
  void* gb[5] = {};
  size_t gb_bytes[5] = {};
};

namespace {
enum { kGbDist = 0, kGbFlags, kGbStats, kGbOrder, kGbSmallTab };
// The context's grow-only buffer `k` with at least `bytes`, or null.
void* ctx_buffer(odba_ctx* c, int k, size_t bytes) {
  if (bytes <= c->gb_bytes[k]) return c->gb[k];
  if (c->gb[k]) hipFree(c->gb[k]);
  c->gb[k] = nullptr;
  c->gb_bytes[k] = 0;
  const size_t want = std::max<size_t>(bytes + bytes / 4, 4096);
  if (hipMalloc(&c->gb[k], want) != hipSuccess) return nullptr;
  c->gb_bytes[k] = want;
  return c->gb[k];
}
}  // namespace

struct odba_seqset {
  odba_ctx* ctx = nullptr;
  int dtype = 0;
  int n = 0;
  void* data = nullptr;        // packed, padded
  long long* offs = nullptr;   // device
  int* lens = nullptr;         // device
  std::vector<long long> h_offs;
  std::vector<int> h_lens;
  int maxlen = 0;
  bool int_no_wrap = true;     // integer costs provably cannot wrap
  void* order = nullptr;       // device pair order (int2 per pair), largest cell count first
  SmallPairRec* small_tab = nullptr;  // device pair records of the small-pair kernel
  int small_tab_order = -1;           // the order they were built in
  bool borrowed = false;              // a view: order and small_tab live in the context's buffers
};

// Fixed device scratch of odba_dtw_all_pairs_noalloc: the pair counter in
// the first 256 bytes, then the edge columns of the resident waves. Each
// device holds its own copy from the module's load until the process ends.
// In host code `scratch` names a host-side stand-in, not a device address:
// only the per-device address looked up by odba_ctx_create is passed on.
static __device__ __align__(256) unsigned char scratch[ODBA_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};  // device address of `scratch`, per device

size_t dtype_size(int dt) {
  switch (dt) {
    case ODBA_I32: case ODBA_U32: case ODBA_F32: return 4;
    case ODBA_U64: case ODBA_F64: return 8;
  }
  return 0;
}

int fail(odba_ctx* c, int code, const char* fmt, ...) {
  if (c) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    c->err = buf;
  }
  return code;
}

#define HIPC(ctx, call)                                                                   \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess) return fail(ctx, ODBA_E_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

#define ODBA_DISPATCH(dtype, CALL)                        \
  switch (dtype) {                                        \
    case ODBA_I32: { typedef int32_t T; return CALL; }    \
    case ODBA_U32: { typedef uint32_t T; return CALL; }   \
    case ODBA_U64: { typedef uint64_t T; return CALL; }   \
    case ODBA_F32: { typedef float T; return CALL; }      \
    case ODBA_F64: { typedef double T; return CALL; }     \
    default: return fail(c, ODBA_E_INVALID, "unknown dtype %d", dtype); \
  }


int ensure_scratch(odba_ctx* c, size_t bytes) {
  if (bytes <= c->esc_bytes) return ODBA_OK;
  if (c->esc) hipFree(c->esc);
  c->esc = nullptr;
  c->esc_bytes = 0;
  if (hipMalloc(&c->esc, bytes) != hipSuccess) return fail(c, ODBA_E_NOMEM, "cannot allocate %zu bytes of edge scratch", bytes);
  c->esc_bytes = bytes;
  return ODBA_OK;
}

template <typename T> bool is_float() { return std::is_floating_point<T>::value; }

// Integer costs cannot wrap when every path cost stays below the type's max:
// cost <= (ly + lx) * range^2.
template <typename T>
bool int_cannot_wrap(const T* const* seqs, const uint64_t* lens, int n) {
  if (is_float<T>()) return true;
  long double lo = 0, hi = 0, maxlen = 0;
  bool first = true;
  for (int i = 0; i < n; i++) {
    for (uint64_t k = 0; k < lens[i]; k++) {
      long double v = (long double)seqs[i][k];
      if (first) { lo = hi = v; first = false; }
      lo = std::min(lo, v);
      hi = std::max(hi, v);
    }
    maxlen = std::max(maxlen, (long double)lens[i]);
  }
  long double r = hi - lo;
  long double bound = 2.0L * maxlen * r * r + r * r;
  return bound < (long double)std::numeric_limits<T>::max() / 2;
}

// Device allocations freed together when the owner goes out of scope.
struct DevBufs {
  std::vector<void*> p;
  ~DevBufs() { for (void* x : p) if (x) hipFree(x); }
  template <typename P> int get(odba_ctx* c, P*& out, size_t elems) {
    void* q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(elems, 1) * sizeof(P)) != hipSuccess)
      return fail(c, ODBA_E_NOMEM, "cannot allocate %zu bytes", elems * sizeof(P));
    p.push_back(q);
    out = (P*)q;
    return ODBA_OK;
  }
};

// The free-memory figure of ODBA_PATH_AUTO: the stated one, or with none
// (0) the device's free memory when the session's sequences are resident,
// capped at the context's share; the reference reads cudaMemGetInfo before
// each sequence's path (dba.hpp:523-524).
uint64_t auto_free_mem(odba_ctx* c, int policy, uint64_t stated) {
  if (policy != ODBA_PATH_AUTO || stated) return stated;
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
    (void)hipGetLastError();
    return c->mem_cap;
  }
  return std::min<uint64_t>((uint64_t)fr, c->mem_cap);
}

int grid_blocks(odba_ctx* c, long long work_items) {
  long long want = (work_items + 3) / 4;
  long long cap = (long long)c->cus * 8;  // 8 workgroups of 4 waves per CU
  return (int)std::max(1LL, std::min(want, cap));
}

}  // namespace

extern "C" {

int odba_ctx_create(int device, odba_ctx** out) {
  if (!out) return ODBA_E_INVALID;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return ODBA_E_NO_DEVICE;
  if (device < 0 || device >= ndev) return ODBA_E_NO_DEVICE;
  odba_ctx* c = new odba_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) { delete c; return ODBA_E_HIP; }
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, device) != hipSuccess) { delete c; return ODBA_E_HIP; }
  c->cus = p.multiProcessorCount;
  c->clock_khz = p.clockRate;
  c->mem = p.totalGlobalMem;
  c->mem_cap = c->mem;
  c->name = p.gcnArchName;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return ODBA_E_HIP; }
  if (hipMalloc(&c->small, 256) != hipSuccess) { hipStreamDestroy(c->stream); delete c; return ODBA_E_NOMEM; }
  if (device < kMaxDevices) {
    // once per device, with that device current: the fixed scratch's device
    // address (a failure leaves it unset; odba_dtw_all_pairs_noalloc then
    // reports it and every other entry point works as before)
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_dev[device]) {
      void* p = nullptr;
      if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess) g_scratch_dev[device] = p;
      else (void)hipGetLastError();
    }
  }
  hipEventCreate(&c->ev0);
  hipEventCreate(&c->ev1);
  *out = c;
  return ODBA_OK;
}

int odba_ctx_destroy(odba_ctx* c) {
  if (!c) return ODBA_OK;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  if (c->esc) hipFree(c->esc);
  if (c->small) hipFree(c->small);
  for (void* p : c->gb)
    if (p) hipFree(p);
  if (c->pinned) hipHostFree(c->pinned);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  hipStreamDestroy(c->stream);
  delete c;
  return ODBA_OK;
}

const char* odba_last_error(const odba_ctx* c) { return c ? c->err.c_str() : "null context"; }

int odba_device_info(odba_ctx* c, int* cus, int* clock_khz, uint64_t* mem, char* name, int name_len) {
  if (!c) return ODBA_E_INVALID;
  if (cus) *cus = c->cus;
  if (clock_khz) *clock_khz = c->clock_khz;
  if (mem) *mem = c->mem;
  if (name && name_len > 0) {
    std::strncpy(name, c->name.c_str(), name_len - 1);
    name[name_len - 1] = 0;
  }
  return ODBA_OK;
}

int odba_seqset_create(odba_ctx* c, int dtype, const void* const* seqs, const uint64_t* lens, int n, odba_seqset** out) {
  if (!c || !seqs || !lens || !out || n <= 0) return fail(c, ODBA_E_INVALID, "invalid seqset arguments");
  const size_t es = dtype_size(dtype);
  if (!es) return fail(c, ODBA_E_INVALID, "unknown dtype %d", dtype);
  HIPC(c, hipSetDevice(c->device));
  auto* s = new odba_seqset();
  s->ctx = c;
  s->dtype = dtype;
  s->n = n;
  s->h_offs.resize(n);
  s->h_lens.resize(n);
  long long total = 0;
  for (int i = 0; i < n; i++) {
    if (lens[i] == 0 || lens[i] > (uint64_t)INT32_MAX / 2) { delete s; return fail(c, ODBA_E_INVALID, "sequence %d has length %llu", i, (unsigned long long)lens[i]); }
    s->h_offs[i] = total;
    s->h_lens[i] = (int)lens[i];
    s->maxlen = std::max(s->maxlen, (int)lens[i]);
    total += padded_len((long long)lens[i]);
  }
  std::vector<unsigned char> host((size_t)total * es, 0);
  for (int i = 0; i < n; i++) std::memcpy(&host[(size_t)s->h_offs[i] * es], seqs[i], (size_t)lens[i] * es);
  switch (dtype) {
    case ODBA_I32: s->int_no_wrap = int_cannot_wrap<int32_t>((const int32_t* const*)seqs, lens, n); break;
    case ODBA_U32: s->int_no_wrap = int_cannot_wrap<uint32_t>((const uint32_t* const*)seqs, lens, n); break;
    case ODBA_U64: s->int_no_wrap = int_cannot_wrap<uint64_t>((const uint64_t* const*)seqs, lens, n); break;
    default: s->int_no_wrap = true;
  }
  if (hipMalloc(&s->data, host.size()) != hipSuccess || hipMalloc(&s->offs, n * sizeof(long long)) != hipSuccess ||
      hipMalloc(&s->lens, n * sizeof(int)) != hipSuccess) {
    odba_seqset_destroy(s);
    return fail(c, ODBA_E_NOMEM, "cannot allocate device sequence set (%zu bytes)", host.size());
  }
  HIPC(c, hipMemcpyAsync(s->data, host.data(), host.size(), hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemcpyAsync(s->offs, s->h_offs.data(), n * sizeof(long long), hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemcpyAsync(s->lens, s->h_lens.data(), n * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  *out = s;
  return ODBA_OK;
}

int odba_seqset_destroy(odba_seqset* s) {
  if (!s) return ODBA_OK;
  if (s->order) hipFree(s->order);
  if (s->small_tab) hipFree(s->small_tab);
  if (s->data) hipFree(s->data);
  if (s->offs) hipFree(s->offs);
  if (s->lens) hipFree(s->lens);
  delete s;
  return ODBA_OK;
}

}  // extern "C"

namespace {

// Pair schedule of the all-vs-all: one wave per pair (many pairs), or bands
// of column strips swept by workgroups (too few pairs to give every SIMD
// work). A schedule the autotuner chose (or the caller pinned) for sets of
// this shape wins over the built-in rule. ODBA_PAIR_SCHEDULE=wave|strip,
// ODBA_BAND_WAVES and ODBA_BAND_RING override both (tests).
struct PairSchedule {
  bool strips = false;
  bool small = false;  // lock-step strips of small pairs (C = 1, 2 or 4, `waves` waves per pair,
                       // sub-blocks of `ring` = 16 or 32 steps)
  int order = 0;
  int C = 16, waves = 4, nblocks = 1, ring = 2048;
  // mixed: pairs of more than mixed_cells cells run first as bands of
  // strips, the rest one wave per pair, largest first
  bool mixed = false;
  long long mixed_cells = 0;
};

int env_int(const char* name, int dflt) {
  const char* v = std::getenv(name);
  return (v && *v) ? std::atoi(v) : dflt;
}

int set_minlen(const odba_seqset* s) {
  int minlen = s->h_lens.empty() ? 0 : s->h_lens[0];
  for (int L : s->h_lens) minlen = std::min(minlen, L);
  return minlen;
}

bool bands_eligible(const odba_seqset* s) {
  return (s->dtype == ODBA_F32 || ((s->dtype == ODBA_I32 || s->dtype == ODBA_U32) && s->int_no_wrap)) &&
         set_minlen(s) >= 64;
}

// The lock-step small-pair kernel: float, every sequence 64..1024 samples
// (one swath, so no swath-width rule of the open end applies).
bool small_eligible(const odba_seqset* s) {
  return s->dtype == ODBA_F32 && set_minlen(s) >= 64 && s->maxlen <= ODBA_W;
}

// Built-in small-pair shape: columns per lane and steps per sub-block.
constexpr int kSmallC = 4, kSmallB = 32;

// Waves per pair of the small-pair kernel at C columns per lane.
int small_waves(const odba_seqset* s, int C) { return std::max(1, (s->maxlen + 64 * C - 1) / (64 * C)); }

std::string shape_key(const odba_seqset* s, int os, int oe) {
  long long sum = 0;
  for (int L : s->h_lens) sum += L;
  char b[160];
  std::snprintf(b, sizeof(b), "%d/%d%d/%d/%d/%d/%lld", s->dtype, os, oe, s->n, s->maxlen, set_minlen(s), sum);
  return b;
}

PairSchedule from_public(odba_ctx* c, const odba_schedule& u, long long npairs) {
  PairSchedule ps;
  ps.strips = u.kind == ODBA_SCHED_BANDS;
  ps.small = u.kind == ODBA_SCHED_SMALL;
  ps.order = u.order;
  ps.C = u.band_c;
  ps.waves = u.band_waves;
  ps.ring = ps.small && u.ring == 0 ? 16 : u.ring;
  // small: a workgroup per pair, or, for more pairs than that holds, three
  // persistent workgroups per CU (the LDS of one is about 41 KB)
  if (ps.small) ps.nblocks = npairs <= 64LL * c->cus ? (int)std::max(1LL, npairs) : 3 * c->cus;
  else if (!ps.strips) ps.nblocks = grid_blocks(c, npairs);
  return ps;
}

odba_schedule to_public(const PairSchedule& ps) {
  odba_schedule u;
  u.kind = ps.small ? ODBA_SCHED_SMALL : (ps.strips ? ODBA_SCHED_BANDS : ODBA_SCHED_WAVE);
  u.order = ps.order;
  u.band_c = ps.C;
  u.band_waves = ps.waves;
  u.ring = ps.ring;
  return u;
}

PairSchedule choose_schedule(odba_ctx* c, const odba_seqset* s, int os, int oe, long long npairs) {
  PairSchedule ps;
  const bool eligible = bands_eligible(s);
  auto t = c->tuned.find(shape_key(s, os, oe));
  const bool tuned_ok = t != c->tuned.end() &&
                        (t->second.kind == ODBA_SCHED_WAVE || (t->second.kind == ODBA_SCHED_BANDS && eligible) ||
                         (t->second.kind == ODBA_SCHED_SMALL && small_eligible(s)));
  if (tuned_ok) {
    ps = from_public(c, t->second, npairs);
  } else if (small_eligible(s) && npairs < 4LL * c->cus) {
    // Fewer pairs than SIMDs, none longer than one swath: one wave per pair
    // leaves each pair's sweep as the latency of a single wave; strips of
    // 256 columns in lock step cut it, largest pair first.
    odba_schedule u{ODBA_SCHED_SMALL, 1, kSmallC, small_waves(s, kSmallC), kSmallB};
    ps = from_public(c, u, npairs);
  } else {
    // Fewer pairs than one wave per SIMD, and long enough that the strip
    // pipeline's fill (about 128 rows per strip) is small against |Y|.
    ps.strips = eligible && npairs < 4LL * c->cus && set_minlen(s) >= 4096;
    if (ps.strips) {
      ps.C = s->maxlen <= 16 * 64 ? 1 : (s->maxlen <= 16 * 256 ? 4 : 16);
      const long long smax = (s->maxlen + 64LL * ps.C - 1) / (64LL * ps.C);
      ps.waves = (int)std::min<long long>(16, std::max<long long>(1, smax));
    } else {
      ps.nblocks = grid_blocks(c, npairs);
    }
  }
  const char* env = std::getenv("ODBA_PAIR_SCHEDULE");
  // A few pairs far larger than the rest: one wave per pair would leave the
  // largest sweeping alone long after the others are done (a 45,000 x 48,000
  // pair is 47 swaths on one wave). A pair of one share (all cells over the
  // resident waves) ends about when the whole launch could; when the largest
  // pair holds more than that, the pairs above half a share run first as
  // bands over the whole device and the rest one wave per pair (largest
  // first). With the open-end stop (open end without open start) the bands
  // sweep every strip while a wave stops at the stop swath, so there the
  // largest pair must hold twice a share and the bands take the pairs
  // above one share.
  if (env && !std::strcmp(env, "small") && small_eligible(s)) {
    odba_schedule u{ODBA_SCHED_SMALL, 1, kSmallC, small_waves(s, kSmallC), kSmallB};
    ps = from_public(c, u, npairs);
  }
  if (env && std::strcmp(env, "small") && ps.small) {  // another schedule asked for
    ps = PairSchedule();
    ps.nblocks = grid_blocks(c, npairs);
  }
  if (!ps.strips && !ps.small && (!tuned_ok || (env && !std::strcmp(env, "mixed"))) && eligible &&
      !(env && std::strcmp(env, "mixed"))) {
    long double sum = 0, sq = 0;
    std::vector<int> L(s->h_lens.begin(), s->h_lens.end());
    for (int v : L) { sum += v; sq += (long double)v * v; }
    std::sort(L.begin(), L.end());
    const long double total = (sum * sum - sq) / 2;
    const long double share = total / (16.0L * c->cus);  // 4 waves per SIMD resident
    const long double largest = L.size() >= 2 ? (long double)L[L.size() - 1] * L[L.size() - 2] : 0;
    const long double f = (oe && !os) ? 2.0L : 1.0L;
    if ((largest > f * share && largest >= (long double)ODBA_W * 4096) || (env && !std::strcmp(env, "mixed"))) {
      ps.mixed = true;
      ps.mixed_cells = (long long)std::max<long double>(f * share / 2, 1.0L);
    }
  }
  if (env && !std::strcmp(env, "wave") && ps.strips) {
    ps = PairSchedule();
    ps.nblocks = grid_blocks(c, npairs);
  }
  if (env && !std::strcmp(env, "strip") && !ps.strips && eligible) {
    ps.strips = true;
    ps.C = s->maxlen <= 16 * 64 ? 1 : (s->maxlen <= 16 * 256 ? 4 : 16);
    const long long smax = (s->maxlen + 64LL * ps.C - 1) / (64LL * ps.C);
    ps.waves = (int)std::min<long long>(16, std::max<long long>(1, smax));
  }
  if (ps.strips) {
    ps.waves = std::max(1, std::min(16, env_int("ODBA_BAND_WAVES", ps.waves)));
    ps.ring = env_int("ODBA_BAND_RING", ps.ring);
  }
  return ps;
}

// Pair order with the largest cell count first (ties: the built-in order),
// built once per set.
int ensure_pair_order(odba_ctx* c, odba_seqset* s, long long npairs) {
  if (s->order) return ODBA_OK;
  std::vector<int2> ord;
  ord.reserve((size_t)npairs);
  for (int cc = s->n - 1; cc >= 1; cc--)
    for (int r = cc - 1; r >= 0; r--) ord.push_back(make_int2(r, cc));
  std::stable_sort(ord.begin(), ord.end(), [&](const int2& a, const int2& b) {
    return (long long)s->h_lens[a.x] * s->h_lens[a.y] > (long long)s->h_lens[b.x] * s->h_lens[b.y];
  });
  if (s->borrowed) s->order = ctx_buffer(c, kGbOrder, ord.size() * sizeof(int2));
  else if (hipMalloc(&s->order, ord.size() * sizeof(int2)) != hipSuccess) s->order = nullptr;
  if (!s->order) return fail(c, ODBA_E_NOMEM, "cannot allocate the pair order");
  HIPC(c, hipMemcpyAsync(s->order, ord.data(), ord.size() * sizeof(int2), hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return ODBA_OK;
}

// Pair records of the small-pair kernel in launch order (order 0: the
// one-wave kernel's, second index descending then first index descending;
// order 1: largest cell count first), built once per set and order.
int ensure_small_tab(odba_ctx* c, odba_seqset* s, int order, long long npairs) {
  if (s->small_tab && s->small_tab_order == order) return ODBA_OK;
  std::vector<int2> ord;
  ord.reserve((size_t)npairs);
  for (int cc = s->n - 1; cc >= 1; cc--)
    for (int r = cc - 1; r >= 0; r--) ord.push_back(make_int2(r, cc));
  if (order == 1)
    std::stable_sort(ord.begin(), ord.end(), [&](const int2& a, const int2& b) {
      return (long long)s->h_lens[a.x] * s->h_lens[a.y] > (long long)s->h_lens[b.x] * s->h_lens[b.y];
    });
  std::vector<SmallPairRec> tab(ord.size());
  const long long n = s->n;
  for (size_t i = 0; i < ord.size(); i++) {
    const int r = ord[i].x, cc = ord[i].y;
    tab[i] = SmallPairRec{s->h_offs[r], s->h_offs[cc], (long long)r * (2 * n - r - 1) / 2 + (cc - r - 1),
                          s->h_lens[r], s->h_lens[cc]};
  }
  if (s->borrowed) {
    s->small_tab = (SmallPairRec*)ctx_buffer(c, kGbSmallTab, tab.size() * sizeof(SmallPairRec));
    if (!s->small_tab) return fail(c, ODBA_E_NOMEM, "cannot allocate the pair records");
  } else if (!s->small_tab && hipMalloc(&s->small_tab, tab.size() * sizeof(SmallPairRec)) != hipSuccess) {
    s->small_tab = nullptr;
    return fail(c, ODBA_E_NOMEM, "cannot allocate the pair records");
  }
  HIPC(c, hipMemcpyAsync(s->small_tab, tab.data(), tab.size() * sizeof(SmallPairRec), hipMemcpyHostToDevice,
                         c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  s->small_tab_order = order;
  return ODBA_OK;
}

// Bands of strips (k_all_pairs_bands): claim table in the one-wave kernel's
// pair order, scratch, launch, hand-off check.
int all_pairs_bands(odba_ctx* c, odba_seqset* s, const PairSchedule& ps, int os, int oe, long long npairs, void* dist,
                    int* flags, const std::vector<int2>* pairs = nullptr, bool events = true) {
  const int n = s->n;
  const long long SW = 64LL * ps.C;
  struct Tab { int r, c; long long band0; };
  static_assert(sizeof(Tab) == 16, "claim table entry");
  if (band_pair_size() != sizeof(Tab)) return fail(c, ODBA_E_INVALID, "claim table layout mismatch");
  std::vector<Tab> tab;
  tab.reserve(pairs ? pairs->size() : (size_t)npairs);
  long long nclaims = 0;
  auto add = [&](int r, int cc) {
    tab.push_back(Tab{r, cc, nclaims});
    const long long S = ((long long)s->h_lens[cc] + SW - 1) / SW;
    nclaims += (S + ps.waves - 1) / ps.waves;
  };
  if (pairs) {
    for (const int2& pr : *pairs) add(pr.x, pr.y);
  } else {
    for (int cc = n - 1; cc >= 1; cc--)
      for (int r = cc - 1; r >= 0; r--) add(r, cc);
  }
  const long long per_cu = std::max(1, 16 / ps.waves);
  const int nblocks = (int)std::max<long long>(1, std::min<long long>(nclaims, per_cu * c->cus));
  const int R = 2 * nblocks + 2;
  const size_t es = dtype_size(s->dtype);
  const long long gstride = s->maxlen + 256;
  const size_t ring_elems = (size_t)nblocks * ps.waves * ps.ring;
  int rc = ensure_scratch(c, (ring_elems + (size_t)R * gstride) * es);
  if (rc) return rc;
  void* rings = c->esc;
  void* gslots = (char*)c->esc + ring_elems * es;
  DevBufs B;
  int* err;
  unsigned char *dtab, *sync;
  if ((rc = B.get(c, err, 1)) || (rc = B.get(c, dtab, tab.size() * sizeof(Tab))) ||
      (rc = B.get(c, sync, band_sync_bytes(R, nclaims, (int)tab.size()))))
    return rc;
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device) != hipSuccess || khz <= 0) khz = 100000;
  const long long limit = (long long)env_int("ODBA_BAND_WAIT_S", 60) * 1000LL * khz;  // per wait
  HIPC(c, hipMemcpyAsync(dtab, tab.data(), tab.size() * sizeof(Tab), hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemsetAsync(err, 0, sizeof(int), c->stream));
  unsigned long long* counter = (unsigned long long*)c->small;
  HIPC(c, hipMemsetAsync(counter, 0, sizeof(unsigned long long), c->stream));
  if (events) HIPC(c, hipEventRecord(c->ev0, c->stream));
  hipError_t e = launch_all_pairs_bands(s->dtype, ps.C, ps.waves, os, oe, nblocks, c->stream, s->data, s->offs, s->lens,
                                        n, dtab, (int)tab.size(), nclaims, counter, dist, rings, ps.ring, gslots, gstride,
                                        sync, R, err, limit, flags);
  if (e != hipSuccess) return fail(c, ODBA_E_HIP, "band launch: %s", hipGetErrorString(e));
  if (events) HIPC(c, hipEventRecord(c->ev1, c->stream));
  int herr = 0;
  HIPC(c, hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  if (herr & 8) return fail(c, ODBA_E_HIP, "band edge hand-off timed out");
  return ODBA_OK;
}

// One all-vs-all launch under a given schedule into device buffers; the
// kernel time (HIP events on the launch stream) in *ms.
int run_all_pairs(odba_ctx* c, odba_seqset* s, const PairSchedule& ps, int os, int oe, long long npairs, void* dist,
                  int* flags, double* ms) {
  int rc;
  if (ps.mixed) {
    // the pairs above the threshold as bands of strips, then the rest one
    // wave per pair, largest first; one event pair brackets both launches
    std::vector<int2> big, rest;
    rest.reserve((size_t)npairs);
    int maxx = 0;
    for (int cc = s->n - 1; cc >= 1; cc--)
      for (int r = cc - 1; r >= 0; r--) {
        if ((long long)s->h_lens[r] * s->h_lens[cc] > ps.mixed_cells) {
          big.push_back(make_int2(r, cc));
          maxx = std::max(maxx, s->h_lens[cc]);
        } else {
          rest.push_back(make_int2(r, cc));
        }
      }
    std::stable_sort(rest.begin(), rest.end(), [&](const int2& a, const int2& b) {
      return (long long)s->h_lens[a.x] * s->h_lens[a.y] > (long long)s->h_lens[b.x] * s->h_lens[b.y];
    });
    PairSchedule bs;
    bs.strips = true;
    bs.C = maxx <= 16 * 64 ? 1 : (maxx <= 16 * 256 ? 4 : 16);
    const long long smax = (maxx + 64LL * bs.C - 1) / (64LL * bs.C);
    bs.waves = (int)std::min<long long>(16, std::max<long long>(1, smax));
    bs.ring = std::max(128, env_int("ODBA_BAND_RING", 2048));
    HIPC(c, hipEventRecord(c->ev0, c->stream));
    if (!big.empty() && (rc = all_pairs_bands(c, s, bs, os, oe, npairs, dist, flags, &big, false))) return rc;
    if (!rest.empty()) {
      DevBufs B;
      int2* ord = nullptr;
      if ((rc = B.get(c, ord, rest.size()))) return rc;
      HIPC(c, hipMemcpyAsync(ord, rest.data(), rest.size() * sizeof(int2), hipMemcpyHostToDevice, c->stream));
      const long long stride = s->maxlen + 256;
      const int nb = grid_blocks(c, (long long)rest.size());
      rc = ensure_scratch(c, (size_t)nb * 4 * 2 * stride * dtype_size(s->dtype));
      if (rc) return rc;
      unsigned long long* counter = (unsigned long long*)c->small;
      HIPC(c, hipMemsetAsync(counter, 0, sizeof(unsigned long long), c->stream));
      hipError_t e = launch_all_pairs(s->dtype, s->int_no_wrap, os, oe, nb, c->stream, s->data, s->offs, s->lens, s->n,
                                      (long long)rest.size(), counter, dist, c->esc, stride, flags, ord);
      if (e != hipSuccess) return fail(c, ODBA_E_HIP, "all-pairs launch: %s", hipGetErrorString(e));
      HIPC(c, hipEventRecord(c->ev1, c->stream));
      HIPC(c, hipStreamSynchronize(c->stream));  // before B frees the order
    } else {
      HIPC(c, hipEventRecord(c->ev1, c->stream));
      HIPC(c, hipStreamSynchronize(c->stream));
    }
  } else if (ps.small) {
    if ((rc = ensure_small_tab(c, s, ps.order, npairs))) return rc;
    unsigned long long* counter = nullptr;  // one workgroup per pair: no claims
    if (ps.nblocks < npairs) {
      counter = (unsigned long long*)c->small;
      HIPC(c, hipMemsetAsync(counter, 0, sizeof(unsigned long long), c->stream));
    }
    HIPC(c, hipEventRecord(c->ev0, c->stream));
    hipError_t e = launch_all_pairs_small(ps.C, ps.ring, ps.waves, os, oe, ps.nblocks, c->stream,
                                          (const float*)s->data, s->small_tab, npairs, counter, (float*)dist, flags);
    if (e != hipSuccess) return fail(c, ODBA_E_HIP, "small-pair launch: %s", hipGetErrorString(e));
    HIPC(c, hipEventRecord(c->ev1, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
  } else if (ps.strips) {
    rc = all_pairs_bands(c, s, ps, os, oe, npairs, dist, flags);
    if (rc) return rc;
  } else {
    const long long stride = s->maxlen + 256;
    rc = ensure_scratch(c, (size_t)ps.nblocks * 4 * 2 * stride * dtype_size(s->dtype));
    if (rc) return rc;
    if (ps.order == 1 && (rc = ensure_pair_order(c, s, npairs))) return rc;
    unsigned long long* counter = (unsigned long long*)c->small;
    HIPC(c, hipMemsetAsync(counter, 0, sizeof(unsigned long long), c->stream));
    HIPC(c, hipEventRecord(c->ev0, c->stream));
    hipError_t e = launch_all_pairs(s->dtype, s->int_no_wrap, os, oe, ps.nblocks, c->stream, s->data, s->offs, s->lens,
                                    s->n, npairs, counter, dist, c->esc, stride, flags, ps.order == 1 ? s->order : nullptr);
    if (e != hipSuccess) return fail(c, ODBA_E_HIP, "all-pairs launch: %s", hipGetErrorString(e));
    HIPC(c, hipEventRecord(c->ev1, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
  }
  if (ms) {
    float f = 0;
    hipEventElapsedTime(&f, c->ev0, c->ev1);
    *ms = f;
  }
  return ODBA_OK;
}

}  // namespace

extern "C" {

int odba_ctx_set_graphs(odba_ctx* c, int on) {
  if (!c) return ODBA_E_INVALID;
  c->graphs = on != 0;
  return ODBA_OK;
}

int odba_ctx_set_mem_cap(odba_ctx* c, uint64_t bytes) {
  if (!c || bytes == 0) return ODBA_E_INVALID;
  c->mem_cap = std::min<uint64_t>(bytes, c->mem);
  return ODBA_OK;
}

int odba_dtw_all_pairs_resident(odba_ctx* c, odba_seqset* s, int os, int oe, void* dist_out, int* flags_out,
                                double* kernel_ms) {
  if (!c || !s || s->ctx != c) return fail(c, ODBA_E_INVALID, "invalid sequence set");
  HIPC(c, hipSetDevice(c->device));
  const int n = s->n;
  if (n < 2) return fail(c, ODBA_E_NOT_ENOUGH_SEQUENCES, "need at least two sequences");
  const long long npairs = (long long)n * (n - 1) / 2;
  const size_t es = dtype_size(s->dtype);
  const PairSchedule ps = choose_schedule(c, s, os, oe, npairs);
  void* dist = ctx_buffer(c, kGbDist, (size_t)npairs * es);
  if (!dist) return fail(c, ODBA_E_NOMEM, "cannot allocate %lld distances", npairs);
  int* flags = nullptr;
  if (flags_out && !(flags = (int*)ctx_buffer(c, kGbFlags, (size_t)npairs * sizeof(int))))
    return fail(c, ODBA_E_NOMEM, "cannot allocate flags");
  int rc = run_all_pairs(c, s, ps, os, oe, npairs, dist, flags, kernel_ms);
  // copies on the context's stream: a plain hipMemcpy runs on the null
  // stream and would wait for every other context's work on the device
  if (!rc && dist_out && hipMemcpyAsync(dist_out, dist, npairs * es, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
    rc = fail(c, ODBA_E_HIP, "copying distances");
  if (!rc && flags_out && hipMemcpyAsync(flags_out, flags, npairs * sizeof(int), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
    rc = fail(c, ODBA_E_HIP, "copying flags");
  if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(c, ODBA_E_HIP, "stream synchronisation");
  return rc;
}

int odba_dtw_all_pairs_noalloc(odba_ctx* c, odba_seqset* s, int os, int oe, void* dist_dev, int* flags_dev) {
  if (!c || !s || s->ctx != c) return fail(c, ODBA_E_INVALID, "invalid sequence set");
  if (!dist_dev) return fail(c, ODBA_E_INVALID, "no distance buffer");
  HIPC(c, hipSetDevice(c->device));
  const int n = s->n;
  if (n < 2) return fail(c, ODBA_E_NOT_ENOUGH_SEQUENCES, "need at least two sequences");
  void* base = c->device < kMaxDevices ? g_scratch_dev[c->device] : nullptr;
  if (!base) return fail(c, ODBA_E_HIP, "the fixed scratch of device %d is unavailable", c->device);
  const long long npairs = (long long)n * (n - 1) / 2;
  // one wave per pair, as many resident waves as the fixed scratch holds
  // (two edge columns of stride values per wave, four waves per workgroup)
  const long long stride = s->maxlen + 256;
  const size_t per_block = (size_t)4 * 2 * stride * dtype_size(s->dtype);
  const size_t avail = (size_t)ODBA_NOALLOC_SCRATCH_BYTES - 256;
  if (per_block > avail)
    return fail(c, ODBA_E_INVALID,
                "longest sequence (%d samples) needs %zu bytes of edge columns per workgroup; the fixed scratch holds %zu",
                s->maxlen, per_block, avail);
  const int nb = (int)std::min<long long>(grid_blocks(c, npairs), (long long)(avail / per_block));
  unsigned long long* counter = (unsigned long long*)base;
  // cleared by a kernel, not a memset node, so the call can sit in a graph capture
  HIPC(c, launch_fill_u32(c->stream, counter, 0u, 2));
  hipError_t e = launch_all_pairs(s->dtype, s->int_no_wrap, os, oe, nb, c->stream, s->data, s->offs, s->lens, n, npairs,
                                  counter, dist_dev, (char*)base + 256, stride, flags_dev, nullptr);
  if (e != hipSuccess) return fail(c, ODBA_E_HIP, "all-pairs launch: %s", hipGetErrorString(e));
  HIPC(c, hipStreamSynchronize(c->stream));
  return ODBA_OK;
}

int odba_tune_all_pairs(odba_ctx* c, odba_seqset* s, int os, int oe, int reps, odba_schedule* best, double* best_ms,
                        int* tried) {
  if (!c || !s || s->ctx != c) return fail(c, ODBA_E_INVALID, "invalid sequence set");
  if (s->n < 2) return fail(c, ODBA_E_NOT_ENOUGH_SEQUENCES, "need at least two sequences");
  HIPC(c, hipSetDevice(c->device));
  reps = std::max(1, reps);
  const long long npairs = (long long)s->n * (s->n - 1) / 2;
  std::vector<PairSchedule> cand;
  for (int order = 0; order < 2; order++) {
    PairSchedule w;
    w.order = order;
    w.nblocks = grid_blocks(c, npairs);
    cand.push_back(w);
  }
  if (small_eligible(s) && npairs < 16LL * c->cus) {
    for (int C : {1, 2, 4})
      for (int B : {16, 32})
        for (int order = 0; order < 2; order++) {
          if (small_waves(s, C) > 8 || 64 * C * small_waves(s, C) > ODBA_W) continue;
          odba_schedule u{ODBA_SCHED_SMALL, order, C, small_waves(s, C), B};
          cand.push_back(from_public(c, u, npairs));
        }
  }
  if (bands_eligible(s) && npairs < 16LL * c->cus) {
    for (int C : {1, 4, 16}) {
      const long long smax = (s->maxlen + 64LL * C - 1) / (64LL * C);
      if (C > 1 && s->maxlen < 64 * C) continue;
      for (int w : {2, 4, 8, 16}) {
        if (w > std::max<long long>(2, smax)) continue;
        PairSchedule b;
        b.strips = true;
        b.C = C;
        b.waves = w;
        cand.push_back(b);
      }
    }
  }
  void* dist = nullptr;
  if (hipMalloc(&dist, npairs * dtype_size(s->dtype)) != hipSuccess) return fail(c, ODBA_E_NOMEM, "cannot allocate %lld distances", npairs);
  double bms = 1e300;
  PairSchedule bps = cand[0];
  int rc = ODBA_OK;
  for (const PairSchedule& ps : cand) {
    std::vector<double> t;
    for (int k = 0; k < reps && !rc; k++) {
      double ms = 0;
      rc = run_all_pairs(c, s, ps, os, oe, npairs, dist, nullptr, &ms);
      t.push_back(ms);
    }
    if (rc) break;
    std::sort(t.begin(), t.end());
    const double med = t[t.size() / 2];
    if (std::getenv("ODBA_TUNE_VERBOSE"))
      std::fprintf(stderr, "odba tune: %s order %d C %d waves %d ring %d: %.4f ms\n",
                   ps.small ? "small" : (ps.strips ? "bands" : "wave"), ps.order,
                   ps.C, ps.waves, ps.ring, med);
    if (med < bms) { bms = med; bps = ps; }
  }
  hipFree(dist);
  if (rc) return rc;
  c->tuned[shape_key(s, os, oe)] = to_public(bps);
  if (best) *best = to_public(bps);
  if (best_ms) *best_ms = bms;
  if (tried) *tried = (int)cand.size();
  return ODBA_OK;
}

int odba_get_schedule(odba_ctx* c, odba_seqset* s, int os, int oe, odba_schedule* out) {
  if (!c || !s || s->ctx != c || !out) return fail(c, ODBA_E_INVALID, "invalid arguments");
  const long long npairs = (long long)s->n * (s->n - 1) / 2;
  *out = to_public(choose_schedule(c, s, os, oe, npairs));
  return ODBA_OK;
}

int odba_set_schedule(odba_ctx* c, odba_seqset* s, int os, int oe, const odba_schedule* u) {
  if (!c || !s || s->ctx != c) return fail(c, ODBA_E_INVALID, "invalid sequence set");
  const std::string key = shape_key(s, os, oe);
  if (!u) {
    c->tuned.erase(key);
    return ODBA_OK;
  }
  if (u->kind == ODBA_SCHED_WAVE) {
    if (u->order != 0 && u->order != 1) return fail(c, ODBA_E_INVALID, "pair order %d", u->order);
  } else if (u->kind == ODBA_SCHED_BANDS) {
    if (!bands_eligible(s)) return fail(c, ODBA_E_INVALID, "bands need float (or integers that cannot wrap) and sequences of 64+ samples");
    if ((u->band_c != 1 && u->band_c != 4 && u->band_c != 16) || u->band_waves < 1 || u->band_waves > 16 ||
        u->ring < 128 || (u->ring & (u->ring - 1)))
      return fail(c, ODBA_E_INVALID, "band schedule (C %d, waves %d, ring %d)", u->band_c, u->band_waves, u->ring);
  } else if (u->kind == ODBA_SCHED_SMALL) {
    if (!small_eligible(s)) return fail(c, ODBA_E_INVALID, "the small-pair schedule needs float sequences of 64..1024 samples");
    if ((u->band_c != 1 && u->band_c != 2 && u->band_c != 4) || u->band_waves != small_waves(s, u->band_c) ||
        u->band_waves > 8 || 64 * u->band_c * u->band_waves > ODBA_W || (u->order != 0 && u->order != 1) ||
        (u->ring != 0 && u->ring != 16 && u->ring != 32))
      return fail(c, ODBA_E_INVALID, "small-pair schedule (C %d, waves %d, order %d, sub-block %d)", u->band_c,
                  u->band_waves, u->order, u->ring);
  } else {
    return fail(c, ODBA_E_INVALID, "schedule kind %d", u->kind);
  }
  c->tuned[key] = *u;
  return ODBA_OK;
}

int odba_dtw_all_pairs(odba_ctx* c, int dtype, const void* const* seqs, const uint64_t* lens, int n, int os, int oe,
                       int ref_w, void* dist_out) {
  if (!c) return ODBA_E_INVALID;
  if (ref_w != ODBA_W) return fail(c, ODBA_E_UNSUPPORTED, "swath width %d: only %d is supported", ref_w, ODBA_W);
  if (n < 2) return fail(c, ODBA_E_NOT_ENOUGH_SEQUENCES, "need at least two sequences, got %d", n);
  odba_seqset* s = nullptr;
  int rc = odba_seqset_create(c, dtype, seqs, lens, n, &s);
  if (rc) return rc;
  rc = odba_dtw_all_pairs_resident(c, s, os, oe, dist_out, nullptr, nullptr);
  odba_seqset_destroy(s);
  return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Path-mode machinery shared by odba_dtw_path and the DBA round.
// ---------------------------------------------------------------------------
namespace {

struct PathPlan {
  std::vector<unsigned char> jobs;  // PathJob<T> records
  std::vector<TbJob> tb;
  long long mv_words = 0, top_words = 0;
};

template <typename T> struct PJ {
  const T* y;
  const T* x;
  int ly, lx;
  long long mv_off, top_off;
};

// Geometry of one path job; returns words needed.
void plan_job(int ly, int lx, int os, int oe, long long& mv_words, long long& top_words, long long& mv_off,
              long long& top_off) {
  const int K = (lx + ODBA_W - 1) / ODBA_W;
  mv_off = mv_words;
  top_off = top_words;
  mv_words += (long long)K * swath_words(ly, os, oe);
  top_words += (long long)K * 64;
}

// Mean swath count (1024-column swaths of X) over a list of path jobs.
template <typename J> double mean_swaths(const std::vector<J>& jobs) {
  if (jobs.empty()) return 0.0;
  double t = 0;
  for (const J& j : jobs) t += (j.lx + ODBA_W - 1) / ODBA_W;
  return t / (double)jobs.size();
}

// Longest job's row-swaths over the mean (1: even jobs). One wave per job
// ends when its longest job does, so a large ratio means a tail.
template <typename J> double path_tail(const std::vector<J>& jobs) {
  if (jobs.empty()) return 0.0;
  double t = 0, mx = 0;
  for (const J& j : jobs) {
    const double w = (double)j.ly * (double)((j.lx + ODBA_W - 1) / ODBA_W);
    t += w;
    mx = std::max(mx, w);
  }
  return t > 0 ? mx / (t / (double)jobs.size()) : 0.0;
}

// Path jobs: one wave per job, or (float, few long jobs) a workgroup of
// waves per job sweeping its swaths as a pipeline. ODBA_PATH_SCHEDULE =
// wave|strip overrides the choice.
// bands_per_cu > 0: take the swath-band schedule (float, with an error
// word) whatever the job count, with that many workgroups per CU.
int run_path_jobs(odba_ctx* c, int dtype, bool minfirst, int os, int oe, const void* jobs_dev, int njobs, int maxly,
                  uint32_t* mv, uint32_t* mtop, int* abr, int maxlx = 0, int minly = 0, int* err = nullptr,
                  double mean_swaths = 0.0, double tail = 0.0, int bands_per_cu = 0) {
  const size_t es = dtype_size(dtype);
  const long long stride = maxly + 256;
  unsigned int* counter = (unsigned int*)((char*)c->small + 64);
  HIPC(c, launch_fill_u32(c->stream, counter, 0u, 1));  // a kernel: this may run inside a graph capture
  // Few jobs (float): a workgroup per job, one wave per swath or (pairs of
  // at most four swaths) one wave per quarter swath.
  // ODBA_PATH_SCHEDULE = wave|strip|quarter.
  // The band schedule (launch_path_bands: a job's swaths on several
  // workgroups with global edge slots) is switched off: its first open-end
  // run on the device made an illegal memory access whose cause has not
  // been found, so nothing selects it and asking for it is an error.
  const char* env = std::getenv("ODBA_PATH_SCHEDULE");
  const int kmax = (maxlx + ODBA_W - 1) / ODBA_W;
  // With many jobs a wave per job fills the device, except when the jobs
  // average eight or more swaths each (a long centroid on the columns):
  // then the longest runs alone on one SIMD for the whole round, and the
  // workgroup-per-job schedule (longest first, its waves busy) is faster.
  // Uneven jobs: the longest holds at least twice the mean work, so one
  // wave per job would end with that job alone on its SIMD (2,000 squiggles
  // of 2k-20k against a 4,603-sample centroid: 41.5 ms per round as one wave
  // per job, 19.7 ms as strips).
  bool strips = (njobs < 4 * c->cus || mean_swaths >= 8.0 || tail >= 2.0) && maxlx > 256 && minly >= 64;
  bool quarters = kmax <= 4;
  const bool bands = false;
  // Few jobs (at most 1,024) of which the longest has more swaths than a
  // strip workgroup has waves (16): its swaths would run one after another
  // on one CU. Bands of whole swaths spread them over several workgroups.
  bool swath_bands = njobs < 4 * c->cus && njobs <= 1024 && kmax > 16 && minly >= 64;
  // Open end without open start, any number of jobs wider than 16 swaths:
  // the second-phase claims leave most bands past each job's stop with
  // nothing to sweep, which strips (pipelined inside one workgroup) cannot
  // do. 2,000 squiggles against an 18-swath centroid: 43.4 ms per round at
  // four bands per CU, 45.9 ms as strips.
  const bool many_open_end = !swath_bands && oe && !os && kmax > 16 && minly >= 64;
  if (many_open_end) swath_bands = true;
  if (bands_per_cu > 0) swath_bands = maxlx > 0 && minly >= 64;
  if (env && std::strcmp(env, "swathband")) swath_bands = false;
  if (env && !std::strcmp(env, "swathband")) swath_bands = maxlx > 0;
  swath_bands = swath_bands && dtype == ODBA_F32 && err != nullptr;  // err: the hand-off time-out word
  if (env && !std::strcmp(env, "band"))
    return fail(c, ODBA_E_INVALID, "the path band schedule is switched off (an open-end run faulted; cause not found)");
  if (env && !std::strcmp(env, "wave")) strips = false;
  if (env && !std::strcmp(env, "strip")) strips = maxlx > 0 && minly >= 64, quarters = false;
  if (env && !std::strcmp(env, "quarter")) strips = maxlx > 0 && minly >= 64, quarters = true;
  strips = strips && dtype == ODBA_F32;
  (void)err;
  hipError_t e;
  if (swath_bands) {
    const int nw = std::max(1, std::min(kPathStripWaves, env_int("ODBA_PATH_BAND_WAVES", 4)));
    const int per_cu =
        std::max(1, bands_per_cu > 0 ? bands_per_cu : env_int("ODBA_PATH_BANDS_PER_CU", many_open_end ? 4 : 2));
    const long long max_claims = (long long)njobs * ((kmax + nw - 1) / nw);
    const int nblocks = (int)std::max(1LL, std::min(max_claims, (long long)per_cu * c->cus));
    const int R = 2 * nblocks + 2;
    const int ring = std::max(128, env_int("ODBA_BAND_RING", 2048));
    const long long gstride = maxly + 256;
    const size_t ring_bytes = (size_t)nblocks * nw * ring * es, slot_bytes = (size_t)R * gstride * es;
    // open end without open start: one boundary edge column per job
    const size_t bnd_bytes = (oe && !os) ? (size_t)njobs * gstride * es : 0;
    const size_t sync_bytes = path_band_sync_bytes(R, max_claims, njobs);
    int rc = ensure_scratch(c, ring_bytes + slot_bytes + bnd_bytes + sync_bytes + 256);
    if (rc) return rc;
    char* base = (char*)c->esc;
    void* bslots = bnd_bytes ? (void*)(base + ring_bytes + slot_bytes) : nullptr;
    void* sync = (void*)(((uintptr_t)(base + ring_bytes + slot_bytes + bnd_bytes) + 255) & ~(uintptr_t)255);
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device) != hipSuccess || khz <= 0) khz = 100000;
    const long long limit = (long long)env_int("ODBA_BAND_WAIT_S", 60) * 1000LL * khz;
    unsigned long long* counter64 = (unsigned long long*)((char*)c->small + 64);
    HIPC(c, launch_fill_u32(c->stream, (unsigned int*)counter64, 0u, 2));
    e = launch_path_swath_bands(nw, os, oe, nblocks, c->stream, jobs_dev, njobs, max_claims, counter64, mv, mtop, abr,
                                base, ring, base + ring_bytes, gstride, sync, R, err, limit, bslots);
  } else if (bands) {
    const long long max_claims = (long long)njobs * std::max(1, kmax);
    const int nblocks = (int)std::max(1LL, std::min(max_claims, 4LL * c->cus));
    const int R = 2 * nblocks + 2;
    const int ring = std::max(128, env_int("ODBA_BAND_RING", 2048));
    const long long gstride = maxly + 256;
    const size_t ring_bytes = (size_t)nblocks * 4 * ring * es, slot_bytes = (size_t)R * gstride * es;
    const size_t sync_bytes = path_band_sync_bytes(R, max_claims, njobs);
    int rc = ensure_scratch(c, ring_bytes + slot_bytes + sync_bytes + 256);
    if (rc) return rc;
    char* base = (char*)c->esc;
    void* sync = (void*)(((uintptr_t)(base + ring_bytes + slot_bytes) + 255) & ~(uintptr_t)255);
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device) != hipSuccess || khz <= 0) khz = 100000;
    const long long limit = (long long)env_int("ODBA_BAND_WAIT_S", 60) * 1000LL * khz;
    unsigned long long* counter64 = (unsigned long long*)((char*)c->small + 64);
    HIPC(c, launch_fill_u32(c->stream, (unsigned int*)counter64, 0u, 2));
    e = launch_path_bands(os, oe, nblocks, c->stream, jobs_dev, njobs, max_claims, counter64, mv, mtop, abr, base, ring,
                          base + ring_bytes, gstride, sync, R, err, limit);
  } else if (strips) {
    const int waves = std::min(kPathStripWaves, quarters ? 4 * kmax : kmax);
    const int per_cu = std::max(1, kPathStripWaves / waves);
    const int nblocks = std::max(1, std::min(njobs, per_cu * c->cus));
    int rc = ensure_scratch(c, (size_t)nblocks * 2 * waves * stride * es);
    if (rc) return rc;
    e = launch_path_strips(waves, os, oe, nblocks, c->stream, jobs_dev, njobs, counter, mv, mtop, abr, c->esc, stride,
                           quarters ? 1 : 0);
  } else {
    const int nblocks = grid_blocks(c, njobs);
    int rc = ensure_scratch(c, (size_t)nblocks * 4 * 2 * stride * es);
    if (rc) return rc;
    e = launch_path_jobs(dtype, minfirst, os, oe, nblocks, c->stream, jobs_dev, njobs, counter, mv, mtop, abr, c->esc,
                         stride);
  }
  if (e != hipSuccess) return fail(c, ODBA_E_HIP, "path launch: %s", hipGetErrorString(e));
  return ODBA_OK;
}

template <typename T>
int dtw_path_t(odba_ctx* c, const T* y, int ly, const T* x, int lx, int os, int oe, uint8_t* moves_out, T* dist_out) {
  // upload both sequences padded
  const long long py = padded_len(ly), px = padded_len(lx);
  std::vector<T> hb((size_t)(py + px), T(0));
  std::memcpy(hb.data(), y, ly * sizeof(T));
  std::memcpy(hb.data() + py, x, lx * sizeof(T));
  T* d = nullptr;
  HIPC(c, hipMalloc(&d, hb.size() * sizeof(T)));
  HIPC(c, hipMemcpyAsync(d, hb.data(), hb.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
  long long mvw = 0, tw = 0, mo, to;
  plan_job(ly, lx, os, oe, mvw, tw, mo, to);
  PJ<T> job{d, d + py, ly, lx, mo, to};
  void* jd = nullptr;
  uint32_t *mv = nullptr, *mtop = nullptr;
  int* abr = nullptr;
  HIPC(c, hipMalloc(&jd, sizeof(job)));
  HIPC(c, hipMalloc((void**)&mv, std::max(1LL, mvw) * 4));
  HIPC(c, hipMalloc((void**)&mtop, std::max(1LL, tw) * 4));
  HIPC(c, hipMalloc((void**)&abr, 2 * sizeof(int)));  // [abrogated swath, band hand-off error]
  HIPC(c, hipMemsetAsync(abr + 1, 0, sizeof(int), c->stream));
  // Step words of swaths past an open-end stop may stay unwritten; they are
  // filled with code 3 (decoded as RIGHT) so that a read of one would show
  // in the returned step matrix instead of passing on zeroed memory.
  HIPC(c, launch_fill_u32(c->stream, mv, 0xFFFFFFFFu, std::max(1LL, mvw)));
  HIPC(c, launch_fill_u32(c->stream, mtop, 0xFFFFFFFFu, std::max(1LL, tw)));
  HIPC(c, hipMemcpyAsync(jd, &job, sizeof(job), hipMemcpyHostToDevice, c->stream));
  std::vector<const T*> sp{y, x};
  std::vector<uint64_t> sl{(uint64_t)ly, (uint64_t)lx};
  bool minfirst = int_cannot_wrap<T>(sp.data(), sl.data(), 2);
  int rc = run_path_jobs(c, std::is_same<T, float>::value ? ODBA_F32 : std::is_same<T, double>::value ? ODBA_F64
                              : std::is_same<T, int32_t>::value ? ODBA_I32 : std::is_same<T, uint32_t>::value ? ODBA_U32 : ODBA_U64,
                         minfirst, os, oe, jd, 1, ly, mv, mtop, abr, lx, ly, abr + 1);
  if (rc) return rc;
  int ab[2] = {-1, 0};
  HIPC(c, hipMemcpyAsync(ab, abr, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  if (ab[1] & 8) return fail(c, ODBA_E_HIP, "path band edge hand-off timed out");
  if (moves_out) {
    uint8_t* md = nullptr;
    HIPC(c, hipMalloc((void**)&md, (size_t)ly * lx));
    HIPC(c, launch_decode_moves(c->stream, mv, mtop, ab[0], ly, lx, os, oe, md));
    HIPC(c, hipMemcpyAsync(moves_out, md, (size_t)ly * lx, hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    hipFree(md);
  }
  hipFree(jd); hipFree(mv); hipFree(mtop); hipFree(abr); hipFree(d);
  if (dist_out) {
    // the distance is computed by the all-pairs machinery for exactness of the
    // abrogation path
    const void* ptrs[2] = {y, x};
    uint64_t lens[2] = {(uint64_t)ly, (uint64_t)lx};
    return odba_dtw_all_pairs(c, std::is_same<T, float>::value ? ODBA_F32 : std::is_same<T, double>::value ? ODBA_F64
                                 : std::is_same<T, int32_t>::value ? ODBA_I32 : std::is_same<T, uint32_t>::value ? ODBA_U32 : ODBA_U64,
                              ptrs, lens, 2, os, oe, ODBA_W, dist_out);
  }
  return ODBA_OK;
}

template <typename T> int dtype_of() {
  return std::is_same<T, float>::value ? ODBA_F32 : std::is_same<T, double>::value ? ODBA_F64
         : std::is_same<T, int32_t>::value ? ODBA_I32 : std::is_same<T, uint32_t>::value ? ODBA_U32 : ODBA_U64;
}

// Device-resident state of a DBA cluster: member sequences (cluster order),
// the centroid, and per-round work buffers.
template <typename T> struct DbaState {
  odba_ctx* c = nullptr;
  int n = 0, os = 0, oe = 0;
  bool minfirst = true;
  T* seqdata = nullptr;          // packed member sequences
  long long* offs = nullptr;     // device offsets (cluster order)
  std::vector<long long> h_offs;
  std::vector<int> h_lens;
  int maxlen = 0;
  int lc = 0;
  T* cen = nullptr;              // padded centroid
  T* newc = nullptr;
  T* acc = nullptr;
  uint32_t* counts = nullptr;
  int4* runs = nullptr;
  size_t runs_cap = 0;
  uint32_t* mv = nullptr;
  size_t mv_cap = 0;
  uint32_t* mtop = nullptr;
  size_t mtop_cap = 0;
  int* abr = nullptr;
  void* jobs = nullptr;
  TbJob* tbj = nullptr;
  int* err = nullptr;
  unsigned long long* dbits = nullptr;
  long long budget_words = 0;    // per-batch move-word budget
  int policy = ODBA_PATH_FULL;
  uint64_t free_mem = 0;
  struct StripeJobH {             // layout of StripeJob<T> in the kernels
    const T* y;
    const T* x;
    int H, LX;
    long long slot_off, mv_off, top_off;
  };
  T* slots = nullptr;
  size_t slots_cap = 0;
  uint32_t* smv = nullptr;
  size_t smv_cap = 0;
  uint32_t* stop = nullptr;
  size_t stop_cap = 0;
  StripeState* sstate = nullptr;
  size_t sstate_cap = 0;
  StripeJobH* sjobs = nullptr;
  size_t sjobs_cap = 0;
  TbJob* stbj = nullptr;
  size_t stbj_cap = 0;
  // Checkpointed global paths (float): row and column checkpoints, block
  // step words, strip slots, rings, jobs and walk states.
  float* ck_rowck = nullptr;
  size_t ck_rowck_cap = 0;
  float* ck_colck = nullptr;
  size_t ck_colck_cap = 0;
  int* ck_cnt = nullptr;
  size_t ck_cnt_cap = 0;
  uint32_t* ck_mv = nullptr;
  size_t ck_mv_cap = 0;
  float* ck_slots = nullptr;
  size_t ck_slots_cap = 0;
  float* ck_rings = nullptr;
  size_t ck_rings_cap = 0;
  unsigned char* ck_jobs = nullptr;
  size_t ck_jobs_cap = 0;
  unsigned char* ck_states = nullptr;
  size_t ck_states_cap = 0;
  // Checkpointed stripes in the open modes (float): whole edge columns,
  // per-swath (top, column min), band-start minima, job records.
  float* sck_slots = nullptr;
  size_t sck_slots_cap = 0;
  float2* sck_tc = nullptr;
  size_t sck_tc_cap = 0;
  float* sck_pm = nullptr;
  size_t sck_pm_cap = 0;
  SckJob* sck_jobs = nullptr;
  size_t sck_jobs_cap = 0;
  float* sck_q = nullptr;  // quarter strips' edges over a block, three per job
  size_t sck_q_cap = 0;
  // Traceback recording (the cells writeDTWPath prints): per sequence a
  // count slot then up to |seq| + |centroid| packed steps.
  bool rec = false;
  bool rec_valid = false;
  unsigned long long* pbuf = nullptr;
  size_t pbuf_cap = 0;
  std::vector<long long> h_poff;
  // Graph replay of a round (graph_round): the jobs of the last eager round
  // when it was a single batch of full-matrix paths without recording.
  bool g_ok = false;
  int g_eager = 0;                         // consecutive replayable eager rounds
  std::vector<PJ<T>> g_pj;
  std::vector<char> g_flip;
  int g_maxly = 0, g_maxlx = 0, g_minly = 0;
  T* cbuf[2] = {nullptr, nullptr};         // the two centroid buffers
  void* gjobs[2] = {nullptr, nullptr};     // job tables reading cbuf[p]
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  std::vector<void*> gsig[2];              // device pointers baked into each graph
  int* h_err = nullptr;                    // pinned outputs of a replay (in the context's staging)
  unsigned long long* h_bits = nullptr;
  T* h_cen = nullptr;

  void drop_graphs() {
    for (int p = 0; p < 2; p++) {
      if (gexec[p]) hipGraphExecDestroy(gexec[p]);
      gexec[p] = nullptr;
      gsig[p].clear();
    }
  }

  ~DbaState() {
    drop_graphs();
    for (void* p : gjobs) if (p) hipFree(p);
    if (!own_seqs) seqdata = nullptr, offs = nullptr;
    for (void* p : {(void*)seqdata, (void*)offs, (void*)cen, (void*)newc, (void*)acc, (void*)counts, (void*)runs,
                    (void*)mv, (void*)mtop, (void*)abr, (void*)jobs, (void*)tbj, (void*)err, (void*)dbits,
                    (void*)slots, (void*)smv, (void*)stop, (void*)sstate, (void*)sjobs, (void*)stbj, (void*)pbuf,
                    (void*)ck_rowck, (void*)ck_colck, (void*)ck_cnt, (void*)ck_mv, (void*)ck_slots, (void*)ck_rings,
                    (void*)ck_jobs, (void*)ck_states, (void*)sck_slots, (void*)sck_tc, (void*)sck_pm, (void*)sck_jobs,
                    (void*)sck_q})
      if (p) hipFree(p);
  }

  bool own_seqs = true;

  // Members already resident on the device (padded layout, as a seqset):
  // the session reads them in place and does not free them.
  int init_shared(odba_ctx* ctx, T* d, long long* doffs, const std::vector<long long>& ho,
                  const std::vector<int>& hl, bool no_wrap, int os_, int oe_) {
    c = ctx; n = (int)hl.size(); os = os_; oe = oe_;
    minfirst = no_wrap;
    own_seqs = false;
    seqdata = d;
    offs = doffs;
    h_offs = ho;
    h_lens = hl;
    for (int L : hl) maxlen = std::max(maxlen, L);
    return alloc_common();
  }

  int alloc_common() {
    HIPC(c, hipMalloc((void**)&abr, n * sizeof(int)));
    HIPC(c, hipMalloc(&jobs, n * sizeof(PJ<T>)));
    HIPC(c, hipMalloc((void**)&tbj, n * sizeof(TbJob)));
    HIPC(c, hipMalloc((void**)&err, sizeof(int)));
    HIPC(c, hipMalloc((void**)&dbits, sizeof(unsigned long long)));
    // move-word budget: a quarter of the context's memory, at least one job
    budget_words = (long long)(c->mem_cap / 4 / 4);
    HIPC(c, hipStreamSynchronize(c->stream));
    return ODBA_OK;
  }

  int init(odba_ctx* ctx, const T* const* seqs, const uint64_t* lens, int nn, int os_, int oe_) {
    c = ctx; n = nn; os = os_; oe = oe_;
    minfirst = int_cannot_wrap<T>(seqs, lens, nn);
    h_offs.resize(n);
    h_lens.resize(n);
    long long total = 0;
    for (int i = 0; i < n; i++) {
      h_offs[i] = total;
      h_lens[i] = (int)lens[i];
      maxlen = std::max(maxlen, (int)lens[i]);
      total += padded_len((long long)lens[i]);
    }
    std::vector<T> hb((size_t)total, T(0));
    for (int i = 0; i < n; i++) std::memcpy(&hb[h_offs[i]], seqs[i], lens[i] * sizeof(T));
    HIPC(c, hipMalloc((void**)&seqdata, hb.size() * sizeof(T)));
    HIPC(c, hipMalloc((void**)&offs, n * sizeof(long long)));
    HIPC(c, hipMemcpyAsync(seqdata, hb.data(), hb.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipMemcpyAsync(offs, h_offs.data(), n * sizeof(long long), hipMemcpyHostToDevice, c->stream));
    return alloc_common();
  }

  int set_centroid(const T* h, int len) {
    if (len != lc) {
      drop_graphs();
      g_ok = false;
      if (cen) hipFree(cen);
      if (newc) hipFree(newc);
      if (acc) hipFree(acc);
      if (counts) hipFree(counts);
      cen = newc = acc = nullptr;
      counts = nullptr;
      lc = len;
      const long long pl = padded_len(lc);
      HIPC(c, hipMalloc((void**)&cen, pl * sizeof(T)));
      HIPC(c, hipMalloc((void**)&newc, pl * sizeof(T)));
      HIPC(c, hipMalloc((void**)&acc, lc * sizeof(T)));
      HIPC(c, hipMalloc((void**)&counts, lc * sizeof(uint32_t)));
      HIPC(c, hipMemsetAsync(cen, 0, pl * sizeof(T), c->stream));
      HIPC(c, hipMemsetAsync(newc, 0, pl * sizeof(T), c->stream));
      cbuf[0] = cen;
      cbuf[1] = newc;
    }
    if (h) HIPC(c, hipMemcpyAsync(cen, h, lc * sizeof(T), hipMemcpyHostToDevice, c->stream));
    return ODBA_OK;
  }

  int set_centroid_device(const T* d, int len) {
    int rc = set_centroid(nullptr, len);
    if (rc) return rc;
    HIPC(c, hipMemcpyAsync(cen, d, lc * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
    return ODBA_OK;
  }

  template <typename P> static int grow(odba_ctx* c, P*& p, size_t& cap, size_t need) {
    if (need <= cap) return ODBA_OK;
    if (p) hipFree(p);
    p = nullptr;
    cap = 0;
    if (hipMalloc((void**)&p, need * sizeof(P)) != hipSuccess) return fail(c, ODBA_E_NOMEM, "cannot allocate %zu bytes", need * sizeof(P));
    cap = need;
    return ODBA_OK;
  }

  // One DBA round: new centroid in newc, counts, delta. Swaps cen/newc.
  // Stripe or full path matrix for sequence b, by the reference's rule
  // (dba.hpp:523-548): stripe when the stated free memory is below the
  // 1-byte-per-cell move matrix plus 5% and the edge buffer.
  bool use_stripe(int b) const {
    if (policy == ODBA_PATH_STRIPE) return true;
    if (policy != ODBA_PATH_AUTO) return false;
    const uint64_t ls = (uint64_t)h_lens[b];
    const bool flip = oe && (uint64_t)lc < ls;
    const double need = (double)((flip ? (uint64_t)lc : ls) * sizeof(T)) + (double)ls * (double)lc * 1.05;
    return (double)free_mem < need;
  }

  // In global mode a stripe-mode path equals the full matrix's, so float
  // jobs that take stripe mode use checkpointed block recomputes instead,
  // which scale to million-sample sequences. ODBA_CKPT=0 keeps the stripe
  // recompute; ODBA_CKPT_RSHIFT (rows per band = 2^k) and ODBA_CKPT_EW
  // (1024-column strips per block, 2 or 16) set the block size.
  bool use_ckpt() const {
    const char* e = std::getenv("ODBA_CKPT");
    return std::is_same<T, float>::value && !os && !oe && !(e && !std::strcmp(e, "0"));
  }

  int run_ckpt_batch(const std::vector<int>& idx, int b0) {
    if constexpr (!std::is_same<T, float>::value) {
      return fail(c, ODBA_E_UNSUPPORTED, "checkpointed paths are float only");
    } else {
      const int nj = (int)idx.size();
      if (!nj) return ODBA_OK;
      const int rshift = std::max(6, std::min(20, env_int("ODBA_CKPT_RSHIFT", 14)));
      const int ew = env_int("ODBA_CKPT_EW", 16) == 2 ? 2 : 16;
      const long long R = 1LL << rshift, E = (long long)ew * ODBA_W;
      struct Job {
        const float* y; const float* x; int ly, lx;
        long long rowck_off, colck_off, cnt_off, claim0, runs_off, path_off;
      };
      struct State {
        int i, j, curc, lo1, hi1, lo2, hi2, gap;
        long long np;
        int done, err, bi, pad;
      };
      if (sizeof(Job) != ckpt_job_size() || sizeof(State) != ckpt_state_size())
        return fail(c, ODBA_E_INVALID, "checkpoint record layout mismatch");
      std::vector<Job> jobs(nj);
      std::vector<State> sts(nj);
      long long nrow = 0, ncol = 0, ncnt = 0, nclaims = 0;
      int maxit = 0;
      for (int q = 0; q < nj; q++) {
        const int b = idx[q];
        const int ly = h_lens[b], lx = lc;
        const long long nb = (ly + R - 1) / R, nm = (lx + E - 1) / E;
        jobs[q] = Job{seqdata + h_offs[b], cen, ly, lx, nrow, ncol, ncnt, nclaims, (long long)(b - b0) * lc,
                      rec ? h_poff[b] : -1};
        nrow += (nb - 1) * lx;
        ncol += (nm - 1) * ((long long)ly + 256);
        ncnt += std::max(0LL, nm - 1);
        nclaims += nm;
        maxit = std::max<int>(maxit, (int)(nb + nm + 1));
        sts[q] = State{ly - 1, lx - 1, lx - 1, 1, 0, 1, 0, 0, 0, 0, 0, 0, 0};
      }
      const long long job_words = (long long)ew * (R + 63) * 64;
      const int ring = 2048;
      const int nblocks = (int)std::max<long long>(1, std::min<long long>(nclaims, (long long)std::max(1, 16 / ew) * c->cus));
      int rc;
      if ((rc = grow(c, ck_rowck, ck_rowck_cap, (size_t)std::max(1LL, nrow))) ||
          (rc = grow(c, ck_colck, ck_colck_cap, (size_t)std::max(1LL, ncol))) ||
          (rc = grow(c, ck_cnt, ck_cnt_cap, (size_t)std::max(1LL, ncnt))) ||
          (rc = grow(c, ck_mv, ck_mv_cap, (size_t)nj * job_words)) ||
          (rc = grow(c, ck_slots, ck_slots_cap, (size_t)nj * ew * (R + 256))) ||
          (rc = grow(c, ck_rings, ck_rings_cap, (size_t)nblocks * ew * ring)) ||
          (rc = grow(c, ck_jobs, ck_jobs_cap, (size_t)nj * sizeof(Job))) ||
          (rc = grow(c, ck_states, ck_states_cap, (size_t)nj * sizeof(State))))
        return rc;
      HIPC(c, hipMemcpyAsync(ck_jobs, jobs.data(), nj * sizeof(Job), hipMemcpyHostToDevice, c->stream));
      HIPC(c, hipMemcpyAsync(ck_states, sts.data(), nj * sizeof(State), hipMemcpyHostToDevice, c->stream));
      HIPC(c, hipMemsetAsync(ck_cnt, 0, std::max(1LL, ncnt) * sizeof(int), c->stream));
      unsigned long long* counter = (unsigned long long*)c->small;
      int* ferr = (int*)((char*)c->small + 192);
      HIPC(c, hipMemsetAsync(counter, 0, sizeof(unsigned long long), c->stream));
      HIPC(c, hipMemsetAsync(ferr, 0, sizeof(int), c->stream));
      int khz = 0;
      if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device) != hipSuccess || khz <= 0) khz = 100000;
      const long long limit = (long long)env_int("ODBA_BAND_WAIT_S", 60) * 1000LL * khz;
      HIPC(c, launch_ckpt_forward(ew, nblocks, c->stream, ck_jobs, nj, nclaims, counter, ck_rowck, ck_colck, ck_cnt,
                                  ck_rings, ring, rshift, ferr, limit));
      for (int it = 0; it < maxit; it++) {
        HIPC(c, launch_ckpt_block(ew, c->stream, ck_jobs, nj, ck_states, ck_rowck, ck_colck, ck_mv, job_words,
                                  ck_slots, R + 256, rshift));
        HIPC(c, launch_ckpt_trace(c->stream, ck_jobs, nj, ck_states, ck_mv, job_words, runs, rshift, ew,
                                  rec ? pbuf : nullptr));
      }
      int herr = 0;
      HIPC(c, hipMemcpyAsync(&herr, ferr, sizeof(int), hipMemcpyDeviceToHost, c->stream));
      HIPC(c, hipMemcpyAsync(sts.data(), ck_states, nj * sizeof(State), hipMemcpyDeviceToHost, c->stream));
      HIPC(c, hipStreamSynchronize(c->stream));
      if (herr & 8) return fail(c, ODBA_E_HIP, "checkpoint forward pass: edge hand-off timed out");
      for (int q = 0; q < nj; q++) {
        if (sts[q].err & 1) return fail(c, ODBA_E_TRACEBACK, "traceback left the matrix before reaching (0,0)");
        if (sts[q].err & 4) return fail(c, ODBA_E_TRACEBACK, "traceback produced a third counted run for one centroid position");
        if (!sts[q].done) return fail(c, ODBA_E_TRACEBACK, "checkpointed traceback of sequence %d did not finish", idx[q]);
      }
      return ODBA_OK;
    }
  }

  // In the open modes the stripe path depends on each stripe's truncated
  // height (the open-end top-row rule at the recompute's last row), so float
  // jobs there keep the reference's stripe semantics but recompute only the
  // band of rows holding the walk's current cell, from cost rows stored by a
  // forward pass over the whole device (k_sck_forward); ODBA_CKPT=0 keeps
  // the recompute from row 0, ODBA_CKPT_RSHIFT sets the band height 2^k
  // (default 4,096 rows: each step recomputes at most a band, so shorter
  // bands pay for their checkpoint rows, a quarter of the edge columns'
  // memory; 8 open-end pairs of 1M-4M samples: 44.3 s per round against
  // 59.4 s with 16,384-row bands).
  bool use_sck() const {
    const char* e = std::getenv("ODBA_CKPT");
    return std::is_same<T, float>::value && (os || oe) && !(e && !std::strcmp(e, "0"));
  }
  static int ck_rshift() { return std::max(6, std::min(20, env_int("ODBA_CKPT_RSHIFT", 12))); }
  // Device 4-byte units of one checkpointed open-mode stripe job.
  static long long sck_need(int H, int LX, int rshift) {
    const long long R = 1LL << rshift, K = (LX + ODBA_W - 1) / ODBA_W, nb = (H + R - 1) / R;
    return K * ((long long)H + 256) + (nb - 1) * (long long)LX + 2 * K + K * nb + K + (R + 127) * 64 + 64 +
           3 * (R + 1 + 256);
  }

  int run_sck_batch(const std::vector<int>& idx, int b0) {
    if constexpr (!std::is_same<T, float>::value) {
      return fail(c, ODBA_E_UNSUPPORTED, "checkpointed stripe paths are float only");
    } else {
      const int nj = (int)idx.size();
      if (!nj) return ODBA_OK;
      const int rshift = ck_rshift();
      const long long R = 1LL << rshift;
      const long long job_words = (R + 127) * 64;  // a block of up to R + 1 rows
      const long long qstride = R + 1 + 256;       // a quarter strip's edge over a block
      std::vector<SckJob> sj(nj);
      std::vector<TbJob> tb(nj);
      long long nslot = 0, nrow = 0, ntc = 0, npm = 0, ncnt = 0, nclaims = 0, words = 0;
      int kmax = 1, maxit = 0;
      for (int q = 0; q < nj; q++) {
        const int b = idx[q];
        const int ls = h_lens[b];
        const int flip = (oe && lc < ls) ? 1 : 0;
        const int H = flip ? lc : ls, LX = flip ? ls : lc;
        const int K = (LX + ODBA_W - 1) / ODBA_W;
        const long long nb = (H + R - 1) / R;
        const T* sq = seqdata + h_offs[b];
        sj[q] = SckJob{flip ? cen : sq, flip ? sq : cen, H, LX, nslot, nrow, ntc, npm, ncnt, nclaims, words};
        tb[q].ly = H; tb[q].lx = LX; tb[q].flip = flip; tb[q].seq = b - b0; tb[q].mv_off = words;
        tb[q].top_off = (long long)q * 64;
        tb[q].path_off = rec ? h_poff[b] : -1;
        tb[q].runs_off = (long long)(b - b0) * lc;
        tb[q].rp = nullptr;
        nslot += (long long)K * ((long long)H + 256);
        nrow += (nb - 1) * (long long)LX;
        ntc += K;
        npm += (long long)K * nb;
        ncnt += K;
        nclaims += (K + kSckBandSwaths - 1) / kSckBandSwaths;
        words += job_words;
        kmax = std::max(kmax, K);
        maxit = std::max<int>(maxit, K + (int)nb + 1);  // each step leaves a stripe or a band, or ends
      }
      int rc;
      if ((rc = grow(c, sck_slots, sck_slots_cap, (size_t)std::max(1LL, nslot))) ||
          (rc = grow(c, ck_rowck, ck_rowck_cap, (size_t)std::max(1LL, nrow))) ||
          (rc = grow(c, sck_tc, sck_tc_cap, (size_t)ntc)) || (rc = grow(c, sck_pm, sck_pm_cap, (size_t)npm)) ||
          (rc = grow(c, ck_cnt, ck_cnt_cap, (size_t)ncnt)) || (rc = grow(c, smv, smv_cap, (size_t)words)) ||
          (rc = grow(c, stop, stop_cap, (size_t)nj * 64)) || (rc = grow(c, sstate, sstate_cap, (size_t)nj)) ||
          (rc = grow(c, sck_jobs, sck_jobs_cap, (size_t)nj)) || (rc = grow(c, stbj, stbj_cap, (size_t)nj)) ||
          (rc = grow(c, sck_q, sck_q_cap, (size_t)nj * 3 * (size_t)qstride)))
        return rc;
      HIPC(c, hipMemcpyAsync(sck_jobs, sj.data(), nj * sizeof(SckJob), hipMemcpyHostToDevice, c->stream));
      HIPC(c, hipMemcpyAsync(stbj, tb.data(), nj * sizeof(TbJob), hipMemcpyHostToDevice, c->stream));
      HIPC(c, hipMemsetAsync(ck_cnt, 0, (size_t)ncnt * sizeof(int), c->stream));
      unsigned long long* counter = (unsigned long long*)c->small;
      int* ferr = (int*)((char*)c->small + 192);
      HIPC(c, hipMemsetAsync(counter, 0, sizeof(unsigned long long), c->stream));
      HIPC(c, hipMemsetAsync(ferr, 0, sizeof(int), c->stream));
      int khz = 0;
      if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device) != hipSuccess || khz <= 0) khz = 100000;
      const long long limit = (long long)env_int("ODBA_BAND_WAIT_S", 60) * 1000LL * khz;
      const int nblocks = (int)std::max<long long>(1, std::min<long long>(nclaims, c->cus));
      HIPC(c, launch_sck_forward(os, oe, nblocks, c->stream, sck_jobs, nj, nclaims, counter, sck_slots, ck_rowck, sck_tc,
                                 ck_cnt, rshift, ferr, limit));
      HIPC(c, launch_sck_init(os, oe, c->stream, sck_jobs, nj, kmax, sck_slots, sck_tc, sck_pm, sstate, rshift));
      for (int it = 0; it < maxit; it++) {
        HIPC(c, launch_sck_step(os, oe, c->stream, sck_jobs, nj, sck_slots, ck_rowck, sck_pm, smv, stop, sck_q, qstride,
                                sstate, rshift));
        HIPC(c, launch_sck_traceback(c->stream, stbj, nj, smv, stop, sstate, runs, os, oe, err, rec ? pbuf : nullptr));
      }
      std::vector<StripeState> sts(nj);
      int herr = 0;
      HIPC(c, hipMemcpyAsync(&herr, ferr, sizeof(int), hipMemcpyDeviceToHost, c->stream));
      HIPC(c, hipMemcpyAsync(sts.data(), sstate, nj * sizeof(StripeState), hipMemcpyDeviceToHost, c->stream));
      HIPC(c, hipStreamSynchronize(c->stream));
      if (herr & 8) return fail(c, ODBA_E_HIP, "checkpointed stripe forward pass: edge hand-off timed out");
      for (int q = 0; q < nj; q++)
        if (!sts[q].done && !sts[q].err)
          return fail(c, ODBA_E_TRACEBACK, "checkpointed stripe traceback of sequence %d did not finish", idx[q]);
      return ODBA_OK;
    }
  }

  int run_stripe_batch(const std::vector<int>& idx, int b0) {
    const int nj = (int)idx.size();
    if (!nj) return ODBA_OK;
    std::vector<StripeJobH> sj(nj);
    std::vector<TbJob> tb(nj);
    long long slot_elems = 0, words = 0;
    int maxK = 0;
    for (int q = 0; q < nj; q++) {
      const int b = idx[q];
      const int ls = h_lens[b];
      const int flip = (oe && lc < ls) ? 1 : 0;
      const int H = flip ? lc : ls, LX = flip ? ls : lc;
      const int K = (LX + ODBA_W - 1) / ODBA_W;
      const T* sq = seqdata + h_offs[b];
      sj[q].y = flip ? cen : sq;
      sj[q].x = flip ? sq : cen;
      sj[q].H = H;
      sj[q].LX = LX;
      sj[q].slot_off = slot_elems;
      sj[q].mv_off = words;
      sj[q].top_off = (long long)q * 64;
      tb[q].ly = H; tb[q].lx = LX; tb[q].flip = flip; tb[q].seq = b - b0; tb[q].mv_off = words; tb[q].top_off = (long long)q * 64;
      tb[q].path_off = rec ? h_poff[b] : -1;
      tb[q].runs_off = (long long)(b - b0) * lc;
      tb[q].rp = nullptr;
      slot_elems += (long long)K * (H + 256);
      words += std::max(1LL, swath_words(H, os, oe));
      maxK = std::max(maxK, K);
    }
    int rc = grow(c, slots, slots_cap, (size_t)std::max(1LL, slot_elems));
    if (rc) return rc;
    rc = grow(c, smv, smv_cap, (size_t)words);
    if (rc) return rc;
    rc = grow(c, stop, stop_cap, (size_t)nj * 64);
    if (rc) return rc;
    rc = grow(c, sstate, sstate_cap, (size_t)nj);
    if (rc) return rc;
    rc = grow(c, sjobs, sjobs_cap, (size_t)nj);
    if (rc) return rc;
    rc = grow(c, stbj, stbj_cap, (size_t)nj);
    if (rc) return rc;
    HIPC(c, hipMemcpyAsync(sjobs, sj.data(), nj * sizeof(StripeJobH), hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipMemcpyAsync(stbj, tb.data(), nj * sizeof(TbJob), hipMemcpyHostToDevice, c->stream));
    unsigned int* counter = (unsigned int*)((char*)c->small + 128);
    HIPC(c, hipMemsetAsync(counter, 0, sizeof(unsigned int), c->stream));
    HIPC(c, launch_stripe(dtype_of<T>(), minfirst, true, os, oe, grid_blocks(c, nj), c->stream, sjobs, nj, 0, counter,
                          slots, smv, stop, sstate));
    const int steps_blocks = (nj + 3) / 4;
    for (int m = 0; m < maxK; m++) {
      HIPC(c, launch_stripe(dtype_of<T>(), minfirst, false, os, oe, steps_blocks, c->stream, sjobs, nj, m, counter,
                            slots, smv, stop, sstate));
      HIPC(c, launch_stripe_traceback(c->stream, stbj, nj, m, smv, stop, sstate, runs, lc, os, oe, err,
                                      rec ? pbuf : nullptr));
    }
    HIPC(c, hipStreamSynchronize(c->stream));
    return ODBA_OK;
  }

  bool graphs_enabled() const {
    const char* e = std::getenv("ODBA_DBA_GRAPH");
    if (e && !std::strcmp(e, "1")) return true;
    return c->graphs && !(e && !std::strcmp(e, "0"));
  }
  // Capturing and instantiating a round costs about as much as a few
  // eager rounds' launch overhead, so replays start after a few rounds.
  static int graph_after() {
    const char* e = std::getenv("ODBA_DBA_GRAPH_AFTER");
    return (e && *e) ? std::max(1, std::atoi(e)) : 4;
  }

  std::vector<PJ<T>> keep_pj;  // a one-batch round's host job tables, alive until the round's last sync
  std::vector<TbJob> keep_tb;

  // One DBA round. With host_out, the new centroid is also copied there.
  int round(double* delta, T* host_out = nullptr) {
    if (g_ok && !rec && graphs_enabled() && g_eager >= graph_after()) return graph_round(delta, host_out);
    if (!g_ok) g_eager = 0;
    g_ok = false;
    HIPC(c, hipMemsetAsync(err, 0, sizeof(int), c->stream));
    HIPC(c, hipMemsetAsync(dbits, 0, sizeof(unsigned long long), c->stream));
    int rc = grow(c, runs, runs_cap, (size_t)n * lc);
    if (rc) return rc;
    rec_valid = false;
    if (rec) {
      h_poff.resize(n);
      long long tot = 0;
      for (int b = 0; b < n; b++) {
        h_poff[b] = tot;
        tot += 1 + (long long)h_lens[b] + lc;
      }
      rc = grow(c, pbuf, pbuf_cap, (size_t)tot);
      if (rc) return rc;
    }
    // batches of sequences whose move words fit the budget (sequence order kept)
    int b0 = 0;
    bool first = true;
    while (b0 < n) {
      std::vector<PJ<T>> pj;
      std::vector<TbJob> tb;
      std::vector<int> stripe_idx;
      long long mvw = 0, tw = 0, sbytes = 0;
      int maxly = 1, maxlx = 0, minly = INT32_MAX;
      int b1 = b0;
      while (b1 < n) {
        const int ls = h_lens[b1];
        const int flip = (oe && lc < ls) ? 1 : 0;
        const int ly = flip ? lc : ls, lx = flip ? ls : lc;
        const bool st = use_stripe(b1);
        long long need = st ? ((long long)((lx + ODBA_W - 1) / ODBA_W) * (ly + 256) * (long long)sizeof(T) + swath_words(ly, os, oe) * 4) / 4
                            : (long long)((lx + ODBA_W - 1) / ODBA_W) * swath_words(ly, os, oe);
        if (st && use_sck()) need = sck_need(ly, lx, ck_rshift());
        if (st && use_ckpt()) {  // checkpoint rows and columns plus one block's words (4-byte units)
          const long long R = 1LL << std::max(6, std::min(20, env_int("ODBA_CKPT_RSHIFT", 14)));
          const long long E = (long long)(env_int("ODBA_CKPT_EW", 16) == 2 ? 2 : 16) * ODBA_W;
          need = ((ly + R - 1) / R) * (long long)lx + ((lx + E - 1) / E) * ((long long)ly + 256) + (E / ODBA_W) * (R + 63) * 64 * 2;
        }
        if (b1 > b0 && mvw + sbytes + need > budget_words) break;
        if (st) {
          stripe_idx.push_back(b1);
          sbytes += need;
        } else {
          long long mo, to;
          plan_job(ly, lx, os, oe, mvw, tw, mo, to);
          const T* sq = seqdata + h_offs[b1];
          PJ<T> j{flip ? cen : sq, flip ? sq : cen, ly, lx, mo, to};
          pj.push_back(j);
          TbJob t;
          t.ly = ly; t.lx = lx; t.flip = flip; t.seq = b1 - b0; t.mv_off = mo; t.top_off = to;
          t.path_off = rec ? h_poff[b1] : -1;
          t.runs_off = (long long)(b1 - b0) * lc;
          t.rp = nullptr;
          tb.push_back(t);
          maxly = std::max(maxly, ly);
          maxlx = std::max(maxlx, lx);
          minly = std::min(minly, ly);
        }
        b1++;
      }
      const int nfull = (int)pj.size();
      if (b0 == 0 && b1 == n && nfull == n && !rec) {  // replayable as a graph from the next round on
        g_ok = true;
        g_pj = pj;
        g_flip.assign(n, 0);
        for (int q = 0; q < n; q++) g_flip[q] = (char)(pj[q].y == cen);
        g_maxly = maxly; g_maxlx = maxlx; g_minly = minly;
      }
      if (nfull) {
        rc = grow(c, mv, mv_cap, (size_t)std::max(1LL, mvw));
        if (rc) return rc;
        rc = grow(c, mtop, mtop_cap, (size_t)std::max(1LL, tw));
        if (rc) return rc;
        HIPC(c, hipMemcpyAsync(jobs, pj.data(), nfull * sizeof(PJ<T>), hipMemcpyHostToDevice, c->stream));
        HIPC(c, hipMemcpyAsync(tbj, tb.data(), nfull * sizeof(TbJob), hipMemcpyHostToDevice, c->stream));
        rc = run_path_jobs(c, dtype_of<T>(), minfirst, os, oe, jobs, nfull, maxly, mv, mtop, abr, maxlx, minly, err,
                           mean_swaths(pj), path_tail(pj));
        if (rc) return rc;
        HIPC(c, launch_traceback(c->stream, tbj, nfull, mv, mtop, abr, runs, lc, os, oe, err, rec ? pbuf : nullptr));
      }
      rc = use_ckpt() ? run_ckpt_batch(stripe_idx, b0)
                      : (use_sck() ? run_sck_batch(stripe_idx, b0) : run_stripe_batch(stripe_idx, b0));
      if (rc) return rc;
      HIPC(c, launch_accumulate(dtype_of<T>(), c->stream, seqdata, offs + b0, runs, b1 - b0, lc, acc, counts, first ? 1 : 0));
      // The host vectors must outlive the async copies: a round of one batch
      // keeps them until its final synchronisation, others wait here.
      if (b0 == 0 && b1 == n && stripe_idx.empty()) {
        keep_pj.swap(pj);
        keep_tb.swap(tb);
      } else {
        HIPC(c, hipStreamSynchronize(c->stream));
      }
      first = false;
      b0 = b1;
    }
    HIPC(c, launch_mean_delta(dtype_of<T>(), c->stream, acc, counts, cen, newc, lc, dbits, err));
    int herr = 0;
    unsigned long long bits = 0;
    HIPC(c, hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipMemcpyAsync(&bits, dbits, sizeof(bits), hipMemcpyDeviceToHost, c->stream));
    if (host_out) HIPC(c, hipMemcpyAsync(host_out, newc, lc * sizeof(T), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    keep_pj.clear();
    keep_tb.clear();
    if (herr & 8) return fail(c, ODBA_E_HIP, "path band edge hand-off timed out");
    if (herr & 1) return fail(c, ODBA_E_TRACEBACK, "traceback left the matrix before reaching (0,0)");
    if (herr & 4) return fail(c, ODBA_E_TRACEBACK, "traceback produced a third counted run for one centroid position");
    if (herr & 2) return fail(c, ODBA_E_DIV_ZERO, "integer centroid position with zero count");
    double d;
    std::memcpy(&d, &bits, sizeof(d));
    if (delta) *delta = d;
    std::swap(cen, newc);
    rec_valid = rec;
    g_eager = g_ok ? g_eager + 1 : 0;
    return ODBA_OK;
  }

  // The round as one hipGraph launch: clears by fill kernels (no memset
  // nodes), path jobs, traceback, accumulation, mean/delta, and the error
  // word, delta and new centroid copied to pinned host memory. One graph per
  // centroid buffer the round reads; each is captured on first use and
  // captured again when a device buffer it uses has moved.
  int graph_round(double* delta, T* host_out) {
    const int p = (cen == cbuf[0]) ? 0 : 1;
    const int nj = (int)g_pj.size();
    const size_t need = 64 + (size_t)std::max(1, lc) * sizeof(T);
    if (c->pinned_bytes < need) {
      if (c->pinned) hipHostFree(c->pinned);
      c->pinned = nullptr;
      c->pinned_bytes = 0;
      const size_t sz = std::max<size_t>(need, (size_t)1 << 20);
      HIPC(c, hipHostMalloc(&c->pinned, sz));
      c->pinned_bytes = sz;
    }
    h_err = (int*)c->pinned;
    h_bits = (unsigned long long*)((char*)c->pinned + 8);
    h_cen = (T*)((char*)c->pinned + 64);
    std::vector<void*> sig = {c->esc, c->small, (void*)mv, (void*)mtop, (void*)runs, (void*)acc, (void*)counts,
                              (void*)abr, (void*)tbj, (void*)err, (void*)dbits, (void*)cbuf[0], (void*)cbuf[1],
                              (void*)seqdata, (void*)offs, c->pinned};
    if (gexec[p] && gsig[p] != sig) {
      hipGraphExecDestroy(gexec[p]);
      gexec[p] = nullptr;
    }
    if (!gexec[p]) {

      if (!gjobs[p]) HIPC(c, hipMalloc(&gjobs[p], nj * sizeof(PJ<T>)));
      std::vector<PJ<T>> pj = g_pj;
      for (int q = 0; q < nj; q++) {
        if (g_flip[q]) pj[q].y = cbuf[p];
        else pj[q].x = cbuf[p];
      }
      HIPC(c, hipMemcpyAsync(gjobs[p], pj.data(), nj * sizeof(PJ<T>), hipMemcpyHostToDevice, c->stream));
      HIPC(c, hipStreamSynchronize(c->stream));
      HIPC(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      int rc = ODBA_OK;
      hipError_t e = launch_fill_u32(c->stream, err, 0u, 1);
      if (e == hipSuccess) e = launch_fill_u32(c->stream, dbits, 0u, 2);
      if (e != hipSuccess) rc = fail(c, ODBA_E_HIP, "graph capture: %s", hipGetErrorString(e));
      if (!rc) rc = run_path_jobs(c, dtype_of<T>(), minfirst, os, oe, gjobs[p], nj, g_maxly, mv, mtop, abr, g_maxlx, g_minly, err,
                                  mean_swaths(g_pj), path_tail(g_pj));
      if (!rc) {
        e = launch_traceback(c->stream, tbj, nj, mv, mtop, abr, runs, lc, os, oe, err, nullptr);
        if (e == hipSuccess) e = launch_accumulate(dtype_of<T>(), c->stream, seqdata, offs, runs, n, lc, acc, counts, 1);
        if (e == hipSuccess) e = launch_mean_delta(dtype_of<T>(), c->stream, acc, counts, cbuf[p], cbuf[1 - p], lc, dbits, err);
        if (e == hipSuccess) e = hipMemcpyAsync(h_err, err, sizeof(int), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(h_bits, dbits, sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(h_cen, cbuf[1 - p], lc * sizeof(T), hipMemcpyDeviceToHost, c->stream);
        if (e != hipSuccess) rc = fail(c, ODBA_E_HIP, "graph capture: %s", hipGetErrorString(e));
      }
      hipGraph_t g = nullptr;
      const hipError_t ee = hipStreamEndCapture(c->stream, &g);
      if (!rc && ee != hipSuccess) rc = fail(c, ODBA_E_HIP, "end of graph capture: %s", hipGetErrorString(ee));
      if (!rc && hipGraphInstantiate(&gexec[p], g, nullptr, nullptr, 0) != hipSuccess) {
        gexec[p] = nullptr;
        rc = fail(c, ODBA_E_HIP, "graph instantiation failed");
      }
      if (g) hipGraphDestroy(g);
      if (rc) return rc;
      gsig[p] = sig;
    }
    HIPC(c, hipGraphLaunch(gexec[p], c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    const int herr = *h_err;
    if (herr & 8) return fail(c, ODBA_E_HIP, "path band edge hand-off timed out");
    if (herr & 1) return fail(c, ODBA_E_TRACEBACK, "traceback left the matrix before reaching (0,0)");
    if (herr & 4) return fail(c, ODBA_E_TRACEBACK, "traceback produced a third counted run for one centroid position");
    if (herr & 2) return fail(c, ODBA_E_DIV_ZERO, "integer centroid position with zero count");
    double d;
    std::memcpy(&d, h_bits, sizeof(d));
    if (delta) *delta = d;
    std::swap(cen, newc);
    rec_valid = false;
    if (host_out) std::memcpy(host_out, h_cen, lc * sizeof(T));
    return ODBA_OK;
  }

  // Recorded traceback of sequence `seq` from the last round.
  int get_path(int seq, uint64_t* out, uint64_t cap, uint64_t* nsteps) {
    if (!rec_valid) return fail(c, ODBA_E_INVALID, "no traceback was recorded in the last DBA step");
    if (seq < 0 || seq >= n) return fail(c, ODBA_E_INVALID, "sequence index %d out of range", seq);
    unsigned long long cnt = 0;
    HIPC(c, hipMemcpyAsync(&cnt, pbuf + h_poff[seq], sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    if (nsteps) *nsteps = cnt;
    if (!out) return ODBA_OK;
    if (cap < cnt) return fail(c, ODBA_E_INVALID, "path buffer holds %llu steps, %llu needed",
                               (unsigned long long)cap, cnt);
    HIPC(c, hipMemcpyAsync(out, pbuf + h_poff[seq] + 1, cnt * sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    return ODBA_OK;
  }
};

template <typename T>
int dba_round_t(odba_ctx* c, const T* C, int lc, const T* const* seqs, const uint64_t* lens, int n, int os, int oe,
                int policy, uint64_t free_mem, T* newc, uint32_t* counts_out, double* delta) {
  DbaState<T> st;
  int rc = st.init(c, seqs, lens, n, os, oe);
  if (rc) return rc;
  st.policy = policy;
  st.free_mem = auto_free_mem(c, policy, free_mem);
  rc = st.set_centroid(C, lc);
  if (rc) return rc;
  rc = st.round(delta);
  if (rc) return rc;
  HIPC(c, hipMemcpyAsync(newc, st.cen, lc * sizeof(T), hipMemcpyDeviceToHost, c->stream));
  if (counts_out) HIPC(c, hipMemcpyAsync(counts_out, st.counts, lc * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return ODBA_OK;
}

template <typename T> bool sanitize_drop_last(T v) { return (double)v >= std::sqrt((double)std::numeric_limits<T>::max()); }
template <> bool sanitize_drop_last<float>(float v) { return v >= std::sqrt(std::numeric_limits<float>::max()); }

template <typename T> T host_d2t(double d) {
  if (std::is_floating_point<T>::value) return (T)d;
  if (d != d) return T(0);
  if (d >= (double)std::numeric_limits<T>::max()) return std::numeric_limits<T>::max();
  if (d <= (double)std::numeric_limits<T>::lowest()) return std::numeric_limits<T>::lowest();
  return (T)d;
}

template <typename T>
int znorm_t(odba_ctx* c, T* const* seqs, const uint64_t* lens, int n, double* mu, double* sigma) {
  std::vector<long long> offs(n);
  std::vector<int> L(n);
  long long total = 0;
  for (int i = 0; i < n; i++) { offs[i] = total; L[i] = (int)lens[i]; total += padded_len((long long)lens[i]); }
  std::vector<T> hb((size_t)total, T(0));
  for (int i = 0; i < n; i++) std::memcpy(&hb[offs[i]], seqs[i], lens[i] * sizeof(T));
  T* d = nullptr;
  long long* doffs = nullptr;
  int* dl = nullptr;
  double *dmu = nullptr, *dsg = nullptr;
  HIPC(c, hipMalloc((void**)&d, hb.size() * sizeof(T)));
  HIPC(c, hipMalloc((void**)&doffs, n * sizeof(long long)));
  HIPC(c, hipMalloc((void**)&dl, n * sizeof(int)));
  HIPC(c, hipMalloc((void**)&dmu, n * sizeof(double)));
  HIPC(c, hipMalloc((void**)&dsg, n * sizeof(double)));
  HIPC(c, hipMemcpyAsync(d, hb.data(), hb.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemcpyAsync(doffs, offs.data(), n * sizeof(long long), hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemcpyAsync(dl, L.data(), n * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPC(c, launch_znorm(dtype_of<T>(), c->stream, d, doffs, dl, n, dmu, dsg));
  HIPC(c, hipMemcpyAsync(hb.data(), d, hb.size() * sizeof(T), hipMemcpyDeviceToHost, c->stream));
  std::vector<double> hmu(n), hsg(n);
  HIPC(c, hipMemcpyAsync(hmu.data(), dmu, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipMemcpyAsync(hsg.data(), dsg, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < n; i++) std::memcpy(seqs[i], &hb[offs[i]], lens[i] * sizeof(T));
  if (mu) std::memcpy(mu, hmu.data(), n * sizeof(double));
  if (sigma) std::memcpy(sigma, hsg.data(), n * sizeof(double));
  hipFree(d); hipFree(doffs); hipFree(dl); hipFree(dmu); hipFree(dsg);
  return ODBA_OK;
}

// Sum of squared distances in T in the host loop order, first strict
// minimum; two members -> the longer, the second on a tie
// (dba.hpp:186-207, 297-351).
template <typename T> int medoid_single(const T* d, const std::vector<uint64_t>& l, int n) {
  std::vector<T> sos(n, T(0));
  size_t idx = 0;
  for (int r = 0; r + 1 < n; r++) {
    for (int cc = r + 1; cc < n; cc++) {
      T v = d[idx + cc - r - 1];
      v = (T)(v * v);
      sos[r] = (T)(sos[r] + v);
      sos[cc] = (T)(sos[cc] + v);
    }
    idx += n - r - 1;
  }
  if (n > 2) {
    int med = -1;
    T low = std::numeric_limits<T>::max();
    for (int i = 0; i < n; i++)
      if (sos[i] < low) { med = i; low = sos[i]; }
    return med;
  }
  if (n == 2) return l[0] > l[1] ? 0 : 1;
  return 0;
}

// The first half of performDBA for one cluster (dba.hpp:863-990): the
// sanitized members, stably sorted by length, are packed and uploaded once,
// z-normalised in place, and that one buffer feeds the all-vs-all; the
// medoid follows. The buffer stays resident for the DBA rounds, which run
// either here (perform_dba_t) or packed with other jobs' rounds (RoundEngine).
template <typename T> struct Prepared {
  int device = 0;
  int n = 0, os = 0, oe = 0, norm = 0;
  T* d = nullptr;            // packed, padded, z-normalised members in sorted order
  long long* o = nullptr;    // device offsets
  int* ln = nullptr;         // device lengths
  std::vector<long long> ho;
  std::vector<int> hl;
  std::vector<uint64_t> l;
  std::vector<double> mu, sg;
  std::vector<int> perm;     // sorted position -> input index
  bool no_wrap = true;
  int med = -1, lc = 0;
  Prepared() = default;
  Prepared(const Prepared&) = delete;
  Prepared& operator=(const Prepared&) = delete;
  void release() {
    if (d || o || ln) hipSetDevice(device);
    for (void* p : {(void*)d, (void*)o, (void*)ln})
      if (p) hipFree(p);
    d = nullptr;
    o = nullptr;
    ln = nullptr;
  }
  ~Prepared() { release(); }
};

template <typename T>
int prepare_t(odba_ctx* c, const T* const* in, const uint64_t* in_lens, int n, int os, int oe, int norm,
              Prepared<T>& P, T* dists_out) {
  if (n < 2) return fail(c, ODBA_E_NOT_ENOUGH_SEQUENCES, "at least two sequences are needed, got %d", n);
  P.device = c->device;
  P.n = n; P.os = os; P.oe = oe; P.norm = norm;
  std::vector<uint64_t> lens(in_lens, in_lens + n);
  for (int i = 0; i < n; i++)
    if (lens[i] && sanitize_drop_last<T>(in[i][lens[i] - 1])) lens[i]--;  // dba.hpp:868-872
  std::vector<int>& perm = P.perm;
  perm.resize(n);
  for (int i = 0; i < n; i++) perm[i] = i;
  std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return lens[a] < lens[b]; });  // dba.hpp:876-885
  std::vector<uint64_t>& l = P.l;
  std::vector<long long>& ho = P.ho;
  std::vector<int>& hl = P.hl;
  l.resize(n);
  ho.resize(n);
  hl.resize(n);
  long long total = 0;
  int maxlen = 0;
  for (int i = 0; i < n; i++) {
    l[i] = lens[perm[i]];
    if (l[i] == 0 || l[i] > (uint64_t)INT32_MAX / 2) return fail(c, ODBA_E_INVALID, "sequence %d has length %llu", perm[i], (unsigned long long)l[i]);
    ho[i] = total;
    hl[i] = (int)l[i];
    maxlen = std::max(maxlen, hl[i]);
    total += padded_len((long long)l[i]);
  }
  std::vector<T> hb((size_t)total, T(0));
  for (int i = 0; i < n; i++) std::memcpy(&hb[ho[i]], in[perm[i]], l[i] * sizeof(T));
  struct Stats {
    double* mu = nullptr; double* sg = nullptr;
  } S;
  HIPC(c, hipMalloc((void**)&P.d, hb.size() * sizeof(T)));
  HIPC(c, hipMalloc((void**)&P.o, n * sizeof(long long)));
  HIPC(c, hipMalloc((void**)&P.ln, n * sizeof(int)));
  HIPC(c, hipMemcpyAsync(P.d, hb.data(), hb.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemcpyAsync(P.o, ho.data(), n * sizeof(long long), hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemcpyAsync(P.ln, hl.data(), n * sizeof(int), hipMemcpyHostToDevice, c->stream));
  P.mu.assign(n, 0.0);
  P.sg.assign(n, 1.0);
  if (norm) {
    S.mu = (double*)ctx_buffer(c, kGbStats, 2 * (size_t)n * sizeof(double));
    if (!S.mu) return fail(c, ODBA_E_NOMEM, "cannot allocate z-norm statistics");
    S.sg = S.mu + n;
    HIPC(c, launch_znorm(dtype_of<T>(), c->stream, P.d, P.o, P.ln, n, S.mu, S.sg));
    HIPC(c, hipMemcpyAsync(P.mu.data(), S.mu, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipMemcpyAsync(P.sg.data(), S.sg, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  P.no_wrap = true;
  if (!is_float<T>()) {  // the wrap bound needs the (normalised) values
    if (norm) HIPC(c, hipMemcpyAsync(hb.data(), P.d, hb.size() * sizeof(T), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    std::vector<const T*> sp(n);
    for (int i = 0; i < n; i++) sp[i] = &hb[ho[i]];
    P.no_wrap = int_cannot_wrap<T>(sp.data(), l.data(), n);
  }
  HIPC(c, hipStreamSynchronize(c->stream));
  // the all-vs-all on the resident buffer (a seqset view that owns nothing)
  odba_seqset view;
  view.ctx = c; view.dtype = dtype_of<T>(); view.n = n; view.data = P.d; view.offs = P.o; view.lens = P.ln;
  view.h_offs = ho; view.h_lens = hl; view.maxlen = maxlen; view.int_no_wrap = P.no_wrap;
  view.borrowed = true;  // its pair order and small-pair records go to the context's buffers
  std::vector<T> d((size_t)n * (n - 1) / 2);
  const int rc = odba_dtw_all_pairs_resident(c, &view, os, oe, d.data(), nullptr, nullptr);
  if (rc) return rc;
  if (dists_out) std::memcpy(dists_out, d.data(), d.size() * sizeof(T));
  P.med = medoid_single<T>(d.data(), l, n);
  if (P.med < 0) return fail(c, ODBA_E_MEDOID, "no medoid found (all sums of squares at the type maximum)");
  P.lc = hl[P.med];
  return ODBA_OK;
}

// performDBA's stop rule for one cluster (dba.hpp:1084-1110), applied after
// round r (0-based) produced nb: true when the loop ends here.
template <typename T> struct StopRule {
  std::vector<T> prev, prev2;
  void init(int lc) {
    prev.assign(lc, T(0));
    prev2.assign(lc, T(0));
  }
  bool done(int r, double delta, const std::vector<T>& nb, int os, int oe) {
    if (delta == 0) return true;
    if (os || oe) {
      if (r >= 1 && !std::memcmp(nb.data(), prev2.data(), sizeof(T) * nb.size())) return true;
      if (r > 1) prev2 = prev;
      if (r > 0) prev = nb;
    }
    return false;
  }
};

// Device-resident pipeline of performDBA for one cluster: preparation, then
// the rounds on the same context until the stop rule, then the rescale.
template <typename T>
int perform_dba_t(odba_ctx* c, const T* const* in, const uint64_t* in_lens, int n, int os, int oe, int norm,
                  int max_rounds, int policy, uint64_t free_mem, T* cen_out, uint64_t* cen_len, int* med_out, int* perm_out, int* rounds_out,
                  T* dists_out) {
  Prepared<T> P;
  int rc = prepare_t<T>(c, in, in_lens, n, os, oe, norm, P, dists_out);
  if (rc) return rc;
  const int med = P.med, lc = P.lc;
  DbaState<T> st;
  rc = st.init_shared(c, P.d, P.o, P.ho, P.hl, P.no_wrap, os, oe);
  if (rc) return rc;
  st.policy = policy;
  st.free_mem = auto_free_mem(c, policy, free_mem);
  rc = st.set_centroid_device(P.d + P.ho[med], lc);
  if (rc) return rc;
  std::vector<T> nb(lc);
  StopRule<T> stop;
  stop.init(lc);
  int rounds = 0;
  for (int r = 0; r < max_rounds; r++) {  // dba.hpp:1084-1110
    double delta = 0;
    rc = st.round(&delta, nb.data());
    if (rc) return rc;
    rounds = r + 1;
    if (stop.done(r, delta, nb, os, oe)) break;
  }
  if (norm)
    for (int i = 0; i < lc; i++) nb[i] = host_d2t<T>(P.mu[med] + nb[i] * P.sg[med]);  // dba.hpp:1117-1125
  std::memcpy(cen_out, nb.data(), lc * sizeof(T));
  if (cen_len) *cen_len = lc;
  if (med_out) *med_out = med;
  if (perm_out) std::memcpy(perm_out, P.perm.data(), n * sizeof(int));
  if (rounds_out) *rounds_out = rounds;
  return ODBA_OK;
}


// performDBA's per-cluster consensus loop (dba.hpp:990-1110) for every
// cluster at once. Each round packs the path jobs of all clusters that have
// not stopped into one path launch and one traceback launch, then one
// accumulation and one mean/delta launch over all their centroid
// positions; each cluster applies its own stop rule (delta == 0, the
// open-mode flip-flop test, max_rounds). Singletons are returned as they
// are (dba.hpp:1004-1041). With mu/sigma given, centroids are rescaled by
// their medoid's (dba.hpp:1117-1125).
template <typename T>
int dba_clusters_t(odba_ctx* c, const T* const* seqs, const uint64_t* lens, int n, const int* memb,
                   const int* medoids, int nc, int os, int oe, int max_rounds, const double* mu, const double* sg,
                   T* cen_out, uint64_t* cen_offs, int* rounds_out) {
  if (n < 1 || nc < 1) return fail(c, ODBA_E_INVALID, "no sequences or clusters");
  std::vector<long long> ho(n);
  std::vector<int> hl(n);
  long long total = 0;
  for (int i = 0; i < n; i++) {
    if (lens[i] == 0 || lens[i] > (uint64_t)INT32_MAX / 2) return fail(c, ODBA_E_INVALID, "sequence %d has length %llu", i, (unsigned long long)lens[i]);
    ho[i] = total;
    hl[i] = (int)lens[i];
    total += padded_len((long long)lens[i]);
  }
  std::vector<std::vector<int>> mem(nc);
  for (int i = 0; i < n; i++) {
    if (memb[i] < 0 || memb[i] >= nc) return fail(c, ODBA_E_INVALID, "membership %d of sequence %d out of range", memb[i], i);
    mem[memb[i]].push_back(i);
  }
  std::vector<int> lcs(nc);
  uint64_t off = 0;
  for (int k = 0; k < nc; k++) {
    if (medoids[k] < 0 || medoids[k] >= n || memb[medoids[k]] != k) return fail(c, ODBA_E_MEDOID, "medoid %d of cluster %d is not a member", medoids[k], k);
    lcs[k] = hl[medoids[k]];
    cen_offs[k] = off;
    off += (uint64_t)lcs[k];
  }
  cen_offs[nc] = off;
  // accumulation positions: each cluster starts at a multiple of 64
  long long apos = 0;
  for (int k = 0; k < nc; k++) apos += round64(lcs[k]);
  const bool no_wrap = int_cannot_wrap<T>(seqs, lens, n);
  // resident data, member offsets per cluster, centroid pools
  std::vector<T> hb((size_t)total, T(0));
  for (int i = 0; i < n; i++) std::memcpy(&hb[ho[i]], seqs[i], lens[i] * sizeof(T));
  std::vector<long long> moffs, mem_off(nc), cen_off(nc), runs_off(nc);
  long long coff = 0, roff = 0;
  for (int k = 0; k < nc; k++) {
    mem_off[k] = (long long)moffs.size();
    for (int b : mem[k]) moffs.push_back(ho[b]);
    cen_off[k] = coff;
    coff += padded_len(lcs[k]);
    runs_off[k] = roff;
    roff += (long long)mem[k].size() * lcs[k];
  }
  DevBufs B;
  T *data, *cen, *newc, *acc;
  long long* dmoffs;
  uint32_t* counts;
  int4* runs;
  MultiSlot* dslots;
  unsigned long long* dbits;
  int* derr;
  int* abr;
  TbJob* tbj;
  void* jobs;
  int rc;
  if ((rc = B.get(c, data, hb.size())) || (rc = B.get(c, dmoffs, moffs.size())) || (rc = B.get(c, cen, coff)) ||
      (rc = B.get(c, newc, coff)) || (rc = B.get(c, acc, apos)) || (rc = B.get(c, counts, apos)) ||
      (rc = B.get(c, runs, roff)) || (rc = B.get(c, dslots, nc)) || (rc = B.get(c, dbits, nc)) ||
      (rc = B.get(c, derr, 1)) || (rc = B.get(c, abr, n)) || (rc = B.get(c, tbj, n)))
    return rc;
  {
    unsigned char* q;
    if ((rc = B.get(c, q, (size_t)n * sizeof(PJ<T>)))) return rc;
    jobs = q;
  }
  HIPC(c, hipMemcpyAsync(data, hb.data(), hb.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemcpyAsync(dmoffs, moffs.data(), std::max<size_t>(moffs.size(), 1) * sizeof(long long), hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemsetAsync(cen, 0, coff * sizeof(T), c->stream));
  HIPC(c, hipMemsetAsync(newc, 0, coff * sizeof(T), c->stream));
  for (int k = 0; k < nc; k++)
    HIPC(c, hipMemcpyAsync(cen + cen_off[k], data + ho[medoids[k]], lcs[k] * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
  // host state per cluster
  std::vector<std::vector<T>> nb(nc), prev(nc), prev2(nc);
  std::vector<char> active(nc, 0);
  for (int k = 0; k < nc; k++) {
    rounds_out[k] = 0;
    nb[k].assign(seqs[medoids[k]], seqs[medoids[k]] + lcs[k]);
    if (mem[k].size() > 1) {
      active[k] = 1;
      prev[k].assign(lcs[k], T(0));
      prev2[k].assign(lcs[k], T(0));
    }
  }
  const long long budget_words = (long long)(c->mem_cap / 4 / 4);
  std::vector<unsigned long long> bits(nc);
  uint32_t* mv = nullptr;
  uint32_t* mtop = nullptr;
  size_t mv_cap = 0, mtop_cap = 0;
  struct FreeWords {
    uint32_t*& a; uint32_t*& b;
    ~FreeWords() { if (a) hipFree(a); if (b) hipFree(b); }
  } free_words{mv, mtop};
  for (int r = 0; r < max_rounds; r++) {
    std::vector<int> act;
    for (int k = 0; k < nc; k++)
      if (active[k]) act.push_back(k);
    if (act.empty()) break;
    // groups of whole clusters whose step words fit the budget
    size_t a0 = 0;
    while (a0 < act.size()) {
      std::vector<PJ<T>> pj;
      std::vector<TbJob> tb;
      std::vector<MultiSlot> slots;
      long long mvw = 0, tw = 0, pos = 0;
      int maxly = 1, maxlx = 0, minly = INT32_MAX;
      size_t a1 = a0;
      for (; a1 < act.size(); a1++) {
        const int k = act[a1];
        long long need = 0;
        for (int b : mem[k]) {
          const int flip = (oe && lcs[k] < hl[b]) ? 1 : 0;
          const int ly = flip ? lcs[k] : hl[b], lx = flip ? hl[b] : lcs[k];
          need += (long long)((lx + ODBA_W - 1) / ODBA_W) * swath_words(ly, os, oe);
        }
        if (a1 > a0 && mvw + need > budget_words) break;
        MultiSlot sl;
        sl.pos0 = pos; sl.data = data; sl.moffs = dmoffs + mem_off[k]; sl.runs = runs + runs_off[k];
        sl.cen = cen + cen_off[k]; sl.newc = newc + cen_off[k];
        sl.lc = lcs[k]; sl.nmem = (int)mem[k].size(); sl.slot = k; sl.pad = 0;
        slots.push_back(sl);
        pos += round64(lcs[k]);
        for (size_t q = 0; q < mem[k].size(); q++) {
          const int b = mem[k][q];
          const int flip = (oe && lcs[k] < hl[b]) ? 1 : 0;
          const int ly = flip ? lcs[k] : hl[b], lx = flip ? hl[b] : lcs[k];
          long long mo, to;
          plan_job(ly, lx, os, oe, mvw, tw, mo, to);
          const T* sq = data + ho[b];
          const T* cp = cen + cen_off[k];
          pj.push_back(PJ<T>{flip ? cp : sq, flip ? sq : cp, ly, lx, mo, to});
          TbJob t;
          t.ly = ly; t.lx = lx; t.flip = flip; t.seq = (int)q; t.mv_off = mo; t.top_off = to; t.path_off = -1;
          t.runs_off = runs_off[k] + (long long)q * lcs[k];
          t.rp = nullptr;
          tb.push_back(t);
          maxly = std::max(maxly, ly);
          maxlx = std::max(maxlx, lx);
          minly = std::min(minly, ly);
        }
      }
      const int nj = (int)pj.size();
      if ((size_t)mvw > mv_cap || (size_t)tw > mtop_cap) {  // grow-only step-word buffers
        if (mv) hipFree(mv);
        if (mtop) hipFree(mtop);
        mv = nullptr;
        mtop = nullptr;
        mv_cap = std::max<size_t>(mv_cap, (size_t)std::max(1LL, mvw));
        mtop_cap = std::max<size_t>(mtop_cap, (size_t)std::max(1LL, tw));
        if (hipMalloc((void**)&mv, mv_cap * 4) != hipSuccess || hipMalloc((void**)&mtop, mtop_cap * 4) != hipSuccess)
          return fail(c, ODBA_E_NOMEM, "cannot allocate %zu step words", mv_cap);
      }
      HIPC(c, hipMemcpyAsync(jobs, pj.data(), nj * sizeof(PJ<T>), hipMemcpyHostToDevice, c->stream));
      HIPC(c, hipMemcpyAsync(tbj, tb.data(), nj * sizeof(TbJob), hipMemcpyHostToDevice, c->stream));
      HIPC(c, hipMemcpyAsync(dslots, slots.data(), slots.size() * sizeof(MultiSlot), hipMemcpyHostToDevice, c->stream));
      HIPC(c, hipMemsetAsync(dbits, 0, nc * sizeof(unsigned long long), c->stream));
      HIPC(c, hipMemsetAsync(derr, 0, sizeof(int), c->stream));
      rc = run_path_jobs(c, dtype_of<T>(), no_wrap, os, oe, jobs, nj, maxly, mv, mtop, abr, maxlx, minly, derr,
                         mean_swaths(pj), path_tail(pj));
      if (rc) return rc;
      HIPC(c, launch_traceback(c->stream, tbj, nj, mv, mtop, abr, runs, 0, os, oe, derr, nullptr));
      HIPC(c, launch_multi_mean(dtype_of<T>(), c->stream, dslots, (int)slots.size(), pos, acc, counts, dbits, derr));
      int herr = 0;
      HIPC(c, hipMemcpyAsync(&herr, derr, sizeof(int), hipMemcpyDeviceToHost, c->stream));
      HIPC(c, hipMemcpyAsync(bits.data(), dbits, nc * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
      for (size_t q = a0; q < a1; q++) {
        const int k = act[q];
        HIPC(c, hipMemcpyAsync(nb[k].data(), newc + cen_off[k], lcs[k] * sizeof(T), hipMemcpyDeviceToHost, c->stream));
      }
      HIPC(c, hipStreamSynchronize(c->stream));
      if (herr & 8) return fail(c, ODBA_E_HIP, "path band edge hand-off timed out");
    if (herr & 1) return fail(c, ODBA_E_TRACEBACK, "traceback left the matrix before reaching (0,0)");
      if (herr & 4) return fail(c, ODBA_E_TRACEBACK, "traceback produced a third counted run for one centroid position");
      if (herr & 2) return fail(c, ODBA_E_DIV_ZERO, "integer centroid position with zero count");
      for (size_t q = a0; q < a1; q++) {  // dba.hpp:1084-1110, per cluster
        const int k = act[q];
        rounds_out[k] = r + 1;
        double d;
        std::memcpy(&d, &bits[k], sizeof(d));
        // the next round starts from the new centroid
        HIPC(c, hipMemcpyAsync(cen + cen_off[k], newc + cen_off[k], lcs[k] * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
        if (d == 0) { active[k] = 0; continue; }
        if (os || oe) {
          if (r >= 1 && !std::memcmp(nb[k].data(), prev2[k].data(), sizeof(T) * lcs[k])) { active[k] = 0; continue; }
          if (r > 1) prev2[k] = prev[k];
          if (r > 0) prev[k] = nb[k];
        }
      }
      a0 = a1;
    }
  }
  HIPC(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < nc; k++) {
    T* out = cen_out + cen_offs[k];
    const int m = medoids[k];
    for (int i = 0; i < lcs[k]; i++)
      out[i] = (mu && sg) ? host_d2t<T>(mu[m] + nb[k][i] * sg[m]) : nb[k][i];
  }
  return ODBA_OK;
}

}  // namespace

// A prepared DBA job (odba_job_prepare) and its result once its rounds ended.
struct odba_job {
  int dtype = 0;
  int device = 0;
  int max_rounds = 250;
  void* prep = nullptr;          // Prepared<T>*
  std::vector<void*> dev;        // the job's round buffers (step words, centroids, run records), freed with the job
  uint32_t* words = nullptr;     // step words of one round (odba_job_words)
  uint64_t word_bytes = 0;
  std::vector<unsigned char> cen;
  uint64_t clen = 0;
  int med = -1, rounds = 0, rc = ODBA_OK;
  bool done = false;
};

namespace {

// ---------------------------------------------------------------------------
// Packed DBA rounds across jobs (continuous batching of a server). Jobs of
// one value type and alignment mode, each prepared on its own (sequences
// resident, medoid found), run their rounds together: every step packs one
// round of every active job into shared launches - path matrices,
// tracebacks, then accumulation and mean/delta over all their centroid
// positions - in batches of whole jobs whose step words fit the engine's
// budget. Each job keeps its own centroid buffers, run records and stop
// rule (dba.hpp:1084-1110), and the accumulation adds each job's values in
// its own sequence order, so a job's result equals perform_dba_t's for the
// same input whatever else is packed with it. Jobs join before a step and
// leave when their stop rule ends them.
// ---------------------------------------------------------------------------
// Geometry of member q of a prepared job: flip, rows, columns (dba.hpp:497-501).
template <typename T> void member_shape(const Prepared<T>& P, int q, int& flip, int& ly, int& lx) {
  flip = (P.oe && P.lc < P.hl[q]) ? 1 : 0;
  ly = flip ? P.lc : P.hl[q];
  lx = flip ? P.hl[q] : P.lc;
}

// Step words of one round of a prepared job: every member's path step
// words (one swath_words block per 1024-column swath) in member order, then
// every member's open-end top-row words (64 per swath). mo / to (optional)
// receive each member's word offsets; returns the total in words.
template <typename T> long long job_word_layout(const Prepared<T>& P, std::vector<long long>* mo,
                                                std::vector<long long>* to) {
  long long mvw = 0, topw = 0;
  if (mo) mo->assign(P.n, 0);
  if (to) to->assign(P.n, 0);
  for (int q = 0; q < P.n; q++) {
    int flip, ly, lx;
    member_shape(P, q, flip, ly, lx);
    const int K = (lx + ODBA_W - 1) / ODBA_W;
    if (mo) (*mo)[q] = mvw;
    if (to) (*to)[q] = topw;
    mvw += (long long)K * swath_words(ly, P.os, P.oe);
    topw += (long long)K * 64;
  }
  if (to)
    for (long long& t : *to) t += mvw;
  return mvw + topw;
}

template <typename T> struct EJob {
  odba_job* h = nullptr;
  Prepared<T>* P = nullptr;
  std::vector<long long> mo, to;  // member word offsets in the job's word buffer
  T* cb[2] = {nullptr, nullptr};  // centroid buffers (padded): a round reads cb[cur] and writes cb[cur ^ 1]
  int cur = 0;
  int4* runs = nullptr;           // run records [member][position]
  int rounds = 0;
  std::vector<T> nb;
  StopRule<T> stop;
};

template <typename T> struct RoundEngine {
  // Path jobs are launched in groups of similar swath count, each on its
  // own stream with its traceback behind it: a strip workgroup is as wide
  // as the widest job of its launch, so one-swath jobs packed with
  // six-swath jobs would leave most of their workgroup's waves idle.
  // Jobs of more than 16 swaths (more than a strip workgroup's waves) go to
  // the swath-band schedule, so the longest pair of a step spreads over
  // CUs instead of setting the step's length on one workgroup.
  static constexpr int kGroups = 4;
  static int group_of(int K) { return K <= 1 ? 0 : (K <= 8 ? 1 : (K <= 16 ? 2 : 3)); }
  // Swath bands (four per CU) for the group wider than 16 swaths and, in
  // open end without open start, for the group of 9-16 swaths too: their
  // second-phase claims skip bands past each job's stop, which strips
  // inside one workgroup cannot (13 open-end serving jobs: 82-83 ms per
  // step alone against 85-87 ms with that group as strips).
  int bands_per_cu(int g) const { return (g == 3 || (g == 2 && oe && !os)) ? 4 : 0; }

  odba_ctx* c = nullptr;
  odba_ctx* sub[kGroups] = {};
  hipEvent_t ev_go = nullptr, ev_done[kGroups] = {};
  int os = 0, oe = 0;
  std::vector<std::unique_ptr<EJob<T>>> act;
  std::vector<odba_job*> finished;
  bool dirty = true;
  long long nstep = 0, s0 = 0;
  struct Group {
    int nj = 0;
    int maxly = 1, maxlx = 0, minly = INT32_MAX;
    double msw = 0, tail = 0;
    double est = 0;  // expected length: the longest member's row-swaths, or all of them over four waves per CU
    bool minfirst = true;
    size_t pj_at[2] = {0, 0}, tb_at = 0;  // byte offsets into the table arena
  };
  Group grp[kGroups];
  int order[kGroups] = {0, 1, 2, 3};  // launch order: longest expected group first
  int ns = 0;
  long long npos = 0;
  size_t sl_at[2] = {0, 0};
  uint32_t* base = nullptr;          // lowest job word buffer: the launches' word base
  std::vector<unsigned char> htab;  // host image of the table arena (kept until the next rebuild)
  unsigned char* tab = nullptr;
  size_t tab_cap = 0;
  int* abr[kGroups] = {};
  size_t abr_cap[kGroups] = {};
  T* acc = nullptr;
  size_t acc_cap = 0;
  uint32_t* counts = nullptr;
  size_t counts_cap = 0;
  unsigned long long* dbits = nullptr;
  size_t dbits_cap = 0;
  int* derr = nullptr;
  int* herr = nullptr;               // pinned
  unsigned long long* hbits = nullptr;
  size_t hbits_cap = 0;

  int init(odba_ctx* ctx, int os_, int oe_) {
    c = ctx;
    os = os_;
    oe = oe_;
    for (int g = 0; g < kGroups; g++) {
      int rc = odba_ctx_create(c->device, &sub[g]);
      if (rc) return fail(c, rc, "cannot create a path stream: %s", sub[g] ? sub[g]->err.c_str() : "");
      sub[g]->graphs = false;
      HIPC(c, hipEventCreateWithFlags(&ev_done[g], hipEventDisableTiming));
    }
    HIPC(c, hipEventCreateWithFlags(&ev_go, hipEventDisableTiming));
    return ODBA_OK;
  }

  ~RoundEngine() {
    if (c) hipSetDevice(c->device);
    for (int g = 0; g < kGroups; g++) {
      if (abr[g]) hipFree(abr[g]);
      if (ev_done[g]) hipEventDestroy(ev_done[g]);
      if (sub[g]) odba_ctx_destroy(sub[g]);
    }
    if (ev_go) hipEventDestroy(ev_go);
    for (void* p : {(void*)tab, (void*)acc, (void*)counts, (void*)dbits, (void*)derr})
      if (p) hipFree(p);
    if (herr) hipHostFree(herr);
    if (hbits) hipHostFree(hbits);
  }

  // Grow-only device buffers, with headroom: a shrinking and growing job
  // mix would otherwise free (and so synchronise the device) every rebuild.
  template <typename P> int grow(P*& p, size_t& cap, size_t need) {
    if (need <= cap) return ODBA_OK;
    if (p) hipFree(p);
    p = nullptr;
    cap = 0;
    const size_t want = need + need / 4;
    if (hipMalloc((void**)&p, want * sizeof(P)) != hipSuccess) return fail(c, ODBA_E_NOMEM, "cannot allocate %zu bytes", want * sizeof(P));
    cap = want;
    return ODBA_OK;
  }

  int add(odba_job* h) {
    auto* P = (Prepared<T>*)h->prep;
    if (P->os != os || P->oe != oe) return fail(c, ODBA_E_INVALID, "job mode differs from the engine's");
    if (!h->words && odba_job_words(h, 1, nullptr) != ODBA_OK)
      return fail(c, ODBA_E_NOMEM, "cannot allocate %llu bytes of step words", (unsigned long long)h->word_bytes);
    auto J = std::make_unique<EJob<T>>();
    J->h = h;
    J->P = P;
    job_word_layout(*P, &J->mo, &J->to);
    const long long pl = padded_len(P->lc);
    for (int k = 0; k < 2; k++) {
      HIPC(c, hipMalloc((void**)&J->cb[k], pl * sizeof(T)));
      h->dev.push_back(J->cb[k]);
      HIPC(c, hipMemsetAsync(J->cb[k], 0, pl * sizeof(T), c->stream));
    }
    HIPC(c, hipMalloc((void**)&J->runs, std::max<size_t>(1, (size_t)P->n * P->lc) * sizeof(int4)));
    h->dev.push_back(J->runs);
    HIPC(c, hipMemcpyAsync(J->cb[0], P->d + P->ho[P->med], P->lc * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
    J->nb.assign(P->lc, T(0));
    J->stop.init(P->lc);
    act.push_back(std::move(J));
    dirty = true;
    return ODBA_OK;
  }

  void finish(size_t k, int rc) {
    EJob<T>& J = *act[k];
    odba_job* h = J.h;
    Prepared<T>& P = *J.P;
    if (P.norm)
      for (int i = 0; i < P.lc; i++) J.nb[i] = host_d2t<T>(P.mu[P.med] + J.nb[i] * P.sg[P.med]);  // dba.hpp:1117-1125
    h->cen.resize((size_t)P.lc * sizeof(T));
    std::memcpy(h->cen.data(), J.nb.data(), h->cen.size());
    h->clen = (uint64_t)P.lc;
    h->med = P.med;
    h->rounds = J.rounds;
    h->rc = rc;
    h->done = true;
    finished.push_back(h);
  }

  // Tables of one step: inside each group the path jobs are listed by
  // ascending work, so the path kernels (which claim from the end of the
  // list) start the longest first. Every job's step words lie in its own
  // buffer; the launches take the lowest of those buffers as their word
  // base and each member's word offsets from it.
  int rebuild() {
    htab.clear();
    size_t njmax[kGroups];
    for (int g = 0; g < kGroups; g++) {
      njmax[g] = 1;
      grp[g] = Group();
      order[g] = g;
    }
    struct Item { double cost; int j, q; };
    auto align = [&]() { htab.resize((htab.size() + 255) & ~(size_t)255); return htab.size(); };
    auto put = [&](const void* p, size_t bytes) {
      const size_t at = align();
      htab.insert(htab.end(), (const unsigned char*)p, (const unsigned char*)p + bytes);
      return at;
    };
    base = nullptr;
    for (const auto& J : act)
      if (!base || (uintptr_t)J->h->words < (uintptr_t)base) base = J->h->words;
    std::vector<Item> items[kGroups];
    for (size_t j = 0; j < act.size(); j++) {
      const Prepared<T>& P = *act[j]->P;
      for (int q = 0; q < P.n; q++) {
        int flip, ly, lx;
        member_shape(P, q, flip, ly, lx);
        const int K = (lx + ODBA_W - 1) / ODBA_W;
        items[group_of(K)].push_back(Item{(double)ly * (double)K, (int)j, q});
      }
    }
    for (int g = 0; g < kGroups; g++) {
      Group& G = grp[g];
      std::stable_sort(items[g].begin(), items[g].end(), [](const Item& a, const Item& z) { return a.cost < z.cost; });
      std::vector<PJ<T>> pj[2];
      std::vector<TbJob> tb;
      double total = 0, longest = 0;
      for (const Item& it : items[g]) {
        total += it.cost;
        longest = std::max(longest, it.cost);
        EJob<T>& J = *act[it.j];
        const Prepared<T>& P = *J.P;
        int flip, ly, lx;
        member_shape(P, it.q, flip, ly, lx);
        const long long at = (long long)(((uintptr_t)J.h->words - (uintptr_t)base) / sizeof(uint32_t));
        const long long mo = at + J.mo[it.q], to = at + J.to[it.q];
        const T* sq = P.d + P.ho[it.q];
        for (int v = 0; v < 2; v++) {
          const T* cp = J.cb[J.cur ^ v];
          pj[v].push_back(PJ<T>{flip ? cp : sq, flip ? sq : cp, ly, lx, mo, to});
        }
        TbJob t;
        t.ly = ly; t.lx = lx; t.flip = flip; t.seq = it.q; t.mv_off = mo; t.top_off = to; t.path_off = -1;
        t.runs_off = 0;
        t.rp = J.runs + (size_t)it.q * P.lc;
        tb.push_back(t);
        G.maxly = std::max(G.maxly, ly);
        G.maxlx = std::max(G.maxlx, lx);
        G.minly = std::min(G.minly, ly);
        G.minfirst = G.minfirst && P.no_wrap;
      }
      G.nj = (int)tb.size();
      if (!G.nj) continue;
      G.est = std::max(longest, total / (4.0 * std::max(1, c->cus)));
      G.msw = mean_swaths(pj[0]);
      G.tail = path_tail(pj[0]);
      for (int v = 0; v < 2; v++) G.pj_at[v] = put(pj[v].data(), pj[v].size() * sizeof(PJ<T>));
      G.tb_at = put(tb.data(), tb.size() * sizeof(TbJob));
      njmax[g] = std::max(njmax[g], (size_t)G.nj);
    }
    std::vector<MultiSlot> sl[2];
    long long pos = 0;
    for (size_t j = 0; j < act.size(); j++) {
      EJob<T>& J = *act[j];
      for (int v = 0; v < 2; v++) {
        MultiSlot m;
        m.pos0 = pos; m.data = J.P->d; m.moffs = J.P->o; m.runs = J.runs;
        m.cen = J.cb[J.cur ^ v]; m.newc = J.cb[J.cur ^ v ^ 1];
        m.lc = J.P->lc; m.nmem = J.P->n; m.slot = (int)j; m.pad = 0;
        sl[v].push_back(m);
      }
      pos += round64(J.P->lc);
    }
    // The group that takes longest starts first: a group launched earlier
    // takes CUs that a later, longer group then waits for.
    std::stable_sort(order, order + kGroups, [&](int a, int z) { return grp[a].est > grp[z].est; });
    ns = (int)act.size();
    npos = pos;
    for (int v = 0; v < 2; v++) sl_at[v] = put(sl[v].data(), sl[v].size() * sizeof(MultiSlot));
    int rc;
    if ((rc = grow(tab, tab_cap, std::max<size_t>(htab.size(), 256))) ||
        (rc = grow(acc, acc_cap, std::max<size_t>(pos, 1))) || (rc = grow(counts, counts_cap, std::max<size_t>(pos, 1))) ||
        (rc = grow(dbits, dbits_cap, act.size())))
      return rc;
    for (int g = 0; g < kGroups; g++)
      if ((rc = grow(abr[g], abr_cap[g], njmax[g]))) return rc;
    if (!derr) HIPC(c, hipMalloc((void**)&derr, sizeof(int)));
    if (!herr) HIPC(c, hipHostMalloc((void**)&herr, sizeof(int)));
    if (hbits_cap < act.size()) {
      if (hbits) hipHostFree(hbits);
      hbits = nullptr;
      hbits_cap = 0;
      HIPC(c, hipHostMalloc((void**)&hbits, act.size() * 2 * sizeof(unsigned long long)));
      hbits_cap = act.size() * 2;
    }
    HIPC(c, hipMemcpyAsync(tab, htab.data(), htab.size(), hipMemcpyHostToDevice, c->stream));
    return ODBA_OK;
  }

  // One round of every active job. Jobs whose stop rule ends them move to
  // `finished` with their results; an error fails every active job.
  int step() {
    if (act.empty()) return ODBA_OK;
    int rc = ODBA_OK;
    if (dirty) {
      rc = rebuild();
      if (rc) return fail_all(rc);
      s0 = nstep;
      dirty = false;
    }
    const int v = (int)((nstep - s0) & 1);
    rc = launch_round(v);
    if (rc) return fail_all(rc);
    const int herr_v = *herr;
    if (herr_v) {
      if (herr_v & 8) rc = fail(c, ODBA_E_HIP, "path band edge hand-off timed out");
      else if (herr_v & 1) rc = fail(c, ODBA_E_TRACEBACK, "traceback left the matrix before reaching (0,0)");
      else if (herr_v & 4) rc = fail(c, ODBA_E_TRACEBACK, "traceback produced a third counted run for one centroid position");
      else rc = fail(c, ODBA_E_DIV_ZERO, "integer centroid position with zero count");
      return fail_all(rc);
    }
    std::vector<std::unique_ptr<EJob<T>>> keep;
    for (size_t k = 0; k < act.size(); k++) {
      EJob<T>& J = *act[k];
      J.cur ^= 1;
      const int r = J.rounds++;
      double d;
      std::memcpy(&d, &hbits[k], sizeof(d));
      if (J.stop.done(r, d, J.nb, os, oe) || J.rounds >= J.h->max_rounds) {
        finish(k, ODBA_OK);
        dirty = true;
      } else {
        keep.push_back(std::move(act[k]));
      }
    }
    act.swap(keep);
    nstep++;
    return ODBA_OK;
  }

  int launch_round(int v) {
    HIPC(c, launch_fill_u32(c->stream, derr, 0u, 1));
    HIPC(c, launch_fill_u32(c->stream, dbits, 0u, 2 * (long long)act.size()));
    HIPC(c, hipEventRecord(ev_go, c->stream));
    for (int gi = 0; gi < kGroups; gi++) {
      const int g = order[gi];
      const Group& G = grp[g];
      if (!G.nj) continue;
      odba_ctx* sc = sub[g];
      HIPC(c, hipStreamWaitEvent(sc->stream, ev_go, 0));
      int rc = run_path_jobs(sc, dtype_of<T>(), G.minfirst, os, oe, tab + G.pj_at[v], G.nj, G.maxly, base, base, abr[g],
                             G.maxlx, G.minly, derr, G.msw, G.tail, bands_per_cu(g));
      if (rc) return fail(c, rc, "%s", sc->err.c_str());
      HIPC(c, launch_traceback(sc->stream, (const TbJob*)(tab + G.tb_at), G.nj, base, base, abr[g], nullptr, 0, os, oe,
                               derr, nullptr));
      HIPC(c, hipEventRecord(ev_done[g], sc->stream));
      HIPC(c, hipStreamWaitEvent(c->stream, ev_done[g], 0));
    }
    HIPC(c, launch_multi_mean(dtype_of<T>(), c->stream, (const MultiSlot*)(tab + sl_at[v]), ns, npos, acc, counts,
                              dbits, derr));
    HIPC(c, hipMemcpyAsync(herr, derr, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipMemcpyAsync(hbits, dbits, act.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
    for (auto& J : act)
      HIPC(c, hipMemcpyAsync(J->nb.data(), J->cb[J->cur ^ 1], J->P->lc * sizeof(T), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    return ODBA_OK;
  }

  int fail_all(int rc) {
    for (size_t k = 0; k < act.size(); k++) finish(k, rc);
    act.clear();
    dirty = true;
    return rc;
  }
};

}  // namespace

struct odba_rounds {
  odba_ctx* ctx = nullptr;
  int dtype = 0;
  void* eng = nullptr;  // RoundEngine<T>*
};

namespace {
template <typename T> int job_prepare_t(odba_ctx* c, const T* const* seqs, const uint64_t* lens, int n, int os, int oe,
                                        int norm, odba_job* h) {
  auto* P = new Prepared<T>();
  h->prep = P;
  return prepare_t<T>(c, seqs, lens, n, os, oe, norm, *P, nullptr);
}
template <typename T> void job_free_t(odba_job* h) { delete (Prepared<T>*)h->prep; }
template <typename T> int rounds_create_t(odba_ctx* c, int os, int oe, void** out) {
  auto* e = new RoundEngine<T>();
  const int rc = e->init(c, os, oe);
  if (rc) {
    delete e;
    return rc;
  }
  *out = e;
  return ODBA_OK;
}
template <typename T> int rounds_add_t(void* e, odba_job* h) { return ((RoundEngine<T>*)e)->add(h); }
template <typename T> int rounds_step_t(void* e, odba_job** fin, int cap, int* nfin, int* nact) {
  auto* E = (RoundEngine<T>*)e;
  const int rc = E->step();
  int k = 0;
  while (k < cap && !E->finished.empty()) {
    fin[k++] = E->finished.front();
    E->finished.erase(E->finished.begin());
  }
  if (nfin) *nfin = k;
  if (nact) *nact = (int)E->act.size() + (int)E->finished.size();
  return rc;
}
template <typename T> void rounds_free_t(void* e) { delete (RoundEngine<T>*)e; }
}  // namespace

extern "C" {

int odba_job_prepare(odba_ctx* c, int dtype, const void* const* seqs, const uint64_t* lens, int n, int use_open_start,
                     int use_open_end, int norm, int max_rounds, odba_job** out) {
  if (!c || !seqs || !lens || !out || max_rounds < 1) return fail(c, ODBA_E_INVALID, "invalid job arguments");
  if (dtype_size(dtype) == 0) return fail(c, ODBA_E_INVALID, "unknown dtype %d", dtype);
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, ODBA_E_HIP, "hipSetDevice");
  auto* h = new odba_job();
  h->dtype = dtype;
  h->device = c->device;
  h->max_rounds = max_rounds;
  int rc;
  switch (dtype) {
    case ODBA_I32: rc = job_prepare_t<int32_t>(c, (const int32_t* const*)seqs, lens, n, use_open_start, use_open_end, norm, h); break;
    case ODBA_U32: rc = job_prepare_t<uint32_t>(c, (const uint32_t* const*)seqs, lens, n, use_open_start, use_open_end, norm, h); break;
    case ODBA_U64: rc = job_prepare_t<uint64_t>(c, (const uint64_t* const*)seqs, lens, n, use_open_start, use_open_end, norm, h); break;
    case ODBA_F32: rc = job_prepare_t<float>(c, (const float* const*)seqs, lens, n, use_open_start, use_open_end, norm, h); break;
    default: rc = job_prepare_t<double>(c, (const double* const*)seqs, lens, n, use_open_start, use_open_end, norm, h); break;
  }
  if (rc) {
    odba_job_destroy(h);
    return rc;
  }
  *out = h;
  return ODBA_OK;
}

int odba_job_result(const odba_job* h, void* centroid_out, uint64_t capacity, uint64_t* centroid_len, int* medoid,
                    int* rounds, int* job_rc) {
  if (!h) return ODBA_E_INVALID;
  if (!h->done) return ODBA_E_BUSY;
  if (centroid_out) {
    if (capacity < h->clen) return ODBA_E_INVALID;
    std::memcpy(centroid_out, h->cen.data(), h->cen.size());
  }
  if (centroid_len) *centroid_len = h->clen;
  if (medoid) *medoid = h->med;
  if (rounds) *rounds = h->rounds;
  if (job_rc) *job_rc = h->rc;
  return ODBA_OK;
}

int odba_job_words(odba_job* h, int allocate, uint64_t* word_bytes) {
  if (!h || !h->prep) return ODBA_E_INVALID;
  if (!h->word_bytes) {
    long long w = 0;
    switch (h->dtype) {
      case ODBA_I32: w = job_word_layout(*(Prepared<int32_t>*)h->prep, nullptr, nullptr); break;
      case ODBA_U32: w = job_word_layout(*(Prepared<uint32_t>*)h->prep, nullptr, nullptr); break;
      case ODBA_U64: w = job_word_layout(*(Prepared<uint64_t>*)h->prep, nullptr, nullptr); break;
      case ODBA_F32: w = job_word_layout(*(Prepared<float>*)h->prep, nullptr, nullptr); break;
      default: w = job_word_layout(*(Prepared<double>*)h->prep, nullptr, nullptr); break;
    }
    h->word_bytes = (uint64_t)std::max(1LL, w) * 4;
  }
  if (word_bytes) *word_bytes = h->word_bytes;
  if (allocate && !h->words) {
    if (hipSetDevice(h->device) != hipSuccess) return ODBA_E_HIP;
    void* p = nullptr;
    if (hipMalloc(&p, h->word_bytes) != hipSuccess) return ODBA_E_NOMEM;
    h->words = (uint32_t*)p;
    h->dev.push_back(p);
  }
  return ODBA_OK;
}

int odba_job_shape(const odba_job* h, uint64_t* centroid_len, int* max_swaths) {
  if (!h || !h->prep) return ODBA_E_INVALID;
  int lc = 0, kmax = 0;
  auto shape = [&](const auto& P) {
    lc = P.lc;
    for (int q = 0; q < P.n; q++) {
      int flip, ly, lx;
      member_shape(P, q, flip, ly, lx);
      kmax = std::max(kmax, (lx + ODBA_W - 1) / ODBA_W);
    }
  };
  switch (h->dtype) {
    case ODBA_I32: shape(*(const Prepared<int32_t>*)h->prep); break;
    case ODBA_U32: shape(*(const Prepared<uint32_t>*)h->prep); break;
    case ODBA_U64: shape(*(const Prepared<uint64_t>*)h->prep); break;
    case ODBA_F32: shape(*(const Prepared<float>*)h->prep); break;
    default: shape(*(const Prepared<double>*)h->prep); break;
  }
  if (centroid_len) *centroid_len = (uint64_t)lc;
  if (max_swaths) *max_swaths = kmax;
  return ODBA_OK;
}

int odba_job_destroy(odba_job* h) {
  if (!h) return ODBA_OK;
  hipSetDevice(h->device);
  for (void* p : h->dev)
    if (p) hipFree(p);
  switch (h->dtype) {
    case ODBA_I32: job_free_t<int32_t>(h); break;
    case ODBA_U32: job_free_t<uint32_t>(h); break;
    case ODBA_U64: job_free_t<uint64_t>(h); break;
    case ODBA_F32: job_free_t<float>(h); break;
    case ODBA_F64: job_free_t<double>(h); break;
  }
  delete h;
  return ODBA_OK;
}

int odba_rounds_create(odba_ctx* c, int dtype, int use_open_start, int use_open_end, odba_rounds** out) {
  if (!c || !out) return fail(c, ODBA_E_INVALID, "invalid rounds-engine arguments");
  if (dtype_size(dtype) == 0) return fail(c, ODBA_E_INVALID, "unknown dtype %d", dtype);
  auto* r = new odba_rounds();
  r->ctx = c;
  r->dtype = dtype;
  int rc;
  switch (dtype) {
    case ODBA_I32: rc = rounds_create_t<int32_t>(c, use_open_start, use_open_end, &r->eng); break;
    case ODBA_U32: rc = rounds_create_t<uint32_t>(c, use_open_start, use_open_end, &r->eng); break;
    case ODBA_U64: rc = rounds_create_t<uint64_t>(c, use_open_start, use_open_end, &r->eng); break;
    case ODBA_F32: rc = rounds_create_t<float>(c, use_open_start, use_open_end, &r->eng); break;
    default: rc = rounds_create_t<double>(c, use_open_start, use_open_end, &r->eng); break;
  }
  if (rc) {
    delete r;
    return rc;
  }
  *out = r;
  return ODBA_OK;
}

int odba_rounds_add(odba_rounds* r, odba_job* job) {
  if (!r || !job || job->dtype != r->dtype || job->done) return ODBA_E_INVALID;
  odba_ctx* c = r->ctx;
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, ODBA_E_HIP, "hipSetDevice");
  ODBA_DISPATCH(r->dtype, rounds_add_t<T>(r->eng, job))
}

int odba_rounds_step(odba_rounds* r, odba_job** finished, int capacity, int* n_finished, int* n_active) {
  if (!r || (capacity > 0 && !finished)) return ODBA_E_INVALID;
  odba_ctx* c = r->ctx;
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, ODBA_E_HIP, "hipSetDevice");
  ODBA_DISPATCH(r->dtype, rounds_step_t<T>(r->eng, finished, capacity, n_finished, n_active))
}

int odba_rounds_destroy(odba_rounds* r) {
  if (!r) return ODBA_OK;
  hipSetDevice(r->ctx->device);
  switch (r->dtype) {
    case ODBA_I32: rounds_free_t<int32_t>(r->eng); break;
    case ODBA_U32: rounds_free_t<uint32_t>(r->eng); break;
    case ODBA_U64: rounds_free_t<uint64_t>(r->eng); break;
    case ODBA_F32: rounds_free_t<float>(r->eng); break;
    case ODBA_F64: rounds_free_t<double>(r->eng); break;
  }
  delete r;
  return ODBA_OK;
}

}  // extern "C"

namespace {

// chopPrefixFromSequences (dba.hpp:1192-1395): every sequence is aligned
// against the prefix (prefix on Y, open end, W = 1024) as one batch of path
// jobs, and the DBA traceback with the prefix as the "centroid" on Y gives,
// per prefix row, the counted sequence columns. Its counts are the
// reference's leader-path histogram (every path cell except the top row's
// open-right run, plus (0,0); dba.hpp:1317-1336) and the first counted
// column of the top row is the chop position. With norm, the prefix (a
// copy) and the sequences (in place) are z-normalised first
// (dba.hpp:1208-1216).
template <typename T>
int prefix_chop_t(odba_ctx* c, const T* prefix_in, int P, T* const* seqs, const uint64_t* lens, int n, int norm,
                  uint64_t* chop, int32_t* hist) {
  if (P < 2) return fail(c, ODBA_E_INVALID, "prefix length %d: at least 2 is needed", P);
  if (n < 1) return fail(c, ODBA_E_INVALID, "no sequences");
  for (int i = 0; i < n; i++)
    if (lens[i] == 0 || lens[i] > (uint64_t)INT32_MAX / 2) return fail(c, ODBA_E_INVALID, "sequence %d has length %llu", i, (unsigned long long)lens[i]);
  std::vector<T> pv(prefix_in, prefix_in + P);
  if (norm) {
    int rc = znorm_t<T>(c, seqs, lens, n, nullptr, nullptr);
    if (rc) return rc;
    T* pp = pv.data();
    const uint64_t pl = (uint64_t)P;
    rc = znorm_t<T>(c, &pp, &pl, 1, nullptr, nullptr);
    if (rc) return rc;
  }
  bool no_wrap = true;
  if (!is_float<T>()) {
    std::vector<const T*> all(seqs, seqs + n);
    std::vector<uint64_t> al(lens, lens + n);
    all.push_back(pv.data());
    al.push_back((uint64_t)P);
    no_wrap = int_cannot_wrap<T>(all.data(), al.data(), n + 1);
  }
  std::vector<long long> ho(n);
  long long total = padded_len(P);
  int maxlx = 0;
  for (int i = 0; i < n; i++) {
    ho[i] = total;
    total += padded_len((long long)lens[i]);
    maxlx = std::max(maxlx, (int)lens[i]);
  }
  std::vector<T> hb((size_t)total, T(0));
  std::memcpy(hb.data(), pv.data(), P * sizeof(T));
  for (int i = 0; i < n; i++) std::memcpy(&hb[ho[i]], seqs[i], lens[i] * sizeof(T));
  DevBufs B;
  T* data;
  int4* runs;
  int *abr, *derr;
  TbJob* tbj;
  unsigned char* jobs;
  int rc;
  if ((rc = B.get(c, data, hb.size())) || (rc = B.get(c, runs, (size_t)n * P)) || (rc = B.get(c, abr, n)) ||
      (rc = B.get(c, derr, 1)) || (rc = B.get(c, tbj, n)) || (rc = B.get(c, jobs, (size_t)n * sizeof(PJ<T>))))
    return rc;
  HIPC(c, hipMemcpyAsync(data, hb.data(), hb.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
  const long long budget_words = (long long)(c->mem_cap / 4 / 4);
  uint32_t* mv = nullptr;
  uint32_t* mtop = nullptr;
  size_t mv_cap = 0, mtop_cap = 0;
  struct FreeWords {
    uint32_t*& a; uint32_t*& b;
    ~FreeWords() { if (a) hipFree(a); if (b) hipFree(b); }
  } free_words{mv, mtop};
  std::vector<int4> hr((size_t)n * P);
  int b0 = 0;
  while (b0 < n) {  // batches whose step words fit the budget
    std::vector<PJ<T>> pj;
    std::vector<TbJob> tb;
    long long mvw = 0, tw = 0;
    int b1 = b0, bmaxlx = 0;
    for (; b1 < n; b1++) {
      const int lx = (int)lens[b1];
      const long long need = (long long)((lx + ODBA_W - 1) / ODBA_W) * swath_words(P, 0, 1);
      if (b1 > b0 && mvw + need > budget_words) break;
      long long mo, to;
      plan_job(P, lx, 0, 1, mvw, tw, mo, to);
      pj.push_back(PJ<T>{data, data + ho[b1], P, lx, mo, to});
      TbJob t;
      t.ly = P; t.lx = lx; t.flip = 1; t.seq = b1 - b0; t.mv_off = mo; t.top_off = to; t.path_off = -1;
      t.runs_off = (long long)(b1 - b0) * P;
      t.rp = nullptr;
      tb.push_back(t);
      bmaxlx = std::max(bmaxlx, lx);
    }
    const int nj = b1 - b0;
    if ((size_t)mvw > mv_cap || (size_t)tw > mtop_cap) {  // grow-only step-word buffers
      if (mv) hipFree(mv);
      if (mtop) hipFree(mtop);
      mv = nullptr;
      mtop = nullptr;
      mv_cap = std::max<size_t>(mv_cap, (size_t)std::max(1LL, mvw));
      mtop_cap = std::max<size_t>(mtop_cap, (size_t)std::max(1LL, tw));
      if (hipMalloc((void**)&mv, mv_cap * 4) != hipSuccess || hipMalloc((void**)&mtop, mtop_cap * 4) != hipSuccess)
        return fail(c, ODBA_E_NOMEM, "cannot allocate %zu step words", mv_cap);
    }
    HIPC(c, hipMemcpyAsync(jobs, pj.data(), nj * sizeof(PJ<T>), hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipMemcpyAsync(tbj, tb.data(), nj * sizeof(TbJob), hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipMemsetAsync(derr, 0, sizeof(int), c->stream));
    rc = run_path_jobs(c, dtype_of<T>(), no_wrap, 0, 1, jobs, nj, P, mv, mtop, abr, bmaxlx, P, derr, mean_swaths(pj), path_tail(pj));
    if (rc) return rc;
    HIPC(c, launch_traceback(c->stream, tbj, nj, mv, mtop, abr, runs, P, 0, 1, derr, nullptr));
    int herr = 0;
    HIPC(c, hipMemcpyAsync(&herr, derr, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipMemcpyAsync(&hr[(size_t)b0 * P], runs, (size_t)nj * P * sizeof(int4), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    if (herr & 8) return fail(c, ODBA_E_HIP, "path band edge hand-off timed out");
    if (herr & 1) return fail(c, ODBA_E_TRACEBACK, "traceback left the matrix before reaching (0,0)");
    if (herr & 4) return fail(c, ODBA_E_TRACEBACK, "traceback produced a third counted run for one prefix row");
    b0 = b1;
  }
  for (int b = 0; b < n; b++) {
    const int4* r = &hr[(size_t)b * P];
    for (int i = 0; i < P; i++)
      hist[(size_t)b * P + i] = (r[i].y >= r[i].x ? r[i].y - r[i].x + 1 : 0) + (r[i].w >= r[i].z ? r[i].w - r[i].z + 1 : 0);
    if (r[P - 1].y < r[P - 1].x) return fail(c, ODBA_E_TRACEBACK, "sequence %d: no counted cell in the prefix's last row", b);
    chop[b] = (uint64_t)r[P - 1].y;
  }
  return ODBA_OK;
}

}  // namespace

extern "C" {

int odba_dtw_path(odba_ctx* c, int dtype, const void* y, uint64_t ly, const void* x, uint64_t lx, int os, int oe,
                  uint8_t* moves_out, void* dist_out) {
  if (!c || !y || !x || ly == 0 || lx == 0) return fail(c, ODBA_E_INVALID, "invalid path arguments");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, ODBA_E_HIP, "hipSetDevice");
  ODBA_DISPATCH(dtype, dtw_path_t<T>(c, (const T*)y, (int)ly, (const T*)x, (int)lx, os, oe, moves_out, (T*)dist_out))
}

int odba_dba_round(odba_ctx* c, int dtype, const void* centroid, uint64_t lc, const void* const* seqs,
                   const uint64_t* lens, int n, int os, int oe, int path_policy, uint64_t ref_free_mem,
                   void* new_centroid, uint32_t* counts_out, double* max_delta) {
  if (!c || !centroid || !seqs || !lens || n <= 0 || lc == 0 || !new_centroid) return fail(c, ODBA_E_INVALID, "invalid DBA round arguments");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, ODBA_E_HIP, "hipSetDevice");
  if (path_policy < ODBA_PATH_AUTO || path_policy > ODBA_PATH_STRIPE) return fail(c, ODBA_E_INVALID, "unknown path policy %d", path_policy);
  ODBA_DISPATCH(dtype, dba_round_t<T>(c, (const T*)centroid, (int)lc, (const T* const*)seqs, lens, n, os, oe,
                                      path_policy, ref_free_mem, (T*)new_centroid, counts_out, max_delta))
}

int odba_dba_clusters(odba_ctx* c, int dtype, const void* const* seqs, const uint64_t* lens, int n,
                      const int* memberships, const int* medoids, int num_clusters, int use_open_start,
                      int use_open_end, int max_rounds, const double* mu, const double* sigma, void* centroids_out,
                      uint64_t* centroid_offs, int* rounds_out) {
  if (!c || !seqs || !lens || !memberships || !medoids || !centroids_out || !centroid_offs || !rounds_out || n <= 0 ||
      num_clusters <= 0 || max_rounds < 0)
    return fail(c, ODBA_E_INVALID, "invalid cluster DBA arguments");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, ODBA_E_HIP, "hipSetDevice");
  ODBA_DISPATCH(dtype, dba_clusters_t<T>(c, (const T* const*)seqs, lens, n, memberships, medoids, num_clusters,
                                         use_open_start, use_open_end, max_rounds, mu, sigma, (T*)centroids_out,
                                         centroid_offs, rounds_out))
}

int odba_znorm(odba_ctx* c, int dtype, void* const* seqs, const uint64_t* lens, int n, double* mu, double* sigma) {
  if (!c || !seqs || !lens || n <= 0) return fail(c, ODBA_E_INVALID, "invalid z-norm arguments");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, ODBA_E_HIP, "hipSetDevice");
  ODBA_DISPATCH(dtype, znorm_t<T>(c, (T* const*)seqs, lens, n, mu, sigma))
}

int odba_prefix_chop(odba_ctx* c, int dtype, const void* prefix, uint64_t prefix_len, void* const* seqs,
                     const uint64_t* lens, int n, int norm, uint64_t* chop_pos, int32_t* hist) {
  if (!c || !prefix || !seqs || !lens || !chop_pos || !hist) return fail(c, ODBA_E_INVALID, "invalid prefix-chop arguments");
  if (prefix_len > (uint64_t)INT32_MAX / 2) return fail(c, ODBA_E_INVALID, "prefix too long");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, ODBA_E_HIP, "hipSetDevice");
  ODBA_DISPATCH(dtype, prefix_chop_t<T>(c, (const T*)prefix, (int)prefix_len, (T* const*)seqs, lens, n, norm, chop_pos, hist))
}

int odba_perform_dba(odba_ctx* c, int dtype, const void* const* seqs, const uint64_t* lens, int n, int os, int oe,
                     int norm, int max_rounds, int path_policy, uint64_t ref_free_mem, void* centroid_out,
                     uint64_t* centroid_len, int* medoid_sorted, int* perm_out, int* rounds_out, void* dists_out) {
  if (!c || !seqs || !lens || !centroid_out) return fail(c, ODBA_E_INVALID, "invalid performDBA arguments");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, ODBA_E_HIP, "hipSetDevice");
  if (path_policy < ODBA_PATH_AUTO || path_policy > ODBA_PATH_STRIPE) return fail(c, ODBA_E_INVALID, "unknown path policy %d", path_policy);
  ODBA_DISPATCH(dtype, perform_dba_t<T>(c, (const T* const*)seqs, lens, n, os, oe, norm, max_rounds, path_policy,
                                        ref_free_mem, (T*)centroid_out,
                                        centroid_len, medoid_sorted, perm_out, rounds_out, (T*)dists_out))
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Adaptive k-segmentation (adaptive_segmentation<T>, segmentation.hpp:339-583)
// ---------------------------------------------------------------------------
namespace {

int kseg_mk(int minseg) {
  const int ds = (minseg + 2) / 3;
  const float x = ((float)minseg) / ds;
  return (int)((256 + x - 1) / x);  // DIV_ROUNDUP with a float divisor (segmentation.hpp:350)
}

// Host post-pass over one sequence's padded medians (segmentation.hpp:486-578):
// epsilon = smallest non-zero step between neighbours produced together,
// equal neighbours collapsed, chunk-boundary neighbours closer than epsilon
// averaged, unused (max-valued) slots dropped, a trailing half-max value
// dropped, then the prefix skip.
template <typename T>
long kseg_post(T* seg, size_t SL, int prefix_skip, long* skipped_out) {
  const T MX = std::numeric_limits<T>::max();
  T eps = MX;
  T prev = seg[0];
  for (size_t j = 1; j < SL; ++j) {
    const T cv = seg[j];
    if (cv != MX && prev != MX && ((cv <= prev && (T)(prev - cv) < eps) || (cv > prev && (T)(cv - prev) < eps))) {
      if ((T)(cv - prev) == 0) {
        for (size_t k = j; k < SL; k++) {
          seg[k - 1] = seg[k];
          if (seg[k] == MX) break;
        }
      } else {
        eps = cv < prev ? (T)(prev - cv) : (T)(cv - prev);
      }
    }
    prev = cv;
  }
  bool undef = false;
  long cursor = 0;
  for (size_t j = 0; j < SL; ++j) {
    if (seg[j] == MX) { undef = true; continue; }
    if (undef) {
      undef = false;
      if (cursor >= 1 && seg[cursor - 1] < (T)(seg[j] + eps) && seg[cursor - 1] > (T)(seg[j] - eps))
        seg[cursor - 1] = (T)((seg[cursor - 1] + seg[j]) / 2);
    }
    if ((size_t)cursor != j) seg[cursor] = seg[j];
    cursor++;
  }
  if (cursor >= 1 && seg[cursor - 1] >= MX / 2) cursor--;
  long skipped = 0;
  if (prefix_skip) {
    if (cursor <= prefix_skip) { skipped = cursor - 1; cursor = 1; }
    else { skipped = prefix_skip; cursor -= prefix_skip; }
  }
  if (cursor < 0) cursor = 0;
  *skipped_out = skipped < 0 ? 0 : skipped;
  return cursor;
}

template <typename T>
int kseg_t(odba_ctx* c, const T* const* seqs, const uint64_t* lens, int n, int minseg, int prefix_skip, int lds_bytes,
           T* out, uint64_t cap, uint64_t* out_lens, uint64_t* out_offs) {
  const int ds = (minseg + 2) / 3;
  const int spb = 256 * ds;
  const int mk = kseg_mk(minseg);
  std::vector<long long> offs(n), obase(n), pbase(n);
  std::vector<int> L(n);
  long long total = 0, ototal = 0, ptotal = 0;
  int maxch = 1;
  for (int i = 0; i < n; i++) {
    offs[i] = total;
    L[i] = (int)lens[i];
    total += padded_len((long long)lens[i]) + spb;
    const int nch = (int)((lens[i] + spb - 1) / spb);
    maxch = std::max(maxch, nch);
    obase[i] = ototal;
    ototal += (long long)nch * mk;
    pbase[i] = ptotal;
    ptotal += (long long)nch * 256 * mk;
  }
  std::vector<T> hb((size_t)total, T(0));
  for (int i = 0; i < n; i++) std::memcpy(&hb[offs[i]], seqs[i], lens[i] * sizeof(T));
  T *d = nullptr, *dout = nullptr;
  long long *doffs = nullptr, *dob = nullptr, *dpb = nullptr;
  int* dl = nullptr;
  unsigned short* kp = nullptr;
  HIPC(c, hipMalloc((void**)&d, hb.size() * sizeof(T)));
  HIPC(c, hipMalloc((void**)&dout, std::max(1LL, ototal) * sizeof(T)));
  HIPC(c, hipMalloc((void**)&doffs, n * sizeof(long long)));
  HIPC(c, hipMalloc((void**)&dob, n * sizeof(long long)));
  HIPC(c, hipMalloc((void**)&dpb, n * sizeof(long long)));
  HIPC(c, hipMalloc((void**)&dl, n * sizeof(int)));
  HIPC(c, hipMalloc((void**)&kp, std::max(1LL, ptotal) * sizeof(unsigned short)));
  HIPC(c, hipMemcpyAsync(d, hb.data(), hb.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemcpyAsync(doffs, offs.data(), n * sizeof(long long), hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemcpyAsync(dob, obase.data(), n * sizeof(long long), hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemcpyAsync(dpb, pbase.data(), n * sizeof(long long), hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemcpyAsync(dl, L.data(), n * sizeof(int), hipMemcpyHostToDevice, c->stream));
  const int e = kseg_extent_host(mk, 256, (int)sizeof(T), lds_bytes);
  if (e < 1) return fail(c, ODBA_E_INVALID, "lds_bytes %d leaves no room for the segmentation tables", lds_bytes);
  const size_t dyn = std::max((size_t)256 * e * 6, (size_t)(spb + 2) * sizeof(T));
  if (dyn > 150 * 1024) return fail(c, ODBA_E_UNSUPPORTED, "k-segmentation needs %zu bytes of LDS", dyn);
  HIPC(c, launch_kseg(dtype_of<T>(), c->stream, d, doffs, dl, n, maxch, dob, dpb, spb, mk, minseg, lds_bytes, kp, dout, dyn));
  std::vector<T> hm((size_t)std::max(1LL, ototal));
  HIPC(c, hipMemcpyAsync(hm.data(), dout, std::max(1LL, ototal) * sizeof(T), hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  hipFree(d); hipFree(dout); hipFree(doffs); hipFree(dob); hipFree(dpb); hipFree(dl); hipFree(kp);
  uint64_t cur = 0;
  for (int i = 0; i < n; i++) {
    const size_t SL = (size_t)((lens[i] + spb - 1) / spb) * mk;
    long skipped = 0;
    const long len = kseg_post<T>(&hm[obase[i]], SL, prefix_skip, &skipped);
    if (cur + (uint64_t)len > cap) return fail(c, ODBA_E_INVALID, "k-segmentation output capacity %llu too small", (unsigned long long)cap);
    std::memcpy(out + cur, &hm[obase[i] + skipped], (size_t)len * sizeof(T));
    out_offs[i] = cur;
    out_lens[i] = (uint64_t)len;
    cur += (uint64_t)len;
  }
  return ODBA_OK;
}

}  // namespace

extern "C" {

int odba_kseg(odba_ctx* c, int dtype, const void* const* seqs, const uint64_t* lens, int n, int min_segment_length,
              int prefix_skip, int lds_bytes, void* out, uint64_t out_capacity, uint64_t* out_lens,
              uint64_t* out_offs) {
  if (!c || !seqs || !lens || n <= 0 || !out || !out_lens || !out_offs || min_segment_length <= 0)
    return fail(c, ODBA_E_INVALID, "invalid k-segmentation arguments");
  for (int i = 0; i < n; i++)
    if (lens[i] == 0) return fail(c, ODBA_E_INVALID, "sequence %d is empty", i);
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, ODBA_E_HIP, "hipSetDevice");
  ODBA_DISPATCH(dtype, kseg_t<T>(c, (const T* const*)seqs, lens, n, min_segment_length, prefix_skip, lds_bytes,
                                 (T*)out, out_capacity, out_lens, out_offs))
}

uint64_t odba_kseg_capacity(const uint64_t* lens, int n, int min_segment_length) {
  if (!lens || n <= 0 || min_segment_length <= 0) return 0;
  const int ds = (min_segment_length + 2) / 3;
  const uint64_t spb = 256ull * ds;
  const uint64_t mk = (uint64_t)kseg_mk(min_segment_length);
  uint64_t t = 0;
  for (int i = 0; i < n; i++) t += mk * ((lens[i] + spb - 1) / spb);
  return t;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Device-resident DBA session: member sequences and the centroid stay in
// HBM across rounds (the reference re-reads managed memory every round).
// ---------------------------------------------------------------------------
struct odba_dba {
  odba_ctx* ctx = nullptr;
  int dtype = 0;
  void* st = nullptr;  // DbaState<T>*
};

namespace {
template <typename T> int dba_create_t(odba_ctx* c, const T* const* seqs, const uint64_t* lens, int n, int os, int oe,
                                       int policy, uint64_t free_mem, void** out) {
  auto* st = new DbaState<T>();
  int rc = st->init(c, seqs, lens, n, os, oe);
  if (rc) { delete st; return rc; }
  st->policy = policy;
  st->free_mem = auto_free_mem(c, policy, free_mem);
  *out = st;
  return ODBA_OK;
}
template <typename T> int dba_destroy_t(void* p) { delete (DbaState<T>*)p; return ODBA_OK; }
template <typename T> int dba_set_t(void* p, const T* c, int lc) { return ((DbaState<T>*)p)->set_centroid(c, lc); }
template <typename T> int dba_step_t(odba_ctx* c, void* p, T* newc, uint32_t* counts, double* delta) {
  auto* st = (DbaState<T>*)p;
  int rc = st->round(delta);
  if (rc) return rc;
  if (newc) HIPC(c, hipMemcpyAsync(newc, st->cen, st->lc * sizeof(T), hipMemcpyDeviceToHost, c->stream));
  if (counts) HIPC(c, hipMemcpyAsync(counts, st->counts, st->lc * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return ODBA_OK;
}
}  // namespace

extern "C" {

int odba_dba_create(odba_ctx* c, int dtype, const void* const* seqs, const uint64_t* lens, int n, int use_open_start,
                    int use_open_end, int path_policy, uint64_t ref_free_mem, odba_dba** out) {
  if (!c || !seqs || !lens || n <= 0 || !out) return fail(c, ODBA_E_INVALID, "invalid DBA session arguments");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, ODBA_E_HIP, "hipSetDevice");
  auto* d = new odba_dba();
  d->ctx = c;
  d->dtype = dtype;
  int rc;
  switch (dtype) {
    case ODBA_I32: rc = dba_create_t<int32_t>(c, (const int32_t* const*)seqs, lens, n, use_open_start, use_open_end, path_policy, ref_free_mem, &d->st); break;
    case ODBA_U32: rc = dba_create_t<uint32_t>(c, (const uint32_t* const*)seqs, lens, n, use_open_start, use_open_end, path_policy, ref_free_mem, &d->st); break;
    case ODBA_U64: rc = dba_create_t<uint64_t>(c, (const uint64_t* const*)seqs, lens, n, use_open_start, use_open_end, path_policy, ref_free_mem, &d->st); break;
    case ODBA_F32: rc = dba_create_t<float>(c, (const float* const*)seqs, lens, n, use_open_start, use_open_end, path_policy, ref_free_mem, &d->st); break;
    case ODBA_F64: rc = dba_create_t<double>(c, (const double* const*)seqs, lens, n, use_open_start, use_open_end, path_policy, ref_free_mem, &d->st); break;
    default: rc = fail(c, ODBA_E_INVALID, "unknown dtype %d", dtype);
  }
  if (rc) { delete d; return rc; }
  *out = d;
  return ODBA_OK;
}

int odba_dba_destroy(odba_dba* d) {
  if (!d) return ODBA_OK;
  switch (d->dtype) {
    case ODBA_I32: dba_destroy_t<int32_t>(d->st); break;
    case ODBA_U32: dba_destroy_t<uint32_t>(d->st); break;
    case ODBA_U64: dba_destroy_t<uint64_t>(d->st); break;
    case ODBA_F32: dba_destroy_t<float>(d->st); break;
    case ODBA_F64: dba_destroy_t<double>(d->st); break;
  }
  delete d;
  return ODBA_OK;
}

int odba_dba_set_centroid(odba_dba* d, const void* centroid, uint64_t lc) {
  if (!d || !centroid || lc == 0) return ODBA_E_INVALID;
  odba_ctx* c = d->ctx;
  ODBA_DISPATCH(d->dtype, dba_set_t<T>(d->st, (const T*)centroid, (int)lc))
}

int odba_dba_step(odba_dba* d, void* new_centroid, uint32_t* counts_out, double* max_delta) {
  if (!d) return ODBA_E_INVALID;
  odba_ctx* c = d->ctx;
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, ODBA_E_HIP, "hipSetDevice");
  ODBA_DISPATCH(d->dtype, dba_step_t<T>(c, d->st, (T*)new_centroid, counts_out, max_delta))
}

int odba_dba_record_paths(odba_dba* d, int on) {
  if (!d) return ODBA_E_INVALID;
  odba_ctx* c = d->ctx;
  ODBA_DISPATCH(d->dtype, (((DbaState<T>*)d->st)->rec = (on != 0), ODBA_OK))
}

int odba_dba_get_path(odba_dba* d, int seq, uint64_t* steps, uint64_t capacity, uint64_t* n_steps) {
  if (!d) return ODBA_E_INVALID;
  odba_ctx* c = d->ctx;
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, ODBA_E_HIP, "hipSetDevice");
  ODBA_DISPATCH(d->dtype, ((DbaState<T>*)d->st)->get_path(seq, steps, capacity, n_steps))
}

}  // extern "C"
