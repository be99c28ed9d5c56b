// DTW swath kernels for gfx950 (CDNA4): all-vs-all distances and
// path-storing DTW for the DBA update.
//
// Semantics follow DTWDistance<T> (dtw.hpp:39-259) exactly, including the
// swath-width-dependent open-end rules; the work decomposition is this
// file's own:
//
//   * One wavefront (64 lanes) sweeps one pair. The cost matrix is cut into
//     vertical swaths 1024 columns wide, aligned like the reference's
//     launches (swath k covers columns [1024k, 1024k+1024)). Inside a swath,
//     lane l owns C consecutive columns (C = 16 for full swaths, fewer for
//     the last, partial one) and the lanes advance one row per step with a
//     one-step skew (lane l works on row s - l at step s). The left
//     neighbour value crosses lanes with a DPP wave_shr:1; lane 0 takes the
//     previous swath's right edge column instead. Each step costs 4 VALU ops
//     per cell (sub, mul, min3, add) plus one DPP move per lane.
//   * The row values y[i] and the left edge column stream through a
//     128-entry LDS window per wave; the swath's right edge column is
//     collected in LDS and written to HBM every 64 rows. Nothing else of the
//     cost matrix leaves registers.
//   * The open-end top row (dtw.hpp:196-200, 172-186) is a prefix minimum
//     along the row, computed after the swath body as an in-lane scan plus a
//     cross-lane exclusive scan.
//   * The open-end early abrogation (dtw.hpp:74-129) is evaluated between
//     swaths from the running minimum of the edge column.
//   * A persistent grid pulls pairs from an atomic counter, largest second
//     sequence first.
//
// Path mode additionally writes a 2-bit step code per cell (DIAG=0,
// RIGHT=1, UP=2), one 32-bit word per lane per step, so every store of a
// step is a coalesced 256-byte row. OPEN_RIGHT, NIL and NIL_OPEN_RIGHT are
// positional (open-end top row, open-start row 0, abrogated swaths, the
// corner) and are reconstructed by the traceback.

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <algorithm>
#include <type_traits>
#include <float.h>
#include <limits.h>

#include "odba_dtw.hpp"

namespace odba {

// ---------------------------------------------------------------------------
// Per-type arithmetic
// ---------------------------------------------------------------------------
template <typename T> struct Ar;

template <> struct Ar<float> {
  static constexpr bool kFloat = true;
  __device__ static float maxv() { return FLT_MAX; }
  __device__ static float rowinf() { return __builtin_inff(); }
  __device__ static float mn(float a, float b) { return __builtin_fminf(a, b); }
  __device__ static float mn3(float a, float b, float c) { return __builtin_fminf(__builtin_fminf(a, b), c); }
  __device__ static float shr(float old, float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x138, 0xf, 0xf, false));
  }
  __device__ static float tofloat(float v) { return v; }
  __device__ static float fromfloat(float f) { return f; }
};

template <> struct Ar<double> {
  static constexpr bool kFloat = true;
  __device__ static double maxv() { return DBL_MAX; }
  __device__ static double rowinf() { return __builtin_inf(); }
  __device__ static double mn(double a, double b) { return __builtin_fmin(a, b); }
  __device__ static double mn3(double a, double b, double c) { return __builtin_fmin(__builtin_fmin(a, b), c); }
  __device__ static double shr(double old, double v) {
    int2 o = *reinterpret_cast<int2*>(&old), s = *reinterpret_cast<int2*>(&v), r;
    r.x = __builtin_amdgcn_update_dpp(o.x, s.x, 0x138, 0xf, 0xf, false);
    r.y = __builtin_amdgcn_update_dpp(o.y, s.y, 0x138, 0xf, 0xf, false);
    return *reinterpret_cast<double*>(&r);
  }
  __device__ static float tofloat(double v) { return (float)v; }
  __device__ static double fromfloat(float f) { return (double)f; }
};

template <> struct Ar<int32_t> {
  static constexpr bool kFloat = false;
  __device__ static int32_t maxv() { return INT_MAX; }
  __device__ static int32_t rowinf() { return INT_MAX; }
  __device__ static int32_t mn(int32_t a, int32_t b) { return a < b ? a : b; }
  __device__ static int32_t mn3(int32_t a, int32_t b, int32_t c) { return mn(mn(a, b), c); }
  __device__ static int32_t shr(int32_t old, int32_t v) { return __builtin_amdgcn_update_dpp(old, v, 0x138, 0xf, 0xf, false); }
  __device__ static float tofloat(int32_t v) { return (float)v; }
  __device__ static int32_t fromfloat(float f) {
    if (f != f) return 0;
    if (f >= 2147483648.0f) return INT_MAX;
    if (f <= -2147483648.0f) return INT_MIN;
    return (int32_t)f;
  }
};

template <> struct Ar<uint32_t> {
  static constexpr bool kFloat = false;
  __device__ static uint32_t maxv() { return UINT_MAX; }
  __device__ static uint32_t rowinf() { return UINT_MAX; }
  __device__ static uint32_t mn(uint32_t a, uint32_t b) { return a < b ? a : b; }
  __device__ static uint32_t mn3(uint32_t a, uint32_t b, uint32_t c) { return mn(mn(a, b), c); }
  __device__ static uint32_t shr(uint32_t old, uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, 0x138, 0xf, 0xf, false);
  }
  __device__ static float tofloat(uint32_t v) { return (float)v; }
  __device__ static uint32_t fromfloat(float f) {
    if (f != f || f <= 0.0f) return 0;
    if (f >= 4294967296.0f) return UINT_MAX;
    return (uint32_t)f;
  }
};

template <> struct Ar<uint64_t> {
  static constexpr bool kFloat = false;
  __device__ static uint64_t maxv() { return ULLONG_MAX; }
  __device__ static uint64_t rowinf() { return ULLONG_MAX; }
  __device__ static uint64_t mn(uint64_t a, uint64_t b) { return a < b ? a : b; }
  __device__ static uint64_t mn3(uint64_t a, uint64_t b, uint64_t c) { return mn(mn(a, b), c); }
  __device__ static uint64_t shr(uint64_t old, uint64_t v) {
    int2 o = *reinterpret_cast<int2*>(&old), s = *reinterpret_cast<int2*>(&v), r;
    r.x = __builtin_amdgcn_update_dpp(o.x, s.x, 0x138, 0xf, 0xf, false);
    r.y = __builtin_amdgcn_update_dpp(o.y, s.y, 0x138, 0xf, 0xf, false);
    return *reinterpret_cast<uint64_t*>(&r);
  }
  __device__ static float tofloat(uint64_t v) { return (float)v; }
  __device__ static uint64_t fromfloat(float f) {
    if (f != f || f <= 0.0f) return 0;
    if (f >= 18446744073709551616.0f) return ULLONG_MAX;
    return (uint64_t)f;
  }
};

// Wavefront-scope ordering point for LDS traffic between lanes of one wave:
// LDS operations of a wave execute in order, this only stops the compiler
// from moving memory operations across it.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wave-uniform copies (SGPR) of values that are equal in every lane. Loop
// bounds and work indices derived from them are then scalar, so the
// compiler keeps all 64 lanes of the wave in the same loop iteration
// around the cross-lane operations (a work index broadcast with a shuffle
// is not known to be uniform, and a loop exiting on it gets structured
// per lane).
__device__ __forceinline__ int uniform_i32(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ unsigned long long uniform_u64(unsigned long long v) {
  unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v & 0xffffffffull));
  unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v >> 32));
  return ((unsigned long long)hi << 32) | lo;
}

template <typename T> __device__ __forceinline__ T wave_min(T v) {
  for (int d = 32; d > 0; d >>= 1) v = Ar<T>::mn(v, __shfl_xor(v, d, 64));
  return v;
}

// Exclusive prefix minimum across the 64 lanes, seeded by `seed` at lane 0.
template <typename T> __device__ __forceinline__ T wave_excl_prefix_min(T tot, T seed, int lane) {
  T v = tot;
  for (int d = 1; d < 64; d <<= 1) {
    T t = __shfl_up(v, d, 64);
    if (lane >= d) v = Ar<T>::mn(v, t);
  }
  T e = __shfl_up(v, 1, 64);
  return lane == 0 ? seed : Ar<T>::mn(e, seed);
}

// ---------------------------------------------------------------------------
// One cell update. MINFIRST: min3(up,left,diag) + dd, bit-identical to the
// reference's add-then-compare for floats (rounding is monotone) and for
// integers that cannot wrap. Otherwise the add-then-min form with explicit
// missing-candidate flags (integer wrap-around keeps the reference's order).
// PATH: the comparison is on the post-addition values, as the reference's
// step selection is (dtw.hpp:202-227); the 2-bit code is returned.
// ---------------------------------------------------------------------------
struct CellFlags {
  bool up_ok, left_ok, diag_ok;
};

template <typename T, bool MINFIRST, bool PATH>
__device__ __forceinline__ T cell(T up, T left, T diag, T dd, uint32_t& code) {
  if (MINFIRST && !PATH) {
    return Ar<T>::mn3(up, left, diag) + dd;
  } else {
    T U = up + dd, R = left + dd, G = diag + dd;
    T v = Ar<T>::mn3(U, R, G);
    if (PATH) code = (G == v) ? 0u : ((U == v) ? 2u : 1u);
    return v;
  }
}

// Add-then-min with per-candidate presence (integer types that may wrap).
template <typename T, bool PATH>
__device__ __forceinline__ T cell_wrap(T up, T left, T diag, T dd, CellFlags f, uint32_t& code) {
  const T M = Ar<T>::maxv();
  T U = f.up_ok ? (T)(up + dd) : M;
  T R = f.left_ok ? (T)(left + dd) : M;
  T G = f.diag_ok ? (T)(diag + dd) : M;
  // reference selection (dtw.hpp:202-227)
  T v;
  uint32_t c;
  if (G > U) {
    if (U > R) { v = R; c = 1u; } else { v = U; c = 2u; }
  } else {
    if (G > R) { v = R; c = 1u; } else { v = G; c = 0u; }
  }
  if (PATH) code = c;
  return v;
}


// Path cell with its step bits in the two planes of plane_code
// (odba_dtw.hpp): each plane is built by shifting the lane's word left and
// adding the compare's lane bit as the carry, one v_addc per plane and cell
// (packing 2-bit codes took two selects and a shift-or per cell).
template <typename T, bool MINFIRST>
__device__ __forceinline__ T cell_planes(T up, T left, T diag, T dd, uint32_t& wa, uint32_t& wb) {
  const T U = up + dd, R = left + dd, G = diag + dd;
  const T v = Ar<T>::mn3(U, R, G);
  if constexpr (std::is_same<T, float>::value) {
    const unsigned long long na = __builtin_amdgcn_fcmpf(G, v, 14);  // unordered or not equal: not DIAG
    const unsigned long long nb = __builtin_amdgcn_fcmpf(U, v, 1);   // ordered and equal: UP on a tie
    unsigned long long co;
    asm("v_addc_co_u32_e64 %0, %1, %2, %2, %3" : "=v"(wa), "=s"(co) : "v"(wa), "s"(na));
    asm("v_addc_co_u32_e64 %0, %1, %2, %2, %3" : "=v"(wb), "=s"(co) : "v"(wb), "s"(nb));
  } else {
    wa = (wa << 1) | ((G != v) ? 1u : 0u);
    wb = (wb << 1) | ((U == v) ? 1u : 0u);
  }
  return v;
}

// The lane's step word from its planes after C cells (column q at bit C-1-q
// of each plane): columns moved to bits 15 - q and 31 - q, or, for a quarter
// strip's lane k of a group of four (columns kC .. kC+C-1 of the word), to
// bits 15 - kC - q and 31 - kC - q.
template <int C> __device__ __forceinline__ uint32_t plane_word(uint32_t wa, uint32_t wb, int k = 0) {
  const int sh = 16 - (k + 1) * C;
  return (wb << (16 + sh)) | (wa << sh);
}

// ---------------------------------------------------------------------------
// Per-wave LDS staging window.
// ---------------------------------------------------------------------------
// Separate objects so the compiler can move window reads across the
// edge-slot writes of earlier steps.
template <typename T> struct WaveWin {
  T* __restrict__ y;  // 128 entries: y rows [base-64, base+64)
  T* __restrict__ e;  // 64 entries: left edge column rows [base, base+64)
  T* __restrict__ o;  // 128 entries: right edge column slots rows [base-64, base+64)
};

template <typename T> struct SwathIO {
  const T* y;      // first sequence (rows), padded
  int ly;          // number of rows (may be a truncated height)
  const T* x;      // second sequence (columns), padded
  int lx;
  int off;         // swath left column
  const T* ein;    // previous swath right edge column (null for off == 0)
  T* eout;         // this swath's right edge column (null: not needed)
  uint32_t* mv;    // path words for this swath (null: distance only)
  uint32_t* mtop;  // path words of the open-end top row (64 per swath)
  int ref_left;    // 1: `off` is a reference swath start (a multiple of 1024)
  int imask = -1;  // row index mask of ein: -1 (whole column) or ring rows - 1
  int omask = -1;  // row index mask of eout
  int lbase = 0;   // quarter strips: word-lane of this wave's lane 0 in the swath's word layout
  // Checkpointed global paths: `top` (when set) is the cost row above the
  // first row, indexed by column (top[off-1] included when off > 0); with
  // the CKPT template flag, the rows whose global index (row + ck_rbase) is
  // one below a multiple of 2^ck_rshift, and below ck_rlimit, are stored to
  // rowck[(global row >> ck_rshift) * ck_stride + column].
  const T* top = nullptr;
  T* rowck = nullptr;
  long long ck_stride = 0;
  int ck_rbase = 0, ck_rshift = 0, ck_rlimit = 0;
};

// Pipeline hooks of run_swath. A wave that sweeps a whole pair needs none.
// A wave that sweeps one column strip of a pair shared by several waves
// waits until its left neighbour has published the edge rows it is about to
// read, and publishes its own edge rows as they are flushed.
struct NoPipe {
  __device__ bool ready(int) const { return true; }  // input rows below the argument are published
  __device__ void wait(int) const {}
  __device__ void publish(int) const {}
  __device__ void consumed(int) const {}  // input rows below the argument have been read
  __device__ void reserve(int) const {}   // output rows below the argument may be written now
  __device__ bool cancelled() const { return false; }  // the rest of the swath is not needed
};

template <typename T> struct SwathRes {
  T colmin;  // min over the right edge column (valid when eout != null)
  T top;     // right edge value of the top row
  T last;    // D(ly-1, lx-1) when the swath holds column lx-1
};

// Runs one swath of C columns per lane. Rows r0..r1 go through the skewed
// main loop; with open_end the top row goes through the prefix-min scan.
// Lane i receives v of lane i - K within its row of 16 lanes (DPP row_shr).
template <int K> __device__ __forceinline__ uint32_t row_shr_u32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x110 + K, 0xF, 0xF, false);
}

// QW (quarter strips): the wave covers a quarter of a swath, C = C_swath / 4
// columns per lane, so four consecutive lanes share one step word of the
// swath's layout (word-lane a.lbase + lane / 4). Lane 4m+k computes row r
// k steps after lane 4m, so the fourth lane of each group ORs the three
// earlier fragments of row r (kept for 1..3 steps, fetched with DPP) into
// its own and stores the whole word where the one-wave swath would.
template <typename T, int C, bool OS, bool OE, bool MINFIRST, bool PATH, class Pipe = NoPipe, bool CKPT = false,
          bool QW = false>
__device__ SwathRes<T> run_swath(const SwathIO<T>& a, const WaveWin<T>* W, int lane, const Pipe& pipe = Pipe()) {
  const T MAXV = Ar<T>::maxv();
  const T RINF = Ar<T>::rowinf();
  const int ly = a.ly, off = a.off;
  const int r0 = OS ? 1 : 0;
  const int r1 = (OE && ly >= 2) ? ly - 2 : ly - 1;
  const int nrows = r1 - r0 + 1;  // may be 0
  const int steps = nrows > 0 ? nrows + 63 : 0;
  const int nblk = (steps + 63) >> 6;
  const bool first_swath = (off == 0);
  const bool edge_out = (a.eout != nullptr);
  const int col0 = off + lane * C;  // first global column of this lane

  T x[C], D[C];
#pragma unroll
  for (int q = 0; q < C; q++) {
    x[q] = a.x[col0 + q];  // padded: always in bounds
    D[q] = OS ? T(0) : RINF;
  }
  // diagonal input of the lane's first row (with open start, row 0 of every
  // edge column is 0)
  T diag_in;
  if (lane == 0) {
    if (first_swath) diag_in = OS ? MAXV : T(0);
    else diag_in = OS ? T(0) : RINF;
  } else {
    diag_in = OS ? T(0) : RINF;
  }
  if (a.top) {  // a block below a stored cost row (global mode); columns past lx-1 are padding
#pragma unroll
    for (int q = 0; q < C; q++) D[q] = (col0 + q < a.lx) ? a.top[col0 + q] : MAXV;
    diag_in = (col0 > 0 && col0 - 1 < a.lx) ? a.top[col0 - 1] : MAXV;
  }
  const int ck_mask = (1 << a.ck_rshift) - 1;
  auto ckpt = [&](int row, bool on) {  // store a checkpoint row of this lane's columns
    const int grow = row + a.ck_rbase;
    if (on && ((grow + 1) & ck_mask) == 0 && grow < a.ck_rlimit) {
      T* dst = a.rowck + (long long)(grow >> a.ck_rshift) * a.ck_stride + col0;
#pragma unroll
      for (int q = 0; q < C; q++)
        if (col0 + q < a.lx) dst[q] = D[q];  // the row holds columns 0 .. lx-1
    }
  };

  T runmin = MAXV;
  uint32_t h1 = 0, h2 = 0, h3 = 0;  // QW: this lane's fragments 1, 2, 3 steps back
  // QW: `word` is this lane's fragment, already at its bits of the whole
  // word (plane_word with k = lane & 3), so the group's last lane ORs the
  // three earlier fragments of the same row into its own.
  auto put_word = [&](int s, uint32_t word, bool act) {
    if constexpr (QW) {
      const uint32_t f1 = row_shr_u32<1>(h1), f2 = row_shr_u32<2>(h2), f3 = row_shr_u32<3>(h3);
      const uint32_t full = f3 | f2 | f1 | word;
      h3 = h2;
      h2 = h1;
      h1 = word;
      const int l = a.lbase + (lane >> 2);
      if ((lane & 3) == 3 && act) a.mv[(size_t)(s - lane + l) * 64 + l] = full;
    } else {
      if (act) a.mv[(size_t)s * 64 + lane] = word;
    }
  };

  // prefetch block 0 inputs
  auto load_y = [&](int b) -> T { return a.y[r0 + 64 * b + lane]; };
  auto load_e = [&](int b) -> T {
    if (first_swath) return MAXV;
    int r = r0 + 64 * b + lane;
    return a.ein[r & a.imask];  // edge buffers are padded by 128 entries (rings wrap)
  };
  auto rows_of_block = [&](int b) -> int { return min(r0 + 64 * b + 64, r1 + 1); };
  // The next block's edge values are loaded one block early whenever the
  // left neighbour has already published them (always, without a pipe);
  // otherwise the wave waits for them when it stages that block.
  T ynext = nblk > 0 ? load_y(0) : T(0);
  T enext = T(0);
  bool e_ready = false;
  if (nblk > 0 && (first_swath || pipe.ready(rows_of_block(0)))) {
    enext = load_e(0);
    e_ready = true;
  }

  SwathRes<T> res;
  res.colmin = MAXV; res.top = MAXV; res.last = T(0);
  for (int b = 0; b < nblk; b++) {
    if (pipe.cancelled()) {  // no result of this swath is needed any more
      __builtin_amdgcn_s_waitcnt(0);
      return res;
    }
    // --- stage the window for block b
    if (!e_ready) {
      pipe.wait(rows_of_block(b));
      enext = load_e(b);
    }
    T yshift = W->y[64 + lane];
    wave_sync();
    W->y[lane] = yshift;
    W->y[64 + lane] = ynext;
    W->e[lane] = enext;
    wave_sync();
    if (!first_swath) pipe.consumed(rows_of_block(b));
    e_ready = false;
    if (b + 1 < nblk) {
      ynext = load_y(b + 1);
      if (first_swath || pipe.ready(rows_of_block(b + 1))) {
        enext = load_e(b + 1);
        e_ready = true;
      }
    }

    const int sbase = 64 * b;
    const bool steady = (sbase >= 63) && (sbase + 63 < nrows);
    if (steady) {
#pragma unroll 8
      for (int u = 0; u < 64; u++) {
        const int s = sbase + u;
        T yv = W->y[u - lane + 64];
        T ev = W->e[u];
        T left_in = Ar<T>::shr(ev, D[C - 1]);
        T dg = diag_in;
        diag_in = left_in;
        T l = left_in;
        uint32_t wa = 0, wb = 0;
#pragma unroll
        for (int q = 0; q < C; q++) {
          T up = D[q];
          T df = yv - x[q];
          T dd = df * df;
          uint32_t code = 0;
          T v;
          if (MINFIRST || Ar<T>::kFloat) {
            if (PATH) v = cell_planes<T, MINFIRST>(up, l, dg, dd, wa, wb);
            else v = cell<T, MINFIRST, PATH>(up, l, dg, dd, code);
          } else {
            const int row = r0 + s - lane;
            CellFlags f{row > 0, (q > 0) || (lane > 0) || !first_swath || row == 0, (row > 0) && ((q > 0) || (lane > 0) || !first_swath)};
            T lv = (q == 0 && lane == 0 && first_swath) ? T(0) : l;
            v = cell_wrap<T, PATH>(up, lv, dg, dd, f, code);
            if (PATH) {
              wa = (wa << 1) | (code != 0u ? 1u : 0u);
              wb = (wb << 1) | (code == 2u ? 1u : 0u);
            }
          }
          dg = up;
          D[q] = v;
          l = v;
        }
        if (PATH) put_word(s, plane_word<C>(wa, wb, QW ? (lane & 3) : 0), true);
        if (CKPT) ckpt(r0 + s - lane, true);
        W->o[u - lane + 64] = D[C - 1];
      }
    } else {
      // Skew prologue / epilogue: lanes outside their rows keep their state
      // through selects, so the block stays straight-line and unrolled.
#pragma unroll 8
      for (int u = 0; u < 64; u++) {
        const int s = sbase + u;
        T yv = W->y[u - lane + 64];
        T ev = W->e[u];
        T left_in = Ar<T>::shr(ev, D[C - 1]);
        const int row = r0 + s - lane;
        const bool active = (unsigned)(s - lane) < (unsigned)nrows;
        T dg = diag_in;
        diag_in = active ? left_in : diag_in;
        T l = left_in;
        uint32_t wa = 0, wb = 0;
#pragma unroll
        for (int q = 0; q < C; q++) {
          T up = D[q];
          T df = yv - x[q];
          T dd = df * df;
          uint32_t code = 0;
          T v;
          if (MINFIRST || Ar<T>::kFloat) {
            if (PATH) v = cell_planes<T, MINFIRST>(up, l, dg, dd, wa, wb);
            else v = cell<T, MINFIRST, PATH>(up, l, dg, dd, code);
          } else {
            CellFlags f{row > 0, (q > 0) || (lane > 0) || !first_swath || row == 0, (row > 0) && ((q > 0) || (lane > 0) || !first_swath)};
            T lv = (q == 0 && lane == 0 && first_swath) ? T(0) : l;
            v = cell_wrap<T, PATH>(up, lv, dg, dd, f, code);
            if (PATH) {
              wa = (wa << 1) | (code != 0u ? 1u : 0u);
              wb = (wb << 1) | (code == 2u ? 1u : 0u);
            }
          }
          dg = up;
          D[q] = active ? v : up;
          l = v;
        }
        if (PATH) put_word(s, plane_word<C>(wa, wb, QW ? (lane & 3) : 0), active);
        if (CKPT) ckpt(row, active);
        W->o[u - lane + 64] = D[C - 1];
      }
    }
    wave_sync();
    // --- flush final right-edge values of rows [r0+sbase-64, r0+sbase)
    if (edge_out) {
      const int row = r0 + sbase - 64 + lane;
      T v = W->o[lane];
      pipe.reserve(min(r0 + sbase, r1 + 1));
      if (row >= r0 && row <= r1) {
        a.eout[row & a.omask] = v;
        runmin = Ar<T>::mn(runmin, v);
      }
      pipe.publish(min(r0 + sbase, r1 + 1));
      T sh = W->o[64 + lane];
      wave_sync();
      W->o[lane] = sh;
      wave_sync();
    }
  }
  if (edge_out && nblk > 0) {
    const int row = r0 + 64 * nblk - 64 + lane;
    T v = W->o[lane];
    pipe.reserve(r1 + 1);
    if (row >= r0 && row <= r1) {
      a.eout[row & a.omask] = v;
      runmin = Ar<T>::mn(runmin, v);
    }
  }
  if (edge_out) pipe.publish(r1 + 1);
  if (OS && edge_out && lane == 0 && a.omask == -1) a.eout[0] = T(0);  // never read back from a ring
  if (OS) runmin = Ar<T>::mn(runmin, T(0));

  // --- open-end top row: D(t,j) = min(min(U,G), D(t,j-1)) with U,G carrying
  // the local cost, except at the swath-left column (dtw.hpp:175-186).
  if (pipe.cancelled()) {
    __builtin_amdgcn_s_waitcnt(0);
    return res;
  }
  if (OE && ly >= 2) {
    const int t = ly - 1;
    const T yv = a.y[t];
    if (!first_swath) pipe.wait(ly);
    const T lane0_diag = first_swath ? MAXV : a.ein[(t - 1) & a.imask];
    const T left0 = first_swath ? MAXV : a.ein[t & a.imask];
    const T nb = Ar<T>::shr(lane0_diag, D[C - 1]);
    // pass 1: m[q] = running min over the lane of min(U, G); U = up + dd,
    // G = diag + dd except at a reference swath-left column where the
    // reference adds no local cost to the diagonal (and at column 0 the
    // diagonal is MAX).
    const bool bare0 = (lane == 0) && (a.ref_left != 0);
    T m[C];
    T run = T(0);
#pragma unroll
    for (int q = 0; q < C; q++) {
      const T df = yv - x[q];
      const T dd = df * df;
      const T U = D[q] + dd;
      const T G = (q == 0 && bare0) ? nb : ((q == 0 ? nb : D[q - 1]) + dd);
      const T mm = (G > U) ? U : G;
      run = (q == 0) ? mm : Ar<T>::mn(run, mm);
      m[q] = run;
    }
    const T pre = wave_excl_prefix_min<T>(m[C - 1], left0, lane);
    // pass 2: values and step codes (reference tie order: DIAG, UP, RIGHT)
    uint32_t word = 0;
    T dprev_old = nb;
#pragma unroll
    for (int q = 0; q < C; q++) {
      const T df = yv - x[q];
      const T dd = df * df;
      const T U = D[q] + dd;
      const T G = (q == 0 && bare0) ? nb : (dprev_old + dd);
      const T v = Ar<T>::mn(pre, m[q]);
      if (PATH) word |= ((G == v) ? 0u : ((U == v) ? 2u : 1u)) << (2 * q);
      dprev_old = D[q];
      D[q] = v;
    }
    if (PATH) {
      if constexpr (QW) {
        const uint32_t full = row_shr_u32<3>(word) | (row_shr_u32<2>(word) << (2 * C)) |
                              (row_shr_u32<1>(word) << (4 * C)) | (word << (6 * C));
        if ((lane & 3) == 3) a.mtop[a.lbase + (lane >> 2)] = full;
      } else {
        a.mtop[lane] = word;
      }
    }
    if (edge_out) {
      T top = __shfl(D[C - 1], 63, 64);
      pipe.reserve(ly);
      if (lane == 0) a.eout[t & a.omask] = top;
      runmin = Ar<T>::mn(runmin, top);
      pipe.publish(ly);
    }
  }
  if (edge_out) {
    res.colmin = wave_min<T>(runmin);
    res.top = __shfl(D[C - 1], 63, 64);  // lane 63 holds the last row at the right edge
  }
  // value at the pair's last column (only meaningful for the swath holding
  // it): the owning lane parks its C values in the LDS slot window, every
  // lane reads the wanted one (static register indices only).
  {
    const int tl = a.lx - 1 - off;
    res.last = T(0);
    if (tl >= 0 && tl < 64 * C) {
      const int ll = tl / C, qq = tl - (tl / C) * C;
      wave_sync();
      if (lane == ll) {
#pragma unroll
        for (int k = 0; k < C; k++) W->o[k] = D[k];
      }
      wave_sync();
      res.last = W->o[qq];
      wave_sync();
    }
  }
  if (edge_out || a.mv) __builtin_amdgcn_s_waitcnt(0);
  return res;
}

// ---------------------------------------------------------------------------
// Whole pair: swaths left to right with the abrogation check between them.
// ---------------------------------------------------------------------------
// Correctly rounded float sqrt and division (the IEEE results nvcc's
// sqrtf and / give by default): evaluated in double and rounded once to
// float, which is exact for these operations because 53 >= 2*24 + 2. The
// single-precision hardware sqrt is only accurate to 1 ulp.
__device__ __forceinline__ float sqrt_rn_f32(float x) { return (float)__builtin_sqrt((double)x); }
__device__ __forceinline__ float div_rn_f32(float a, float b) { return (float)((double)a / (double)b); }

template <typename T> __device__ __forceinline__ T pair_distance(T cost, int ly, bool normalise) {
  float f = sqrt_rn_f32(Ar<T>::tofloat(cost));
  if (normalise) f = div_rn_f32(f, (float)ly);
  return Ar<T>::fromfloat(f);
}

template <typename T, bool OS, bool OE, bool MINFIRST, bool PATH, int C>
__device__ SwathRes<T> swath_c(const SwathIO<T>& io, const WaveWin<T>* W, int lane) {
  return run_swath<T, C, OS, OE, MINFIRST, PATH>(io, W, lane);
}

template <typename T, bool OS, bool OE, bool MINFIRST, bool PATH>
__device__ SwathRes<T> swath_dispatch(int C, const SwathIO<T>& io, const WaveWin<T>* W, int lane) {
  switch (C) {
    case 1: return swath_c<T, OS, OE, MINFIRST, PATH, 1>(io, W, lane);
    case 2: return swath_c<T, OS, OE, MINFIRST, PATH, 2>(io, W, lane);
    case 4: return swath_c<T, OS, OE, MINFIRST, PATH, 4>(io, W, lane);
    case 8: return swath_c<T, OS, OE, MINFIRST, PATH, 8>(io, W, lane);
    case 12: return swath_c<T, OS, OE, MINFIRST, PATH, 12>(io, W, lane);
    default: return swath_c<T, OS, OE, MINFIRST, PATH, 16>(io, W, lane);
  }
}

template <typename T, bool OS, bool OE, bool PATH, class Pipe>
__device__ SwathRes<T> swath_dispatch_pipe(int C, const SwathIO<T>& io, const WaveWin<T>* W, int lane, const Pipe& pp) {
  switch (C) {
    case 1: return run_swath<T, 1, OS, OE, true, PATH, Pipe>(io, W, lane, pp);
    case 2: return run_swath<T, 2, OS, OE, true, PATH, Pipe>(io, W, lane, pp);
    case 4: return run_swath<T, 4, OS, OE, true, PATH, Pipe>(io, W, lane, pp);
    case 8: return run_swath<T, 8, OS, OE, true, PATH, Pipe>(io, W, lane, pp);
    case 12: return run_swath<T, 12, OS, OE, true, PATH, Pipe>(io, W, lane, pp);
    default: return run_swath<T, 16, OS, OE, true, PATH, Pipe>(io, W, lane, pp);
  }
}

// Quarter strips of a path swath: C = 1, 2, 3, 4 columns per lane (a
// quarter of a swath of 4, 8, 12, 16).
template <typename T, bool OS, bool OE, class Pipe>
__device__ SwathRes<T> swath_dispatch_quarter(int C, const SwathIO<T>& io, const WaveWin<T>* W, int lane,
                                              const Pipe& pp) {
  switch (C) {
    case 1: return run_swath<T, 1, OS, OE, true, true, Pipe, false, true>(io, W, lane, pp);
    case 2: return run_swath<T, 2, OS, OE, true, true, Pipe, false, true>(io, W, lane, pp);
    case 3: return run_swath<T, 3, OS, OE, true, true, Pipe, false, true>(io, W, lane, pp);
    default: return run_swath<T, 4, OS, OE, true, true, Pipe, false, true>(io, W, lane, pp);
  }
}

__host__ __device__ inline int last_swath_c(int wcols) {
  if (wcols <= 64) return 1;
  if (wcols <= 128) return 2;
  if (wcols <= 256) return 4;
  if (wcols <= 512) return 8;
  if (wcols <= 768) return 12;
  return 16;
}

struct PairOut {
  int abrogated;  // swath index where the open-end search stopped, -1 if none
  int q3;         // 1 when the reference would read an unwritten LDS slot (ly == 1)
};

template <typename T, bool OS, bool OE, bool MINFIRST, bool PATH>
__device__ T dtw_one_pair(const T* y, int ly, const T* x, int lx, T* ebuf0, T* ebuf1, uint32_t* mvbase,
                          uint32_t* mtop, const WaveWin<T>* W, int lane, PairOut& po) {
  const int K = (lx + ODBA_W - 1) / ODBA_W;
  const int nrows_main = ((OE && ly >= 2) ? ly - 2 : ly - 1) - (OS ? 1 : 0) + 1;
  const size_t swath_words = (size_t)(nrows_main > 0 ? nrows_main + 63 : 0) * 64;
  po.abrogated = -1;
  po.q3 = 0;
  T colmin = T(0), top = T(0);
  for (int k = 0; k < K; k++) {
    const int off = k * ODBA_W;
    if (k > 0 && OE && !OS && off > ly) {
      if (top == colmin) {  // dtw.hpp:106-127
        po.abrogated = k;
        return pair_distance<T>(ly == 1 ? Ar<T>::maxv() : top, ly, true);
      }
    }
    const bool last = (k == K - 1);
    const int wcols = last ? lx - off : ODBA_W;
    SwathIO<T> io;
    io.y = y; io.ly = ly; io.x = x; io.lx = lx; io.off = off;
    io.ein = k > 0 ? (((k - 1) & 1) ? ebuf1 : ebuf0) : nullptr;
    io.eout = last ? nullptr : ((k & 1) ? ebuf1 : ebuf0);
    io.mv = PATH ? mvbase + (size_t)k * swath_words : nullptr;
    io.mtop = PATH ? mtop + (size_t)k * 64 : nullptr;
    io.ref_left = 1;
    const int C = last ? last_swath_c(wcols) : 16;
    SwathRes<T> r = swath_dispatch<T, OS, OE, MINFIRST, PATH>(C, io, W, lane);
    if (!last) { colmin = r.colmin; top = r.top; }
    if (last) {
      if (ly == 1 && ((wcols - 1) % 3) != 0) po.q3 = 1;
      T cost = r.last;
      if (OS && ly == 1) cost = T(0);
      return pair_distance<T>(cost, ly, OE != OS);
    }
  }
  return T(0);
}

// ---------------------------------------------------------------------------
// All-vs-all kernel. Pair order: second sequence index c descending, then
// first index r descending (longest X first). Pair (r,c) has Y = seq r.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void decode_pair(long long p, long long npairs, int& r, int& c) {
  long long q = npairs - 1 - p;  // reverse index: q = c(c-1)/2 + r
  long long cc = (long long)((1.0 + sqrt(1.0 + 8.0 * (double)q)) * 0.5);
  while (cc * (cc - 1) / 2 > q) cc--;
  while ((cc + 1) * cc / 2 <= q) cc++;
  c = (int)cc;
  r = (int)(q - cc * (cc - 1) / 2);
}

// Float kernels are held to 128 VGPRs (4 waves per SIMD): the cell update's
// VALU mix only reaches its full issue rate with two or more waves resident
// per SIMD, and at 129+ VGPRs only 3 fit. 8-byte types would spill under the
// same limit and keep the compiler's choice.
template <typename T> struct WavesPerEU { static constexpr int value = 1; };
template <> struct WavesPerEU<float> { static constexpr int value = 4; };

template <typename T, bool OS, bool OE, bool MINFIRST>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WavesPerEU<T>::value))) k_all_pairs(const T* __restrict__ data, const long long* __restrict__ offs,
                                                   const int* __restrict__ lens, int n, long long npairs,
                                                   unsigned long long* __restrict__ counter, T* __restrict__ dist,
                                                   T* __restrict__ escratch, long long escratch_stride,
                                                   int* __restrict__ flags, const int2* __restrict__ order) {
  __shared__ T s_y[4][128];
  __shared__ T s_e[4][64];
  __shared__ T s_o[4][128];
  const int lane = threadIdx.x & 63;
  const int wid = uniform_i32((int)(threadIdx.x >> 6));
  const long long gw = (long long)blockIdx.x * 4 + wid;
  T* e0 = escratch + gw * 2 * escratch_stride;
  T* e1 = e0 + escratch_stride;
  const WaveWin<T> win{s_y[wid], s_e[wid], s_o[wid]};
  const WaveWin<T>* W = &win;
  for (;;) {
    // Every lane takes part in the fetch (only lane 0 adds 1): no divergent
    // branch around the atomic, so the loop stays one wave-uniform loop.
    const unsigned long long p = uniform_u64(atomicAdd(counter, lane == 0 ? 1ull : 0ull));
    if ((long long)p >= npairs) break;
    int r, c;
    if (order) {  // an explicit packing order (r < c)
      r = uniform_i32(order[p].x);
      c = uniform_i32(order[p].y);
    } else {
      decode_pair((long long)p, npairs, r, c);
    }
    PairOut po;
    T d = dtw_one_pair<T, OS, OE, MINFIRST, false>(data + offs[r], lens[r], data + offs[c], lens[c], e0, e1, nullptr,
                                                    nullptr, W, lane, po);
    if (lane == 0) {
      long long idx = (long long)r * (2LL * n - r - 1) / 2 + (c - r - 1);
      dist[idx] = d;
      if (flags) flags[idx] = (po.abrogated >= 0 ? (1 | (po.abrogated << 8)) : 0) | (po.q3 ? 2 : 0);
    }
  }
}

// ---------------------------------------------------------------------------
// Pipeline hooks for a workgroup of waves sweeping one pair in column
// strips (k_path_strips): strip s reads the right-edge column of strip s-1
// from a global edge slot; the slot's owner and the number of rows published
// so far are LDS words, so the hand-off costs one LDS poll per 64-row block.
// Slots are reused round-robin (2*NW of them); a producer claims a slot only
// after the previous owner's consumer released it. Waiting only ever goes to
// a lower strip index, which is already running or done, so the pipeline
// cannot deadlock.
//
// The open-end early stop (dtw.hpp:74-129) is not taken by strip schedules:
// every strip is swept, and the stop is evaluated afterwards at each
// reference swath start. This gives the same distance: when the column
// minimum at a swath boundary equals its top value, every later cell is at
// least that minimum (costs only grow by non-negative terms; min-first types
// only), while the free right moves of the open-end top row carry exactly
// that value to the last column. The flags report the swath the reference
// stops at.
// ---------------------------------------------------------------------------
struct StripPipe {
  int* in_owner;  // LDS: strip whose edge is in the input slot
  int* in_cnt;    // LDS: edge rows published into the input slot
  int in_strip;   // the left neighbour's strip index
  int* out_cnt;   // LDS: rows published of this strip's edge (null: none)
  int lane;
  // Open-end early stop: LDS word holding the lowest swath decided so far
  // at which the sweep stops (INT_MAX: none yet), and this strip's swath.
  // A strip at or past a decided stop is cancelled: nothing reads its cells.
  const int* stop = nullptr;
  int swath = 0;
  __device__ bool cancelled() const {
    return stop && uniform_i32(__hip_atomic_load(stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) <= swath;
  }
  __device__ bool ready(int rows) const {
    const int o = uniform_i32(__hip_atomic_load(in_owner, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
    const int c = uniform_i32(__hip_atomic_load(in_cnt, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
    if (o != in_strip || c < rows) return false;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    return true;
  }
  __device__ void wait(int rows) const {
    for (;;) {
      const int o = uniform_i32(__hip_atomic_load(in_owner, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
      const int c = uniform_i32(__hip_atomic_load(in_cnt, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
      if (o == in_strip && c >= rows) break;
      if (cancelled()) return;  // the producer (at or past the stop too) may never publish
      __builtin_amdgcn_s_sleep(2);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  }
  __device__ void publish(int rows) const {
    if (!out_cnt) return;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if (lane == 0) __hip_atomic_store(out_cnt, rows, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __device__ void consumed(int) const {}
  __device__ void reserve(int) const {}
};

constexpr int kMaxStripWaves = 16;

// ---------------------------------------------------------------------------
// All-vs-all for a few very long pairs: bands of strips across workgroups.
// A pair's columns are cut into strips 64*C wide and the strips into bands
// of NW consecutive strips. Workgroups (NW waves) claim bands from a global
// counter in pair order, band order; wave w of a band sweeps its strip w.
//
//   * Inside a band, the edge column between strips w and w+1 goes through
//     a ring of `ring` rows in the workgroup's scratch: the producer waits
//     before overwriting rows its consumer has not read (LDS counters).
//   * Between bands, the band's last strip writes the whole edge column to
//     a global slot taken from a pool of free slots (a bitmap), publishes
//     the slot in a per-claim word and its progress in a per-slot word, 256
//     rows at a time; the next band's first strip (the next claim) reads it
//     and returns the slot to the pool when done. A slot is held from its
//     producer's start until its consumer ends, so at most twice the number
//     of running claims plus one are held at once: a pool of
//     2 * workgroups + 2 slots is never empty.
//
// No deadlock: every wait targets a lower claim or a lower strip of the
// same claim, except the ring wait, which targets the next strip of the
// same band; the band's last strip never waits on a ring (its output slot
// is a whole column), and a claim is only taken by a running workgroup.
// Every wait also ends after `tick_limit` wall-clock ticks, setting bit 8
// of *err, after which every other wait returns at once: a broken hand-off
// cannot hang the device.
//
// Distances are written by the pair's last strip. The open-end early stop
// is reconstructed after the fact as for every strip schedule: the per-pair
// stop swath is an atomic minimum in `abr`, turned into flags by
// k_band_flags once every band has finished.
// ---------------------------------------------------------------------------
struct BandPair {
  int r, c;         // Y = sequence r, X = sequence c
  long long band0;  // claim index of the pair's first band
};

constexpr int kGlobalGrain = 256;  // rows per global publish

__device__ __forceinline__ bool band_timed_out(long long t0, long long limit, int* err, int lane) {
  if (wall_clock64() - t0 <= limit) return false;
  if (lane == 0) atomicOr(err, 8);
  return true;
}

// After one wait has timed out every other wait returns at once, so the
// launch drains quickly (with garbage results, reported through *err).
__device__ __forceinline__ bool band_failed(int* err) {
  return (uniform_i32(__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) & 8) != 0;
}

struct BandPipe {
  const int* in_cnt;    // rows published into the input (LDS, or a global word)
  int in_global;
  int* in_used;         // ring input: LDS word of rows read so far
  int* out_cnt;         // rows published of the output (null: none)
  int out_global;
  const int* out_used;  // ring output: the consumer's rows read
  int ring;
  int main_end;         // rows of the main loop (r1 + 1): always published
  int lane;
  int* err;
  long long limit;
  mutable int last_pub = 0;
  // Open-end stop of path jobs (k_path_swath_bands): the job's global stop
  // word (lowest swath decided so far at which the sweep stops; INT_MAX:
  // none) and this strip's swath. Null: no cancellation.
  const int* stop = nullptr;
  int swath = 0;
  __device__ int stop_swath() const {
    return uniform_i32(__hip_atomic_load(stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  }

  __device__ bool ready(int rows) const {
    if (in_global) {
      const int c = uniform_i32(__hip_atomic_load(in_cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      if (c < rows) return false;
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    } else {
      const int c = uniform_i32(__hip_atomic_load(in_cnt, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
      if (c < rows) return false;
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
    return true;
  }
  __device__ void wait(int rows) const {
    if (band_failed(err)) return;
    const long long t0 = wall_clock64();
    while (!ready(rows)) {
      if (cancelled()) return;  // the producer is at or past the stop too
      if (band_timed_out(t0, limit, err, lane)) return;
      if (in_global) __builtin_amdgcn_s_sleep(2);
      else __builtin_amdgcn_s_sleep(1);
    }
  }
  __device__ void publish(int rows) const {
    if (!out_cnt) return;
    if (out_global) {
      if (rows < main_end && rows - last_pub < kGlobalGrain) return;
      last_pub = rows;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      if (lane == 0) __hip_atomic_store(out_cnt, rows, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      if (lane == 0) __hip_atomic_store(out_cnt, rows, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
  }
  __device__ void consumed(int rows) const {
    if (!in_used) return;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if (lane == 0) __hip_atomic_store(in_used, rows, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __device__ bool cancelled() const { return stop && stop_swath() <= swath; }
  __device__ void reserve(int rows) const {
    if (!out_used) return;
    const int need = rows - ring;
    if (need <= 0) return;
    if (band_failed(err)) return;
    const long long t0 = wall_clock64();
    for (;;) {
      const int u = uniform_i32(__hip_atomic_load(out_used, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
      if (u >= need) break;
      if (stop && stop_swath() <= swath + 1) return;  // the consumer is cancelled: nothing reads the ring
      if (band_timed_out(t0, limit, err, lane)) return;
      __builtin_amdgcn_s_sleep(1);
    }
  }
};

// Takes a free global slot (lane 0 works the bitmap; the pool is sized so
// that one is always free).
__device__ int band_pop_slot(unsigned long long* bits, int nwords, int* err, long long limit, int lane) {
  int got = -1;
  if (lane == 0) {
    const long long t0 = wall_clock64();
    while (got < 0) {
      for (int w = 0; w < nwords && got < 0; w++) {
        unsigned long long v = __hip_atomic_load(&bits[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while (v && got < 0) {
          const int b = __builtin_ctzll(v);
          const unsigned long long m = 1ull << b;
          const unsigned long long old = atomicAnd(&bits[w], ~m);
          if (old & m) got = w * 64 + b;
          else v = old & ~m;
        }
      }
      if (got < 0 && wall_clock64() - t0 > limit) {
        atomicOr(err, 8);
        got = 0;
      }
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // after the last reader's release
  return uniform_i32(got);
}

template <typename T, int C, bool OS, bool OE>
__global__ void __launch_bounds__(1024) k_all_pairs_bands(const T* __restrict__ data, const long long* __restrict__ offs,
                                                          const int* __restrict__ lens, int n,
                                                          const BandPair* __restrict__ tab, int ntab, long long nclaims,
                                                          unsigned long long* __restrict__ counter, T* __restrict__ dist,
                                                          T* __restrict__ rings, int ring, T* __restrict__ gslots,
                                                          long long gstride, int* __restrict__ g_cnt,
                                                          unsigned long long* __restrict__ g_free, int R,
                                                          int* __restrict__ slot_of, int* __restrict__ abr,
                                                          int* __restrict__ err, long long limit) {
  __shared__ T s_y[kMaxStripWaves][128];
  __shared__ T s_e[kMaxStripWaves][64];
  __shared__ T s_o[kMaxStripWaves][128];
  __shared__ int s_cnt[kMaxStripWaves];   // rows published into ring w
  __shared__ int s_used[kMaxStripWaves];  // rows of ring w read by wave w+1
  __shared__ unsigned long long s_q;
  constexpr int SW = 64 * C;
  const int lane = threadIdx.x & 63;
  const int wid = uniform_i32((int)(threadIdx.x >> 6));
  const int nw = uniform_i32((int)(blockDim.x >> 6));
  T* rbase = rings + (long long)blockIdx.x * nw * ring;
  const WaveWin<T> win{s_y[wid], s_e[wid], s_o[wid]};
  for (;;) {
    __syncthreads();  // every wave is done with the previous claim
    if (threadIdx.x == 0) s_q = atomicAdd(counter, 1ull);
    if (threadIdx.x < kMaxStripWaves) {
      s_cnt[threadIdx.x] = 0;
      s_used[threadIdx.x] = 0;
    }
    __syncthreads();
    const long long q = (long long)uniform_u64(s_q);
    if (q >= nclaims) break;
    int lo = 0, hi = ntab - 1;  // the pair: last table entry with band0 <= q
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tab[mid].band0 <= q) lo = mid;
      else hi = mid - 1;
    }
    const int t = uniform_i32(lo);
    const int r = uniform_i32(tab[t].r), c = uniform_i32(tab[t].c);
    const int b = uniform_i32((int)(q - tab[t].band0));
    const int ly = lens[r], lx = lens[c];
    const int S = (lx + SW - 1) / SW;
    const int s0 = b * nw;
    const int slast = min(s0 + nw, S) - 1;
    const int s = s0 + wid;
    if (s > slast) continue;  // no strip for this wave in the pair's last band
    const int off = s * SW;
    const bool has_out = s < S - 1;
    const bool in_global = (wid == 0) && s > 0;
    const bool out_global = (s == slast) && has_out;
    int gout = 0, gin = 0;
    if (out_global) {  // take a slot, reset its progress word, then announce it
      gout = band_pop_slot(g_free, (R + 63) / 64, err, limit, lane);
      if (lane == 0) {
        __hip_atomic_store(&g_cnt[gout], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&slot_of[q], gout, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    if (in_global && !band_failed(err)) {  // the left band's slot
      const long long t0 = wall_clock64();
      for (;;) {
        gin = uniform_i32(__hip_atomic_load(&slot_of[q - 1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT));
        if (gin >= 0) break;
        if (band_timed_out(t0, limit, err, lane)) { gin = 0; break; }
        __builtin_amdgcn_s_sleep(2);
      }
    }
    BandPipe pp;
    pp.in_global = in_global;
    pp.in_cnt = in_global ? &g_cnt[gin] : (s > 0 ? &s_cnt[wid - 1] : nullptr);
    pp.in_used = (s > 0 && !in_global) ? &s_used[wid - 1] : nullptr;
    pp.out_cnt = has_out ? (out_global ? &g_cnt[gout] : &s_cnt[wid]) : nullptr;
    pp.out_global = out_global;
    pp.out_used = (has_out && !out_global) ? &s_used[wid] : nullptr;
    pp.ring = ring;
    pp.main_end = (OE && ly >= 2) ? ly - 1 : ly;
    pp.lane = lane;
    pp.err = err;
    pp.limit = limit;
    SwathIO<T> io;
    io.y = data + offs[r]; io.ly = ly; io.x = data + offs[c]; io.lx = lx; io.off = off;
    io.ein = s == 0 ? nullptr : (in_global ? gslots + (long long)gin * gstride : rbase + (long long)(wid - 1) * ring);
    io.imask = in_global ? -1 : ring - 1;
    io.eout = has_out ? (out_global ? gslots + (long long)gout * gstride : rbase + (long long)wid * ring) : nullptr;
    io.omask = out_global ? -1 : ring - 1;
    io.mv = nullptr; io.mtop = nullptr;
    io.ref_left = (off % ODBA_W) == 0;
    SwathRes<T> res = run_swath<T, C, OS, OE, true, false, BandPipe>(io, &win, lane, pp);
    if (in_global && lane == 0) {  // every read of the slot has returned: back to the pool
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      atomicOr(&g_free[gin >> 6], 1ull << (gin & 63));
    }
    if (OE && !OS && has_out) {
      const int next = off + SW;
      if ((next % ODBA_W) == 0 && next > ly && res.top == res.colmin && lane == 0) atomicMin(&abr[t], next / ODBA_W);
    }
    if (s == S - 1 && lane == 0) {
      const long long idx = (long long)r * (2LL * n - r - 1) / 2 + (c - r - 1);
      dist[idx] = pair_distance<T>(res.last, ly, OE != OS);
    }
  }
}

// ---------------------------------------------------------------------------
// Small pairs (float, every length 64..1024, so each pair is one swath and
// none of the swath-width rules of the open end applies): a workgroup per
// pair, wave k sweeping the pair's strip of 64*C columns [64*C*k, 64*C*(k+1)),
// all waves in lock step: B-step sub-blocks separated by workgroup
// barriers, strip k running 64 + B steps behind strip k-1. Its right-edge
// value of row r (lane 63, 63 steps after lane 0 took the row) lands in an
// LDS column of the strip B + 1 steps before strip k+1's lane 0 needs it, in
// an earlier sub-block, so strips hand over every sub-block instead of
// every 64 or 256 rows. Y is staged in LDS once per pair, between guard
// rows, and each strip's edge column has a row of LDS per row of the pair:
// no index wraps, so every LDS access of a sub-block is one base register
// plus an immediate offset. The recurrence and its initial values are
// run_swath's for a first swath, with C columns per lane; a smaller C
// shortens each step's instruction stream (4C + 2) at the price of more
// strips. With open end the top row (a prefix minimum along the row) runs
// strip by strip after the main rows. One strip per wave needs
// C*64*waves >= |X|.
// ---------------------------------------------------------------------------
constexpr int kSmallMaxLen = 1024;  // longest sequence (one swath)
constexpr int kSmallMaxWaves = 8;
constexpr int kSmallYPad = 64;  // guard rows before Y: lane l reads row s - l
// Y with its guards: steps run to nrows + 63 + B - 1 (rounded up to a
// sub-block), lane 0 reading row r0 + step.
constexpr int kSmallYLen = kSmallYPad + kSmallMaxLen + 64 + 32;
// Edge column of a strip: row r at index r + kSmallEdgePad, so lane 63's
// rows of a sub-block (s0 - 63 .. s0 - 63 + B - 1, s0 a multiple of 16) start
// 16-byte aligned and every row it sweeps, before the first (>= -63) or
// past the last, has a slot; lane 0 of the next strip reads rows up to its
// last step.
constexpr int kSmallEdgePad = 67;
constexpr int kSmallEdgeLen = kSmallEdgePad + kSmallMaxLen + 64 + 32 + 1;

// One B-step sub-block of one strip. Every LDS read of the sub-block (its
// rows of Y, and the left neighbour's edge rows, written in earlier
// sub-blocks) is issued first, so their latency is paid once per sub-block;
// the steps are then register-only, and the right-edge values of lane 63 go
// to the strip's edge column afterwards, all B of them: a row before the
// first lands in a guard slot nobody reads, a row past the last is read only
// by a lane that is past its last row too. Strip 0 reads the last wave's
// column, which no strip writes and holds MAXV (no column -1).
//
// Lane l takes row s - l at step s. Before its first row (s < l) a lane
// must leave its state alone, and after its last row too. pro_free says the
// first need not be enforced: without open start the row above the first
// row is +inf in every column but the corner, and Y's guard rows before the
// first row hold +inf, so with X finite (and not +inf) a lane's steps before
// its first row compute min3(...) + inf = +inf, exactly the values of the row
// above that it holds, and hand +inf on as left and diagonal. Then only a
// sub-block that reaches past the last row masks, and only the state update.
template <int C, int B, bool OS, bool OE>
__device__ __forceinline__ void small_steps(int lb, int nrows, int r0, bool has_out, bool pro_free, int lane,
                                            const float* __restrict__ sy, const float* __restrict__ edge_in,
                                            float* __restrict__ edge_out, const float (&x)[C], float (&D)[C],
                                            float& diag_in) {
  const int s0 = B * lb;
  float yv[B], ev[B], ov[B];
  const float* yb = sy + (kSmallYPad + r0 + s0 - lane);  // lane l at step s0 + u: row r0 + s0 + u - l
  const float* eb = edge_in + (kSmallEdgePad + s0);      // lane 0 at step s0 + u: edge row s0 + u
#pragma unroll
  for (int u = 0; u < B; u++) yv[u] = yb[u];
#pragma unroll
  for (int u = 0; u < B; u++) ev[u] = eb[u];
  const bool head = s0 < 63;             // some lane is before its first row
  const bool tail = s0 + B - 1 >= nrows;  // some lane passes its last row
  uint32_t code = 0;
  if (!tail && (!head || pro_free)) {
#pragma unroll
    for (int u = 0; u < B; u++) {
      const float left_in = Ar<float>::shr(ev[u], D[C - 1]);
      float dg = diag_in;
      diag_in = left_in;
      float l = left_in;
#pragma unroll
      for (int q = 0; q < C; q++) {
        const float up = D[q];
        const float df = yv[u] - x[q];
        const float v = cell<float, true, false>(up, l, dg, df * df, code);
        dg = up;
        D[q] = v;
        l = v;
      }
      ov[u] = D[C - 1];
    }
  } else if (pro_free) {
#pragma unroll
    for (int u = 0; u < B; u++) {
      const bool live = lane > s0 + u - nrows;  // row s0 + u - lane is not past the last
      const float left_in = Ar<float>::shr(ev[u], D[C - 1]);
      float dg = diag_in;
      diag_in = left_in;
      float l = left_in;
#pragma unroll
      for (int q = 0; q < C; q++) {
        const float up = D[q];
        const float df = yv[u] - x[q];
        const float v = cell<float, true, false>(up, l, dg, df * df, code);
        dg = up;
        D[q] = live ? v : up;
        l = v;
      }
      ov[u] = D[C - 1];
    }
  } else {
#pragma unroll
    for (int u = 0; u < B; u++) {
      const int s = s0 + u;
      const bool active = (unsigned)(s - lane) < (unsigned)nrows;
      const float left_in = Ar<float>::shr(ev[u], D[C - 1]);
      float dg = diag_in;
      diag_in = active ? left_in : diag_in;
      float l = left_in;
#pragma unroll
      for (int q = 0; q < C; q++) {
        const float up = D[q];
        const float df = yv[u] - x[q];
        const float v = cell<float, true, false>(up, l, dg, df * df, code);
        dg = up;
        D[q] = active ? v : up;
        l = v;
      }
      ov[u] = D[C - 1];
    }
  }
  if (has_out && lane == 63) {
    float* ob = edge_out + (kSmallEdgePad - 63 + s0);  // lane 63 at step s0 + u: row s0 + u - 63
#pragma unroll
    for (int u = 0; u < B; u += 4) *reinterpret_cast<float4*>(ob + u) = make_float4(ov[u], ov[u + 1], ov[u + 2], ov[u + 3]);
  }
}

// counter == nullptr: the grid holds one workgroup per pair (pair =
// blockIdx.x); otherwise workgroups claim pairs from *counter.
template <int C, int B, bool OS, bool OE>
__global__ void __launch_bounds__(64 * kSmallMaxWaves) k_all_pairs_small(
    const float* __restrict__ data, const SmallPairRec* __restrict__ tab, long long npairs,
    unsigned long long* __restrict__ counter, float* __restrict__ dist, int* __restrict__ flags) {
  constexpr int kLag = (64 + B) / B;  // sub-blocks between neighbouring strips
  static_assert((64 + B) % B == 0, "sub-block length divides 64");
  __shared__ float s_y[kSmallYLen];
  __shared__ __attribute__((aligned(16))) float s_edge[kSmallMaxWaves][kSmallEdgeLen];
  static_assert(kSmallEdgeLen % 4 == 0 && (kSmallEdgePad - 63) % 4 == 0, "aligned edge rows");
  __shared__ float s_top[kSmallMaxWaves];
  __shared__ unsigned long long s_p;
  const float MAXV = Ar<float>::maxv();
  const float RINF = Ar<float>::rowinf();
  const int lane = threadIdx.x & 63;
  const int k = uniform_i32((int)(threadIdx.x >> 6));
  const int nw = (int)(blockDim.x >> 6);
  for (int i = threadIdx.x; i < kSmallEdgeLen; i += blockDim.x) s_edge[nw - 1][i] = MAXV;
  // Y's guard rows before the first row: +inf (the prologue without masks;
  // without open start lanes read guard rows kSmallYPad - 63 .. kSmallYPad - 1)
  for (int i = threadIdx.x; i < kSmallYPad; i += blockDim.x) s_y[i] = RINF;
  for (long long it = 0;; ++it) {
    long long p;
    if (counter) {
      __syncthreads();  // every wave is done with the previous pair
      if (threadIdx.x == 0) s_p = atomicAdd(counter, 1ull);
      __syncthreads();
      p = (long long)uniform_u64(s_p);
    } else {
      p = it == 0 ? (long long)blockIdx.x : npairs;
    }
    if (p >= npairs) break;
    const SmallPairRec P = tab[p];
    const int ly = uniform_i32(P.ly), lx = uniform_i32(P.lx);
    const float* yp = data + P.oy;
    for (int i = threadIdx.x; i < ly; i += blockDim.x) s_y[kSmallYPad + i] = yp[i];
    const int S = (lx + 64 * C - 1) / (64 * C);
    const int off = k * 64 * C;
    const bool mine = k < S;
    const bool first = (k == 0);
    const bool has_out = k < S - 1;
    float x[C], D[C];
    const float* xp = data + P.ox + off + lane * C;  // padded: in bounds
    bool xfin = true;
#pragma unroll
    for (int q = 0; q < C; q++) {
      x[q] = xp[q];
      D[q] = OS ? 0.0f : RINF;
      xfin = xfin && (x[q] < RINF);  // false for +inf and NaN
    }
    const bool pro_free = !OS && __all(xfin);
    float diag_in = (lane == 0 && first) ? (OS ? MAXV : 0.0f) : (OS ? 0.0f : RINF);
    const int r0 = OS ? 1 : 0;
    const int r1 = (OE && ly >= 2) ? ly - 2 : ly - 1;
    const int nrows = r1 - r0 + 1;
    const int nsub = nrows > 0 ? (nrows + 63 + B - 1) / B : 0;
    const int G = nsub + (S - 1) * kLag;
    const float* edge_in = first ? s_edge[nw - 1] : s_edge[k - 1];
    float* edge_out = s_edge[k];
    __syncthreads();  // Y staged
    for (int g = 0; g < G; g++) {
      const int lb = g - k * kLag;
      if (mine && lb >= 0 && lb < nsub)
        small_steps<C, B, OS, OE>(lb, nrows, r0, has_out, pro_free, lane, s_y, edge_in, edge_out, x, D, diag_in);
      __syncthreads();
    }
    if (OE && ly >= 2) {  // the top row, strip after strip (dtw.hpp:172-200)
      for (int kk = 0; kk < S; kk++) {
        if (k == kk) {
          const int t = ly - 1;
          const float yv = s_y[kSmallYPad + t];
          const float lane0_diag = first ? MAXV : edge_in[kSmallEdgePad + nrows - 1];  // row t - 1 of the left strip
          const float left0 = first ? MAXV : s_top[k - 1];
          const float nb = Ar<float>::shr(lane0_diag, D[C - 1]);
          const bool bare0 = (lane == 0) && first;  // the pair's column 0: no diagonal
          float m[C];
          float run = 0.0f;
#pragma unroll
          for (int q = 0; q < C; q++) {
            const float df = yv - x[q];
            const float dd = df * df;
            const float U = D[q] + dd;
            const float Gd = (q == 0 && bare0) ? nb : ((q == 0 ? nb : D[q - 1]) + dd);
            const float mm = (Gd > U) ? U : Gd;
            run = (q == 0) ? mm : Ar<float>::mn(run, mm);
            m[q] = run;
          }
          const float pre = wave_excl_prefix_min<float>(m[C - 1], left0, lane);
#pragma unroll
          for (int q = 0; q < C; q++) D[q] = Ar<float>::mn(pre, m[q]);
          const float top = __shfl(D[C - 1], 63, 64);
          if (lane == 0) s_top[k] = top;
        }
        __syncthreads();
      }
    }
    if (k == S - 1) {  // the pair's last column is in this strip
      const int tl = lx - 1 - off;
      const int ll = tl / C, qq = tl - ll * C;
      float last = 0.0f;
#pragma unroll
      for (int q = 0; q < C; q++) {
        const float v = __shfl(D[q], ll, 64);
        if (q == qq) last = v;
      }
      if (lane == 0) {
        dist[P.idx] = pair_distance<float>(last, ly, OE != OS);
        if (flags) flags[P.idx] = 0;
      }
    }
    if (!counter) break;
  }
}

__global__ void k_band_init(int* __restrict__ g_cnt, unsigned long long* __restrict__ g_free, int R,
                            int* __restrict__ slot_of, long long nclaims, int* __restrict__ abr, int ntab,
                            int* __restrict__ bcnt = nullptr) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < R) g_cnt[i] = 0;
  if (bcnt && i < ntab) bcnt[i] = 0;
  if (i < (R + 63) / 64) {
    const int rem = R - (int)i * 64;
    g_free[i] = rem >= 64 ? ~0ull : ((1ull << rem) - 1);
  }
  if (i < nclaims) slot_of[i] = -1;
  if (i < ntab) abr[i] = INT_MAX;
}

__global__ void k_band_flags(const BandPair* __restrict__ tab, int ntab, int n, const int* __restrict__ abr,
                             int* __restrict__ flags) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntab) return;
  const int r = tab[t].r, c = tab[t].c;
  const long long idx = (long long)r * (2LL * n - r - 1) / 2 + (c - r - 1);
  flags[idx] = abr[t] != INT_MAX ? (1 | (abr[t] << 8)) : 0;
}

// ---------------------------------------------------------------------------
// Path-DTW jobs (DBA update): job j aligns Y=ys[j] against X=xs[j] and stores
// step words at mv + mv_off[j], top-row words at mtop + top_off[j].
// ---------------------------------------------------------------------------
template <typename T> struct PathJob {
  const T* y;
  const T* x;
  int ly, lx;
  long long mv_off;   // in 32-bit words
  long long top_off;  // in 32-bit words
};

template <typename T, bool OS, bool OE, bool MINFIRST>
__global__ void __launch_bounds__(256) k_path_jobs(const PathJob<T>* __restrict__ jobs, int njobs,
                                                   unsigned int* __restrict__ counter, uint32_t* __restrict__ mv,
                                                   uint32_t* __restrict__ mtop, int* __restrict__ abrogated,
                                                   T* __restrict__ escratch, long long escratch_stride) {
  __shared__ T s_y[4][128];
  __shared__ T s_e[4][64];
  __shared__ T s_o[4][128];
  const int lane = threadIdx.x & 63;
  const int wid = uniform_i32((int)(threadIdx.x >> 6));
  const long long gw = (long long)blockIdx.x * 4 + wid;
  T* e0 = escratch + gw * 2 * escratch_stride;
  T* e1 = e0 + escratch_stride;
  const WaveWin<T> win{s_y[wid], s_e[wid], s_o[wid]};
  const WaveWin<T>* W = &win;
  for (;;) {
    const unsigned int j = (unsigned int)uniform_i32((int)atomicAdd(counter, lane == 0 ? 1u : 0u));
    if ((int)j >= njobs) break;
    PathJob<T> jb = jobs[j];
    PairOut po;
    (void)dtw_one_pair<T, OS, OE, MINFIRST, true>(jb.y, jb.ly, jb.x, jb.lx, e0, e1, mv + jb.mv_off, mtop + jb.top_off,
                                                  W, lane, po);
    if (lane == 0) abrogated[j] = po.abrogated;
  }
}

// ---------------------------------------------------------------------------
// Path jobs with a workgroup of waves per job (few long jobs: the DBA round
// of a cluster with tens to hundreds of members). The strips are the
// reference's 1024-column swaths, so the step words land exactly where the
// one-wave kernel puts them; the edge hand-off is that of the strip
// schedules. The open-end stop is decided at each swath's end as the strips
// finish, and once it is decided the strips at or past it stop sweeping
// (and later ones never start): a short sequence against a long centroid
// sweeps a few swaths past its own length, not the whole centroid. Words of
// swaths at or past the stop may be left unwritten or partly written; the
// traceback never reads them (it treats those cells positionally).
// ---------------------------------------------------------------------------
// With QW every swath of 4+ columns per lane is split into four quarter
// strips (C/4 columns per lane, one wave each), so a pair of one to four
// swaths (a long sequence against a short centroid) still spreads over four
// times as many waves; the step words land in the same layout.
template <typename T, bool OS, bool OE, bool QW>
__global__ void __launch_bounds__(64 * kPathStripWaves) k_path_strips(const PathJob<T>* __restrict__ jobs, int njobs,
                                                      unsigned int* __restrict__ counter, uint32_t* __restrict__ mv,
                                                      uint32_t* __restrict__ mtop, int* __restrict__ abrogated,
                                                      T* __restrict__ escratch, long long stride) {
  __shared__ T s_y[kPathStripWaves][128];
  __shared__ T s_e[kPathStripWaves][64];
  __shared__ T s_o[kPathStripWaves][128];
  __shared__ int s_owner[2 * kPathStripWaves];
  __shared__ int s_cnt[2 * kPathStripWaves];
  __shared__ int s_free[2 * kPathStripWaves];
  __shared__ int s_job;
  __shared__ int s_abr;
  const int lane = threadIdx.x & 63;
  const int wid = uniform_i32((int)(threadIdx.x >> 6));
  const int nw = uniform_i32((int)(blockDim.x >> 6));
  const int nslot = 2 * nw;
  T* ebase = escratch + (long long)blockIdx.x * nslot * stride;
  const WaveWin<T> win{s_y[wid], s_e[wid], s_o[wid]};
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) {
      s_job = (int)atomicAdd(counter, 1u);
      s_abr = INT_MAX;
    }
    if (threadIdx.x < 2 * kPathStripWaves) {
      s_owner[threadIdx.x] = -1;
      s_cnt[threadIdx.x] = 0;
      s_free[threadIdx.x] = 1;
    }
    __syncthreads();
    const int claim = uniform_i32(s_job);
    if (claim >= njobs) break;
    // Claims go from the last job to the first: callers list jobs by
    // ascending sequence length, so the longest start first and the short
    // ones fill in behind them.
    const int jix = njobs - 1 - claim;
    const PathJob<T> jb = jobs[jix];
    const int ly = jb.ly, lx = jb.lx;
    const int K = (lx + ODBA_W - 1) / ODBA_W;
    const int nrows_main = ((OE && ly >= 2) ? ly - 2 : ly - 1) - (OS ? 1 : 0) + 1;
    const size_t swath_words = (size_t)(nrows_main > 0 ? nrows_main + 63 : 0) * 64;
    const int Cl = last_swath_c(lx - (K - 1) * ODBA_W);
    const int qlast = (QW && Cl >= 4) ? 4 : 1;  // strips of the last swath
    const int NS = QW ? 4 * (K - 1) + qlast : K;
    for (int s = wid; s < NS; s += nw) {
      const int k = QW ? (s >> 2) : s, q = QW ? (s & 3) : 0;
      const bool split = QW && (k < K - 1 || qlast == 4);
      const int C = (k < K - 1) ? (split ? 4 : 16) : (split ? Cl / 4 : Cl);
      const int off = k * ODBA_W + q * 64 * C;
      const int so = s % nslot, si = (s + nslot - 1) % nslot;
      const bool has_out = (s < NS - 1);
      StripPipe pp{&s_owner[si], &s_cnt[si], s - 1, has_out ? &s_cnt[so] : nullptr, lane};
      if (OE && !OS) {  // strips at or past a decided open-end stop are cancelled
        pp.stop = &s_abr;
        pp.swath = k;
      }
      // A cancelled strip claims no slot (a slot's previous reader may still
      // be running) and sweeps nothing. Strips are cancelled in strip order,
      // so every strip that waits on a cancelled one is cancelled too and
      // leaves its waits (run_swath polls the stop word in them).
      bool skip = pp.cancelled();
      if (has_out && !skip) {
        for (;;) {
          const int f = uniform_i32(__hip_atomic_load(&s_free[so], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
          if (f) break;
          if (pp.cancelled()) {
            skip = true;
            break;
          }
          __builtin_amdgcn_s_sleep(2);
        }
        if (!skip && lane == 0) {
          __hip_atomic_store(&s_free[so], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          __hip_atomic_store(&s_cnt[so], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          __hip_atomic_store(&s_owner[so], s, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      }
      SwathIO<T> io;
      io.y = jb.y; io.ly = ly; io.x = jb.x; io.lx = lx; io.off = off;
      io.ein = s > 0 ? ebase + (long long)si * stride : nullptr;
      io.eout = has_out ? ebase + (long long)so * stride : nullptr;
      io.mv = mv + jb.mv_off + (size_t)k * swath_words;
      io.mtop = mtop + jb.top_off + (size_t)k * 64;
      io.ref_left = (q == 0);
      io.lbase = q * 16;
      SwathRes<T> res;
      if (!skip)
        res = split ? swath_dispatch_quarter<T, OS, OE, StripPipe>(C, io, &win, lane, pp)
                    : swath_dispatch_pipe<T, OS, OE, true, StripPipe>(C, io, &win, lane, pp);
      if (s > 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        if (lane == 0) __hip_atomic_store(&s_free[si], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
      // the open-end stop is decided on a swath's right edge column
      // (dtw.hpp:74-129); a strip cancelled meanwhile holds no edge, and a
      // stop already decided at or below its swath is the lower one anyway
      const bool swath_end = !split || q == 3;
      if (OE && !OS && has_out && swath_end && !skip && !pp.cancelled() && k * ODBA_W + ODBA_W > ly &&
          res.top == res.colmin && lane == 0)
        atomicMin(&s_abr, k + 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) abrogated[jix] = (s_abr != INT_MAX) ? s_abr : -1;
  }
}

// ---------------------------------------------------------------------------
// Path jobs as bands (float): one band per 1024-column swath of a job, swept
// by a workgroup of four waves, one per quarter swath (or wave 0 alone for
// a last swath of 1 or 2 columns per lane), so the swaths of one job run on
// as many CUs as it has swaths. Claims go in job order from the last job to
// the first (callers list jobs by ascending length: the longest start first)
// and swath order within a job; inside a band the edge column passes
// through LDS-counted rings, between bands through whole-column global slots
// from a free pool, exactly as in k_all_pairs_bands, so every wait targets a
// lower claim held by a running workgroup and times out into err bit 8. The
// step words land in the one-wave layout; a swath's last quarter decides the
// open-end stop (dtw.hpp:74-129) into abr[job] (INT_MAX: none).
// band0[p] = first claim of the p-th job in claim order; band0[njobs] = all.
// ---------------------------------------------------------------------------
// Bands of a job claimed in the first phase. In open end without open
// start the sweep may stop after any swath whose right edge passes the last
// row ((s + 1) * W > ly, dtw.hpp:74-129); a job's bands after the one
// holding the first such swath are claimed in a second phase, after every
// job's first-phase bands, so that most of them find the stop already
// decided and sweep nothing (claimed left to right right away, they ran a
// few rows behind their left neighbours and were nearly done by the time
// the stop was known).
__host__ __device__ inline int first_phase_bands(int ly, int K, int per_band, bool stop_rule) {
  const int nb = (K + per_band - 1) / per_band;
  if (!stop_rule) return nb;
  const int ks = ly / ODBA_W;  // first swath whose right edge passes the last row
  return min(nb, ks / per_band + 1);
}

// band0[p] = first claim of the p-th job in claim order (its first-phase
// bands), band0[njobs] = first-phase claims; with band1, band1[p] = first
// claim of its second-phase bands (all after band0[njobs]) and
// band1[njobs] = all claims.
__global__ void __launch_bounds__(64) k_path_band0(const PathJob<float>* __restrict__ jobs, int njobs,
                                                     long long* __restrict__ band0, int per_band,
                                                     long long* __restrict__ band1 = nullptr, int stop_rule = 0) {
  // One wave (it needs one free wave slot, not a whole CU, so it starts as
  // soon as any wave of another stream's kernel retires): thread t sums the
  // bands of its chunk of claim-order jobs, the chunk sums are scanned, then
  // each thread writes its chunk's prefixes; once per phase.
  __shared__ long long s[64];
  const int t = threadIdx.x;
  const int per = (njobs + 63) / 64;
  const int p0 = min(njobs, t * per), p1 = min(njobs, p0 + per);
  auto bands = [&](int p, int ph) -> long long {
    const PathJob<float>& jb = jobs[njobs - 1 - p];
    const int K = (jb.lx + ODBA_W - 1) / ODBA_W;  // swaths of the job
    const int nb = (K + per_band - 1) / per_band;
    const int n1 = band1 ? first_phase_bands(jb.ly, K, per_band, stop_rule != 0) : nb;
    return (long long)(ph ? nb - n1 : n1);
  };
  long long base = 0;
  for (int ph = 0; ph < (band1 ? 2 : 1); ph++) {
    long long* out = ph ? band1 : band0;
    long long mine = 0;
    for (int p = p0; p < p1; p++) mine += bands(p, ph);
    __syncthreads();  // every thread has read the previous phase's total
    s[t] = mine;
    __syncthreads();
    for (int o = 1; o < 64; o <<= 1) {
      const long long a = t >= o ? s[t - o] : 0;
      __syncthreads();
      s[t] += a;
      __syncthreads();
    }
    long long run = base + s[t] - mine;
    for (int p = p0; p < p1; p++) {
      out[p] = run;
      run += bands(p, ph);
    }
    if (t == 63) out[njobs] = base + s[t];
    base += s[63];
  }
}

__global__ void k_abr_index(int* __restrict__ abr, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && abr[i] == INT_MAX) abr[i] = -1;
}

template <bool OS, bool OE>
__global__ void __launch_bounds__(256) k_path_bands(const PathJob<float>* __restrict__ jobs, int njobs,
                                                    const long long* __restrict__ band0,
                                                    unsigned long long* __restrict__ counter,
                                                    uint32_t* __restrict__ mv, uint32_t* __restrict__ mtop,
                                                    float* __restrict__ rings, int ring, float* __restrict__ gslots,
                                                    long long gstride, int* __restrict__ g_cnt,
                                                    unsigned long long* __restrict__ g_free, int R,
                                                    int* __restrict__ slot_of, int* __restrict__ abr,
                                                    int* __restrict__ err, long long limit) {
  using T = float;
  constexpr int NW = 4;
  __shared__ T s_y[NW][128];
  __shared__ T s_e[NW][64];
  __shared__ T s_o[NW][128];
  __shared__ int s_cnt[NW];   // rows published into ring w
  __shared__ int s_used[NW];  // rows of ring w read by wave w+1
  __shared__ unsigned long long s_q;
  const int lane = threadIdx.x & 63;
  const int wid = uniform_i32((int)(threadIdx.x >> 6));
  T* rbase = rings + (long long)blockIdx.x * NW * ring;
  const WaveWin<T> win{s_y[wid], s_e[wid], s_o[wid]};
  const long long nclaims = band0[njobs];
  for (;;) {
    __syncthreads();  // every wave is done with the previous claim
    if (threadIdx.x == 0) s_q = atomicAdd(counter, 1ull);
    if (threadIdx.x < NW) {
      s_cnt[threadIdx.x] = 0;
      s_used[threadIdx.x] = 0;
    }
    __syncthreads();
    const long long q = (long long)uniform_u64(s_q);
    if (q >= nclaims) break;
    int lo = 0, hi = njobs - 1;  // claim-order position: last with band0 <= q
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (band0[mid] <= q) lo = mid;
      else hi = mid - 1;
    }
    const int pos = uniform_i32(lo);
    const int j = njobs - 1 - pos;
    const PathJob<T> jb = jobs[j];
    const int ly = jb.ly, lx = jb.lx;
    const int K = (lx + ODBA_W - 1) / ODBA_W;
    const int k = uniform_i32((int)(q - band0[pos]));  // this band's swath
    const int Cl = last_swath_c(lx - (K - 1) * ODBA_W);
    const bool split = (k < K - 1) || Cl >= 4;
    const int C = (k < K - 1) ? 4 : (split ? Cl / 4 : Cl);
    const int nstr = split ? 4 : 1;
    if (wid >= nstr) continue;  // a last swath of 1 or 2 columns per lane: wave 0 alone
    const int off = k * ODBA_W + wid * 64 * C;
    const bool has_out = !(k == K - 1 && wid == nstr - 1);
    const bool in_global = (wid == 0) && k > 0;
    const bool out_global = (wid == nstr - 1) && has_out;
    int gout = 0, gin = 0;
    if (out_global) {  // take a slot, reset its progress word, then announce it
      gout = band_pop_slot(g_free, (R + 63) / 64, err, limit, lane);
      if (lane == 0) {
        __hip_atomic_store(&g_cnt[gout], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&slot_of[q], gout, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    if (in_global && !band_failed(err)) {  // the left swath's slot
      const long long t0 = wall_clock64();
      for (;;) {
        gin = uniform_i32(__hip_atomic_load(&slot_of[q - 1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT));
        if (gin >= 0) break;
        if (band_timed_out(t0, limit, err, lane)) { gin = 0; break; }
        __builtin_amdgcn_s_sleep(2);
      }
    }
    BandPipe pp;
    pp.in_global = in_global;
    pp.in_cnt = in_global ? &g_cnt[gin] : (wid > 0 ? &s_cnt[wid - 1] : nullptr);
    pp.in_used = (wid > 0) ? &s_used[wid - 1] : nullptr;
    pp.out_cnt = has_out ? (out_global ? &g_cnt[gout] : &s_cnt[wid]) : nullptr;
    pp.out_global = out_global;
    pp.out_used = (has_out && !out_global) ? &s_used[wid] : nullptr;
    pp.ring = ring;
    pp.main_end = (OE && ly >= 2) ? ly - 1 : ly;
    pp.lane = lane;
    pp.err = err;
    pp.limit = limit;
    const int nrows_main = ((OE && ly >= 2) ? ly - 2 : ly - 1) - (OS ? 1 : 0) + 1;
    const size_t swath_words = (size_t)(nrows_main > 0 ? nrows_main + 63 : 0) * 64;
    SwathIO<T> io;
    io.y = jb.y; io.ly = ly; io.x = jb.x; io.lx = lx; io.off = off;
    io.ein = (k == 0 && wid == 0) ? nullptr : (in_global ? gslots + (long long)gin * gstride : rbase + (long long)(wid - 1) * ring);
    io.imask = in_global ? -1 : ring - 1;
    io.eout = has_out ? (out_global ? gslots + (long long)gout * gstride : rbase + (long long)wid * ring) : nullptr;
    io.omask = out_global ? -1 : ring - 1;
    io.mv = mv + jb.mv_off + (size_t)k * swath_words;
    io.mtop = mtop + jb.top_off + (size_t)k * 64;
    io.ref_left = (wid == 0);
    io.lbase = wid * 16;
    SwathRes<T> res = split ? swath_dispatch_quarter<T, OS, OE, BandPipe>(C, io, &win, lane, pp)
                            : swath_dispatch_pipe<T, OS, OE, true, BandPipe>(C, io, &win, lane, pp);
    if (in_global && lane == 0) {  // every read of the slot has returned: back to the pool
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      atomicOr(&g_free[gin >> 6], 1ull << (gin & 63));
    }
    if (OE && !OS && out_global && (k + 1) * ODBA_W > ly && res.top == res.colmin && lane == 0)
      atomicMin(&abr[j], k + 1);
  }
}

// ---------------------------------------------------------------------------
// Path jobs as bands of whole swaths (float): the band hand-off of
// k_all_pairs_bands with step words. A job's 1024-column swaths are cut
// into bands of nw consecutive swaths; workgroups of nw waves claim bands
// from one counter, jobs from the last (longest, as callers list them by
// ascending length) to the first and bands left to right, and wave w sweeps
// swath w of its band as one strip of C = 16 columns per lane (the last
// swath with its own C), so the step words land exactly where the one-wave
// kernel puts them. Inside a band the edge column goes through LDS-counted
// rings; between bands through whole-column global slots from a free pool,
// published every 256 rows. A long pair's swaths thus run on as many
// workgroups (and CUs) as it has bands instead of on one.
//
// No deadlock: every wait targets a lower claim, or a lower strip of the
// same claim (the ring wait of a producer targets the next strip of its
// band, whose band's last strip never waits on a ring); claims are taken in
// order by running workgroups. Every wait still ends after `limit` ticks,
// setting bit 8 of *err.
//
// The open-end stop (dtw.hpp:74-129) is decided at each swath's end into
// abr[job] (an atomic minimum; INT_MAX until k_abr_index turns it into -1);
// swaths past it are swept anyway and their words are never read.
// band0[p] = first claim of the p-th job in claim order; band0[njobs] = all.
// ---------------------------------------------------------------------------
template <bool OS, bool OE>
__global__ void __launch_bounds__(64 * kPathStripWaves) k_path_swath_bands(const PathJob<float>* __restrict__ jobs, int njobs,
                                                           const long long* __restrict__ band0,
                                                           unsigned long long* __restrict__ counter,
                                                           uint32_t* __restrict__ mv, uint32_t* __restrict__ mtop,
                                                           float* __restrict__ rings, int ring,
                                                           float* __restrict__ gslots, long long gstride,
                                                           int* __restrict__ g_cnt, unsigned long long* __restrict__ g_free,
                                                           int R, int* __restrict__ slot_of, int* __restrict__ abr,
                                                           int* __restrict__ err, long long limit,
                                                           const long long* __restrict__ band1,
                                                           float* __restrict__ bslots, int* __restrict__ bcnt) {
  using T = float;
  constexpr bool kStop = OE && !OS;  // the open-end stop rule applies
  __shared__ T s_y[kPathStripWaves][128];
  __shared__ T s_e[kPathStripWaves][64];
  __shared__ T s_o[kPathStripWaves][128];
  __shared__ int s_cnt[kPathStripWaves];   // rows published into ring w
  __shared__ int s_used[kPathStripWaves];  // rows of ring w read by wave w+1
  __shared__ unsigned long long s_q;
  const int lane = threadIdx.x & 63;
  const int wid = uniform_i32((int)(threadIdx.x >> 6));
  const int nw = uniform_i32((int)(blockDim.x >> 6));
  T* rbase = rings + (long long)blockIdx.x * nw * ring;
  const WaveWin<T> win{s_y[wid], s_e[wid], s_o[wid]};
  const long long nfirst = band0[njobs];                    // first-phase claims
  const long long nclaims = kStop ? band1[njobs] : nfirst;  // all claims
  for (;;) {
    __syncthreads();  // every wave is done with the previous claim
    if (threadIdx.x == 0) s_q = atomicAdd(counter, 1ull);
    if (threadIdx.x < kPathStripWaves) {
      s_cnt[threadIdx.x] = 0;
      s_used[threadIdx.x] = 0;
    }
    __syncthreads();
    const long long q = (long long)uniform_u64(s_q);
    if (q >= nclaims) break;
    // claim-order position: last with (first- or second-phase) start <= q
    const bool second = q >= nfirst;
    const long long* tabq = second ? band1 : band0;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tabq[mid] <= q) lo = mid;
      else hi = mid - 1;
    }
    const int pos = uniform_i32(lo);
    const int j = njobs - 1 - pos;
    const PathJob<T> jb = jobs[j];
    const int ly = jb.ly, lx = jb.lx;
    const int K = (lx + ODBA_W - 1) / ODBA_W;
    const int nb = (K + nw - 1) / nw;
    const int n1 = first_phase_bands(ly, K, nw, kStop);                   // the job's first-phase bands
    const int b = uniform_i32((int)(q - tabq[pos]) + (second ? n1 : 0));  // band of the job
    // The hand-off between the job's last first-phase band and its first
    // second-phase band goes through the job's own boundary slot (edge
    // column and progress word), not the shared pool: the two claims are far
    // apart, and pool slots held that long could leave the pool empty.
    const bool bnd_out = kStop && b == n1 - 1 && n1 < nb;
    const bool bnd_in = kStop && b == n1 && b > 0;
    const int s0 = b * nw;
    const int slast = min(s0 + nw, K) - 1;
    const int s = s0 + wid;  // this wave's swath
    if (s > slast) continue;  // no swath for this wave in the job's last band
    const int C = (s < K - 1) ? 16 : last_swath_c(lx - (K - 1) * ODBA_W);
    const bool has_out = s < K - 1;
    const bool in_global = (wid == 0) && s > 0;
    const bool out_global = (s == slast) && has_out;
    BandPipe pp;
    if (OE && !OS) {  // strips at or past a decided open-end stop are cancelled
      pp.stop = &abr[j];
      pp.swath = s;
    }
    // A strip already past the stop sweeps nothing; as a producer it takes
    // no slot and announces none (-2), as a consumer it still returns its
    // producer's slot once that producer is done with it.
    bool skip = pp.cancelled();
    int gout = 0, gin = 0;
    // took: this strip announced a slot, so its consumer waits for the
    // slot's done mark; it must write that mark even if it turns to skip
    // below (one wave per band is both the consumer of the left band's slot
    // and the producer of its own, and may learn only after announcing that
    // its left neighbour was past the stop)
    const bool pool_out = out_global && !bnd_out;
    const bool pool_in = in_global && !bnd_in;
    const bool took = pool_out && !skip;
    if (pool_out) {  // take a slot, reset its progress word, then announce it
      if (took) gout = band_pop_slot(g_free, (R + 63) / 64, err, limit, lane);
      if (lane == 0) {
        if (took) __hip_atomic_store(&g_cnt[gout], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&slot_of[q], took ? gout : -2, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    bool had_slot = false;
    if (pool_in && !band_failed(err)) {  // the left band's slot, or -2: none
      const long long t0 = wall_clock64();
      for (;;) {
        gin = uniform_i32(__hip_atomic_load(&slot_of[q - 1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT));
        if (gin != -1) break;
        if (band_timed_out(t0, limit, err, lane)) { gin = 0; break; }
        __builtin_amdgcn_s_sleep(2);
      }
      had_slot = gin >= 0;
      if (!had_slot) {  // a producer past the stop: so is this strip
        skip = true;
        gin = 0;
      }
    }
    // A boundary consumer waits on the boundary progress word and leaves
    // that wait when its producer is past the stop: the stop word then
    // reaches this strip's swath too.
    pp.in_global = in_global;
    pp.in_cnt = in_global ? (bnd_in ? &bcnt[pos] : &g_cnt[gin]) : (s > 0 ? &s_cnt[wid - 1] : nullptr);
    pp.in_used = (s > 0 && !in_global) ? &s_used[wid - 1] : nullptr;
    pp.out_cnt = has_out ? (out_global ? (bnd_out ? &bcnt[pos] : &g_cnt[gout]) : &s_cnt[wid]) : nullptr;
    pp.out_global = out_global;
    pp.out_used = (has_out && !out_global) ? &s_used[wid] : nullptr;
    pp.ring = ring;
    pp.main_end = (OE && ly >= 2) ? ly - 1 : ly;
    pp.lane = lane;
    pp.err = err;
    pp.limit = limit;
    const int nrows_main = ((OE && ly >= 2) ? ly - 2 : ly - 1) - (OS ? 1 : 0) + 1;
    const size_t swath_words = (size_t)(nrows_main > 0 ? nrows_main + 63 : 0) * 64;
    SwathIO<T> io;
    io.y = jb.y; io.ly = ly; io.x = jb.x; io.lx = lx; io.off = s * ODBA_W;
    float* const bslot = kStop ? bslots + (long long)pos * gstride : nullptr;
    io.ein = s == 0 ? nullptr
                    : (in_global ? (bnd_in ? bslot : gslots + (long long)gin * gstride)
                                 : rbase + (long long)(wid - 1) * ring);
    io.imask = in_global ? -1 : ring - 1;
    io.eout = has_out ? (out_global ? (bnd_out ? bslot : gslots + (long long)gout * gstride)
                                    : rbase + (long long)wid * ring)
                      : nullptr;
    io.omask = out_global ? -1 : ring - 1;
    io.mv = mv + jb.mv_off + (size_t)s * swath_words;
    io.mtop = mtop + jb.top_off + (size_t)s * 64;
    io.ref_left = 1;
    SwathRes<T> res;
    if (!skip) res = swath_dispatch_pipe<T, OS, OE, true, BandPipe>(C, io, &win, lane, pp);
    if (OE && !OS && took && lane == 0) {
      // done with the slot (every row, stopped past the stop, or never
      // started): a consumer that stopped early returns the slot only after
      // this mark
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      __hip_atomic_store(&g_cnt[gout], INT_MAX, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (OE && !OS && had_slot && !band_failed(err)) {
      // the producer may still be writing the slot if this strip stopped
      // early (or never started): wait for its done mark
      const long long t0 = wall_clock64();
      for (;;) {
        const int c = uniform_i32(__hip_atomic_load(&g_cnt[gin], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (c == INT_MAX) break;
        if (band_timed_out(t0, limit, err, lane)) break;
        __builtin_amdgcn_s_sleep(2);
      }
    }
    if (had_slot && lane == 0) {  // every read of the slot has returned: back to the pool
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      atomicOr(&g_free[gin >> 6], 1ull << (gin & 63));
    }
    if (OE && !OS && has_out && !skip && !pp.cancelled() && (s + 1) * ODBA_W > ly && res.top == res.colmin &&
        lane == 0)
      atomicMin(&abr[j], s + 1);
  }
}

// ---------------------------------------------------------------------------
// Checkpointed paths (global mode, float): the step codes of a long pair
// without its whole step matrix. In global mode the path does not depend on
// how the matrix is cut (every cell's code follows from exact cost values),
// so the stripe mode's result is the full matrix's, and any exact recompute
// gives it.
//
//   * Forward (k_ckpt_forward): the whole matrix once, as bands of 16
//     strips of 1024 columns (one block of E columns per claim, claims in
//     job order), keeping two kinds of checkpoint: the cost row at the
//     bottom of every band of R rows (rowck[b][j] = D(bR+R-1, j)) and the
//     cost column at the right edge of every column block
//     (colck[m][i] = D(i, mE+E-1), which is also the hand-off to the next
//     block, published in a per-block progress word).
//   * Then, block by block along the path (k_ckpt_block, one workgroup per
//     job): the block that holds the path's current cell is recomputed from
//     its top checkpoint row and left checkpoint column, rows up to the
//     current cell's, columns up to its column, storing step codes; the
//     traceback (k_ckpt_trace) walks it until the path leaves the block
//     (down or left), carrying the walk state to the next launch.
// ---------------------------------------------------------------------------
struct CkJob {
  const float* y;  // rows
  const float* x;  // columns
  int ly, lx;
  long long rowck_off;  // elements: (nb - 1) rows of lx
  long long colck_off;  // elements: (nm - 1) columns of ly + 256
  long long cnt_off;    // ints: nm - 1 progress words
  long long claim0;     // the job's first claim (column block 0)
  long long runs_off;   // its run records (one per column of X)
  long long path_off;   // its recorded cells, or -1
};

template <int EW>
__global__ void __launch_bounds__(1024) k_ckpt_forward(const CkJob* __restrict__ jobs, int njobs, long long nclaims,
                                                       unsigned long long* __restrict__ counter,
                                                       float* __restrict__ rowck, float* __restrict__ colck,
                                                       int* __restrict__ cnt, float* __restrict__ rings, int ring,
                                                       int rshift, int* __restrict__ err, long long limit) {
  __shared__ float s_y[kMaxStripWaves][128];
  __shared__ float s_e[kMaxStripWaves][64];
  __shared__ float s_o[kMaxStripWaves][128];
  __shared__ int s_cnt[kMaxStripWaves];
  __shared__ int s_used[kMaxStripWaves];
  __shared__ unsigned long long s_q;
  const int lane = threadIdx.x & 63;
  const int wid = uniform_i32((int)(threadIdx.x >> 6));
  float* rbase = rings + (long long)blockIdx.x * EW * ring;
  const WaveWin<float> win{s_y[wid], s_e[wid], s_o[wid]};
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_q = atomicAdd(counter, 1ull);
    if (threadIdx.x < kMaxStripWaves) {
      s_cnt[threadIdx.x] = 0;
      s_used[threadIdx.x] = 0;
    }
    __syncthreads();
    const long long q = (long long)uniform_u64(s_q);
    if (q >= nclaims) break;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (jobs[mid].claim0 <= q) lo = mid;
      else hi = mid - 1;
    }
    const int t = uniform_i32(lo);
    const CkJob J = jobs[t];
    const int m = uniform_i32((int)(q - J.claim0));
    const int S = (J.lx + ODBA_W - 1) / ODBA_W;
    const int s0 = m * EW, slast = min(s0 + EW, S) - 1;
    const int s = s0 + wid;
    if (wid >= EW || s > slast) continue;
    const int off = s * ODBA_W;
    const bool has_out = s < S - 1;
    const bool in_global = (wid == 0) && m > 0;
    const bool out_global = (s == slast) && has_out;
    const long long cstride = (long long)J.ly + 256;
    BandPipe pp;
    pp.in_global = in_global;
    pp.in_cnt = in_global ? &cnt[J.cnt_off + m - 1] : (wid > 0 ? &s_cnt[wid - 1] : nullptr);
    pp.in_used = (wid > 0) ? &s_used[wid - 1] : nullptr;
    pp.out_cnt = has_out ? (out_global ? &cnt[J.cnt_off + m] : &s_cnt[wid]) : nullptr;
    pp.out_global = out_global;
    pp.out_used = (has_out && !out_global) ? &s_used[wid] : nullptr;
    pp.ring = ring;
    pp.main_end = J.ly;
    pp.lane = lane;
    pp.err = err;
    pp.limit = limit;
    SwathIO<float> io;
    io.y = J.y; io.ly = J.ly; io.x = J.x; io.lx = J.lx; io.off = off;
    io.ein = s == 0 ? nullptr : (in_global ? colck + J.colck_off + (long long)(m - 1) * cstride : rbase + (long long)(wid - 1) * ring);
    io.imask = in_global ? -1 : ring - 1;
    io.eout = has_out ? (out_global ? colck + J.colck_off + (long long)m * cstride : rbase + (long long)wid * ring) : nullptr;
    io.omask = out_global ? -1 : ring - 1;
    io.mv = nullptr; io.mtop = nullptr;
    io.ref_left = 1;
    io.rowck = rowck + J.rowck_off;
    io.ck_stride = J.lx;
    io.ck_rbase = 0;
    io.ck_rshift = rshift;
    const int nb = (J.ly + (1 << rshift) - 1) >> rshift;
    io.ck_rlimit = (nb - 1) << rshift;
    const int wcols = min(ODBA_W, J.lx - off);
    const int C = (s == S - 1) ? last_swath_c(wcols) : 16;
    switch (C) {
      case 1: run_swath<float, 1, false, false, true, false, BandPipe, true>(io, &win, lane, pp); break;
      case 2: run_swath<float, 2, false, false, true, false, BandPipe, true>(io, &win, lane, pp); break;
      case 4: run_swath<float, 4, false, false, true, false, BandPipe, true>(io, &win, lane, pp); break;
      case 8: run_swath<float, 8, false, false, true, false, BandPipe, true>(io, &win, lane, pp); break;
      case 12: run_swath<float, 12, false, false, true, false, BandPipe, true>(io, &win, lane, pp); break;
      default: run_swath<float, 16, false, false, true, false, BandPipe, true>(io, &win, lane, pp); break;
    }
  }
}

// Walk state of one checkpointed path between launches.
struct CkState {
  int i, j;          // current cell
  int curc;          // current centroid position
  int lo1, hi1, lo2, hi2, gap;
  long long np;      // recorded cells so far
  int done, err;
  int bi;            // the block's top row when it was recomputed (its rows bR .. bi)
  int pad;
};

// Recomputes, for every job not done, the block holding its current cell:
// rows bR .. i, columns mE .. j, step words [strip][step][lane] (strip =
// 1024 columns, steps over rows bR .. i) at mv + t * job_words. Strip w
// hands its right edge to strip w+1 through slot w (LDS progress words);
// strip 0 reads the left checkpoint column, which is complete.
template <int EW>
__global__ void __launch_bounds__(1024) k_ckpt_block(const CkJob* __restrict__ jobs, CkState* __restrict__ st,
                                                     const float* __restrict__ rowck, const float* __restrict__ colck,
                                                     uint32_t* __restrict__ mv, long long job_words,
                                                     float* __restrict__ slots, long long slot_stride, int rshift) {
  __shared__ float s_y[kMaxStripWaves][128];
  __shared__ float s_e[kMaxStripWaves][64];
  __shared__ float s_o[kMaxStripWaves][128];
  __shared__ int s_owner[kMaxStripWaves + 1];
  __shared__ int s_cnt[kMaxStripWaves + 1];
  const int lane = threadIdx.x & 63;
  const int wid = uniform_i32((int)(threadIdx.x >> 6));
  const int t = blockIdx.x;
  const CkJob J = jobs[t];
  const int done = uniform_i32(st[t].done | st[t].err);
  const int i = uniform_i32(st[t].i), j = uniform_i32(st[t].j);
  if (threadIdx.x <= kMaxStripWaves) {
    s_owner[threadIdx.x] = threadIdx.x < kMaxStripWaves ? (int)threadIdx.x : -1;
    s_cnt[threadIdx.x] = threadIdx.x < kMaxStripWaves ? 0 : INT_MAX;  // the last one: always ready
  }
  __syncthreads();
  if (done || i < 0 || j < 0) return;
  const int R = 1 << rshift;
  const int b = i >> rshift;
  const int m = j / (EW * ODBA_W);
  const int bR = b * R;
  const int Rb = i - bR + 1;
  const int s = m * EW + wid;  // global strip (swath) index
  const int off = s * ODBA_W;
  if (wid >= EW || off > j) return;
  const int S = (J.lx + ODBA_W - 1) / ODBA_W;
  const bool has_out = off + ODBA_W <= j;  // strip w+1 of this block is needed
  const int in_slot = wid > 0 ? wid - 1 : kMaxStripWaves;
  StripPipe pp{&s_owner[in_slot], &s_cnt[in_slot], wid > 0 ? wid - 1 : -1, has_out ? &s_cnt[wid] : nullptr, lane};
  const WaveWin<float> win{s_y[wid], s_e[wid], s_o[wid]};
  SwathIO<float> io;
  io.y = J.y + bR; io.ly = Rb; io.x = J.x; io.lx = J.lx; io.off = off;
  const long long cstride = (long long)J.ly + 256;
  if (wid == 0) io.ein = m > 0 ? colck + J.colck_off + (long long)(m - 1) * cstride + bR : nullptr;
  else io.ein = slots + ((long long)t * EW + wid - 1) * slot_stride;
  io.eout = has_out ? slots + ((long long)t * EW + wid) * slot_stride : nullptr;
  io.mv = mv + (long long)t * job_words + (long long)wid * ((long long)(Rb + 63) * 64);
  io.mtop = nullptr;
  io.ref_left = 1;
  io.top = b > 0 ? rowck + J.rowck_off + (long long)(b - 1) * J.lx : nullptr;
  const int wcols = min(ODBA_W, J.lx - off);
  const int C = (s == S - 1) ? last_swath_c(wcols) : 16;
  (void)swath_dispatch_pipe<float, false, false, true, StripPipe>(C, io, &win, lane, pp);
}

// ---------------------------------------------------------------------------
// Stripe mode (dba.hpp:523-548, 646-763). Forward: every swath's right edge
// column goes to its own slot (slots + k*H), no step codes. Backward: the
// host steps m = 0, 1, ...; job j recomputes its stripe k = K_j-1-m with the
// height cut to the current traceback row (state.h), exactly as the
// reference relaunches DTWDistance with first_seq_length = backtrace rows.
// ---------------------------------------------------------------------------
template <typename T> struct StripeJob {
  const T* y;          // Y (rows), padded
  const T* x;          // X (columns), padded
  int H, LX;           // full height, width
  long long slot_off;  // element offset of this job's edge slots (stride H + 256)
  long long mv_off;    // word offset of this job's stripe step words
  long long top_off;   // word offset of this job's 64 top-row words
};

template <typename T, bool OS, bool OE, bool MINFIRST>
__global__ void __launch_bounds__(256) k_stripe_forward(const StripeJob<T>* __restrict__ jobs, int njobs,
                                                        unsigned int* __restrict__ counter, T* __restrict__ slots,
                                                        StripeState* __restrict__ st) {
  __shared__ T s_y[4][128];
  __shared__ T s_e[4][64];
  __shared__ T s_o[4][128];
  const int lane = threadIdx.x & 63;
  const int wid = uniform_i32((int)(threadIdx.x >> 6));
  const WaveWin<T> win{s_y[wid], s_e[wid], s_o[wid]};
  for (;;) {
    const unsigned int j = (unsigned int)uniform_i32((int)atomicAdd(counter, lane == 0 ? 1u : 0u));
    if ((int)j >= njobs) break;
    const StripeJob<T> jb = jobs[j];
    T* sl = slots + jb.slot_off;
    const int K = (jb.LX + ODBA_W - 1) / ODBA_W;
    int ab = -1;
    T colmin = T(0), top = T(0);
    for (int k = 0; k < K - 1; k++) {  // the last swath's edge is never read back
      const int off = k * ODBA_W;
      if (k > 0 && OE && !OS && off > jb.H && top == colmin) { ab = k; break; }
      SwathIO<T> io;
      io.y = jb.y; io.ly = jb.H; io.x = jb.x; io.lx = jb.LX; io.off = off;
      io.ein = k > 0 ? sl + (long long)(k - 1) * (jb.H + 256) : nullptr;
      io.eout = sl + (long long)k * (jb.H + 256);
      io.mv = nullptr; io.mtop = nullptr; io.ref_left = 1;
      SwathRes<T> r = run_swath<T, 16, OS, OE, MINFIRST, false>(io, &win, lane);
      colmin = r.colmin; top = r.top;
    }
    if (ab < 0 && K > 1) {
      const int off = (K - 1) * ODBA_W;
      if (OE && !OS && off > jb.H && top == colmin) ab = K - 1;
    }
    if (lane == 0) {
      StripeState z;
      z.i = jb.H - 1; z.j = jb.LX - 1; z.h = jb.H; z.curc = -1; z.lo = 1; z.hi = 0; z.lo2 = 1; z.hi2 = 0; z.gap = 0; z.npath = 0;
      z.fab = ab; z.sab = 0; z.done = 0; z.err = 0; z.br = 0;
      st[j] = z;
    }
  }
}

// One stripe per job: recompute with the truncated height (or mark the
// stripe open-right when the reference's abrogation test fires at that
// height), step codes into the job's stripe buffer.
template <typename T, bool OS, bool OE, bool MINFIRST>
__global__ void __launch_bounds__(256) k_stripe_step(const StripeJob<T>* __restrict__ jobs, int njobs, int m,
                                                     const T* __restrict__ slots, uint32_t* __restrict__ mv,
                                                     uint32_t* __restrict__ mtop, StripeState* __restrict__ st) {
  __shared__ T s_y[4][128];
  __shared__ T s_e[4][64];
  __shared__ T s_o[4][128];
  const int lane = threadIdx.x & 63;
  const int wid = uniform_i32((int)(threadIdx.x >> 6));
  const int j = blockIdx.x * 4 + wid;
  if (j >= njobs) return;
  const WaveWin<T> win{s_y[wid], s_e[wid], s_o[wid]};
  const StripeJob<T> jb = jobs[j];
  StripeState S = st[j];
  const int K = (jb.LX + ODBA_W - 1) / ODBA_W;
  const int k = K - 1 - m;
  if (k < 0 || S.done || S.err) return;
  const int h = S.h;
  if (S.fab >= 0 && k >= S.fab) {  // already abandoned in the forward pass
    if (lane == 0) st[j].sab = 1;
    return;
  }
  const int off = k * ODBA_W;
  const T* prev = k > 0 ? slots + jb.slot_off + (long long)(k - 1) * (jb.H + 256) : nullptr;
  int sab = 0;
  if (OE && !OS && off > h) {  // dtw.hpp:74-129 at the truncated height
    const T MAXV = Ar<T>::maxv();
    if (prev[0] == MAXV) sab = 1;
    else {
      T mn = MAXV;
      for (int r = lane; r < h; r += 64) mn = Ar<T>::mn(mn, prev[r]);
      mn = wave_min<T>(mn);
      if (prev[h - 1] == mn) sab = 1;
    }
  }
  if (!sab) {
    SwathIO<T> io;
    io.y = jb.y; io.ly = h; io.x = jb.x; io.lx = jb.LX; io.off = off;
    io.ein = prev; io.eout = nullptr;
    io.mv = mv + jb.mv_off; io.mtop = mtop + jb.top_off; io.ref_left = 1;
    const int wcols = (k == K - 1) ? jb.LX - off : ODBA_W;
    const int C = (k == K - 1) ? last_swath_c(wcols) : 16;
    (void)swath_dispatch<T, OS, OE, MINFIRST, true>(C, io, &win, lane);
  }
  if (lane == 0) st[j].sab = sab;
}

// ---------------------------------------------------------------------------
// Checkpointed stripe paths in the open modes (float). The reference's stripe
// mode (dba.hpp:646-763) recomputes each 1024-column stripe from row 0 up to
// the current traceback row h-1, the open-end top-row rule applied at row
// h-1 (the recompute's last row), and walks the stripe from its top-right
// cell until the path leaves it on the left. The rows below h-1 of such a
// recompute are the forward pass's rows (same recurrence, same left edge
// column), so only the band of rows that holds the walk's current cell needs
// recomputing, from the cost row stored below it:
//   * k_sck_forward: the whole matrix once, as column blocks of 16 swaths
//     claimed over the device, every swath's right edge column kept whole
//     (the reference's dtwCostSoFar; each swath waits on its left
//     neighbour's progress word), the cost row at the bottom of every band of
//     R rows, and each swath's (top, column min) for the forward pass's
//     open-end stop (dtw.hpp:74-129).
//   * k_sck_init: the edge columns' minima over every band start, the stop
//     swath of the forward pass and the walk states.
//   * k_sck_step: per job, the stop test at the truncated height when the
//     walk enters a stripe, else the block of rows bR .. i of the stripe
//     (bR the band start at or below the current row i) from the stored row
//     above it, the top-row rule only when i is the stripe's top row.
//   * k_sck_traceback: the stripe walk through that block, until the path
//     leaves the stripe (the next stripe's height is the row reached + 1) or
//     the block (the next step recomputes the band below).
// ---------------------------------------------------------------------------
template <bool OS, bool OE>
__global__ void __launch_bounds__(1024) k_sck_forward(const SckJob* __restrict__ jobs, int njobs, long long nclaims,
                                                      unsigned long long* __restrict__ counter,
                                                      float* __restrict__ slots, float* __restrict__ rowck,
                                                      float2* __restrict__ tc, int* __restrict__ cnt, int rshift,
                                                      int* __restrict__ err, long long limit) {
  constexpr int EW = kSckBandSwaths;
  static_assert(EW <= kMaxStripWaves, "one wave per swath of a claim");
  __shared__ float s_y[EW][128];
  __shared__ float s_e[EW][64];
  __shared__ float s_o[EW][128];
  __shared__ unsigned long long s_q;
  const int lane = threadIdx.x & 63;
  const int wid = uniform_i32((int)(threadIdx.x >> 6));
  const WaveWin<float> win{s_y[wid], s_e[wid], s_o[wid]};
  for (;;) {
    __syncthreads();  // every wave is done with the previous claim
    if (threadIdx.x == 0) s_q = atomicAdd(counter, 1ull);
    __syncthreads();
    const long long q = (long long)uniform_u64(s_q);
    if (q >= nclaims) break;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (jobs[mid].claim0 <= q) lo = mid;
      else hi = mid - 1;
    }
    const SckJob J = jobs[uniform_i32(lo)];
    const int m = uniform_i32((int)(q - J.claim0));
    const int S = (J.LX + ODBA_W - 1) / ODBA_W;
    const int s = m * EW + wid;
    if (s >= S) continue;
    const int off = s * ODBA_W;
    const bool has_out = s < S - 1;
    const long long sstride = (long long)J.H + 256;
    // Every hand-off, inside the block of 16 too, goes through the swath's
    // whole edge column and its global progress word: each wait targets the
    // left neighbour, a lower wave of this workgroup or a lower claim.
    BandPipe pp;
    pp.in_global = s > 0;
    pp.in_cnt = s > 0 ? &cnt[J.cnt_off + s - 1] : nullptr;
    pp.in_used = nullptr;
    pp.out_cnt = has_out ? &cnt[J.cnt_off + s] : nullptr;
    pp.out_global = has_out;
    pp.out_used = nullptr;
    pp.ring = 0;
    pp.main_end = (OE && J.H >= 2) ? J.H - 1 : J.H;
    pp.lane = lane;
    pp.err = err;
    pp.limit = limit;
    SwathIO<float> io;
    io.y = J.y; io.ly = J.H; io.x = J.x; io.lx = J.LX; io.off = off;
    io.ein = s == 0 ? nullptr : slots + J.slot_off + (long long)(s - 1) * sstride;
    io.eout = has_out ? slots + J.slot_off + (long long)s * sstride : nullptr;
    io.mv = nullptr; io.mtop = nullptr;
    io.ref_left = 1;
    io.rowck = rowck + J.rowck_off;
    io.ck_stride = J.LX;
    io.ck_rbase = 0;
    io.ck_rshift = rshift;
    const int nb = (J.H + (1 << rshift) - 1) >> rshift;
    io.ck_rlimit = (nb - 1) << rshift;
    // Every swath with 16 columns per lane, the job's last one too: its
    // columns past LX read the sequences' padding (1,024 values at least)
    // and influence no column before them; nothing reads them back (no
    // edge column, checkpoint rows stop at LX). One form of the sweep keeps
    // it inline.
    const SwathRes<float> r = run_swath<float, 16, OS, OE, true, false, BandPipe, true>(io, &win, lane, pp);
    if (has_out && lane == 0) tc[J.tc_off + s] = make_float2(r.top, r.colmin);
  }
}

// Grid (column groups, job groups), a wave per edge column k < K-1 of each
// job t = blockIdx.y, blockIdx.y + gridDim.y, ...: pm[k][b] = min of its rows
// 0 .. b*R-1 (MAXV for b = 0). The first wave of each job's first group
// then sets the job's stop swath and walk state, as the reference's forward
// pass leaves them.
template <bool OS, bool OE>
__global__ void __launch_bounds__(256) k_sck_init(const SckJob* __restrict__ jobs, int njobs,
                                                  const float* __restrict__ slots, const float2* __restrict__ tc,
                                                  float* __restrict__ pm, StripeState* __restrict__ st, int rshift) {
  const float MAXV = Ar<float>::maxv();
  const int lane = threadIdx.x & 63;
  const int wid = uniform_i32((int)(threadIdx.x >> 6));
  for (int t = blockIdx.y; t < njobs; t += gridDim.y) {
    const SckJob J = jobs[t];
    const int K = (J.LX + ODBA_W - 1) / ODBA_W;
    const int R = 1 << rshift;
    const int nb = (J.H + R - 1) >> rshift;
    const int k = blockIdx.x * 4 + wid;
    if (OE && !OS && k < K - 1) {
      const float* col = slots + J.slot_off + (long long)k * ((long long)J.H + 256);
      float* out = pm + J.pm_off + (long long)k * nb;
      float run = MAXV;
      if (lane == 0) out[0] = MAXV;
      for (int b = 1; b < nb; b++) {
        const int r1 = b * R;
        for (int r = (b - 1) * R + lane; r < r1; r += 64) run = Ar<float>::mn(run, col[r]);
        const float w = wave_min<float>(run);
        if (lane == 0) out[b] = w;
      }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      int ab = -1;
      if (OE && !OS) {
        for (int s = 1; s < K - 1; s++)
          if (s * ODBA_W > J.H && tc[J.tc_off + s - 1].x == tc[J.tc_off + s - 1].y) { ab = s; break; }
        if (ab < 0 && K > 1 && (K - 1) * ODBA_W > J.H && tc[J.tc_off + K - 2].x == tc[J.tc_off + K - 2].y) ab = K - 1;
      }
      StripeState z;
      z.i = J.H - 1; z.j = J.LX - 1; z.h = J.H; z.curc = -1; z.lo = 1; z.hi = 0; z.lo2 = 1; z.hi2 = 0; z.gap = 0; z.npath = 0;
      z.fab = ab; z.sab = 0; z.done = 0; z.err = 0; z.br = 0;
      st[t] = z;
    }
  }
}

// The block's sweep as four quarter strips (a wave each, 4 columns per
// lane), the last stripe too, so a stripe's step words always have the
// whole-swath layout (the quarters join their fragments into it).
template <bool OS, bool OE>
__device__ __forceinline__ void sck_block(const SwathIO<float>& io, const WaveWin<float>* W, int lane,
                                          const StripPipe& pp) {
  (void)run_swath<float, 4, OS, OE, true, true, StripPipe, false, true>(io, W, lane, pp);
}

// A workgroup of four waves per job: see the comment of k_sck_forward. Wave
// q sweeps columns [256q, 256q + 256) of the stripe; wave 0 reads the
// stored edge column of the left stripe (complete), waves 1-3 their left
// neighbour's edge through qslots (rows of the block, LDS progress words).
template <bool OS, bool OE>
__global__ void __launch_bounds__(256) k_sck_step(const SckJob* __restrict__ jobs, int njobs,
                                                  const float* __restrict__ slots, const float* __restrict__ rowck,
                                                  const float* __restrict__ pm, uint32_t* __restrict__ mv,
                                                  uint32_t* __restrict__ mtop, float* __restrict__ qslots,
                                                  long long qstride, StripeState* __restrict__ st, int rshift) {
  __shared__ float s_y[4][128];
  __shared__ float s_e[4][64];
  __shared__ float s_o[4][128];
  __shared__ int s_owner[5];
  __shared__ int s_cnt[5];
  const int lane = threadIdx.x & 63;
  const int q = uniform_i32((int)(threadIdx.x >> 6));
  const int t = blockIdx.x;
  if (threadIdx.x < 5) {
    s_owner[threadIdx.x] = threadIdx.x < 4 ? (int)threadIdx.x : -1;
    s_cnt[threadIdx.x] = threadIdx.x < 4 ? 0 : INT_MAX;  // the last one: wave 0's input, always ready
  }
  __syncthreads();
  const WaveWin<float> win{s_y[q], s_e[q], s_o[q]};
  const SckJob jb = jobs[t];
  const StripeState S = st[t];
  if (S.done || S.err || S.j < 0) return;
  const int k = uniform_i32(S.j) / ODBA_W;
  const int h = uniform_i32(S.h), i = uniform_i32(S.i);
  int sab = 0, br = 0;
  if (S.fab >= 0 && k >= S.fab) {
    sab = 1;  // abandoned in the forward pass: the stripe's top row is all OPEN_RIGHT
  } else {
    const int off = k * ODBA_W;
    const bool entry = (i == h - 1);  // the walk is on the stripe's top row
    const float* prev = k > 0 ? slots + jb.slot_off + (long long)(k - 1) * ((long long)jb.H + 256) : nullptr;
    if (OE && !OS && entry && off > h) {  // dtw.hpp:74-129 at the truncated height (every wave alike)
      const float MAXV = Ar<float>::maxv();
      if (prev[0] == MAXV) {
        sab = 1;
      } else {
        const int nb = (jb.H + (1 << rshift) - 1) >> rshift;
        const int bh = (h - 1) >> rshift;
        float mn = pm[jb.pm_off + (long long)(k - 1) * nb + bh];  // rows below bh*R
        for (int r = (bh << rshift) + lane; r < h; r += 64) mn = Ar<float>::mn(mn, prev[r]);
        mn = wave_min<float>(mn);
        if (prev[h - 1] == mn) sab = 1;
      }
    }
    if (!sab) {
      int b = i >> rshift;
      // A top row alone in its band takes the band below too: the top-row
      // rule reads the row under it.
      if (OE && entry && b > 0 && i == (b << rshift)) b--;
      const int bR = b << rshift;
      br = bR;
      StripPipe pp{&s_owner[q > 0 ? q - 1 : 4], &s_cnt[q > 0 ? q - 1 : 4], q > 0 ? q - 1 : -1,
                   q < 3 ? &s_cnt[q] : nullptr, lane};
      SwathIO<float> io;
      io.y = jb.y + bR; io.ly = i - bR + 1; io.x = jb.x; io.lx = jb.LX; io.off = off + q * 256;
      io.ein = q == 0 ? (prev ? prev + bR : nullptr) : qslots + ((long long)t * 3 + q - 1) * qstride;
      io.eout = q < 3 ? qslots + ((long long)t * 3 + q) * qstride : nullptr;
      io.mv = mv + jb.mv_off; io.mtop = mtop + (long long)t * 64;
      io.ref_left = (q == 0);
      io.lbase = q * 16;
      io.top = b > 0 ? rowck + jb.rowck_off + (long long)(b - 1) * jb.LX : nullptr;
      const bool os_b = OS && b == 0, oe_b = OE && entry;
      if (os_b && oe_b) sck_block<true, true>(io, &win, lane, pp);
      else if (os_b) sck_block<true, false>(io, &win, lane, pp);
      else if (oe_b) sck_block<false, true>(io, &win, lane, pp);
      else sck_block<false, false>(io, &win, lane, pp);
    }
  }
  if (q == 0 && lane == 0) {
    st[t].sab = sab;
    st[t].br = br;
  }
}

// ---------------------------------------------------------------------------
// Host-side launch helpers (instantiation dispatch).
// ---------------------------------------------------------------------------
template <typename T, bool MINFIRST>
static hipError_t launch_all_pairs_t(int os, int oe, dim3 grid, hipStream_t st, const T* data, const long long* offs,
                                     const int* lens, int n, long long npairs, unsigned long long* counter, T* dist,
                                     T* esc, long long stride, int* flags, const int2* order) {
  if (os && oe) hipLaunchKernelGGL((k_all_pairs<T, true, true, MINFIRST>), grid, dim3(256), 0, st, data, offs, lens, n, npairs, counter, dist, esc, stride, flags, order);
  else if (os) hipLaunchKernelGGL((k_all_pairs<T, true, false, MINFIRST>), grid, dim3(256), 0, st, data, offs, lens, n, npairs, counter, dist, esc, stride, flags, order);
  else if (oe) hipLaunchKernelGGL((k_all_pairs<T, false, true, MINFIRST>), grid, dim3(256), 0, st, data, offs, lens, n, npairs, counter, dist, esc, stride, flags, order);
  else hipLaunchKernelGGL((k_all_pairs<T, false, false, MINFIRST>), grid, dim3(256), 0, st, data, offs, lens, n, npairs, counter, dist, esc, stride, flags, order);
  return hipGetLastError();
}

template <typename T, bool MINFIRST>
static hipError_t launch_path_jobs_t(int os, int oe, dim3 grid, hipStream_t st, const void* jobs, int njobs,
                                     unsigned int* counter, uint32_t* mv, uint32_t* mtop, int* abr, T* esc,
                                     long long stride) {
  const PathJob<T>* J = (const PathJob<T>*)jobs;
  if (os && oe) hipLaunchKernelGGL((k_path_jobs<T, true, true, MINFIRST>), grid, dim3(256), 0, st, J, njobs, counter, mv, mtop, abr, esc, stride);
  else if (os) hipLaunchKernelGGL((k_path_jobs<T, true, false, MINFIRST>), grid, dim3(256), 0, st, J, njobs, counter, mv, mtop, abr, esc, stride);
  else if (oe) hipLaunchKernelGGL((k_path_jobs<T, false, true, MINFIRST>), grid, dim3(256), 0, st, J, njobs, counter, mv, mtop, abr, esc, stride);
  else hipLaunchKernelGGL((k_path_jobs<T, false, false, MINFIRST>), grid, dim3(256), 0, st, J, njobs, counter, mv, mtop, abr, esc, stride);
  return hipGetLastError();
}

hipError_t launch_all_pairs(int dtype, bool minfirst, int os, int oe, int nblocks, hipStream_t st, const void* data,
                            const long long* offs, const int* lens, int n, long long npairs,
                            unsigned long long* counter, void* dist, void* esc, long long stride, int* flags,
                            const void* order) {
  const int2* O = (const int2*)order;
  dim3 g(nblocks);
  switch (dtype) {
    case ODBA_F32: return launch_all_pairs_t<float, true>(os, oe, g, st, (const float*)data, offs, lens, n, npairs, counter, (float*)dist, (float*)esc, stride, flags, O);
    case ODBA_F64: return launch_all_pairs_t<double, true>(os, oe, g, st, (const double*)data, offs, lens, n, npairs, counter, (double*)dist, (double*)esc, stride, flags, O);
    case ODBA_I32:
      if (minfirst) return launch_all_pairs_t<int32_t, true>(os, oe, g, st, (const int32_t*)data, offs, lens, n, npairs, counter, (int32_t*)dist, (int32_t*)esc, stride, flags, O);
      return launch_all_pairs_t<int32_t, false>(os, oe, g, st, (const int32_t*)data, offs, lens, n, npairs, counter, (int32_t*)dist, (int32_t*)esc, stride, flags, O);
    case ODBA_U32:
      if (minfirst) return launch_all_pairs_t<uint32_t, true>(os, oe, g, st, (const uint32_t*)data, offs, lens, n, npairs, counter, (uint32_t*)dist, (uint32_t*)esc, stride, flags, O);
      return launch_all_pairs_t<uint32_t, false>(os, oe, g, st, (const uint32_t*)data, offs, lens, n, npairs, counter, (uint32_t*)dist, (uint32_t*)esc, stride, flags, O);
    case ODBA_U64:
      if (minfirst) return launch_all_pairs_t<uint64_t, true>(os, oe, g, st, (const uint64_t*)data, offs, lens, n, npairs, counter, (uint64_t*)dist, (uint64_t*)esc, stride, flags, O);
      return launch_all_pairs_t<uint64_t, false>(os, oe, g, st, (const uint64_t*)data, offs, lens, n, npairs, counter, (uint64_t*)dist, (uint64_t*)esc, stride, flags, O);
  }
  return hipErrorInvalidValue;
}

template <int C, int B>
static void launch_small_cb(int os, int oe, dim3 g, dim3 b, hipStream_t st, const float* data,
                            const SmallPairRec* tab, long long npairs, unsigned long long* counter, float* dist,
                            int* flags) {
#define ODBA_SMALL(A, Z) hipLaunchKernelGGL((k_all_pairs_small<C, B, A, Z>), g, b, 0, st, data, tab, npairs, counter, dist, flags)
  if (os && oe) ODBA_SMALL(true, true); else if (os) ODBA_SMALL(true, false); else if (oe) ODBA_SMALL(false, true); else ODBA_SMALL(false, false);
#undef ODBA_SMALL
}

hipError_t launch_all_pairs_small(int C, int B, int nw, int os, int oe, int nblocks, hipStream_t st,
                                  const float* data, const SmallPairRec* tab, long long npairs,
                                  unsigned long long* counter, float* dist, int* flags) {
  // every strip inside the first 1024 columns, which a padded sequence holds
  if (nw < 1 || nw > kSmallMaxWaves || 64 * C * nw > kSmallMaxLen || nblocks < 1) return hipErrorInvalidValue;
  if (!counter && nblocks < npairs) return hipErrorInvalidValue;  // without claims, a workgroup per pair
  const dim3 g(nblocks), b(64 * nw);
#define ODBA_SMALL_CB(CC, BB) launch_small_cb<CC, BB>(os, oe, g, b, st, data, tab, npairs, counter, dist, flags)
  if (B == 16) {
    switch (C) {
      case 1: ODBA_SMALL_CB(1, 16); break;
      case 2: ODBA_SMALL_CB(2, 16); break;
      case 4: ODBA_SMALL_CB(4, 16); break;
      default: return hipErrorInvalidValue;
    }
  } else if (B == 32) {
    switch (C) {
      case 1: ODBA_SMALL_CB(1, 32); break;
      case 2: ODBA_SMALL_CB(2, 32); break;
      case 4: ODBA_SMALL_CB(4, 32); break;
      default: return hipErrorInvalidValue;
    }
  } else {
    return hipErrorInvalidValue;
  }
#undef ODBA_SMALL_CB
  return hipGetLastError();
}

template <typename T, int C>
static void launch_bands_c(int os, int oe, dim3 g, dim3 b, hipStream_t st, const T* data, const long long* offs,
                           const int* lens, int n, const BandPair* tab, int ntab, long long nclaims,
                           unsigned long long* counter, T* dist, T* rings, int ring, T* gslots, long long gstride,
                           int* g_cnt, unsigned long long* g_free, int R, int* slot_of, int* abr, int* err,
                           long long limit) {
#define ODBA_BANDS(A, B) hipLaunchKernelGGL((k_all_pairs_bands<T, C, A, B>), g, b, 0, st, data, offs, lens, n, tab, ntab, nclaims, counter, dist, rings, ring, gslots, gstride, g_cnt, g_free, R, slot_of, abr, err, limit)
  if (os && oe) ODBA_BANDS(true, true); else if (os) ODBA_BANDS(true, false); else if (oe) ODBA_BANDS(false, true); else ODBA_BANDS(false, false);
#undef ODBA_BANDS
}

template <typename T>
static hipError_t launch_bands_t(int C, int os, int oe, dim3 g, dim3 b, hipStream_t st, const T* data,
                                 const long long* offs, const int* lens, int n, const BandPair* tab, int ntab,
                                 long long nclaims, unsigned long long* counter, T* dist, T* rings, int ring, T* gslots,
                                 long long gstride, int* g_cnt, unsigned long long* g_free, int R, int* slot_of,
                                 int* abr, int* err, long long limit) {
  switch (C) {
    case 1: launch_bands_c<T, 1>(os, oe, g, b, st, data, offs, lens, n, tab, ntab, nclaims, counter, dist, rings, ring, gslots, gstride, g_cnt, g_free, R, slot_of, abr, err, limit); break;
    case 4: launch_bands_c<T, 4>(os, oe, g, b, st, data, offs, lens, n, tab, ntab, nclaims, counter, dist, rings, ring, gslots, gstride, g_cnt, g_free, R, slot_of, abr, err, limit); break;
    case 16: launch_bands_c<T, 16>(os, oe, g, b, st, data, offs, lens, n, tab, ntab, nclaims, counter, dist, rings, ring, gslots, gstride, g_cnt, g_free, R, slot_of, abr, err, limit); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

size_t band_sync_bytes(int R, long long nclaims, int ntab) {
  return (size_t)R * 4 + (size_t)((R + 63) / 64) * 8 + (size_t)nclaims * 4 + (size_t)ntab * 4 + 64;
}

hipError_t launch_all_pairs_bands(int dtype, int C, int waves, int os, int oe, int nblocks, hipStream_t st,
                                  const void* data, const long long* offs, const int* lens, int n, const void* tab,
                                  int ntab, long long nclaims, unsigned long long* counter, void* dist, void* rings,
                                  int ring, void* gslots, long long gstride, void* sync, int R, int* err,
                                  long long tick_limit, int* flags) {
  if (waves < 1 || waves > kMaxStripWaves || ring < 128 || (ring & (ring - 1)) || R < 2 * nblocks + 2) return hipErrorInvalidValue;
  // sync: g_free (u64 words), g_cnt[R], slot_of[nclaims], abr[ntab]
  unsigned long long* g_free = (unsigned long long*)sync;
  int* g_cnt = (int*)(g_free + (R + 63) / 64);
  int* slot_of = g_cnt + R;
  int* abr = slot_of + nclaims;
  const long long ninit = std::max<long long>(std::max<long long>(R, nclaims), ntab);
  hipLaunchKernelGGL(k_band_init, dim3((unsigned)((ninit + 255) / 256)), dim3(256), 0, st, g_cnt, g_free, R, slot_of,
                     nclaims, abr, ntab);
  dim3 g(nblocks), b(64 * waves);
  const BandPair* T_ = (const BandPair*)tab;
  hipError_t e = hipErrorInvalidValue;
  switch (dtype) {
    case ODBA_F32: e = launch_bands_t<float>(C, os, oe, g, b, st, (const float*)data, offs, lens, n, T_, ntab, nclaims, counter, (float*)dist, (float*)rings, ring, (float*)gslots, gstride, g_cnt, g_free, R, slot_of, abr, err, tick_limit); break;
    case ODBA_I32: e = launch_bands_t<int32_t>(C, os, oe, g, b, st, (const int32_t*)data, offs, lens, n, T_, ntab, nclaims, counter, (int32_t*)dist, (int32_t*)rings, ring, (int32_t*)gslots, gstride, g_cnt, g_free, R, slot_of, abr, err, tick_limit); break;
    case ODBA_U32: e = launch_bands_t<uint32_t>(C, os, oe, g, b, st, (const uint32_t*)data, offs, lens, n, T_, ntab, nclaims, counter, (uint32_t*)dist, (uint32_t*)rings, ring, (uint32_t*)gslots, gstride, g_cnt, g_free, R, slot_of, abr, err, tick_limit); break;
  }
  if (e != hipSuccess) return e;
  if (flags) hipLaunchKernelGGL(k_band_flags, dim3((ntab + 255) / 256), dim3(256), 0, st, T_, ntab, n, abr, flags);
  return hipGetLastError();
}

size_t band_pair_size() { return sizeof(BandPair); }

size_t path_band_sync_bytes(int R, long long max_claims, int njobs) {
  // g_free, g_cnt, slot_of, then (64-aligned) band0 and band1, then bcnt
  return (size_t)((R + 63) / 64) * 8 + (size_t)R * 4 + (size_t)max_claims * 4 + 64 + (size_t)(njobs + 1) * 16 +
         (size_t)njobs * 4;
}

hipError_t launch_path_bands(int os, int oe, int nblocks, hipStream_t st, const void* jobs, int njobs,
                             long long max_claims, unsigned long long* counter, uint32_t* mv, uint32_t* mtop,
                             int* abr, void* rings, int ring, void* gslots, long long gstride, void* sync, int R,
                             int* err, long long tick_limit) {
  if (njobs < 1 || njobs > 1024 || ring < 128 || (ring & (ring - 1)) || R < 2 * nblocks + 2) return hipErrorInvalidValue;
  // sync: g_free (u64 words), g_cnt[R], slot_of[max_claims], band0[njobs + 1] (8-aligned)
  unsigned long long* g_free = (unsigned long long*)sync;
  int* g_cnt = (int*)(g_free + (R + 63) / 64);
  int* slot_of = g_cnt + R;
  long long* band0 = (long long*)(((uintptr_t)(slot_of + max_claims) + 63) & ~(uintptr_t)63);
  const PathJob<float>* J = (const PathJob<float>*)jobs;
  hipLaunchKernelGGL(k_path_band0, dim3(1), dim3(64), 0, st, J, njobs, band0, 1);
  const long long ninit = std::max<long long>(std::max<long long>(R, max_claims), njobs);
  hipLaunchKernelGGL(k_band_init, dim3((unsigned)((ninit + 255) / 256)), dim3(256), 0, st, g_cnt, g_free, R, slot_of,
                     max_claims, abr, njobs);
  dim3 g(nblocks), b(256);
  float* rg = (float*)rings;
  float* gs = (float*)gslots;
#define ODBA_PBANDS(A, B) hipLaunchKernelGGL((k_path_bands<A, B>), g, b, 0, st, J, njobs, band0, counter, mv, mtop, rg, ring, gs, gstride, g_cnt, g_free, R, slot_of, abr, err, tick_limit)
  if (os && oe) ODBA_PBANDS(true, true); else if (os) ODBA_PBANDS(true, false); else if (oe) ODBA_PBANDS(false, true); else ODBA_PBANDS(false, false);
#undef ODBA_PBANDS
  hipLaunchKernelGGL(k_abr_index, dim3((njobs + 255) / 256), dim3(256), 0, st, abr, njobs);
  return hipGetLastError();
}

hipError_t launch_sck_forward(int os, int oe, int nblocks, hipStream_t st, const SckJob* jobs, int njobs,
                              long long nclaims, unsigned long long* counter, float* slots, float* rowck, float2* tc,
                              int* cnt, int rshift, int* err, long long tick_limit) {
  if (njobs < 1 || nblocks < 1) return hipErrorInvalidValue;
  dim3 g(nblocks), b(64 * kSckBandSwaths);
#define ODBA_SCKF(A, B) hipLaunchKernelGGL((k_sck_forward<A, B>), g, b, 0, st, jobs, njobs, nclaims, counter, slots, rowck, tc, cnt, rshift, err, tick_limit)
  if (os && oe) ODBA_SCKF(true, true); else if (os) ODBA_SCKF(true, false); else if (oe) ODBA_SCKF(false, true); else ODBA_SCKF(false, false);
#undef ODBA_SCKF
  return hipGetLastError();
}

hipError_t launch_sck_init(int os, int oe, hipStream_t st, const SckJob* jobs, int njobs, int kmax, const float* slots,
                           const float2* tc, float* pm, StripeState* ss, int rshift) {
  if (njobs < 1) return hipSuccess;
  dim3 g((unsigned)std::max(1, (kmax - 1 + 3) / 4), (unsigned)std::min(njobs, 65535)), b(256);
#define ODBA_SCKI(A, B) hipLaunchKernelGGL((k_sck_init<A, B>), g, b, 0, st, jobs, njobs, slots, tc, pm, ss, rshift)
  if (os && oe) ODBA_SCKI(true, true); else if (os) ODBA_SCKI(true, false); else if (oe) ODBA_SCKI(false, true); else ODBA_SCKI(false, false);
#undef ODBA_SCKI
  return hipGetLastError();
}

hipError_t launch_sck_step(int os, int oe, hipStream_t st, const SckJob* jobs, int njobs, const float* slots,
                           const float* rowck, const float* pm, uint32_t* mv, uint32_t* mtop, float* qslots,
                           long long qstride, StripeState* ss, int rshift) {
  if (njobs < 1) return hipSuccess;
  dim3 g((unsigned)njobs), b(256);
#define ODBA_SCKS(A, B) hipLaunchKernelGGL((k_sck_step<A, B>), g, b, 0, st, jobs, njobs, slots, rowck, pm, mv, mtop, qslots, qstride, ss, rshift)
  if (os && oe) ODBA_SCKS(true, true); else if (os) ODBA_SCKS(true, false); else if (oe) ODBA_SCKS(false, true); else ODBA_SCKS(false, false);
#undef ODBA_SCKS
  return hipGetLastError();
}

hipError_t launch_path_swath_bands(int nw, int os, int oe, int nblocks, hipStream_t st, const void* jobs, int njobs,
                                   long long max_claims, unsigned long long* counter, uint32_t* mv, uint32_t* mtop,
                                   int* abr, void* rings, int ring, void* gslots, long long gstride, void* sync, int R,
                                   int* err, long long tick_limit, void* bslots) {
  const bool stop_rule = oe && !os;
  if (njobs < 1 || nw < 1 || nw > kPathStripWaves || ring < 128 || (ring & (ring - 1)) ||
      R < 2 * nblocks + 2 || (stop_rule && !bslots))
    return hipErrorInvalidValue;
  // sync: g_free (u64 words), g_cnt[R], slot_of[max_claims], band0[njobs + 1]
  // (64-aligned), band1[njobs + 1], bcnt[njobs]; bslots: njobs boundary
  // edge columns of gstride values (open end without open start only)
  unsigned long long* g_free = (unsigned long long*)sync;
  int* g_cnt = (int*)(g_free + (R + 63) / 64);
  int* slot_of = g_cnt + R;
  long long* band0 = (long long*)(((uintptr_t)(slot_of + max_claims) + 63) & ~(uintptr_t)63);
  long long* band1 = band0 + njobs + 1;
  int* bcnt = (int*)(band1 + njobs + 1);
  const PathJob<float>* J = (const PathJob<float>*)jobs;
  hipLaunchKernelGGL(k_path_band0, dim3(1), dim3(64), 0, st, J, njobs, band0, nw, band1, stop_rule ? 1 : 0);
  const long long ninit = std::max<long long>(std::max<long long>(R, max_claims), njobs);
  hipLaunchKernelGGL(k_band_init, dim3((unsigned)((ninit + 255) / 256)), dim3(256), 0, st, g_cnt, g_free, R, slot_of,
                     max_claims, abr, njobs, bcnt);
  dim3 g(nblocks), b(64 * nw);
  float* rg = (float*)rings;
  float* gs = (float*)gslots;
  float* bs = (float*)bslots;
#define ODBA_PSB(A, B) hipLaunchKernelGGL((k_path_swath_bands<A, B>), g, b, 0, st, J, njobs, band0, counter, mv, mtop, rg, ring, gs, gstride, g_cnt, g_free, R, slot_of, abr, err, tick_limit, band1, bs, bcnt)
  if (os && oe) ODBA_PSB(true, true); else if (os) ODBA_PSB(true, false); else if (oe) ODBA_PSB(false, true); else ODBA_PSB(false, false);
#undef ODBA_PSB
  hipLaunchKernelGGL(k_abr_index, dim3((njobs + 255) / 256), dim3(256), 0, st, abr, njobs);
  return hipGetLastError();
}

hipError_t launch_ckpt_forward(int ew, int nblocks, hipStream_t st, const void* jobs, int njobs, long long nclaims,
                               unsigned long long* counter, float* rowck, float* colck, int* cnt, float* rings,
                               int ring, int rshift, int* err, long long tick_limit) {
  if (ring < 128 || (ring & (ring - 1))) return hipErrorInvalidValue;
  const CkJob* J = (const CkJob*)jobs;
  const dim3 g(nblocks), b(64 * ew);
  switch (ew) {
    case 2: hipLaunchKernelGGL(k_ckpt_forward<2>, g, b, 0, st, J, njobs, nclaims, counter, rowck, colck, cnt, rings, ring, rshift, err, tick_limit); break;
    case 16: hipLaunchKernelGGL(k_ckpt_forward<16>, g, b, 0, st, J, njobs, nclaims, counter, rowck, colck, cnt, rings, ring, rshift, err, tick_limit); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_ckpt_block(int ew, hipStream_t st, const void* jobs, int njobs, void* states, const float* rowck,
                             const float* colck, uint32_t* mv, long long job_words, float* slots,
                             long long slot_stride, int rshift) {
  const CkJob* J = (const CkJob*)jobs;
  CkState* S = (CkState*)states;
  const dim3 g(njobs), b(64 * ew);
  switch (ew) {
    case 2: hipLaunchKernelGGL(k_ckpt_block<2>, g, b, 0, st, J, S, rowck, colck, mv, job_words, slots, slot_stride, rshift); break;
    case 16: hipLaunchKernelGGL(k_ckpt_block<16>, g, b, 0, st, J, S, rowck, colck, mv, job_words, slots, slot_stride, rshift); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

size_t ckpt_job_size() { return sizeof(CkJob); }
size_t ckpt_state_size() { return sizeof(CkState); }

hipError_t launch_path_strips(int waves, int os, int oe, int nblocks, hipStream_t st, const void* jobs, int njobs,
                              unsigned int* counter, uint32_t* mv, uint32_t* mtop, int* abr, void* esc,
                              long long stride, int quarters) {
  if (waves < 1 || waves > kPathStripWaves) return hipErrorInvalidValue;
  dim3 g(nblocks), b(64 * waves);
  const PathJob<float>* J = (const PathJob<float>*)jobs;
  float* E = (float*)esc;
#define ODBA_PSTRIPS(A, B, Q) hipLaunchKernelGGL((k_path_strips<float, A, B, Q>), g, b, 0, st, J, njobs, counter, mv, mtop, abr, E, stride)
  if (quarters) {
    if (os && oe) ODBA_PSTRIPS(true, true, true); else if (os) ODBA_PSTRIPS(true, false, true); else if (oe) ODBA_PSTRIPS(false, true, true); else ODBA_PSTRIPS(false, false, true);
  } else {
    if (os && oe) ODBA_PSTRIPS(true, true, false); else if (os) ODBA_PSTRIPS(true, false, false); else if (oe) ODBA_PSTRIPS(false, true, false); else ODBA_PSTRIPS(false, false, false);
  }
#undef ODBA_PSTRIPS
  return hipGetLastError();
}

hipError_t launch_path_jobs(int dtype, bool minfirst, int os, int oe, int nblocks, hipStream_t st, const void* jobs,
                            int njobs, unsigned int* counter, uint32_t* mv, uint32_t* mtop, int* abr, void* esc,
                            long long stride) {
  dim3 g(nblocks);
  switch (dtype) {
    case ODBA_F32: return launch_path_jobs_t<float, true>(os, oe, g, st, jobs, njobs, counter, mv, mtop, abr, (float*)esc, stride);
    case ODBA_F64: return launch_path_jobs_t<double, true>(os, oe, g, st, jobs, njobs, counter, mv, mtop, abr, (double*)esc, stride);
    case ODBA_I32: return launch_path_jobs_t<int32_t, false>(os, oe, g, st, jobs, njobs, counter, mv, mtop, abr, (int32_t*)esc, stride);
    case ODBA_U32: return launch_path_jobs_t<uint32_t, false>(os, oe, g, st, jobs, njobs, counter, mv, mtop, abr, (uint32_t*)esc, stride);
    case ODBA_U64: return launch_path_jobs_t<uint64_t, false>(os, oe, g, st, jobs, njobs, counter, mv, mtop, abr, (uint64_t*)esc, stride);
  }
  return hipErrorInvalidValue;
}

template <typename T, bool MINFIRST>
static hipError_t launch_stripe_t(bool forward, int os, int oe, int nblocks, hipStream_t st, const void* jobs, int njobs,
                                  int m, unsigned int* counter, void* slots, uint32_t* mv, uint32_t* mtop,
                                  StripeState* ss) {
  const StripeJob<T>* J = (const StripeJob<T>*)jobs;
  dim3 g(nblocks), b(256);
#define ODBA_SFWD(A, B) hipLaunchKernelGGL((k_stripe_forward<T, A, B, MINFIRST>), g, b, 0, st, J, njobs, counter, (T*)slots, ss)
#define ODBA_SSTEP(A, B) hipLaunchKernelGGL((k_stripe_step<T, A, B, MINFIRST>), g, b, 0, st, J, njobs, m, (const T*)slots, mv, mtop, ss)
  if (forward) {
    if (os && oe) ODBA_SFWD(true, true); else if (os) ODBA_SFWD(true, false); else if (oe) ODBA_SFWD(false, true); else ODBA_SFWD(false, false);
  } else {
    if (os && oe) ODBA_SSTEP(true, true); else if (os) ODBA_SSTEP(true, false); else if (oe) ODBA_SSTEP(false, true); else ODBA_SSTEP(false, false);
  }
#undef ODBA_SFWD
#undef ODBA_SSTEP
  return hipGetLastError();
}

hipError_t launch_stripe(int dtype, bool minfirst, bool forward, int os, int oe, int nblocks, hipStream_t st,
                         const void* jobs, int njobs, int m, unsigned int* counter, void* slots, uint32_t* mv,
                         uint32_t* mtop, StripeState* ss) {
  switch (dtype) {
    case ODBA_F32: return launch_stripe_t<float, true>(forward, os, oe, nblocks, st, jobs, njobs, m, counter, slots, mv, mtop, ss);
    case ODBA_F64: return launch_stripe_t<double, true>(forward, os, oe, nblocks, st, jobs, njobs, m, counter, slots, mv, mtop, ss);
    case ODBA_I32: return launch_stripe_t<int32_t, false>(forward, os, oe, nblocks, st, jobs, njobs, m, counter, slots, mv, mtop, ss);
    case ODBA_U32: return launch_stripe_t<uint32_t, false>(forward, os, oe, nblocks, st, jobs, njobs, m, counter, slots, mv, mtop, ss);
    case ODBA_U64: return launch_stripe_t<uint64_t, false>(forward, os, oe, nblocks, st, jobs, njobs, m, counter, slots, mv, mtop, ss);
  }
  return hipErrorInvalidValue;
}

size_t path_job_size(int dtype) {
  switch (dtype) {
    case ODBA_F32: return sizeof(PathJob<float>);
    case ODBA_F64: return sizeof(PathJob<double>);
    case ODBA_I32: return sizeof(PathJob<int32_t>);
    case ODBA_U32: return sizeof(PathJob<uint32_t>);
    case ODBA_U64: return sizeof(PathJob<uint64_t>);
  }
  return 0;
}

}  // namespace odba
