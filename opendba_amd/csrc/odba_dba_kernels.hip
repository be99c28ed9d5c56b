// DBA update kernels (traceback, ordered accumulation, mean/delta) and the
// z-normalisation kernel for gfx950.
//
// The reference accumulates every sequence's traceback into the centroid
// sums one sequence after the other, each in traceback order
// (updateCentroid, dba.hpp:374-423, serialised per sequence at
// dba.hpp:641-643). Along a monotone path the cells that feed one centroid
// position are visited in descending sequence-index order and form at most
// two contiguous counted runs: a block, then (only when the position is a
// row that became the top row of a height-cut stripe, whose RIGHT steps are
// OPEN_RIGHT and not counted) the single cell that leaves it. The traceback
// kernel emits (lo1, hi1, lo2, hi2) per (sequence, centroid position); the
// accumulation kernel then walks, for each centroid position in parallel,
// the sequences in cluster order and each run from hi down to lo. The float additions happen in exactly the
// reference's order, so the sums are bit-identical, with no atomics.

#include <hip/hip_runtime.h>
#include <float.h>
#include <limits.h>
#include <stdint.h>

#include "odba_dtw.hpp"

namespace odba {

struct Geo {
  int ly, lx, os, oe, ab, K, lastC, r0;
  long long S64;  // words per swath
};

__device__ __forceinline__ Geo make_geo(int ly, int lx, int os, int oe, int ab) {
  Geo g;
  g.ly = ly; g.lx = lx; g.os = os; g.oe = oe; g.ab = ab;
  g.K = (lx + ODBA_W - 1) / ODBA_W;
  int wl = lx - (g.K - 1) * ODBA_W;
  g.lastC = wl <= 64 ? 1 : wl <= 128 ? 2 : wl <= 256 ? 4 : wl <= 512 ? 8 : wl <= 768 ? 12 : 16;
  g.r0 = os ? 1 : 0;
  int nr = ((oe && ly >= 2) ? ly - 2 : ly - 1) - g.r0 + 1;
  g.S64 = nr > 0 ? (long long)(nr + 63) * 64 : 0;
  return g;
}


// Reference step code of cell (i, j) (dtw.hpp:10-16). Cells the reference
// never writes (below the top row of an abrogated swath) come back as 0.
__device__ __forceinline__ uint8_t step_code(const Geo& g, const uint32_t* __restrict__ mv,
                                             const uint32_t* __restrict__ mtop, int i, int j, long long& caddr,
                                             uint32_t& cword) {
  if (i == 0 && j == 0) return g.os ? ODBA_NIL_OPEN_RIGHT : ODBA_NIL;
  if (g.ab >= 0 && j >= g.ab * ODBA_W) return (i == g.ly - 1) ? ODBA_OPEN_RIGHT : 0;
  if (i == 0) return g.os ? ODBA_OPEN_RIGHT : ODBA_RIGHT;
  const int k = j / ODBA_W, t = j - k * ODBA_W;
  const int C = (k == g.K - 1) ? g.lastC : 16;
  const int lane = t / C, q = t - lane * C;
  if (g.oe && g.ly >= 2 && i == g.ly - 1) {
    uint32_t w = mtop[(long long)k * 64 + lane];
    uint32_t c = (w >> (2 * q)) & 3u;
    return c == 0 ? ODBA_DIAGONAL : (c == 2 ? ODBA_UP : ODBA_OPEN_RIGHT);
  }
  const long long s = (long long)(i - g.r0) + lane;
  const long long addr = (long long)k * g.S64 + s * 64 + lane;
  uint32_t w;
  if (addr == caddr) w = cword;
  else { w = mv[addr]; caddr = addr; cword = w; }
  const uint32_t c = plane_code(w, q);
  return c == 0 ? ODBA_DIAGONAL : (c == 2 ? ODBA_UP : ODBA_RIGHT);
}

// Counted runs of one centroid position: block 1, then after an uncounted
// gap block 2 (empty runs have lo > hi). add() returns false when a third
// counted block would start, which a monotone path cannot produce.
struct RunRec {
  int lo1, hi1, lo2, hi2, gap;
  __device__ void reset() { lo1 = 1; hi1 = 0; lo2 = 1; hi2 = 0; gap = 0; }
  __device__ bool add(bool counted, int sidx) {
    if (!counted) { if (hi1 >= lo1) gap = 1; return true; }
    if (!gap) {
      if (hi1 < lo1) { hi1 = sidx; lo1 = sidx; } else lo1 = sidx;
    } else {
      if (hi2 < lo2) { hi2 = sidx; lo2 = sidx; }
      else if (lo2 == sidx + 1) lo2 = sidx;
      else return false;
    }
    return true;
  }
  __device__ int4 out() const { return make_int4(lo1, hi1, lo2, hi2); }
};

__device__ __forceinline__ int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// The traceback walk and the ordered accumulation are single dependent
// chains per wave (one path, one position's additions in order) and use few
// waves. When other jobs' path kernels share the SIMDs (several DBA jobs
// served at once), those waves would get a fraction of the issue slots and
// each chain would stretch accordingly; raising the wave's issue priority
// inside the SIMD lets the chain issue first. Throughput kernels keep the
// default priority, and results do not depend on it.
__device__ __forceinline__ void chain_priority() { __builtin_amdgcn_s_setprio(3); }

// s_waitcnt immediate for vmcnt(0) with expcnt and lgkmcnt left at their
// maxima (gfx9 encoding: vmcnt[3:0] | expcnt[6:4] | lgkmcnt[11:8] | vmcnt_hi[15:14]).
constexpr int kWaitVmcnt0 = 0x0F70;

// Wave-cooperative cache of path words for a traceback that walks up and to
// the left: lane t holds the words of row i0 - t for the word-lanes l0,
// l0-1, l0-2, l0-3 of swath k0 (64 rows x 4*C columns). The walk reads a
// word with readlane; a tile is refilled only when the path leaves it, so a
// path costs one round of loads per 64 rows or 4*C columns instead of one
// dependent load per row.
struct WordTile {
  uint32_t w0, w1, w2, w3;
  int i0, l0, k0;
  __device__ void init() { i0 = -1; l0 = -1; k0 = -1; w0 = w1 = w2 = w3 = 0; }
  // word of main-region cell row i, word-lane l of swath k; step s = row - r0 + l
  __device__ uint32_t get(const uint32_t* __restrict__ M, long long S64, int r0, int k, int i, int l, int lane) {
    if (!(k == k0 && i <= i0 && i > i0 - 64 && l <= l0 && l > l0 - 4)) {
      i0 = i; l0 = l; k0 = k;
      const int row = i - lane;
      const long long base = (long long)k * S64;
      w0 = (row >= r0) ? M[base + (long long)(row - r0 + l) * 64 + l] : 0u;
      w1 = (row >= r0 && l >= 1) ? M[base + (long long)(row - r0 + l - 1) * 64 + (l - 1)] : 0u;
      w2 = (row >= r0 && l >= 2) ? M[base + (long long)(row - r0 + l - 2) * 64 + (l - 2)] : 0u;
      w3 = (row >= r0 && l >= 3) ? M[base + (long long)(row - r0 + l - 3) * 64 + (l - 3)] : 0u;
      // Wait for the refill here: otherwise the wait lands after the branch,
      // on every step, and since vmcnt also counts stores every step would
      // wait for the previous steps' run-record stores to complete.
      __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
    }
    const int dl = l0 - l;
    const uint32_t sel = dl == 0 ? w0 : (dl == 1 ? w1 : (dl == 2 ? w2 : w3));
    return (uint32_t)__builtin_amdgcn_readlane((int)sel, i0 - i);
  }
};

// Column j of the aligned matrix as (swath k, word-lane l, column q inside
// the lane's word), stepped left without divisions. Every swath but the
// last is full (16 columns per lane).
struct ColCursor {
  int k, l, q, C;
  __device__ void set(int j, int K, int lastC) {
    k = j >> 10;
    const int t = j & (ODBA_W - 1);
    C = (k == K - 1) ? lastC : 16;
    l = t / C;
    q = t - l * C;
  }
  __device__ void left() {
    if (q > 0) { q--; return; }
    if (l > 0) { l--; q = C - 1; return; }
    k--; C = 16; l = 63; q = 15;
  }
};

// Even bits of w packed down (bit 2q -> bit q): the low or high bits of 16
// two-bit step codes (open-end top-row words) as a 16-bit mask.
__device__ __forceinline__ uint32_t even_bits(uint32_t x) {
  x &= 0x55555555u;
  x = (x | (x >> 1)) & 0x33333333u;
  x = (x | (x >> 2)) & 0x0F0F0F0Fu;
  x = (x | (x >> 4)) & 0x00FF00FFu;
  x = (x | (x >> 8)) & 0x0000FFFFu;
  return x;
}

__device__ __forceinline__ uint64_t readlane64(uint64_t v, int l) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}

// Move masks of a 64-row tile for the run-skipping walk: lane t holds row
// i0 - t; bit x of rm / um is set when the cell in tile column x (absolute
// column cb + x, word-lanes l-3..l of one swath, 4*C columns) steps RIGHT /
// UP; a cell with neither bit steps DIAG. Columns below xmin (word-lanes of
// the previous swath) and lanes past tmax (rows below the main region) are
// outside the tile.
struct MoveTile {
  uint64_t rm, um;          // per lane
  int i0, cb, xmin, tmax;   // wave-uniform
  __device__ void init() { rm = um = 0; i0 = -1; cb = 0; xmin = 0; tmax = -1; }
  __device__ bool covers(int i, int j) const {
    const int t = i0 - i, x = j - cb;
    return t >= 0 && t <= tmax && x >= xmin && x < 64;
  }
  // Place cell (i, j) at lane 0 in the tile's rightmost word-lane.
  __device__ void fill(const Geo& g, const uint32_t* __restrict__ M, int i, int j, int rmin, int lane) {
    const int k = j >> 10, tt = j & (ODBA_W - 1);
    const int C = (k == g.K - 1) ? g.lastC : 16;
    const int l = tt / C, q = tt - l * C;
    const int row = i - lane;
    const uint32_t* base = M + (long long)k * g.S64;
    uint32_t w[4];
#pragma unroll
    for (int d = 0; d < 4; d++) {
      const int lw = l - 3 + d;
      w[d] = (row >= g.r0 && lw >= 0) ? base[(long long)(row - g.r0 + lw) * 64 + lw] : 0u;
    }
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
    const uint32_t cm = (1u << C) - 1u;
    uint64_t r = 0, u = 0;
#pragma unroll
    for (int d = 0; d < 4; d++) {
      // reversed, column q's bits land at 16 + q (not DIAG) and q (UP on a tie)
      const uint32_t rv = __builtin_bitreverse32(w[d]);
      const uint32_t nd = (rv >> 16) & cm, up = rv & cm;
      r |= (uint64_t)(nd & ~up) << (d * C);  // steps RIGHT
      u |= (uint64_t)(nd & up) << (d * C);   // steps UP
    }
    rm = r; um = u;
    i0 = i; cb = j - 3 * C - q;
    xmin = l >= 3 ? 0 : (3 - l) * C;
    tmax = min(63, i - rmin);
  }
};

// First column at or left of jj whose open-end top-row code is DIAG or UP
// (not OPEN_RIGHT), or -1: lane l of swath k holds the codes of columns
// k*W + l*C .. + C-1, so one ballot per swath finds the word and a bit scan
// the column.
__device__ int top_run_end(const Geo& g, const uint32_t* __restrict__ TP, int jj, int lane) {
  int k = jj >> 10;
  const int tt = jj & (ODBA_W - 1);
  int C = (k == g.K - 1) ? g.lastC : 16;
  int l = tt / C, q = tt - l * C;
  for (; k >= 0; k--) {
    const uint32_t w = TP[(long long)k * 64 + lane];
    uint32_t nr = ~even_bits(w) & ((1u << C) - 1u);  // codes 0 (DIAG) and 2 (UP)
    if (lane > l) nr = 0u;
    else if (lane == l) nr &= (2u << q) - 1u;
    const uint64_t b = (uint64_t)__ballot(nr != 0u);
    if (b) {
      const int lh = 63 - __builtin_clzll(b);
      const uint32_t m = (uint32_t)__builtin_amdgcn_readlane((int)nr, lh);
      return k * ODBA_W + lh * C + (31 - __builtin_clz(m));
    }
    C = 16; l = 63; q = 15;
  }
  return -1;
}

// One wavefront per aligned sequence (all lanes walk the same path; lane 0
// stores). runs[seq][c] = counted runs of sequence indices for centroid
// position c. With a path buffer and J.path_off >= 0 every visited cell is
// also recorded in traceback order (the cells and codes writeDTWPath prints,
// io_utils.hpp:285-323), the count first. Cells of the main region (rows
// with stored step words) take a short path; row 0, the open-end top row,
// abrogated swaths and the corner are decided positionally.
template <bool REC>
__global__ void __launch_bounds__(256) k_traceback(const TbJob* __restrict__ jobs, int njobs,
                                                   const uint32_t* __restrict__ mv, const uint32_t* __restrict__ mtop,
                                                   const int* __restrict__ abr, int4* __restrict__ runs, int lc, int os,
                                                   int oe, int* __restrict__ err, unsigned long long* __restrict__ pbuf) {
  chain_priority();
  const int lane = threadIdx.x & 63;
  const int jid = wave_uniform((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
  if (jid >= njobs) return;
  const TbJob J = jobs[jid];
  const Geo g = make_geo(J.ly, J.lx, os, oe, abr[jid]);
  const uint32_t* M = mv + J.mv_off;
  const uint32_t* TP = mtop + J.top_off;
  int4* R = J.rp ? J.rp : runs + J.runs_off;
  unsigned long long* P = (REC && J.path_off >= 0) ? pbuf + J.path_off : nullptr;
  const bool flip = J.flip != 0;
  const int top = (g.oe && g.ly >= 2) ? g.ly - 1 : INT32_MAX;  // open-end top row (its own words)
  const int rmin = g.r0 > 1 ? g.r0 : 1;                        // rows below are positional
  const int abcol = g.ab >= 0 ? g.ab * ODBA_W : INT32_MAX;
  int i = wave_uniform(J.ly - 1), j = wave_uniform(J.lx - 1);
  ColCursor cc;
  cc.set(j, g.K, g.lastC);
  int curc = flip ? i : j;
  RunRec rr;
  rr.reset();
  WordTile tile;
  tile.init();
  MoveTile mt;
  mt.init();
  int topk = -1;
  uint32_t topw = 0;
  long long np = 0;
  int e = 0;
  for (;;) {
    // Main region: every cell has a stored word and is counted; the run of
    // the current centroid position closes when the path leaves its column
    // (or its row, when the centroid is on the rows).
    if (i >= rmin && i < top && j < abcol) {
      const int c = flip ? i : j;
      if (c != curc) {
        if (lane == 0) R[curc] = rr.out();
        curc = c;
        rr.reset();
      }
    }
    // Without a recorded path, the main region is walked a run at a time: a
    // run is a maximal stretch of one move (RIGHT along the row, UP down the
    // column, DIAG down the diagonal) inside a MoveTile, found with one mask
    // scan or ballot. A run either stays in the current centroid position
    // (UP on columns, RIGHT on rows), extending its counted block, or leaves
    // it after its first cell, and then each further cell is a position of
    // its own with a one-cell block; those records are stored one per lane.
    // The records are the ones the cell-by-cell walk below writes.
    if (!REC && rr.gap == 0) {
      while (i >= rmin && i < top && j < abcol) {
        i = wave_uniform(i); j = wave_uniform(j); curc = wave_uniform(curc);
        rr.lo1 = wave_uniform(rr.lo1); rr.hi1 = wave_uniform(rr.hi1);
        mt.i0 = wave_uniform(mt.i0); mt.cb = wave_uniform(mt.cb);
        mt.xmin = wave_uniform(mt.xmin); mt.tmax = wave_uniform(mt.tmax);
        if (!mt.covers(i, j)) mt.fill(g, M, i, j, rmin, lane);
        const int t = wave_uniform(mt.i0 - i), x = wave_uniform(j - mt.cb);
        const uint64_t rw = readlane64(mt.rm, t), uw = readlane64(mt.um, t);
        int L, di, dj;
        if ((rw >> x) & 1u) {  // RIGHT: scan the row's mask down to the first other move
          const uint64_t keep = (x >= 63 ? ~0ull : ((2ull << x) - 1ull)) & (~0ull << mt.xmin);
          const uint64_t stop = ~rw & keep;
          const int xs = stop ? 63 - __builtin_clzll(stop) : mt.xmin - 1;
          L = x - xs; di = 0; dj = 1;
        } else if ((uw >> x) & 1u) {  // UP: first lane below whose cell in column x is not UP
          const bool st = lane > mt.tmax || !((mt.um >> x) & 1u);
          const uint64_t s = (uint64_t)__ballot(st) & (~0ull << t);
          L = (s ? __builtin_ctzll(s) : 64) - t; di = 1; dj = 0;
        } else {  // DIAG: lane t + s holds the run's cell s at column x - s
          const int pos = x + t - lane;
          const bool dg = lane >= t && lane <= mt.tmax && pos >= mt.xmin && pos <= 63 &&
                          !(((mt.rm | mt.um) >> (pos & 63)) & 1u);
          const uint64_t s = (uint64_t)__ballot(!dg) & (~0ull << t);
          L = (s ? __builtin_ctzll(s) : 64) - t; di = 1; dj = 1;
        }
        L = wave_uniform(L);
        const int si = flip ? j : i;  // sequence index of the run's first cell
        if (flip ? di == 0 : dj == 0) {  // the whole run is in the current position
          if (rr.hi1 < rr.lo1) rr.hi1 = si;
          rr.lo1 = si - L + 1;
        } else {
          if (rr.hi1 < rr.lo1) rr.hi1 = si;
          rr.lo1 = si;
          const int ds = flip ? dj : di;  // sequence-index step along the run
          if (lane < L)
            R[curc - lane] = lane == 0 ? rr.out() : make_int4(si - lane * ds, si - lane * ds, 1, 0);
          rr.reset();
          curc -= L;
        }
        i -= L * di;
        j -= L * dj;
        if (j < 0) { e |= 1; break; }
      }
      if (e & 1) break;
      cc.set(j, g.K, g.lastC);
    }
    while (i >= rmin && i < top && j < abcol) {
      // The walk state is the same in every lane; saying so keeps it in
      // scalar registers and the loop's branches scalar.
      i = wave_uniform(i); j = wave_uniform(j); curc = wave_uniform(curc);
      cc.k = wave_uniform(cc.k); cc.l = wave_uniform(cc.l); cc.q = wave_uniform(cc.q); cc.C = wave_uniform(cc.C);
      tile.i0 = wave_uniform(tile.i0); tile.l0 = wave_uniform(tile.l0); tile.k0 = wave_uniform(tile.k0);
      rr.lo1 = wave_uniform(rr.lo1); rr.hi1 = wave_uniform(rr.hi1); rr.lo2 = wave_uniform(rr.lo2);
      rr.hi2 = wave_uniform(rr.hi2); rr.gap = wave_uniform(rr.gap);
      const uint32_t c2 = plane_code(tile.get(M, g.S64, g.r0, cc.k, i, cc.l, lane), cc.q);  // 0 DIAG 1 RIGHT 2 UP
      if (REC) {
        if (P && lane == 0) P[1 + np] = pack_step(i, j, c2 == 0 ? ODBA_DIAGONAL : (c2 == 2 ? ODBA_UP : ODBA_RIGHT));
        np++;
      }
      if (!rr.add(true, flip ? j : i)) e |= 4;
      const bool up = c2 != 1u, lf = c2 != 2u;
      if (flip ? up : lf) {  // leaving the centroid position
        if (lane == 0) R[curc] = rr.out();
        rr.reset();
        curc = flip ? i - 1 : j - 1;
      }
      i -= up ? 1 : 0;
      if (lf) { j--; cc.left(); }
      if (i < 0 || j < 0) { e |= 1; break; }
    }
    if (e & 1) break;
    // Without a recorded path, stretches of positional moves are skipped in
    // one go, with the records the cell-by-cell steps below would write:
    // the open-end top row's OPEN_RIGHT cells up to its first DIAG or UP,
    // and row 0 from column j to the corner.
    if (!REC && rr.gap == 0 && i == top && top != INT32_MAX) {
      const int js = top_run_end(g, TP, min(j, abcol - 1), lane);  // the run is columns j .. js + 1
      const int L = j - js;
      if (L > 0 && (flip || j != curc || rr.hi1 < rr.lo1)) {
        const int c = flip ? i : j;
        if (c != curc) {
          if (lane == 0) R[curc] = rr.out();
          curc = c;
          rr.reset();
        }
        if (flip && rr.hi1 >= rr.lo1) rr.gap = 1;
        if (!flip) {  // every column of the run is left holding only its uncounted cell
          for (int b = 0; b < L - 1; b += 64)
            if (b + lane < L - 1) R[j - b - lane] = make_int4(1, 0, 1, 0);
          curc = js + 1;
        }  // on the rows the uncounted cells leave the (empty) block of the top row as it is
        j = js;
        if (j < 0) { e |= 1; break; }
        cc.set(j, g.K, g.lastC);
      }
    } else if (!REC && rr.gap == 0 && i == 0 && j > 0 && j < abcol) {
      const bool cnt = !g.os;  // RIGHT is counted, OPEN_RIGHT is not (dba.hpp:395-404)
      const int c = flip ? 0 : j;
      if (c != curc) {
        if (lane == 0) R[curc] = rr.out();
        curc = c;
        rr.reset();
      }
      if (flip) {  // columns j .. 1 all feed row 0
        if (cnt) {
          if (rr.hi1 < rr.lo1) rr.hi1 = j;
          rr.lo1 = 1;
        } else if (rr.hi1 >= rr.lo1) {
          rr.gap = 1;
        }
      } else {  // column j closes with cell (0, j); columns j-1 .. 1 hold one cell each
        if (cnt) {
          if (rr.hi1 < rr.lo1) rr.hi1 = 0;
          rr.lo1 = 0;
        }
        const int4 one = cnt ? make_int4(0, 0, 1, 0) : make_int4(1, 0, 1, 0);
        const int4 first = rr.out();
        // per component: a select of whole int4 values by lane compiles to a
        // two-entry private array and a scratch load per store
        for (int b = 0; b < j; b += 64)
          if (b + lane < j) {
            const bool f = b + lane == 0;
            R[j - b - lane] = make_int4(f ? first.x : one.x, f ? first.y : one.y, f ? first.z : one.z,
                                        f ? first.w : one.w);
          }
        curc = 0;
        rr.reset();
      }
      j = 0;
      cc.set(0, g.K, g.lastC);
    }
    uint32_t code;
    if (i == 0 && j == 0) {
      code = g.os ? ODBA_NIL_OPEN_RIGHT : ODBA_NIL;
    } else if (j >= abcol) {
      code = (i == g.ly - 1) ? ODBA_OPEN_RIGHT : 0;
    } else if (i == 0) {
      code = g.os ? ODBA_OPEN_RIGHT : ODBA_RIGHT;
    } else {  // open-end top row
      if (cc.k != topk) {
        topk = cc.k;
        topw = TP[(long long)cc.k * 64 + lane];
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
      }
      const uint32_t c2 = ((uint32_t)__builtin_amdgcn_readlane((int)topw, cc.l) >> (2 * cc.q)) & 3u;
      code = c2 == 0 ? ODBA_DIAGONAL : (c2 == 2 ? ODBA_UP : ODBA_OPEN_RIGHT);
    }
    if (REC) {
      if (P && lane == 0) P[1 + np] = pack_step(i, j, code);
      np++;
    }
    const int c = flip ? i : j;
    if (c != curc) {
      if (lane == 0) R[curc] = rr.out();
      curc = c;
      rr.reset();
    }
    const bool corner = (code == ODBA_NIL || code == ODBA_NIL_OPEN_RIGHT);
    const bool counted = corner ? (code == ODBA_NIL) : (code != ODBA_OPEN_RIGHT);  // dba.hpp:395-415
    if (!rr.add(counted, flip ? j : i)) e |= 4;
    if (corner) break;
    if (code == ODBA_DIAGONAL) { i--; j--; cc.left(); }
    else if (code == ODBA_UP) { i--; }
    else if (code == ODBA_RIGHT || code == ODBA_OPEN_RIGHT) { j--; cc.left(); }
    else { e |= 1; break; }
    if (i < 0 || j < 0) { e |= 1; break; }
  }
  if (lane == 0) {
    R[curc] = rr.out();
    if (REC && P) P[0] = (unsigned long long)np;
    if (e) atomicOr(err, e);
  }
}

// ---------------------------------------------------------------------------
// Checkpointed global paths: the walk inside the block k_ckpt_block has just
// recomputed (rows bR .. bi, columns mE .. j0; bi and j0 are the state's
// cell at launch), cell by cell as k_traceback walks a full matrix, until
// the path leaves the block downwards or to the left (the state then holds
// the first cell outside it) or reaches the corner. Global mode: every cell
// is counted, row 0 is positional (RIGHT, then NIL at the corner).
// ---------------------------------------------------------------------------
struct CkJobD {  // layout of CkJob in the swath kernels
  const float* y;
  const float* x;
  int ly, lx;
  long long rowck_off, colck_off, cnt_off, claim0, runs_off, path_off;
};
struct CkStateD {  // layout of CkState
  int i, j, curc, lo1, hi1, lo2, hi2, gap;
  long long np;
  int done, err, bi, pad;
};

template <bool REC>
__global__ void __launch_bounds__(256) k_ckpt_trace(const CkJobD* __restrict__ jobs, int njobs,
                                                    CkStateD* __restrict__ st, const uint32_t* __restrict__ mv,
                                                    long long job_words, int4* __restrict__ runs, int rshift, int ew,
                                                    unsigned long long* __restrict__ pbuf) {
  const int lane = threadIdx.x & 63;
  const int t = wave_uniform((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
  if (t >= njobs) return;
  const CkJobD J = jobs[t];
  CkStateD S = st[t];
  if (wave_uniform(S.done | S.err)) return;
  int i = wave_uniform(S.i), j = wave_uniform(S.j);
  const int bR = (i >> rshift) << rshift;
  const int E = ew * ODBA_W;
  const int mE = (j / E) * E;
  const int Rb = i - bR + 1;
  const long long S64 = (long long)(Rb + 63) * 64;
  const int K = (J.lx + ODBA_W - 1) / ODBA_W;
  const int wl = J.lx - (K - 1) * ODBA_W;
  const int lastC = wl <= 64 ? 1 : wl <= 128 ? 2 : wl <= 256 ? 4 : wl <= 512 ? 8 : wl <= 768 ? 12 : 16;
  const uint32_t* M = mv + (long long)t * job_words;
  int4* R = runs + J.runs_off;
  unsigned long long* P = (REC && J.path_off >= 0) ? pbuf + J.path_off : nullptr;
  RunRec rr;
  rr.lo1 = S.lo1; rr.hi1 = S.hi1; rr.lo2 = S.lo2; rr.hi2 = S.hi2; rr.gap = S.gap;
  int curc = wave_uniform(S.curc);
  long long np = S.np;
  int e = 0, done = 0;
  ColCursor cc;
  cc.set(j, K, lastC);
  const int k0 = mE / ODBA_W;  // the block's first swath
  WordTile tile;
  tile.init();
  for (;;) {
    if (i == 0) {  // row 0 to the corner: RIGHT steps, then NIL, all counted
      for (; j >= 0; j--) {
        if (REC) {
          if (P && lane == 0) P[1 + np] = pack_step(0, j, j == 0 ? ODBA_NIL : ODBA_RIGHT);
          np++;
        }
        if (j != curc) {
          if (lane == 0) R[curc] = rr.out();
          curc = j;
          rr.reset();
        }
        if (!rr.add(true, 0)) e |= 4;
      }
      done = 1;
      break;
    }
    if (i < bR || j < mE) break;  // left the block: the next launch recomputes the next one
    i = wave_uniform(i); j = wave_uniform(j); curc = wave_uniform(curc);
    cc.k = wave_uniform(cc.k); cc.l = wave_uniform(cc.l); cc.q = wave_uniform(cc.q); cc.C = wave_uniform(cc.C);
    tile.i0 = wave_uniform(tile.i0); tile.l0 = wave_uniform(tile.l0); tile.k0 = wave_uniform(tile.k0);
    rr.lo1 = wave_uniform(rr.lo1); rr.hi1 = wave_uniform(rr.hi1); rr.lo2 = wave_uniform(rr.lo2);
    rr.hi2 = wave_uniform(rr.hi2); rr.gap = wave_uniform(rr.gap);
    if (j != curc) {
      if (lane == 0) R[curc] = rr.out();
      curc = j;
      rr.reset();
    }
    const uint32_t c2 = plane_code(tile.get(M, S64, 0, cc.k - k0, i - bR, cc.l, lane), cc.q);  // 0 DIAG 1 RIGHT 2 UP
    if (REC) {
      if (P && lane == 0) P[1 + np] = pack_step(i, j, c2 == 0 ? ODBA_DIAGONAL : (c2 == 2 ? ODBA_UP : ODBA_RIGHT));
      np++;
    }
    if (!rr.add(true, i)) e |= 4;
    const bool up = c2 != 1u, lf = c2 != 2u;
    if (lf) {
      if (lane == 0) R[curc] = rr.out();
      rr.reset();
      curc = j - 1;
    }
    i -= up ? 1 : 0;
    if (lf) { j--; cc.left(); }
    if (i < 0 || j < 0) { e |= 1; break; }
  }
  if (lane == 0) {
    if (done) R[curc] = rr.out();
    S.i = i; S.j = j; S.curc = curc;
    S.lo1 = rr.lo1; S.hi1 = rr.hi1; S.lo2 = rr.lo2; S.hi2 = rr.hi2; S.gap = rr.gap;
    S.np = np;
    S.done = done;
    S.err = e;
    st[t] = S;
    if (REC && P && done) P[0] = (unsigned long long)np;
  }
}

hipError_t launch_ckpt_trace(hipStream_t st, const void* jobs, int njobs, void* states, const uint32_t* mv,
                             long long job_words, int4* runs, int rshift, int ew, unsigned long long* pbuf) {
  if (njobs <= 0) return hipSuccess;
  const dim3 g((njobs + 3) / 4), b(256);
  if (pbuf)
    hipLaunchKernelGGL(k_ckpt_trace<true>, g, b, 0, st, (const CkJobD*)jobs, njobs, (CkStateD*)states, mv, job_words,
                       runs, rshift, ew, pbuf);
  else
    hipLaunchKernelGGL(k_ckpt_trace<false>, g, b, 0, st, (const CkJobD*)jobs, njobs, (CkStateD*)states, mv, job_words,
                       runs, rshift, ew, pbuf);
  return hipGetLastError();
}

// Stripe-mode traceback for stripe k = K-1-m of every job (updateCentroid
// with column_offset / stripe_rows, dba.hpp:374-423 launched at 733), one
// wavefront per job: walk from the carried cell until the path leaves the
// stripe on the left (the new height is the row reached + 1) or reaches the
// (0,0) anchor. BLK (checkpointed stripes): the stripe is the one holding
// the carried cell, its step words cover rows br .. of the block last
// recomputed, and a walk that goes below row br stops there (the next step
// recomputes the band below).
template <bool BLK>
__global__ void __launch_bounds__(256) k_stripe_traceback(const TbJob* __restrict__ jobs, int njobs, int m,
                                                          const uint32_t* __restrict__ mv,
                                                          const uint32_t* __restrict__ mtop,
                                                          StripeState* __restrict__ ss, int4* __restrict__ runs,
                                                          int lc, int os, int oe, int* __restrict__ err,
                                                          unsigned long long* __restrict__ pbuf) {
  const int lane = threadIdx.x & 63;
  const int jid = wave_uniform((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
  if (jid >= njobs) return;
  const TbJob J = jobs[jid];
  StripeState S = ss[jid];
  const int K = (J.lx + ODBA_W - 1) / ODBA_W;
  if (S.done || S.err || (BLK && S.j < 0)) return;
  const int k = BLK ? wave_uniform(S.j) / ODBA_W : K - 1 - m;
  if (k < 0) return;
  const int left = k * ODBA_W;
  const int h = S.h;
  const int wl = J.lx - left;
  // BLK: every stripe's words have the whole-swath layout (16 columns per lane)
  const int C = (k == K - 1 && !BLK) ? (wl <= 64 ? 1 : wl <= 128 ? 2 : wl <= 256 ? 4 : wl <= 512 ? 8 : wl <= 768 ? 12 : 16) : 16;
  const int br = (BLK && !S.sab) ? wave_uniform(S.br) : 0;  // the step words' first row
  const int r0 = (os && br == 0) ? 1 : 0;
  const uint32_t* M = mv + J.mv_off;
  const uint32_t* TP = mtop + J.top_off;
  int4* R = J.rp ? J.rp : runs + J.runs_off;
  unsigned long long* P = (pbuf && J.path_off >= 0) ? pbuf + J.path_off : nullptr;
  int i = wave_uniform(S.i), j = wave_uniform(S.j);
  RunRec rr;
  rr.lo1 = S.lo; rr.hi1 = S.hi; rr.lo2 = S.lo2; rr.hi2 = S.hi2; rr.gap = S.gap;
  if (S.curc < 0) { S.curc = J.flip ? i : j; rr.reset(); }
  WordTile tile;
  tile.init();
  const uint32_t topw = TP[lane];
  __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
  bool done = false;
  int e = 0;
  // word-lane l and column q of j inside this stripe, stepped left
  int l = 0, q = 0;
  if (j >= left) {
    l = (j - left) / C;
    q = (j - left) - l * C;
  }
  while (j >= left && i >= br) {
    uint8_t code;
    if (i == 0 && j == 0) code = os ? ODBA_NIL_OPEN_RIGHT : ODBA_NIL;
    else if (S.sab && i == h - 1) code = ODBA_OPEN_RIGHT;
    else if (i == 0) code = os ? ODBA_OPEN_RIGHT : ODBA_RIGHT;
    else {
      if (oe && h >= 2 && i == h - 1) {
        const uint32_t c = ((uint32_t)__builtin_amdgcn_readlane((int)topw, l) >> (2 * q)) & 3u;
        code = c == 0 ? ODBA_DIAGONAL : (c == 2 ? ODBA_UP : ODBA_OPEN_RIGHT);
      } else {
        const uint32_t c = plane_code(tile.get(M, 0, r0, 0, i - br, l, lane), q);
        code = c == 0 ? ODBA_DIAGONAL : (c == 2 ? ODBA_UP : ODBA_RIGHT);
      }
    }
    if (P && lane == 0) P[1 + S.npath] = pack_step(i, j, code);
    S.npath++;
    const int c = J.flip ? i : j;
    const int sidx = J.flip ? j : i;
    if (c != S.curc) {
      if (lane == 0) R[S.curc] = rr.out();
      S.curc = c;
      rr.reset();
    }
    const bool corner = (code == ODBA_NIL || code == ODBA_NIL_OPEN_RIGHT);
    const bool counted = corner ? (code == ODBA_NIL) : (code != ODBA_OPEN_RIGHT);
    if (!rr.add(counted, sidx)) e |= 4;
    if (corner) { done = true; break; }
    const bool left_move = (code != ODBA_UP);
    if (code == ODBA_DIAGONAL) { i--; j--; }
    else if (code == ODBA_UP) { i--; }
    else { j--; }
    if (left_move) {
      if (q > 0) q--;
      else { l--; q = C - 1; }
    }
    if (i < 0) { S.err = 1; e |= 1; break; }
  }
  if (done) {
    if (lane == 0) R[S.curc] = rr.out();
    S.done = 1;
    if (P && lane == 0) P[0] = (unsigned long long)S.npath;
  } else if (!S.err && j < left) {  // (BLK: else below the block, still in the stripe)
    if (j != left - 1) { S.err = 1; e |= 1; }
    S.h = i + 1;
  }
  S.i = i;
  S.j = j;
  S.lo = rr.lo1; S.hi = rr.hi1; S.lo2 = rr.lo2; S.hi2 = rr.hi2; S.gap = rr.gap;
  if (lane == 0) {
    ss[jid] = S;
    if (e) atomicOr(err, e);
  }
}

// a += p[hi], p[hi-1], ..., p[lo], in that order. The values are loaded
// eight at a time before they are added, so a long run costs one memory
// latency per eight values instead of one per value.
template <typename T, typename P = const T*>
__device__ __forceinline__ T add_run_desc(T a, P __restrict__ p, int hi, int lo) {
  for (int k = hi; k >= lo; k -= 8) {
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = (k - u >= lo) ? p[k - u] : T(0);
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (k - u >= lo) a = (T)(a + v[u]);
  }
  return a;
}

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

// A 4-byte value of type T from its bits.
template <typename T> __device__ __forceinline__ T bits_as(uint32_t b) { return __builtin_bit_cast(T, b); }

// Value of v in lane t (wave-uniform t), for 4- and 8-byte types.
template <typename T> __device__ __forceinline__ T bcast(T v, int t) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), t));
  } else {
    return __builtin_bit_cast(T, readlane64(__builtin_bit_cast(uint64_t, v), t));
  }
}

// Sum of one centroid position's runs over sequences 0..nseq-1 in order,
// each run from hi down to lo (runs[s * stride] is sequence s's record).
// Eight sequences at a time: while their first eight run values are in
// flight the records and offsets of the next eight are loaded, so a batch
// costs about one memory latency; only runs longer than eight values and
// second blocks load more. The additions stay in the reference's order.
template <typename T>
__device__ __forceinline__ void accumulate_runs(T& a, uint32_t& n, const int4* __restrict__ runs, long long stride,
                                                int nseq, const T* __restrict__ data,
                                                const long long* __restrict__ offs) {
  constexpr int B = 8, V = 8;
  // Every lane of the wave takes part in every batch (the long-run loads
  // need all 64), up to the wave's largest sequence count.
  int nmax = nseq;
  for (int o = 32; o > 0; o >>= 1) nmax = max(nmax, __shfl_xor(nmax, o));
  int4 rn[B];
  long long on[B];
#pragma unroll
  for (int u = 0; u < B; u++) {
    rn[u] = (u < nseq) ? runs[(long long)u * stride] : make_int4(1, 0, 1, 0);
    on[u] = (u < nseq) ? offs[u] : 0;
  }
  for (int s0 = 0; s0 < nmax; s0 += B) {
    int4 r[B];
    const T* p[B];
#pragma unroll
    for (int u = 0; u < B; u++) {
      r[u] = rn[u];
      p[u] = data + on[u];
    }
    T v[B][V];
#pragma unroll
    for (int u = 0; u < B; u++)
#pragma unroll
      for (int k = 0; k < V; k++) v[u][k] = (r[u].y - k >= r[u].x) ? p[u][r[u].y - k] : T(0);
#pragma unroll
    for (int u = 0; u < B; u++) {
      const int s = s0 + B + u;
      rn[u] = (s < nseq) ? runs[(long long)s * stride] : make_int4(1, 0, 1, 0);
      on[u] = (s < nseq) ? offs[s] : 0;
    }
#pragma unroll
    for (int u = 0; u < B; u++) {
      if (r[u].y >= r[u].x) {
#pragma unroll
        for (int k = 0; k < V; k++)
          if (r[u].y - k >= r[u].x) a = (T)(a + v[u][k]);
        if (r[u].y - V >= r[u].x) a = add_run_desc<T>(a, p[u], r[u].y - V, r[u].x);
      }
      if (r[u].y >= r[u].x) n += (uint32_t)(r[u].y - r[u].x + 1);
      if (r[u].w >= r[u].z) {
        a = add_run_desc<T>(a, p[u], r[u].w, r[u].z);
        n += (uint32_t)(r[u].w - r[u].z + 1);
      }
    }
  }
}

// Minimum / maximum over the 64 lanes as a wave-uniform value, by DPP
// within the VALU: pairs, quads, half rows and rows combined in registers
// (quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror, row_mirror),
// then row 0 into row 1 and row 2 into row 3 (row_bcast:15), rows 0-1 into
// rows 2-3 (row_bcast:31); lane 63 then holds the whole wave's result. A
// lane that a masked row broadcast does not write keeps its own value.
// Every lane of the wave must call it.
template <bool MAX> __device__ __forceinline__ int wave_reduce_i(int v) {
  auto op = [](int a, int b) { return MAX ? max(a, b) : min(a, b); };
  v = op(v, __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, false));
  v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, false));
  v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x141, 0xF, 0xF, false));
  v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x140, 0xF, 0xF, false));
  v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x142, 0xA, 0xF, false));
  v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x143, 0xC, 0xF, false));
  return __builtin_amdgcn_readlane(v, 63);
}
__device__ __forceinline__ int wave_min_i(int v) { return wave_reduce_i<false>(v); }
__device__ __forceinline__ int wave_max_i(int v) { return wave_reduce_i<true>(v); }
// The same by cross-lane shuffles (LDS permutes), the result in every lane.
__device__ __forceinline__ int wave_min_shfl(int v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ int wave_max_shfl(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

constexpr int kStageB = 8;    // sequences per batch of the staged accumulation
constexpr int kStageW = 256;  // values per staged window (4 per lane)
// windows loaded together per batch (4 values per lane each, in registers)
template <typename T> constexpr int stage_windows() { return sizeof(T) == 4 ? 32 : 16; }

// Sum of the runs of a wave's 64 consecutive centroid positions, one lane
// per position, over sequences 0..nall-1 of one cluster (runs[s * stride]
// is this lane's record of sequence s; lanes with nseq == 0 have none).
// Along a monotone path the positions' runs tile one contiguous range of
// each sequence, so the wave loads that range with all 64 lanes into LDS,
// 256 values per window, and each lane then adds its own run from LDS, hi
// down to lo, in sequence order: the reference's order of additions. The
// windows of a batch of eight sequences, in sequence order, are loaded
// together up to stage_windows<T>() of them (one memory latency per
// batch); a batch whose ranges need more loads the rest one window at a
// time, each one window ahead of the adds. buf holds (stage_windows<T>() + 1) * kStageW
// values; every lane of the block's one wave must call this. It is inlined
// into its kernels so that buf is known to be LDS (ds_read / ds_write, not
// flat accesses through a generic pointer).
// Global-memory pointers for data read through pointers held in memory
// (the multi-cluster slot table): without the address space the compiler
// must use flat accesses with 64-bit lane addresses.
template <typename T> using gptr = const T __attribute__((address_space(1)))*;

// DPPRED: the batch's range reductions by DPP (wave-uniform results) or by
// shuffles. Measured on gfx950, DPP took the single-cluster kernel from
// 4.67 to 4.37 ms on a 2,000-sequence round, and shuffles kept the
// multi-cluster kernel at 6.1 ms against 6.8 ms with DPP on a 19-job
// round, so each kernel takes its faster form (the sums are the same).
// VEC (4-byte types, buf 16-byte aligned): a run's values are read from
// LDS 16 bytes at a time. Measured on gfx950, that took the multi-cluster
// kernel from 6.15 to 5.65 ms on a 19-job step, but left the single-cluster
// kernel at 4.66-4.68 ms against 4.37 ms on a 2,000-sequence round, so
// only the multi-cluster kernel takes it.
template <typename T, typename DP = const T*, typename OP = const long long*, bool DPPRED = true, bool VEC = false>
__device__ __forceinline__ void accumulate_staged(T& a, uint32_t& n, const int4* __restrict__ runs, long long stride,
                                                  int nseq, int nall, DP __restrict__ data, OP __restrict__ offs,
                                                  T* __restrict__ buf) {
  constexpr int B = kStageB, SW = kStageW, NW = stage_windows<T>();
  const int lane = lane_id();
  int4 rn[B];
  long long on[B];
#pragma unroll
  for (int u = 0; u < B; u++) {
    rn[u] = (u < nseq) ? runs[(long long)u * stride] : make_int4(1, 0, 1, 0);
    on[u] = (u < nall) ? offs[u] : 0;
  }
  for (int s0 = 0; s0 < nall; s0 += B) {
    int4 r[B];
    long long o[B];
    int lo[B], hi[B], wb0[B], nw[B];
    int tot = 0;
#pragma unroll
    for (int u = 0; u < B; u++) {
      r[u] = rn[u];
      o[u] = on[u];
      const bool ne = r[u].y >= r[u].x;
      lo[u] = DPPRED ? wave_min_i(ne ? r[u].x : INT32_MAX) : wave_min_shfl(ne ? r[u].x : INT32_MAX);
      hi[u] = DPPRED ? wave_max_i(ne ? r[u].y : INT32_MIN) : wave_max_shfl(ne ? r[u].y : INT32_MIN);
      nw[u] = hi[u] >= lo[u] ? (hi[u] - lo[u]) / SW + 1 : 0;
      wb0[u] = tot;
      tot += nw[u];
    }
    // window k of the batch: sequence u with wb0[u] <= k < wb0[u] + nw[u],
    // values top, top-1, ... of that sequence (selected without indexing)
    auto load_window = [&](int k, T* x) {
      int ktop = -1, klo = 0;
      long long koff = 0;
#pragma unroll
      for (int u = 0; u < B; u++)
        if (k >= wb0[u] && k < wb0[u] + nw[u]) {
          ktop = hi[u] - (k - wb0[u]) * SW;
          klo = lo[u];
          koff = o[u];
        }
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int idx = ktop - (lane * 4 + e);
        x[e] = (k < tot && idx >= klo) ? data[koff + idx] : T(0);
      }
    };
    T w[NW][4];
#pragma unroll
    for (int k = 0; k < NW; k++) load_window(k, w[k]);
    // Windows past the first NW go through one more LDS window, each loaded
    // into registers one window ahead: the load of window k + 1 is in
    // flight while the lanes add window k.
    T xn[4];
    if (tot > NW) load_window(NW, xn);
#pragma unroll
    for (int u = 0; u < B; u++) {
      const int sn = s0 + B + u;
      rn[u] = (sn < nseq) ? runs[(long long)sn * stride] : make_int4(1, 0, 1, 0);
      on[u] = (sn < nall) ? offs[sn] : 0;
    }
    __syncthreads();  // the previous batch's reads of buf are done
#pragma unroll
    for (int k = 0; k < NW; k++)
#pragma unroll
      for (int e = 0; e < 4; e++) buf[k * SW + lane * 4 + e] = w[k][e];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < B; u++) {
      int top = hi[u];
      for (int wi = 0; wi < nw[u]; wi++) {  // wave-uniform
        const T* wb;                        // wb[t] = value top - t
        const int slot = wb0[u] + wi;
        if (slot < NW) {
          wb = buf + slot * SW;
        } else {  // past the batch's loaded windows: the one loaded ahead
          __syncthreads();  // every lane is done with the previous extra window
#pragma unroll
          for (int e = 0; e < 4; e++) buf[NW * SW + lane * 4 + e] = xn[e];
          __syncthreads();
          wb = buf + NW * SW;
          if (slot + 1 < tot) load_window(slot + 1, xn);
        }
        const int wlo = max(lo[u], top - SW + 1);
        if (r[u].y >= r[u].x) {
          // A run can be tens of thousands of values (a path along row 0 or
          // column 0), and its dependent adds are the kernel's longest
          // chain: the LDS reads of the next eight values are issued before
          // the adds of the current eight, so the chain does not wait on
          // LDS latency.
          const int k1 = min(r[u].y, top), k0 = max(r[u].x, wlo);
          const T* q = wb + (top - k1);
          int m = k1 - k0 + 1;
          if constexpr (VEC && sizeof(T) == 4) {
            {
              // Whole 16-byte LDS reads (buf is 16-byte aligned): single
              // values up to an aligned address, then eight values per step
              // (the next eight read before these are added), then the rest;
              // the additions keep their order.
              for (; m > 0 && ((uintptr_t)q & 15u); m--, q++) a = (T)(a + *q);
              const uint4* q4 = (const uint4*)q;
              const int n8 = m >> 3;
              auto add8 = [&](const uint4& c0, const uint4& c1) {
                a = (T)(a + bits_as<T>(c0.x)); a = (T)(a + bits_as<T>(c0.y));
                a = (T)(a + bits_as<T>(c0.z)); a = (T)(a + bits_as<T>(c0.w));
                a = (T)(a + bits_as<T>(c1.x)); a = (T)(a + bits_as<T>(c1.y));
                a = (T)(a + bits_as<T>(c1.z)); a = (T)(a + bits_as<T>(c1.w));
              };
              if (n8 > 0) {
                // x holds block t - 1; each block is read while the one
                // before it is added, the two register pairs alternating
                uint4 x0 = q4[0], x1 = q4[1];
                int t = 1;
                for (; t + 1 < n8; t += 2) {
                  const uint4 y0 = q4[2 * t], y1 = q4[2 * t + 1];
                  add8(x0, x1);
                  x0 = q4[2 * t + 2];
                  x1 = q4[2 * t + 3];
                  add8(y0, y1);
                }
                if (t < n8) {
                  const uint4 y0 = q4[2 * t], y1 = q4[2 * t + 1];
                  add8(x0, x1);
                  x0 = y0;
                  x1 = y1;
                }
                add8(x0, x1);
                q += 8 * n8;
                m -= 8 * n8;
              }
            }
          }
          if (m >= 8) {
            T v[8];
#pragma unroll
            for (int e = 0; e < 8; e++) v[e] = q[e];
            q += 8;
            m -= 8;
#pragma unroll 2
            for (; m >= 8; m -= 8, q += 8) {
              T nx[8];
#pragma unroll
              for (int e = 0; e < 8; e++) nx[e] = q[e];
#pragma unroll
              for (int e = 0; e < 8; e++) a = (T)(a + v[e]);
#pragma unroll
              for (int e = 0; e < 8; e++) v[e] = nx[e];
            }
#pragma unroll
            for (int e = 0; e < 8; e++) a = (T)(a + v[e]);
          }
          for (; m > 0; m--, q++) a = (T)(a + *q);
        }
        top = wlo - 1;
      }
      if (r[u].y >= r[u].x) n += (uint32_t)(r[u].y - r[u].x + 1);
      if (r[u].w >= r[u].z) {  // second block: one cell of a height-cut stripe row
        a = add_run_desc<T>(a, data + o[u], r[u].w, r[u].z);
        n += (uint32_t)(r[u].w - r[u].z + 1);
      }
    }
  }
}

// One 64-lane block per 64 centroid positions (no early exit: lanes past lc
// take part with no records).
template <typename T>
__global__ void __launch_bounds__(64) k_accumulate(const T* __restrict__ data, const long long* __restrict__ offs,
                                                   const int4* __restrict__ runs, int nseq, int lc,
                                                   T* __restrict__ acc, uint32_t* __restrict__ counts, int first) {
  __shared__ T buf[(stage_windows<T>() + 1) * kStageW];
  chain_priority();
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const bool on = c < lc;
  T a = (first || !on) ? T(0) : acc[c];
  uint32_t n = (first || !on) ? 0u : counts[c];
  accumulate_staged<T>(a, n, runs + (on ? c : 0), lc, on ? nseq : 0, nseq, data, offs, buf);
  if (on) {
    acc[c] = a;
    counts[c] = n;
  }
}

// mean = sum / count with the host's T / unsigned arithmetic (dba.hpp:820-822)
template <typename T> __device__ __forceinline__ T div_count(T s, uint32_t c) { return (T)(s / (T)c); }
// A zero count gives 0/0 on the host in the reference; x86 returns its
// default NaN (sign bit set) for that, which is reproduced here.
template <> __device__ __forceinline__ float div_count<float>(float s, uint32_t c) {
  if (c == 0u && s == 0.0f) return __uint_as_float(0xFFC00000u);
  // correctly rounded: double quotient rounded once to float (53 >= 2*24 + 2)
  return (float)((double)s / (double)(float)c);
}
template <> __device__ __forceinline__ double div_count<double>(double s, uint32_t c) {
  if (c == 0u && s == 0.0) return __longlong_as_double((long long)0xFFF8000000000000ull);
  return s / (double)c;
}
template <> __device__ __forceinline__ int32_t div_count<int32_t>(int32_t s, uint32_t c) { return (int32_t)((uint32_t)s / c); }
template <> __device__ __forceinline__ uint64_t div_count<uint64_t>(uint64_t s, uint32_t c) { return s / (uint64_t)c; }

template <typename T> struct IsFloat { static constexpr bool v = false; };
template <> struct IsFloat<float> { static constexpr bool v = true; };
template <> struct IsFloat<double> { static constexpr bool v = true; };

// The max delta of a round as the bits of a non-negative double (their
// integer order is the numeric order). NaN deltas are skipped, as the
// reference's comparison skips them. With `uniform` (every lane of the
// wave targets the same word) the wave's maximum is formed first and lane 0
// issues the atomic; otherwise each lane does (positions of several
// clusters in one wave).
// d < 0: this lane has no delta. Every lane of the wave must call it.
__device__ __forceinline__ void delta_max(unsigned long long* word, double d, bool uniform) {
  const bool has = d == d && d >= 0.0;
  if (uniform) {
    double w = has ? d : -1.0;
    for (int o = 32; o > 0; o >>= 1) w = fmax(w, __shfl_xor(w, o));
    if (w >= 0.0 && (threadIdx.x & 63) == 0) atomicMax(word, (unsigned long long)__double_as_longlong(w));
  } else if (has) {
    atomicMax(word, (unsigned long long)__double_as_longlong(d));
  }
}

template <typename T>
__global__ void k_mean_delta(const T* __restrict__ acc, const uint32_t* __restrict__ counts,
                             const T* __restrict__ oldc, T* __restrict__ newc, int lc,
                             unsigned long long* __restrict__ delta_bits, int* __restrict__ err) {
  // No early exit: the delta's wave reduction needs every lane.
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  double d = -1.0;  // no delta
  if (c < lc) {
    T m;
    if (!IsFloat<T>::v && counts[c] == 0) {
      atomicOr(err, 2);
      m = T(0);
    } else {
      m = div_count<T>(acc[c], counts[c]);
    }
    newc[c] = m;
    d = fabs((double)(T)(oldc[c] - m));  // dba.hpp:830-836
  }
  delta_max(delta_bits, d, true);
}

// ---------------------------------------------------------------------------
// Several clusters' DBA rounds in one launch each: a table of the active
// clusters (sorted by first position) maps a global centroid position to
// its cluster; each position is accumulated and averaged exactly as in the
// single-cluster kernels above. Callers start each cluster at a multiple of
// 64 positions, so no 64-lane block straddles two clusters; the positions
// between a cluster's lc and the next cluster's start hold nothing.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ int slot_of(const MultiSlot* __restrict__ sl, int ns, long long g) {
  int lo = 0, hi = ns - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (sl[mid].pos0 <= g) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// One 64-lane block per 64 global positions: a block inside one cluster
// takes the staged path (a long run costs one memory latency per 256-value
// window), one that straddles clusters adds lane by lane (one latency per
// eight values of a run).
template <typename T>
__global__ void __launch_bounds__(64) k_accumulate_multi(const MultiSlot* __restrict__ sl, int ns, long long total,
                                                         T* __restrict__ acc, uint32_t* __restrict__ counts) {
  __shared__ __attribute__((aligned(16))) T buf[(stage_windows<T>() + 1) * kStageW];
  chain_priority();
  // No early exit: lanes past total take part with no sequences.
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool on = g < total;
  const int si = slot_of<T>(sl, ns, on ? g : total - 1);
  const int su = wave_min_i(si);  // wave-uniform
  T a = T(0);
  uint32_t n = 0;
  if (su == wave_max_i(si)) {
    // One cluster for the whole wave: its slot is read once into scalar
    // registers. Lanes in the cluster's padding (positions lc .. the next
    // multiple of 64) hold no sequences.
    const MultiSlot S = sl[su];
    const int c = on ? (int)(g - S.pos0) : 0;
    const bool live = on && c < S.lc;
    accumulate_staged<T, gptr<T>, gptr<long long>, false, true>(a, n, S.runs + (live ? c : 0), S.lc, live ? S.nmem : 0,
                                                          S.nmem, (gptr<T>)S.data, (gptr<long long>)S.moffs, buf);
  } else {
    const MultiSlot S = sl[si];
    const int c = on ? (int)(g - S.pos0) : 0;
    const bool live = on && c < S.lc;
    accumulate_runs<T>(a, n, S.runs + (live ? c : 0), S.lc, live ? S.nmem : 0, (const T*)S.data, S.moffs);
  }
  if (on) {
    acc[g] = a;
    counts[g] = n;
  }
}

template <typename T>
__global__ void k_mean_delta_multi(const T* __restrict__ acc, const uint32_t* __restrict__ counts,
                                   const MultiSlot* __restrict__ sl, int ns, long long total,
                                   unsigned long long* __restrict__ delta_bits, int* __restrict__ err) {
  // No early exit: lanes past total take the last position's cluster and
  // no delta, so the delta's wave reduction sees every lane.
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool on = g < total;
  const MultiSlot S = sl[slot_of<T>(sl, ns, on ? g : total - 1)];
  double d = -1.0;  // no delta
  const int c = on ? (int)(g - S.pos0) : 0;
  if (on && c < S.lc) {  // not in the cluster's padding
    T m;
    if (!IsFloat<T>::v && counts[g] == 0) {
      atomicOr(err, 2);
      m = T(0);
    } else {
      m = div_count<T>(acc[g], counts[g]);
    }
    ((T*)S.newc)[c] = m;
    d = fabs((double)(T)(((const T*)S.cen)[c] - m));  // dba.hpp:830-836
  }
  delta_max(&delta_bits[S.slot], d, wave_min_i(S.slot) == wave_max_i(S.slot));
}

// ---------------------------------------------------------------------------
// Z-normalisation (gpu_utils.hpp:34-129), one 1024-thread workgroup per
// sequence: chunk sums with the 32-lane shuffle-down tree and the same tree
// over the 32 group sums, chunks added in ascending order (deterministic;
// the reference adds them with atomics), then the double sum of squares the
// same way, then x <- (T)(0 + ((x - mu)/sigma) * 1).
// ---------------------------------------------------------------------------
template <typename T> __device__ __forceinline__ T d2t(double d) { return (T)d; }
template <> __device__ __forceinline__ int32_t d2t<int32_t>(double d) {
  if (d != d) return 0;
  if (d >= 2147483647.0) return INT_MAX;
  if (d <= -2147483648.0) return INT_MIN;
  return (int32_t)d;
}
template <> __device__ __forceinline__ uint32_t d2t<uint32_t>(double d) {
  if (d != d || d <= 0.0) return 0;
  if (d >= 4294967295.0) return UINT_MAX;
  return (uint32_t)d;
}
template <> __device__ __forceinline__ uint64_t d2t<uint64_t>(double d) {
  if (d != d || d <= 0.0) return 0;
  if (d >= 18446744073709551615.0) return ULLONG_MAX;
  return (uint64_t)d;
}

template <typename A> __device__ __forceinline__ A tree32(A v) {
  for (int off = 16; off > 0; off >>= 1) v = (A)(v + __shfl_down(v, off, 32));
  return v;
}

template <typename T>
__global__ void __launch_bounds__(1024) k_znorm(T* __restrict__ data, const long long* __restrict__ offs,
                                                const int* __restrict__ lens, int n, double* __restrict__ mu,
                                                double* __restrict__ sigma) {
  __shared__ T wsum[32];
  __shared__ double wsos[32];
  __shared__ T total_s;
  __shared__ double sos_s;
  const int seq = blockIdx.x;
  const int L = lens[seq];
  T* x = data + offs[seq];
  const int tid = threadIdx.x;
  const int nch = (L + 1023) / 1024;
  T total = T(0);
  for (int ch = 0; ch < nch; ch++) {
    const int p = ch * 1024 + tid;
    T v = p < L ? x[p] : T(0);
    v = tree32(v);
    if ((tid & 31) == 0) wsum[tid >> 5] = v;
    __syncthreads();
    if (tid < 32) {
      T w = wsum[tid];
      w = tree32(w);
      if (tid == 0) total = (T)(total + w);
    }
    __syncthreads();
  }
  if (tid == 0) total_s = total;
  __syncthreads();
  total = total_s;
  const double mean = (double)total / (double)L;
  double sos = 0.0;
  for (int ch = 0; ch < nch; ch++) {
    const int p = ch * 1024 + tid;
    double v = 0.0;
    if (p < L) {
      v = (double)x[p] - ((double)total / (double)L);
      v *= v;
    }
    v = tree32(v);
    if ((tid & 31) == 0) wsos[tid >> 5] = v;
    __syncthreads();
    if (tid < 32) {
      double w = wsos[tid];
      w = tree32(w);
      if (tid == 0) sos += w;
    }
    __syncthreads();
  }
  if (tid == 0) sos_s = sos;
  __syncthreads();
  sos = sos_s;
  double sd = sqrt(sos / (double)L);
  if (tid == 0) {
    if (mu) mu[seq] = mean;
    if (sigma) sigma[seq] = sd;
  }
  if (sd == 0) sd = 1;
  for (int p = tid; p < L; p += 1024) x[p] = d2t<T>(0.0 + (((double)x[p] - mean) / sd) * 1.0);
}

// Reference step-code matrix of one pair (ly x lx, row-major).
__global__ void k_decode_moves(const uint32_t* __restrict__ mv, const uint32_t* __restrict__ mtop, int ab, int ly,
                               int lx, int os, int oe, uint8_t* __restrict__ out) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)ly * lx) return;
  const int i = (int)(idx / lx), j = (int)(idx % lx);
  const Geo g = make_geo(ly, lx, os, oe, ab);
  long long caddr = -1;
  uint32_t cword = 0;
  out[idx] = step_code(g, mv, mtop, i, j, caddr, cword);
}

// ---------------------------------------------------------------------------
hipError_t launch_traceback(hipStream_t st, const TbJob* jobs, int njobs, const uint32_t* mv, const uint32_t* mtop,
                            const int* abr, int4* runs, int lc, int os, int oe, int* err, unsigned long long* pbuf) {
  if (njobs <= 0) return hipSuccess;
  if (pbuf)
    hipLaunchKernelGGL(k_traceback<true>, dim3((njobs + 3) / 4), dim3(256), 0, st, jobs, njobs, mv, mtop, abr, runs, lc,
                       os, oe, err, pbuf);
  else
    hipLaunchKernelGGL(k_traceback<false>, dim3((njobs + 3) / 4), dim3(256), 0, st, jobs, njobs, mv, mtop, abr, runs, lc,
                       os, oe, err, pbuf);
  return hipGetLastError();
}

hipError_t launch_stripe_traceback(hipStream_t st, const TbJob* jobs, int njobs, int m, const uint32_t* mv,
                                   const uint32_t* mtop, StripeState* ss, int4* runs, int lc, int os, int oe,
                                   int* err, unsigned long long* pbuf) {
  if (njobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_stripe_traceback<false>, dim3((njobs + 3) / 4), dim3(256), 0, st, jobs, njobs, m, mv, mtop, ss, runs,
                     lc, os, oe, err, pbuf);
  return hipGetLastError();
}

hipError_t launch_sck_traceback(hipStream_t st, const TbJob* jobs, int njobs, const uint32_t* mv,
                                const uint32_t* mtop, StripeState* ss, int4* runs, int os, int oe, int* err,
                                unsigned long long* pbuf) {
  if (njobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_stripe_traceback<true>, dim3((njobs + 3) / 4), dim3(256), 0, st, jobs, njobs, 0, mv, mtop, ss,
                     runs, 0, os, oe, err, pbuf);
  return hipGetLastError();
}

template <typename T>
static hipError_t acc_t(hipStream_t st, const void* data, const long long* offs, const int4* runs, int nseq, int lc,
                        void* acc, uint32_t* counts, int first) {
  hipLaunchKernelGGL(k_accumulate<T>, dim3((lc + 63) / 64), dim3(64), 0, st, (const T*)data, offs, runs, nseq, lc,
                     (T*)acc, counts, first);
  return hipGetLastError();
}

hipError_t launch_accumulate(int dtype, hipStream_t st, const void* data, const long long* offs, const int4* runs,
                             int nseq, int lc, void* acc, uint32_t* counts, int first) {
  switch (dtype) {
    case ODBA_F32: return acc_t<float>(st, data, offs, runs, nseq, lc, acc, counts, first);
    case ODBA_F64: return acc_t<double>(st, data, offs, runs, nseq, lc, acc, counts, first);
    case ODBA_I32: return acc_t<int32_t>(st, data, offs, runs, nseq, lc, acc, counts, first);
    case ODBA_U32: return acc_t<uint32_t>(st, data, offs, runs, nseq, lc, acc, counts, first);
    case ODBA_U64: return acc_t<uint64_t>(st, data, offs, runs, nseq, lc, acc, counts, first);
  }
  return hipErrorInvalidValue;
}

template <typename T>
static hipError_t md_t(hipStream_t st, const void* acc, const uint32_t* counts, const void* oldc, void* newc, int lc,
                       unsigned long long* delta_bits, int* err) {
  hipLaunchKernelGGL(k_mean_delta<T>, dim3((lc + 255) / 256), dim3(256), 0, st, (const T*)acc, counts, (const T*)oldc,
                     (T*)newc, lc, delta_bits, err);
  return hipGetLastError();
}

template <typename T>
static hipError_t multi_t(hipStream_t st, const MultiSlot* sl, int ns, long long total, void* acc, uint32_t* counts,
                          unsigned long long* delta_bits, int* err) {
  const unsigned nb = (unsigned)((total + 255) / 256);
  hipLaunchKernelGGL(k_accumulate_multi<T>, dim3((unsigned)((total + 63) / 64)), dim3(64), 0, st, sl, ns, total,
                     (T*)acc, counts);
  hipLaunchKernelGGL(k_mean_delta_multi<T>, dim3(nb), dim3(256), 0, st, (const T*)acc, counts, sl, ns, total,
                     delta_bits, err);
  return hipGetLastError();
}

hipError_t launch_multi_mean(int dtype, hipStream_t st, const MultiSlot* sl, int ns, long long total, void* acc,
                             uint32_t* counts, unsigned long long* delta_bits, int* err) {
  if (ns <= 0 || total <= 0) return hipSuccess;
  switch (dtype) {
    case ODBA_F32: return multi_t<float>(st, sl, ns, total, acc, counts, delta_bits, err);
    case ODBA_F64: return multi_t<double>(st, sl, ns, total, acc, counts, delta_bits, err);
    case ODBA_I32: return multi_t<int32_t>(st, sl, ns, total, acc, counts, delta_bits, err);
    case ODBA_U32: return multi_t<uint32_t>(st, sl, ns, total, acc, counts, delta_bits, err);
    case ODBA_U64: return multi_t<uint64_t>(st, sl, ns, total, acc, counts, delta_bits, err);
  }
  return hipErrorInvalidValue;
}

// Fills n 32-bit words (flags, counters) with a kernel rather than
// hipMemsetAsync, so the work can sit inside a captured graph.
__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

hipError_t launch_fill_u32(hipStream_t st, void* p, uint32_t v, long long n) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (uint32_t*)p, v, n);
  return hipGetLastError();
}

hipError_t launch_mean_delta(int dtype, hipStream_t st, const void* acc, const uint32_t* counts, const void* old_c,
                             void* new_c, int lc, unsigned long long* delta_bits, int* err) {
  switch (dtype) {
    case ODBA_F32: return md_t<float>(st, acc, counts, old_c, new_c, lc, delta_bits, err);
    case ODBA_F64: return md_t<double>(st, acc, counts, old_c, new_c, lc, delta_bits, err);
    case ODBA_I32: return md_t<int32_t>(st, acc, counts, old_c, new_c, lc, delta_bits, err);
    case ODBA_U32: return md_t<uint32_t>(st, acc, counts, old_c, new_c, lc, delta_bits, err);
    case ODBA_U64: return md_t<uint64_t>(st, acc, counts, old_c, new_c, lc, delta_bits, err);
  }
  return hipErrorInvalidValue;
}

template <typename T>
static hipError_t zn_t(hipStream_t st, void* data, const long long* offs, const int* lens, int n, double* mu,
                       double* sigma) {
  hipLaunchKernelGGL(k_znorm<T>, dim3(n), dim3(1024), 0, st, (T*)data, offs, lens, n, mu, sigma);
  return hipGetLastError();
}

hipError_t launch_znorm(int dtype, hipStream_t st, void* data, const long long* offs, const int* lens, int n,
                        double* mu, double* sigma) {
  if (n <= 0) return hipSuccess;
  switch (dtype) {
    case ODBA_F32: return zn_t<float>(st, data, offs, lens, n, mu, sigma);
    case ODBA_F64: return zn_t<double>(st, data, offs, lens, n, mu, sigma);
    case ODBA_I32: return zn_t<int32_t>(st, data, offs, lens, n, mu, sigma);
    case ODBA_U32: return zn_t<uint32_t>(st, data, offs, lens, n, mu, sigma);
    case ODBA_U64: return zn_t<uint64_t>(st, data, offs, lens, n, mu, sigma);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_decode_moves(hipStream_t st, const uint32_t* mv, const uint32_t* mtop, int ab, int ly, int lx, int os,
                               int oe, uint8_t* out) {
  long long cells = (long long)ly * lx;
  if (cells <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_decode_moves, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, st, mv, mtop, ab, ly, lx, os,
                     oe, out);
  return hipGetLastError();
}

}  // namespace odba
