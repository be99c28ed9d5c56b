/*
 * CPU oracle for the OpenDBA DTW/DBA hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This file restates, on the CPU and in plain C++,
 * the semantics of the reference's CUDA kernels and host loops so that the
 * HIP implementation can be checked against it. Only the test suite, the
 * smoke entry point and the benchmark's CPU-baseline leg load it; the product
 * library never links or calls it.
 *
 * Build: g++ -O3 -march=x86-64-v3 -std=c++17 -ffp-contract=off -fwrapv (oracle/Makefile).
 * -ffp-contract=off keeps `cost + diff*diff` as a multiply then an add, the
 * order the DTW recurrence is specified with; -fwrapv gives signed int
 * arithmetic the two's-complement wrap the GPU has.
 *
 * Every function cites the reference lines it follows. Cells of the DTW cost
 * matrix are indexed (i, j): i = row = position in the "first" sequence Y,
 * j = column = position in the "second" sequence X (dtw.hpp:36-38).
 *
 * What pins each function (tests/test_oracle.py, test_oracle_open_modes.py,
 * test_oracle_kseg_dp.py):
 *   dtw_swath / dtw_pair      textbook brute force (global mode, bit-exact),
 *                             transposition symmetry, the openDBA_test.cu KATs,
 *                             the test1 golden centroids; open end, open start
 *                             and open against closed forms over textbook and
 *                             subsequence DTW matrices (distances and step
 *                             codes), including the swath-left columns at
 *                             small swath widths; the open-end early stop
 *                             against the same run with the stop switched off
 *                             (same distance and traceback counts)
 *   KsegDP                    brute-force enumeration of every cut into k
 *                             segments of span < e (optimal cost and the
 *                             first-strict-minimum backtrack's choice)
 *   kseg (search, medians)    closed forms only (constant and step signals,
 *                             medians are input values): the k search and the
 *                             median read past a segment's end are unpinned
 *                             (the reference's threads sort neighbouring
 *                             segments concurrently, so the value it reads
 *                             past the end is a race there; this restatement
 *                             reads the values as loaded)
 *   traceback_full            test1 golden centroids (exact in double),
 *                             Sigma count = path length minus open moves
 *   dba_update / perform_dba  openDBA_test.cu KATs (3 decimals) and the
 *                             test1 golden vectors output_results*.txt (exact)
 *   znorm                     seq*_test_znorm.txt (population-SD z-norm)
 *   stripe traceback          equals the full-matrix path in global mode
 *                             (brute-force comparison); the open-end height
 *                             quirk has no fixture: parity unpinned there
 *   prefix_chop               closed forms: an exact prefix match (chop at
 *                             P-1, one cell per prefix row) and a prefix
 *                             whose values each occur twice (hand-derived
 *                             chop and histogram); histogram sum = path
 *                             length minus open moves
 */
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../include/odba/odba_types.h"

namespace {

/* cudahack::numeric_limits<T>::max() (limits.hpp:25-48). */
template <typename T> struct Lim;
template <> struct Lim<int32_t> { static int32_t max() { return INT_MAX; } };
template <> struct Lim<uint32_t> { static uint32_t max() { return UINT_MAX; } };
template <> struct Lim<uint64_t> { static uint64_t max() { return ULLONG_MAX; } };
template <> struct Lim<float> { static float max() { return FLT_MAX; } };
template <> struct Lim<double> { static double max() { return DBL_MAX; } };

/* Float -> T with the GPU conversion rule: round toward zero, saturate,
 * NaN -> 0 (what cvt.rzi.* does for the casts at dtw.hpp:121,251,254). */
template <typename T> T f2t(float f);
template <> float f2t<float>(float f) { return f; }
template <> double f2t<double>(float f) { return (double)f; }
template <> int32_t f2t<int32_t>(float f) {
  if (f != f) return 0;
  if (f >= 2147483648.0f) return INT_MAX;
  if (f <= -2147483648.0f) return INT_MIN;
  return (int32_t)f;
}
template <> uint32_t f2t<uint32_t>(float f) {
  if (f != f || f <= 0.0f) return 0;
  if (f >= 4294967296.0f) return UINT_MAX;
  return (uint32_t)f;
}
template <> uint64_t f2t<uint64_t>(float f) {
  if (f != f || f <= 0.0f) return 0;
  if (f >= 18446744073709551616.0f) return ULLONG_MAX;
  return (uint64_t)f;
}
/* Double -> T, same rule (gpu_utils.hpp:127 and the host rescale at
 * dba.hpp:1122, where the host cast of an in-range value agrees). */
template <typename T> T d2t(double d);
template <> float d2t<float>(double d) { return (float)d; }
template <> double d2t<double>(double d) { return d; }
template <> int32_t d2t<int32_t>(double d) {
  if (d != d) return 0;
  if (d >= 2147483647.0) return INT_MAX;
  if (d <= -2147483648.0) return INT_MIN;
  return (int32_t)d;
}
template <> uint32_t d2t<uint32_t>(double d) {
  if (d != d || d <= 0.0) return 0;
  if (d >= 4294967295.0) return UINT_MAX;
  return (uint32_t)d;
}
template <> uint64_t d2t<uint64_t>(double d) {
  if (d != d || d <= 0.0) return 0;
  if (d >= 18446744073709551615.0) return ULLONG_MAX;
  return (uint64_t)d;
}

/* moveI / moveJ step tables (dtw.hpp:20-21). */
const int kMoveI[] = {-1, -1, 0, -1, 0, 0, 0};
const int kMoveJ[] = {-1, -1, -1, 0, -1, -1, -1};

/* Final pair distance from the accumulated cost (dtw.hpp:242-256 and the
 * abrogation branch dtw.hpp:121): computed in float for every T. */
template <typename T> T pair_distance(T cost, size_t ly, bool normalise) {
  float f = sqrtf((float)cost);
  if (normalise) f = f / (float)ly;
  return f2t<T>(f);
}

struct SwathResult {
  int kind = 0;          /* 0 computed, 1 skipped (already abrogated), 2 abrogated now, 3 past end */
  bool q3_undefined = false;
};

/*
 * One launch of DTWDistance<T> for one pair (dtw.hpp:39-259).
 *   y, ly   first sequence (rows); ly may be a truncated height (stripe
 *           recompute, dba.hpp:708,716)
 *   x, lx   second sequence (columns)
 *   off     offset_within_second_seq (swath left column)
 *   prev    dtwCostSoFar: previous swath's right edge, ly entries
 *   next    newDtwCostSoFar or nullptr; on entry it must hold what the
 *           reference's buffer holds (the D2D copy of prev in the medoid and
 *           full-path flows, dba.hpp:150,603,613)
 *   mv      callback (row, t, code) for path storage; t is the swath-local
 *           column; the caller maps it to absolute when next != nullptr
 *   dist    dtwPairwiseDistances slot or nullptr
 */
template <typename T, bool kMoves, typename MoveFn>
SwathResult dtw_swath(const T* y, size_t ly, const T* x, size_t lx, size_t W, size_t off,
                      int os, int oe, const T* prev, T* next, MoveFn&& mv, T* dist,
                      std::vector<T>& colprev, std::vector<T>& colcur, std::vector<T>& row0,
                      bool abrogate = true) {
  SwathResult res;
  const T MAXV = Lim<T>::max();
  if (off >= lx) { res.kind = 3; return res; }                         /* dtw.hpp:50-52 */
  const size_t wcols = std::min(W, lx - off);
  /* Open-end early abrogation, checked only at swath starts (dtw.hpp:74-129).
     abrogate == false is a test hook (the reference has no such switch): the
     check is skipped so a test can compare the full recurrence with it. */
  if (abrogate && off > ly && oe && !os) {
    if (prev[0] == MAXV) {                                               /* dtw.hpp:76-81 */
      if (kMoves) for (size_t t = 0; t < wcols; t++) mv(ly - 1, t, (uint8_t)ODBA_OPEN_RIGHT);
      res.kind = 1;
      return res;
    }
    T mn = prev[0];                                                      /* dtw.hpp:84-103 */
    for (size_t i = 1; i < ly; i++)
      if (prev[i] < mn) mn = prev[i];
    if (prev[ly - 1] == mn) {                                            /* dtw.hpp:106-127 */
      if (kMoves) for (size_t t = 0; t < wcols; t++) mv(ly - 1, t, (uint8_t)ODBA_OPEN_RIGHT);
      if (next) {
        next[0] = MAXV;
        if (dist) *dist = pair_distance<T>(next[ly - 1], ly, true);
      }
      res.kind = 2;
      return res;
    }
  }
  /* Row 0, serial on thread 0 (dtw.hpp:131-156). */
  if (row0.size() < wcols) row0.resize(wcols);
  const T first = y[0];
  T c0 = (off == 0) ? T(0) : prev[0];
  if (os) c0 = (T)(c0 + T(0));
  else { T d = (T)(first - x[off]); c0 = (T)(c0 + (T)(d * d)); }
  row0[0] = c0;
  if (kMoves) mv(0, 0, (uint8_t)(off == 0 ? (os ? ODBA_NIL_OPEN_RIGHT : ODBA_NIL) : (os ? ODBA_OPEN_RIGHT : ODBA_RIGHT)));
  size_t col;
  for (col = 1; col < W && off + col < lx; col++) {
    T d = os ? T(0) : (T)(first - x[off + col]);
    row0[col] = (T)(row0[col - 1] + (T)(d * d));
    if (kMoves) mv(0, col, (uint8_t)(os ? ODBA_OPEN_RIGHT : ODBA_RIGHT));
  }
  if (next) {
    /* dtw.hpp:155 reads costs[col-1] without the ring offset; that is the
       right cell only when (col-1) % 3 == 0 (always so for a full 1024-wide
       swath). Otherwise the reference reads an unwritten LDS slot. */
    next[0] = row0[col - 1];
    if ((col - 1) % 3 != 0) res.q3_undefined = true;
  }
  /* Wavefront body (dtw.hpp:159-239), evaluated column by column. */
  if (colprev.size() < ly) { colprev.resize(ly); colcur.resize(ly); }
  for (size_t t = 0; t < wcols; t++) {
    const size_t j = off + t;
    const T xv = x[j];
    colcur[0] = row0[t];
    for (size_t i = 1; i < ly; i++) {
      T df = (T)(y[i] - xv);
      T dd = (T)(df * df);
      T up = (T)(colcur[i - 1] + dd);
      T right = MAXV, diag = MAXV;
      bool open = false;
      if (t == 0) {
        if (off != 0) {                                                  /* dtw.hpp:175-186 */
          right = prev[i];
          diag = prev[i - 1];
          if (i < ly - 1 || !oe) { right = (T)(right + dd); diag = (T)(diag + dd); }
          else open = true;
        }
      } else {                                                           /* dtw.hpp:190-192 */
        right = (T)(colprev[i] + dd);
        diag = (T)(colprev[i - 1] + dd);
      }
      if (oe && i == ly - 1 && t != 0) { right = colprev[i]; open = true; } /* dtw.hpp:196-200 */
      T v; uint8_t code;                                                 /* dtw.hpp:202-227 */
      if (diag > up) {
        if (up > right) { v = right; code = open ? ODBA_OPEN_RIGHT : ODBA_RIGHT; }
        else { v = up; code = ODBA_UP; }
      } else {
        if (diag > right) { v = right; code = open ? ODBA_OPEN_RIGHT : ODBA_RIGHT; }
        else { v = diag; code = ODBA_DIAGONAL; }
      }
      colcur[i] = v;
      if (kMoves) mv(i, t, code);
    }
    if (next && (t == W - 1 || j == lx - 1))                             /* dtw.hpp:232-234 */
      for (size_t i = 1; i < ly; i++) next[i] = colcur[i];
    std::swap(colprev, colcur);
  }
  /* Final distance once the last swath is done (dtw.hpp:242-256). */
  if (off + W >= lx && dist && next) *dist = pair_distance<T>(next[ly - 1], ly, (oe != 0) != (os != 0));
  return res;
}

/*
 * Whole-pair DTW following the host loop of the medoid search
 * (dba.hpp:138-155) or of a full-path DBA round (dba.hpp:589-630): swaths
 * left to right, the edge buffer copied new -> old after every launch.
 * moves, when given, is a row-major ly x lx matrix of step codes (0 where
 * the reference leaves the cell unwritten).
 */
template <typename T>
int dtw_pair(const T* y, size_t ly, const T* x, size_t lx, size_t W, int os, int oe,
             uint8_t* moves, T* dist, int* abrogated_swath, int* q3_flag, bool abrogate = true) {
  if (ly == 0 || lx == 0 || W == 0) return ODBA_E_INVALID;
  std::vector<T> prev(ly, T(0)), next(ly, T(0)), cp, cc, r0;
  if (abrogated_swath) *abrogated_swath = -1;
  bool q3_last = false;
  for (size_t off = 0; off < lx; off += W) {
    SwathResult r;
    if (moves) {
      auto mv = [&](size_t i, size_t t, uint8_t c) { moves[i * lx + off + t] = c; };
      r = dtw_swath<T, true>(y, ly, x, lx, W, off, os, oe, prev.data(), next.data(), mv, dist, cp, cc, r0, abrogate);
    } else {
      auto mv = [](size_t, size_t, uint8_t) {};
      r = dtw_swath<T, false>(y, ly, x, lx, W, off, os, oe, prev.data(), next.data(), mv, dist, cp, cc, r0, abrogate);
    }
    if (r.kind == 2 && abrogated_swath) *abrogated_swath = (int)(off / W);
    q3_last = r.q3_undefined;
    prev = next; /* cudaMemcpyAsync new -> old (dba.hpp:150) */
  }
  /* The Q3 slot is consumed as the distance only when ly == 1. */
  if (q3_flag) *q3_flag = (ly == 1 && q3_last && !(abrogated_swath && *abrogated_swath >= 0)) ? 1 : 0;
  return ODBA_OK;
}

/*
 * updateCentroid<T> with column_offset == 0 (dba.hpp:374-423): traceback
 * from the top-right corner of a full move matrix, adding the sequence value
 * into the centroid position for every step except OPEN_RIGHT, then the
 * (0,0) anchor unless it is NIL_OPEN_RIGHT. rows x cols is the matrix shape
 * (Y length x X length). flip: centroid on Y.
 */
template <typename T>
int traceback_full(const uint8_t* M, size_t rows, size_t cols, int flip, const T* seq, T* sums,
                   uint32_t* counts, size_t* path_len) {
  long i = (long)rows - 1, j = (long)cols - 1;
  uint8_t mv = M[i * cols + j];
  size_t steps = 0;
  while (j >= 0 && mv != ODBA_NIL && mv != ODBA_NIL_OPEN_RIGHT) {
    if (mv != ODBA_OPEN_RIGHT) {
      size_t ci = flip ? (size_t)i : (size_t)j, si = flip ? (size_t)j : (size_t)i;
      sums[ci] = (T)(sums[ci] + seq[si]);
      counts[ci]++;
    }
    if (mv > 5) return ODBA_E_TRACEBACK;
    i += kMoveI[mv];
    j += kMoveJ[mv];
    steps++;
    if (j < 0 || i < 0) break;
    mv = M[i * cols + j];
  }
  if (j != 0 || i != 0) return ODBA_E_TRACEBACK; /* dba.hpp:405-411 (the reference checks j) */
  if (mv != ODBA_NIL_OPEN_RIGHT) {                /* dba.hpp:412-415 */
    sums[0] = (T)(sums[0] + seq[0]);
    counts[0]++;
  }
  if (path_len) *path_len = steps + 1;
  return ODBA_OK;
}

/*
 * chopPrefixFromSequences<T> for one (already normalised) sequence
 * (dba.hpp:1192-1395): open-end DTW with the prefix on Y (rows) and the
 * sequence on X, W = 1024 (dba.hpp:1219-1224, 1264-1280), then the host walk
 * of dba.hpp:1317-1336: from the top-right corner, step left over the top
 * row's OPEN_RIGHT run; the column reached is the chop position. hist[i] then
 * counts the path cells in prefix row i from that cell down to (0,0); the
 * first cell is set to 1 (dba.hpp:1329). hist starts at zero (the reference's
 * buffer is fresh pinned pages).
 */
template <typename T>
int prefix_chop(const T* prefix, size_t P, const T* seq, size_t L, uint64_t* chop, int32_t* hist) {
  if (P == 0 || L == 0) return ODBA_E_INVALID;
  std::vector<uint8_t> M(P * L, 0);
  T dist;
  int rc = dtw_pair<T>(prefix, P, seq, L, 1024, 0, 1, M.data(), &dist, nullptr, nullptr);
  if (rc) return rc;
  long i = (long)P - 1, j = (long)L - 1;
  uint8_t mv = M[i * L + j];
  while (mv == ODBA_OPEN_RIGHT) {
    i += kMoveI[mv];
    j += kMoveJ[mv];
    if (j < 0) return ODBA_E_TRACEBACK;
    mv = M[i * L + j];
  }
  *chop = (uint64_t)j;
  for (size_t k = 0; k < P; k++) hist[k] = 0;
  hist[i] = 1;
  while (mv != ODBA_NIL) {
    if (mv == 0 || mv > 5) return ODBA_E_TRACEBACK;  // an unwritten cell: never reached
    i += kMoveI[mv];
    j += kMoveJ[mv];
    if (i < 0 || j < 0) return ODBA_E_TRACEBACK;
    hist[i]++;
    mv = M[i * L + j];
  }
  return ODBA_OK;
}

/*
 * Stripe-mode path for one sequence (dba.hpp:523-548, 589-630, 646-763):
 * forward pass storing every swath's right edge column, then right-to-left
 * recomputation of each stripe with the height truncated to the current
 * traceback row, and a partial traceback per stripe (updateCentroid with
 * column_offset / stripe_rows). Edge slots are zero before the forward pass
 * writes them (fresh managed pages).
 */
template <typename T>
int stripe_update(const T* Y, size_t H, const T* X, size_t LX, size_t W, int os, int oe, int flip,
                  const T* seq, T* sums, uint32_t* counts) {
  const size_t K = (LX + W - 1) / W;
  std::vector<T> slots(K * H, T(0));
  std::vector<T> cp, cc, r0;
  auto nomv = [](size_t, size_t, uint8_t) {};
  for (size_t off = 0; off < LX; off += W) {
    size_t k = off / W;
    T* newc = &slots[k * H];
    const T* oldc = k ? &slots[(k - 1) * H] : nullptr;
    std::vector<T> dummy;
    if (!oldc) { dummy.assign(H, T(0)); oldc = dummy.data(); }
    dtw_swath<T, false>(Y, H, X, LX, W, off, os, oe, oldc, newc, nomv, (T*)nullptr, cp, cc, r0);
  }
  size_t offset = LX;
  long h = (long)H;
  std::vector<uint8_t> S;
  while (offset) {
    size_t left = (size_t)((std::ceil(((float)offset) / W) - 1) * W);  /* dba.hpp:699 */
    size_t width = offset % W;
    if (width == 0) width = W;
    const T* existing = left ? &slots[(left / W - 1) * H] : nullptr;
    std::vector<T> dummy;
    if (!existing) { dummy.assign(H, T(0)); existing = dummy.data(); }
    S.assign((size_t)h * W, 0);
    auto mv = [&](size_t i, size_t t, uint8_t c) { S[i * W + t] = c; };
    dtw_swath<T, true>(Y, (size_t)h, X, LX, W, left, os, oe, existing, (T*)nullptr, mv, (T*)nullptr, cp, cc, r0);
    offset -= width;
    size_t co = offset;
    /* updateCentroid with stripe_rows (dba.hpp:374-423, launched at 733) */
    long i = h - 1, j = (long)width - 1;
    uint8_t m = S[i * W + j];
    while (j >= 0 && m != ODBA_NIL && m != ODBA_NIL_OPEN_RIGHT) {
      if (m != ODBA_OPEN_RIGHT) {
        size_t ci = flip ? (size_t)i : (size_t)j + co, si = flip ? (size_t)j + co : (size_t)i;
        sums[ci] = (T)(sums[ci] + seq[si]);
        counts[ci]++;
      }
      if (m > 5) return ODBA_E_TRACEBACK;
      i += kMoveI[m];
      j += kMoveJ[m];
      if (j < 0 || i < 0) break;
      m = S[i * W + j];
    }
    if (co == 0) {
      if (j != 0) return ODBA_E_TRACEBACK;
      if (m != ODBA_NIL_OPEN_RIGHT) { sums[0] = (T)(sums[0] + seq[0]); counts[0]++; }
    } else if (j != -1 || i < 0) {
      return ODBA_E_TRACEBACK;
    }
    h = i + 1;
  }
  return ODBA_OK;
}

/*
 * One DBA round (DBAUpdate<T>, dba.hpp:432-848) on a single device:
 * sequences in order, each aligned against the centroid (flipped when
 * open_end and the centroid is shorter, dba.hpp:497-501), traceback
 * accumulation in sequence order, mean = sum / count in T / unsigned
 * (dba.hpp:820-822), delta = max |old - new| with the subtraction in T
 * (dba.hpp:830-836). path_policy / free_mem select full or stripe mode per
 * sequence with the reference's rule (dba.hpp:523-548).
 */
template <typename T> T div_count(T s, uint32_t c) { return (T)(s / c); }
template <> int32_t div_count<int32_t>(int32_t s, uint32_t c) { return (int32_t)((uint32_t)s / c); }
template <> float div_count<float>(float s, uint32_t c) { return s / (float)c; }
template <> double div_count<double>(double s, uint32_t c) { return s / (double)c; }

template <typename T> bool is_integral_t() { return !std::is_floating_point<T>::value; }

template <typename T>
int dba_update(const T* C, size_t lc, const T* const* seqs, const uint64_t* lens, size_t n, int os, int oe,
               int path_policy, uint64_t free_mem, T* out, uint32_t* counts_out, double* delta_out) {
  const size_t W = ODBA_REF_W;
  std::vector<T> sums(lc, T(0));
  std::vector<uint32_t> counts(lc, 0u);
  std::vector<uint8_t> M;
  for (size_t s = 0; s < n; s++) {
    const size_t ls = lens[s];
    const int flip = (oe && lc < ls) ? 1 : 0;
    const T* Y = flip ? C : seqs[s];
    const size_t H = flip ? lc : ls;
    const T* X = flip ? seqs[s] : C;
    const size_t LX = flip ? ls : lc;
    bool stripe = false;
    if (path_policy == ODBA_PATH_STRIPE) stripe = true;
    else if (path_policy == ODBA_PATH_AUTO) {
      double path_bytes = (double)ls * (double)lc;
      double edge_bytes = (double)sizeof(T) * (double)H;
      stripe = (double)free_mem < edge_bytes + path_bytes * 1.05;
    }
    int rc;
    if (stripe) {
      rc = stripe_update<T>(Y, H, X, LX, W, os, oe, flip, seqs[s], sums.data(), counts.data());
    } else {
      M.assign(H * LX, 0);
      rc = dtw_pair<T>(Y, H, X, LX, W, os, oe, M.data(), nullptr, nullptr, nullptr);
      if (rc == ODBA_OK) rc = traceback_full<T>(M.data(), H, LX, flip, seqs[s], sums.data(), counts.data(), nullptr);
    }
    if (rc != ODBA_OK) return rc;
  }
  double md = 0.0;
  for (size_t t = 0; t < lc; t++) {
    if (counts[t] == 0 && is_integral_t<T>()) return ODBA_E_DIV_ZERO;
    out[t] = div_count<T>(sums[t], counts[t]);
    double d = std::fabs((double)(T)(C[t] - out[t]));
    if (d > md) md = d;
  }
  if (counts_out) std::memcpy(counts_out, counts.data(), lc * sizeof(uint32_t));
  if (delta_out) *delta_out = md;
  return ODBA_OK;
}

/*
 * Z-normalisation (calc_sums / calc_sum_of_squares / rescale_sequences,
 * gpu_utils.hpp:34-129): sums in T over 1024-element chunks with the 32-lane
 * shuffle-down tree, then the same tree over the 32 warp sums, chunks added
 * in ascending order (the reference adds them with atomics, so its order is
 * unspecified for L > 1024); sum of squares in double the same way;
 * x <- (T)(0 + ((x - mu)/sigma)*1) with population sigma, sigma == 0 -> 1.
 */
template <typename A> A warp_tree32(A* a) {
  for (int off = 16; off > 0; off /= 2) {
    A b[32];
    for (int i = 0; i < 32; i++) b[i] = (A)(a[i] + ((i + off < 32) ? a[i + off] : a[i]));
    std::memcpy(a, b, sizeof(b));
  }
  return a[0];
}
template <typename A> A block_tree1024(const A* v) {
  A warp[32];
  for (int w = 0; w < 32; w++) {
    A a[32];
    for (int l = 0; l < 32; l++) a[l] = v[w * 32 + l];
    warp[w] = warp_tree32(a);
  }
  return warp_tree32(warp);
}
template <typename T>
void znorm_one(T* x, size_t L, double* mu_out, double* sigma_out) {
  const size_t chunks = (L + 1023) / 1024;
  T sum = T(0);
  std::vector<T> buf(1024);
  for (size_t c = 0; c < chunks; c++) {
    for (size_t k = 0; k < 1024; k++) { size_t p = c * 1024 + k; buf[k] = p < L ? x[p] : T(0); }
    sum = (T)(sum + block_tree1024(buf.data()));
  }
  double sos = 0.0;
  std::vector<double> dbuf(1024);
  for (size_t c = 0; c < chunks; c++) {
    for (size_t k = 0; k < 1024; k++) {
      size_t p = c * 1024 + k;
      double v = 0.0;
      if (p < L) { v = (double)x[p] - ((double)sum / (double)L); v *= v; }
      dbuf[k] = v;
    }
    sos += block_tree1024(dbuf.data());
  }
  double mean = (double)sum / (double)L;
  double sd = std::sqrt(sos / (double)L);
  if (mu_out) *mu_out = mean;
  if (sigma_out) *sigma_out = sd;
  if (sd == 0) sd = 1;
  for (size_t p = 0; p < L; p++) x[p] = d2t<T>(0.0 + (((double)x[p] - mean) / sd) * 1.0);
}

/* Sanitize (dba.hpp:868-872). */
template <typename T> bool sanitize_drop_last(T v) { return (double)v >= std::sqrt((double)Lim<T>::max()); }
template <> bool sanitize_drop_last<float>(float v) { return v >= std::sqrt(FLT_MAX); }

/*
 * All-vs-all distances in the given order (the medoid search kernel loop,
 * dba.hpp:67-158): pair (r, c > r) has Y = sequence r. Output is the
 * row-major upper triangle (PAIRWISE_DIST_ROW, clustering.cuh:6).
 */
template <typename T>
int all_pairs(const T* const* seqs, const uint64_t* lens, size_t n, int os, int oe, size_t W, T* out,
              int nthreads, const int64_t* pair_subset, size_t n_subset) {
  const size_t npairs = n * (n - 1) / 2;
  std::vector<std::pair<uint32_t, uint32_t>> pairs;
  if (pair_subset) {
    for (size_t k = 0; k < n_subset; k++) {
      int64_t p = pair_subset[k];
      size_t r = 0, base = 0;
      while (base + (n - r - 1) <= (size_t)p) { base += n - r - 1; r++; }
      pairs.emplace_back((uint32_t)r, (uint32_t)(r + 1 + (p - base)));
    }
  } else {
    pairs.reserve(npairs);
    for (size_t r = 0; r + 1 < n; r++)
      for (size_t c = r + 1; c < n; c++) pairs.emplace_back((uint32_t)r, (uint32_t)c);
  }
  std::atomic<size_t> next{0};
  std::atomic<int> err{0};
  auto worker = [&]() {
    size_t k;
    while ((k = next.fetch_add(1)) < pairs.size()) {
      size_t r = pairs[k].first, c = pairs[k].second;
      size_t idx = n * (n - 1) / 2 - (n - r) * (n - r - 1) / 2 + (c - r - 1);
      T d = T(0);
      int rc = dtw_pair<T>(seqs[r], lens[r], seqs[c], lens[c], W, os, oe, nullptr, &d, nullptr, nullptr);
      if (rc) err = rc;
      out[pair_subset ? k : idx] = d;
    }
  };
  if (nthreads < 1) nthreads = 1;
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; t++) th.emplace_back(worker);
  for (auto& t : th) t.join();
  (void)npairs;
  return err.load();
}

/*
 * Medoid of one all-in-one cluster (dba.hpp:186-207, 297-351): sum of squared
 * distances in T accumulated in the host loop order, first strict minimum;
 * two members -> the longer, the second on a tie (dba.hpp:336-339).
 */
template <typename T>
int medoid_single(const T* dists, const uint64_t* lens, size_t n, T* sos_out) {
  std::vector<T> sos(n, T(0));
  size_t idx = 0;
  for (size_t r = 0; r + 1 < n; r++) {
    for (size_t c = r + 1; c < n; c++) {
      T d = dists[idx + c - r - 1];
      d = (T)(d * d);
      sos[r] = (T)(sos[r] + d);
      sos[c] = (T)(sos[c] + d);
    }
    idx += n - r - 1;
  }
  if (sos_out) std::memcpy(sos_out, sos.data(), n * sizeof(T));
  int med = -1;
  if (n > 2) {
    T low = Lim<T>::max();
    for (size_t i = 0; i < n; i++)
      if (sos[i] < low) { med = (int)i; low = sos[i]; }
  } else if (n == 2) {
    med = lens[0] > lens[1] ? 0 : 1;
  } else {
    med = 0;
  }
  return med;
}

/*
 * performDBA<T> for a single cluster (cdist == 1, CLUSTER_AND_CONSENSUS;
 * dba.hpp:863-1189): sanitize, stable length sort, optional z-norm,
 * all-vs-all + medoid, up to max_rounds DBA rounds with the delta == 0 stop
 * and the open-mode flip-flop stop (dba.hpp:1084-1110), then the rescale by
 * the medoid's mu and sigma (dba.hpp:1117-1125). The flip-flop buffers start
 * zeroed (fresh pinned pages). Outputs: the final centroid (T, len = medoid
 * length), the medoid index in sorted order, the sort permutation and the
 * number of rounds run.
 */
template <typename T>
int perform_dba(T* const* in_seqs, const uint64_t* in_lens, size_t n, int os, int oe, int norm, int max_rounds,
                size_t medoid_w, int path_policy, uint64_t free_mem, T* centroid_out, uint64_t* centroid_len,
                int* medoid_sorted, int* perm_out, int* rounds_out, T* dists_out) {
  std::vector<std::vector<T>> seqs(n);
  std::vector<uint64_t> lens(n);
  for (size_t i = 0; i < n; i++) {
    lens[i] = in_lens[i];
    seqs[i].assign(in_seqs[i], in_seqs[i] + in_lens[i]);
    if (lens[i] && sanitize_drop_last<T>(seqs[i][lens[i] - 1])) lens[i]--;
  }
  std::vector<int> perm(n);
  for (size_t i = 0; i < n; i++) perm[i] = (int)i;
  std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return lens[a] < lens[b]; });
  std::vector<std::vector<T>> s(n);
  std::vector<uint64_t> l(n);
  for (size_t i = 0; i < n; i++) { s[i] = seqs[perm[i]]; s[i].resize(lens[perm[i]]); l[i] = lens[perm[i]]; }
  std::vector<double> mu(n, 0.0), sg(n, 1.0);
  if (norm)
    for (size_t i = 0; i < n; i++) znorm_one<T>(s[i].data(), l[i], &mu[i], &sg[i]);
  std::vector<const T*> ptr(n);
  for (size_t i = 0; i < n; i++) ptr[i] = s[i].data();
  std::vector<T> d(n * (n - 1) / 2);
  int rc = all_pairs<T>(ptr.data(), l.data(), n, os, oe, medoid_w, d.data(), (int)std::max(1u, std::thread::hardware_concurrency()), nullptr, 0);
  if (rc) return rc;
  if (dists_out) std::memcpy(dists_out, d.data(), d.size() * sizeof(T));
  int med = medoid_single<T>(d.data(), l.data(), n, nullptr);
  if (med < 0) return ODBA_E_MEDOID;
  const size_t lc = l[med];
  std::vector<T> cen(s[med]), nb(lc), prev(lc, T(0)), prev2(lc, T(0));
  int rounds = 0;
  for (int r = 0; r < max_rounds; r++) {
    double delta = 0;
    rc = dba_update<T>(cen.data(), lc, ptr.data(), l.data(), n, os, oe, path_policy, free_mem, nb.data(), nullptr, &delta);
    if (rc) return rc;
    rounds = r + 1;
    if (delta == 0) break;
    if (os || oe) {
      if (r >= 1 && !std::memcmp(nb.data(), prev2.data(), sizeof(T) * lc)) break;
      if (r > 1) prev2 = prev;
      if (r > 0) prev = nb;
    }
    cen = nb;
  }
  if (norm)
    for (size_t i = 0; i < lc; i++) nb[i] = d2t<T>(mu[med] + nb[i] * sg[med]);
  std::memcpy(centroid_out, nb.data(), lc * sizeof(T));
  if (centroid_len) *centroid_len = lc;
  if (medoid_sorted) *medoid_sorted = med;
  if (perm_out) std::memcpy(perm_out, perm.data(), n * sizeof(int));
  if (rounds_out) *rounds_out = rounds;
  return ODBA_OK;
}

/*
 * Adaptive k-segmentation (adaptive_device_segmentation<T>,
 * segmentation.hpp:48-337, one call per (sequence, chunk) block, and the
 * host post-pass of adaptive_segmentation, segmentation.hpp:486-578).
 * Thread-level steps are restated serially; the warp reductions are
 * simulated lane by lane (32-lane shuffle-down trees) because NaN values
 * produced by the downsample bound (segmentation.hpp:114) make their order
 * visible. lds_bytes is the dynamic shared memory size the reference asks
 * for (the device's opt-in maximum, segmentation.hpp:458-462): it fixes the
 * maximum segment extent e. Float -> unsigned char uses the GPU conversion
 * rule (round toward zero, saturate, NaN -> 0); inputs that reach it with
 * NaN or out of range are parity unpinned.
 */
template <typename T> T lim_min();
template <> int32_t lim_min<int32_t>() { return INT_MIN; }
template <> uint32_t lim_min<uint32_t>() { return 0; }
template <> uint64_t lim_min<uint64_t>() { return 0; }
template <> float lim_min<float>() { return FLT_MIN; }
template <> double lim_min<double>() { return DBL_MIN; }

template <typename T> T warp_reduce(const T* v32, bool is_max) {
  T a[32];
  for (int i = 0; i < 32; i++) a[i] = v32[i];
  for (int off = 16; off > 0; off /= 2) {
    T b[32];
    for (int i = 0; i < 32; i++) {
      T tmp = (i + off < 32) ? a[i + off] : a[i];
      T val = a[i];
      if (is_max) { if (tmp > val) val = tmp; }
      else { if (tmp < val) val = tmp; }
      b[i] = val;
    }
    std::memcpy(a, b, sizeof(a));
  }
  return a[0];
}

inline unsigned char d2u8(double d) {
  if (d != d) return 0;
  if (d <= 0.0) return 0;
  if (d >= 255.0) return 255;
  return (unsigned char)d;
}

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

/*
 * The Bellman k-segmentation DP of one chunk for one segment count k
 * (segmentation.hpp:172-190 tables, 194-262 recurrence and backtrack) on
 * the chunk's 8-bit levels q[0..N_ds):
 *   SUMS(l, r) / SQUARES(l, r): u16 / u32 running sums over spans r - l < e;
 *   DIST(l, r) = SQUARES - SUMS*SUMS / (r - l + 1) (1-based, int arithmetic);
 *   KD(1, n) = DIST(1, n) when n <= e, else INT_MAX / 2;
 *   KD(p, n) = min over the last segment's start idx in
 *              [max(n - e + 1, p), n] of KD(p - 1, idx - 1) + DIST(idx, n),
 *              the first strict minimum in ascending idx, and
 *   KP(p, n - 1) = idx - 1; the backtrack starts at bp[k] = N_ds.
 * run() returns KD(k, N_ds) and writes bp[0..k] (bp[0] = 0).
 */
struct KsegDP {
  int N_ds, e;
  std::vector<unsigned short> sums;
  std::vector<uint32_t> squares;
  std::vector<uint32_t> kdist, options;
  std::vector<unsigned short> kpath;
  KsegDP(const unsigned char* q, int n_ds, int e_) : N_ds(n_ds), e(e_) {
    sums.assign((size_t)std::max(1, N_ds) * std::max(1, e), 0);
    squares.assign((size_t)std::max(1, N_ds) * std::max(1, e), 0);
    for (int t = 0; t < N_ds; t++) {
      S(t, t) = q[t];
      Q(t, t) = (uint32_t)S(t, t) * (uint32_t)S(t, t);
    }
    for (int t = 0; t < N_ds; t++)
      for (int sp = 1; sp < e; sp++)
        if (t + sp < N_ds) {
          S(t, t + sp) = (unsigned short)(S(t, t + sp - 1) + S(t + sp, t + sp));
          Q(t, t + sp) = Q(t, t + sp - 1) + Q(t + sp, t + sp);
        }
    kdist.assign((size_t)2 * (N_ds + 1), 0);
    options.assign((size_t)N_ds + 1, 0);
  }
  unsigned short& S(int l, int r) { return sums[(size_t)(r - l) + (size_t)e * l]; }
  uint32_t& Q(int l, int r) { return squares[(size_t)(r - l) + (size_t)e * l]; }
  uint32_t DIST(int l, int r) {  // 1-based (segmentation.hpp:31-35)
    int a = l - 1, b = r - 1;
    int s = S(a, b);
    int ms = s * s / (b - a + 1);
    return Q(a, b) - (uint32_t)ms;
  }
  uint32_t& KD(int p, int n) { return kdist[(size_t)N_ds * (p % 2) + n]; }
  unsigned short& KP(int p, int n) { return kpath[(size_t)N_ds * (p - 1) + n]; }
  uint32_t run(int test_k, unsigned short* bp) {
    if (kpath.size() < (size_t)std::max(1, N_ds) * (size_t)(test_k + 1))
      kpath.resize((size_t)std::max(1, N_ds) * (size_t)(test_k + 1), 0);
    for (int t = 0; t < N_ds; t++) {
      KD(1, t + 1) = (t + 1 <= e) ? DIST(1, t + 1) : (uint32_t)(INT_MAX / 2);
      if (t < test_k) KP(t + 1, t) = (unsigned short)(test_k - 1);
    }
    for (int p = 2; p <= test_k; p++) {
      for (int n = p; n <= N_ds; n++) {
        for (int t = 0; t < N_ds; t++)
          if (t + 1 >= p && t + 1 <= n && n - t - 1 < e) options[t + 1] = KD(p - 1, t) + DIST(t + 1, n);
        int minval = INT_MAX, minidx = -1;
        for (int idx = (n - e + 1 > p) ? n - e + 1 : p, start = idx; idx - start < e && idx <= n; idx++)
          if (options[idx] < (uint32_t)minval) { minval = (int)options[idx]; minidx = idx; }
        KD(p, n) = (uint32_t)minval;
        KP(p, n - 1) = (unsigned short)((unsigned short)minidx - 1);
      }
    }
    bp[test_k] = (unsigned short)N_ds;
    for (int p = test_k - 1; p >= 1; p--) {
      int n = bp[p + 1] - 1;
      bp[p] = (n >= 0 && n < N_ds) ? KP(p + 1, n) : 0;
    }
    bp[0] = 0;
    return N_ds >= 1 ? KD(test_k, N_ds) : 0;
  }
};

template <typename T>
void kseg_block(const T* series, int orig_N, int chunk, int spb, int mk, int minseg, int lds_bytes,
                T* out /* mk slots of this chunk */) {
  const int ds = (minseg + 2) / 3;
  const int N = ((chunk + 1) * spb > orig_N) ? orig_N % spb : spb;
  int expected_k = mk;
  if ((chunk + 1) * spb > orig_N) expected_k = ((orig_N % spb) + ((spb + mk - 1) / mk) - 1) / ((spb + mk - 1) / mk);
  const int N_ds = N / ds;
  const T* ser = series + (size_t)chunk * spb;
  // downsample (segmentation.hpp:112-121), including the +threadIdx.x bound
  std::vector<T> dq(std::max(N_ds, 1), T(0));
  for (int t = 0; t < 256; t++) {
    if (t * ds <= N - ds) {
      T acc = T(0);
      int cnt = 0;
      for (int i = 0; i < ds && ds * t + t + i < N; i++) { acc = (T)(acc + ser[ds * t + i]); cnt++; }
      if (std::is_floating_point<T>::value) dq[t] = (T)(acc / (T)cnt);
      else dq[t] = cnt ? (T)(acc / (T)cnt) : T(0);  /* integer 0/0 has no defined value: parity unpinned */
    }
  }
  // min / max with warp trees over full warps only (segmentation.hpp:126-150)
  const int nw = N_ds / 32;
  std::vector<T> wmax(8), wmin(8);
  for (int w = 0; w < 8; w++) {
    T vmx[32], vmn[32];
    for (int l = 0; l < 32; l++) {
      int t = w * 32 + l;
      vmx[l] = t < N_ds ? dq[t] : lim_min<T>();
      vmn[l] = t < N_ds ? dq[t] : Lim<T>::max();
    }
    wmax[w] = warp_reduce(vmx, true);
    wmin[w] = warp_reduce(vmn, false);
  }
  T bmx[32], bmn[32];
  for (int l = 0; l < 32; l++) {
    bmx[l] = l < nw ? wmax[l] : lim_min<T>();
    bmn[l] = l < nw ? wmin[l] : Lim<T>::max();
  }
  const T Tmax = warp_reduce(bmx, true);
  const T Tmin = warp_reduce(bmn, false);
  // layout of the dynamic shared memory (segmentation.hpp:93-107, 158-165)
  size_t off = (size_t)(mk + 1) * 2;                  // breakpoints
  off = align_up(off, 4) + (size_t)4 * (N_ds + 1);    // options
  off += (size_t)4 * 2 * (N_ds + 1);                  // k_seg_dist
  off = align_up(off, sizeof(T)) + sizeof(T) * (size_t)(N_ds + 2);  // downsample_qtype + max/min
  off = align_up(off, 4);                             // squares start
  const int e = N_ds > 0 ? (int)(((long long)lds_bytes - (long long)off) / ((long long)N_ds * 6)) : 0;
  // 8-bit rescale (segmentation.hpp:166-170)
  std::vector<unsigned char> q((size_t)std::max(1, N_ds), 0);
  for (int t = 0; t < N_ds; t++) {
    /* quotient in T: integer division for integer T (segmentation.hpp:168) */
    const T qt = (T)((T)(dq[t] - Tmin) / (T)(Tmax - Tmin + 1));
    q[t] = d2u8(256.0 * (double)qt);
  }
  KsegDP dp(q.data(), N_ds, e);
  std::vector<unsigned short> bp((size_t)mk + 1, 0);
  bool ex = false;
  short smallest_noisy = (short)(expected_k + 1);
  short test_k = (short)((expected_k + 1) / 2);
  short delta = (short)(test_k / 2);
  while (!ex) {
    dp.run(test_k, bp.data());
    bool noisy = false;
    for (int i = 1; i <= test_k; i++)
      if ((int)bp[i] - (int)bp[i - 1] < minseg / ds) noisy = true;
    if (delta == 0 || test_k + delta > mk) ex = true;
    else if (test_k != 1 && noisy) { smallest_noisy = test_k; test_k = (short)(test_k - delta); }
    else {
      if (test_k == 1 || smallest_noisy == test_k + 1) ex = true;
      else test_k = (short)(test_k + delta);
    }
    delta = (short)((delta + 1) / 2);
    if (ex) {
      std::vector<T> orig((size_t)spb + 2, T(0));
      for (int t = 0; t < 256; t++)
        if (t * ds + chunk * spb < orig_N)
          for (int i = 0; i < ds && t * ds + i < spb && t * ds + chunk * spb + i < orig_N; i++)
            orig[(size_t)t * ds + i] = ser[(size_t)t * ds + i];
      for (int t = 1; t <= 255; t++) {
        if (t > test_k && t <= mk) { out[t - 1] = Lim<T>::max(); }
        else if (t > 0 && t <= test_k) {
          int right = bp[t] * ds;
          if (right > N) right = N;
          int left = bp[t - 1] * ds;
          for (int i = left; i < right - 1; ++i)
            for (int j = i + 1; j < right; ++j)
              if (orig[i] > orig[j]) std::swap(orig[i], orig[j]);
          if ((left - right) % 2) out[t - 1] = orig[(left + right) / 2];
          else out[t - 1] = (T)((orig[(left + right) / 2] + orig[(left + right) / 2 + 1]) / T(2));
        }
      }
    }
  }
}

template <typename T>
int kseg(const T* const* seqs, const uint64_t* lens, size_t n, int minseg, int prefix_skip, int lds_bytes,
         T* out, uint64_t* out_lens, uint64_t* out_offs) {
  if (minseg <= 0) return ODBA_E_INVALID;
  const int ds = (minseg + 2) / 3;
  const int spb = 256 * ds;
  const int mk = (int)((256 + ((float)minseg) / ds - 1) / (((float)minseg) / ds));
  uint64_t cur = 0;
  for (size_t s = 0; s < n; s++) {
    const int L = (int)lens[s];
    const int nch = (L + spb - 1) / spb;
    std::vector<T> seg((size_t)nch * mk, T(0));
    for (int ch = 0; ch < nch; ch++) kseg_block<T>(seqs[s], L, ch, spb, mk, minseg, lds_bytes, &seg[(size_t)ch * mk]);
    // host post-pass (segmentation.hpp:486-578)
    const T MX = Lim<T>::max();
    T eps = MX;
    T prev = seg[0];
    const size_t SL = seg.size();
    for (size_t j = 1; j < SL; ++j) {
      T cv = seg[j];
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
    std::memcpy(out + cur, seg.data() + skipped, (size_t)cursor * sizeof(T));
    out_offs[s] = cur;
    out_lens[s] = (uint64_t)cursor;
    cur += (uint64_t)cursor;
  }
  return ODBA_OK;
}

}  // namespace

#define ORC_DISPATCH(dtype, CALL)                            \
  switch (dtype) {                                           \
    case ODBA_I32: { typedef int32_t T; return CALL; }       \
    case ODBA_U32: { typedef uint32_t T; return CALL; }      \
    case ODBA_U64: { typedef uint64_t T; return CALL; }      \
    case ODBA_F32: { typedef float T; return CALL; }         \
    case ODBA_F64: { typedef double T; return CALL; }        \
    default: return ODBA_E_INVALID;                          \
  }

extern "C" {

int orc_dtw_pair(int dtype, const void* y, uint64_t ly, const void* x, uint64_t lx, uint64_t w, int os, int oe,
                 uint8_t* moves, void* dist, int* abrogated_swath, int* q3_flag) {
  ORC_DISPATCH(dtype, dtw_pair<T>((const T*)y, ly, (const T*)x, lx, w, os, oe, moves, (T*)dist, abrogated_swath, q3_flag))
}

/* Test hook: dtw_pair with the open-end abrogation check switched off
 * (abrogate = 0), everything else as orc_dtw_pair. */
int orc_dtw_pair_abr(int dtype, const void* y, uint64_t ly, const void* x, uint64_t lx, uint64_t w, int os, int oe,
                     int abrogate, uint8_t* moves, void* dist, int* abrogated_swath) {
  ORC_DISPATCH(dtype, dtw_pair<T>((const T*)y, ly, (const T*)x, lx, w, os, oe, moves, (T*)dist, abrogated_swath,
                                  (int*)nullptr, abrogate != 0))
}

int orc_all_pairs(int dtype, const void* const* seqs, const uint64_t* lens, uint64_t n, int os, int oe, uint64_t w,
                  void* out, int nthreads, const int64_t* subset, uint64_t nsubset) {
  ORC_DISPATCH(dtype, all_pairs<T>((const T* const*)seqs, lens, n, os, oe, w, (T*)out, nthreads, subset, nsubset))
}

int orc_traceback_full(int dtype, const uint8_t* moves, uint64_t rows, uint64_t cols, int flip, const void* seq,
                       void* sums, uint32_t* counts, uint64_t* path_len) {
  size_t pl = 0;
  int rc;
  switch (dtype) {
    case ODBA_I32: rc = traceback_full<int32_t>(moves, rows, cols, flip, (const int32_t*)seq, (int32_t*)sums, counts, &pl); break;
    case ODBA_U32: rc = traceback_full<uint32_t>(moves, rows, cols, flip, (const uint32_t*)seq, (uint32_t*)sums, counts, &pl); break;
    case ODBA_U64: rc = traceback_full<uint64_t>(moves, rows, cols, flip, (const uint64_t*)seq, (uint64_t*)sums, counts, &pl); break;
    case ODBA_F32: rc = traceback_full<float>(moves, rows, cols, flip, (const float*)seq, (float*)sums, counts, &pl); break;
    case ODBA_F64: rc = traceback_full<double>(moves, rows, cols, flip, (const double*)seq, (double*)sums, counts, &pl); break;
    default: return ODBA_E_INVALID;
  }
  if (path_len) *path_len = pl;
  return rc;
}

int orc_dba_update(int dtype, const void* C, uint64_t lc, const void* const* seqs, const uint64_t* lens, uint64_t n,
                   int os, int oe, int path_policy, uint64_t free_mem, void* out, uint32_t* counts, double* delta) {
  ORC_DISPATCH(dtype, dba_update<T>((const T*)C, lc, (const T* const*)seqs, lens, n, os, oe, path_policy, free_mem,
                                    (T*)out, counts, delta))
}

int orc_znorm(int dtype, void* const* seqs, const uint64_t* lens, uint64_t n, double* mu, double* sigma) {
  ORC_DISPATCH(dtype, (([&]() {
                 for (size_t i = 0; i < n; i++) znorm_one<T>((T*)seqs[i], lens[i], mu ? &mu[i] : nullptr, sigma ? &sigma[i] : nullptr);
                 return (int)ODBA_OK;
               })()))
}

int orc_medoid_single(int dtype, const void* dists, const uint64_t* lens, uint64_t n, void* sos_out) {
  ORC_DISPATCH(dtype, medoid_single<T>((const T*)dists, lens, n, (T*)sos_out))
}

int orc_perform_dba(int dtype, void* const* seqs, const uint64_t* lens, uint64_t n, int os, int oe, int norm,
                    int max_rounds, uint64_t medoid_w, int path_policy, uint64_t free_mem, void* centroid_out,
                    uint64_t* centroid_len, int* medoid_sorted, int* perm_out, int* rounds_out, void* dists_out) {
  ORC_DISPATCH(dtype, perform_dba<T>((T* const*)seqs, lens, n, os, oe, norm, max_rounds, medoid_w, path_policy, free_mem,
                                     (T*)centroid_out, centroid_len, medoid_sorted, perm_out, rounds_out, (T*)dists_out))
}

int orc_prefix_chop(int dtype, const void* prefix, uint64_t P, const void* seq, uint64_t L, uint64_t* chop,
                    int32_t* hist) {
  ORC_DISPATCH(dtype, prefix_chop<T>((const T*)prefix, P, (const T*)seq, L, chop, hist))
}

/* out must hold sum over sequences of mk * ceil(L / (256 * ds)) values. */
/* Test hook: the k-segmentation DP of one chunk (KsegDP) on given 8-bit
 * levels q[0..n_ds) with extent e and segment count k; bp needs k + 1 slots. */
int orc_kseg_dp(const unsigned char* q, int n_ds, int e, int k, uint32_t* cost, unsigned short* bp) {
  if (n_ds < 1 || e < 1 || k < 1 || k > n_ds) return ODBA_E_INVALID;
  KsegDP dp(q, n_ds, e);
  *cost = dp.run(k, bp);
  return ODBA_OK;
}

int orc_kseg(int dtype, const void* const* seqs, const uint64_t* lens, uint64_t n, int minseg, int prefix_skip,
             int lds_bytes, void* out, uint64_t* out_lens, uint64_t* out_offs) {
  ORC_DISPATCH(dtype, kseg<T>((const T* const*)seqs, lens, n, minseg, prefix_skip, lds_bytes, (T*)out, out_lens, out_offs))
}

}  // extern "C"
