// Internal declarations shared by the HIP kernel files and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/odba/odba_types.h"

#define ODBA_W 1024  // swath width; equals the reference's blockDim.x (dba.hpp:41-48, 453-457)

namespace odba {

// Padded element count of a stored sequence: every swath may read one full
// 1024-wide column block and the row window 128 entries past the end.
inline long long padded_len(long long L) { return ((L + 63) / 64) * 64 + 1024; }
// Positions of one cluster in a packed accumulation: lc rounded up to a
// whole 64-lane block.
inline long long round64(long long L) { return (L + 63) / 64 * 64; }

// Waves of a path-strip or swath-band workgroup at most: 16 caps a path
// kernel at 128 VGPRs, four waves per SIMD. With the step bits in two
// planes (plane_code) the whole-swath variants fit that without spills
// (2-bit code packing needed 160-168 VGPRs and spilled 46-74 at this cap).
constexpr int kPathStripWaves = 16;

// Path step words. A lane's word for one step holds its C <= 16 cells in
// two bit planes: bit 15 - q of the low half is set when column q did not
// take the diagonal, bit 31 - q of the high half when its up candidate
// holds the minimum. The reference picks DIAG, then UP, then RIGHT
// (dtw.hpp:202-227), so the code is DIAG when the low bit is clear, else UP
// when the high bit is set, else RIGHT. Returns 0 DIAG, 1 RIGHT, 2 UP.
// (The open-end top row's own words keep 2-bit codes at bits 2q, 2q + 1.)
__host__ __device__ inline uint32_t plane_code(uint32_t w, int q) {
  const uint32_t a = (w >> (15 - q)) & 1u, b = (w >> (31 - q)) & 1u;
  return a ? (b ? 2u : 1u) : 0u;
}

inline int last_swath_c_host(int wcols) {
  if (wcols <= 64) return 1;
  if (wcols <= 128) return 2;
  if (wcols <= 256) return 4;
  if (wcols <= 512) return 8;
  if (wcols <= 768) return 12;
  return 16;
}

// Rows covered by the skewed main loop of a swath (the rest is positional
// or the open-end top-row scan).
inline int main_rows(int ly, int os, int oe) { return ((oe && ly >= 2) ? ly - 2 : ly - 1) - (os ? 1 : 0) + 1; }
inline long long swath_words(int ly, int os, int oe) {
  int nr = main_rows(ly, os, oe);
  return nr > 0 ? (long long)(nr + 63) * 64 : 0;
}

// order (optional): npairs int2 {r, c} (r < c) giving the packing order;
// null: second index descending, then first index descending.
hipError_t launch_all_pairs(int dtype, bool minfirst, int os, int oe, int nblocks, hipStream_t st, const void* data,
                            const long long* offs, const int* lens, int n, long long npairs,
                            unsigned long long* counter, void* dist, void* esc, long long stride, int* flags,
                            const void* order);
// Small pairs (float, every length 64..1024): a workgroup of nw waves per
// pair, C = 1, 2 or 4 columns per lane, nw * 64 * C >= the longest length,
// sub-blocks of B = 16 or 32 steps between barriers. counter == nullptr:
// one workgroup per pair (nblocks >= npairs); otherwise persistent
// workgroups claim pairs from *counter (zeroed by the caller).
// The pairs in launch order, one record each: element offsets of Y and X in
// the packed data, the pair's index in the distance array, the lengths.
struct SmallPairRec {
  long long oy, ox, idx;
  int ly, lx;
};
hipError_t launch_all_pairs_small(int C, int B, int nw, int os, int oe, int nblocks, hipStream_t st,
                                  const float* data, const SmallPairRec* tab, long long npairs,
                                  unsigned long long* counter, float* dist, int* flags);
hipError_t launch_path_jobs(int dtype, bool minfirst, int os, int oe, int nblocks, hipStream_t st, const void* jobs,
                            int njobs, unsigned int* counter, uint32_t* mv, uint32_t* mtop, int* abr, void* esc,
                            long long stride);
size_t path_job_size(int dtype);

// Traceback job: geometry of one aligned pair in path mode.
struct TbJob {
  int ly, lx;          // Y rows, X columns
  int flip;            // 1: centroid on Y (dba.hpp:497-501)
  int seq;             // index of the sequence in the cluster order
  long long mv_off;    // word offset of the pair's step words
  long long top_off;   // word offset of the open-end top-row words
  long long path_off;  // -1, or the slot in the path buffer: [count][steps...]
  long long runs_off;  // element offset of this job's run records (one per centroid position)
  int4* rp;            // when set: this job's run records (runs_off unused)
};

// One visited cell of a recorded traceback: row i (31 bits), column j
// (30 bits), step code (3 bits; 0 = NIL, otherwise the reference code 1..5).
__host__ __device__ inline unsigned long long pack_step(int i, int j, int code) {
  return ((unsigned long long)(unsigned)i << 33) | ((unsigned long long)(unsigned)j << 3) |
         (unsigned long long)(code == 255 ? 0 : code);
}

// Per-job traceback state of stripe mode, carried across stripes.
struct StripeState {
  int i, j;      // current cell (global column)
  int h;         // current truncated height (rows still in play)
  int curc, lo, hi;  // open runs of the current centroid position
  int lo2, hi2, gap;
  int npath;     // recorded traceback steps so far
  int fab;       // swath where the forward pass abandoned the search, -1 none
  int sab;       // 1: this stripe's top row is all OPEN_RIGHT
  int done, err;
  int br;        // checkpointed stripes: first row of the block last recomputed
};

// Swaths per claim of the checkpointed open-mode stripe forward pass (one
// wave each).
constexpr int kSckBandSwaths = 16;

// Checkpointed stripe paths in the open modes (float): one record per job.
// slot_off: K-1 edge columns of H + 256 floats (every swath's right edge, as
// the reference's dtwCostSoFar); rowck_off: nb-1 checkpoint rows of LX
// floats (nb = row bands of 2^rshift); tc_off: K (top, column min) pairs of
// the forward pass; pm_off: K*nb prefix minima of the edge columns at the
// band starts; cnt_off / claim0: progress words and first claim of the
// job's column blocks of 16 swaths; mv_off: words of one recomputed block.
struct SckJob {
  const float* y;
  const float* x;
  int H, LX;
  long long slot_off, rowck_off, tc_off, pm_off, cnt_off, claim0, mv_off;
};

// A few very long pairs: bands of `waves` strips (64*C columns each) claimed
// by workgroups in pair order. tab: one {int r, c; long long band0} per pair
// in claim order (band0 = first claim index of the pair), nclaims bands in
// all. rings: nblocks*waves*ring elements (ring a power of two >= 128);
// gslots: R >= 2*nblocks+2 slots of gstride elements (>= the longest Y +
// 256); sync: band_sync_bytes(R, nclaims, ntab) bytes; err: bit 8 set when a
// hand-off wait timed out after tick_limit wall-clock ticks. flags
// (optional) as for launch_all_pairs.
hipError_t launch_all_pairs_bands(int dtype, int C, int waves, int os, int oe, int nblocks, hipStream_t st,
                                  const void* data, const long long* offs, const int* lens, int n, const void* tab,
                                  int ntab, long long nclaims, unsigned long long* counter, void* dist, void* rings,
                                  int ring, void* gslots, long long gstride, void* sync, int R, int* err,
                                  long long tick_limit, int* flags);
size_t band_sync_bytes(int R, long long nclaims, int ntab);
size_t band_pair_size();

// Float path jobs with a workgroup of `waves` waves per job (strips = swaths).
// Path jobs as bands of quarter-swath strips over the whole device (float);
// njobs <= 1024, max_claims >= total swaths; timeouts set err bit 8.
size_t path_band_sync_bytes(int R, long long max_claims, int njobs);
hipError_t launch_path_bands(int os, int oe, int nblocks, hipStream_t st, const void* jobs, int njobs,
                             long long max_claims, unsigned long long* counter, uint32_t* mv, uint32_t* mtop,
                             int* abr, void* rings, int ring, void* gslots, long long gstride, void* sync, int R,
                             int* err, long long tick_limit);
// Path jobs as bands of nw whole swaths over the whole device (float);
// max_claims >= total bands; timeouts set err bit 8. In open end without
// open start, bslots holds njobs boundary edge columns of gstride values
// (null otherwise).
hipError_t launch_path_swath_bands(int nw, int os, int oe, int nblocks, hipStream_t st, const void* jobs, int njobs,
                                   long long max_claims, unsigned long long* counter, uint32_t* mv, uint32_t* mtop,
                                   int* abr, void* rings, int ring, void* gslots, long long gstride, void* sync, int R,
                                   int* err, long long tick_limit, void* bslots);
hipError_t launch_path_strips(int waves, int os, int oe, int nblocks, hipStream_t st, const void* jobs, int njobs,
                              unsigned int* counter, uint32_t* mv, uint32_t* mtop, int* abr, void* esc,
                              long long stride, int quarters);

// Checkpointed global float paths (see k_ckpt_forward in the swath kernels):
// jobs are CkJob records (ckpt_job_size() bytes), states CkState records
// (ckpt_state_size() bytes); ew = strips of 1024 columns per column block
// (2 or 16), rows per band = 2^rshift.
hipError_t launch_ckpt_forward(int ew, int nblocks, hipStream_t st, const void* jobs, int njobs, long long nclaims,
                               unsigned long long* counter, float* rowck, float* colck, int* cnt, float* rings,
                               int ring, int rshift, int* err, long long tick_limit);
hipError_t launch_ckpt_block(int ew, hipStream_t st, const void* jobs, int njobs, void* states, const float* rowck,
                             const float* colck, uint32_t* mv, long long job_words, float* slots,
                             long long slot_stride, int rshift);
hipError_t launch_ckpt_trace(hipStream_t st, const void* jobs, int njobs, void* states, const uint32_t* mv,
                             long long job_words, int4* runs, int rshift, int ew, unsigned long long* pbuf);
size_t ckpt_job_size();
size_t ckpt_state_size();

// Checkpointed stripe paths in the open modes (float; see k_sck_forward):
// the forward pass over the whole device, then per step the block of the
// path's current cell and its traceback, until every path is done.
hipError_t launch_sck_forward(int os, int oe, int nblocks, hipStream_t st, const SckJob* jobs, int njobs,
                              long long nclaims, unsigned long long* counter, float* slots, float* rowck, float2* tc,
                              int* cnt, int rshift, int* err, long long tick_limit);
hipError_t launch_sck_init(int os, int oe, hipStream_t st, const SckJob* jobs, int njobs, int kmax, const float* slots,
                           const float2* tc, float* pm, StripeState* ss, int rshift);
hipError_t launch_sck_step(int os, int oe, hipStream_t st, const SckJob* jobs, int njobs, const float* slots,
                           const float* rowck, const float* pm, uint32_t* mv, uint32_t* mtop, float* qslots,
                           long long qstride, StripeState* ss, int rshift);
hipError_t launch_sck_traceback(hipStream_t st, const TbJob* jobs, int njobs, const uint32_t* mv,
                                const uint32_t* mtop, StripeState* ss, int4* runs, int os, int oe, int* err,
                                unsigned long long* pbuf);

hipError_t launch_stripe(int dtype, bool minfirst, bool forward, int os, int oe, int nblocks, hipStream_t st,
                         const void* jobs, int njobs, int m, unsigned int* counter, void* slots, uint32_t* mv,
                         uint32_t* mtop, StripeState* ss);
hipError_t launch_stripe_traceback(hipStream_t st, const TbJob* jobs, int njobs, int m, const uint32_t* mv,
                                   const uint32_t* mtop, StripeState* ss, int4* runs, int lc, int os, int oe,
                                   int* err, unsigned long long* pbuf);

hipError_t launch_traceback(hipStream_t st, const TbJob* jobs, int njobs, const uint32_t* mv, const uint32_t* mtop,
                            const int* abr, int4* runs, int lc, int os, int oe, int* err, unsigned long long* pbuf);
hipError_t launch_accumulate(int dtype, hipStream_t st, const void* data, const long long* offs, const int4* runs,
                             int nseq, int lc, void* acc, uint32_t* counts, int first_batch);
hipError_t launch_mean_delta(int dtype, hipStream_t st, const void* acc, const uint32_t* counts, const void* old_c,
                             void* new_c, int lc, unsigned long long* delta_bits, int* err);
// One active cluster of a packed multi-cluster DBA round.
struct MultiSlot {
  long long pos0;            // first global centroid position of the cluster
  const void* data;          // its sequence data (padded, packed)
  const long long* moffs;    // its members' offsets into data, in cluster order
  const int4* runs;          // its run records: [member][position]
  const void* cen;           // its current centroid
  void* newc;                // its next centroid
  int lc, nmem, slot;        // centroid length, members, index of its delta word
  int pad;
};
hipError_t launch_multi_mean(int dtype, hipStream_t st, const MultiSlot* sl, int ns, long long total, void* acc,
                             uint32_t* counts, unsigned long long* delta_bits, int* err);

hipError_t launch_fill_u32(hipStream_t st, void* p, uint32_t v, long long n);
hipError_t launch_znorm(int dtype, hipStream_t st, void* data, const long long* offs, const int* lens, int n,
                        double* mu, double* sigma);
hipError_t launch_decode_moves(hipStream_t st, const uint32_t* mv, const uint32_t* mtop, int ab, int ly, int lx, int os,
                               int oe, uint8_t* out);

hipError_t launch_kseg(int dtype, hipStream_t st, const void* data, const long long* offs, const int* lens, int n,
                       int maxch, const long long* out_base, const long long* path_base, int spb, int mk, int minseg,
                       int lds_bytes, unsigned short* kpath, void* out, size_t dyn);
int kseg_extent_host(int mk, int N_ds, int tsize, int lds_bytes);

}  // namespace odba
