// Setup-phase kernels of BoomerAMG (PMIS + extended+i + Galerkin RAP).
//
// The hierarchy must match hypre's host path bit for bit (C/F split, sparsity
// patterns, and -- because truncation breaks ties by an unstable quicksort on
// |w| -- the exact floating-point values).  Every per-row quantity is therefore
// accumulated in the reference's traversal order:
//   * one wavefront owns one row; the order-defining traversal (S row, A row,
//     R row, RA row) is a uniform loop, and the work inside one step (an
//     entire neighbour row) is spread over the 64 lanes, whose updates always
//     touch distinct accumulators, so no two updates of one accumulator are
//     ever reordered;
//   * first-seen column order is kept with a per-wave LDS hash table whose
//     new entries get positions from a ballot prefix count in lane order;
//   * ordered sums (ext+i denominators) are folded left-to-right with
//     readlane over the ballot of contributing lanes.
// The file is compiled with -ffp-contract=off.  Row tables that would exceed
// the LDS capacity of a launch are deferred to a launch with larger tables.
#include "hx_internal.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <climits>

namespace hx {

constexpr int WAVE = 64;

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ unsigned long long lanemask_lt(int lane)
{
   return lane ? ((~0ULL) >> (64 - lane)) : 0ULL;
}

__device__ __forceinline__ double readlane_d(double v, int src)
{
   long long b = __double_as_longlong(v);
   int lo = __builtin_amdgcn_readlane((int) (b & 0xffffffffLL), src);
   int hi = __builtin_amdgcn_readlane((int) (b >> 32), src);
   return __longlong_as_double(((long long) hi << 32) | (unsigned int) lo);
}

__device__ __forceinline__ int readlane_i(int v, int src) { return __builtin_amdgcn_readlane(v, src); }

constexpr __host__ __device__ int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v >> 1); }

// Slot of column c in an LDS table of 2^bits slots for columns that come in
// runs of consecutive indices: a run takes consecutive slots (its flushes add
// to consecutive counters), and each 2^bits-aligned block of columns is
// rotated by 3077 slots, so runs a multiple of the table size apart (grid
// planes of 320^2 = 25 * 4096 points, say) do not pile onto one slot.  With
// 3077 no two columns of a 256-row block of the 7- or 27-point operators of
// the 200^3 and 320^3 grids share a slot (checked on the host for the
// rotations tried; the low bits alone collide on 2/3 of the keys at 320^3).
__device__ __forceinline__ unsigned run_slot(unsigned c, int bits)
{
   return (c + (c >> bits) * 3077u) & ((1u << bits) - 1u);
}

__device__ __forceinline__ void wave_sync() { __syncthreads(); }   // blocks are one wave

static int launch_blocks(long long work_items)
{
   long long b = std::min<long long>(std::max<long long>(work_items, 1), 256 * 32);
   if (b >= 64) { b = b / 8 * 8; }
   return (int) b;
}

// HYPRE_MI355X_VERBOSE=1 prints how many rows each setup stage hands on
static bool setup_verbose()
{
   static const bool v = [] {
      const char *e = std::getenv("HYPRE_MI355X_VERBOSE");
      return e && std::atoi(e) > 0;
   }();
   return v;
}

// one-wave blocks that all fit on the chip at once (256 CUs, 160 KB LDS and
// 32 wave slots per CU), so that grid-strided rows advance together
static int resident_blocks(long long work_items, size_t lds_per_block)
{
   long long per_cu = lds_per_block ? std::min<long long>(32, 163840 / (long long) lds_per_block) : 32;
   long long b = std::min<long long>(std::max<long long>(work_items, 1), 256 * std::max<long long>(per_cu, 1));
   if (b >= 64) { b = b / 8 * 8; }
   return (int) b;
}

__device__ __forceinline__ int xcd_blk(int bid, int nb)
{
   return (nb & 7) ? bid : (bid & 7) * (nb >> 3) + (bid >> 3);
}

// XCD (accelerator complex die) the wave runs on, 0..7
__device__ __forceinline__ int xcc_id()
{
   unsigned v;
   asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
   return (int) (v & 7);
}

// Work items handed out in order from eight counters, one per XCD: XCD x
// owns items [x*chunk, (x+1)*chunk) and its waves claim them one at a time,
// so the rows in flight on an XCD form a narrow window of consecutive rows
// (their neighbour rows stay in that XCD's L2) however unevenly rows cost.
// A wave whose range is used up claims from the next XCD's range.  Every wave
// ends after at most eight failed claims.  ctr: 8 counters 32 ints apart,
// zero at launch.  Called by all lanes of the wave together.
struct RowClaim
{
   int *ctr;
   int  n, chunk, x, t, v;

   __device__ RowClaim(int *c, int nwork) : ctr(c), n(nwork), chunk((nwork + 7) >> 3), x(xcc_id()), t(0), v(0) {}

   // start claiming the next item; the result is read by take()
   __device__ void issue(int lane)
   {
      v = 0;
      if (lane == 0) { v = atomicAdd(ctr + ((x + t) & 7) * 32, 1); }
   }

   // the claimed item, or -1 once every range is used up
   __device__ int take(int lane)
   {
      for (;;) {
         const int cx = (x + t) & 7;
         const int lo = cx * chunk;
         const int got = readlane_i(v, 0);
         if (got < min(chunk, n - lo)) { return lo + got; }
         if (++t >= 8) { return -1; }
         issue(lane);
      }
   }
};

static int *claim_counters()
{
   int *c = device_flags() + 512;
   dzero(c, 256 * sizeof(int));
   return c;
}

// inclusive prefix sum over the 64 lanes of a wave
__device__ __forceinline__ int wave_incl_scan(int v, int lane)
{
   for (int d = 1; d < WAVE; d <<= 1) {
      const int t = __shfl_up(v, d, WAVE);
      if (lane >= d) { v += t; }
   }
   return v;
}

// last index s in [0, nseg) with off[s] <= q (off non-decreasing, off[0] = 0 <= q)
__device__ __forceinline__ int seg_find(const int *off, int nseg, int q)
{
   int lo = 0, hi = nseg - 1;
   while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (off[mid] <= q) { lo = mid; } else { hi = mid - 1; }
   }
   return lo;
}

// lanes per row for row-parallel kernels, from the mean row length
static int lanes_per_row(long long nnz, int n)
{
   double m = n > 0 ? (double) nnz / n : 0.0;
   if (m <= 6) { return 4; }
   if (m <= 32) { return 8; }
   if (m <= 64) { return 16; }
   return 32;
}

// the PMIS rounds and the strength recount do little per entry: more rows
// per wave pay off (measured on the 27-point fine level)
static int lanes_per_row_light(long long nnz, int n)
{
   double m = n > 0 ? (double) nnz / n : 0.0;
   if (m <= 32) { return 4; }
   return lanes_per_row(nnz, n);
}

template <class F>
static void dispatch_k(int K, F f)
{
   switch (K) {
      case 4: f(std::integral_constant<int, 4>{}); break;
      case 8: f(std::integral_constant<int, 8>{}); break;
      case 16: f(std::integral_constant<int, 16>{}); break;
      default: f(std::integral_constant<int, 32>{}); break;
   }
}

// Left-to-right sum of the values held by the K lanes of a group, one value
// per lane and step; `acc` of every lane of the group receives the fold.
template <int K>
__device__ __forceinline__ double group_fold(double acc, double v, bool valid)
{
   for (int l = 0; l < K; l++) {
      const double x = __shfl(v, l, K);
      const int ok = __shfl((int) valid, l, K);
      if (ok) { acc += x; }
   }
   return acc;
}

__global__ void k_rowlen_max(int n, const int *__restrict__ ia, int *__restrict__ out)
{
   int m = 0;
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      m = max(m, ia[i + 1] - ia[i]);
   }
   for (int d = 32; d > 0; d >>= 1) { m = max(m, __shfl_xor(m, d, WAVE)); }
   if ((threadIdx.x & 63) == 0 && m > 0) { atomicMax(out, m); }
}

static int max_row_len(const Csr &A)
{
   int *flag = device_flags() + 56;
   dzero(flag, sizeof(int));
   int nb = std::max(1, std::min((A.nrows + 255) / 256, 2048));
   k_rowlen_max<<<nb, 256, 0, stream()>>>(A.nrows, A.ia, flag);
   HX_LAUNCH_CHECK();
   int h = 0;
   d2h(&h, flag, sizeof(int));
   return h;
}

// ================================================================== strength
// src/parcsr_ls/par_strength.c:75-525 (CreateSHost, num_functions 1, one rank):
// row_scale is a min/max (order-free), the row sum is folded in column order,
// S keeps the column order of A.
// ------------------------------------------------------------------ strength, fused passes
// k_row_stats: one row per lane, 64 rows per wave, the values streamed in
// coalesced windows of RS_SW entries through LDS (as the solve-phase SpMV);
// each lane walks its row in storage order and produces what k_strength and
// k_l1 derive per row: the strength threshold theta * row_scale with the
// diagonal-sign and all-weak flags (par_strength.c:260-330), the l1 norm
// (hypre_ParCSRComputeL1Norms option 1 with the sign of the diagonal), and,
// when the row lies in one window, its number of strong entries (-1 when it
// does not: k_strength_recount counts those rows).
constexpr int RS_SW = 1024;
constexpr int RS_WPB = 4;

__global__ __launch_bounds__(RS_WPB * WAVE) void k_row_stats(int n, const int *__restrict__ ia,
                                                               const double *__restrict__ a,
                                                               double theta, double mrs,
                                                               double *__restrict__ thr,
                                                               int *__restrict__ flg,
                                                               int *__restrict__ cnt,
                                                               double *__restrict__ l1,
                                                               int *__restrict__ l1_zero)
{
   constexpr int U = RS_SW / WAVE;
   __shared__ double sb_all[RS_WPB * RS_SW];
   const int nb = gridDim.x;
   const int blk = xcd_blk(blockIdx.x, nb);
   const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
   double *sb = sb_all + wv * RS_SW;
   const int ng = (n + WAVE - 1) / WAVE;
   const int gpb = (ng + nb - 1) / nb;
   const int g0 = blk * gpb, g1 = min(ng, g0 + gpb);
   for (int g = g0 + wv; g < g1; g += RS_WPB) {
      const int r0 = g * WAVE;
      const int nr = min(WAVE, n - r0);
      int rs = 0, re = 0;
      if (lane < nr) {
         rs = ia[r0 + lane];
         re = ia[r0 + lane + 1];
      }
      const int base = __shfl(rs, 0, WAVE);
      const int end = __shfl(re, nr - 1, WAVE);
      double diag = 0.0, scale = 0.0, rsum = 0.0, l1s = 0.0;
      bool neg = false, allweak = false;
      double t = 0.0;
      int c = 0;
      double an[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
         const int j = base + u * WAVE + lane;
         an[u] = (j < end) ? __builtin_nontemporal_load(a + j) : 0.0;
      }
      for (int wb = base; wb < end; wb += RS_SW) {
         const int we = min(end, wb + RS_SW);
#pragma unroll
         for (int u = 0; u < U; u++) { sb[u * WAVE + lane] = an[u]; }
         const int nxt = wb + RS_SW;
#pragma unroll
         for (int u = 0; u < U; u++) {
            const int j = nxt + u * WAVE + lane;
            an[u] = (j < end) ? __builtin_nontemporal_load(a + j) : 0.0;
         }
         __builtin_amdgcn_wave_barrier();
         if (lane < nr) {
            const int s = max(rs, wb), e = min(re, we);
            for (int k = s; k < e; k++) {
               const double v = sb[k - wb];
               if (k == rs) {
                  diag = v;
                  neg = v < 0;
                  rsum = v;
                  l1s = 0.0 + fabs(v);
               } else {
                  scale = neg ? ((scale < v) ? v : scale) : ((scale < v) ? scale : v);
                  rsum += v;
                  l1s += fabs(v);
               }
            }
            if (re > wb && re <= we) {
               // the row ends in this window
               allweak = (fabs(rsum) > fabs(diag) * mrs) && (mrs < 1.0);
               t = theta * scale;
               if (rs >= wb) {
                  if (!allweak) {
                     for (int k = rs + 1; k < re; k++) {
                        const double v = sb[k - wb];
                        c += (neg ? !(v <= t) : !(v >= t)) ? 1 : 0;
                     }
                  }
               } else {
                  c = -1;
               }
            }
         }
         __builtin_amdgcn_wave_barrier();
      }
      if (lane < nr) {
         const int row = r0 + lane;
         thr[row] = t;
         flg[row] = (neg ? 1 : 0) | (allweak ? 2 : 0);
         cnt[row] = c;
         if (l1) {
            double s = l1s;
            if (re > rs && diag < 0.0) { s = -s; }
            if (fabs(s) == 0.0) { *l1_zero = 1; }
            l1[row] = s;
         }
      }
   }
}

// strong entries of the rows that k_row_stats left uncounted (cnt < 0)
template <int K>
__global__ void k_strength_recount(int n, const int *__restrict__ ia, const double *__restrict__ a,
                                   const double *__restrict__ thr, const int *__restrict__ flg,
                                   int *__restrict__ cnt)
{
   const int gl = threadIdx.x & (K - 1);
   const int rows_per_block = blockDim.x / K;
   for (long long rb = (long long) blockIdx.x * rows_per_block; rb < n;
        rb += (long long) gridDim.x * rows_per_block) {
      const long long ir = rb + threadIdx.x / K;
      const bool act = ir < n && cnt[ir] < 0;
      if (!__any(act)) { continue; }
      const int i = act ? (int) ir : 0;
      const int s = act ? ia[i] : 0, e = act ? ia[i + 1] : 0;
      const double t = act ? thr[i] : 0.0;
      const int fl = act ? flg[i] : 2;
      const bool neg = fl & 1;
      int c = 0;
      if (!(fl & 2)) {
         for (int k = s + 1 + gl; k < e; k += K) {
            const double v = a[k];
            c += (neg ? !(v <= t) : !(v >= t)) ? 1 : 0;
         }
      }
      for (int d = K / 2; d > 0; d >>= 1) { c += __shfl_xor(c, d, K); }
      if (act && gl == 0) { cnt[i] = c; }
   }
}

// S rows (strong columns in storage order) at the offsets of S_ia; with stc,
// also the number of rows in which each column is strong (the columns of
// S^T), aggregated per block in an LDS table over SF_ROWS consecutive rows.
constexpr int SF_ROWS = 256;
constexpr int SF_SLOTS = 4096;
constexpr int SF_U = 4;

template <int K, bool RUNS>
__global__ __launch_bounds__(256) void k_strength_fill(int n, const int *__restrict__ ia,
                                                       const int *__restrict__ ja,
                                                       const double *__restrict__ a,
                                                       const double *__restrict__ thr,
                                                       const int *__restrict__ flg,
                                                       const int *__restrict__ S_ia,
                                                       int *__restrict__ S_ja,
                                                       int *__restrict__ stc)
{
   __shared__ int hkey[SF_SLOTS];
   __shared__ int hcnt[SF_SLOTS];
   const int gl = threadIdx.x & (K - 1);
   const int rows_per_iter = 256 / K;
   const int gshift = (threadIdx.x & 63) & ~(K - 1);
   const unsigned long long gmask = (K == 64) ? ~0ULL : ((1ULL << K) - 1);
   if (stc) {
      for (int s = threadIdx.x; s < SF_SLOTS; s += 256) { hkey[s] = -1; hcnt[s] = 0; }
      __syncthreads();
   }
   const long long r_begin = (long long) blockIdx.x * SF_ROWS;
   const long long r_end = min((long long) n, r_begin + SF_ROWS);
   // a row's own operands are loaded one iteration ahead, while the rows
   // before it run
   int n_s = 0, n_e = 0, n_fl = 2, n_o = 0;
   double n_t = 0.0;
   auto load_row = [&](long long rb) {
      const long long ir = rb + threadIdx.x / K;
      n_s = n_e = n_o = 0;
      n_fl = 2;
      n_t = 0.0;
      if (ir < r_end) {
         const int i = (int) ir;
         n_s = ia[i];
         n_e = ia[i + 1];
         n_t = thr[i];
         n_fl = flg[i];
         n_o = S_ia[i];
      }
   };
   if (r_begin < r_end) { load_row(r_begin); }
   for (long long rb = r_begin; rb < r_end; rb += rows_per_iter) {
      const int s = n_s, e = n_e;
      const double t = n_t;
      const int fl = n_fl;
      const bool neg = fl & 1, allweak = fl & 2;
      const int o = n_o;
      if (rb + rows_per_iter < r_end) { load_row(rb + rows_per_iter); }
      int c = 0;
      // SF_U strided steps of the row per batch: every load of a batch is
      // issued before any is used (positions past the row end read its last
      // entry and are masked)
      const int klast = max(e - 1, 0);
      for (int kb = s + 1; __any(kb < e); kb += SF_U * K) {
         double vv[SF_U];
         int jj[SF_U];
#pragma unroll
         for (int u = 0; u < SF_U; u++) {
            const int kq = min(kb + gl + u * K, klast);
            vv[u] = a[kq];
            jj[u] = ja[kq];
         }
#pragma unroll
         for (int u = 0; u < SF_U; u++) {
            const int k = kb + gl + u * K;
            bool strong = false;
            const int j = jj[u];
            if (k < e && !allweak) {
               const double v = vv[u];
               strong = neg ? !(v <= t) : !(v >= t);
            }
            const unsigned long long m = (__ballot(strong) >> gshift) & gmask;
            if (strong) {
               S_ja[o + c + __popcll(m & ((1ULL << gl) - 1))] = j;
               if (stc) {
                  // RUNS (a fine-grid stencil operator in its natural order): a
                  // block's columns are a few runs of consecutive indices; by
                  // run_slot they fill consecutive slots, and the flush below
                  // adds to consecutive counters (one memory transaction for
                  // many lanes).  Otherwise (coarse operators,
                  // scattered columns) a multiplicative hash spreads them.
                  unsigned h = RUNS ? run_slot((unsigned) j, ilog2(SF_SLOTS))
                                    : (((unsigned) j * 2654435761u) >> (32 - ilog2(SF_SLOTS)));
                  bool done = false;
                  for (int probe = 0; probe < 32; probe++) {
                     const int old = atomicCAS(&hkey[h], -1, j);
                     if (old == -1 || old == j) {
                        atomicAdd(&hcnt[h], 1);
                        done = true;
                        break;
                     }
                     h = (h + 1) & (SF_SLOTS - 1);
                  }
                  if (!done) { atomicAdd(&stc[j], 1); }
               }
            }
            c += __popcll(m);
         }
      }
   }
   if (stc) {
      __syncthreads();
      for (int s = threadIdx.x; s < SF_SLOTS; s += 256) {
         if (hkey[s] >= 0) { atomicAdd(&stc[hkey[s]], hcnt[s]); }
      }
   }
}

void create_s(const Csr &A, double theta, double max_row_sum, Csr &S, int *st_count, double *l1,
              int *l1_zero, double *thr_out, int *flg_out, bool column_runs)
{
   int n = A.nrows;
   TimedLaunch tl("setup_strength", 0.0);
   S.nrows = n;
   S.ncols = A.ncols;
   S.ia = dalloc_n<int>((size_t) n + 1);
   S.a = nullptr;
   const int K = lanes_per_row(A.nnz, n);
   double *thr = thr_out ? thr_out : dalloc_n<double>((size_t) std::max(n, 1));
   int *flg = flg_out ? flg_out : dalloc_n<int>((size_t) std::max(n, 1));
   if (st_count) { dzero(st_count, sizeof(int) * (size_t) std::max(n, 1)); }
   if (n > 0) {
      const long long ng = ((long long) n + WAVE - 1) / WAVE;
      long long nbs = std::min<long long>((ng + RS_WPB - 1) / RS_WPB, 8192);
      if (nbs >= 64) { nbs = (nbs + 7) / 8 * 8; }
      k_row_stats<<<(int) std::max<long long>(nbs, 1), RS_WPB * WAVE, 0, stream()>>>(
          n, A.ia, A.a, theta, max_row_sum, thr, flg, S.ia, l1, l1_zero);
      HX_LAUNCH_CHECK();
      const int KR = lanes_per_row_light(A.nnz, n);
      const int nb = (int) std::max<long long>(1, std::min<long long>(((long long) n * KR + 255) / 256, 16384));
      dispatch_k(KR, [&](auto kc) {
         k_strength_recount<decltype(kc)::value><<<nb, 256, 0, stream()>>>(n, A.ia, A.a, thr, flg,
                                                                           S.ia);
      });
      HX_LAUNCH_CHECK();
   }
   S.nnz = exclusive_scan(S.ia, S.ia, n);
   S.ja = dalloc_n<int>((size_t) std::max<long long>(S.nnz, 1));
   if (n > 0) {
      const int nbf = (int) (((long long) n + SF_ROWS - 1) / SF_ROWS);
      dispatch_k(K, [&](auto kc) {
         if (column_runs) {
            k_strength_fill<decltype(kc)::value, true><<<nbf, 256, 0, stream()>>>(
               n, A.ia, A.ja, A.a, thr, flg, S.ia, S.ja, st_count);
         } else {
            k_strength_fill<decltype(kc)::value, false><<<nbf, 256, 0, stream()>>>(
               n, A.ia, A.ja, A.a, thr, flg, S.ia, S.ja, st_count);
         }
      });
      HX_LAUNCH_CHECK();
   }
   if (!thr_out) { dfree(thr); }
   if (!flg_out) { dfree(flg); }
}

// An off-diagonal entry of A whose sign is opposite to its row's diagonal and
// whose column is a C point: absolute entry index, column, value (16 bytes,
// one load).
struct __align__(16) OCEntry
{
   int    k;
   int    j;
   double a;
};

// One pass, K lanes per row: the strong C neighbours of the row in S order
// (SC, stored from the row's offset in S) and its OC entries in storage order
// (stored from the row's offset in A); list lengths to SC_cnt / OC_cnt.
template <int K>
__global__ void k_c_lists(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                          const double *__restrict__ a, const int *__restrict__ S_ia,
                          const int *__restrict__ S_ja, const int *__restrict__ CF,
                          int *__restrict__ SC_cnt, int *__restrict__ SC_ja,
                          int *__restrict__ OC_cnt, OCEntry *__restrict__ OC)
{
   const int gl = threadIdx.x & (K - 1);
   const int rows_per_block = blockDim.x / K;
   const int gshift = (threadIdx.x & 63) & ~(K - 1);
   const unsigned long long gmask = (K == 64) ? ~0ULL : ((1ULL << K) - 1);
   const unsigned long long below = (1ULL << gl) - 1;
   for (long long rb = (long long) blockIdx.x * rows_per_block; rb < n;
        rb += (long long) gridDim.x * rows_per_block) {
      const long long ir = rb + threadIdx.x / K;
      const bool act = ir < n;
      const int i = act ? (int) ir : 0;
      // strong C neighbours
      const int ss = act ? S_ia[i] : 0, se = act ? S_ia[i + 1] : 0;
      int csc = 0;
      for (int kb = ss; __any(kb < se); kb += K) {
         const int k = kb + gl;
         int j = -1;
         bool isc = false;
         if (k < se) {
            j = S_ja[k];
            isc = CF[j] >= 0;
         }
         const unsigned long long m = (__ballot(isc) >> gshift) & gmask;
         if (isc) { SC_ja[ss + csc + __popcll(m & below)] = j; }
         csc += __popcll(m);
      }
      // opposite-sign C entries
      const int s = act ? ia[i] : 0, e = act ? ia[i + 1] : 0;
      const bool neg = act && e > s && a[s] < 0;
      int coc = 0;
      for (int kb = s + 1; __any(kb < e); kb += K) {
         const int k = kb + gl;
         bool take = false;
         int j = 0;
         double v = 0.0;
         if (k < e) {
            v = a[k];
            if (neg ? (v > 0) : (v < 0)) {
               j = ja[k];
               take = CF[j] >= 0;
            }
         }
         const unsigned long long m = (__ballot(take) >> gshift) & gmask;
         if (take) { OC[s + coc + __popcll(m & below)] = OCEntry{k, j, v}; }
         coc += __popcll(m);
      }
      if (act && gl == 0) {
         SC_cnt[i] = csc;
         OC_cnt[i] = coc;
      }
   }
}

// The lists of k_c_lists from A and the strength threshold of each row (kept
// from k_row_stats), K lanes per row and SF_ROWS consecutive rows per block:
// SF_U strided steps of the row per batch, every load of a batch (values,
// columns, then the columns' C/F flags) issued before any is used; strong
// entries are those the threshold marks, in A order, i.e. S order, and both
// lists are compacted in storage order by ballots over the row's lanes.
template <int K>
__global__ __launch_bounds__(256) void k_c_lists_t(int n, const int *__restrict__ ia,
                                                   const int *__restrict__ ja,
                                                   const double *__restrict__ a,
                                                   const double *__restrict__ thr,
                                                   const int *__restrict__ flg,
                                                   const int *__restrict__ S_ia,
                                                   const int *__restrict__ CF,
                                                   int *__restrict__ SC_cnt,
                                                   int *__restrict__ SC_ja,
                                                   int *__restrict__ OC_cnt,
                                                   OCEntry *__restrict__ OC)
{
   const int gl = threadIdx.x & (K - 1);
   const int rows_per_iter = 256 / K;
   const int gshift = (threadIdx.x & 63) & ~(K - 1);
   const unsigned long long gmask = (K == 64) ? ~0ULL : ((1ULL << K) - 1);
   const unsigned long long below = (1ULL << gl) - 1;
   const long long r_begin = (long long) blockIdx.x * SF_ROWS;
   const long long r_end = min((long long) n, r_begin + SF_ROWS);
   // a row's own operands are loaded one iteration ahead, while the rows
   // before it run
   int n_s = 0, n_e = 0, n_fl = 2, n_so = 0;
   double n_t = 0.0;
   auto load_row = [&](long long rb) {
      const long long ir = rb + threadIdx.x / K;
      n_s = n_e = n_so = 0;
      n_fl = 2;
      n_t = 0.0;
      if (ir < r_end) {
         const int i = (int) ir;
         n_s = ia[i];
         n_e = ia[i + 1];
         n_t = thr[i];
         n_fl = flg[i];
         n_so = S_ia[i];
      }
   };
   if (r_begin < r_end) { load_row(r_begin); }
   for (long long rb = r_begin; rb < r_end; rb += rows_per_iter) {
      const long long ir = rb + threadIdx.x / K;
      const bool act = ir < r_end;
      const int i = act ? (int) ir : 0;
      const int s = n_s, e = n_e;
      const double t = n_t;
      const int fl = n_fl;
      const bool neg = fl & 1, allweak = fl & 2;
      const int so = n_so;
      if (rb + rows_per_iter < r_end) { load_row(rb + rows_per_iter); }
      // positions past the row end read its last entry and are masked (a
      // batch is only entered when some row of the wave has an entry left)
      const int klast = max(e - 1, 0);
      int coc = 0, csc = 0;
      for (int kb = s + 1; __any(kb < e); kb += SF_U * K) {
         double vv[SF_U];
         int jj[SF_U], cc[SF_U];
#pragma unroll
         for (int u = 0; u < SF_U; u++) {
            const int kq = min(kb + gl + u * K, klast);
            vv[u] = __builtin_nontemporal_load(a + kq);
            jj[u] = __builtin_nontemporal_load(ja + kq);
         }
#pragma unroll
         for (int u = 0; u < SF_U; u++) { cc[u] = CF[jj[u]]; }
#pragma unroll
         for (int u = 0; u < SF_U; u++) {
            const int k = kb + gl + u * K;
            bool toc = false, tsc = false;
            if (k < e && cc[u] >= 0) {
               const double v = vv[u];
               toc = neg ? (v > 0) : (v < 0);
               tsc = !allweak && (neg ? !(v <= t) : !(v >= t));
            }
            const unsigned long long mo = (__ballot(toc) >> gshift) & gmask;
            const unsigned long long ms = (__ballot(tsc) >> gshift) & gmask;
            if (toc) { OC[s + coc + __popcll(mo & below)] = OCEntry{k, jj[u], vv[u]}; }
            if (tsc) { SC_ja[so + csc + __popcll(ms & below)] = jj[u]; }
            coc += __popcll(mo);
            csc += __popcll(ms);
         }
      }
      if (act && gl == 0) {
         SC_cnt[i] = csc;
         OC_cnt[i] = coc;
      }
   }
}

// SC and OC lists of one level (storage as in k_c_lists)
struct CLists
{
   int     *sc_cnt = nullptr, *sc_ja = nullptr, *oc_cnt = nullptr;
   OCEntry *oc = nullptr;
   void     free_all()
   {
      dfree(sc_cnt); dfree(sc_ja); dfree(oc_cnt); dfree(oc);
      sc_cnt = sc_ja = oc_cnt = nullptr;
      oc = nullptr;
   }
};

static void c_lists(const Csr &A, const Csr &S, const int *CF, CLists &L, const double *thr,
                    const int *flg)
{
   const int n = A.nrows;
   L.sc_cnt = dalloc_n<int>((size_t) std::max(n, 1));
   L.oc_cnt = dalloc_n<int>((size_t) std::max(n, 1));
   L.sc_ja = dalloc_n<int>((size_t) std::max<long long>(S.nnz, 1));
   L.oc = dalloc_n<OCEntry>((size_t) std::max<long long>(A.nnz, 1));
   if (n == 0) { return; }
   const int K = lanes_per_row(A.nnz, n);
   if (thr && flg) {
      const int nbf = (int) (((long long) n + SF_ROWS - 1) / SF_ROWS);
      dispatch_k(K, [&](auto kc) {
         k_c_lists_t<decltype(kc)::value><<<nbf, 256, 0, stream()>>>(
             n, A.ia, A.ja, A.a, thr, flg, S.ia, CF, L.sc_cnt, L.sc_ja, L.oc_cnt, L.oc);
      });
      HX_LAUNCH_CHECK();
      return;
   }
   const int nb = (int) std::max<long long>(1, std::min<long long>(((long long) n * K + 255) / 256, 16384));
   dispatch_k(K, [&](auto kc) {
      k_c_lists<decltype(kc)::value><<<nb, 256, 0, stream()>>>(n, A.ia, A.ja, A.a, S.ia, S.ja, CF,
                                                               L.sc_cnt, L.sc_ja, L.oc_cnt, L.oc);
   });
   HX_LAUNCH_CHECK();
}

// ================================================================== l1 norms
// src/parcsr_ls/ams.c:4523-4600 option 1 + ams.c:770-797; K lanes per row,
// |a_ij| summed in column order.
template <int K>
__global__ void k_l1(int n, const int *__restrict__ ia, const double *__restrict__ a,
                     double *__restrict__ l1, int *__restrict__ zero_flag)
{
   const int gl = threadIdx.x & (K - 1);
   const int rows_per_block = blockDim.x / K;
   for (long long rb = (long long) blockIdx.x * rows_per_block; rb < n;
        rb += (long long) gridDim.x * rows_per_block) {
      const long long ir = rb + threadIdx.x / K;
      const bool act = ir < n;
      const int i = act ? (int) ir : 0;
      const int b = act ? ia[i] : 0, e = act ? ia[i + 1] : 0;
      double s = 0.0;
      for (int kb = b; __any(kb < e); kb += K) {
         const int k = kb + gl;
         const bool valid = k < e;
         s = group_fold<K>(s, valid ? fabs(a[k]) : 0.0, valid);
      }
      if (act && gl == 0) {
         if (a[b] < 0.0) { s = -s; }
         if (fabs(s) == 0.0) { *zero_flag = 1; }
         l1[i] = s;
      }
   }
}

// src/parcsr_ls/ams.c:649-676 option 4 on one rank (no offd block): |a_ii|
// of the last entry with j == i (csr_matop.c:2091-2117), sign of a_ii
// restored (ams.c:776-795); K lanes per row, the largest matching position
// found by a group max.
template <int K>
__global__ void k_l1_diag(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                          const double *__restrict__ a, double *__restrict__ l1,
                          int *__restrict__ zero_flag)
{
   const int gl = threadIdx.x & (K - 1);
   const int rows_per_block = blockDim.x / K;
   for (long long rb = (long long) blockIdx.x * rows_per_block; rb < n;
        rb += (long long) gridDim.x * rows_per_block) {
      const long long ir = rb + threadIdx.x / K;
      const bool act = ir < n;
      const int i = act ? (int) ir : 0;
      const int b = act ? ia[i] : 0, e = act ? ia[i + 1] : 0;
      int pos = -1;
      for (int k = b + gl; k < e; k += K) {
         if (ja[k] == i) { pos = k; }
      }
      for (int o = K / 2; o > 0; o >>= 1) { pos = max(pos, __shfl_xor(pos, o, K)); }
      if (act && gl == 0) {
         const double d = pos >= 0 ? a[pos] : 0.0;
         const double ad = fabs(d);
         const double s = d < 0.0 ? -ad : ad;
         if (fabs(s) == 0.0) { *zero_flag = 1; }
         l1[i] = s;
      }
   }
}

// row k of R is row n-1-k of A with its entries in storage order (the
// backward hybrid Gauss-Seidel sweep streams R forward); K lanes per row
template <int K>
__global__ void k_rev_rows(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                           const double *__restrict__ a, const float *__restrict__ a32,
                           int *__restrict__ ria, int *__restrict__ rja, double *__restrict__ ra,
                           float *__restrict__ ra32)
{
   const int gl = threadIdx.x & (K - 1);
   const int rows_per_block = blockDim.x / K;
   const int nnz = ia[n];
   for (long long rb = (long long) blockIdx.x * rows_per_block; rb < n;
        rb += (long long) gridDim.x * rows_per_block) {
      const long long ir = rb + threadIdx.x / K;
      if (ir >= n) { continue; }
      const int r = (int) ir;
      const int s = ia[r], e = ia[r + 1];
      const int d = nnz - e;
      if (gl == 0) {
         ria[n - 1 - r] = d;
         if (r == 0) { ria[n] = nnz; }
      }
      for (int q = s + gl; q < e; q += K) {
         rja[d + (q - s)] = ja[q];
         if (ra) { ra[d + (q - s)] = a[q]; }
         if (ra32) { ra32[d + (q - s)] = a32[q]; }
      }
   }
}

void csr_reverse_rows(const Csr &A, Csr &R)
{
   R.free_all();
   R.alloc(A.nrows, A.ncols, A.nnz, A.a != nullptr);
   if (A.a32) { R.a32 = dalloc_n<float>((size_t) std::max<long long>(A.nnz, 1)); }
   const int n = A.nrows;
   if (n <= 0) {
      dzero(R.ia, sizeof(int));
      return;
   }
   const int K = lanes_per_row(A.nnz, n);
   const int nb = (int) std::max<long long>(1, std::min<long long>(((long long) n * K + 255) / 256, 16384));
   dispatch_k(K, [&](auto kc) {
      k_rev_rows<decltype(kc)::value><<<nb, 256, 0, stream()>>>(n, A.ia, A.ja, A.a, A.a32, R.ia,
                                                                R.ja, R.a, R.a32);
   });
   HX_LAUNCH_CHECK();
}

void l1_norms_diag(const Csr &A, double *l1)
{
   int *flag = device_flags() + 40;
   dzero(flag, sizeof(int));
   const int n = A.nrows;
   const int K = lanes_per_row(A.nnz, n);
   const int nb = (int) std::max<long long>(1, std::min<long long>(((long long) n * K + 255) / 256, 16384));
   dispatch_k(K, [&](auto kc) {
      k_l1_diag<decltype(kc)::value><<<nb, 256, 0, stream()>>>(n, A.ia, A.ja, A.a, l1, flag);
   });
   HX_LAUNCH_CHECK();
   int h = 0;
   d2h(&h, flag, sizeof(int));
   if (h) { g_error_arg = 1; set_error(HYPRE_ERROR_ARG, "zero l1 norm"); }
}

void l1_norms(const Csr &A, double *l1)
{
   int *flag = device_flags() + 40;
   dzero(flag, sizeof(int));
   const int n = A.nrows;
   const int K = lanes_per_row(A.nnz, n);
   const int nb = (int) std::max<long long>(1, std::min<long long>(((long long) n * K + 255) / 256, 16384));
   dispatch_k(K, [&](auto kc) {
      k_l1<decltype(kc)::value><<<nb, 256, 0, stream()>>>(n, A.ia, A.a, l1, flag);
   });
   HX_LAUNCH_CHECK();
   int h = 0;
   d2h(&h, flag, sizeof(int));
   if (h) { g_error_arg = 1; set_error(HYPRE_ERROR_ARG, "zero l1 norm"); }
}

// ================================================================== PMIS
// Park-Miller stream (src/utilities/random.c): draw k (0-based) after
// hypre_SeedRand(seed) is 16807^(k+1) * seed mod (2^31-1); computed by
// jump-ahead, bit-identical to the serial Schrage recurrence.
__device__ __forceinline__ unsigned long long pm_mulmod(unsigned long long x, unsigned long long y)
{
   const unsigned long long M = 2147483647ULL;
   unsigned long long p = x * y;
   p = (p & M) + (p >> 31);
   p = (p & M) + (p >> 31);
   if (p >= M) { p -= M; }
   return p;
}

__device__ __forceinline__ unsigned long long pm_pow(unsigned long long b, unsigned long long e)
{
   unsigned long long r = 1;
   while (e) {
      if (e & 1) { r = pm_mulmod(r, b); }
      b = pm_mulmod(b, b);
      e >>= 1;
   }
   return r;
}

__device__ __forceinline__ double pm_draw(int seed, long long k)
{
   unsigned long long x = pm_mulmod(pm_pow(16807ULL, (unsigned long long) k + 1), (unsigned long long) seed);
   return (double) (int) x / (double) 2147483647;
}

__global__ void k_pm_uniform(int n, int seed, double *__restrict__ out)
{
   for (long long i = (long long) blockIdx.x * blockDim.x + threadIdx.x; i < n;
        i += (long long) gridDim.x * blockDim.x) {
      out[i] = pm_draw(seed, i);
   }
}

void pm_uniform(int n, int seed, double *out_dev)
{
   if (seed < 1) { seed = 1; }
   if (seed >= 2147483647) { seed = 2147483646; }
   int nb = std::max(1, std::min((n + 255) / 256, 8192));
   k_pm_uniform<<<nb, 256, 0, stream()>>>(n, seed, out_dev);
   HX_LAUNCH_CHECK();
}

__global__ void k_col_count(long long nnz, const int *__restrict__ ja, int *__restrict__ cnt)
{
   for (long long k = (long long) blockIdx.x * blockDim.x + threadIdx.x; k < nnz;
        k += (long long) gridDim.x * blockDim.x) {
      atomicAdd(&cnt[ja[k]], 1);
   }
}

// par_coarsen.c:2248-2407 + par_indepset.c:50-75, first-round candidate flags
// CF_init 3 (par_coarsen.c:2396-2399): rows without strong connections
// become C points, and no candidate is flagged (no independent set is picked
// in the first round, par_coarsen.c:2487)
__global__ void k_pmis_init(int n, const int *__restrict__ S_ia, const int *__restrict__ cnt,
                            const double *__restrict__ given, int seed, double *__restrict__ m,
                            int *__restrict__ CF, int cf_init)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      double mi = given ? given[i] : (double) cnt[i] + pm_draw(seed, i);
      if (S_ia[i + 1] == S_ia[i]) {
         CF[i] = (cf_init == 3) ? 1 : -3;   // C_PT : SF_PT
         m[i] = 0.0;
      } else {
         m[i] = mi;
         CF[i] = (cf_init == 0 && mi > 1) ? 1 : 0;
      }
   }
}

// CF_init 1 (HMIS, par_coarsen.c:2346-2374): CF holds the Ruge-Stueben first
// pass's split (1 C, -1 F, -2 Z, -3 SF); F points become undecided, Z points
// stay in the graph only when their measure is >= 1 or they depend on
// someone, and no candidate flags are set (the first pass's C points are the
// first independent set).  Rows outside the graph get measure 0.
__global__ void k_pmis_init_cf1(int n, const int *__restrict__ S_ia, const int *__restrict__ cnt,
                                const double *__restrict__ given, int seed,
                                double *__restrict__ m, int *__restrict__ CF)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      const double mi = given ? given[i] : (double) cnt[i] + pm_draw(seed, i);
      int c = CF[i];
      double mo = mi;
      if (c == -3) {
         mo = 0.0;
      } else {
         if (c == -1) { c = 0; }
         if (c == -2) {
            if (mi >= 1.0 || S_ia[i + 1] > S_ia[i]) {
               c = 0;
            } else {
               c = -1;
               mo = 0.0;
            }
         }
      }
      m[i] = mo;
      CF[i] = c;
   }
}

constexpr int PM_U = 4;   // strided steps of a row whose loads are issued together

// par_coarsen.c:2506-2550: drop candidates with a larger-measure S-neighbour
// (K lanes per row; every write is a 0, so their order does not matter)
template <int K>
__global__ void k_pmis_remove(int nw, const int *__restrict__ list, const int *__restrict__ S_ia,
                              const int *__restrict__ S_ja, const double *__restrict__ m, int *CF)
{
   const int gl = threadIdx.x & (K - 1);
   const int rows_per_block = blockDim.x / K;
   const int gshift = (threadIdx.x & 63) & ~(K - 1);
   const unsigned long long gmask = (K == 64) ? ~0ULL : ((1ULL << K) - 1);
   for (long long rb = (long long) blockIdx.x * rows_per_block; rb < nw;
        rb += (long long) gridDim.x * rows_per_block) {
      const long long ir = rb + threadIdx.x / K;
      const bool act = ir < nw;
      const int i = act ? (list ? list[ir] : (int) ir) : 0;
      const double mi = act ? m[i] : 0.0;
      const bool cand = mi > 1;
      const int s = (act && cand) ? S_ia[i] : 0, e = (act && cand) ? S_ia[i + 1] : 0;
      bool lose = false;
      // PM_U strided steps per batch: their columns, then their measures, are
      // loaded before any is used (positions past the row end read its last
      // entry and are masked)
      const int klast = max(e - 1, 0);
      for (int kb = s; __any(kb < e); kb += PM_U * K) {
         int jv[PM_U];
         double mv[PM_U];
#pragma unroll
         for (int u = 0; u < PM_U; u++) { jv[u] = S_ja[min(kb + gl + u * K, klast)]; }
#pragma unroll
         for (int u = 0; u < PM_U; u++) { mv[u] = m[jv[u]]; }
#pragma unroll
         for (int u = 0; u < PM_U; u++) {
            if (kb + gl + u * K < e && mv[u] > 1) {
               if (mi > mv[u]) { CF[jv[u]] = 0; }
               else if (mv[u] > mi) { lose = true; }
            }
         }
      }
      const unsigned long long lm = (__ballot(lose) >> gshift) & gmask;
      if (lm && gl == 0) { CF[i] = 0; }
   }
}

// par_coarsen.c:2603-2660: C points, and F points depending strongly on a C
// point.  A neighbour's "CF > 0" does not change within this pass (survivors
// stay 1, everything else is 0 or -1), so the concurrent updates are safe.
template <int K>
__global__ void k_pmis_cf(int nw, const int *__restrict__ list, const int *__restrict__ S_ia,
                          const int *__restrict__ S_ja, const double *__restrict__ m, int *CF)
{
   const int gl = threadIdx.x & (K - 1);
   const int rows_per_block = blockDim.x / K;
   const int gshift = (threadIdx.x & 63) & ~(K - 1);
   const unsigned long long gmask = (K == 64) ? ~0ULL : ((1ULL << K) - 1);
   for (long long rb = (long long) blockIdx.x * rows_per_block; rb < nw;
        rb += (long long) gridDim.x * rows_per_block) {
      const long long ir = rb + threadIdx.x / K;
      const bool act = ir < nw;
      const int i = act ? (list ? list[ir] : (int) ir) : 0;
      const double mi = act ? m[i] : 0.0;
      const bool in_graph = act && mi > 0;
      int c = in_graph ? CF[i] : 0;
      if (in_graph && mi < 1) { c = -1; }
      const bool scan = in_graph && !(c > 0);
      const int s = scan ? S_ia[i] : 0, e = scan ? S_ia[i + 1] : 0;
      bool hitc = false;
      const int klast = max(e - 1, 0);
      for (int kb = s; __any(kb < e); kb += PM_U * K) {
         int jv[PM_U], cv[PM_U];
#pragma unroll
         for (int u = 0; u < PM_U; u++) { jv[u] = S_ja[min(kb + gl + u * K, klast)]; }
#pragma unroll
         for (int u = 0; u < PM_U; u++) { cv[u] = CF[jv[u]]; }
#pragma unroll
         for (int u = 0; u < PM_U; u++) {
            if (kb + gl + u * K < e && cv[u] > 0) { hitc = true; }
         }
      }
      const unsigned long long hm = (__ballot(hitc) >> gshift) & gmask;
      if (in_graph && gl == 0) {
         if (c > 0) { c = 1; } else if (hm) { c = -1; }
         CF[i] = c;
      }
   }
}

// par_coarsen.c:2690-2750 graph update, fused with the next round's
// candidate flags (par_coarsen.c:2485-2495).  The rows still in the graph
// (m > 0) are appended to `next` (in no particular order: a round's outcome
// does not depend on the order of its rows); *remaining counts them.
__global__ __launch_bounds__(256) void k_pmis_update(int nw, const int *__restrict__ list,
                                                     double *__restrict__ m, int *__restrict__ CF,
                                                     int *__restrict__ next,
                                                     int *__restrict__ remaining)
{
   __shared__ int wsum[4];
   __shared__ int bbase;
   const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
   // each block owns a contiguous chunk of the work list: update and count,
   // one atomic for the chunk's place in `next`, then append in chunk order
   const int chunk = (nw + gridDim.x - 1) / gridDim.x;
   const int w0 = blockIdx.x * chunk, w1 = min(nw, w0 + chunk);
   int kept = 0;
   for (int w = w0 + threadIdx.x; w < w1; w += blockDim.x) {
      const int i = list ? list[w] : w;
      const double mi = m[i];
      if (mi > 0) {
         if (CF[i] != 0) {
            m[i] = 0.0;
         } else {
            kept++;
            if (mi > 1) { CF[i] = 1; }
         }
      }
   }
   for (int d = 32; d > 0; d >>= 1) { kept += __shfl_xor(kept, d, 64); }
   if (lane == 0) { wsum[wv] = kept; }
   __syncthreads();
   if (threadIdx.x == 0) {
      const int tot = wsum[0] + wsum[1] + wsum[2] + wsum[3];
      bbase = tot ? atomicAdd(remaining, tot) : 0;
   }
   __syncthreads();
   int run = bbase;
   for (int wb = w0; wb < w1; wb += blockDim.x) {
      const int w = wb + threadIdx.x;
      bool keep = false;
      int i = -1;
      if (w < w1) {
         i = list ? list[w] : w;
         keep = m[i] > 0;
      }
      const unsigned long long km = __ballot(keep);
      __syncthreads();
      if (lane == 0) { wsum[wv] = __popcll(km); }
      __syncthreads();
      int off = run;
      for (int v = 0; v < wv; v++) { off += wsum[v]; }
      if (keep) { next[off + __popcll(km & ((1ULL << lane) - 1))] = i; }
      run += wsum[0] + wsum[1] + wsum[2] + wsum[3];
   }
}

int pmis(const Csr &S, const double *measure_in_dev, int *CF, double *measure_keep_dev,
         const int *st_count, int cf_init)
{
   int n = S.nrows;
   if (n == 0) { return 0; }
   TimedLaunch tl("setup_pmis", 0.0);
   hipStream_t st = stream();
   int nb = std::max(1, std::min((n + 255) / 256, 8192));
   double *m = dalloc_n<double>(n);
   int *cnt = nullptr;
   const int *stc = st_count;
   if (!measure_in_dev && !stc) {
      cnt = dalloc_n<int>(n);
      dzero(cnt, sizeof(int) * n);
      int nbe = std::max<long long>(1, std::min<long long>((S.nnz + 255) / 256, 8192));
      if (S.nnz > 0) { k_col_count<<<nbe, 256, 0, st>>>(S.nnz, S.ja, cnt); }
      HX_LAUNCH_CHECK();
      stc = cnt;
   }
   if (cf_init == 1) {
      k_pmis_init_cf1<<<nb, 256, 0, st>>>(n, S.ia, stc, measure_in_dev, 2747, m, CF);
   } else {
      k_pmis_init<<<nb, 256, 0, st>>>(n, S.ia, stc, measure_in_dev, 2747, m, CF, cf_init);
   }
   HX_LAUNCH_CHECK();
   if (measure_keep_dev) { d2d(measure_keep_dev, m, sizeof(double) * n); }
   int *remaining = device_flags() + 44;
   int *h = static_cast<int *>(pinned_scratch(64));
   int rounds = 0;
   const int K = lanes_per_row_light(S.nnz, n);
   // rows in the graph: all rows in the first round, then those left by the update
   int *lists[2] = {dalloc_n<int>((size_t) n), dalloc_n<int>((size_t) n)};
   const int *list = nullptr;
   int nw = n;
   while (true) {
      dzero(remaining, sizeof(int));
      const int nbk = (int) std::max<long long>(1, std::min<long long>(((long long) nw * K + 255) / 256, 16384));
      const int nbu = std::max(1, std::min((nw + 255) / 256, 8192));
      // par_coarsen.c:2487: with CF_init 1 the first round keeps the given
      // C points as its independent set; with CF_init 3 it picks none
      const bool pick = !(cf_init != 0 && rounds == 0);
      dispatch_k(K, [&](auto kc) {
         if (pick) { k_pmis_remove<decltype(kc)::value><<<nbk, 256, 0, st>>>(nw, list, S.ia, S.ja, m, CF); }
         k_pmis_cf<decltype(kc)::value><<<nbk, 256, 0, st>>>(nw, list, S.ia, S.ja, m, CF);
      });
      int *next = lists[rounds & 1];
      k_pmis_update<<<nbu, 256, 0, st>>>(nw, list, m, CF, next, remaining);
      HX_LAUNCH_CHECK();
      rounds++;
      HX_HIP(hipMemcpyAsync(h, remaining, sizeof(int), hipMemcpyDeviceToHost, st));
      sync();
      if (h[0] == 0) { break; }
      if (rounds > 100000) { fail(HYPRE_ERROR_GENERIC, "PMIS did not terminate"); }
      list = next;
      nw = h[0];
   }
   dfree(lists[0]);
   dfree(lists[1]);
   dfree(m);
   dfree(cnt);
   return rounds;
}

// ================================================================== coarse map
__global__ void k_cflag(int n, const int *__restrict__ CF, int *__restrict__ flag)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      flag[i] = (CF[i] == 1) ? 1 : 0;
   }
}

int coarse_map(int n, const int *CF, int *f2c)
{
   int nb = std::max(1, std::min((n + 255) / 256, 8192));
   k_cflag<<<nb, 256, 0, stream()>>>(n, CF, f2c);
   HX_LAUNCH_CHECK();
   return (int) exclusive_scan(f2c, f2c, n);
}

// ================================================================== row tables
// Per-wave open-addressing hash (int keys -> int values) with a runtime
// power-of-two capacity; callers keep the load <= 1/2 so probing terminates.
// Storage is dynamic LDS for the common launches and a per-wave slab of global
// memory for the few rows that outgrow LDS (coarse levels).
struct RowHash
{
   int *keys;
   int *vals;
   int  cap;
   int  bits;
   __device__ __forceinline__ void clear(int lane)
   {
      for (int s = lane; s < cap; s += WAVE) { keys[s] = -1; }
   }
   __device__ __forceinline__ unsigned h0(int key) const
   {
      return ((unsigned) key * 2654435761u) >> (32 - bits);
   }
   // the probe loops have one exit (a lane leaves when its slot is settled),
   // which compiles to far fewer scalar mask operations per probe than
   // returning from inside the loop
   __device__ __forceinline__ int insert(int key, bool &isnew)
   {
      unsigned h = h0(key);
      int slot = -1, old = 0;
      for (int probe = 0; probe < cap; probe++) {
         old = atomicCAS(&keys[h], -1, key);
         if (old == -1 || old == key) {
            slot = (int) h;
            break;
         }
         h = (h + 1) & (cap - 1);
      }
      isnew = slot >= 0 && old == -1;
      return slot;
   }
   __device__ __forceinline__ int find(int key) const
   {
      unsigned h = h0(key);
      int slot = -1;
      for (int probe = 0; probe < cap; probe++) {
         const int k = keys[h];
         if (k == key || k == -1) {
            slot = (k == key) ? (int) h : -1;
            break;
         }
         h = (h + 1) & (cap - 1);
      }
      return slot;
   }
};

// bytes of one table set: cap keys + cap vals + cap/2 ints + cap/2 doubles
__host__ __device__ inline size_t table_bytes(int cap) { return (size_t) cap * 8 + (size_t) (cap / 2) * 12; }

__device__ __forceinline__ void carve(char *base, int cap, int *&keys, int *&vals, int *&col,
                                      double *&val)
{
   // doubles first for alignment
   val = reinterpret_cast<double *>(base);
   keys = reinterpret_cast<int *>(base + (size_t) (cap / 2) * 8);
   vals = keys + cap;
   col = vals + cap;
}

// ================================================================== ext+i
// src/parcsr_ls/par_lr_interp.c:1024-1900 (BuildExtPIInterpHost) followed by
// src/parcsr_mv/par_csr_matrix.c:2943-3340 (Truncate, rescale 1, inf-norm) and
// src/utilities/qsort.c:452 (hypre_qsort2_abs).
enum { MODE_SYM = 0, MODE_NUM = 1 };

struct InterpArgs
{
   int           n;
   const int    *A_ia, *A_ja;
   const double *A_a;
   const int    *S_ia, *S_ja;
   const int    *CF, *f2c;
   double        trunc_factor;
   int           max_elmts;
   const int    *rowlist;        // rows to process (nullptr: 0..n-1)
   int           nwork;
   const int    *base;           // MODE_NUM: row slot base (nullptr: i*max_elmts)
   int          *out_j;
   double       *out_a;
   int          *cnt;
   int          *defer_list;
   int          *ndefer;
   int           cap;            // hash capacity (power of two)
   char         *gslab;          // nullptr: tables in dynamic LDS
   size_t        slab_stride;
};

// serial partial quicksort in the order of qsort.c:452 (hypre_qsort2_abs):
// stops refining once the first `keep` positions are final; explicit stack,
// larger side pushed first.  Disjoint sub-ranges are independent, so the
// processing order does not change the result.
__device__ void qsort2_abs_partial(int *v, double *w, int cnt, int keep)
{
   int stl[48], str[48];
   int sp = 0;
   stl[sp] = 0; str[sp] = cnt - 1; sp++;
   while (sp > 0) {
      sp--;
      int left = stl[sp], right = str[sp];
      if (left >= right || left >= keep) { continue; }
      int mid = (left + right) / 2;
      { int t = v[left]; v[left] = v[mid]; v[mid] = t; double tw = w[left]; w[left] = w[mid]; w[mid] = tw; }
      int last = left;
      double piv = fabs(w[left]);
      for (int i = left + 1; i <= right; i++) {
         if (fabs(w[i]) > piv) {
            ++last;
            int t = v[last]; v[last] = v[i]; v[i] = t;
            double tw = w[last]; w[last] = w[i]; w[i] = tw;
         }
      }
      { int t = v[left]; v[left] = v[last]; v[last] = t; double tw = w[left]; w[left] = w[last]; w[last] = tw; }
      int l1 = left, r1 = last - 1, l2 = last + 1, r2 = right;
      if ((r1 - l1) > (r2 - l2)) {
         stl[sp] = l1; str[sp] = r1; sp++;
         stl[sp] = l2; str[sp] = r2; sp++;
      } else {
         stl[sp] = l2; str[sp] = r2; sp++;
         stl[sp] = l1; str[sp] = r1; sp++;
      }
   }
}

template <int MODE>
__global__ __launch_bounds__(WAVE) void k_extpi(InterpArgs p)
{
   extern __shared__ double dyn_lds[];
   const int lane = threadIdx.x;
   const int nb = gridDim.x;
   const int blk = xcd_blk(blockIdx.x, nb);
   char *base = p.gslab ? p.gslab + (size_t) blk * p.slab_stride : reinterpret_cast<char *>(dyn_lds);
   int *keys, *vals, *s_col;
   double *s_w;
   carve(base, p.cap, keys, vals, s_col, s_w);
   RowHash H{keys, vals, p.cap, ilog2(p.cap)};
   const int capc = p.cap / 2;

   // rows strided over the grid: co-resident waves work on one window of
   // neighbouring rows, whose neighbour rows stay in the XCD's L2
   for (int wi = blk; wi < p.nwork; wi += nb) {
      const int i = p.rowlist ? p.rowlist[wi] : wi;
      const int cfi = p.CF[i];
      if (cfi >= 0 || cfi == -3) {
         if (lane == 0) {
            if (MODE == MODE_SYM) {
               p.cnt[i] = (cfi >= 0) ? 1 : 0;
            } else {
               int b = p.base ? p.base[i] : i * p.max_elmts;
               if (cfi >= 0) {
                  p.out_j[b] = p.f2c[i];
                  p.out_a[b] = 1.0;
                  p.cnt[i] = 1;
               } else {
                  p.cnt[i] = 0;
               }
            }
         }
         continue;
      }
      // ---- F row: C-hat in first-seen order (par_lr_interp.c:1505-1560)
      H.clear(lane);
      wave_sync();
      int nkeys = 0, nch = 0;
      bool overflow = false;
      const int s0 = p.S_ia[i], s1 = p.S_ia[i + 1];
      for (int jj = s0; jj < s1 && !overflow; jj++) {
         const int i1 = p.S_ja[jj];
         const int c1 = p.CF[i1];
         if (c1 >= 0) {
            if (nkeys + 1 > capc) { overflow = true; break; }
            int isn = 0;
            if (lane == 0) {
               bool isnew;
               int slot = H.insert(i1, isnew);
               if (isnew && slot >= 0) {
                  vals[slot] = nch;
                  s_col[nch] = p.f2c[i1];
                  s_w[nch] = 0.0;
               }
               isn = isnew ? 1 : 0;
            }
            isn = readlane_i(isn, 0);
            nch += isn;
            nkeys += isn;
            wave_sync();
         } else if (c1 != -3) {
            // strong F neighbour i1: mark it, then add its strong C neighbours
            if (nkeys + 1 > capc) { overflow = true; break; }
            int isn = 0;
            if (lane == 0) {
               bool isnew;
               int slot = H.insert(i1, isnew);
               if (slot >= 0) { vals[slot] = -2; }
               isn = isnew ? 1 : 0;
            }
            nkeys += readlane_i(isn, 0);
            wave_sync();
            const int k0 = p.S_ia[i1], k1e = p.S_ia[i1 + 1];
            for (int kb = k0; kb < k1e; kb += WAVE) {
               const int kk = kb + lane;
               int k1 = -1;
               bool want = false;
               if (kk < k1e) {
                  k1 = p.S_ja[kk];
                  want = p.CF[k1] >= 0;
               }
               const int nw = __popcll(__ballot(want));
               if (nkeys + nw > capc) { overflow = true; break; }
               bool isnew = false;
               int slot = -1;
               if (want) { slot = H.insert(k1, isnew); }
               const unsigned long long nm = __ballot(isnew);
               if (isnew && slot >= 0) {
                  const int pos = nch + __popcll(nm & lanemask_lt(lane));
                  vals[slot] = pos;
                  s_col[pos] = p.f2c[k1];
                  s_w[pos] = 0.0;
               }
               nch += __popcll(nm);
               nkeys += __popcll(nm);
               wave_sync();
            }
         }
      }
      if (overflow) {
         if (lane == 0) {
            int d = atomicAdd(p.ndefer, 1);
            p.defer_list[d] = i;
         }
         wave_sync();
         continue;
      }
      if (MODE == MODE_SYM) {
         if (lane == 0) { p.cnt[i] = nch; }
         wave_sync();
         continue;
      }
      // ---- weights (par_lr_interp.c:1640-1805)
      const int a0 = p.A_ia[i], a1 = p.A_ia[i + 1];
      double diagonal = p.A_a[a0];
      for (int jj = a0 + 1; jj < a1; jj++) {
         const int i1 = p.A_ja[jj];
         const double aij = p.A_a[jj];
         const int slot = H.find(i1);
         const int v = (slot >= 0) ? vals[slot] : -1;
         if (v >= 0) {
            // C-hat member: a_ij goes into its weight
            if (lane == 0) { s_w[v] += aij; }
            wave_sync();
         } else if (v == -2) {
            // strong F neighbour: distribute a_ij over C-hat (sign-opposite
            // entries of row i1), the part landing on i itself to the diagonal
            const int b1 = p.A_ia[i1], e1 = p.A_ia[i1 + 1];
            const int sgn = (p.A_a[b1] < 0) ? -1 : 1;
            double sum = 0.0;
            for (int kb = b1 + 1; kb < e1; kb += WAVE) {
               const int k = kb + lane;
               bool qual = false;
               double a2 = 0.0;
               if (k < e1) {
                  const int i2 = p.A_ja[k];
                  a2 = p.A_a[k];
                  const int sl = H.find(i2);
                  const int vi2 = (sl >= 0) ? vals[sl] : -1;
                  qual = ((vi2 >= 0) || i2 == i) && ((double) sgn * a2) < 0;
               }
               unsigned long long qm = __ballot(qual);
               while (qm) {                       // left-to-right fold
                  const int t = __ffsll((long long) qm) - 1;
                  sum += readlane_d(a2, t);
                  qm &= qm - 1;
               }
            }
            if (sum != 0) {
               const double distribute = aij / sum;
               for (int kb = b1 + 1; kb < e1; kb += WAVE) {
                  const int k = kb + lane;
                  bool dq = false;
                  double dv = 0.0;
                  if (k < e1) {
                     const int i2 = p.A_ja[k];
                     const double a2 = p.A_a[k];
                     if (((double) sgn * a2) < 0) {
                        const int sl = H.find(i2);
                        const int vi2 = (sl >= 0) ? vals[sl] : -1;
                        if (vi2 >= 0) { s_w[vi2] += distribute * a2; }
                        if (i2 == i) { dq = true; dv = distribute * a2; }
                     }
                  }
                  const unsigned long long dm = __ballot(dq);
                  if (dm) { diagonal += readlane_d(dv, __ffsll((long long) dm) - 1); }
               }
               wave_sync();
            } else {
               diagonal += aij;
            }
         } else if (p.CF[i1] != -3) {
            diagonal += aij;                       // weak connection
         }
      }
      wave_sync();
      if (diagonal != 0.0) {
         for (int k = lane; k < nch; k += WAVE) { s_w[k] /= -diagonal; }
      }
      wave_sync();
      // ---- truncation (par_csr_matrix.c:3040-3320), serial in lane 0
      const int b = p.base ? p.base[i] : i * p.max_elmts;
      int kept = nch;
      if (lane == 0) {
         if (p.trunc_factor > 0) {
            double row_nrm = 0.0;
            for (int k = 0; k < nch; k++) {
               row_nrm = (row_nrm < fabs(s_w[k])) ? fabs(s_w[k]) : row_nrm;
            }
            const double drop_coeff = p.trunc_factor * row_nrm;
            double row_sum = 0.0, scale = 0.0;
            int o = 0;
            for (int k = 0; k < nch; k++) {
               const double w = s_w[k];
               row_sum += w;
               if (fabs(w) < drop_coeff) { continue; }
               scale += w;
               s_w[o] = w;
               s_col[o] = s_col[k];
               o++;
            }
            if (scale != 0.) {
               if (scale != row_sum) {
                  scale = row_sum / scale;
                  for (int k = 0; k < o; k++) { s_w[k] *= scale; }
               }
            }
            kept = o;
         }
         if (p.max_elmts > 0 && kept > p.max_elmts) {
            double row_sum = 0.0;
            for (int k = 0; k < kept; k++) { row_sum += s_w[k]; }
            qsort2_abs_partial(s_col, s_w, kept, p.max_elmts);
            double scale = 0.0;
            for (int k = 0; k < p.max_elmts; k++) { scale += s_w[k]; }
            if (scale != 0.) {
               if (scale != row_sum) {
                  scale = row_sum / scale;
                  for (int k = 0; k < p.max_elmts; k++) { s_w[k] *= scale; }
               }
            }
            kept = p.max_elmts;
         }
         p.cnt[i] = kept;
      }
      kept = readlane_i(kept, 0);
      wave_sync();
      for (int k = lane; k < kept; k += WAVE) {
         p.out_j[b + k] = s_col[k];
         p.out_a[b + k] = s_w[k];
      }
      wave_sync();
   }
}

// Shared driver of the interpolation and RAP passes: run over all rows with
// small LDS tables, then re-run deferred rows with larger LDS tables, then with
// global-memory slabs.
template <class Args, class Launch>
static void deferred_passes(Args a, int nwork_all, int *ndefer, const int *lds_caps, int nlds,
                            int max_global_cap, int ntables, Launch launch,
                            const int *init_list = nullptr, int init_n = 0)
{
   int *h = static_cast<int *>(pinned_scratch(64));
   const int nl = init_list ? init_n : nwork_all;
   int *lists[2] = {dalloc_n<int>((size_t) nl + 1), dalloc_n<int>((size_t) nl + 1)};
   a.rowlist = init_list;
   a.nwork = nl;
   int cur = 0;
   int cap = lds_caps[0];
   int stage = 0;
   char *slab = nullptr;
   while (true) {
      dzero(ndefer, sizeof(int));
      a.defer_list = lists[cur];
      a.ndefer = ndefer;
      a.cap = cap;
      a.gslab = nullptr;
      a.slab_stride = 0;
      int nb = launch_blocks(a.nwork);
      if (stage >= nlds) {
         // global slabs: bound the number of concurrent waves by memory
         size_t per = (size_t) ntables * table_bytes(cap) + 256;
         long long waves = std::min<long long>(a.nwork, std::max<long long>(8, (1LL << 31) / (long long) per));
         nb = (int) std::max<long long>(1, std::min<long long>(waves, 2048));
         if (nb >= 64) { nb = nb / 8 * 8; }
         slab = static_cast<char *>(dalloc(per * (size_t) nb));
         a.gslab = slab;
         a.slab_stride = per;
      }
      launch(a, nb);
      HX_HIP(hipMemcpyAsync(h, ndefer, sizeof(int), hipMemcpyDeviceToHost, stream()));
      sync();
      if (slab) { dfree(slab); slab = nullptr; }
      if (h[0] == 0) { break; }
      stage++;
      if (stage < nlds) {
         cap = lds_caps[stage];
      } else {
         cap = (stage == nlds) ? std::max(cap * 2, 8192) : cap * 4;
         if (cap > max_global_cap) {
            dfree(lists[0]); dfree(lists[1]);
            fail(HYPRE_ERROR_GENERIC, "setup row exceeds the table capacity");
         }
      }
      a.rowlist = lists[cur];
      a.nwork = h[0];
      cur ^= 1;
   }
   dfree(lists[0]);
   dfree(lists[1]);
}

template <int MODE>
static void extpi_pass(InterpArgs a, const int *init_list = nullptr, int init_n = 0)
{
   static const int caps[2] = {512, 4096};
   deferred_passes(a, a.n, device_flags() + 48, caps, 2, 1 << 22, 1,
                   [](InterpArgs &x, int nb) {
                      size_t lds = x.gslab ? 0 : table_bytes(x.cap);
                      k_extpi<MODE><<<nb, WAVE, lds, stream()>>>(x);
                      HX_LAUNCH_CHECK();
                   },
                   init_list, init_n);
}

// ------------------------------------------------------------------ ext+i, fast path
// Same arithmetic as k_extpi, restructured so that a row costs a few dependent
// global loads instead of one chain per neighbour, with ~2.5 KB of LDS:
//   * a row's work is cut into segments held one per lane (segment s <-> lane
//     s); the flattened entries of all segments are read 64 per chunk and
//     XF_U chunks per batch of independent loads; a lane finds its segment by
//     a binary search over the lanes' offsets (__shfl);
//   * C-hat is built over "strong neighbours, each strong F neighbour
//     expanded to its strong row": the first occurrence of a key is the
//     position that wins an atomicMin on the slot's value (used as a stamp
//     until the key gets its C-hat position), and first occurrences get
//     positions by ballot prefix, i.e. the serial first-seen order;
//   * weights: segments are the off-diagonal entries of row i (a C-hat entry
//     is a one-entry segment, a strong F neighbour i1 spans row i1).  Pass A
//     folds each denominator left to right over the qualifying entries of row
//     i1; pass B adds the distributed values into the weights one segment per
//     round, so every weight receives its terms in the serial order (the
//     columns within one segment are distinct);
//   * truncation runs afterwards in k_trunc_rows, one thread per row.
// Rows with more than 64 strong neighbours or more than XF_MC C-hat points
// are handed to k_extpi.
constexpr int XF_HC = 256;   // hash slots (<= 64 F markers + <= XF_MC + 64 C keys)
constexpr int XF_MC = 32;    // C-hat points per row on the fast path
constexpr int XF_U = 3;      // chunks of 64 entries per batch of loads
constexpr int XF_QC = 176;   // C-hat terms of 64 entries of row i kept for pass B
// (LDS of the fast kernel stays under 5 KB: 32 waves per CU)

struct ExtpiFastArgs
{
   int           n;
   const int    *A_ia, *A_ja;
   const double *A_a;
   const int    *S_ia, *S_ja;
   const int    *SC_cnt, *SC_ja;  // strong C neighbours of row r: SC_ja[S_ia[r] + t], t < SC_cnt[r]
   const int    *OC_cnt;          // opposite-sign C entries of row r: OC[A_ia[r] + t], t < OC_cnt[r]
   const OCEntry *OC;
   const int    *CF, *f2c;
   const int    *rowlist;     // rows to process (nullptr: 0..n-1)
   int           nwork;
   int          *raw_j;       // work item w at w * MC
   double       *raw_a;
   int          *cnt;         // entries of the raw row of work item w, -1 = deferred
   int          *defer_list;
   int          *ndefer;
   int          *claim;       // RowClaim counters
};

// last segment s in [0, nseg) with off_s <= q; segment s's offset is held by
// lane s (non-decreasing, off_0 = 0 <= q).  Uniform: every lane runs 6 steps.
__device__ __forceinline__ int seg_find_lanes(int off_reg, int nseg, int q)
{
   int lo = 0, hi = nseg - 1;
#pragma unroll
   for (int it = 0; it < 6; it++) {
      const int mid = (lo + hi + 1) >> 1;
      const int om = __shfl(off_reg, mid, WAVE);
      if (lo < hi) {
         if (om <= q) { lo = mid; } else { hi = mid - 1; }
      }
   }
   return lo;
}

__device__ __forceinline__ double shfl_d(double v, int src)
{
   return __shfl(v, src, WAVE);
}

template <int MC, int QC>
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(8))) void k_extpi_fast(ExtpiFastArgs p)
{
   __shared__ int keys[XF_HC];
   __shared__ int vals[XF_HC];
   __shared__ double wgt[MC];
   __shared__ int chat_key[MC];
   __shared__ int chat_col[MC];
   __shared__ int mins[MC];
   __shared__ int seg_ki[WAVE];
   __shared__ double q_a[QC];
   __shared__ int q_ts[QC];   // weight index << 9 | 0x100 (C-hat entry) | segment lane
   const int lane = threadIdx.x;
   RowHash H{keys, vals, XF_HC, ilog2(XF_HC)};
   RowClaim rc(p.claim, p.nwork);
   rc.issue(lane);
   for (int wi = rc.take(lane); wi >= 0; wi = rc.take(lane)) {
      rc.issue(lane);   // the next row's claim is in flight while this row runs
      // raw row and count are indexed by row, or by list position on a row list
      const int i = p.rowlist ? p.rowlist[wi] : wi;
      const size_t ro = (size_t) wi * MC;
      const int cfi = p.CF[i];
      if (cfi >= 0 || cfi == -3) {
         if (lane == 0) {
            if (cfi >= 0) {
               p.raw_j[ro] = p.f2c[i];
               p.raw_a[ro] = 1.0;
               p.cnt[wi] = 1;
            } else {
               p.cnt[wi] = 0;
            }
         }
         continue;
      }
      const int s0 = p.S_ia[i], ns = p.S_ia[i + 1] - s0;
      const int a0 = p.A_ia[i], na = p.A_ia[i + 1] - a0 - 1;
      bool defer = ns > WAVE;
      if (!defer) {
         for (int s = lane; s < XF_HC; s += WAVE) { keys[s] = -1; vals[s] = 0x7fffffff; }
         // ---- strong neighbours: kind 1 = C, 2 = F (par_lr_interp.c:1505-1560)
         int i1 = -1, kind = 0, kb = 0, klen = 0;
         if (lane < ns) {
            i1 = p.S_ja[s0 + lane];
            const int c1 = p.CF[i1];
            if (c1 >= 0) {
               kind = 1; klen = 1; kb = i1;
            } else if (c1 != -3) {
               // the candidates of a strong F neighbour are its strong C neighbours
               kind = 2; kb = p.S_ia[i1]; klen = p.SC_cnt[i1];
            }
         }
         wave_sync();
         if (kind == 2) {
            bool isn;
            const int sl = H.insert(i1, isn);
            if (sl >= 0) { vals[sl] = -2; }
         }
         const int incl = wave_incl_scan(klen, lane);
         const int L = readlane_i(incl, WAVE - 1);
         const int off = incl - klen;
         // up to 2 QC candidates: candidate q's source tabulated by its
         // segment's lane in the term space (free until pass B): ~column for
         // a strong C neighbour, else its position in SC_ja; longer lists
         // search the segment offsets
         int *ctab = reinterpret_cast<int *>(q_a);
         const bool ctab_on = L <= 2 * QC;
         if (ctab_on) {
            for (int t = 0; t < klen; t++) { ctab[off + t] = (kind == 1) ? ~kb : kb + t; }
         }
         wave_sync();
         // ---- C-hat in first-seen order.  vals[] of a C key holds the
         // smallest position seen (>= 64 after the first chunk) until the key
         // receives its C-hat position (< MC).
         int nch = 0;
         const int nseg = max(ns, 1);
         for (int cb = 0; cb < L && !defer; cb += WAVE * XF_U) {
            int key[XF_U];
#pragma unroll
            for (int u = 0; u < XF_U; u++) {
               const int q = cb + u * WAVE + lane;
               key[u] = -1;
               if (ctab_on) {
                  if (q < L) {
                     const int e = ctab[q];
                     key[u] = (e < 0) ? ~e : p.SC_ja[e];
                  }
               } else {
                  const int s = seg_find_lanes(off, nseg, q);
                  const int kind_s = __shfl(kind, s, WAVE);
                  const int b_s = __shfl(kb, s, WAVE);
                  const int off_s = __shfl(off, s, WAVE);
                  if (q < L) {
                     if (kind_s == 1) {
                        key[u] = b_s;
                     } else if (kind_s == 2) {
                        key[u] = p.SC_ja[b_s + (q - off_s)];
                     }
                  }
               }
            }
#pragma unroll
            for (int u = 0; u < XF_U; u++) {
               if (cb + u * WAVE >= L || defer) { break; }
               const int q = cb + u * WAVE + lane;
               int sl = -1;
               if (key[u] >= 0) {
                  bool isn;
                  sl = H.insert(key[u], isn);
                  if (sl >= 0) { atomicMin(&vals[sl], q); }
               }
               wave_sync();
               const bool first = (sl >= 0) && vals[sl] == q;
               const unsigned long long fm = __ballot(first);
               const int nnew = __popcll(fm);
               if (nch + nnew > MC) {
                  defer = true;
               } else if (first) {
                  const int pos = nch + __popcll(fm & lanemask_lt(lane));
                  vals[sl] = pos;
                  chat_key[pos] = key[u];
               }
               nch += nnew;
               wave_sync();
            }
         }
         if (!defer) {
            if (lane < nch) {
               chat_col[lane] = p.f2c[chat_key[lane]];
               wgt[lane] = 0.0;
            }
            if (lane < MC) { mins[lane] = 0x7fffffff; }
            wave_sync();
            // ---- weights (par_lr_interp.c:1640-1805), 64 entries of row i per block
            double diag = p.A_a[a0];
            for (int jb = 0; jb < na; jb += WAVE) {
               const int j = jb + lane;
               // entry j of row i: kind 1 C-hat (weight index v), 2 strong F
               // neighbour c1, 3 weak, 0 none
               int c1 = 0, v = -1, sk = 0;
               int ib = 0, ilen = 0;          // row c1 without its diagonal (i search)
               int ob = 0, olen = 0;          // opposite-sign C entries of row c1
               bool sneg = false;
               double aij = 0.0;
               if (j < na) {
                  c1 = p.A_ja[a0 + 1 + j];
                  aij = p.A_a[a0 + 1 + j];
                  const int sl = H.find(c1);
                  v = (sl >= 0) ? vals[sl] : -1;
                  if (v >= 0) {
                     sk = 1; olen = 1;
                  } else if (v == -2) {
                     const int b1 = p.A_ia[c1];
                     sk = 2;
                     ib = b1 + 1;
                     ilen = p.A_ia[c1 + 1] - b1 - 1;
                     sneg = p.A_a[b1] < 0;
                     ob = b1;
                     olen = p.OC_cnt[c1];
                  } else if (p.CF[c1] != -3) {
                     sk = 3;
                  }
               }
               const int nseg2 = min(WAVE, na - jb);
               // pass I: the entry of column i in each strong F neighbour row
               const int iincl = wave_incl_scan(ilen, lane);
               const int IL = readlane_i(iincl, WAVE - 1);
               const int ioff = iincl - ilen;
               seg_ki[lane] = -1;
               wave_sync();
               for (int cb = 0; cb < IL; cb += WAVE * XF_U) {
                  int kk[XF_U], sq[XF_U], cv[XF_U];
#pragma unroll
                  for (int u = 0; u < XF_U; u++) {
                     const int q = cb + u * WAVE + lane;
                     const int s = seg_find_lanes(ioff, nseg2, q);
                     const int b_s = __shfl(ib, s, WAVE);
                     const int o_s = __shfl(ioff, s, WAVE);
                     sq[u] = s;
                     kk[u] = b_s + (q - o_s);
                     cv[u] = (q < IL) ? p.A_ja[kk[u]] : -1;   // compared after all loads
                  }
#pragma unroll
                  for (int u = 0; u < XF_U; u++) {
                     if (cv[u] == i) { seg_ki[sq[u]] = kk[u]; }   // one hit per row at most
                  }
               }
               wave_sync();
               const int ki = seg_ki[lane];
               double sai = 0.0;
               bool shas_i = false;
               if (sk == 2 && ki >= 0) {
                  sai = p.A_a[ki];
                  shas_i = sneg ? (sai > 0) : (sai < 0);
               }
               // pass A, one lane per strong F neighbour c1: the denominator
               // sums, in column order, the opposite-sign entries of row c1
               // whose column is in C-hat, with the entry of column i merged
               // in by its storage index.  The C-hat terms for pass B are
               // counted here and written below at prefix-sum offsets.
               double ssum = 0.0;
               int nq = (sk == 1) ? 1 : 0;
               if (sk == 2) {
                  double acc = 0.0;
                  bool pend = shas_i;
                  for (int e = ob; e < ob + olen; e++) {
                     const OCEntry en = p.OC[e];
                     const int s2 = H.find(en.j);
                     if (s2 >= 0 && vals[s2] >= 0) {
                        if (pend && ki < en.k) {
                           acc += sai;
                           pend = false;
                        }
                        acc += en.a;
                        nq++;
                     }
                  }
                  if (pend) { acc += sai; }
                  ssum = acc;
               }
               const int qincl = wave_incl_scan(nq, lane);
               const int qn = readlane_i(qincl, WAVE - 1);
               if (qn > QC) { defer = true; break; }
               {
                  int pos = qincl - nq;
                  if (sk == 1) {
                     q_ts[pos] = (v << 9) | 0x100 | lane;
                     q_a[pos] = aij;
                  } else if (sk == 2) {
                     for (int e = ob; e < ob + olen; e++) {
                        const OCEntry en = p.OC[e];
                        const int s2 = H.find(en.j);
                        const int v2 = (s2 >= 0) ? vals[s2] : -1;
                        if (v2 >= 0) {
                           q_ts[pos] = (v2 << 9) | lane;
                           q_a[pos] = en.a;
                           pos++;
                        }
                     }
                  }
               }
               // per entry: distribution factor and diagonal term, in entry order
               bool hasd = false, sdist = false;
               double dval = 0.0, dist = 0.0;
               if (sk == 2) {
                  if (ssum != 0) {
                     dist = aij / ssum;
                     sdist = true;
                     if (shas_i) { hasd = true; dval = dist * sai; }
                  } else {
                     hasd = true;
                     dval = aij;
                  }
               } else if (sk == 3) {
                  hasd = true;                        // weak connection
                  dval = aij;
               }
               unsigned long long dm = __ballot(hasd);
               while (dm) {
                  const int t = __ffsll((long long) dm) - 1;
                  diag += readlane_d(dval, t);
                  dm &= dm - 1;
               }
               // pass B: the C-hat terms collected in pass A, in order: a_ij of a
               // C-hat entry, distribute * a of a strong F neighbour with a
               // nonzero denominator.  Terms of one step that hit the same weight
               // are applied lowest lane first (atomicMin on mins[weight]).
               wave_sync();
               for (int qb = 0; qb < qn; qb += WAVE) {
                  const int e = qb + lane;
                  const bool valid = e < qn;
                  const int ts = valid ? q_ts[e] : 0;
                  const int tg = ts >> 9;
                  const int sgw = ts & 0x1ff;
                  const int sg = sgw & 0xff;
                  const bool dr = (sgw & 0x100) != 0;
                  const double a = valid ? q_a[e] : 0.0;
                  const int d_s = __shfl((int) sdist, sg, WAVE);
                  const double dist_s = shfl_d(dist, sg);
                  bool pending = valid && (dr || d_s);
                  const double x = dr ? a : dist_s * a;
                  while (__ballot(pending)) {
                     if (pending) { atomicMin(&mins[tg], lane); }
                     wave_sync();
                     if (pending && mins[tg] == lane) {
                        wgt[tg] += x;
                        mins[tg] = 0x7fffffff;
                        pending = false;
                     }
                     wave_sync();
                  }
               }
               wave_sync();
            }
            wave_sync();
            if (!defer) {
               if (lane < nch) {
                  double w = wgt[lane];
                  if (diag != 0.0) { w /= -diag; }
                  p.raw_j[ro + lane] = chat_col[lane];
                  p.raw_a[ro + lane] = w;
               }
               if (lane == 0) { p.cnt[wi] = nch; }
            }
         }
      }
      if (defer && lane == 0) {
         p.cnt[wi] = -1;
         const int d = atomicAdd(p.ndefer, 1);
         p.defer_list[d] = i;
      }
      wave_sync();
   }
}

// ------------------------------------------------------------------ ext+i, rows of <= 64 entries
// The common case (<= 64 strong neighbours, <= 64 off-diagonal entries, the
// weight terms in <= XR_OL positions) with the arithmetic of k_extpi_fast but
// about five dependent rounds of global loads per row instead of ~20:
//   T1 row offsets and C/F state; T2 strong neighbours (lane s <-> S entry s)
//   and off-diagonal entries (lane j <-> entry j); T3 everything indexed by
//   those columns, read whether or not it is used; T4 the strong-C
//   candidates (C-hat, first-seen order), and for each neighbour row c1 the
//   entry at the mirror of c1's position in row i, which holds column i when
//   a symmetric stencil is stored in a symmetric order ("pass I"; rows where
//   it does not are searched afterwards); T5 the terms: one per C-hat entry
//   of row i, and the opposite-sign C entries of the row of every strong F
//   neighbour, in entry order.
// The denominators fold each row's terms left to right out of LDS, with the
// entry of column i merged in by storage index; the weights receive the terms
// straight from registers, 64 per step, lowest lane first on a shared weight.
// Rows outside these limits are deferred to k_extpi_fast<64, ...>.
constexpr int XR_U = 4;      // chunks of 64 per batch of the candidate and pass I searches
constexpr int XR_OL = 192;   // term positions of a row (3 chunks)

template <int MC>
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(8))) void k_extpi_row64(ExtpiFastArgs p)
{
   __shared__ int keys[XF_HC];
   __shared__ int vals[XF_HC];   // C key: stamp, then C-hat position; F neighbour: -2 - S index
   __shared__ double wgt[MC];
   __shared__ int chat_key[MC];
   __shared__ unsigned long long wmask[MC];   // lanes of a step that hit a weight
   __shared__ int seg_ki[WAVE];     // entry j: index of column i in row A_ja[j], or -1
   __shared__ double st_a[XR_OL];   // term values
   __shared__ int st_k[XR_OL];      // term storage index if its column is in C-hat, else -1
   const int lane = threadIdx.x;
   RowHash H{keys, vals, XF_HC, ilog2(XF_HC)};
   // Rows are pipelined two deep: while row wi runs, the claim of the row
   // after the next is in flight and the next row's T1 values are loading.
   RowClaim rc(p.claim, p.nwork);
   rc.issue(lane);
   int wn = rc.take(lane);
   int n_i = 0, n_cf = 0, n_s0 = 0, n_s1 = 0, n_a0 = 0, n_a1 = 0;
   auto load_t1 = [&](int w) {
      n_i = p.rowlist ? p.rowlist[w] : w;
      n_cf = p.CF[n_i];
      n_s0 = p.S_ia[n_i];
      n_s1 = p.S_ia[n_i + 1];
      n_a0 = p.A_ia[n_i];
      n_a1 = p.A_ia[n_i + 1];
   };
   if (wn >= 0) {
      rc.issue(lane);
      load_t1(wn);
   }
   for (int wi = wn; wi >= 0; wi = wn) {
      // T1 of this row (loaded during the previous row)
      const int i = n_i;
      const int cfi = n_cf;
      const int s0 = n_s0, s1 = n_s1;
      const int a0 = n_a0, a1 = n_a1;
      wn = rc.take(lane);
      if (wn >= 0) {
         rc.issue(lane);
         load_t1(wn);
      }
      const size_t ro = (size_t) wi * MC;
      if (cfi >= 0 || cfi == -3) {
         if (lane == 0) {
            if (cfi >= 0) {
               p.raw_j[ro] = p.f2c[i];
               p.raw_a[ro] = 1.0;
               p.cnt[wi] = 1;
            } else {
               p.cnt[wi] = 0;
            }
         }
         continue;
      }
      const int ns = s1 - s0, na = a1 - a0 - 1;
      bool defer = ns > WAVE || na > WAVE;
      if (!defer) {
         // T2
         int i1 = -1, c1 = -1;
         double aij = 0.0;
         if (lane < ns) { i1 = p.S_ja[s0 + lane]; }
         if (lane < na) {
            c1 = p.A_ja[a0 + 1 + lane];
            aij = p.A_a[a0 + 1 + lane];
         }
         double diag = p.A_a[a0];
         for (int s = lane; s < XF_HC; s += WAVE) { keys[s] = -1; vals[s] = 0x7fffffff; }
         // T3
         int cf_i1 = 0, scb = 0, sce = 0;
         if (lane < ns) {
            cf_i1 = p.CF[i1];
            scb = p.S_ia[i1];
            sce = scb + p.SC_cnt[i1];
         }
         int cf_c1 = 0, cb1 = 0, ce1 = 0, ob = 0, oe = 0;
         if (lane < na) {
            cf_c1 = p.CF[c1];
            cb1 = p.A_ia[c1];
            ce1 = p.A_ia[c1 + 1];
            ob = cb1;
            oe = cb1 + p.OC_cnt[c1];
         }
         // strong neighbours: kind 1 = C, 2 = F (par_lr_interp.c:1505-1560); a
         // strong F neighbour contributes its strong C neighbours as candidates
         int kind = 0, kb = 0, klen = 0;
         if (lane < ns) {
            if (cf_i1 >= 0) {
               kind = 1; klen = 1; kb = i1;
            } else if (cf_i1 != -3) {
               kind = 2; kb = scb; klen = sce - scb;
            }
         }
         wave_sync();
         if (kind == 2) {
            bool isn;
            const int sl = H.insert(i1, isn);
            if (sl >= 0) { vals[sl] = -2 - lane; }
         }
         const int incl = wave_incl_scan(klen, lane);
         const int L = readlane_i(incl, WAVE - 1);
         const int off = incl - klen;
         const int nseg = max(ns, 1);
         // up to 2 XR_OL candidates: candidate q's source tabulated by its
         // segment's lane in the term space (free until the terms are
         // staged): ~column for a strong C neighbour, else its position in
         // SC_ja; longer lists search the segment offsets
         int *ctab = reinterpret_cast<int *>(st_a);
         const bool ctab_on = L <= 2 * XR_OL;
         if (ctab_on) {
            for (int t = 0; t < klen; t++) { ctab[off + t] = (kind == 1) ? ~kb : kb + t; }
         }
         // with the first batch: the diagonal of each neighbour row c1 (its
         // sign) and the entry at the mirror of c1's position in row i, where
         // a symmetric stencil stored in a symmetric order keeps column i
         double dg1 = 0.0;
         int gk = -1, gv = -1;
         if (lane < na && ce1 > cb1) {
            dg1 = p.A_a[cb1];
            const int g = ce1 - 1 - lane;
            if (g > cb1) {
               gk = g;
               gv = p.A_ja[g];
            }
         }
         wave_sync();
         // T4: the candidates
         int nch = 0;
         for (int cb = 0; cb < L && !defer; cb += WAVE * XR_U) {
            int key[XR_U];
#pragma unroll
            for (int u = 0; u < XR_U; u++) {
               const int c0 = cb + u * WAVE;   // uniform: the chunk's first position
               const int q = c0 + lane;
               key[u] = -1;
               if (ctab_on) {
                  if (q < L) {
                     const int e = ctab[q];
                     key[u] = (e < 0) ? ~e : p.SC_ja[e];
                  }
               } else if (c0 < L) {
                  const int sa = seg_find_lanes(off, nseg, q);
                  const int kind_s = __shfl(kind, sa, WAVE);
                  const int b_s = __shfl(kb, sa, WAVE);
                  const int o_s = __shfl(off, sa, WAVE);
                  if (q < L) {
                     if (kind_s == 1) {
                        key[u] = b_s;
                     } else if (kind_s == 2) {
                        key[u] = p.SC_ja[b_s + (q - o_s)];
                     }
                  }
               }
            }
#pragma unroll
            for (int u = 0; u < XR_U; u++) {
               if (cb + u * WAVE >= L || defer) { break; }
               const int q = cb + u * WAVE + lane;
               int sl = -1;
               if (key[u] >= 0) {
                  bool isn;
                  sl = H.insert(key[u], isn);
                  if (sl >= 0) { atomicMin(&vals[sl], q); }
               }
               wave_sync();
               const bool first = (sl >= 0) && vals[sl] == q;
               const unsigned long long fm = __ballot(first);
               const int nnew = __popcll(fm);
               if (nch + nnew > MC) {
                  defer = true;
               } else if (first) {
                  const int pos = nch + __popcll(fm & lanemask_lt(lane));
                  vals[sl] = pos;
                  chat_key[pos] = key[u];
               }
               nch += nnew;
               wave_sync();
            }
         }
         if (!defer) {
            // T5
            int ccol = 0;
            if (lane < nch) {
               ccol = p.f2c[chat_key[lane]];
               wgt[lane] = 0.0;
            }
            if (lane < MC) { wmask[lane] = 0ull; }
            // entry j: kind 1 C-hat (weight v), 2 strong F neighbour, 3 weak, 0 none
            int v = -1, sk = 0;
            if (lane < na) {
               const int sl = H.find(c1);
               v = (sl >= 0) ? vals[sl] : -1;
               if (v >= 0) {
                  sk = 1;
               } else if (v <= -2) {
                  sk = 2;
               } else if (cf_c1 != -3) {
                  sk = 3;
               }
            }
            // pass I: the entry of column i in the row of each strong F
            // neighbour; rows whose mirror guess missed are searched
            int ki = -1;
            const bool miss = (sk == 2) && (gv != i);
            if (sk == 2 && !miss) { ki = gk; }
            const int ilen = miss ? ce1 - cb1 - 1 : 0;
            const int iincl = wave_incl_scan(ilen, lane);
            const int IL = readlane_i(iincl, WAVE - 1);
            if (IL > 0) {
               const int ioff = iincl - ilen;
               const int ib = cb1 + 1;
               const int nsegi = max(na, 1);
               seg_ki[lane] = -1;
               wave_sync();
               for (int cb = 0; cb < IL; cb += WAVE * XR_U) {
                  int hs[XR_U], hk[XR_U], hv[XR_U];
#pragma unroll
                  for (int u = 0; u < XR_U; u++) {
                     const int q = cb + u * WAVE + lane;
                     hs[u] = -1;
                     hk[u] = 0;
                     hv[u] = -1;
                     if (cb + u * WAVE < IL) {   // uniform
                        const int si = seg_find_lanes(ioff, nsegi, q);
                        const int ib_s = __shfl(ib, si, WAVE);
                        const int io_s = __shfl(ioff, si, WAVE);
                        if (q < IL) {
                           hk[u] = ib_s + (q - io_s);
                           hs[u] = si;
                           hv[u] = p.A_ja[hk[u]];   // compared after all loads are issued
                        }
                     }
                  }
#pragma unroll
                  for (int u = 0; u < XR_U; u++) {
                     if (hs[u] >= 0 && hv[u] == i) { seg_ki[hs[u]] = hk[u]; }   // one hit per row
                  }
               }
               wave_sync();
               if (miss) { ki = seg_ki[lane]; }
            }
            double sai = 0.0;
            if (sk == 2 && ki >= 0) { sai = p.A_a[ki]; }
            const int olen = (sk == 1) ? 1 : ((sk == 2) ? oe - ob : 0);
            const int oincl = wave_incl_scan(olen, lane);
            const int OL = readlane_i(oincl, WAVE - 1);
            const int ooff = oincl - olen;
            if (OL > XR_OL) { defer = true; }
            if (!defer) {
               // term q's entry (lane) in st_k and its OC record's index in
               // the term space as ints (-1: the entry's own C-hat weight),
               // read before the terms are staged over them
               wave_sync();
               for (int t = 0; t < olen; t++) {
                  st_k[ooff + t] = lane;
                  ctab[ooff + t] = (sk == 1) ? -1 : ob + t;
               }
               wave_sync();
               int tg[3], ts[3], tcol[3], tk[3];
               bool dr[3];
               double ta[3];
               // the loads of all three chunks first, the lookups after
#pragma unroll
               for (int u = 0; u < 3; u++) {
                  const int q = u * WAVE + lane;
                  ts[u] = 0;
                  tg[u] = -1;
                  tcol[u] = -1;
                  tk[u] = -1;
                  dr[u] = false;
                  ta[u] = 0.0;
                  if (u * WAVE >= OL) { continue; }   // uniform
                  const bool in = q < OL;
                  const int s = in ? st_k[q] : 0;
                  const int e = in ? ctab[q] : -1;
                  const int v_s = __shfl(v, s, WAVE);
                  const double aij_s = shfl_d(aij, s);
                  ts[u] = s;
                  if (in) {
                     if (e < 0) {
                        dr[u] = true;
                        tg[u] = v_s;
                        ta[u] = aij_s;
                     } else {
                        const OCEntry en = p.OC[e];
                        tcol[u] = en.j;
                        tk[u] = en.k;
                        ta[u] = en.a;
                     }
                  }
               }
               wave_sync();   // the tables are read before the terms are staged
#pragma unroll
               for (int u = 0; u < 3; u++) {
                  const int q = u * WAVE + lane;
                  if (tcol[u] >= 0) {
                     const int sl = H.find(tcol[u]);
                     const int v2 = (sl >= 0) ? vals[sl] : -1;
                     tg[u] = (v2 >= 0) ? v2 : -1;
                     st_a[q] = ta[u];
                     st_k[q] = (v2 >= 0) ? tk[u] : -1;
                  }
               }
               wave_sync();
               // denominators (par_lr_interp.c:1700-1760): the terms of row c1
               // whose column is in C-hat, in column order, with the entry of
               // column i merged in by its storage index
               double ssum = 0.0;
               bool shas_i = false;
               if (sk == 2) {
                  shas_i = (ki >= 0) && ((dg1 < 0) ? (sai > 0) : (sai < 0));
                  double acc = 0.0;
                  bool pend = shas_i;
                  // four terms' LDS reads issued together, then folded in order
                  for (int q0 = ooff; q0 < ooff + olen; q0 += 4) {
                     int kq[4];
                     double aq[4];
#pragma unroll
                     for (int v = 0; v < 4; v++) {
                        kq[v] = -1;
                        aq[v] = 0.0;
                        if (q0 + v < ooff + olen) {
                           kq[v] = st_k[q0 + v];
                           aq[v] = st_a[q0 + v];
                        }
                     }
#pragma unroll
                     for (int v = 0; v < 4; v++) {
                        if (kq[v] >= 0) {
                           if (pend && ki < kq[v]) {
                              acc += sai;
                              pend = false;
                           }
                           acc += aq[v];
                        }
                     }
                  }
                  if (pend) { acc += sai; }
                  ssum = acc;
               }
               // per entry: distribution factor and diagonal term, in entry order
               bool hasd = false, sdist = false;
               double dval = 0.0, dist = 0.0;
               if (sk == 2) {
                  if (ssum != 0) {
                     dist = aij / ssum;
                     sdist = true;
                     if (shas_i) { hasd = true; dval = dist * sai; }
                  } else {
                     hasd = true;
                     dval = aij;
                  }
               } else if (sk == 3) {
                  hasd = true;   // weak connection
                  dval = aij;
               }
               // the diagonal terms in entry order, folded by lane 0 from LDS
               // (the term values' space, free once the denominators are
               // formed), four reads in flight
               double *xs = st_a;
               const unsigned long long dm = __ballot(hasd);
               wave_sync();
               xs[lane] = dval;
               wave_sync();
               if (lane == 0) {
                  double acc = diag;
                  for (unsigned long long rest = dm; rest;) {
                     int tq[4];
                     double vq[4];
#pragma unroll
                     for (int v = 0; v < 4; v++) {
                        tq[v] = -1;
                        vq[v] = 0.0;
                        if (rest) {
                           tq[v] = __ffsll((long long) rest) - 1;
                           vq[v] = xs[tq[v]];
                           rest &= rest - 1;
                        }
                     }
#pragma unroll
                     for (int v = 0; v < 4; v++) {
                        if (tq[v] >= 0) { acc += vq[v]; }
                     }
                  }
                  diag = acc;
               }
               diag = readlane_d(diag, 0);
               // weights: a_ij of a C-hat entry, distribute * a of a strong F
               // neighbour with a nonzero denominator; the terms of one step
               // that hit a weight are added by the weight's lowest lane, in
               // lane order, from the step's terms in LDS
               wave_sync();
#pragma unroll
               for (int u = 0; u < 3; u++) {
                  if (u * WAVE >= OL) { continue; }   // uniform
                  const int d_s = __shfl((int) sdist, ts[u], WAVE);
                  const double dist_s = shfl_d(dist, ts[u]);
                  const bool pending = (tg[u] >= 0) && (dr[u] || d_s);
                  const double x = dr[u] ? ta[u] : dist_s * ta[u];
                  if (pending) { atomicOr(&wmask[tg[u]], 1ull << lane); }
                  xs[lane] = x;
                  wave_sync();
                  if (pending) {
                     const unsigned long long m = wmask[tg[u]];
                     if ((m & lanemask_lt(lane)) == 0ull) {
                        double acc = wgt[tg[u]];
                        acc += x;
                        for (unsigned long long rest = m & (m - 1); rest; rest &= rest - 1) {
                           acc += xs[__ffsll((long long) rest) - 1];
                        }
                        wgt[tg[u]] = acc;
                        wmask[tg[u]] = 0ull;
                     }
                  }
                  wave_sync();
               }
               wave_sync();
               if (lane < nch) {
                  double w = wgt[lane];
                  if (diag != 0.0) { w /= -diag; }
                  p.raw_j[ro + lane] = ccol;
                  p.raw_a[ro + lane] = w;
               }
               if (lane == 0) { p.cnt[wi] = nch; }
            }
         }
      }
      if (defer && lane == 0) {
         p.cnt[wi] = -1;
         const int d = atomicAdd(p.ndefer, 1);
         p.defer_list[d] = i;
      }
      wave_sync();
   }
}

// ------------------------------------------------------------------ ext+i, several rows per wave
// k_extpi_row64 for rows of <= G strong neighbours and <= G entries (the fine
// levels of 3-D 7- and 27-point problems): each group of G lanes runs one row
// with the same steps, so a wave keeps 64/G rows' dependent loads in flight.
// Every scan, segment search and shuffle stays inside its group; loops run
// while any group has work, the other groups' lanes idling through them.
// Rows outside these limits (or with more than MC <= G C-hat points or 3G
// terms) are deferred to k_extpi_row64.
template <int G>
__device__ __forceinline__ int group_incl_scan(int v, int gl)
{
#pragma unroll
   for (int d = 1; d < G; d <<= 1) {
      const int t = __shfl_up(v, d, G);
      if (gl >= d) { v += t; }
   }
   return v;
}

// last segment s in [0, nseg) with off_s <= q inside a group of G lanes
template <int G>
__device__ __forceinline__ int group_seg_find(int off_reg, int nseg, int q)
{
   int lo = 0, hi = nseg - 1;
#pragma unroll
   for (int it = 0; it < ilog2(G); it++) {
      const int mid = (lo + hi + 1) >> 1;
      const int om = __shfl(off_reg, mid, G);
      if (lo < hi) {
         if (om <= q) { lo = mid; } else { hi = mid - 1; }
      }
   }
   return lo;
}

template <int G, int MC>
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(5))) void k_extpi_rowg(ExtpiFastArgs p)
{
   static_assert(MC <= G && G <= 32, "one C-hat point per lane of a group");
   constexpr int R = WAVE / G;     // rows per wave
   constexpr int HC = 4 * G;       // hash slots per row
   constexpr int OLC = 3 * G;      // term positions per row (3 chunks)
   __shared__ int keys_all[R * HC];
   __shared__ int vals_all[R * HC];
   __shared__ double wgt_all[R * MC];
   __shared__ int chat_key_all[R * MC];
   __shared__ unsigned long long wmask_all[R * MC];   // lanes of a step that hit a weight
   __shared__ int seg_ki_all[WAVE];
   __shared__ double st_a_all[R * OLC];
   __shared__ int st_k_all[R * OLC];
   const int lane = threadIdx.x;
   const int grp = lane / G, gl = lane % G;
   const unsigned long long gbits = ((1ull << G) - 1) << (grp * G);
   int *keys = keys_all + grp * HC;
   int *vals = vals_all + grp * HC;
   double *wgt = wgt_all + grp * MC;
   int *chat_key = chat_key_all + grp * MC;
   unsigned long long *wmask = wmask_all + grp * MC;
   int *seg_ki = seg_ki_all + grp * G;
   double *st_a = st_a_all + grp * OLC;
   int *st_k = st_k_all + grp * OLC;
   RowHash H{keys, vals, HC, ilog2(HC)};
   // work item w is the set of rows R w .. R w + R - 1; work items are
   // pipelined as in k_extpi_row64 (next item's claim and T1 values in flight)
   const int nitems = (p.nwork + R - 1) / R;
   RowClaim rc(p.claim, nitems);
   rc.issue(lane);
   int wn = rc.take(lane);
   int n_i = 0, n_cf = -1, n_s0 = 0, n_s1 = 0, n_a0 = 0, n_a1 = 0;
   auto load_t1 = [&](int w) {
      const int r = R * w + grp;
      if (r < p.nwork) {
         n_i = p.rowlist ? p.rowlist[r] : r;
         n_cf = p.CF[n_i];
         n_s0 = p.S_ia[n_i];
         n_s1 = p.S_ia[n_i + 1];
         n_a0 = p.A_ia[n_i];
         n_a1 = p.A_ia[n_i + 1];
      }
   };
   if (wn >= 0) {
      rc.issue(lane);
      load_t1(wn);
   }
   for (int wp = wn; wp >= 0; wp = wn) {
      const int wi = R * wp + grp;
      const bool has_row = wi < p.nwork;
      const int i = n_i;
      const int cfi = n_cf;
      const int s0 = n_s0, s1 = n_s1;
      const int a0 = n_a0, a1 = n_a1;
      wn = rc.take(lane);
      if (wn >= 0) {
         rc.issue(lane);
         load_t1(wn);
      }
      const size_t ro = (size_t) wi * MC;
      bool live = has_row;
      if (has_row && (cfi >= 0 || cfi == -3)) {
         if (gl == 0) {
            if (cfi >= 0) {
               p.raw_j[ro] = p.f2c[i];
               p.raw_a[ro] = 1.0;
               p.cnt[wi] = 1;
            } else {
               p.cnt[wi] = 0;
            }
         }
         live = false;
      }
      const int ns = live ? s1 - s0 : 0, na = live ? a1 - a0 - 1 : 0;
      bool defer = live && (ns > G || na > G);
      live = live && !defer;
      // T2
      int i1 = -1, c1 = -1;
      double aij = 0.0, diag = 0.0;
      if (live && gl < ns) { i1 = p.S_ja[s0 + gl]; }
      if (live && gl < na) {
         c1 = p.A_ja[a0 + 1 + gl];
         aij = p.A_a[a0 + 1 + gl];
      }
      if (live) { diag = p.A_a[a0]; }
      for (int s = gl; s < HC; s += G) { keys[s] = -1; vals[s] = 0x7fffffff; }
      // T3
      int cf_i1 = 0, scb = 0, sce = 0;
      if (live && gl < ns) {
         cf_i1 = p.CF[i1];
         scb = p.S_ia[i1];
         sce = scb + p.SC_cnt[i1];
      }
      int cf_c1 = 0, cb1 = 0, ce1 = 0, ob = 0, oe = 0;
      if (live && gl < na) {
         cf_c1 = p.CF[c1];
         cb1 = p.A_ia[c1];
         ce1 = p.A_ia[c1 + 1];
         ob = cb1;
         oe = cb1 + p.OC_cnt[c1];
      }
      // strong neighbours: kind 1 = C, 2 = F (par_lr_interp.c:1505-1560)
      int kind = 0, kb = 0, klen = 0;
      if (live && gl < ns) {
         if (cf_i1 >= 0) {
            kind = 1; klen = 1; kb = i1;
         } else if (cf_i1 != -3) {
            kind = 2; kb = scb; klen = sce - scb;
         }
      }
      wave_sync();
      if (kind == 2) {
         bool isn;
         const int sl = H.insert(i1, isn);
         if (sl >= 0) { vals[sl] = -2 - gl; }
      }
      const int incl = group_incl_scan<G>(klen, gl);
      const int L = __shfl(incl, G - 1, G);
      const int off = incl - klen;
      // candidate q's source, tabulated by its segment's lane in the group's
      // term space (free until the terms are staged): ~column for a strong C
      // neighbour, else the position of the candidate in SC_ja
      int *ctab = reinterpret_cast<int *>(st_a);
      if (live && L > 2 * OLC) {
         defer = true;
         live = false;
      }
      if (live) {
         for (int t = 0; t < klen; t++) { ctab[off + t] = (kind == 1) ? ~kb : kb + t; }
      }
      // the diagonal of each neighbour row and the mirror guess of column i
      double dg1 = 0.0;
      int gk = -1, gv = -1;
      if (live && gl < na && ce1 > cb1) {
         dg1 = p.A_a[cb1];
         const int g = ce1 - 1 - gl;
         if (g > cb1) {
            gk = g;
            gv = p.A_ja[g];
         }
      }
      wave_sync();
      // T4: the candidates, both rows' chunks in the same steps
      int nch = 0;
      for (int cb = 0; __any(live && cb < L); cb += G * XR_U) {
         int key[XR_U];
#pragma unroll
         for (int u = 0; u < XR_U; u++) {
            const int q = cb + u * G + gl;
            key[u] = -1;
            if (live && q < L) {
               const int e = ctab[q];
               key[u] = (e < 0) ? ~e : p.SC_ja[e];
            }
         }
#pragma unroll
         for (int u = 0; u < XR_U; u++) {
            const bool cact = live && (cb + u * G < L);   // uniform within the group
            const int q = cb + u * G + gl;
            int sl = -1;
            if (cact && key[u] >= 0) {
               bool isn;
               sl = H.insert(key[u], isn);
               if (sl >= 0) { atomicMin(&vals[sl], q); }
            }
            wave_sync();
            const bool first = cact && (sl >= 0) && vals[sl] == q;
            const unsigned long long fm = __ballot(first) & gbits;
            const int nnew = __popcll(fm);
            if (cact) {
               if (nch + nnew > MC) {
                  defer = true;
                  live = false;
               } else if (first) {
                  const int pos = nch + __popcll(fm & lanemask_lt(lane));
                  vals[sl] = pos;
                  chat_key[pos] = key[u];
               }
               nch += nnew;
            }
            wave_sync();
         }
      }
      // T5
      int ccol = 0;
      if (live && gl < nch) {
         ccol = p.f2c[chat_key[gl]];
         wgt[gl] = 0.0;
      }
      if (gl < MC) { wmask[gl] = 0ull; }
      // entry j: kind 1 C-hat (weight v), 2 strong F neighbour, 3 weak, 0 none
      int v = -1, sk = 0;
      if (live && gl < na) {
         const int sl = H.find(c1);
         v = (sl >= 0) ? vals[sl] : -1;
         if (v >= 0) {
            sk = 1;
         } else if (v <= -2) {
            sk = 2;
         } else if (cf_c1 != -3) {
            sk = 3;
         }
      }
      // pass I: the mirror guess, searched where it missed
      int ki = -1;
      const bool miss = (sk == 2) && (gv != i);
      if (sk == 2 && !miss) { ki = gk; }
      const int ilen = miss ? ce1 - cb1 - 1 : 0;
      const int iincl = group_incl_scan<G>(ilen, gl);
      const int IL = __shfl(iincl, G - 1, G);
      if (__any(IL > 0)) {
         const int ioff = iincl - ilen;
         const int ib = cb1 + 1;
         const int nsegi = max(na, 1);
         seg_ki[gl] = -1;
         wave_sync();
         for (int cb = 0; __any(cb < IL); cb += G * XR_U) {
            int hs[XR_U], hk[XR_U], hv[XR_U];
#pragma unroll
            for (int u = 0; u < XR_U; u++) {
               const int q = cb + u * G + gl;
               hs[u] = -1;
               hk[u] = 0;
               hv[u] = -1;
               if (cb + u * G < IL) {   // uniform within the group
                  const int si = group_seg_find<G>(ioff, nsegi, q);
                  const int ib_s = __shfl(ib, si, G);
                  const int io_s = __shfl(ioff, si, G);
                  if (q < IL) {
                     hk[u] = ib_s + (q - io_s);
                     hs[u] = si;
                     hv[u] = p.A_ja[hk[u]];
                  }
               }
            }
#pragma unroll
            for (int u = 0; u < XR_U; u++) {
               if (hs[u] >= 0 && hv[u] == i) { seg_ki[hs[u]] = hk[u]; }
            }
         }
         wave_sync();
         if (miss) { ki = seg_ki[gl]; }
      }
      double sai = 0.0;
      if (sk == 2 && ki >= 0) { sai = p.A_a[ki]; }
      const int olen = (sk == 1) ? 1 : ((sk == 2) ? oe - ob : 0);
      const int oincl = group_incl_scan<G>(olen, gl);
      const int OL = __shfl(oincl, G - 1, G);
      const int ooff = oincl - olen;
      if (live && OL > OLC) {
         defer = true;
         live = false;
      }
      // term q's entry (lane) and source, tabulated by the entry's lane:
      // st_k holds the lane, the term space as ints the OC record's index
      // (-1: the entry's own C-hat weight); both are read before the terms
      // are staged over them
      wave_sync();
      if (live) {
         for (int t = 0; t < olen; t++) {
            st_k[ooff + t] = gl;
            ctab[ooff + t] = (sk == 1) ? -1 : ob + t;
         }
      }
      wave_sync();
      int tg[3], ts[3], tcol[3], tk[3];
      bool dr[3];
      double ta[3];
#pragma unroll
      for (int u = 0; u < 3; u++) {
         const int q = u * G + gl;
         ts[u] = 0;
         tg[u] = -1;
         tcol[u] = -1;
         tk[u] = -1;
         dr[u] = false;
         ta[u] = 0.0;
         if (!live || u * G >= OL) { continue; }   // uniform within the group
         const bool in = q < OL;
         const int s = in ? st_k[q] : 0;
         const int e = in ? ctab[q] : -1;
         const int v_s = __shfl(v, s, G);
         const double aij_s = __shfl(aij, s, G);
         ts[u] = s;
         if (in) {
            if (e < 0) {
               dr[u] = true;
               tg[u] = v_s;
               ta[u] = aij_s;
            } else {
               const OCEntry en = p.OC[e];
               tcol[u] = en.j;
               tk[u] = en.k;
               ta[u] = en.a;
            }
         }
      }
      wave_sync();   // the tables are read before the terms are staged
#pragma unroll
      for (int u = 0; u < 3; u++) {
         const int q = u * G + gl;
         if (tcol[u] >= 0) {
            const int sl = H.find(tcol[u]);
            const int v2 = (sl >= 0) ? vals[sl] : -1;
            tg[u] = (v2 >= 0) ? v2 : -1;
            st_a[q] = ta[u];
            st_k[q] = (v2 >= 0) ? tk[u] : -1;
         }
      }
      wave_sync();
      // denominators (par_lr_interp.c:1700-1760)
      double ssum = 0.0;
      bool shas_i = false;
      if (live && sk == 2) {
         shas_i = (ki >= 0) && ((dg1 < 0) ? (sai > 0) : (sai < 0));
         double acc = 0.0;
         bool pend = shas_i;
         // four terms' LDS reads issued together, then folded in order
         for (int q0 = ooff; q0 < ooff + olen; q0 += 4) {
            int kq[4];
            double aq[4];
#pragma unroll
            for (int v = 0; v < 4; v++) {
               kq[v] = -1;
               aq[v] = 0.0;
               if (q0 + v < ooff + olen) {
                  kq[v] = st_k[q0 + v];
                  aq[v] = st_a[q0 + v];
               }
            }
#pragma unroll
            for (int v = 0; v < 4; v++) {
               if (kq[v] >= 0) {
                  if (pend && ki < kq[v]) {
                     acc += sai;
                     pend = false;
                  }
                  acc += aq[v];
               }
            }
         }
         if (pend) { acc += sai; }
         ssum = acc;
      }
      bool hasd = false, sdist = false;
      double dval = 0.0, dist = 0.0;
      if (live) {
         if (sk == 2) {
            if (ssum != 0) {
               dist = aij / ssum;
               sdist = true;
               if (shas_i) { hasd = true; dval = dist * sai; }
            } else {
               hasd = true;
               dval = aij;
            }
         } else if (sk == 3) {
            hasd = true;   // weak connection
            dval = aij;
         }
      }
      // each row's diagonal terms in entry order, folded by the group's first
      // lane from LDS (the term values' space, free once the denominators are
      // formed), four reads in flight
      double *xs = st_a_all;
      const unsigned long long dm = __ballot(hasd) & gbits;
      wave_sync();
      xs[lane] = dval;
      wave_sync();
      if (gl == 0) {
         double acc = diag;
         for (unsigned long long rest = dm; rest;) {
            int tq[4];
            double vq[4];
#pragma unroll
            for (int v = 0; v < 4; v++) {
               tq[v] = -1;
               vq[v] = 0.0;
               if (rest) {
                  tq[v] = __ffsll((long long) rest) - 1;
                  vq[v] = xs[tq[v]];
                  rest &= rest - 1;
               }
            }
#pragma unroll
            for (int v = 0; v < 4; v++) {
               if (tq[v] >= 0) { acc += vq[v]; }
            }
         }
         diag = acc;
      }
      diag = __shfl(diag, 0, G);
      // weights: the terms of one step that hit a weight are added by the
      // weight's lowest lane, in lane order, from the step's terms in LDS
      wave_sync();
#pragma unroll
      for (int u = 0; u < 3; u++) {
         const bool chunk = live && (u * G < OL);   // uniform within the group
         int d_s = 0;
         double dist_s = 0.0;
         if (chunk) {
            d_s = __shfl((int) sdist, ts[u], G);
            dist_s = __shfl(dist, ts[u], G);
         }
         const bool pending = chunk && (tg[u] >= 0) && (dr[u] || d_s);
         const double x = dr[u] ? ta[u] : dist_s * ta[u];
         if (__ballot(pending)) {
            if (pending) { atomicOr(&wmask[tg[u]], 1ull << lane); }
            xs[lane] = x;
            wave_sync();
            if (pending) {
               const unsigned long long m = wmask[tg[u]];
               if ((m & lanemask_lt(lane)) == 0ull) {
                  double acc = wgt[tg[u]];
                  acc += x;
                  for (unsigned long long rest = m & (m - 1); rest; rest &= rest - 1) {
                     acc += xs[__ffsll((long long) rest) - 1];
                  }
                  wgt[tg[u]] = acc;
                  wmask[tg[u]] = 0ull;
               }
            }
            wave_sync();
         }
      }
      wave_sync();
      if (live) {
         if (gl < nch) {
            double w = wgt[gl];
            if (diag != 0.0) { w /= -diag; }
            p.raw_j[ro + gl] = ccol;
            p.raw_a[ro + gl] = w;
         }
         if (gl == 0) { p.cnt[wi] = nch; }
      }
      if (has_row && defer && gl == 0) {
         p.cnt[wi] = -1;
         const int d = atomicAdd(p.ndefer, 1);
         p.defer_list[d] = i;
      }
      wave_sync();
   }
}

// hypre_qsort2_abs (src/utilities/qsort.c:452) on a row stored with stride SS:
// w holds the values, v the positions of the entries in the staged row (one
// byte each), refined only as far as the first `keep` positions need.  The
// stack of (left, right) pairs lives in two registers, 8 bits a level.
template <int SS>
__device__ void qsort2_abs_strided(unsigned char *v, double *w, int cnt, int keep)
{
   unsigned long long stl = 0ull, str = (unsigned long long) (cnt - 1);
   int sp = 1;
   while (sp > 0) {
      sp--;
      const int left = (int) ((stl >> (8 * sp)) & 0xffull);
      const int right = (int) ((str >> (8 * sp)) & 0xffull);
      if (left >= right || left >= keep) { continue; }
      const int mid = (left + right) / 2;
      { unsigned char t = v[left * SS]; v[left * SS] = v[mid * SS]; v[mid * SS] = t;
        double tw = w[left * SS]; w[left * SS] = w[mid * SS]; w[mid * SS] = tw; }
      int last = left;
      const double piv = fabs(w[left * SS]);
      for (int i = left + 1; i <= right; i++) {
         if (fabs(w[i * SS]) > piv) {
            ++last;
            unsigned char t = v[last * SS]; v[last * SS] = v[i * SS]; v[i * SS] = t;
            double tw = w[last * SS]; w[last * SS] = w[i * SS]; w[i * SS] = tw;
         }
      }
      { unsigned char t = v[left * SS]; v[left * SS] = v[last * SS]; v[last * SS] = t;
        double tw = w[left * SS]; w[left * SS] = w[last * SS]; w[last * SS] = tw; }
      const int l1 = left, r1 = last - 1, l2 = last + 1, r2 = right;
      // push the larger side first so the smaller one is refined next (the
      // stack then never holds more than log2(cnt) + 1 pairs); a side that
      // cannot reach the first `keep` positions is dropped here
      const bool big1 = (r1 - l1) > (r2 - l2);
      const int la = big1 ? l1 : l2, ra = big1 ? r1 : r2;
      const int lb = big1 ? l2 : l1, rb = big1 ? r2 : r1;
      if (la < ra && la < keep) {
         const unsigned long long m = ~(0xffull << (8 * sp));
         stl = (stl & m) | ((unsigned long long) la << (8 * sp));
         str = (str & m) | ((unsigned long long) ra << (8 * sp));
         sp++;
      }
      if (lb < rb && lb < keep) {
         const unsigned long long m = ~(0xffull << (8 * sp));
         stl = (stl & m) | ((unsigned long long) lb << (8 * sp));
         str = (str & m) | ((unsigned long long) rb << (8 * sp));
         sp++;
      }
   }
}

// par_csr_matrix.c:3040-3320 (trunc_factor drop + rescale, then P_max_elmts
// with qsort2_abs and rescale), one thread per row.  The values of 64 rows are
// staged in LDS with a stride of 65 so that a wave's accesses fall into
// distinct banks; the columns stay in the raw rows and the sort moves one-byte
// positions instead, which keeps a block at MC x 585 + 256 bytes of LDS (8
// blocks of MC = 32 rows per CU).
constexpr int TR_ROWS = 64;
constexpr int TR_SS = 65;
constexpr int TR_U = 8;        // staging loads in flight per thread

template <int MC>
__global__ __launch_bounds__(TR_ROWS) void k_trunc_rows(int n, const int *__restrict__ rowlist,
                                                        const int *__restrict__ cnt_raw,
                                                        const int *__restrict__ raw_j,
                                                        const double *__restrict__ raw_a,
                                                        double trunc_factor, int max_elmts,
                                                        int out_stride, int *__restrict__ out_cnt,
                                                        int *__restrict__ out_j,
                                                        double *__restrict__ out_a)
{
   static_assert(MC <= 128, "one-byte positions and an 8-level stack");
   __shared__ double sw[MC * TR_SS];
   __shared__ unsigned char sv[MC * TR_SS];
   __shared__ int sc[TR_ROWS];
   const int t = threadIdx.x;
   for (long long base = (long long) blockIdx.x * TR_ROWS; base < n;
        base += (long long) gridDim.x * TR_ROWS) {
      const long long i = base + t;
      sc[t] = (i < n) ? cnt_raw[i] : -1;
      __syncthreads();
      // TR_U loads per thread issued before their LDS stores
#pragma unroll 1
      for (int k0 = t; k0 < TR_ROWS * MC; k0 += TR_ROWS * TR_U) {
         double wv[TR_U];
         bool ok[TR_U];
#pragma unroll
         for (int u = 0; u < TR_U; u++) {
            const int k = k0 + u * TR_ROWS;
            const bool in = k < TR_ROWS * MC;
            const int r = in ? k / MC : 0, e = in ? k % MC : 0;
            ok[u] = in && e < sc[r];
            if (ok[u]) { wv[u] = raw_a[(size_t) (base + r) * MC + e]; }
         }
#pragma unroll
         for (int u = 0; u < TR_U; u++) {
            if (ok[u]) {
               const int k = k0 + u * TR_ROWS;
               const int r = k / MC, e = k % MC;
               sw[e * TR_SS + r] = wv[u];
               sv[e * TR_SS + r] = (unsigned char) e;
            }
         }
      }
      __syncthreads();
      const int c = sc[t];
      if (c >= 0) {
         unsigned char *vj = sv + t;
         double *vw = sw + t;
         int kept = c;
         if (trunc_factor > 0) {
            double row_nrm = 0.0;
            for (int k = 0; k < c; k++) {
               const double a = fabs(vw[k * TR_SS]);
               row_nrm = (row_nrm < a) ? a : row_nrm;
            }
            const double drop_coeff = trunc_factor * row_nrm;
            double row_sum = 0.0, scale = 0.0;
            int o = 0;
            for (int k = 0; k < c; k++) {
               const double w = vw[k * TR_SS];
               row_sum += w;
               if (fabs(w) < drop_coeff) { continue; }
               scale += w;
               vw[o * TR_SS] = w;
               vj[o * TR_SS] = vj[k * TR_SS];
               o++;
            }
            if (scale != 0.) {
               if (scale != row_sum) {
                  scale = row_sum / scale;
                  for (int k = 0; k < o; k++) { vw[k * TR_SS] *= scale; }
               }
            }
            kept = o;
         }
         if (max_elmts > 0 && kept > max_elmts) {
            double row_sum = 0.0;
            for (int k = 0; k < kept; k++) { row_sum += vw[k * TR_SS]; }
            qsort2_abs_strided<TR_SS>(vj, vw, kept, max_elmts);
            double scale = 0.0;
            for (int k = 0; k < max_elmts; k++) { scale += vw[k * TR_SS]; }
            if (scale != 0.) {
               if (scale != row_sum) {
                  scale = row_sum / scale;
                  for (int k = 0; k < max_elmts; k++) { vw[k * TR_SS] *= scale; }
               }
            }
            kept = max_elmts;
         }
         const long long row = rowlist ? rowlist[i] : i;
         out_cnt[row] = kept;
         const int *rj = raw_j + (size_t) i * MC;
         for (int k = 0; k < kept; k++) {
            out_j[(size_t) row * out_stride + k] = rj[vj[k * TR_SS]];
            out_a[(size_t) row * out_stride + k] = vw[k * TR_SS];
         }
      }
      __syncthreads();
   }
}

__global__ void k_compact_rows(int n, const int *__restrict__ base, const int *__restrict__ cnt,
                               const int *__restrict__ out_ia, const int *__restrict__ in_j,
                               const double *__restrict__ in_a, int stride,
                               int *__restrict__ oj, double *__restrict__ oa)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      int b = base ? base[i] : i * stride;
      int o = out_ia[i], c = cnt[i];
      // four entries' loads issued together
      for (int k0 = 0; k0 < c; k0 += 4) {
         int jv[4];
         double av[4];
#pragma unroll
         for (int t = 0; t < 4; t++) {
            if (k0 + t < c) {
               jv[t] = in_j[b + k0 + t];
               av[t] = in_a[b + k0 + t];
            }
         }
#pragma unroll
         for (int t = 0; t < 4; t++) {
            if (k0 + t < c) {
               oj[o + k0 + t] = jv[t];
               oa[o + k0 + t] = av[t];
            }
         }
      }
   }
}

__global__ void k_cf_fix(int n, int *CF)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      if (CF[i] == -3) { CF[i] = -1; }
   }
}

void extpi_interp(const Csr &A, const Csr &S, int *CF, const int *f2c, int nc,
                  double trunc_factor, int max_elmts, Csr &P, const double *s_thr,
                  const int *s_flg)
{
   int n = A.nrows;
   TimedLaunch tl("setup_interp", 0.0);
   InterpArgs a{};
   a.n = n;
   a.A_ia = A.ia; a.A_ja = A.ja; a.A_a = A.a;
   a.S_ia = S.ia; a.S_ja = S.ja;
   a.CF = CF; a.f2c = f2c;
   a.trunc_factor = trunc_factor;
   a.max_elmts = max_elmts;
   int *cnt = dalloc_n<int>((size_t) n + 1);
   int *base = nullptr;
   int *tmp_j = nullptr;
   double *tmp_a = nullptr;
   int stride = 0;
   if (max_elmts > 0 && max_elmts <= XF_MC) {
      // fast path into raw rows, truncation per thread, leftovers by k_extpi
      stride = max_elmts;
      CLists CL;
      c_lists(A, S, CF, CL, s_thr, s_flg);
      ExtpiFastArgs f{};
      f.n = n;
      f.A_ia = A.ia; f.A_ja = A.ja; f.A_a = A.a;
      f.S_ia = S.ia; f.S_ja = S.ja;
      f.SC_cnt = CL.sc_cnt; f.SC_ja = CL.sc_ja;
      f.OC_cnt = CL.oc_cnt; f.OC = CL.oc;
      f.CF = CF; f.f2c = f2c;
      tmp_j = dalloc_n<int>((size_t) std::max(n, 1) * stride);
      tmp_a = dalloc_n<double>((size_t) std::max(n, 1) * stride);
      int *h = static_cast<int *>(pinned_scratch(64));
      int *lists[2] = {dalloc_n<int>((size_t) n + 1), dalloc_n<int>((size_t) n + 1)};
      const int *list = nullptr;
      int nwork = n;
      // passes over the rows left by the previous one: two rows per wave (rows
      // of <= 32 entries; tried where the mean row is that short), one row
      // per wave (C-hat <= 32 points; tried where the mean row has at most
      // 64 entries: on longer rows most C-hats exceed 32 and the pass only
      // defers them, e.g. 82% of the third level of the 27-point problem),
      // then C-hat <= 64 points
      int kinds[3];
      int nk = 0;
      const double mean_row = (double) A.nnz / std::max(n, 1);
      if (mean_row <= 12.0) {
         kinds[nk++] = 3;   // four rows per wave
      } else if (mean_row <= 28.0) {
         kinds[nk++] = 0;   // two rows per wave
      }
      if (mean_row <= 64.0) { kinds[nk++] = 1; }
      kinds[nk++] = 2;
      for (int pass = 0; pass < nk && nwork > 0; pass++) {
         const int kind = kinds[pass];
         const int mc = kind == 2 ? 2 * XF_MC : (kind == 3 ? XF_MC / 2 : XF_MC);
         int *raw_cnt = dalloc_n<int>((size_t) nwork + 1);
         f.rowlist = list;
         f.nwork = nwork;
         f.raw_j = dalloc_n<int>((size_t) nwork * mc);
         f.raw_a = dalloc_n<double>((size_t) nwork * mc);
         f.cnt = raw_cnt;
         f.defer_list = lists[pass & 1];
         f.ndefer = device_flags() + 60;
         f.claim = claim_counters();
         dzero(f.ndefer, sizeof(int));
         const int nbt = (int) std::max<long long>(1, std::min<long long>((nwork + TR_ROWS - 1) / TR_ROWS, 16384));
         if (kind == 3) {
            k_extpi_rowg<16, XF_MC / 2><<<launch_blocks((nwork + 3) / 4), WAVE, 0, stream()>>>(f);
            k_trunc_rows<XF_MC / 2><<<nbt, TR_ROWS, 0, stream()>>>(nwork, list, raw_cnt, f.raw_j,
                                                                   f.raw_a, trunc_factor, max_elmts,
                                                                   stride, cnt, tmp_j, tmp_a);
         } else if (kind == 0) {
            k_extpi_rowg<32, XF_MC><<<launch_blocks((nwork + 1) / 2), WAVE, 0, stream()>>>(f);
            k_trunc_rows<XF_MC><<<nbt, TR_ROWS, 0, stream()>>>(nwork, list, raw_cnt, f.raw_j, f.raw_a,
                                                               trunc_factor, max_elmts, stride, cnt,
                                                               tmp_j, tmp_a);
         } else if (kind == 1) {
            k_extpi_row64<XF_MC><<<launch_blocks(nwork), WAVE, 0, stream()>>>(f);
            k_trunc_rows<XF_MC><<<nbt, TR_ROWS, 0, stream()>>>(nwork, list, raw_cnt, f.raw_j, f.raw_a,
                                                               trunc_factor, max_elmts, stride, cnt,
                                                               tmp_j, tmp_a);
         } else {
            k_extpi_fast<2 * XF_MC, 512><<<launch_blocks(nwork), WAVE, 0, stream()>>>(f);
            k_trunc_rows<2 * XF_MC><<<nbt, TR_ROWS, 0, stream()>>>(nwork, list, raw_cnt, f.raw_j,
                                                                   f.raw_a, trunc_factor, max_elmts,
                                                                   stride, cnt, tmp_j, tmp_a);
         }
         HX_LAUNCH_CHECK();
         HX_HIP(hipMemcpyAsync(h, f.ndefer, sizeof(int), hipMemcpyDeviceToHost, stream()));
         sync();
         if (setup_verbose()) {
            std::fprintf(stderr, "[hypre_mi355x] ext+i %d rows, pass %d (C-hat <= %d): %d left\n", n,
                         pass, mc, h[0]);
         }
         dfree(f.raw_j);
         dfree(f.raw_a);
         dfree(raw_cnt);
         list = lists[pass & 1];
         nwork = h[0];
      }
      CL.free_all();
      if (nwork > 0) {
         a.base = nullptr;
         a.out_j = tmp_j;
         a.out_a = tmp_a;
         a.cnt = cnt;
         extpi_pass<MODE_NUM>(a, list, nwork);
      }
      dfree(lists[0]);
      dfree(lists[1]);
   } else {
      if (max_elmts > 0) {
         stride = max_elmts;
         tmp_j = dalloc_n<int>((size_t) n * stride);
         tmp_a = dalloc_n<double>((size_t) n * stride);
      } else {
         a.cnt = cnt;
         extpi_pass<MODE_SYM>(a);
         base = dalloc_n<int>((size_t) n + 1);
         long long tot = exclusive_scan(cnt, base, n);
         tmp_j = dalloc_n<int>((size_t) std::max<long long>(tot, 1));
         tmp_a = dalloc_n<double>((size_t) std::max<long long>(tot, 1));
      }
      a.base = base;
      a.out_j = tmp_j;
      a.out_a = tmp_a;
      a.cnt = cnt;
      extpi_pass<MODE_NUM>(a);
   }
   P.nrows = n;
   P.ncols = nc;
   P.ia = dalloc_n<int>((size_t) n + 1);
   P.nnz = exclusive_scan(cnt, P.ia, n);
   P.ja = dalloc_n<int>((size_t) std::max<long long>(P.nnz, 1));
   P.a = dalloc_n<double>((size_t) std::max<long long>(P.nnz, 1));
   P.a32 = nullptr;
   int nb = std::max(1, std::min((n + 255) / 256, 8192));
   k_compact_rows<<<nb, 256, 0, stream()>>>(n, base, cnt, P.ia, tmp_j, tmp_a, stride, P.ja, P.a);
   k_cf_fix<<<nb, 256, 0, stream()>>>(n, CF);
   HX_LAUNCH_CHECK();
   dfree(tmp_j);
   dfree(tmp_a);
   dfree(base);
   dfree(cnt);
}

// rows of P staged at a fixed stride for k_trunc_rows
__global__ void k_stage_rows(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                             const double *__restrict__ a, int mc, int *__restrict__ cnt,
                             int *__restrict__ rj, double *__restrict__ ra)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      const int s = ia[i], c = ia[i + 1] - s;
      cnt[i] = c;
      for (int k = 0; k < c; k++) {
         rj[(size_t) i * mc + k] = ja[s + k];
         ra[(size_t) i * mc + k] = a[s + k];
      }
   }
}

void truncate_rows(Csr &P, double trunc_factor, int max_elmts)
{
   const int n = P.nrows;
   if (n == 0) { return; }
   const int mx = max_row_len(P);
   if (mx > 2 * XF_MC) {
      fail(HYPRE_ERROR_GENERIC, "interpolation rows longer than 64 entries cannot be truncated");
   }
   const int mc = mx <= XF_MC / 2 ? XF_MC / 2 : (mx <= XF_MC ? XF_MC : 2 * XF_MC);
   const int stride = max_elmts > 0 ? std::min(max_elmts, mc) : mc;
   int *cnt = dalloc_n<int>((size_t) n);
   int *rj = dalloc_n<int>((size_t) n * mc);
   double *ra = dalloc_n<double>((size_t) n * mc);
   int *ocnt = dalloc_n<int>((size_t) n + 1);
   int *oj = dalloc_n<int>((size_t) n * stride);
   double *oa = dalloc_n<double>((size_t) n * stride);
   const int nb = std::max(1, std::min((n + 255) / 256, 8192));
   k_stage_rows<<<nb, 256, 0, stream()>>>(n, P.ia, P.ja, P.a, mc, cnt, rj, ra);
   const int nbt = (int) std::max<long long>(1, std::min<long long>((n + TR_ROWS - 1) / TR_ROWS, 16384));
   if (mc == XF_MC / 2) {
      k_trunc_rows<XF_MC / 2><<<nbt, TR_ROWS, 0, stream()>>>(n, nullptr, cnt, rj, ra, trunc_factor,
                                                             max_elmts, stride, ocnt, oj, oa);
   } else if (mc == XF_MC) {
      k_trunc_rows<XF_MC><<<nbt, TR_ROWS, 0, stream()>>>(n, nullptr, cnt, rj, ra, trunc_factor,
                                                         max_elmts, stride, ocnt, oj, oa);
   } else {
      k_trunc_rows<2 * XF_MC><<<nbt, TR_ROWS, 0, stream()>>>(n, nullptr, cnt, rj, ra, trunc_factor,
                                                             max_elmts, stride, ocnt, oj, oa);
   }
   HX_LAUNCH_CHECK();
   dfree(rj);
   dfree(ra);
   dfree(cnt);
   const int nc = P.ncols;
   P.free_all();
   P.nrows = n;
   P.ncols = nc;
   P.ia = dalloc_n<int>((size_t) n + 1);
   P.nnz = exclusive_scan(ocnt, P.ia, n);
   P.ja = dalloc_n<int>((size_t) std::max<long long>(P.nnz, 1));
   P.a = dalloc_n<double>((size_t) std::max<long long>(P.nnz, 1));
   k_compact_rows<<<nb, 256, 0, stream()>>>(n, nullptr, ocnt, P.ia, oj, oa, stride, P.ja, P.a);
   HX_LAUNCH_CHECK();
   dfree(oj);
   dfree(oa);
   dfree(ocnt);
}

// ================================================================== transpose
// src/seq_mv/csr_matop.c:1071 (stable counting sort): rows of R list the fine
// rows in ascending order.  Count the entries of each column and fill with
// atomics, then sort each (short) row.
// Both steps aggregate their atomics per chunk in an LDS table
// keyed by column: counts are added once per distinct column of a chunk of
// consecutive entries, and the fill reserves one range per distinct column of
// a chunk of 256 rows, each entry taking its place in the range from an LDS
// counter (entries that find no slot, or past the first TF_E of their row, use
// the global counter directly).  The order inside a row of R is fixed by
// k_t_sort afterwards.
constexpr int TF_SLOTS = 4096;
constexpr int TF_E = 8;
constexpr int TC_U = 8;

// the columns of a chunk are runs of consecutive coarse indices, so they take
// consecutive slots and the flushes add to consecutive counters (one memory
// transaction for many lanes)
__device__ __forceinline__ unsigned tf_hash(int c)
{
   return run_slot((unsigned) c, ilog2(TF_SLOTS));
}

__global__ __launch_bounds__(256) void k_col_count_h(long long nnz, const int *__restrict__ ja,
                                                     int *__restrict__ cnt)
{
   __shared__ int hkey[TF_SLOTS];
   __shared__ int hval[TF_SLOTS];
   constexpr long long CH = 256LL * TC_U;
   for (long long c0 = (long long) blockIdx.x * CH; c0 < nnz; c0 += (long long) gridDim.x * CH) {
      for (int q = threadIdx.x; q < TF_SLOTS; q += 256) {
         hkey[q] = -1;
         hval[q] = 0;
      }
      __syncthreads();
      int jv[TC_U];
#pragma unroll
      for (int u = 0; u < TC_U; u++) {
         const long long k = c0 + threadIdx.x + 256LL * u;
         jv[u] = (k < nnz) ? ja[k] : -1;
      }
#pragma unroll
      for (int u = 0; u < TC_U; u++) {
         const int c = jv[u];
         if (c < 0) { continue; }
         unsigned h = tf_hash(c);
         bool done = false;
         for (int probe = 0; probe < 32; probe++) {
            const int old = atomicCAS(&hkey[h], -1, c);
            if (old == -1 || old == c) {
               atomicAdd(&hval[h], 1);
               done = true;
               break;
            }
            h = (h + 1) & (TF_SLOTS - 1);
         }
         if (!done) { atomicAdd(&cnt[c], 1); }
      }
      __syncthreads();
      for (int q = threadIdx.x; q < TF_SLOTS; q += 256) {
         if (hkey[q] >= 0) { atomicAdd(&cnt[hkey[q]], hval[q]); }
      }
      __syncthreads();
   }
}

__global__ __launch_bounds__(256) void k_t_fill_h(int n, const int *__restrict__ ia,
                                                  const int *__restrict__ ja,
                                                  const double *__restrict__ a,
                                                  int *__restrict__ fillp, int *__restrict__ Rj,
                                                  double *__restrict__ Ra)
{
   __shared__ int hkey[TF_SLOTS];
   __shared__ int hval[TF_SLOTS];
   for (long long r0 = (long long) blockIdx.x * 256; r0 < n; r0 += (long long) gridDim.x * 256) {
      for (int q = threadIdx.x; q < TF_SLOTS; q += 256) {
         hkey[q] = -1;
         hval[q] = 0;
      }
      __syncthreads();
      const long long il = r0 + threadIdx.x;
      const int i = (int) il;
      int s = 0, e = 0;
      if (il < n) {
         s = ia[i];
         e = ia[i + 1];
      }
      int slot[TF_E], li[TF_E];
#pragma unroll
      for (int t = 0; t < TF_E; t++) {
         slot[t] = -1;
         li[t] = 0;
         if (s + t < e) {
            const int c = ja[s + t];
            unsigned h = tf_hash(c);
            for (int probe = 0; probe < 32; probe++) {
               const int old = atomicCAS(&hkey[h], -1, c);
               if (old == -1 || old == c) {
                  slot[t] = (int) h;
                  li[t] = atomicAdd(&hval[h], 1);
                  break;
               }
               h = (h + 1) & (TF_SLOTS - 1);
            }
         }
      }
      __syncthreads();
      // one range per distinct column of the chunk
      for (int q = threadIdx.x; q < TF_SLOTS; q += 256) {
         if (hkey[q] >= 0) { hval[q] = atomicAdd(&fillp[hkey[q]], hval[q]); }
      }
      __syncthreads();
#pragma unroll
      for (int t = 0; t < TF_E; t++) {
         if (s + t < e) {
            const int k = s + t;
            const int pos = (slot[t] >= 0) ? hval[slot[t]] + li[t] : atomicAdd(&fillp[ja[k]], 1);
            Rj[pos] = i;
            Ra[pos] = a[k];
         }
      }
      for (int k = s + TF_E; k < e; k++) {
         const int pos = atomicAdd(&fillp[ja[k]], 1);
         Rj[pos] = i;
         Ra[pos] = a[k];
      }
      __syncthreads();
   }
}

// Rows of R hold distinct fine indices: each entry's final place is its rank
// in the row.  One wave per row, the row staged in LDS; rows longer than the
// stage are sorted by insertion in lane 0.
constexpr int TS_MAX = 512;

__global__ __launch_bounds__(WAVE) void k_t_sort(int nc, const int *__restrict__ Ria,
                                                 int *__restrict__ Rj, double *__restrict__ Ra)
{
   __shared__ int sk[TS_MAX];
   __shared__ double sv[TS_MAX];
   const int lane = threadIdx.x;
   // rows of up to 64 entries are held one entry per lane; the next row's
   // bounds and entries are loaded while the current one is ranked
   int b = 0, e = 0, key = 0;
   double val = 0.0;
   auto fetch = [&](int r) {
      b = Ria[r];
      e = Ria[r + 1];
      if (e - b <= WAVE && lane < e - b) {
         key = Rj[b + lane];
         val = Ra[b + lane];
      }
   };
   if (blockIdx.x < nc) { fetch(blockIdx.x); }
   for (int r = blockIdx.x; r < nc; r += gridDim.x) {
      const int cb = b, ce = e, L = e - b;
      const int ckey = key;
      const double cval = val;
      if (r + (int) gridDim.x < nc) { fetch(r + gridDim.x); }
      if (L <= 1) { continue; }
      if (L <= WAVE) {
         if (lane < L) { sk[lane] = ckey; }
         wave_sync();
         if (lane < L) {
            int rank = 0;
            for (int m = 0; m < L; m++) { rank += (sk[m] < ckey) ? 1 : 0; }
            Rj[cb + rank] = ckey;
            Ra[cb + rank] = cval;
         }
         wave_sync();
         continue;
      }
      if (L > TS_MAX) {
         if (lane == 0) {
            for (int k = cb + 1; k < ce; k++) {
               const int key = Rj[k];
               const double v = Ra[k];
               int m = k - 1;
               while (m >= cb && Rj[m] > key) {
                  Rj[m + 1] = Rj[m];
                  Ra[m + 1] = Ra[m];
                  m--;
               }
               Rj[m + 1] = key;
               Ra[m + 1] = v;
            }
         }
         continue;
      }
      for (int k = lane; k < L; k += WAVE) {
         sk[k] = Rj[cb + k];
         sv[k] = Ra[cb + k];
      }
      wave_sync();
      for (int k = lane; k < L; k += WAVE) {
         const int key = sk[k];
         int rank = 0;
         for (int m = 0; m < L; m++) { rank += (sk[m] < key) ? 1 : 0; }
         Rj[cb + rank] = key;
         Ra[cb + rank] = sv[k];
      }
      wave_sync();
   }
}

void transpose(const Csr &P, Csr &R)
{
   int n = P.nrows, nc = P.ncols;
   TimedLaunch tl("setup_transpose", 0.0);
   R.nrows = nc;
   R.ncols = n;
   R.nnz = P.nnz;
   R.ia = dalloc_n<int>((size_t) nc + 1);
   R.ja = dalloc_n<int>((size_t) std::max<long long>(P.nnz, 1));
   R.a = dalloc_n<double>((size_t) std::max<long long>(P.nnz, 1));
   R.a32 = nullptr;
   dzero(R.ia, sizeof(int) * ((size_t) nc + 1));
   const long long nch = (P.nnz + 256LL * TC_U - 1) / (256LL * TC_U);
   const int nbe = (int) std::max<long long>(1, std::min<long long>(nch, 8192));
   if (P.nnz > 0) { k_col_count_h<<<nbe, 256, 0, stream()>>>(P.nnz, P.ja, R.ia); }
   HX_LAUNCH_CHECK();
   exclusive_scan_nowait(R.ia, R.ia, nc);   // the total is P.nnz
   int *fillp = dalloc_n<int>((size_t) nc + 1);
   d2d(fillp, R.ia, sizeof(int) * ((size_t) nc + 1));
   int nb = std::max(1, std::min((n + 255) / 256, 8192));
   k_t_fill_h<<<nb, 256, 0, stream()>>>(n, P.ia, P.ja, P.a, fillp, R.ja, R.a);
   k_t_sort<<<launch_blocks(nc), WAVE, 0, stream()>>>(nc, R.ia, R.ja, R.a);
   HX_LAUNCH_CHECK();
   dfree(fillp);
}

// ================================================================== RAP
// src/parcsr_ls/par_rap.c:1530-1930 (BuildCoarseOperatorKT, one rank): row ic
// of RA by a first-seen sweep over R(ic,:) x A, then (RA)(ic,:) x P with the
// diagonal entry created first.  Two tables per wave: RA (fine columns) and
// Ac (coarse columns), each of capacity `cap`.
struct RapArgs
{
   int           nc;
   const int    *R_ia, *R_ja;
   const double *R_a;
   const int    *A_ia, *A_ja;
   const double *A_a;
   const int    *P_ia, *P_ja;
   const double *P_a;
   const int    *rowlist;
   int           nwork;
   const int    *out_ia;     // MODE_NUM
   int          *out_j;
   double       *out_a;
   int          *cnt;        // MODE_SYM
   int          *defer_list;
   int          *ndefer;
   int           cap;
   char         *gslab;
   size_t        slab_stride;
};

template <int MODE>
__global__ __launch_bounds__(WAVE) void k_rap(RapArgs p)
{
   extern __shared__ double dyn_lds[];
   const int lane = threadIdx.x;
   const int nb = gridDim.x;
   const int blk = xcd_blk(blockIdx.x, nb);
   char *base = p.gslab ? p.gslab + (size_t) blk * p.slab_stride : reinterpret_cast<char *>(dyn_lds);
   int *ra_keys, *ra_vals, *ra_col, *ac_keys, *ac_vals, *ac_col;
   double *ra_v, *ac_v;
   carve(base, p.cap, ra_keys, ra_vals, ra_col, ra_v);
   carve(base + table_bytes(p.cap), p.cap, ac_keys, ac_vals, ac_col, ac_v);
   const int bits = ilog2(p.cap);
   RowHash HR{ra_keys, ra_vals, p.cap, bits};
   RowHash HC{ac_keys, ac_vals, p.cap, bits};
   const int capc = p.cap / 2;

   // rows strided over the grid: co-resident waves work on one window of
   // neighbouring rows, whose neighbour rows stay in the XCD's L2
   for (int wi = blk; wi < p.nwork; wi += nb) {
      const int ic = p.rowlist ? p.rowlist[wi] : wi;
      HR.clear(lane);
      HC.clear(lane);
      wave_sync();
      bool overflow = false;
      // ---- row ic of RA (par_rap.c:1712-1800)
      int nra = 0;
      for (int jj1 = p.R_ia[ic]; jj1 < p.R_ia[ic + 1] && !overflow; jj1++) {
         const int i1 = p.R_ja[jj1];
         const double r = (MODE == MODE_NUM) ? p.R_a[jj1] : 0.0;
         const int b = p.A_ia[i1], e = p.A_ia[i1 + 1];
         for (int kb = b; kb < e; kb += WAVE) {
            const int k = kb + lane;
            const bool valid = k < e;
            const int nv = min(WAVE, e - kb);
            if (nra + nv > capc) { overflow = true; break; }
            int i2 = -1, slot = -1;
            double prod = 0.0;
            bool isnew = false;
            if (valid) {
               i2 = p.A_ja[k];
               if (MODE == MODE_NUM) { prod = r * p.A_a[k]; }
               slot = HR.insert(i2, isnew);
            }
            const unsigned long long nm = __ballot(isnew);
            if (valid && slot >= 0) {
               if (isnew) {
                  const int pos = nra + __popcll(nm & lanemask_lt(lane));
                  ra_vals[slot] = pos;
                  ra_col[pos] = i2;
                  if (MODE == MODE_NUM) { ra_v[pos] = prod; }
               } else if (MODE == MODE_NUM) {
                  ra_v[ra_vals[slot]] += prod;   // key inserted by an earlier step
               }
            }
            nra += __popcll(nm);
            wave_sync();
         }
      }
      // ---- (RA) x P, diagonal first (par_rap.c:1612-1620, 1860-1920)
      int nac = 0;
      if (!overflow) {
         if (lane == 0) {
            bool isnew;
            int slot = HC.insert(ic, isnew);
            if (slot >= 0) { ac_vals[slot] = 0; }
            ac_col[0] = ic;
            if (MODE == MODE_NUM) { ac_v[0] = 0.0; }
         }
         nac = 1;
         wave_sync();
         for (int k = 0; k < nra && !overflow; k++) {
            const int i1 = ra_col[k];
            const double rav = (MODE == MODE_NUM) ? ra_v[k] : 0.0;
            const int b = p.P_ia[i1], e = p.P_ia[i1 + 1];
            for (int cb = b; cb < e; cb += WAVE) {
               const int len = min(WAVE, e - cb);
               if (nac + len > capc) { overflow = true; break; }
               const bool valid = lane < len;
               int i2 = -1, slot = -1;
               bool isnew = false;
               double prod = 0.0;
               if (valid) {
                  i2 = p.P_ja[cb + lane];
                  if (MODE == MODE_NUM) { prod = rav * p.P_a[cb + lane]; }
                  slot = HC.insert(i2, isnew);
               }
               const unsigned long long nm = __ballot(isnew);
               if (valid && slot >= 0) {
                  if (isnew) {
                     const int pos = nac + __popcll(nm & lanemask_lt(lane));
                     ac_vals[slot] = pos;
                     ac_col[pos] = i2;
                     if (MODE == MODE_NUM) { ac_v[pos] = prod; }
                  } else if (MODE == MODE_NUM) {
                     ac_v[ac_vals[slot]] += prod;
                  }
               }
               nac += __popcll(nm);
               wave_sync();
            }
         }
      }
      if (overflow) {
         if (lane == 0) {
            int d = atomicAdd(p.ndefer, 1);
            p.defer_list[d] = ic;
         }
         wave_sync();
         continue;
      }
      if (MODE == MODE_SYM) {
         if (lane == 0) { p.cnt[ic] = nac; }
      } else {
         const int o = p.out_ia[ic];
         for (int k = lane; k < nac; k += WAVE) {
            p.out_j[o + k] = ac_col[k];
            p.out_a[o + k] = ac_v[k];
         }
      }
      wave_sync();
   }
}

template <int MODE>
static void rap_pass(RapArgs a, const int *init_list = nullptr, int init_n = 0)
{
   static const int caps[2] = {512, 2048};
   deferred_passes(a, a.nc, device_flags() + 52, caps, 2, 1 << 22, 2,
                   [](RapArgs &x, int nb) {
                      size_t lds = x.gslab ? 0 : 2 * table_bytes(x.cap);
                      k_rap<MODE><<<nb, WAVE, lds, stream()>>>(x);
                      HX_LAUNCH_CHECK();
                   },
                   init_list, init_n);
}

// ------------------------------------------------------------------ RAP, fast path
// One numeric pass per coarse row with the arithmetic of k_rap:
//   * RA row: the A rows named by R(ic,:) are flattened and read 64 entries
//     per step; the step is applied one A row (segment) at a time, in R order,
//     so a new column is created by the first product and later products add
//     to it in the serial order; no two lanes of one segment share a column;
//   * Ac row: diagonal first, then the P rows of the RA entries in first-seen
//     order, flattened the same way.  Products of one step that hit the same
//     coarse column are applied in rounds: each round, the lowest pending lane
//     of every column (atomicMin on the slot) applies its product, so each
//     column receives its products in traversal order, and columns created in
//     a step get positions in the order of their first product.
// Output rows go to a buffer with a stride of `acl` entries; rows whose RA row
// exceeds `rcl` entries or whose Ac row exceeds `acl` entries are deferred to
// the next stage (larger tables) and finally to k_rap.
constexpr int RF_U = 4;      // chunks of 64 products per batch of loads
constexpr int RAP_RAREG = 5; // stage 0: RA row held in 5 registers per lane

struct RapFastArgs
{
   int           nc;
   const int    *R_ia, *R_ja;
   const double *R_a;
   const int    *A_ia, *A_ja;
   const double *A_a;
   const int    *P_ia, *P_ja;
   const double *P_a;
   const int    *rowlist;
   int           nwork;
   int           stage;
   int           rc;          // RA hash slots (power of two)
   int           rcl;         // RA entries (<= rc - 64)
   int           ac;          // Ac hash slots (power of two)
   int           acl;         // Ac entries (<= ac - 64), the output stride
   int          *tmp_j;       // row at (stage 0: ic, later: worklist index) * acl
   double       *tmp_a;
   int          *cnt;         // entries of a finished row
   int          *src;         // stage 0: offset >= 0; later stages: -2 - offset; -1 deferred
   int          *defer_list;
   int          *ndefer;
   int          *claim;       // RowClaim counters
};

// LDS of one wave.  RAREG > 0: the finished RA row (<= 64 * RAREG entries) is
// moved to registers before the Ac phase, whose tables then reuse the RA
// tables' space.
static size_t rap_fast_lds(int rc, int rcl, int ac, int acl, int rareg)
{
   const size_t ra = (size_t) rc * 12 + (size_t) rcl * 4;
   const size_t acb = (size_t) ac * 20 + (size_t) acl * 4;
   return rareg > 0 ? std::max(ra, acb) : ra + acb;
}

template <int RAREG>
__global__ __launch_bounds__(WAVE) void k_rap_fast(RapFastArgs p)
{
   extern __shared__ double dyn_lds[];
   const int lane = threadIdx.x;
   double *ra_val = dyn_lds;                    // rc
   int *ra_keys = reinterpret_cast<int *>(ra_val + p.rc);
   int *ra_order = ra_keys + p.rc;              // rcl
   double *ac_val = RAREG > 0 ? dyn_lds : reinterpret_cast<double *>(ra_order + p.rcl);   // ac
   int *ac_keys = reinterpret_cast<int *>(ac_val + p.ac);
   int *ac_pos = ac_keys + p.ac;
   int *ac_mins = ac_pos + p.ac;
   int *ac_order = ac_mins + p.ac;              // acl
   RowHash HR{ra_keys, nullptr, p.rc, ilog2(p.rc)};
   RowHash HC{ac_keys, nullptr, p.ac, ilog2(p.ac)};
   int rk[RAREG > 0 ? RAREG : 1];
   double rx[RAREG > 0 ? RAREG : 1];
   int rpb[RAREG > 0 ? RAREG : 1], rpl[RAREG > 0 ? RAREG : 1];   // P rows of the RA entries

   RowClaim claim(p.claim, p.nwork);
   claim.issue(lane);
   for (int wi = claim.take(lane); wi >= 0; wi = claim.take(lane)) {
      claim.issue(lane);   // the next row's claim is in flight while this row runs
      const int ic = p.rowlist ? p.rowlist[wi] : wi;
      for (int s = lane; s < p.rc; s += WAVE) { ra_keys[s] = -1; }
      if (RAREG == 0) {
         for (int s = lane; s < p.ac; s += WAVE) {
            ac_keys[s] = -1;
            ac_pos[s] = -1;
            ac_mins[s] = 0x7fffffff;
         }
      }
      wave_sync();
      bool defer = false;
      // ---- row ic of RA (par_rap.c:1712-1800)
      int nra = 0;
      const int rb0 = p.R_ia[ic], rb1 = p.R_ia[ic + 1];
      for (int rb = rb0; rb < rb1 && !defer; rb += WAVE) {
         const int jr = rb + lane;
         int len = 0, ab = 0;
         double rv = 0.0;
         if (jr < rb1) {
            const int i1 = p.R_ja[jr];
            rv = p.R_a[jr];
            ab = p.A_ia[i1];
            len = p.A_ia[i1 + 1] - ab;
         }
         const int incl = wave_incl_scan(len, lane);
         const int L = readlane_i(incl, WAVE - 1);
         const int off = incl - len;
         const int nseg = min(WAVE, rb1 - rb);
         for (int cb = 0; cb < L && !defer; cb += WAVE * RF_U) {
            int key[RF_U], sq[RF_U];
            double x[RF_U], av[RF_U], rsv[RF_U];
            // loads first (positions past L read entry L - 1), products after,
            // so that the batch's loads are in flight together
#pragma unroll
            for (int u = 0; u < RF_U; u++) {
               const int q = min(cb + u * WAVE + lane, L - 1);
               const int s = seg_find_lanes(off, nseg, q);
               const int b_s = __shfl(ab, s, WAVE);
               const int o_s = __shfl(off, s, WAVE);
               rsv[u] = shfl_d(rv, s);
               sq[u] = s;
               const int k = b_s + (q - o_s);
               key[u] = p.A_ja[k];
               av[u] = p.A_a[k];
            }
#pragma unroll
            for (int u = 0; u < RF_U; u++) {
               if (cb + u * WAVE + lane >= L) { key[u] = -1; }
               x[u] = rsv[u] * av[u];
            }
#pragma unroll
            for (int u = 0; u < RF_U; u++) {
               if (cb + u * WAVE >= L || defer) { break; }
               const bool valid = cb + u * WAVE + lane < L;
               const int s_lo = readlane_i(sq[u], 0);
               const int s_hi = readlane_i(sq[u], min(WAVE - 1, L - 1 - (cb + u * WAVE)));
               for (int sg = s_lo; sg <= s_hi; sg++) {
                  const bool act = valid && sq[u] == sg;
                  bool isn = false;
                  int sl = -1;
                  if (act) { sl = HR.insert(key[u], isn); }
                  const unsigned long long nm = __ballot(isn);
                  if (nra + __popcll(nm) > p.rcl) { defer = true; break; }
                  if (act && sl >= 0) {
                     if (isn) {
                        const int pos = nra + __popcll(nm & lanemask_lt(lane));
                        ra_order[pos] = sl;
                        ra_val[sl] = x[u];
                     } else {
                        ra_val[sl] += x[u];
                     }
                  }
                  nra += __popcll(nm);
                  wave_sync();
               }
            }
         }
      }
      // ---- (RA) x P, diagonal first (par_rap.c:1612-1620, 1860-1920)
      int nac = 0;
      if (RAREG > 0 && !defer) {
         // RA entry k to lane k % 64, register k / 64; then the Ac tables
         // take over the space
#pragma unroll
         for (int t = 0; t < RAREG; t++) {
            const int k = t * WAVE + lane;
            rk[t] = 0;
            rx[t] = 0.0;
            if (k < nra) {
               const int sl = ra_order[k];
               rk[t] = ra_keys[sl];
               rx[t] = ra_val[sl];
            }
         }
         // the P row offsets of all RA entries in one round of loads
#pragma unroll
         for (int t = 0; t < RAREG; t++) {
            const int k = t * WAVE + lane;
            rpb[t] = 0;
            rpl[t] = 0;
            if (k < nra) {
               rpb[t] = p.P_ia[rk[t]];
               rpl[t] = p.P_ia[rk[t] + 1] - rpb[t];
            }
         }
         wave_sync();
         for (int s = lane; s < p.ac; s += WAVE) {
            ac_keys[s] = -1;
            ac_pos[s] = -1;
            ac_mins[s] = 0x7fffffff;
         }
         wave_sync();
      }
      if (!defer) {
         if (lane == 0) {
            bool isn;
            const int sl = HC.insert(ic, isn);
            ac_pos[sl] = 0;
            ac_order[0] = sl;
            ac_val[sl] = 0.0;
         }
         nac = 1;
         wave_sync();
         for (int kb = 0; kb < nra && !defer; kb += WAVE) {
            const int k = kb + lane;
            int len = 0, pb = 0;
            double rav = 0.0;
            if (k < nra) {
               if (RAREG > 0) {
                  // register kb / 64, selected without dynamic indexing
                  rav = rx[0];
                  pb = rpb[0];
                  len = rpl[0];
#pragma unroll
                  for (int t = 1; t < RAREG; t++) {
                     if (kb == t * WAVE) { rav = rx[t]; pb = rpb[t]; len = rpl[t]; }
                  }
               } else {
                  const int sl = ra_order[k];
                  const int i1 = ra_keys[sl];
                  rav = ra_val[sl];
                  pb = p.P_ia[i1];
                  len = p.P_ia[i1 + 1] - pb;
               }
            }
            const int incl = wave_incl_scan(len, lane);
            const int L = readlane_i(incl, WAVE - 1);
            const int off = incl - len;
            const int nseg = min(WAVE, nra - kb);
            for (int cb = 0; cb < L && !defer; cb += WAVE * RF_U) {
               int key[RF_U];
               double x[RF_U], pv[RF_U], rsv[RF_U];
#pragma unroll
               for (int u = 0; u < RF_U; u++) {
                  const int q = min(cb + u * WAVE + lane, L - 1);
                  const int s = seg_find_lanes(off, nseg, q);
                  const int b_s = __shfl(pb, s, WAVE);
                  const int o_s = __shfl(off, s, WAVE);
                  rsv[u] = shfl_d(rav, s);
                  const int e = b_s + (q - o_s);
                  key[u] = p.P_ja[e];
                  pv[u] = p.P_a[e];
               }
#pragma unroll
               for (int u = 0; u < RF_U; u++) {
                  if (cb + u * WAVE + lane >= L) { key[u] = -1; }
                  x[u] = rsv[u] * pv[u];
               }
#pragma unroll
               for (int u = 0; u < RF_U; u++) {
                  if (cb + u * WAVE >= L || defer) { break; }
                  const bool valid = cb + u * WAVE + lane < L;
                  bool isn = false;
                  int sl = -1;
                  if (valid) { sl = HC.insert(key[u], isn); }
                  if (nac + __popcll(__ballot(isn)) > p.acl) { defer = true; break; }
                  bool pending = valid && sl >= 0;
                  while (__ballot(pending)) {
                     if (pending) { atomicMin(&ac_mins[sl], lane); }
                     wave_sync();
                     const bool apply = pending && ac_mins[sl] == lane;
                     const bool first = apply && ac_pos[sl] < 0;
                     const unsigned long long fm = __ballot(first);
                     if (apply) {
                        if (first) {
                           const int pos = nac + __popcll(fm & lanemask_lt(lane));
                           ac_pos[sl] = pos;
                           ac_order[pos] = sl;
                           ac_val[sl] = x[u];
                        } else {
                           ac_val[sl] += x[u];
                        }
                        ac_mins[sl] = 0x7fffffff;
                        pending = false;
                     }
                     nac += __popcll(fm);
                     wave_sync();
                  }
               }
            }
         }
      }
      if (!defer) {
         const size_t o = (size_t) (p.rowlist ? wi : ic) * p.acl;
         for (int k = lane; k < nac; k += WAVE) {
            const int sl = ac_order[k];
            p.tmp_j[o + k] = ac_keys[sl];
            p.tmp_a[o + k] = ac_val[sl];
         }
         if (lane == 0) {
            p.cnt[ic] = nac;
            // stage 0: the offset; stage 1 or 2: -2 - (2 * offset + stage - 1)
            p.src[ic] = (p.stage == 0) ? (int) o : -2 - (2 * (int) o + p.stage - 1);
         }
      } else if (lane == 0) {
         p.src[ic] = -1;
         const int d = atomicAdd(p.ndefer, 1);
         p.defer_list[d] = ic;
      }
      wave_sync();
   }
}

// ------------------------------------------------------------------ RAP, stage 0 by fixed lane layouts
// k_rap_fast<RAREG>'s arithmetic and outputs without flattened segments (no
// per-chunk segment search, no prefix sums, no rounds):
//   * RA row, MULTI: 64 / WA A rows per step, WA lanes each (WA, a power of
//     two <= 32, at least A's longest row).  All lanes insert at once; a key
//     created during the step is stored with bit 31 set until the first of
//     the step's rows that holds it (in R order) numbers it and clears the
//     bit, and the rows then apply their products one after the other, so
//     every RA entry receives its products in R order.  !MULTI: one A row
//     per step, 64 entries at a time (a row's columns are distinct);
//   * Ac row: 64 / W RA entries per step, W lanes each (W, a power of two, at
//     least P's longest row), lane order = traversal order.  Each lane ORs its
//     bit into its column's mask; the lowest lane of a column adds the
//     column's products of the step in lane order (reading them from LDS) and
//     numbers the column if it is new, so columns are created in the order of
//     their first product and receive their products in traversal order;
//   * the loads of RS_D steps are issued while the previous RS_D steps are
//     applied.
constexpr int RS_D = 4;
constexpr int RS_NEW = (int) 0x80000000;   // RA key created in the current step

__device__ __forceinline__ int rs_insert(int *keys, int cap, int bits, int key, bool &isnew,
                                         bool &fresh)
{
   unsigned h = ((unsigned) key * 2654435761u) >> (32 - bits);
   int slot = -1, old = 0;
   for (int probe = 0; probe < cap; probe++) {
      old = atomicCAS(&keys[h], -1, key | RS_NEW);
      if (old == -1 || (old & 0x7fffffff) == key) {
         slot = (int) h;
         break;
      }
      h = (h + 1) & (cap - 1);
   }
   isnew = slot >= 0 && old == -1;
   fresh = slot >= 0 && (old == -1 || (old & RS_NEW) != 0);
   return slot;
}

static size_t rap_seg_lds(int rc, int rcl, int ac, int acl)
{
   const size_t ra = (size_t) rc * 12 + (size_t) rcl * 4;
   const size_t acb = (size_t) ac * 24 + (size_t) acl * 4 + WAVE * 8;
   return std::max(ra, acb);
}

template <int RAREG, bool MULTI>
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu((RAREG <= RAP_RAREG && !MULTI) ? 5 : 1))) void k_rap_seg(RapFastArgs p, int wsh, int wsa)
{
   extern __shared__ double dyn_lds[];
   const int lane = threadIdx.x;
   double *ra_val = dyn_lds;                    // rc
   int *ra_keys = reinterpret_cast<int *>(ra_val + p.rc);
   int *ra_order = ra_keys + p.rc;              // rcl
   // the Ac tables reuse the RA tables' space once RA is in registers
   double *ac_val = dyn_lds;                    // ac
   unsigned long long *ac_mask = reinterpret_cast<unsigned long long *>(ac_val + p.ac);
   int *ac_keys = reinterpret_cast<int *>(ac_mask + p.ac);
   int *ac_pos = ac_keys + p.ac;
   int *ac_order = ac_pos + p.ac;               // acl
   double *xs = reinterpret_cast<double *>(ac_order + p.acl);   // 64 products of a step
   RowHash HC{ac_keys, nullptr, p.ac, ilog2(p.ac)};
   const int rbits = ilog2(p.rc);
   const int wmask = (1 << wsh) - 1;
   const int epc = WAVE >> wsh;                 // RA entries per Ac step
   const int sub = lane >> wsh, pj = lane & (wmask);
   const int amask_a = (1 << wsa) - 1;
   const int epa = WAVE >> wsa;                 // A rows per RA step (MULTI)
   const int sub_a = lane >> wsa, aj = lane & amask_a;

   RowClaim claim(p.claim, p.nwork);
   claim.issue(lane);
   for (int wi = claim.take(lane); wi >= 0; wi = claim.take(lane)) {
      claim.issue(lane);   // the next row's claim is in flight while this row runs
      const int ic = p.rowlist ? p.rowlist[wi] : wi;
      for (int s = lane; s < p.rc; s += WAVE) { ra_keys[s] = -1; }
      wave_sync();
      bool defer = false;
      // ---- row ic of RA (par_rap.c:1712-1800)
      int nra = 0;
      const int rb0 = p.R_ia[ic], rb1 = p.R_ia[ic + 1];
      for (int rb = rb0; rb < rb1 && !defer; rb += WAVE) {
         // lane s: R entry rb + s and its A row [ab, ae)
         int ab = 0, ae = 0;
         double rv = 0.0;
         if (rb + lane < rb1) {
            const int i1 = p.R_ja[rb + lane];
            rv = p.R_a[rb + lane];
            ab = p.A_ia[i1];
            ae = p.A_ia[i1 + 1];
         }
         const int nseg = min(WAVE, rb1 - rb);
         if (MULTI && __ballot(rb + lane < rb1 && ae - ab > amask_a + 1)) {
            defer = true;   // an A row longer than WA
            break;
         }
         // step cursor (uniform): segment sg (MULTI: first of epa) from entry kb
         int sg = 0, kb = readlane_i(ab, 0), ke = readlane_i(ae, 0);
         // two buffers of RS_D steps, named so that every register index is static
         int k0[RS_D], k1[RS_D];
         double a0[RS_D], a1[RS_D], r0[RS_D], r1[RS_D];
         bool s0[RS_D], s1[RS_D];
         auto fill = [&](int (&key)[RS_D], double (&av)[RS_D], double (&rr)[RS_D], bool (&st)[RS_D]) {
#pragma unroll
            for (int u = 0; u < RS_D; u++) {
               st[u] = sg < nseg;
               key[u] = -1;
               av[u] = 0.0;
               rr[u] = 0.0;
               if (sg < nseg) {
                  if (MULTI) {
                     const int e = sg + sub_a;
                     const int src = min(e, WAVE - 1);
                     const int b_s = __shfl(ab, src, WAVE);
                     const int e_s = __shfl(ae, src, WAVE);
                     rr[u] = shfl_d(rv, src);
                     if (e < nseg && b_s + aj < e_s) {
                        key[u] = p.A_ja[b_s + aj];
                        av[u] = p.A_a[b_s + aj];
                     }
                     sg += epa;
                  } else {
                     const int k = kb + lane;
                     rr[u] = readlane_d(rv, sg);
                     if (k < ke) {
                        key[u] = p.A_ja[k];
                        av[u] = p.A_a[k];
                     }
                     kb += WAVE;
                     if (kb >= ke) {
                        sg++;
                        if (sg < nseg) {
                           kb = readlane_i(ab, sg);
                           ke = readlane_i(ae, sg);
                        }
                     }
                  }
               }
            }
         };
         auto apply = [&](const int (&key)[RS_D], const double (&av)[RS_D], const double (&rr)[RS_D],
                          const bool (&st)[RS_D]) {
#pragma unroll
            for (int u = 0; u < RS_D; u++) {
               if (!st[u] || defer) { break; }
               const bool valid = key[u] >= 0;
               bool isn = false, fresh = false;
               int sl = -1;
               if (valid) { sl = rs_insert(ra_keys, p.rc, rbits, key[u], isn, fresh); }
               const bool on = valid && sl >= 0;
               const double x = rr[u] * av[u];
               // the step's rows in R order (one row when !MULTI)
               for (int h = 0; h < (MULTI ? epa : 1); h++) {
                  const bool mine = on && (!MULTI || sub_a == h);
                  // a key created in this step is numbered by the first row
                  // that holds it: the one that still finds bit 31 set
                  bool create = false;
                  if (mine && fresh) { create = (ra_keys[sl] & RS_NEW) != 0; }
                  const unsigned long long cm = __ballot(create);
                  if (nra + __popcll(cm) > p.rcl) {
                     defer = true;
                     break;
                  }
                  if (mine) {
                     if (create) {
                        const int pos = nra + __popcll(cm & lanemask_lt(lane));
                        ra_order[pos] = sl;
                        ra_val[sl] = x;
                        ra_keys[sl] = key[u];
                     } else {
                        ra_val[sl] += x;   // created by an earlier row
                     }
                  }
                  nra += __popcll(cm);
                  wave_sync();
               }
            }
         };
         fill(k0, a0, r0, s0);
         for (;;) {
            if (!s0[0] || defer) { break; }
            fill(k1, a1, r1, s1);
            apply(k0, a0, r0, s0);
            if (!s1[0] || defer) { break; }
            fill(k0, a0, r0, s0);
            apply(k1, a1, r1, s1);
         }
      }
      // ---- (RA) x P, diagonal first (par_rap.c:1612-1620, 1860-1920)
      int nac = 0;
      int rk[RAREG];
      double rx[RAREG];
      int rpb[RAREG], rpl[RAREG];
      if (!defer) {
         // RA entry k to lane k % 64, register k / 64, and its P row
#pragma unroll
         for (int t = 0; t < RAREG; t++) {
            const int k = t * WAVE + lane;
            rk[t] = 0;
            rx[t] = 0.0;
            if (k < nra) {
               const int sl = ra_order[k];
               rk[t] = ra_keys[sl];
               rx[t] = ra_val[sl];
            }
         }
#pragma unroll
         for (int t = 0; t < RAREG; t++) {
            const int k = t * WAVE + lane;
            rpb[t] = 0;
            rpl[t] = 0;
            if (k < nra) {
               rpb[t] = p.P_ia[rk[t]];
               rpl[t] = p.P_ia[rk[t] + 1] - rpb[t];
            }
         }
         wave_sync();
         for (int s = lane; s < p.ac; s += WAVE) {
            ac_keys[s] = -1;
            ac_pos[s] = -1;
            ac_mask[s] = 0ull;
         }
         wave_sync();
         if (lane == 0) {
            bool isn;
            const int sl = HC.insert(ic, isn);
            ac_pos[sl] = 0;
            ac_order[0] = sl;
            ac_val[sl] = 0.0;
         }
         nac = 1;
         wave_sync();
      }
#pragma unroll
      for (int t = 0; t < RAREG; t++) {
         if (defer || t * WAVE >= nra) { break; }
         const int nb = min(WAVE, nra - t * WAVE);   // RA entries held in register t
         // steps of epc entries: entry e = eb + sub, P entry pj of its row
         int eb = 0;
         int k0[RS_D], k1[RS_D];
         double v0[RS_D], v1[RS_D], r0[RS_D], r1[RS_D];
         bool s0[RS_D], s1[RS_D];
         auto fill = [&](int (&key)[RS_D], double (&pv)[RS_D], double (&rs)[RS_D], bool (&st)[RS_D]) {
#pragma unroll
            for (int u = 0; u < RS_D; u++) {
               st[u] = eb < nb;
               key[u] = -1;
               pv[u] = 0.0;
               rs[u] = 0.0;
               if (eb < nb) {
                  const int e = eb + sub;
                  const int src = min(e, WAVE - 1);
                  const int pb = __shfl(rpb[t], src, WAVE);
                  const int pl = __shfl(rpl[t], src, WAVE);
                  rs[u] = shfl_d(rx[t], src);
                  // a P row longer than W would lose entries: the row is deferred
                  if (__ballot(e < nb && pl > wmask + 1)) { defer = true; }
                  if (e < nb && pj < pl) {
                     key[u] = p.P_ja[pb + pj];
                     pv[u] = p.P_a[pb + pj];
                  }
                  eb += epc;
               }
            }
         };
         auto apply = [&](const int (&key)[RS_D], const double (&pv)[RS_D], const double (&rs)[RS_D],
                          const bool (&st)[RS_D]) {
#pragma unroll
            for (int u = 0; u < RS_D; u++) {
               if (!st[u] || defer) { break; }
               const bool valid = key[u] >= 0;
               bool isn = false;
               int sl = -1;
               if (valid) { sl = HC.insert(key[u], isn); }
               if (nac + __popcll(__ballot(isn)) > p.acl) {
                  defer = true;
                  break;
               }
               const double x = rs[u] * pv[u];
               const bool on = valid && sl >= 0;
               if (on) { atomicOr(&ac_mask[sl], 1ull << lane); }
               xs[lane] = x;
               wave_sync();
               unsigned long long m = 0ull;
               bool lead = false;
               if (on) {
                  m = ac_mask[sl];
                  lead = (m & lanemask_lt(lane)) == 0ull;
               }
               const bool fresh = lead && ac_pos[sl] < 0;
               const unsigned long long fm = __ballot(fresh);
               if (lead) {
                  // the column's products of this step, in lane order
                  double acc = x;
                  if (!fresh) {
                     acc = ac_val[sl];
                     acc += x;
                  }
                  for (unsigned long long rest = m & (m - 1); rest; rest &= rest - 1) {
                     acc += xs[__ffsll((long long) rest) - 1];
                  }
                  ac_val[sl] = acc;
                  ac_mask[sl] = 0ull;
                  if (fresh) {
                     const int pos = nac + __popcll(fm & lanemask_lt(lane));
                     ac_pos[sl] = pos;
                     ac_order[pos] = sl;
                  }
               }
               nac += __popcll(fm);
               wave_sync();
            }
         };
         fill(k0, v0, r0, s0);
         for (;;) {
            if (!s0[0] || defer) { break; }
            fill(k1, v1, r1, s1);
            apply(k0, v0, r0, s0);
            if (!s1[0] || defer) { break; }
            fill(k0, v0, r0, s0);
            apply(k1, v1, r1, s1);
         }
      }
      if (!defer) {
         const size_t o = (size_t) (p.rowlist ? wi : ic) * p.acl;
         for (int k = lane; k < nac; k += WAVE) {
            const int sl = ac_order[k];
            p.tmp_j[o + k] = ac_keys[sl];
            p.tmp_a[o + k] = ac_val[sl];
         }
         if (lane == 0) {
            p.cnt[ic] = nac;
            // stage 0: the offset; stage 1 or 2: -2 - (2 * offset + stage - 1)
            p.src[ic] = (p.stage == 0) ? (int) o : -2 - (2 * (int) o + p.stage - 1);
         }
      } else if (lane == 0) {
         p.src[ic] = -1;
         const int d = atomicAdd(p.ndefer, 1);
         p.defer_list[d] = ic;
      }
      wave_sync();
   }
}

// copy the rows finished by the fast stages into Ac: K lanes per row, the
// loads of RC_U entries per lane issued before their stores
constexpr int RC_U = 4;

template <int K>
__global__ __launch_bounds__(256) void k_rap_compact(int nc, const int *__restrict__ cnt,
                                                     const int *__restrict__ src,
                                                     const int *__restrict__ out_ia,
                                                     const int *__restrict__ tj0,
                                                     const double *__restrict__ ta0,
                                                     const int *__restrict__ tj1,
                                                     const double *__restrict__ ta1,
                                                     const int *__restrict__ tj2,
                                                     const double *__restrict__ ta2,
                                                     int *__restrict__ oj, double *__restrict__ oa)
{
   const int g = threadIdx.x & (K - 1);
   const int rows_per_block = 256 / K;
   for (long long rb = (long long) blockIdx.x * rows_per_block; rb < nc;
        rb += (long long) gridDim.x * rows_per_block) {
      const long long r = rb + threadIdx.x / K;
      if (r >= nc) { continue; }
      const int s = src[r];
      if (s == -1) { continue; }
      const int c = cnt[r];
      const int o = out_ia[r];
      const int later = s >= 0 ? 0 : 1 + ((-2 - s) & 1);
      const int *tj = later == 0 ? tj0 : (later == 1 ? tj1 : tj2);
      const double *ta = later == 0 ? ta0 : (later == 1 ? ta1 : ta2);
      const size_t b = s >= 0 ? (size_t) s : (size_t) ((-2 - s) >> 1);
      for (int k0 = g; k0 < c; k0 += K * RC_U) {
         int jv[RC_U];
         double av[RC_U];
#pragma unroll
         for (int u = 0; u < RC_U; u++) {
            const int k = k0 + u * K;
            if (k < c) {
               jv[u] = tj[b + k];
               av[u] = ta[b + k];
            }
         }
#pragma unroll
         for (int u = 0; u < RC_U; u++) {
            const int k = k0 + u * K;
            if (k < c) {
               oj[o + k] = jv[u];
               oa[o + k] = av[u];
            }
         }
      }
   }
}

void rap(const Csr &R, const Csr &A, const Csr &P, Csr &Ac, int p_row_bound, int a_row_bound,
         int *ac_max_row)
{
   int nc = R.nrows;
   TimedLaunch tl("setup_rap", 0.0);
   RapArgs a{};
   a.nc = nc;
   a.R_ia = R.ia; a.R_ja = R.ja; a.R_a = R.a;
   a.A_ia = A.ia; a.A_ja = A.ja; a.A_a = A.a;
   a.P_ia = P.ia; a.P_ja = P.ja; a.P_a = P.a;
   Ac.nrows = nc;
   Ac.ncols = nc;
   Ac.a32 = nullptr;
   Ac.ia = dalloc_n<int>((size_t) nc + 1);

   // fast path in three stages of growing tables; stage 0 writes its rows
   // indexed by coarse row, later stages by position in their row list
   RapFastArgs f{};
   f.nc = nc;
   f.R_ia = R.ia; f.R_ja = R.ja; f.R_a = R.a;
   f.A_ia = A.ia; f.A_ja = A.ja; f.A_a = A.a;
   f.P_ia = P.ia; f.P_ja = P.ja; f.P_a = P.a;
   // stage 0: RA <= 320 entries (in 5 registers per lane), Ac <= 128;
   // stage 1: RA <= 640 (10 registers), Ac <= 256; stage 2: RA <= 1280, Ac <= 640
   static const int rcs[3] = {512, 1024, 2048}, acs[3] = {256, 512, 1024},
                    acls[3] = {128, 256, 640};
   static const int raregs[3] = {RAP_RAREG, 2 * RAP_RAREG, 0};
   int *cnt = dalloc_n<int>((size_t) nc + 1);
   int *src = dalloc_n<int>((size_t) nc + 1);
   dzero(cnt, sizeof(int) * ((size_t) nc + 1));
   int *tj[3] = {nullptr, nullptr, nullptr};
   double *ta[3] = {nullptr, nullptr, nullptr};
   int *lists[2] = {dalloc_n<int>((size_t) nc + 1), dalloc_n<int>((size_t) nc + 1)};
   int *ndefer = device_flags() + 60;
   int *h = static_cast<int *>(pinned_scratch(64));
   f.cnt = cnt;
   f.src = src;
   f.ndefer = ndefer;
   const int *list = nullptr;
   int nwork = nc;
   int cur = 0;
   // stages 0 and 1 run k_rap_seg when P's rows fit W <= 16 lanes each (log2 W);
   // A rows of at most WA <= 32 entries are taken 64 / WA at a time (log2 WA)
   int pw = -1, aw = 6;
   {
      const int pmax = p_row_bound > 0 ? p_row_bound : csr_max_row(P);
      if (pmax <= 16) {
         pw = 0;
         while ((1 << pw) < pmax) { pw++; }
         const int amax = a_row_bound > 0 ? a_row_bound : csr_max_row(A);
         if (amax <= 32) {
            aw = 2;
            while ((1 << aw) < amax) { aw++; }
         }
      }
   }
   // where A's mean row has more than 64 entries, most RA rows overflow stage
   // 0's tables (two thirds at the third level of the 27-point problem) and
   // the stage would only defer them
   const int first_stage = ((double) A.nnz / std::max(A.nrows, 1) > 64.0) ? 1 : 0;
   for (int stage = first_stage; stage < 3 && nwork > 0; stage++) {
      // offsets are kept as int (times 2 after stage 0)
      if ((long long) nwork * acls[stage] * 2 >= INT_MAX) { break; }
      tj[stage] = dalloc_n<int>((size_t) nwork * acls[stage]);
      ta[stage] = dalloc_n<double>((size_t) nwork * acls[stage]);
      f.rowlist = list;
      f.nwork = nwork;
      f.stage = stage;
      f.rc = rcs[stage];
      f.rcl = rcs[stage] * 5 / 8;
      f.ac = acs[stage];
      f.acl = acls[stage];
      f.tmp_j = tj[stage];
      f.tmp_a = ta[stage];
      f.defer_list = lists[cur];
      dzero(ndefer, sizeof(int));
      f.claim = claim_counters();
      if (raregs[stage] > 0) { f.rcl = raregs[stage] * WAVE; }
      const size_t lds = rap_fast_lds(f.rc, f.rcl, f.ac, f.acl, raregs[stage]);
      if (stage <= 1 && pw >= 0) {
         const size_t lds_s = rap_seg_lds(f.rc, f.rcl, f.ac, f.acl);
         const int nb = launch_blocks(nwork);
         if (stage == 0 && aw <= 5) {
            k_rap_seg<RAP_RAREG, true><<<nb, WAVE, lds_s, stream()>>>(f, pw, aw);
         } else if (stage == 0) {
            k_rap_seg<RAP_RAREG, false><<<nb, WAVE, lds_s, stream()>>>(f, pw, 6);
         } else if (aw <= 5) {
            k_rap_seg<2 * RAP_RAREG, true><<<nb, WAVE, lds_s, stream()>>>(f, pw, aw);
         } else {
            k_rap_seg<2 * RAP_RAREG, false><<<nb, WAVE, lds_s, stream()>>>(f, pw, 6);
         }
      } else if (stage == 0) {
         k_rap_fast<RAP_RAREG><<<launch_blocks(nwork), WAVE, lds, stream()>>>(f);
      } else if (stage == 1) {
         k_rap_fast<2 * RAP_RAREG><<<launch_blocks(nwork), WAVE, lds, stream()>>>(f);
      } else {
         k_rap_fast<0><<<launch_blocks(nwork), WAVE, lds, stream()>>>(f);
      }
      HX_LAUNCH_CHECK();
      HX_HIP(hipMemcpyAsync(h, ndefer, sizeof(int), hipMemcpyDeviceToHost, stream()));
      sync();
      list = lists[cur];
      if (setup_verbose()) {
         std::fprintf(stderr, "[hypre_mi355x] RAP %d coarse rows: stage %d (RA %d, Ac %d) left %d\n",
                      nc, stage, f.rc, f.ac, h[0]);
      }
      nwork = h[0];
      cur ^= 1;
   }
   // rows beyond the fast path: counts by the general kernel
   if (nwork > 0) {
      a.cnt = cnt;
      rap_pass<MODE_SYM>(a, list, nwork);
   }
   Ac.nnz = exclusive_scan(cnt, Ac.ia, nc, ac_max_row);
   Ac.ja = dalloc_n<int>((size_t) std::max<long long>(Ac.nnz, 1));
   Ac.a = dalloc_n<double>((size_t) std::max<long long>(Ac.nnz, 1));
   {
      const int K = lanes_per_row(Ac.nnz, nc);
      const int nbc = (int) std::max<long long>(1, std::min<long long>(((long long) nc * K + 255) / 256, 16384));
      dispatch_k(K, [&](auto kc) {
         k_rap_compact<decltype(kc)::value><<<nbc, 256, 0, stream()>>>(
             nc, cnt, src, Ac.ia, tj[0], ta[0], tj[1], ta[1], tj[2], ta[2], Ac.ja, Ac.a);
      });
      HX_LAUNCH_CHECK();
   }
   if (nwork > 0) {
      a.out_ia = Ac.ia;
      a.out_j = Ac.ja;
      a.out_a = Ac.a;
      rap_pass<MODE_NUM>(a, list, nwork);
   }
   dfree(lists[0]);
   dfree(lists[1]);
   for (int s = 0; s < 3; s++) {
      dfree(tj[s]);
      dfree(ta[s]);
   }
   dfree(src);
   dfree(cnt);
}

// ================================================================== dense coarse matrix
__global__ void k_dense_colmajor(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                 const double *__restrict__ a, double *__restrict__ M)
{
   for (int i = threadIdx.x; i < n * n; i += blockDim.x) { M[i] = 0.0; }
   __syncthreads();
   for (int i = threadIdx.x; i < n; i += blockDim.x) {
      for (int k = ia[i]; k < ia[i + 1]; k++) { M[i + n * ja[k]] = a[k]; }
   }
}

void csr_to_dense_colmajor(const Csr &A, double *M_dev)
{
   k_dense_colmajor<<<1, 256, 0, stream()>>>(A.nrows, A.ia, A.ja, A.a, M_dev);
   HX_LAUNCH_CHECK();
}

} // namespace hx
