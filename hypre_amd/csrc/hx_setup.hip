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

__device__ __forceinline__ void wave_sync() { __syncthreads(); }   // blocks are one wave

static int launch_blocks(long long work_items)
{
   long long b = std::min<long long>(std::max<long long>(work_items, 1), 256 * 32);
   if (b >= 64) { b = b / 8 * 8; }
   return (int) b;
}

__device__ __forceinline__ int xcd_blk(int bid, int nb)
{
   return (nb & 7) ? bid : (bid & 7) * (nb >> 3) + (bid >> 3);
}

// ================================================================== strength
// src/parcsr_ls/par_strength.c:75-525 (CreateSHost, num_functions 1, one rank)
__device__ __forceinline__ void strength_row(int i, const int *__restrict__ ia,
                                             const double *__restrict__ a, double theta,
                                             double max_row_sum, double &thr, int &mode)
{
   // mode: 0 all weak, 1 strong iff a > thr (diag < 0), 2 strong iff a < thr
   int s = ia[i], e = ia[i + 1];
   double diag = a[s];
   double row_scale = 0.0, row_sum = diag;
   if (diag < 0) {
      for (int j = s + 1; j < e; j++) {
         double v = a[j];
         row_scale = (row_scale < v) ? v : row_scale;
         row_sum += v;
      }
   } else {
      for (int j = s + 1; j < e; j++) {
         double v = a[j];
         row_scale = (row_scale < v) ? row_scale : v;
         row_sum += v;
      }
   }
   if ((fabs(row_sum) > fabs(diag) * max_row_sum) && (max_row_sum < 1.0)) {
      mode = 0;
   } else {
      mode = (diag < 0) ? 1 : 2;
   }
   thr = theta * row_scale;
}

__global__ void k_strength_count(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                 const double *__restrict__ a, double theta, double mrs,
                                 int *__restrict__ cnt)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      double thr;
      int mode;
      strength_row(i, ia, a, theta, mrs, thr, mode);
      int c = 0;
      if (mode) {
         for (int j = ia[i] + 1; j < ia[i + 1]; j++) {
            double v = a[j];
            bool strong = (mode == 1) ? !(v <= thr) : !(v >= thr);
            c += strong;
         }
      }
      cnt[i] = c;
   }
}

__global__ void k_strength_fill(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                const double *__restrict__ a, double theta, double mrs,
                                const int *__restrict__ S_ia, int *__restrict__ S_ja)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      double thr;
      int mode;
      strength_row(i, ia, a, theta, mrs, thr, mode);
      if (!mode) { continue; }
      int o = S_ia[i];
      for (int j = ia[i] + 1; j < ia[i + 1]; j++) {
         double v = a[j];
         bool strong = (mode == 1) ? !(v <= thr) : !(v >= thr);
         if (strong) { S_ja[o++] = ja[j]; }
      }
   }
}

void create_s(const Csr &A, double theta, double max_row_sum, Csr &S)
{
   int n = A.nrows;
   TimedLaunch tl("setup_strength", 0.0);
   S.nrows = n;
   S.ncols = A.ncols;
   S.ia = dalloc_n<int>((size_t) n + 1);
   int nb = std::max(1, std::min((n + 255) / 256, 8192));
   k_strength_count<<<nb, 256, 0, stream()>>>(n, A.ia, A.ja, A.a, theta, max_row_sum, S.ia);
   HX_LAUNCH_CHECK();
   S.nnz = exclusive_scan(S.ia, S.ia, n);
   S.ja = dalloc_n<int>((size_t) std::max<long long>(S.nnz, 1));
   S.a = nullptr;
   k_strength_fill<<<nb, 256, 0, stream()>>>(n, A.ia, A.ja, A.a, theta, max_row_sum, S.ia, S.ja);
   HX_LAUNCH_CHECK();
}

// ================================================================== l1 norms
// src/parcsr_ls/ams.c:4523-4600 option 1 + ams.c:770-797
__global__ void k_l1(int n, const int *__restrict__ ia, const double *__restrict__ a,
                     double *__restrict__ l1, int *__restrict__ zero_flag)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      double s = 0.0;
      int b = ia[i], e = ia[i + 1];
      for (int j = b; j < e; j++) { s += fabs(a[j]); }
      if (a[b] < 0.0) { s = -s; }
      if (fabs(s) == 0.0) { *zero_flag = 1; }
      l1[i] = s;
   }
}

void l1_norms(const Csr &A, double *l1)
{
   int *flag = device_flags() + 40;
   dzero(flag, sizeof(int));
   int nb = std::max(1, std::min((A.nrows + 255) / 256, 8192));
   k_l1<<<nb, 256, 0, stream()>>>(A.nrows, A.ia, A.a, l1, flag);
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
__global__ void k_pmis_init(int n, const int *__restrict__ S_ia, const int *__restrict__ cnt,
                            const double *__restrict__ given, int seed, double *__restrict__ m,
                            int *__restrict__ CF)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      double mi = given ? given[i] : (double) cnt[i] + pm_draw(seed, i);
      if (S_ia[i + 1] == S_ia[i]) {
         CF[i] = -3;                 // SF_PT
         m[i] = 0.0;
      } else {
         m[i] = mi;
         CF[i] = (mi > 1) ? 1 : 0;
      }
   }
}

// par_coarsen.c:2506-2550: drop candidates with a larger-measure S-neighbour
__global__ void k_pmis_remove(int n, const int *__restrict__ S_ia, const int *__restrict__ S_ja,
                              const double *__restrict__ m, int *CF)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      double mi = m[i];
      if (!(mi > 1)) { continue; }
      for (int k = S_ia[i]; k < S_ia[i + 1]; k++) {
         int j = S_ja[k];
         double mj = m[j];
         if (mj > 1) {
            if (mi > mj) { CF[j] = 0; }
            else if (mj > mi) { CF[i] = 0; }
         }
      }
   }
}

// par_coarsen.c:2603-2660: C points, and F points depending strongly on a C point
__global__ void k_pmis_cf(int n, const int *__restrict__ S_ia, const int *__restrict__ S_ja,
                          const double *__restrict__ m, int *CF)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      double mi = m[i];
      if (!(mi > 0)) { continue; }   // not in the graph
      int c = CF[i];
      if (mi < 1) { c = -1; }
      if (c > 0) {
         c = 1;
      } else {
         for (int k = S_ia[i]; k < S_ia[i + 1]; k++) {
            if (CF[S_ja[k]] > 0) { c = -1; break; }
         }
      }
      CF[i] = c;
   }
}

// par_coarsen.c:2690-2750 graph update, fused with the next round's
// candidate flags (par_coarsen.c:2485-2495)
__global__ void k_pmis_update(int n, double *__restrict__ m, int *__restrict__ CF,
                              int *__restrict__ remaining)
{
   __shared__ int sh;
   if (threadIdx.x == 0) { sh = 0; }
   __syncthreads();
   int local = 0;
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      double mi = m[i];
      if (!(mi > 0)) { continue; }
      if (CF[i] != 0) {
         m[i] = 0.0;
      } else {
         local++;
         if (mi > 1) { CF[i] = 1; }
      }
   }
   if (local) { atomicAdd(&sh, local); }
   __syncthreads();
   if (threadIdx.x == 0 && sh) { atomicAdd(remaining, sh); }
}

int pmis(const Csr &S, const double *measure_in_dev, int *CF, double *measure_keep_dev)
{
   int n = S.nrows;
   if (n == 0) { return 0; }
   TimedLaunch tl("setup_pmis", 0.0);
   hipStream_t st = stream();
   int nb = std::max(1, std::min((n + 255) / 256, 8192));
   double *m = dalloc_n<double>(n);
   int *cnt = nullptr;
   if (!measure_in_dev) {
      cnt = dalloc_n<int>(n);
      dzero(cnt, sizeof(int) * n);
      int nbe = std::max<long long>(1, std::min<long long>((S.nnz + 255) / 256, 8192));
      if (S.nnz > 0) { k_col_count<<<nbe, 256, 0, st>>>(S.nnz, S.ja, cnt); }
      HX_LAUNCH_CHECK();
   }
   k_pmis_init<<<nb, 256, 0, st>>>(n, S.ia, cnt, measure_in_dev, 2747, m, CF);
   HX_LAUNCH_CHECK();
   if (measure_keep_dev) { d2d(measure_keep_dev, m, sizeof(double) * n); }
   int *remaining = device_flags() + 44;
   int *h = static_cast<int *>(pinned_scratch(64));
   int rounds = 0;
   while (true) {
      dzero(remaining, sizeof(int));
      k_pmis_remove<<<nb, 256, 0, st>>>(n, S.ia, S.ja, m, CF);
      k_pmis_cf<<<nb, 256, 0, st>>>(n, S.ia, S.ja, m, CF);
      k_pmis_update<<<nb, 256, 0, st>>>(n, m, CF, remaining);
      HX_LAUNCH_CHECK();
      rounds++;
      HX_HIP(hipMemcpyAsync(h, remaining, sizeof(int), hipMemcpyDeviceToHost, st));
      sync();
      if (h[0] == 0) { break; }
      if (rounds > 100000) { fail(HYPRE_ERROR_GENERIC, "PMIS did not terminate"); }
   }
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
   __device__ __forceinline__ int insert(int key, bool &isnew)
   {
      unsigned h = h0(key);
      for (int probe = 0; probe < cap; probe++) {
         int old = atomicCAS(&keys[h], -1, key);
         if (old == -1) { isnew = true; return (int) h; }
         if (old == key) { isnew = false; return (int) h; }
         h = (h + 1) & (cap - 1);
      }
      isnew = false;
      return -1;
   }
   __device__ __forceinline__ int find(int key) const
   {
      unsigned h = h0(key);
      for (int probe = 0; probe < cap; probe++) {
         int k = keys[h];
         if (k == key) { return (int) h; }
         if (k == -1) { return -1; }
         h = (h + 1) & (cap - 1);
      }
      return -1;
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
   const int per = (p.nwork + nb - 1) / nb;
   const int w0 = blk * per, w1 = min(p.nwork, w0 + per);

   for (int wi = w0; wi < w1; wi++) {
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
                            int max_global_cap, int ntables, Launch launch)
{
   int *h = static_cast<int *>(pinned_scratch(64));
   int *lists[2] = {dalloc_n<int>((size_t) nwork_all + 1), dalloc_n<int>((size_t) nwork_all + 1)};
   a.rowlist = nullptr;
   a.nwork = nwork_all;
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
static void extpi_pass(InterpArgs a)
{
   static const int caps[2] = {512, 4096};
   deferred_passes(a, a.n, device_flags() + 48, caps, 2, 1 << 22, 1,
                   [](InterpArgs &x, int nb) {
                      size_t lds = x.gslab ? 0 : table_bytes(x.cap);
                      k_extpi<MODE><<<nb, WAVE, lds, stream()>>>(x);
                      HX_LAUNCH_CHECK();
                   });
}

__global__ void k_compact_rows(int n, const int *__restrict__ base, const int *__restrict__ cnt,
                               const int *__restrict__ out_ia, const int *__restrict__ in_j,
                               const double *__restrict__ in_a, int stride,
                               int *__restrict__ oj, double *__restrict__ oa)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      int b = base ? base[i] : i * stride;
      int o = out_ia[i], c = cnt[i];
      for (int k = 0; k < c; k++) {
         oj[o + k] = in_j[b + k];
         oa[o + k] = in_a[b + k];
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
                  double trunc_factor, int max_elmts, Csr &P)
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

// ================================================================== transpose
// src/seq_mv/csr_matop.c:1071 (stable counting sort): rows of R list the fine
// rows in ascending order.  Fill with atomics, then sort each (short) row.
__global__ void k_t_fill(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                         const double *__restrict__ a, int *__restrict__ fillp,
                         int *__restrict__ Rj, double *__restrict__ Ra)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      for (int k = ia[i]; k < ia[i + 1]; k++) {
         int pos = atomicAdd(&fillp[ja[k]], 1);
         Rj[pos] = i;
         Ra[pos] = a[k];
      }
   }
}

__global__ void k_t_sort(int nc, const int *__restrict__ Ria, int *__restrict__ Rj,
                         double *__restrict__ Ra)
{
   for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nc; r += gridDim.x * blockDim.x) {
      int b = Ria[r], e = Ria[r + 1];
      for (int k = b + 1; k < e; k++) {
         int key = Rj[k];
         double v = Ra[k];
         int m = k - 1;
         while (m >= b && Rj[m] > key) {
            Rj[m + 1] = Rj[m];
            Ra[m + 1] = Ra[m];
            m--;
         }
         Rj[m + 1] = key;
         Ra[m + 1] = v;
      }
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
   int nbe = (int) std::max<long long>(1, std::min<long long>((P.nnz + 255) / 256, 8192));
   if (P.nnz > 0) { k_col_count<<<nbe, 256, 0, stream()>>>(P.nnz, P.ja, R.ia); }
   HX_LAUNCH_CHECK();
   exclusive_scan(R.ia, R.ia, nc);
   int *fillp = dalloc_n<int>((size_t) nc + 1);
   d2d(fillp, R.ia, sizeof(int) * ((size_t) nc + 1));
   int nb = std::max(1, std::min((n + 255) / 256, 8192));
   k_t_fill<<<nb, 256, 0, stream()>>>(n, P.ia, P.ja, P.a, fillp, R.ja, R.a);
   int nbc = std::max(1, std::min((nc + 255) / 256, 8192));
   k_t_sort<<<nbc, 256, 0, stream()>>>(nc, R.ia, R.ja, R.a);
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
   const int per = (p.nwork + nb - 1) / nb;
   const int w0 = blk * per, w1 = min(p.nwork, w0 + per);

   for (int wi = w0; wi < w1; wi++) {
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
static void rap_pass(RapArgs a)
{
   static const int caps[2] = {512, 2048};
   deferred_passes(a, a.nc, device_flags() + 52, caps, 2, 1 << 22, 2,
                   [](RapArgs &x, int nb) {
                      size_t lds = x.gslab ? 0 : 2 * table_bytes(x.cap);
                      k_rap<MODE><<<nb, WAVE, lds, stream()>>>(x);
                      HX_LAUNCH_CHECK();
                   });
}

void rap(const Csr &R, const Csr &A, const Csr &P, Csr &Ac)
{
   int nc = R.nrows;
   TimedLaunch tl("setup_rap", 0.0);
   RapArgs a{};
   a.nc = nc;
   a.R_ia = R.ia; a.R_ja = R.ja; a.R_a = R.a;
   a.A_ia = A.ia; a.A_ja = A.ja; a.A_a = A.a;
   a.P_ia = P.ia; a.P_ja = P.ja; a.P_a = P.a;
   int *cnt = dalloc_n<int>((size_t) nc + 1);
   a.cnt = cnt;
   rap_pass<MODE_SYM>(a);
   Ac.nrows = nc;
   Ac.ncols = nc;
   Ac.ia = dalloc_n<int>((size_t) nc + 1);
   Ac.nnz = exclusive_scan(cnt, Ac.ia, nc);
   Ac.ja = dalloc_n<int>((size_t) std::max<long long>(Ac.nnz, 1));
   Ac.a = dalloc_n<double>((size_t) std::max<long long>(Ac.nnz, 1));
   Ac.a32 = nullptr;
   a.out_ia = Ac.ia;
   a.out_j = Ac.ja;
   a.out_a = Ac.a;
   rap_pass<MODE_NUM>(a);
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
