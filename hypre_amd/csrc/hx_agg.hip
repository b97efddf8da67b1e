// Aggressive coarsening (agg_num_levels) with modified multipass
// interpolation (agg_interp_type 8), one rank, num_paths 1.
//
// src/parcsr_ls/par_amg_setup.c:1311-1352: on an aggressive level the C
// points of the first PMIS split are coarsened again, over the strength graph
// of distance one or two between C points (hypre_BoomerAMGCreate2ndS,
// par_strength.c:1795), by PMIS with CF_init 3; par_amg_setup.c:1799-1824
// corrects the first split with the second (hypre_BoomerAMGCorrectCFMarker,
// par_strength.c:3068) and builds P with hypre_BoomerAMGBuildModMultipass
// (par_mod_multi_interp.c:17).
//
// Device layout:
// * k_second_s: one wave per C row.  The row's candidates (for each strong
//   neighbour i2 of i1: i2 itself, then the strong neighbours of i2) are
//   enumerated in the reference's traversal order, 64 at a time; an LDS hash
//   per row keeps each coarse column once and, through an atomicMin on the
//   candidate index, the position of its first occurrence, so the row is
//   written in the reference's first-seen order.
// * passes: one thread per point and pass; a point joins pass p+1 when a
//   strong neighbour is in pass p (a point read while it is being marked
//   holds 0 or p+1, never p, so the marking does not depend on timing).
// * k_multipass_q: one thread per row of a pass, the one-step weights of
//   hypre_GenerateMultipassPi (par_mod_multi_interp.c:530-800).
// * k_matmul_fs: the product of the one-step weights with the previous pass's
//   weights (hypre_ParMatmul, par_csr_matop.c:348), one wave per row: the
//   row's products are staged in LDS in traversal order, the first occurrence
//   of each column found with an LDS hash and atomicMin, and each output
//   value summed from its products left to right (first product stored, the
//   others added), the reference's order.
#include "hx_internal.h"

#include <algorithm>
#include <climits>
#include <vector>

namespace hx {

namespace {

constexpr int AW = 64;           // one wave per block
constexpr int S2_HS = 1024;      // hash slots of a row of the second strength matrix
constexpr int S2_SHIFT = 22;     // 32 - log2(S2_HS)
constexpr int S2_MAXU = 768;     // distinct columns of a row before it is refused
constexpr int MM_T = 1024;       // staged products of a row of the product
constexpr int MM_HS = 2048;      // hash slots of a row of the product (> MM_T)
constexpr int MM_SHIFT = 21;     // 32 - log2(MM_HS)
constexpr int MP_MAX_ITER = 9;   // the reference stops with an error in its 9th pass iteration

__device__ __forceinline__ int wave_excl_scan(int v, int lane, int &total)
{
   int x = v;
   for (int d = 1; d < AW; d <<= 1) {
      const int y = __shfl_up(x, d, AW);
      if (lane >= d) { x += y; }
   }
   total = __shfl(x, AW - 1, AW);
   return x - v;
}

// linear probing; the callers keep the table below 3/4 full
template <int SHIFT, int MASK>
__device__ __forceinline__ int hash_insert(int *keys, int key)
{
   unsigned h = ((unsigned) key * 2654435761u) >> SHIFT;
   while (true) {
      const int old = atomicCAS(&keys[h], -1, key);
      if (old == -1 || old == key) { return (int) h; }
      h = (h + 1) & MASK;
   }
}

template <int SHIFT, int MASK>
__device__ __forceinline__ int hash_find(const int *keys, int key)
{
   unsigned h = ((unsigned) key * 2654435761u) >> SHIFT;
   while (keys[h] != key) { h = (h + 1) & MASK; }
   return (int) h;
}

// segment of item t: the last lane whose segment starts at or before t
// (starts are non-decreasing; lanes without a segment start at the total)
__device__ __forceinline__ int find_segment(const int *start, int t)
{
   int lo = 0, hi = AW - 1;
   while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (start[mid] <= t) { lo = mid; } else { hi = mid - 1; }
   }
   return lo;
}

__global__ void k_c2f(int n, const int *__restrict__ CF, const int *__restrict__ f2c,
                      int *__restrict__ c2f)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      if (CF[i] == 1) { c2f[f2c[i]] = i; }
   }
}

// par_strength.c:2433-2466 (counting, FILL false) and :2740-2790 (filling)
template <bool FILL>
__global__ __launch_bounds__(AW) void k_second_s(int nc, const int *__restrict__ c2f,
                                                 const int *__restrict__ S_ia,
                                                 const int *__restrict__ S_ja,
                                                 const int *__restrict__ CF,
                                                 const int *__restrict__ f2c,
                                                 int *__restrict__ cnt,
                                                 const int *__restrict__ out_ia,
                                                 int *__restrict__ out_ja,
                                                 int *__restrict__ overflow)
{
   __shared__ int keys[S2_HS];
   __shared__ int pos[S2_HS];
   __shared__ int seg_start[AW];
   __shared__ int seg_i2[AW];
   __shared__ int seg_s2[AW];
   const int lane = threadIdx.x;
   for (int ic = blockIdx.x; ic < nc; ic += gridDim.x) {
      for (int h = lane; h < S2_HS; h += AW) {
         keys[h] = -1;
         pos[h] = INT_MAX;
      }
      __syncthreads();
      const int i1 = c2f[ic];
      const int s = S_ia[i1], e = S_ia[i1 + 1];
      int count = 0;       // distinct columns so far (wave-uniform)
      int gbase = 0;       // candidate index of the first item of this chunk
      bool over = false;
      for (int k0 = s; k0 < e && !over; k0 += AW) {
         const int k = k0 + lane;
         int i2 = 0, s2 = 0, len = 0;
         if (k < e) {
            i2 = S_ja[k];
            s2 = S_ia[i2];
            len = 1 + (S_ia[i2 + 1] - s2);
         }
         int T;
         const int off = wave_excl_scan(len, lane, T);
         seg_start[lane] = (k < e) ? off : T;
         seg_i2[lane] = i2;
         seg_s2[lane] = s2;
         __syncthreads();
         for (int t0 = 0; t0 < T && !over; t0 += AW) {
            const int t = t0 + lane;
            int key = -1;
            if (t < T) {
               const int sg = find_segment(seg_start, t);
               const int idx = t - seg_start[sg];
               if (idx == 0) {
                  const int p = seg_i2[sg];
                  if (CF[p] > 0) { key = f2c[p]; }
               } else {
                  const int p = S_ja[seg_s2[sg] + idx - 1];
                  if (CF[p] > 0) {
                     const int c = f2c[p];
                     if (c != ic) { key = c; }
                  }
               }
            }
            int h = -1;
            if (key >= 0) {
               h = hash_insert<S2_SHIFT, S2_HS - 1>(keys, key);
               atomicMin(&pos[h], gbase + t);
            }
            __syncthreads();
            const bool first = key >= 0 && pos[h] == gbase + t;
            const unsigned long long bm = __ballot(first);
            if (FILL && first) {
               out_ja[out_ia[ic] + count + __popcll(bm & ((1ULL << lane) - 1))] = key;
            }
            count += __popcll(bm);
            if (count > S2_MAXU) { over = true; }
            __syncthreads();
         }
         gbase += T;
      }
      if (lane == 0) {
         if (!FILL) { cnt[ic] = over ? 0 : count; }
         if (over) { atomicOr(overflow, 1); }
      }
      __syncthreads();
   }
}

// par_strength.c:3068-3087 (no point carries the mark 2 here: a C point with
// an empty row of S2 has no strong connection there and becomes C in the
// second coarsening, which is what the mark stands for)
__global__ void k_correct_cf(int n, int *__restrict__ CF, const int *__restrict__ f2c1,
                             const int *__restrict__ CFN)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      if (CF[i] == 1) { CF[i] = CFN[f2c1[i]]; }
   }
}

__global__ void k_pass_init(int n, const int *__restrict__ CF, int *__restrict__ pm)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      pm[i] = (CF[i] == 1) ? 1 : 0;
   }
}

// par_mod_multi_interp.c:151-219, one pass
__global__ __launch_bounds__(256) void k_pass_mark(int n, const int *__restrict__ S_ia,
                                                   const int *__restrict__ S_ja, int *pm, int p,
                                                   int *__restrict__ newly)
{
   __shared__ int ws[4];
   int mine = 0;
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      if (pm[i] != 0) { continue; }
      for (int k = S_ia[i]; k < S_ia[i + 1]; k++) {
         if (pm[S_ja[k]] == p) {
            pm[i] = p + 1;
            mine++;
            break;
         }
      }
   }
   for (int d = 32; d > 0; d >>= 1) { mine += __shfl_xor(mine, d, 64); }
   if ((threadIdx.x & 63) == 0) { ws[threadIdx.x >> 6] = mine; }
   __syncthreads();
   if (threadIdx.x == 0) {
      const int tot = ws[0] + ws[1] + ws[2] + ws[3];
      if (tot) { atomicAdd(newly, tot); }
   }
}

__global__ void k_pass_flag(int n, const int *__restrict__ pm, int p, int *__restrict__ flag)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      flag[i] = (pm[i] == p) ? 1 : 0;
   }
}

// rank of each point of pass p among the points of that pass (row order),
// and the points of the pass in that order
__global__ void k_pass_rank(int n, const int *__restrict__ pm, int p,
                            const int *__restrict__ scanned, int *__restrict__ rank,
                            int *__restrict__ order)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      if (pm[i] == p) {
         rank[i] = scanned[i];
         order[scanned[i]] = i;
      }
   }
}

// par_mod_multi_interp.c:221-265: off-diagonal row sums of the F rows
__global__ void k_mp_row_sums(int n, const int *__restrict__ CF, const int *__restrict__ A_ia,
                              const double *__restrict__ A_a, double *__restrict__ rs)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      double s = 0.0;
      if (CF[i] < 0) {
         for (int k = A_ia[i] + 1; k < A_ia[i + 1]; k++) { s += A_a[k]; }
      }
      rs[i] = s;
   }
}

// par_mod_multi_interp.c:692-765: rows of pass color+1 (order[r]), the
// entries a_ij of the strong connections j of pass `color` (S order, the
// entry of A found by a forward search from the diagonal), then scaled by
// -row_sums[i] / (sum_j a_ij * a_ii) (the division only when that product is
// nonzero).  Columns: rank of j within pass `color`.
template <bool FILL>
__global__ void k_multipass_q(int nrows, const int *__restrict__ order, const int *__restrict__ S_ia,
                              const int *__restrict__ S_ja, const int *__restrict__ A_ia,
                              const int *__restrict__ A_ja, const double *__restrict__ A_a,
                              const int *__restrict__ pm, int color, const int *__restrict__ rank,
                              const double *__restrict__ row_sums, int *__restrict__ cnt,
                              const int *__restrict__ q_ia, int *__restrict__ q_ja,
                              double *__restrict__ q_a, int *__restrict__ bad)
{
   for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nrows; r += gridDim.x * blockDim.x) {
      const int i1 = order[r];
      if (!FILL) {
         int c = 0;
         for (int k = S_ia[i1]; k < S_ia[i1 + 1]; k++) { c += (pm[S_ja[k]] == color) ? 1 : 0; }
         cnt[r] = c;
         continue;
      }
      const int a_end = A_ia[i1 + 1];
      int j2 = A_ia[i1];
      int o = q_ia[r];
      const int o0 = o;
      for (int k = S_ia[i1]; k < S_ia[i1 + 1]; k++) {
         const int j1 = S_ja[k];
         while (j2 < a_end && A_ja[j2] != j1) { j2++; }
         if (j2 == a_end) {
            // S is not a subsequence of A's row: cannot happen for S built
            // from A; reported instead of reading past the row
            atomicOr(bad, 1);
            break;
         }
         if (pm[j1] == color) {
            q_ja[o] = rank[j1];
            q_a[o] = A_a[j2];
            o++;
         }
      }
      double sum_c = 0.0;
      for (int e = o0; e < o; e++) { sum_c += q_a[e]; }
      const double value = sum_c * A_a[A_ia[i1]];
      double rs = row_sums[i1];
      if (value != 0) { rs = rs / value; }
      for (int e = o0; e < o; e++) { q_a[e] = -q_a[e] * rs; }
   }
}

// par_csr_matop.c:348-1140 on one rank: row r of Q W, first-seen column order,
// first product stored and the others added in traversal order; a square
// product opens each row with its diagonal position (value 0).
template <bool FILL>
__global__ __launch_bounds__(AW) void k_matmul_fs(int nrows, int allsquare,
                                                  const int *__restrict__ Q_ia,
                                                  const int *__restrict__ Q_ja,
                                                  const double *__restrict__ Q_a,
                                                  const int *__restrict__ W_ia,
                                                  const int *__restrict__ W_ja,
                                                  const double *__restrict__ W_a,
                                                  int *__restrict__ cnt,
                                                  const int *__restrict__ o_ia,
                                                  int *__restrict__ o_ja, double *__restrict__ o_a,
                                                  int *__restrict__ overflow)
{
   __shared__ int pc[MM_T];
   __shared__ double pv[MM_T];
   __shared__ int first_t[MM_T];
   __shared__ int keys[MM_HS];
   __shared__ int pos[MM_HS];
   __shared__ int seg_start[AW];
   __shared__ int seg_w[AW];
   __shared__ double seg_q[AW];
   const int lane = threadIdx.x;
   for (int r = blockIdx.x; r < nrows; r += gridDim.x) {
      for (int h = lane; h < MM_HS; h += AW) {
         keys[h] = -1;
         pos[h] = INT_MAX;
      }
      int T = 0;           // staged products (wave-uniform)
      if (allsquare) {
         if (lane == 0) {
            pc[0] = r;
            pv[0] = 0.0;
         }
         T = 1;
      }
      bool over = false;
      const int qs = Q_ia[r], qe = Q_ia[r + 1];
      for (int q0 = qs; q0 < qe && !over; q0 += AW) {
         const int k = q0 + lane;
         int len = 0, wb = 0;
         double qv = 0.0;
         if (k < qe) {
            const int c = Q_ja[k];
            wb = W_ia[c];
            len = W_ia[c + 1] - wb;
            qv = Q_a[k];
         }
         int tot;
         const int off = wave_excl_scan(len, lane, tot);
         if (T + tot > MM_T) {
            over = true;
            break;
         }
         __syncthreads();
         seg_start[lane] = (k < qe) ? off : tot;
         seg_w[lane] = wb;
         seg_q[lane] = qv;
         __syncthreads();
         for (int t0 = 0; t0 < tot; t0 += AW) {
            const int t = t0 + lane;
            if (t < tot) {
               const int sg = find_segment(seg_start, t);
               const int e = seg_w[sg] + (t - seg_start[sg]);
               pc[T + t] = W_ja[e];
               pv[T + t] = seg_q[sg] * W_a[e];
            }
         }
         T += tot;
      }
      __syncthreads();
      if (over) {
         if (lane == 0) {
            if (!FILL) { cnt[r] = 0; }
            atomicOr(overflow, 1);
         }
         __syncthreads();
         continue;
      }
      for (int t = lane; t < T; t += AW) {
         const int h = hash_insert<MM_SHIFT, MM_HS - 1>(keys, pc[t]);
         atomicMin(&pos[h], t);
      }
      __syncthreads();
      int count = 0;
      for (int t0 = 0; t0 < T; t0 += AW) {
         const int t = t0 + lane;
         bool first = false;
         if (t < T) { first = pos[hash_find<MM_SHIFT, MM_HS - 1>(keys, pc[t])] == t; }
         const unsigned long long bm = __ballot(first);
         if (first) { first_t[count + __popcll(bm & ((1ULL << lane) - 1))] = t; }
         count += __popcll(bm);
      }
      if (!FILL) {
         if (lane == 0) { cnt[r] = count; }
      } else {
         __syncthreads();
         const int ob = o_ia[r];
         for (int u = lane; u < count; u += AW) {
            const int t1 = first_t[u];
            const int c = pc[t1];
            double v = pv[t1];
            for (int t = t1 + 1; t < T; t++) {
               if (pc[t] == c) { v += pv[t]; }
            }
            o_ja[ob + u] = c;
            o_a[ob + u] = v;
         }
      }
      __syncthreads();
   }
}

// the weights of every pass, indexed by pass (2 .. 10)
struct PassWeights
{
   const int    *ia[MP_MAX_ITER + 2];
   const int    *ja[MP_MAX_ITER + 2];
   const double *a[MP_MAX_ITER + 2];
};

// par_mod_multi_interp.c:266-330: C rows are unit rows, the row of a point
// of pass p >= 2 is its row of that pass's weights; unreached points get
// empty rows
__global__ void k_mp_count(int n, const int *__restrict__ pm, const int *__restrict__ rank,
                           PassWeights W, int *__restrict__ cnt)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      const int p = pm[i];
      int c = 0;
      if (p == 1) {
         c = 1;
      } else if (p >= 2) {
         const int r = rank[i];
         c = W.ia[p][r + 1] - W.ia[p][r];
      }
      cnt[i] = c;
   }
}

__global__ void k_mp_fill(int n, const int *__restrict__ pm, const int *__restrict__ rank,
                          PassWeights W, const int *__restrict__ P_ia, int *__restrict__ P_ja,
                          double *__restrict__ P_a)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      const int p = pm[i];
      const int o = P_ia[i];
      if (p == 1) {
         P_ja[o] = rank[i];
         P_a[o] = 1.0;
      } else if (p >= 2) {
         const int r = rank[i];
         const int s = W.ia[p][r], e = W.ia[p][r + 1];
         for (int k = s; k < e; k++) {
            P_ja[o + k - s] = W.ja[p][k];
            P_a[o + k - s] = W.a[p][k];
         }
      }
   }
}

__global__ void k_cf_sf_to_f(int n, int *CF)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      if (CF[i] == -3) { CF[i] = -1; }
   }
}

int grid1(long long n)
{
   return (int) std::max<long long>(1, std::min<long long>((n + 255) / 256, 8192));
}

int grid_rows(long long rows)
{
   return (int) std::max<long long>(1, std::min<long long>(rows, 65536));
}

int read_flag(int *dev)
{
   int h = 0;
   d2h(&h, dev, sizeof(int));
   return h;
}

} // namespace

void create_2nd_s(const Csr &S, const int *CF, const int *f2c, int nc, Csr &S2)
{
   TimedLaunch tl("setup_second_s", 0.0);
   const int n = S.nrows;
   S2.free_all();
   S2.nrows = nc;
   S2.ncols = nc;
   S2.ia = dalloc_n<int>((size_t) nc + 1);
   if (nc == 0) {
      dzero(S2.ia, sizeof(int));
      S2.nnz = 0;
      S2.ja = dalloc_n<int>(1);
      return;
   }
   int *c2f = dalloc_n<int>((size_t) nc);
   k_c2f<<<grid1(n), 256, 0, stream()>>>(n, CF, f2c, c2f);
   HX_LAUNCH_CHECK();
   int *flag = device_flags() + 72;
   dzero(flag, sizeof(int));
   k_second_s<false><<<grid_rows(nc), AW, 0, stream()>>>(nc, c2f, S.ia, S.ja, CF, f2c, S2.ia,
                                                         nullptr, nullptr, flag);
   HX_LAUNCH_CHECK();
   if (read_flag(flag)) {
      dfree(c2f);
      fail(HYPRE_ERROR_GENERIC, "aggressive coarsening: a row of the second strength matrix "
                                "exceeds 768 entries");
   }
   S2.nnz = exclusive_scan(S2.ia, S2.ia, nc);
   S2.ja = dalloc_n<int>((size_t) std::max<long long>(S2.nnz, 1));
   k_second_s<true><<<grid_rows(nc), AW, 0, stream()>>>(nc, c2f, S.ia, S.ja, CF, f2c, nullptr,
                                                        S2.ia, S2.ja, flag);
   HX_LAUNCH_CHECK();
   dfree(c2f);
}

void correct_cf_marker(int n, int *CF, const int *f2c1, const int *CFN)
{
   k_correct_cf<<<grid1(n), 256, 0, stream()>>>(n, CF, f2c1, CFN);
   HX_LAUNCH_CHECK();
}

static void matmul_fs(const Csr &Q, const Csr &W, Csr &out)
{
   out.free_all();
   out.nrows = Q.nrows;
   out.ncols = W.ncols;
   out.ia = dalloc_n<int>((size_t) Q.nrows + 1);
   if (Q.nrows == 0) {
      dzero(out.ia, sizeof(int));
      out.ja = dalloc_n<int>(1);
      out.a = dalloc_n<double>(1);
      return;
   }
   const int allsquare = (Q.nrows == W.ncols) ? 1 : 0;
   int *flag = device_flags() + 73;
   dzero(flag, sizeof(int));
   k_matmul_fs<false><<<grid_rows(Q.nrows), AW, 0, stream()>>>(Q.nrows, allsquare, Q.ia, Q.ja, Q.a,
                                                               W.ia, W.ja, W.a, out.ia, nullptr,
                                                               nullptr, nullptr, flag);
   HX_LAUNCH_CHECK();
   if (read_flag(flag)) {
      fail(HYPRE_ERROR_GENERIC, "multipass interpolation: a row of a pass product exceeds "
                                "1024 products");
   }
   out.nnz = exclusive_scan(out.ia, out.ia, Q.nrows);
   out.ja = dalloc_n<int>((size_t) std::max<long long>(out.nnz, 1));
   out.a = dalloc_n<double>((size_t) std::max<long long>(out.nnz, 1));
   k_matmul_fs<true><<<grid_rows(Q.nrows), AW, 0, stream()>>>(Q.nrows, allsquare, Q.ia, Q.ja, Q.a,
                                                              W.ia, W.ja, W.a, nullptr, out.ia,
                                                              out.ja, out.a, flag);
   HX_LAUNCH_CHECK();
}

void mod_multipass_interp(const Csr &A, const Csr &S, int *CF, int nc, double trunc_factor,
                          int max_elmts, Csr &P)
{
   TimedLaunch tl("setup_interp", 0.0);
   const int n = A.nrows;
   const int nb = grid1(n);
   int *pm = dalloc_n<int>((size_t) std::max(n, 1));
   int *rank = dalloc_n<int>((size_t) std::max(n, 1));
   int *scan = dalloc_n<int>((size_t) n + 1);
   double *rs = dalloc_n<double>((size_t) std::max(n, 1));
   int *newly = device_flags() + 74;
   int *bad = device_flags() + 75;
   dzero(bad, sizeof(int));
   k_pass_init<<<nb, 256, 0, stream()>>>(n, CF, pm);
   HX_LAUNCH_CHECK();
   // par_mod_multi_interp.c:151-219
   int remaining = n - nc;
   int last_pass = 1;
   while (remaining > 0) {
      if (last_pass >= MP_MAX_ITER) {
         dfree(pm); dfree(rank); dfree(scan); dfree(rs);
         fail(HYPRE_ERROR_GENERIC, "multipass interpolation: too many passes");
      }
      dzero(newly, sizeof(int));
      k_pass_mark<<<nb, 256, 0, stream()>>>(n, S.ia, S.ja, pm, last_pass, newly);
      HX_LAUNCH_CHECK();
      const int got = read_flag(newly);
      if (got == 0) { break; }
      remaining -= got;
      last_pass++;
   }
   // ranks within each pass (pass 1: the coarse index) and each pass's rows
   std::vector<int> pass_n(last_pass + 1, 0);
   std::vector<int *> order(last_pass + 1, nullptr);
   for (int p = 1; p <= last_pass; p++) {
      k_pass_flag<<<nb, 256, 0, stream()>>>(n, pm, p, scan);
      HX_LAUNCH_CHECK();
      pass_n[p] = (int) exclusive_scan(scan, scan, n);
      order[p] = dalloc_n<int>((size_t) std::max(pass_n[p], 1));
      k_pass_rank<<<nb, 256, 0, stream()>>>(n, pm, p, scan, rank, order[p]);
      HX_LAUNCH_CHECK();
   }
   if (pass_n[1] != nc) { fail(HYPRE_ERROR_GENERIC, "multipass interpolation: C count mismatch"); }
   k_mp_row_sums<<<nb, 256, 0, stream()>>>(n, CF, A.ia, A.a, rs);
   HX_LAUNCH_CHECK();
   // one-step weights of each pass, chained through the products
   std::vector<Csr> Wp(last_pass + 1);
   for (int p = 2; p <= last_pass; p++) {
      Csr Q;
      const int nr = pass_n[p];
      Q.nrows = nr;
      Q.ncols = pass_n[p - 1];
      Q.ia = dalloc_n<int>((size_t) nr + 1);
      const int nbr = grid1(nr);
      k_multipass_q<false><<<nbr, 256, 0, stream()>>>(nr, order[p], S.ia, S.ja, A.ia, A.ja, A.a, pm,
                                                      p - 1, rank, rs, Q.ia, nullptr, nullptr,
                                                      nullptr, bad);
      HX_LAUNCH_CHECK();
      Q.nnz = exclusive_scan(Q.ia, Q.ia, nr);
      Q.ja = dalloc_n<int>((size_t) std::max<long long>(Q.nnz, 1));
      Q.a = dalloc_n<double>((size_t) std::max<long long>(Q.nnz, 1));
      k_multipass_q<true><<<nbr, 256, 0, stream()>>>(nr, order[p], S.ia, S.ja, A.ia, A.ja, A.a, pm,
                                                     p - 1, rank, rs, nullptr, Q.ia, Q.ja, Q.a, bad);
      HX_LAUNCH_CHECK();
      if (p == 2) {
         Wp[p] = Q;
      } else {
         matmul_fs(Q, Wp[p - 1], Wp[p]);
         Q.free_all();
      }
   }
   if (read_flag(bad)) { fail(HYPRE_ERROR_GENERIC, "multipass interpolation: S is not a sub-pattern of A"); }
   PassWeights W{};
   for (int p = 2; p <= last_pass; p++) {
      W.ia[p] = Wp[p].ia;
      W.ja[p] = Wp[p].ja;
      W.a[p] = Wp[p].a;
   }
   P.free_all();
   P.nrows = n;
   P.ncols = nc;
   P.ia = dalloc_n<int>((size_t) n + 1);
   k_mp_count<<<nb, 256, 0, stream()>>>(n, pm, rank, W, P.ia);
   HX_LAUNCH_CHECK();
   P.nnz = exclusive_scan(P.ia, P.ia, n);
   P.ja = dalloc_n<int>((size_t) std::max<long long>(P.nnz, 1));
   P.a = dalloc_n<double>((size_t) std::max<long long>(P.nnz, 1));
   k_mp_fill<<<nb, 256, 0, stream()>>>(n, pm, rank, W, P.ia, P.ja, P.a);
   HX_LAUNCH_CHECK();
   for (int p = 2; p <= last_pass; p++) { Wp[p].free_all(); }
   for (int p = 1; p <= last_pass; p++) { dfree(order[p]); }
   dfree(pm); dfree(rank); dfree(scan); dfree(rs);
   // par_mod_multi_interp.c:410-419
   if (trunc_factor != 0.0 || max_elmts > 0) { truncate_rows(P, trunc_factor, max_elmts); }
   k_cf_sf_to_f<<<nb, 256, 0, stream()>>>(n, CF);
   HX_LAUNCH_CHECK();
}

} // namespace hx
