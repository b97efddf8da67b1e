// Solve-phase kernels: CSR SpMV (FP64 / FP32 values), l1-Jacobi and Jacobi
// relaxation, deterministic dot products, fused PCG vector updates and the
// coarsest-grid Gaussian elimination.
//
// SpMV and l1-Jacobi layout: a wave owns a group of consecutive rows and
// streams their nonzeros in coalesced windows (products staged in LDS, each
// row summed by one lane in storage order).  Each block owns a contiguous
// slice of row groups and block ids are remapped so that every XCD (blocks are
// dealt round-robin to the 8 XCDs) streams one contiguous eighth of the
// matrix and its x-gathers stay in that XCD's L2.  Weighted Jacobi (relax
// type 0, off the default path) keeps K lanes per row with a shuffle sum.
#include "hx_internal.h"

#include <algorithm>
#include <climits>
#include <cstdlib>

namespace hx {

constexpr int BS = 256;

__device__ __forceinline__ int xcd_block(int bid, int nb)
{
   // bijection on [0, nb) when nb % 8 == 0: XCD (bid % 8) gets a contiguous range
   return (nb & 7) ? bid : (bid & 7) * (nb >> 3) + (bid >> 3);
}

static int pick_k(long long nnz, int n)
{
   double avg = n > 0 ? (double) nnz / n : 1.0;
   if (avg <= 4.5) { return 4; }
   if (avg <= 9.0) { return 8; }
   if (avg <= 18.0) { return 16; }
   if (avg <= 40.0) { return 32; }
   return 64;
}

static int grid_for_rows(int n, int K)
{
   long long groups = ((long long) n + 0) ;
   long long blocks = (groups * K + BS - 1) / BS;
   blocks = std::min<long long>(blocks, 8192);
   blocks = std::max<long long>(blocks, 1);
   if (blocks >= 64) { blocks = (blocks + 7) / 8 * 8; }
   return (int) blocks;
}

template <int K, typename V, typename X>
__device__ __forceinline__ X row_dot(int s, int e, int lane, const int *__restrict__ ja,
                                     const V *__restrict__ a, const X *__restrict__ x)
{
   X t = 0;
   for (int j = s + lane; j < e; j += K) { t += (X) a[j] * x[ja[j]]; }
#pragma unroll
   for (int off = K / 2; off > 0; off >>= 1) { t += __shfl_xor(t, off, K); }
   return t;
}

// ---------------------------------------------------------------- windowed row sums
// A wave owns a contiguous range of rows and streams their nonzeros, one
// contiguous range of the CSR arrays, in windows of SW entries with fully
// coalesced loads: lane l loads entries 2l and 2l+1 of each 128-entry slice
// (8-byte column-index and 16-byte value loads), gathers their x values and
// writes the products to a per-wave LDS buffer.  The rows are taken RW at a
// time; lane l (< RW) owns row l of the current block and adds up its
// products in storage order, i.e. t = ((0 + a_0 x_0) + a_1 x_1) + ..., the
// order of hypre's host CSR matvec (src/seq_mv/csr_matvec.c).  A block whose
// rows end inside a window is finished there and the next block continues in
// the same window, so the stream never restarts between row blocks.
constexpr int SW = 512;          // default window (entries); 1024 selectable
constexpr int WPB = BS / 64;   // waves per block

// rows per block: one row per lane, fewer only for very long rows so that a
// lane's serial sum stays short
static int pick_rw(long long nnz, int n)
{
   double avg = n > 0 ? (double) nnz / n : 1.0;
   int rw = 64;
   while (rw > 1 && rw * avg > 4 * SW) { rw >>= 1; }
   return rw;
}

__device__ __forceinline__ void wave_fence() { __builtin_amdgcn_wave_barrier(); }

// window size of the SpMV / l1-Jacobi kernels: 1024 entries (measured faster
// on the 27-point 200^3 fine level than 512), or 512 when the environment sets
// HYPRE_MI355X_SPMV_WINDOW=512 (read once); 256 or 512 for matrices whose
// 64-row groups hold at most that many entries on average (interpolation,
// 7-point operators), where a wide window would mostly issue masked loads
static int spmv_window(long long nnz, int n)
{
   static int w = [] {
      const char *e = std::getenv("HYPRE_MI355X_SPMV_WINDOW");
      return (e && std::atoi(e) == SW) ? SW : 1024;
   }();
   const double group = n > 0 ? (double) nnz / n * 64.0 : 0.0;
   if (group <= 256.0) { return 256; }
   if (group <= 512.0) { return 512; }
   return w;
}

typedef int hx_int2 __attribute__((ext_vector_type(2)));
typedef double hx_dbl2 __attribute__((ext_vector_type(2)));
typedef float hx_flt2 __attribute__((ext_vector_type(2)));

// entries p and p+1 (p even) of the column indices and values
template <typename V>
__device__ __forceinline__ void load_pair(const int *__restrict__ ja, const V *__restrict__ a, int p,
                                          int (&j)[2], V (&v)[2])
{
   const hx_int2 jj = __builtin_nontemporal_load(reinterpret_cast<const hx_int2 *>(ja + p));
   j[0] = jj.x;
   j[1] = jj.y;
   if constexpr (sizeof(V) == 8) {
      const hx_dbl2 vv = __builtin_nontemporal_load(reinterpret_cast<const hx_dbl2 *>(a + p));
      v[0] = vv.x;
      v[1] = vv.y;
   } else {
      const hx_flt2 vv = __builtin_nontemporal_load(reinterpret_cast<const hx_flt2 *>(a + p));
      v[0] = vv.x;
      v[1] = vv.y;
   }
}

// Rows [R0, R1) of A times x; epi(row, t) is called once per row, by the lane
// that owns it, with t the row's sum.  The matrix is streamed once: its loads
// are non-temporal so that x stays in L2, and the next window's loads are
// issued before the current window's products are summed.  Windows start at
// the even entry at or below the range's first entry; pair loads past the
// range read the last pair inside it (every CSR array is allocated with at
// least one element of slack) and are masked when used.
template <int RW, int SW, typename V, typename X, class Epi>
__device__ __forceinline__ void rows_stream(int R0, int R1, int lane, const int *__restrict__ ia,
                                            const int *__restrict__ ja, const V *__restrict__ a,
                                            const X *__restrict__ x, X *sbuf, Epi epi)
{
   constexpr int G = SW / 128;
   if (R0 >= R1) { return; }
   int rb = R0;
   int nr = min(RW, R1 - rb);
   int rs = 0, re = 0;
   if (lane < nr) {
      rs = ia[rb + lane];
      re = ia[rb + lane + 1];
   }
   // bounds of the next block, loaded one block ahead
   int nr2 = min(RW, R1 - (rb + RW));
   int rs2 = 0, re2 = 0;
   if (lane < nr2) {
      rs2 = ia[rb + RW + lane];
      re2 = ia[rb + RW + lane + 1];
   }
   const int base = __shfl(rs, 0, 64);
   const int end = ia[R1];
   int bend = __shfl(re, nr - 1, 64);
   X t = 0;
   if (base >= end) {
      // no entries: every row sums to zero
      while (nr > 0) {
         if (lane < nr) { epi(rb + lane, t); }
         rb += RW;
         nr = min(RW, R1 - rb);
      }
      return;
   }
   const int wb0 = base & ~1;
   const int last = (end - 1) & ~1;
   // Loads are unconditional (clamped to the last pair) so that the compiler
   // issues each batch of matrix loads and of x gathers back to back and
   // waits once per batch.
   int jn[G][2];
   V an[G][2];
#pragma unroll
   for (int g = 0; g < G; g++) { load_pair<V>(ja, a, min(wb0 + 2 * (g * 64 + lane), last), jn[g], an[g]); }
   for (int wb = wb0; wb < end; wb += SW) {
      const int we = min(end, wb + SW);
      X xv[G][2];
#pragma unroll
      for (int g = 0; g < G; g++) {
#pragma unroll
         for (int c = 0; c < 2; c++) {
            const int q = wb + 2 * (g * 64 + lane) + c;
            xv[g][c] = x[(q < we) ? jn[g][c] : 0];
         }
      }
      // products straight to LDS (no register copy), then the next window's loads
#pragma unroll
      for (int g = 0; g < G; g++) {
#pragma unroll
         for (int c = 0; c < 2; c++) {
            const int q = wb + 2 * (g * 64 + lane) + c;
            sbuf[2 * (g * 64 + lane) + c] = (q < we) ? (X) an[g][c] * xv[g][c] : (X) 0;
         }
      }
      const int nxt = wb + SW;
      if (nxt < end) {
#pragma unroll
         for (int g = 0; g < G; g++) {
            load_pair<V>(ja, a, min(nxt + 2 * (g * 64 + lane), last), jn[g], an[g]);
         }
      }
      wave_fence();
      const X *q = sbuf - wb;
      while (true) {
         if (lane < nr) {
            // four LDS reads per wait, added in storage order
            const int s = max(rs, wb), e = min(re, we);
            int k = s;
            for (; k + 4 <= e; k += 4) {
               const X p0 = q[k], p1 = q[k + 1], p2 = q[k + 2], p3 = q[k + 3];
               t += p0;
               t += p1;
               t += p2;
               t += p3;
            }
            for (; k < e; k++) { t += q[k]; }
         }
         if (bend > we) { break; }
         // the block's rows end in this window
         if (lane < nr) { epi(rb + lane, t); }
         t = 0;
         rb += RW;
         nr = nr2;
         rs = rs2;
         re = re2;
         if (nr <= 0) { break; }
         bend = __shfl(re, nr - 1, 64);
         nr2 = min(RW, R1 - (rb + RW));
         if (lane < nr2) {
            rs2 = ia[rb + RW + lane];
            re2 = ia[rb + RW + lane + 1];
         }
      }
      wave_fence();
   }
}

// rows of wave wv of block blk: ceil(groups / waves) blocks of RW rows each
template <int RW>
__device__ __forceinline__ void wave_rows(int n, int blk, int wv, int nb, int &R0, int &R1)
{
   const long long ng = ((long long) n + RW - 1) / RW;
   const long long nw = (long long) nb * WPB;
   const long long per = (ng + nw - 1) / nw * RW;
   const long long r0 = ((long long) blk * WPB + wv) * per;
   R0 = (int) (r0 < n ? r0 : n);
   R1 = (int) (r0 + per < n ? r0 + per : n);
}

// ZR: also u = (0 + zw y) ./ zl1, the zero-guess l1-Jacobi sweep on y, from
// the value just formed (the bits of a separate sweep reading y back)
template <int RW, int SW, typename V, typename X, bool DOT, bool ZR = false>
__global__ __launch_bounds__(BS) void k_spmv_w(int n, const int *__restrict__ ia,
                                               const int *__restrict__ ja, const V *__restrict__ a,
                                               const X *__restrict__ x, X alpha, X beta,
                                               const X *b, X *y, double *__restrict__ part,
                                               const X *__restrict__ zl1 = nullptr, X zw = (X) 0,
                                               X *__restrict__ zu = nullptr)
{
   __shared__ X sbuf_all[WPB * SW];
   const int nb = gridDim.x;
   const int blk = xcd_block(blockIdx.x, nb);
   const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
   X *sbuf = sbuf_all + wv * SW;
   int R0, R1;
   wave_rows<RW>(n, blk, wv, nb, R0, R1);
   double acc = 0.0;
   rows_stream<RW, SW, V, X>(R0, R1, lane, ia, ja, a, x, sbuf, [&](int row, X t) {
      const X v = (beta == (X) 0) ? alpha * t : alpha * t + beta * b[row];
      y[row] = v;
      if (DOT) { acc += (double) v * (double) x[row]; }
      if (ZR) { zu[row] = (X) 0 + (zw * v) / zl1[row]; }
   });
   if (DOT) {
      __shared__ double sh[BS];
      sh[threadIdx.x] = acc;
      __syncthreads();
      for (int off = BS / 2; off > 0; off >>= 1) {
         if (threadIdx.x < off) { sh[threadIdx.x] += sh[threadIdx.x + off]; }
         __syncthreads();
      }
      if (threadIdx.x == 0) { part[blk] = sh[0]; }
   }
}

__global__ __launch_bounds__(BS) void k_reduce_parts(const double *__restrict__ part, int np,
                                                     double *__restrict__ out)
{
   __shared__ double sh[BS];
   double s = 0.0;
   for (int i = threadIdx.x; i < np; i += BS) { s += part[i]; }
   sh[threadIdx.x] = s;
   __syncthreads();
   for (int off = BS / 2; off > 0; off >>= 1) {
      if (threadIdx.x < off) { sh[threadIdx.x] += sh[threadIdx.x + off]; }
      __syncthreads();
   }
   if (threadIdx.x == 0) { *out = sh[0]; }
}

static double *g_part = nullptr;
static double *parts()
{
   if (!g_part) { g_part = dalloc_n<double>(8192 * 2); }
   return g_part;
}

static int grid_for_groups(int n, int rw)
{
   long long ng = ((long long) n + rw - 1) / rw;
   long long blocks = (ng + WPB - 1) / WPB;
   blocks = std::min<long long>(blocks, 8192);
   blocks = std::max<long long>(blocks, 1);
   if (blocks >= 64) { blocks = (blocks + 7) / 8 * 8; }
   return (int) blocks;
}

// ---------------------------------------------------------------- sliced ELL
// A wave owns a contiguous range of 64-row slices and takes them one at a
// time, lane = row.  Step k of a slice loads entry k of its 64 rows: one
// contiguous index load (128 bytes of 16-bit offsets, or 256 bytes of 32-bit
// columns) and one 512-byte value load per wave (non-temporal: the matrix is
// read once, x stays in L2), then the 64 x gathers, which for a stencil
// operator are neighbours of 64 consecutive rows.  U steps are issued
// together (index and value loads, then the gathers, then the adds), so each
// round trip to memory carries U steps.  Each lane adds its row's products in
// storage order, t = ((0 + a_0 x_0) + a_1 x_1) + ..., the order of hypre's
// host CSR matvec (src/seq_mv/csr_matvec.c:22) and of the windowed CSR
// kernels above, so both formats give the same bits.
#ifndef HX_ELL_U
#define HX_ELL_U 9
#endif
constexpr int ELL_U = HX_ELL_U;

struct EllView
{
   const int *sw, *mode, *jb, *j;
   const unsigned short *j16;
   const unsigned char *j8;
};

static EllView ell_view(const Ell &E) { return EllView{E.sw, E.mode, E.jb, E.j, E.j16, E.j8}; }

// IK: 0 = 32-bit columns, 1 = 16-bit offsets, 2 = 8-bit offsets
template <int U, int IK, typename V, typename X>
__device__ __forceinline__ X ell_steps(int r, int w, int len, const int *__restrict__ ejs,
                                       const unsigned short *__restrict__ e16s,
                                       const unsigned char *__restrict__ e8s,
                                       const int *__restrict__ jbs, const V *__restrict__ eas,
                                       const X *__restrict__ x)
{
   X t = 0;
   for (int k0 = 0; k0 < w; k0 += U) {
      int jj[U];
      V vv[U];
      X xx[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
         // steps past the slice width re-read its last step (in bounds, masked)
         const int ks = min(k0 + u, w - 1);
         if (IK == 2) {
            jj[u] = r + jbs[ks] + (int) __builtin_nontemporal_load(e8s + ks * 64);
         } else if (IK == 1) {
            jj[u] = r + jbs[ks] + (int) __builtin_nontemporal_load(e16s + ks * 64);
         } else {
            jj[u] = __builtin_nontemporal_load(ejs + ks * 64);
         }
         vv[u] = __builtin_nontemporal_load(eas + ks * 64);
      }
#pragma unroll
      for (int u = 0; u < U; u++) { xx[u] = x[(k0 + u < len) ? jj[u] : 0]; }
      // every step's product is formed and selected, so no load is moved
      // into a branch (and waited for on its own)
#pragma unroll
      for (int u = 0; u < U; u++) {
         const X s = t + (X) vv[u] * xx[u];
         t = (k0 + u < len) ? s : t;
      }
   }
   return t;
}

// pre(row) loads the row's own operands of the epilogue (right-hand side,
// l1 norm, ...) before the slice's matrix stream, so they arrive with it
// instead of costing one more round trip after it; epi(row, t, pre values)
template <int U, typename V, typename X, class Pre, class Epi>
__device__ __forceinline__ void ell_slice(int sl, int lane, int n, const int *__restrict__ ia,
                                          const EllView &E, const V *__restrict__ ea,
                                          const X *__restrict__ x, Pre pre, Epi epi)
{
   const int r = sl * 64 + lane;
   const int w0 = E.sw[sl], w = E.sw[sl + 1] - w0;
   const size_t base = (size_t) w0 * 64 + lane;
   int len = 0;
   if (r < n) { len = ia[r + 1] - ia[r]; }
   const auto pv = pre(min(r, n - 1));
   X t;
   const int m = E.mode[sl];
   if (m == 2) {
      t = ell_steps<U, 2, V, X>(r, w, len, nullptr, nullptr, E.j8 + base, E.jb + w0, ea + base, x);
   } else if (m == 1) {
      t = ell_steps<U, 1, V, X>(r, w, len, nullptr, E.j16 + base, nullptr, E.jb + w0, ea + base, x);
   } else {
      t = ell_steps<U, 0, V, X>(r, w, len, E.j + base, nullptr, nullptr, nullptr, ea + base, x);
   }
   if (r < n) { epi(r, t, pv); }
}

template <typename X>
struct RowOps
{
   X a, b, c;
};

// slices of wave wv of block blk: a contiguous range, ranges in block order
__device__ __forceinline__ void wave_slices(int ns, int blk, int wv, int nb, int &S0, int &S1)
{
   const long long nw = (long long) nb * WPB;
   const long long per = (ns + nw - 1) / nw;
   const long long s0 = ((long long) blk * WPB + wv) * per;
   S0 = (int) (s0 < ns ? s0 : ns);
   S1 = (int) (s0 + per < ns ? s0 + per : ns);
}

template <typename V, typename X, bool DOT, bool ZR = false>
__global__ __launch_bounds__(BS) void k_spmv_ell(int n, int ns, const int *__restrict__ ia,
                                                 EllView E,
                                                 const V *__restrict__ ea, const X *__restrict__ x,
                                                 X alpha, X beta, const X *b, X *y,
                                                 double *__restrict__ part,
                                                 const X *__restrict__ zl1 = nullptr,
                                                 X zw = (X) 0, X *__restrict__ zu = nullptr)
{
   const int nb = gridDim.x;
   const int blk = xcd_block(blockIdx.x, nb);
   const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
   int S0, S1;
   wave_slices(ns, blk, wv, nb, S0, S1);
   double acc = 0.0;
   for (int sl = S0; sl < S1; sl++) {
      ell_slice<ELL_U, V, X>(
         sl, lane, n, ia, E, ea, x,
         [&](int row) {
            RowOps<X> o{(X) 0, (X) 0, (X) 0};
            if (beta != (X) 0) { o.a = b[row]; }
            if (DOT) { o.b = x[row]; }
            if (ZR) { o.c = zl1[row]; }
            return o;
         },
         [&](int row, X t, const RowOps<X> &o) {
            const X v = (beta == (X) 0) ? alpha * t : alpha * t + beta * o.a;
            y[row] = v;
            if (DOT) { acc += (double) v * (double) o.b; }
            if (ZR) { zu[row] = (X) 0 + (zw * v) / o.c; }
         });
   }
   if (DOT) {
      __shared__ double sh[BS];
      sh[threadIdx.x] = acc;
      __syncthreads();
      for (int off = BS / 2; off > 0; off >>= 1) {
         if (threadIdx.x < off) { sh[threadIdx.x] += sh[threadIdx.x + off]; }
         __syncthreads();
      }
      if (threadIdx.x == 0) { part[blk] = sh[0]; }
   }
}

template <typename V, typename X>
__global__ __launch_bounds__(BS) void k_relax_l1_ell(int n, int ns, const int *__restrict__ ia,
                                                     EllView E,
                                                     const V *__restrict__ ea,
                                                     const X *__restrict__ f,
                                                     const X *__restrict__ u,
                                                     const X *__restrict__ l1, X w,
                                                     X *__restrict__ uo)
{
   const int nb = gridDim.x;
   const int blk = xcd_block(blockIdx.x, nb);
   const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
   int S0, S1;
   wave_slices(ns, blk, wv, nb, S0, S1);
   for (int sl = S0; sl < S1; sl++) {
      ell_slice<ELL_U, V, X>(
         sl, lane, n, ia, E, ea, u,
         [&](int row) { return RowOps<X>{f[row], u[row], l1[row]}; },
         [&](int row, X t, const RowOps<X> &o) {
            const X v = w * o.a - w * t;
            uo[row] = o.b + v / o.c;
         });
   }
}

template <typename V>
static const V *ell_values(const Ell &E)
{
   if constexpr (sizeof(V) == 8) {
      return E.a;
   } else {
      if (!E.a32) { fail(HYPRE_ERROR_GENERIC, "FP32 sliced-ELL values missing"); }
      return E.a32;
   }
}

// slice widths: the longest row of each 64-row slice
__global__ void k_ell_width(int n, int ns, const int *__restrict__ ia, int *__restrict__ wout)
{
   const int sl = blockIdx.x * blockDim.x + threadIdx.x;
   if (sl >= ns) { return; }
   const int r0 = sl * 64, r1 = min(n, r0 + 64);
   int w = 0;
   for (int r = r0; r < r1; r++) { w = max(w, ia[r + 1] - ia[r]); }
   wout[sl] = w;
}

// one wave per slice, lane = row: entry k of row r goes to sw[s]*64 + k*64 +
// lane; padding holds the row's own index and a zero value (never used).
// The slice's CSR entries are one contiguous range; it is staged through LDS
// in chunks with coalesced loads, and each lane then writes its row's entries
// of the chunk (for rows that start inside the chunk, entry j of every row is
// written by the same instruction: one contiguous 64-word step).
constexpr int EF_CH = 1024;
constexpr int EF_U = 4;

__global__ __launch_bounds__(64) void k_ell_fill(int n, int ns, const int *__restrict__ ia,
                                                 const int *__restrict__ ja,
                                                 const double *__restrict__ a,
                                                 const int *__restrict__ sw, int *__restrict__ ej,
                                                 double *__restrict__ ea)
{
   __shared__ int sj[EF_CH];
   __shared__ double sa[EF_CH];
   const int lane = threadIdx.x;
   for (int sl = blockIdx.x; sl < ns; sl += gridDim.x) {
      const int r = sl * 64 + lane;
      const int w0 = sw[sl], w = sw[sl + 1] - w0;
      int *ejs = ej + (size_t) w0 * 64 + lane;
      double *eas = ea + (size_t) w0 * 64 + lane;
      int rs = 0, re = 0;
      if (r < n) {
         rs = ia[r];
         re = ia[r + 1];
      }
      const int pad_col = (r < n) ? r : 0;
      for (int k = re - rs; k < w; k++) {
         ejs[(size_t) k * 64] = pad_col;
         eas[(size_t) k * 64] = 0.0;
      }
      const int s0 = ia[sl * 64], s1 = ia[min(n, sl * 64 + 64)];
      for (int c0 = s0; c0 < s1; c0 += EF_CH) {
         const int c1 = min(s1, c0 + EF_CH);
         for (int q0 = c0; q0 < c1; q0 += 64 * EF_U) {
            int jv[EF_U];
            double av[EF_U];
#pragma unroll
            for (int u = 0; u < EF_U; u++) {
               const int q = min(q0 + u * 64 + lane, c1 - 1);
               jv[u] = __builtin_nontemporal_load(ja + q);
               av[u] = __builtin_nontemporal_load(a + q);
            }
#pragma unroll
            for (int u = 0; u < EF_U; u++) {
               const int q = q0 + u * 64 + lane;
               if (q < c1) {
                  sj[q - c0] = jv[u];
                  sa[q - c0] = av[u];
               }
            }
         }
         __syncthreads();
         const int q1 = min(re, c1);
         for (int q = max(rs, c0); q < q1; q++) {
            const size_t k = (size_t) (q - rs) * 64;
            ejs[k] = sj[q - c0];
            eas[k] = sa[q - c0];
         }
         __syncthreads();
      }
   }
}

constexpr int EP_U = 9;

// per slice: the least and the largest col - row of each step over the rows
// that have an entry there; a slice whose steps all span at most 255 (65535)
// reads 8-bit (16-bit) offsets from its steps' least values, mode 2 (1)
__global__ __launch_bounds__(64) void k_ell_pack(int n, int ns, const int *__restrict__ ia,
                                                 const int *__restrict__ sw,
                                                 const int *__restrict__ ej,
                                                 unsigned short *__restrict__ e16,
                                                 unsigned char *__restrict__ e8,
                                                 int *__restrict__ jb, int *__restrict__ mode)
{
   const int lane = threadIdx.x;
   for (int sl = blockIdx.x; sl < ns; sl += gridDim.x) {
      const int r = sl * 64 + lane;
      const int w0 = sw[sl], w = sw[sl + 1] - w0;
      int len = 0;
      if (r < n) { len = ia[r + 1] - ia[r]; }
      const int *ejs = ej + (size_t) w0 * 64 + lane;
      unsigned short *e16s = e16 + (size_t) w0 * 64 + lane;
      unsigned char *e8s = e8 + (size_t) w0 * 64 + lane;
      long long span = 0;
      // the columns of EP_U steps are loaded before any is reduced (steps
      // past the slice width re-read its last step and are not used)
      for (int k0 = 0; k0 < w; k0 += EP_U) {
         int cv[EP_U];
#pragma unroll
         for (int u = 0; u < EP_U; u++) { cv[u] = ejs[(size_t) min(k0 + u, w - 1) * 64]; }
#pragma unroll
         for (int u = 0; u < EP_U; u++) {
            const int k = k0 + u;
            if (k >= w) { break; }   // uniform
            const bool in = k < len;
            const int off = cv[u] - r;   // both in [0, 2^31)
            int lo = in ? off : INT_MAX, hi = in ? off : INT_MIN;
            for (int d = 32; d > 0; d >>= 1) {
               lo = min(lo, __shfl_xor(lo, d, 64));
               hi = max(hi, __shfl_xor(hi, d, 64));
            }
            // some row of the slice has step k (w is its longest row)
            span = max(span, (long long) hi - (long long) lo);
            if (lane == 0) { jb[w0 + k] = lo; }
            const unsigned d = in ? (unsigned) (off - lo) : 0u;   // truncated where unused
            e16s[(size_t) k * 64] = (unsigned short) d;
            e8s[(size_t) k * 64] = (unsigned char) d;
         }
      }
      if (lane == 0) { mode[sl] = span <= 255 ? 2 : (span <= 65535 ? 1 : 0); }
   }
}

__global__ void k_d2f(int n, const double *__restrict__ x, float *__restrict__ y);
static int blas_grid(int n);

bool ell_plan(const Csr &A, Ell &E, double max_pad, int min_rows, bool fp32)
{
   E.free_all();
   if (A.nrows < std::max(min_rows, 1) || A.nnz <= 0 || A.ncols <= 0) { return false; }
   const int ns = (A.nrows + 63) / 64;
   int *w = dalloc_n<int>((size_t) ns + 1);
   k_ell_width<<<(ns + BS - 1) / BS, BS, 0, stream()>>>(A.nrows, ns, A.ia, w);
   HX_LAUNCH_CHECK();
   int wmax = 0;
   const long long tot = exclusive_scan(w, w, ns, &wmax);
   const long long stored = 64LL * tot;
   if ((double) stored > (1.0 + max_pad) * (double) A.nnz || stored > 0x7fffffffLL) {
      dfree(w);
      return false;
   }
   E.nslices = ns;
   E.stored = stored;
   E.wmax = wmax;
   E.sw = w;
   E.j = dalloc_n<int>((size_t) stored);
   E.a = dalloc_n<double>((size_t) stored);
   E.j16 = dalloc_n<unsigned short>((size_t) stored);
   E.j8 = dalloc_n<unsigned char>((size_t) stored);
   E.jb = dalloc_n<int>((size_t) (stored / 64) + 1);
   E.mode = dalloc_n<int>((size_t) ns);
   if (fp32) { E.a32 = dalloc_n<float>((size_t) stored); }
   return true;
}

// The copy is filled beside the setup's main stream: 1024 one-wave blocks (4
// per CU) take a small share of the CUs and of HBM bandwidth and still finish
// before the solve (measured at 200^3 27-point: setup 44.4 -> 43.3 ms against
// one block per slice, which slowed the bandwidth-bound kernels after RAP).
constexpr int EF_GRID = 1024;

void ell_fill(const Csr &A, Ell &E, bool fp32, hipStream_t s)
{
   const int ns = E.nslices;
   k_ell_fill<<<std::min(ns, EF_GRID), 64, 0, s>>>(A.nrows, ns, A.ia, A.ja, A.a, E.sw, E.j, E.a);
   HX_LAUNCH_CHECK();
   k_ell_pack<<<std::min(ns, EF_GRID), 64, 0, s>>>(A.nrows, ns, A.ia, E.sw, E.j, E.j16, E.j8, E.jb,
                                                   E.mode);
   HX_LAUNCH_CHECK();
   if (fp32) {
      k_d2f<<<blas_grid((int) E.stored), BS, 0, s>>>((int) E.stored, E.a, E.a32);
      HX_LAUNCH_CHECK();
   }
}


template <class F>
static void dispatch_rw(int rw, F f)
{
   switch (rw) {
      case 1: f(std::integral_constant<int, 1>{}); break;
      case 2: f(std::integral_constant<int, 2>{}); break;
      case 4: f(std::integral_constant<int, 4>{}); break;
      case 8: f(std::integral_constant<int, 8>{}); break;
      case 16: f(std::integral_constant<int, 16>{}); break;
      case 32: f(std::integral_constant<int, 32>{}); break;
      default: f(std::integral_constant<int, 64>{}); break;
   }
}

template <typename V, typename X, bool DOT, bool ZR = false>
static int launch_spmv(const Csr &A, const V *vals, X alpha, const X *x, X beta, const X *b,
                       X *y, double *part, const char *cls, double vbytes,
                       const X *zl1 = nullptr, X zw = (X) 0, X *zu = nullptr)
{
   if (A.nrows <= 0) { return 0; }
   // ZR also reads l1 and writes u on the rows of y
   double bytes = (double) A.nnz * (4.0 + vbytes) + 4.0 * (A.nrows + 1) +
                  (double) sizeof(X) * (A.ncols + A.nrows) + (beta != (X) 0 ? sizeof(X) * (double) A.nrows : 0.0) +
                  (ZR ? 2.0 * sizeof(X) * A.nrows : 0.0);
   if (A.ell) {
      const Ell &E = *A.ell;
      const int nb = grid_for_groups(A.nrows, 64);
      TimedLaunch tl(cls, bytes);
      k_spmv_ell<V, X, DOT, ZR><<<nb, BS, 0, stream()>>>(A.nrows, E.nslices, A.ia, ell_view(E),
                                                        ell_values<V>(E), x, alpha, beta, b, y,
                                                        part, zl1, zw, zu);
      HX_LAUNCH_CHECK();
      return nb;
   }
   const int rw = pick_rw(A.nnz, A.nrows);
   const int nb = grid_for_groups(A.nrows, rw);
   TimedLaunch tl(cls, bytes);
   hipStream_t s = stream();
   const int sw = spmv_window(A.nnz, A.nrows);
   dispatch_rw(rw, [&](auto c) {
      if (sw == 1024) {
         k_spmv_w<decltype(c)::value, 1024, V, X, DOT, ZR><<<nb, BS, 0, s>>>(
            A.nrows, A.ia, A.ja, vals, x, alpha, beta, b, y, part, zl1, zw, zu);
      } else if (sw == 256) {
         k_spmv_w<decltype(c)::value, 256, V, X, DOT, ZR><<<nb, BS, 0, s>>>(
            A.nrows, A.ia, A.ja, vals, x, alpha, beta, b, y, part, zl1, zw, zu);
      } else {
         k_spmv_w<decltype(c)::value, SW, V, X, DOT, ZR><<<nb, BS, 0, s>>>(
            A.nrows, A.ia, A.ja, vals, x, alpha, beta, b, y, part, zl1, zw, zu);
      }
   });
   HX_LAUNCH_CHECK();
   return nb;
}

void spmv(const Csr &A, double alpha, const double *x, double beta, const double *b, double *y,
          const char *cls)
{
   launch_spmv<double, double, false>(A, A.a, alpha, x, beta, b, y, nullptr, cls, 8.0);
}

void spmv32(const Csr &A, float alpha, const float *x, float beta, const float *b, float *y,
            const char *cls)
{
   if (!A.a32) { fail(HYPRE_ERROR_GENERIC, "FP32 values missing"); }
   launch_spmv<float, float, false>(A, A.a32, alpha, x, beta, b, y, nullptr, cls, 4.0);
}

void spmv_zero_relax(const Csr &A, const double *x, double *y, const double *l1, double w,
                     double *u, const char *cls)
{
   launch_spmv<double, double, false, true>(A, A.a, 1.0, x, 0.0, nullptr, y, nullptr, cls, 8.0, l1,
                                            w, u);
}

void spmv32_zero_relax(const Csr &A, const float *x, float *y, const float *l1, float w, float *u,
                       const char *cls)
{
   if (!A.a32) { fail(HYPRE_ERROR_GENERIC, "FP32 values missing"); }
   launch_spmv<float, float, false, true>(A, A.a32, 1.f, x, 0.f, nullptr, y, nullptr, cls, 4.0, l1,
                                          w, u);
}

void spmv_dot(const Csr &A, const double *p, double *s, double *dot_out_dev, const char *cls)
{
   const int nb = launch_spmv<double, double, true>(A, A.a, 1.0, p, 0.0, nullptr, s, parts(), cls, 8.0);
   k_reduce_parts<<<1, BS, 0, stream()>>>(parts(), std::max(nb, 0), dot_out_dev);
   HX_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- relaxation
// l1-Jacobi (src/parcsr_ls/par_relax.c:1178-1240): v = w*f - w*(A u);
// u_out = u + v/l1, in k_relax_l1_w / k_relax_l1_zero below.

// weighted Jacobi, relax type 0 (par_relax.c:180-320, Skip_diag = 1)
template <int K, typename V, typename X>
__global__ __launch_bounds__(BS) void k_relax_jac(int n, const int *__restrict__ ia,
                                                  const int *__restrict__ ja,
                                                  const V *__restrict__ a, const X *__restrict__ f,
                                                  const X *__restrict__ u, X w,
                                                  X *__restrict__ uo)
{
   constexpr int G = BS / K;
   const int nb = gridDim.x;
   const int blk = xcd_block(blockIdx.x, nb);
   const int rows_per_block = (n + nb - 1) / nb;
   const int r0 = blk * rows_per_block;
   const int r1 = min(n, r0 + rows_per_block);
   const int lane = threadIdx.x & (K - 1);
   const int g = threadIdx.x / K;
   for (int row = r0 + g; row < r1; row += G) {
      int s = ia[row], e = ia[row + 1];
      X t = row_dot<K, V, X>(s + 1, e, lane, ja, a, u);
      if (lane == 0) {
         X di = (X) a[s];
         if (di != (X) 0) {
            X res = f[row] - t;
            uo[row] = ((X) 1 - w) * u[row] + w * res / di;
         } else {
            uo[row] = u[row];
         }
      }
   }
}

// windowed l1-Jacobi sweep: t = (A u)_row in storage order, then the update
template <int RW, int SW, typename V, typename X>
__global__ __launch_bounds__(BS) void k_relax_l1_w(int n, const int *__restrict__ ia,
                                                   const int *__restrict__ ja,
                                                   const V *__restrict__ a, const X *__restrict__ f,
                                                   const X *__restrict__ u, const X *__restrict__ l1,
                                                   X w, X *__restrict__ uo)
{
   __shared__ X sbuf_all[WPB * SW];
   const int nb = gridDim.x;
   const int blk = xcd_block(blockIdx.x, nb);
   const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
   X *sbuf = sbuf_all + wv * SW;
   int R0, R1;
   wave_rows<RW>(n, blk, wv, nb, R0, R1);
   rows_stream<RW, SW, V, X>(R0, R1, lane, ia, ja, a, u, sbuf, [&](int row, X t) {
      const X v = w * f[row] - w * t;
      uo[row] = u[row] + v / l1[row];
   });
}

// zero initial guess: u_out = (0 + w f) / l1, one thread per row
template <typename X>
__global__ void k_relax_l1_zero(int n, const X *__restrict__ f, const X *__restrict__ l1, X w,
                                X *__restrict__ uo)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      uo[i] = (X) 0 + (w * f[i]) / l1[i];
   }
}

template <typename V, typename X>
static void launch_relax_l1(const Csr &A, const V *vals, const X *f, const X *u, const X *l1, X w,
                            X *uo, bool zero, const char *cls, double vbytes)
{
   if (A.nrows <= 0) { return; }
   double xs = sizeof(X);
   double bytes = zero ? xs * 3.0 * A.nrows
                       : (double) A.nnz * (4.0 + vbytes) + 4.0 * (A.nrows + 1) + xs * 4.0 * A.nrows;
   TimedLaunch tl(cls, bytes);
   hipStream_t s = stream();
   if (zero) {
      const int nb = std::max(1, std::min((A.nrows + BS - 1) / BS, 16384));
      k_relax_l1_zero<X><<<nb, BS, 0, s>>>(A.nrows, f, l1, w, uo);
   } else if (A.ell) {
      const Ell &E = *A.ell;
      const int nb = grid_for_groups(A.nrows, 64);
      k_relax_l1_ell<V, X><<<nb, BS, 0, s>>>(A.nrows, E.nslices, A.ia, ell_view(E),
                                             ell_values<V>(E), f, u, l1, w, uo);
   } else {
      const int rw = pick_rw(A.nnz, A.nrows);
      const int nb = grid_for_groups(A.nrows, rw);
      const int sw = spmv_window(A.nnz, A.nrows);
      dispatch_rw(rw, [&](auto c) {
         if (sw == 1024) {
            k_relax_l1_w<decltype(c)::value, 1024, V, X><<<nb, BS, 0, s>>>(A.nrows, A.ia, A.ja, vals,
                                                                          f, u, l1, w, uo);
         } else if (sw == 256) {
            k_relax_l1_w<decltype(c)::value, 256, V, X><<<nb, BS, 0, s>>>(A.nrows, A.ia, A.ja, vals,
                                                                         f, u, l1, w, uo);
         } else {
            k_relax_l1_w<decltype(c)::value, SW, V, X><<<nb, BS, 0, s>>>(A.nrows, A.ia, A.ja, vals, f,
                                                                        u, l1, w, uo);
         }
      });
   }
   HX_LAUNCH_CHECK();
}

void relax_l1(const Csr &A, const double *f, const double *u, const double *l1, double w,
              double *u_out, bool zero_guess, const char *cls)
{
   launch_relax_l1<double, double>(A, A.a, f, u, l1, w, u_out, zero_guess, cls, 8.0);
}

void relax_l1_32(const Csr &A, const float *f, const float *u, const float *l1, float w,
                 float *u_out, bool zero_guess, const char *cls)
{
   launch_relax_l1<float, float>(A, A.a32, f, u, l1, w, u_out, zero_guess, cls, 4.0);
}

template <typename V, typename X>
static void launch_relax_jac(const Csr &A, const V *vals, const X *f, const X *u, X w, X *uo)
{
   if (A.nrows <= 0) { return; }
   int K = pick_k(A.nnz, A.nrows);
   int nb = grid_for_rows(A.nrows, K);
   hipStream_t s = stream();
   switch (K) {
      case 4: k_relax_jac<4, V, X><<<nb, BS, 0, s>>>(A.nrows, A.ia, A.ja, vals, f, u, w, uo); break;
      case 8: k_relax_jac<8, V, X><<<nb, BS, 0, s>>>(A.nrows, A.ia, A.ja, vals, f, u, w, uo); break;
      case 16: k_relax_jac<16, V, X><<<nb, BS, 0, s>>>(A.nrows, A.ia, A.ja, vals, f, u, w, uo); break;
      case 32: k_relax_jac<32, V, X><<<nb, BS, 0, s>>>(A.nrows, A.ia, A.ja, vals, f, u, w, uo); break;
      default: k_relax_jac<64, V, X><<<nb, BS, 0, s>>>(A.nrows, A.ia, A.ja, vals, f, u, w, uo); break;
   }
   HX_LAUNCH_CHECK();
}

void relax_jacobi(const Csr &A, const double *f, const double *u, double w, double *u_out)
{
   launch_relax_jac<double, double>(A, A.a, f, u, w, u_out);
}

void relax_jacobi_32(const Csr &A, const float *f, const float *u, float w, float *u_out)
{
   launch_relax_jac<float, float>(A, A.a32, f, u, w, u_out);
}

// ---------------------------------------------------------------- hybrid GS
// src/parcsr_ls/par_relax.c:692-953 (hypre_BoomerAMGRelaxHybridGaussSeidel_core)
// with the row kernels of par_relax.h:13-455.  The reference's threads become
// `parts` contiguous row partitions (utilities.c:99 hypre_partition1D).  A
// group of G lanes owns a partition and walks its rows in sweep order: the
// group loads a row's entries with coalesced loads, gathers the values of
// the iterate before the sweep (the reference's Vtemp), forms the products
// that do not depend on this sweep and stages them in LDS; one lane then
// folds the row in storage order, taking the values this sweep already
// produced in the partition from a register (the previous row) or from LDS.
// A backward sweep runs forward over the rows of a row-reversed copy of the
// matrix, so both directions stream contiguous memory.
bool is_hybrid_gs(int t) { return t == 3 || t == 4 || t == 6 || t == 8 || t == 13 || t == 14; }

bool hgs_needs_reverse(int t) { return t == 4 || t == 6 || t == 8 || t == 14; }

int hgs_parts(int n, int parts, int rows)
{
   if (parts > 0) { return parts; }
   // partitions of `rows` rows, fewer rows (down to 8) on levels too small
   // to keep ~128K partitions in flight
   const long long want = ((long long) n + 131071) / 131072;
   const int r = (int) std::min<long long>(std::max(rows, 1), std::max<long long>(want, 8));
   if (n <= r) { return 1; }
   return (int) (((long long) n + r - 1) / r);
}

constexpr int HGS_EPL = 8;     // staged entries per lane and chunk (chunk = 8 G entries)
constexpr int HGS_CMAX = 16;   // partitions of up to this many rows keep new values in LDS

// group width: one chunk (8 G entries) covers a typical row; levels small
// enough to be latency-bound take a whole wave per partition when their
// longest rows would need several chunks
static int hgs_group(long long nnz, int n, int max_row)
{
   const double avg = n > 0 ? (double) nnz / n : 1.0;
   int G = (avg <= 28.0) ? 4 : 16;
   if (max_row > HGS_EPL * G && n <= 262144) { G = 64; }
   return G;
}

// categories of an entry (low 2 bits of the staged code; the partition-local
// row index above them): 0 other partition (product with the copy's value),
// 1 the row this sweep produced last, 2 an earlier row of this sweep, 3 a row
// of this partition not reached yet in this sweep (product with its value)
template <typename V, int G>
__device__ __forceinline__ void hgs_load_ja(int c0, int re, int gl, const int *__restrict__ ja,
                                            const V *__restrict__ a, int (&jv)[HGS_EPL],
                                            V (&av)[HGS_EPL])
{
#pragma unroll
   for (int t = 0; t < HGS_EPL; t++) {
      const int q = min(c0 + gl + t * G, re - 1);
      jv[t] = ja[q];
      av[t] = a[q];
   }
}

// values of the copy (xo) and, in the second sweep of a symmetric pair, the
// first sweep's values of this partition (x1)
template <typename X, bool ZERO, bool SECOND>
__device__ __forceinline__ void hgs_gather(const int (&jv)[HGS_EPL], int i, int n, int ks, int ke,
                                           bool rev, const X *__restrict__ uo, const X *un,
                                           X (&xo)[HGS_EPL], X (&x1)[HGS_EPL])
{
#pragma unroll
   for (int t = 0; t < HGS_EPL; t++) {
      const int j = jv[t];
      xo[t] = ZERO ? (X) 0 : uo[j];
      if (SECOND) {
         const int sj = rev ? n - 1 - j : j;
         const bool in = sj >= ks && sj < ke;
         x1[t] = un[in ? j : i];
      } else {
         x1[t] = xo[t];
      }
   }
}

template <typename V, typename X, bool NS, bool SECOND, bool PRE, int G>
__device__ __forceinline__ void hgs_stage(int kr, int n, int ks, int ke, bool rev, int gl,
                                          const int (&jv)[HGS_EPL], const V (&av)[HGS_EPL],
                                          const X (&xo)[HGS_EPL], const X (&x1)[HGS_EPL], X *v0,
                                          X *v1, int *cc, const X *su)
{
   // every slot of the chunk is written (entries past the row end hold values
   // the fold never reads), so the staging has no branches.  PRE: the rows of
   // this partition that the sweep already produced are in LDS, so every
   // product is formed here and the fold is a plain ordered sum.
#pragma unroll
   for (int t = 0; t < HGS_EPL; t++) {
      const int e = gl + t * G;
      const int j = jv[t];
      const int sj = rev ? n - 1 - j : j;
      const bool in = sj >= ks && sj < ke;
      const bool done = in && sj < kr;
      const X A_ = (X) av[t];
      if (PRE) {
         const X xv = done ? su[(sj - ks) & (HGS_CMAX - 1)] : (in ? x1[t] : xo[t]);
         v0[e] = A_ * xv;
         if (!NS) {
            v1[e] = A_ * xo[t];
            cc[e] = in ? 1 : 0;
         }
      } else {
         const int cat = !in ? 0 : (done ? (sj == kr - 1 ? 1 : 2) : 3);
         cc[e] = cat | ((in ? sj - ks : 0) << 2);
         v0[e] = done ? A_ : A_ * (in ? x1[t] : xo[t]);
         if (!NS) { v1[e] = A_ * xo[t]; }
      }
   }
}

// One term of a row's fold.  Without weights every term is res -= v * x with
// x = 1 for a staged product (exact: v * 1 == v), so that path has no branch.
template <typename X, bool NS>
__device__ __forceinline__ void hgs_fold_one(int code, X pv, X p1, X prev_val, const X *su,
                                             bool lds_u, int ks, int n, bool rev, const X *un,
                                             X &res, X &res0, X &res2)
{
   const int cat = code & 3;
   X xc = (X) 1;
   if (cat == 1) { xc = prev_val; }
   if (cat == 2) {
      if (lds_u) {
         xc = su[(code >> 2) & (HGS_CMAX - 1)];
      } else {
         const int sj = ks + (code >> 2);
         xc = un[rev ? n - 1 - sj : sj];
      }
   }
   if (NS) {
      res -= pv * xc;
   } else if (cat == 0) {
      res -= pv;
   } else {
      res0 -= pv * xc;
      res2 += p1;
   }
}

// MULTI: rows may be longer than one chunk (further chunks are loaded while
// folding); otherwise every row fits one chunk and the loads of the next two
// rows are kept in flight: row kr+2's columns and coefficients and row kr+1's
// gathered values are issued while row kr is folded.
template <typename V, typename X, bool NS, bool SKIP, bool ZERO, bool REV, bool SECOND, bool MULTI,
          bool PRE, int G>
__global__ __launch_bounds__(BS) void k_hgs(int n, int parts, const int *__restrict__ ia,
                                            const int *__restrict__ ja, const V *__restrict__ a,
                                            const X *__restrict__ f, const X *__restrict__ l1,
                                            const X *__restrict__ uo, X *un, X w, X omega, X prod,
                                            X omw)
{
   constexpr int NG = 64 / G;
   constexpr int CH = HGS_EPL * G;
   constexpr int NV1 = NS ? 2 : CH;
   constexpr int SK = SKIP ? 1 : 0;
   __shared__ __align__(16) X s_v0[WPB * NG * CH];
   __shared__ __align__(16) X s_v1[WPB * NG * NV1];
   __shared__ __align__(16) int s_c[WPB * NG * CH];
   __shared__ X s_u[WPB * NG * HGS_CMAX];
   const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
   const int g = lane / G, gl = lane % G;
   const int slot = wv * NG + g;
   X *v0 = s_v0 + slot * CH;
   X *v1 = s_v1 + slot * NV1;
   int *cc = s_c + slot * CH;
   X *su = s_u + slot * HGS_CMAX;
   // each XCD takes a contiguous range of partitions, so the rows its resident
   // groups walk (and their x neighbourhoods) stay close together in its L2
   const int blk = xcd_block(blockIdx.x, gridDim.x);
   const long long pl = ((long long) blk * WPB + wv) * NG + g;
   if (pl >= parts) { return; }
   const int p = (int) pl;
   int ns, ne;
   if (parts == 1) {
      ns = 0;
      ne = n;
   } else {
      const int size = n / parts, rest = n - size * parts;
      if (p < rest) {
         ns = p * (size + 1);
         ne = ns + size + 1;
      } else {
         ns = p * size + rest;
         ne = ns + size;
      }
   }
   if (ns >= ne) { return; }
   // rows in stream order: k = i forward, k = n-1-i over the reversed copy
   const int ks = REV ? n - ne : ns, ke = REV ? n - ns : ne;
   const bool lds_u = (ne - ns) <= HGS_CMAX;
   auto row_of = [&](int k) { return REV ? n - 1 - k : k; };
   auto bound = [&](int k) { return ia[min(k, ke)]; };
   X prev_val = (X) 0;
   // bounds b0..b3 of rows kr..kr+2; pipeline registers
   int b0 = bound(ks), b1 = bound(ks + 1), b2 = bound(ks + 2), b3 = bound(ks + 3);
   int j0[HGS_EPL], j1[HGS_EPL];
   V a0[HGS_EPL], a1[HGS_EPL];
   X xo0[HGS_EPL], x10[HGS_EPL];
   if (b0 + SK < b1) {
      hgs_load_ja<V, G>(b0 + SK, b1, gl, ja, a, j0, a0);
      hgs_gather<X, ZERO, SECOND>(j0, row_of(ks), n, ks, ke, REV, uo, un, xo0, x10);
   }
   if (!MULTI && ks + 1 < ke && b1 + SK < b2) { hgs_load_ja<V, G>(b1 + SK, b2, gl, ja, a, j1, a1); }
   // metadata of row kr (divisor, right-hand side, current value)
   auto meta = [&](int k, int rs, int re, X &d, X &fi, X &ci) {
      const int i = row_of(k);
      d = l1 ? l1[i] : (re > rs ? (X) a[rs] : (X) 0);
      fi = (X) 0;
      ci = (X) 0;
      if (gl == 0) {
         fi = f[i];
         ci = SECOND ? un[i] : (ZERO ? (X) 0 : uo[i]);
      }
   };
   X d0, f0, c0v;
   meta(ks, b0, b1, d0, f0, c0v);
   for (int kr = ks; kr < ke; kr++) {
      const int i = row_of(kr);
      const int rs = b0, re = b1;
      const bool has = rs + SK < re;
      if (has) {
         hgs_stage<V, X, NS, SECOND, PRE, G>(kr, n, ks, ke, REV, gl, j0, a0, xo0, x10, v0, v1, cc, su);
      }
      wave_fence();
      // issue the loads of the next rows
      const int b4 = (kr + 3 < ke) ? ia[kr + 4] : b3;
      X d1 = (X) 0, f1 = (X) 0, c1v = (X) 0;
      if (kr + 1 < ke) {
         meta(kr + 1, b1, b2, d1, f1, c1v);
         if (MULTI) {
            if (b1 + SK < b2) {
               hgs_load_ja<V, G>(b1 + SK, b2, gl, ja, a, j0, a0);
               hgs_gather<X, ZERO, SECOND>(j0, row_of(kr + 1), n, ks, ke, REV, uo, un, xo0, x10);
            }
         } else {
            if (b1 + SK < b2) {
               hgs_gather<X, ZERO, SECOND>(j1, row_of(kr + 1), n, ks, ke, REV, uo, un, xo0, x10);
#pragma unroll
               for (int t = 0; t < HGS_EPL; t++) {
                  j0[t] = j1[t];
                  a0[t] = a1[t];
               }
            }
            if (kr + 2 < ke && b2 + SK < b3) { hgs_load_ja<V, G>(b2 + SK, b3, gl, ja, a, j1, a1); }
         }
      }
      wave_fence();
      X nv = c0v;
      if (d0 != (X) 0) {
         X res = f0, res0 = (X) 0, res2 = (X) 0;
         for (int cb = rs + SK; cb < re; cb += CH) {
            if (MULTI && cb != rs + SK) {
               // rows longer than one chunk: the further chunks are loaded here
               int jw[HGS_EPL];
               V aw[HGS_EPL];
               X xw[HGS_EPL], xw1[HGS_EPL];
               wave_fence();
               hgs_load_ja<V, G>(cb, re, gl, ja, a, jw, aw);
               hgs_gather<X, ZERO, SECOND>(jw, i, n, ks, ke, REV, uo, un, xw, xw1);
               hgs_stage<V, X, NS, SECOND, PRE, G>(kr, n, ks, ke, REV, gl, jw, aw, xw, xw1, v0, v1, cc,
                                                   su);
               wave_fence();
            }
            if (gl == 0 && PRE) {
               const int ce = min(CH, re - cb);
               int e = 0;
               for (; e + 4 <= ce; e += 4) {
                  const X q0 = v0[e], q1 = v0[e + 1], q2 = v0[e + 2], q3 = v0[e + 3];
                  if (NS) {
                     res -= q0;
                     res -= q1;
                     res -= q2;
                     res -= q3;
                  } else {
                     const int4 c4 = *reinterpret_cast<const int4 *>(cc + e);
                     const X r0 = v1[e], r1 = v1[e + 1], r2 = v1[e + 2], r3 = v1[e + 3];
                     if (c4.x) { res0 -= q0; res2 += r0; } else { res -= q0; }
                     if (c4.y) { res0 -= q1; res2 += r1; } else { res -= q1; }
                     if (c4.z) { res0 -= q2; res2 += r2; } else { res -= q2; }
                     if (c4.w) { res0 -= q3; res2 += r3; } else { res -= q3; }
                  }
               }
               for (; e < ce; e++) {
                  if (NS || !cc[e]) { res -= v0[e]; }
                  else { res0 -= v0[e]; res2 += v1[e]; }
               }
            } else if (gl == 0) {
               const int ce = min(CH, re - cb);
               int e = 0;
               for (; e + 4 <= ce; e += 4) {
                  const int4 c4 = *reinterpret_cast<const int4 *>(cc + e);
                  const X q0 = v0[e], q1 = v0[e + 1], q2 = v0[e + 2], q3 = v0[e + 3];
                  X r0 = (X) 0, r1 = (X) 0, r2 = (X) 0, r3 = (X) 0;
                  if (!NS) { r0 = v1[e]; r1 = v1[e + 1]; r2 = v1[e + 2]; r3 = v1[e + 3]; }
                  hgs_fold_one<X, NS>(c4.x, q0, r0, prev_val, su, lds_u, ks, n, REV, un, res, res0, res2);
                  hgs_fold_one<X, NS>(c4.y, q1, r1, prev_val, su, lds_u, ks, n, REV, un, res, res0, res2);
                  hgs_fold_one<X, NS>(c4.z, q2, r2, prev_val, su, lds_u, ks, n, REV, un, res, res0, res2);
                  hgs_fold_one<X, NS>(c4.w, q3, r3, prev_val, su, lds_u, ks, n, REV, un, res, res0, res2);
               }
               for (; e < ce; e++) {
                  hgs_fold_one<X, NS>(cc[e], v0[e], NS ? (X) 0 : v1[e], prev_val, su, lds_u, ks, n, REV,
                                      un, res, res0, res2);
               }
            }
         }
         if (gl == 0) {
            if (NS) {
               nv = SKIP ? res / d0 : c0v + res / d0;
            } else {
               const X base = SKIP ? c0v * prod : c0v;
               nv = base + w * (omega * res + res0 + omw * res2) / d0;
            }
         }
      }
      if (gl == 0) {
         un[i] = nv;
         if (lds_u) { su[kr - ks] = nv; }
         prev_val = nv;
      }
      b0 = b1;
      b1 = b2;
      b2 = b3;
      b3 = b4;
      d0 = d1;
      f0 = f1;
      c0v = c1v;
      wave_fence();
   }
}

template <typename V, typename X, int G>
static void launch_hgs_sweep(bool multi, bool pre, bool ns, bool skip, bool zero, bool rev,
                             bool second, const Csr &M, const V *vals, const X *f, const X *l1,
                             const X *u, X *uo, int parts, X w, X omega, X prod, X omw)
{
   constexpr int NG = 64 / G;
   const long long waves = ((long long) parts + NG - 1) / NG;
   int nb = (int) ((waves + WPB - 1) / WPB);
   if (nb >= 64) { nb = (nb + 7) / 8 * 8; }     // xcd_block needs a multiple of 8
   hipStream_t s = stream();
   using T = std::true_type;
   using F = std::false_type;
   auto go = [&](auto nsc, auto skc, auto zc, auto rc, auto sc, auto mc) {
      auto k = [&](auto pc) {
         k_hgs<V, X, decltype(nsc)::value, decltype(skc)::value, decltype(zc)::value,
               decltype(rc)::value, decltype(sc)::value, decltype(mc)::value, decltype(pc)::value, G>
            <<<nb, BS, 0, s>>>(M.nrows, parts, M.ia, M.ja, vals, f, l1, u, uo, w, omega, prod, omw);
      };
      if (pre) { k(T{}); } else { k(F{}); }
   };
   auto l5 = [&](auto nsc, auto skc, auto zc, auto rc, auto sc) {
      if (multi) { go(nsc, skc, zc, rc, sc, T{}); } else { go(nsc, skc, zc, rc, sc, F{}); }
   };
   auto l4 = [&](auto nsc, auto skc, auto zc) {
      if (!rev) { l5(nsc, skc, zc, F{}, F{}); }
      else if (!second) { l5(nsc, skc, zc, T{}, F{}); }
      else { l5(nsc, skc, zc, T{}, T{}); }
   };
   auto l3 = [&](auto nsc, auto skc) {
      if (zero) { l4(nsc, skc, T{}); } else { l4(nsc, skc, F{}); }
   };
   // weights other than 1 always skip the diagonal (par_relax.c:1266-1383)
   if (ns) {
      if (skip) { l3(T{}, T{}); } else { l3(T{}, F{}); }
   } else {
      l3(F{}, T{});
   }
   HX_LAUNCH_CHECK();
}

template <typename V, typename X>
static void hgs_sweep(int max_row, bool ns, bool skip, bool zero, bool rev, bool second,
                      const Csr &M, const V *vals, const X *f, const X *l1, const X *u, X *uo,
                      int parts, X w, X omega, X prod, X omw)
{
   const int G = hgs_group(M.nnz, M.nrows, max_row);
   const bool multi = max_row - (skip ? 1 : 0) > HGS_EPL * G;
   // partitions hold at most ceil(n / parts) rows
   const bool pre = ((long long) M.nrows + parts - 1) / parts <= HGS_CMAX;
   if (!ns && !skip) { fail(HYPRE_ERROR_GENERIC, "weighted hybrid Gauss-Seidel keeps the diagonal"); }
   if (G == 4) {
      launch_hgs_sweep<V, X, 4>(multi, pre, ns, skip, zero, rev, second, M, vals, f, l1, u, uo,
                                parts, w, omega, prod, omw);
   } else if (G == 16) {
      launch_hgs_sweep<V, X, 16>(multi, pre, ns, skip, zero, rev, second, M, vals, f, l1, u, uo,
                                 parts, w, omega, prod, omw);
   } else {
      launch_hgs_sweep<V, X, 64>(multi, pre, ns, skip, zero, rev, second, M, vals, f, l1, u, uo,
                                 parts, w, omega, prod, omw);
   }
}

__global__ void k_row_len_max(int n, const int *__restrict__ ia, int *__restrict__ out)
{
   int m = 0;
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      m = max(m, ia[i + 1] - ia[i]);
   }
   for (int o = 32; o > 0; o >>= 1) { m = max(m, __shfl_xor(m, o, 64)); }
   if ((threadIdx.x & 63) == 0) { atomicMax(out, m); }
}

int csr_max_row(const Csr &A)
{
   if (A.nrows <= 0) { return 0; }
   int *d = device_flags() + 64;
   dzero(d, sizeof(int));
   const int nb = std::max(1, std::min((A.nrows + BS - 1) / BS, 2048));
   k_row_len_max<<<nb, BS, 0, stream()>>>(A.nrows, A.ia, d);
   HX_LAUNCH_CHECK();
   int h = 0;
   d2h(&h, d, sizeof(int));
   return h;
}

template <typename V, typename X>
static void launch_hgs(int type, const Csr &A, const Csr *Arev, int max_row, const V *vals,
                       const V *rvals, const X *f, const X *u, const X *l1, X w, X omega, X prod,
                       X omw, int parts, X *uo, bool zero, const char *cls)
{
   if (A.nrows <= 0) { return; }
   if (!is_hybrid_gs(type)) { fail(HYPRE_ERROR_ARG, "not a hybrid Gauss-Seidel type"); }
   const bool l1t = (type == 8 || type == 13 || type == 14);
   const bool ns = (w == (X) 1 && omega == (X) 1);
   // par_relax.c:946-989, 1160-1171 (3/4/6: Skip_diag 1, no l1 norms) and
   // 1266-1383 (8/13/14: Skip_diag 0 only without weights)
   const bool skip = l1t ? !ns : true;
   const int mode = (type == 6 || type == 8) ? 2 : (type == 4 || type == 14) ? 1 : 0;
   if (l1t && !l1) { fail(HYPRE_ERROR_GENERIC, "hybrid l1 Gauss-Seidel without l1 norms"); }
   if (mode != 0 && (!Arev || Arev->nrows != A.nrows || !rvals)) {
      fail(HYPRE_ERROR_GENERIC, "backward hybrid Gauss-Seidel without the row-reversed matrix");
   }
   parts = std::max(parts, 1);
   if (max_row < 0) { max_row = csr_max_row(A); }
   const double xs = sizeof(X);
   // algorithmic bytes: the matrix once, the copy u once (not with a zero
   // guess), f, l1 and the new u
   const double bytes = ((double) A.nnz * (4.0 + sizeof(V)) + 4.0 * (A.nrows + 1) +
                         xs * (zero ? 3.0 : 4.0) * A.nrows) *
                        (mode == 2 ? 2.0 : 1.0);
   TimedLaunch tl(cls, bytes);
   const X *l1p = l1t ? l1 : nullptr;
   if (mode == 0) {
      hgs_sweep<V, X>(max_row, ns, skip, zero, false, false, A, vals, f, l1p, u, uo, parts, w,
                      omega, prod, omw);
   } else if (mode == 1) {
      hgs_sweep<V, X>(max_row, ns, skip, zero, true, false, *Arev, rvals, f, l1p, u, uo, parts,
                      w, omega, prod, omw);
   } else {
      hgs_sweep<V, X>(max_row, ns, skip, zero, false, false, A, vals, f, l1p, u, uo, parts, w,
                      omega, prod, omw);
      hgs_sweep<V, X>(max_row, ns, skip, zero, true, true, *Arev, rvals, f, l1p, u, uo, parts,
                      w, omega, prod, omw);
   }
}

void relax_hgs(int type, const Csr &A, const Csr *Arev, int max_row, const double *f,
               const double *u, const double *l1, double w, double omega, int parts, double *u_out,
               bool zero_guess, const char *cls)
{
   // prod and 1 - omega as the reference forms them on the host
   launch_hgs<double, double>(type, A, Arev, max_row, A.a, Arev ? Arev->a : nullptr, f, u, l1, w,
                              omega,
                              1.0 - w * omega, 1.0 - omega, parts, u_out, zero_guess, cls);
}

void relax_hgs_32(int type, const Csr &A, const Csr *Arev, int max_row, const float *f,
                  const float *u, const float *l1, float w, float omega, int parts, float *u_out,
                  bool zero_guess, const char *cls)
{
   launch_hgs<float, float>(type, A, Arev, max_row, A.a32, Arev ? Arev->a32 : nullptr, f, u, l1, w,
                            omega,
                            1.f - w * omega, 1.f - omega, parts, u_out, zero_guess, cls);
}

// ---------------------------------------------------------------- BLAS-1
static int blas_grid(int n)
{
   int nb = (int) std::min<long long>(((long long) n + BS * 4 - 1) / (BS * 4), 2048);
   return std::max(nb, 1);
}

__global__ __launch_bounds__(BS) void k_dot_part(int n, const double *__restrict__ x,
                                                 const double *__restrict__ y,
                                                 double *__restrict__ part)
{
   __shared__ double sh[BS];
   double s = 0.0;
   for (long long i = (long long) blockIdx.x * BS + threadIdx.x; i < n; i += (long long) gridDim.x * BS) {
      s += x[i] * y[i];
   }
   sh[threadIdx.x] = s;
   __syncthreads();
   for (int off = BS / 2; off > 0; off >>= 1) {
      if (threadIdx.x < off) { sh[threadIdx.x] += sh[threadIdx.x + off]; }
      __syncthreads();
   }
   if (threadIdx.x == 0) { part[blockIdx.x] = sh[0]; }
}

__global__ __launch_bounds__(BS) void k_dot2_part(int n, const double *__restrict__ r,
                                                  const double *__restrict__ s,
                                                  double *__restrict__ part)
{
   __shared__ double sh[2][BS];
   double a = 0.0, b = 0.0;
   for (long long i = (long long) blockIdx.x * BS + threadIdx.x; i < n; i += (long long) gridDim.x * BS) {
      double ri = r[i];
      a += ri * s[i];
      b += ri * ri;
   }
   sh[0][threadIdx.x] = a;
   sh[1][threadIdx.x] = b;
   __syncthreads();
   for (int off = BS / 2; off > 0; off >>= 1) {
      if (threadIdx.x < off) {
         sh[0][threadIdx.x] += sh[0][threadIdx.x + off];
         sh[1][threadIdx.x] += sh[1][threadIdx.x + off];
      }
      __syncthreads();
   }
   if (threadIdx.x == 0) {
      part[blockIdx.x] = sh[0][0];
      part[8192 + blockIdx.x] = sh[1][0];
   }
}

void dot_dev(int n, const double *x, const double *y, double *out_dev)
{
   int nb = blas_grid(n);
   TimedLaunch tl("blas", 16.0 * n);
   k_dot_part<<<nb, BS, 0, stream()>>>(n, x, y, parts());
   k_reduce_parts<<<1, BS, 0, stream()>>>(parts(), nb, out_dev);
   HX_LAUNCH_CHECK();
}

void dot2_dev(int n, const double *r, const double *s, double *out2_dev)
{
   int nb = blas_grid(n);
   TimedLaunch tl("blas", 16.0 * n);
   k_dot2_part<<<nb, BS, 0, stream()>>>(n, r, s, parts());
   k_reduce_parts<<<1, BS, 0, stream()>>>(parts(), nb, out2_dev);
   k_reduce_parts<<<1, BS, 0, stream()>>>(parts() + 8192, nb, out2_dev + 1);
   HX_LAUNCH_CHECK();
}

double dot(int n, const double *x, const double *y)
{
   double *d = reinterpret_cast<double *>(device_flags() + 32);
   dot_dev(n, x, y, d);
   double *h = static_cast<double *>(pinned_scratch(64));
   HX_HIP(hipMemcpyAsync(h, d, sizeof(double), hipMemcpyDeviceToHost, stream()));
   sync();
   return h[0];
}

__global__ void k_axpy(int n, double alpha, const double *__restrict__ x, double *__restrict__ y)
{
   for (long long i = (long long) blockIdx.x * BS + threadIdx.x; i < n; i += (long long) gridDim.x * BS) {
      y[i] += alpha * x[i];
   }
}
__global__ void k_scale(int n, double alpha, double *__restrict__ x)
{
   for (long long i = (long long) blockIdx.x * BS + threadIdx.x; i < n; i += (long long) gridDim.x * BS) {
      x[i] *= alpha;
   }
}
template <typename T>
__global__ void k_fill(int n, T v, T *__restrict__ x)
{
   for (long long i = (long long) blockIdx.x * BS + threadIdx.x; i < n; i += (long long) gridDim.x * BS) {
      x[i] = v;
   }
}
__global__ void k_copy(int n, const double *__restrict__ x, double *__restrict__ y)
{
   for (long long i = (long long) blockIdx.x * BS + threadIdx.x; i < n; i += (long long) gridDim.x * BS) {
      y[i] = x[i];
   }
}
// x += alpha p ; r += (-alpha) s   (two hypre_ParKrylovAxpy calls, pcg.c)
__global__ void k_pcg_xr(int n, const double *__restrict__ ad, const double *__restrict__ p,
                         const double *__restrict__ s, double *__restrict__ x,
                         double *__restrict__ r)
{
   double alpha = *ad, malpha = -alpha;
   for (long long i = (long long) blockIdx.x * BS + threadIdx.x; i < n; i += (long long) gridDim.x * BS) {
      x[i] += alpha * p[i];
      r[i] += malpha * s[i];
   }
}
// p = beta p ; p += 1.0 s   (ScaleVector + Axpy, pcg.c)
__global__ void k_pcg_p(int n, const double *__restrict__ bd, const double *__restrict__ s,
                        double *__restrict__ p)
{
   double beta = *bd;
   for (long long i = (long long) blockIdx.x * BS + threadIdx.x; i < n; i += (long long) gridDim.x * BS) {
      double t = beta * p[i];
      p[i] = t + s[i];
   }
}
__global__ void k_d2f(int n, const double *__restrict__ x, float *__restrict__ y)
{
   for (long long i = (long long) blockIdx.x * BS + threadIdx.x; i < n; i += (long long) gridDim.x * BS) {
      y[i] = (float) x[i];
   }
}
__global__ void k_f2d(int n, const float *__restrict__ x, double *__restrict__ y)
{
   for (long long i = (long long) blockIdx.x * BS + threadIdx.x; i < n; i += (long long) gridDim.x * BS) {
      y[i] = (double) x[i];
   }
}
__global__ void k_sdiv(const double *num, const double *den, double *out) { *out = *num / *den; }

void axpy(int n, double alpha, const double *x, double *y)
{
   if (n <= 0) { return; }
   k_axpy<<<blas_grid(n), BS, 0, stream()>>>(n, alpha, x, y);
   HX_LAUNCH_CHECK();
}
void scale(int n, double alpha, double *x)
{
   if (n <= 0) { return; }
   k_scale<<<blas_grid(n), BS, 0, stream()>>>(n, alpha, x);
   HX_LAUNCH_CHECK();
}
void fill(int n, double v, double *x)
{
   if (n <= 0) { return; }
   k_fill<double><<<blas_grid(n), BS, 0, stream()>>>(n, v, x);
   HX_LAUNCH_CHECK();
}
void fill32(int n, float v, float *x)
{
   if (n <= 0) { return; }
   k_fill<float><<<blas_grid(n), BS, 0, stream()>>>(n, v, x);
   HX_LAUNCH_CHECK();
}
void copy(int n, const double *x, double *y)
{
   if (n <= 0) { return; }
   k_copy<<<blas_grid(n), BS, 0, stream()>>>(n, x, y);
   HX_LAUNCH_CHECK();
}
void pcg_update_xr(int n, const double *alpha_dev, const double *p, const double *s, double *x,
                   double *r)
{
   TimedLaunch tl("blas", 40.0 * n);
   k_pcg_xr<<<blas_grid(n), BS, 0, stream()>>>(n, alpha_dev, p, s, x, r);
   HX_LAUNCH_CHECK();
}
void pcg_update_p(int n, const double *beta_dev, const double *s, double *p)
{
   TimedLaunch tl("blas", 24.0 * n);
   k_pcg_p<<<blas_grid(n), BS, 0, stream()>>>(n, beta_dev, s, p);
   HX_LAUNCH_CHECK();
}
void d2f(int n, const double *x, float *y)
{
   if (n <= 0) { return; }
   k_d2f<<<blas_grid(n), BS, 0, stream()>>>(n, x, y);
   HX_LAUNCH_CHECK();
}
void f2d(int n, const float *x, double *y)
{
   if (n <= 0) { return; }
   k_f2d<<<blas_grid(n), BS, 0, stream()>>>(n, x, y);
   HX_LAUNCH_CHECK();
}
void scalar_div(const double *num, const double *den, double *out)
{
   k_sdiv<<<1, 1, 0, stream()>>>(num, den, out);
   HX_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- coarsest solve
// hypre_gselim (src/utilities/gselim.h) on the column-major matrix of
// par_gauss_elim.c:312-316, one block: the updates of one elimination step
// are independent element-wise and are spread over the threads; each element
// sees the same operation sequence as the serial loop.
constexpr int GE_MAX = 64;

template <typename X>
__global__ __launch_bounds__(BS) void k_gselim(int n, const double *__restrict__ Am,
                                               const X *__restrict__ f, X *__restrict__ u)
{
   __shared__ double A[GE_MAX * GE_MAX];
   __shared__ double x[GE_MAX];
   for (int i = threadIdx.x; i < n * n; i += BS) { A[i] = Am[i]; }
   for (int i = threadIdx.x; i < n; i += BS) { x[i] = (double) f[i]; }
   __syncthreads();
   if (n == 1) {
      if (threadIdx.x == 0) {
         if (A[0] != 0.0) { x[0] = x[0] / A[0]; }
         u[0] = (X) x[0];
      }
      return;
   }
   for (int k = 0; k < n - 1; k++) {
      double akk = A[k * n + k];
      if (akk != 0.0) {
         double divA = 1.0 / akk;
         int cnt = (n - k - 1) * (n - k);     // rows j>k, columns m in (k, n) plus the rhs
         for (int t = threadIdx.x; t < cnt; t += BS) {
            int j = k + 1 + t / (n - k);
            int m = k + 1 + t % (n - k);      // m == n encodes the right-hand side
            double ajk = A[j * n + k];
            if (ajk != 0.0) {
               double factor = ajk * divA;
               if (m < n) { A[j * n + m] -= factor * A[k * n + m]; }
               else { x[j] -= factor * x[k]; }
            }
         }
      }
      __syncthreads();
   }
   for (int k = n - 1; k > 0; --k) {
      __syncthreads();
      double akk = A[k * n + k];
      if (akk != 0.0) {
         if (threadIdx.x == 0) { x[k] /= akk; }
         __syncthreads();
         double xk = x[k];
         for (int j = threadIdx.x; j < k; j += BS) {
            if (A[j * n + k] != 0.0) { x[j] -= xk * A[j * n + k]; }
         }
      }
   }
   __syncthreads();
   if (threadIdx.x == 0) {
      if (A[0] != 0.0) { x[0] /= A[0]; }
   }
   __syncthreads();
   for (int i = threadIdx.x; i < n; i += BS) { u[i] = (X) x[i]; }
}

void gselim_solve(int n, const double *Amat_dev, const double *f, double *u)
{
   if (n > GE_MAX) { fail(HYPRE_ERROR_GENERIC, "coarse grid too large for the dense solve"); }
   k_gselim<double><<<1, BS, 0, stream()>>>(n, Amat_dev, f, u);
   HX_LAUNCH_CHECK();
}

void gselim_solve32(int n, const double *Amat_dev, const float *f, float *u)
{
   if (n > GE_MAX) { fail(HYPRE_ERROR_GENERIC, "coarse grid too large for the dense solve"); }
   k_gselim<float><<<1, BS, 0, stream()>>>(n, Amat_dev, f, u);
   HX_LAUNCH_CHECK();
}

} // namespace hx
