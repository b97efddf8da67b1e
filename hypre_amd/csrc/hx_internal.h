// Internal declarations of libhypre_mi355x.so (MI355X / gfx950, single rank).
//
// Everything the solver touches lives in HBM: CSR operators with int32 row
// pointers / columns and FP64 values (optional FP32 value copies for the
// mixed-precision V-cycle), FP64 level vectors.  One library stream orders all
// work; host code only sees scalars (dot products, level sizes, counters).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

extern "C" {
#include "HYPRE.h"
#include "hypre_mi355x.h"
}

namespace hx {

// ------------------------------------------------------------------ errors
// hypre's sticky global error bitmask (src/utilities/error.c).
extern int g_error_flag;
extern int g_error_arg;

struct Error : public std::runtime_error
{
   int code;
   Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_error(int code, const char *msg);
[[noreturn]] void fail(int code, const std::string &msg);

#define HX_HIP(call)                                                                   \
   do {                                                                                \
      hipError_t _e = (call);                                                          \
      if (_e != hipSuccess) {                                                          \
         ::hx::fail(HYPRE_ERROR_GENERIC, std::string("HIP error ") +                   \
                    hipGetErrorString(_e) + " at " __FILE__ ":" + std::to_string(__LINE__)); \
      }                                                                                \
   } while (0)

#define HX_LAUNCH_CHECK() HX_HIP(hipGetLastError())

// API wrapper: runs body, converts exceptions into the error flag.
template <class F>
inline int guard(F &&f)
{
   try {
      f();
   } catch (const Error &e) {
      set_error(e.code, e.what());
   } catch (const std::bad_alloc &) {
      set_error(HYPRE_ERROR_MEMORY, "host allocation failed");
   } catch (const std::exception &e) {
      set_error(HYPRE_ERROR_GENERIC, e.what());
   }
   return g_error_flag;
}

// ------------------------------------------------------------------ runtime
void        ensure_init();
// takes the address of the fixed scratch array of the allocation-free entry
// points on the current device (called once per device by ensure_init)
void        noalloc_attach_device();
hipStream_t stream();
void        sync();
hipStream_t side_stream();
void side_after_main();
void main_after_side();
void side_sync();                       // hipStreamSynchronize(stream())
void       *dalloc(size_t bytes);         // caching device allocator
void        dfree(void *p);
void        release_cache();
void       *pinned_scratch(size_t bytes); // host-pinned staging for scalars
int        *device_flags();               // 1024 ints of device status words (512..767: row claim counters)

template <class T>
inline T *dalloc_n(size_t n) { return static_cast<T *>(dalloc(n * sizeof(T) + 16)); }

void h2d(void *d, const void *h, size_t bytes);
void d2h(void *h, const void *d, size_t bytes);
void d2d(void *dst, const void *src, size_t bytes);
void dzero(void *d, size_t bytes);

// exclusive scan of n ints -> out[0..n], out[n] = total (in may alias out);
// waits for the total, and with max_in also returns the largest input
long long exclusive_scan(const int *in, int *out, int n, int *max_in = nullptr);
// the same scan without waiting (the caller knows or does not need the total)
void exclusive_scan_nowait(const int *in, int *out, int n);
// reductions returning to host
long long reduce_sum_int(const int *in, int n);
int       reduce_max_int(const int *in, int n);

// ------------------------------------------------------------------ timing
bool timing_on();
struct TimedLaunch
{
   int   slot = -1;
   TimedLaunch(const char *cls, double bytes);
   ~TimedLaunch();
};

// Profiling ranges named as the reference's GPU profiling ranges
// (src/utilities/device_markers.c), emitted through roctx while enabled
// (HYPRE_MI355X_SetProfilingRanges or HYPRE_MI355X_ROCTX=1 in the
// environment); closing a range waits for the library stream, as the
// reference's StreamSync0, so that a range covers its kernels.  When disabled
// a range costs one branch.
bool ranges_on();
void range_push(const char *name);
void range_pop();
struct Range
{
   bool on;
   explicit Range(const char *name) : on(ranges_on())
   {
      if (on) { range_push(name); }
   }
   Range(const char *fmt, int v);   // "AMG Level-%d" and the like
   ~Range()
   {
      if (on) { range_pop(); }
   }
   Range(const Range &) = delete;
   Range &operator=(const Range &) = delete;
};

// ------------------------------------------------------------------ CSR
// Sliced-ELL copy of an operator for the solve phase: slices of 64
// consecutive rows; a slice of width w (its longest row) stores 64*w entries
// column-major (entry k of its 64 rows are 64 consecutive words), each row's
// entries in CSR storage order, padding masked by the CSR row length.
// Columns of a slice whose every step k spans at most 255 (at most 65535)
// in col - row are kept as 8-bit (16-bit) offsets: col = row + jb[k] + off,
// jb[k] the step's least col - row (stencil operators: 8-bit in slices of
// interior rows, 16-bit in most boundary slices); the other slices read the
// 32-bit columns.
struct Ell
{
   int       nslices = 0;
   long long stored = 0;          // 64 * sum of slice widths
   int      *sw = nullptr;        // exclusive prefix of slice widths (nslices + 1)
   int      *j = nullptr;         // 32-bit columns (all slices)
   unsigned short *j16 = nullptr; // 16-bit column offsets (slices with mode 1)
   unsigned char  *j8 = nullptr;  // 8-bit column offsets (slices with mode 2)
   int      *jb = nullptr;        // per slice step: least col - row (sum of widths)
   int      *mode = nullptr;      // per slice: 2 = 8-bit offsets, 1 = 16-bit, 0 = 32-bit columns
   int       wmax = 0;            // widest slice: the operator's longest row
   double   *a = nullptr;
   float    *a32 = nullptr;
   void      free_all();
};

struct Csr
{
   int       nrows = 0, ncols = 0;
   long long nnz = 0;
   int      *ia = nullptr, *ja = nullptr;
   double   *a = nullptr;
   float    *a32 = nullptr;
   // not owned: when set, SpMV and l1-Jacobi run on this copy (same results)
   const Ell *ell = nullptr;
   void      free_all();
   void      alloc(int n, int nc, long long nz, bool values = true);
};

// The sliced-ELL copy E of A, built when the padding stays within max_pad
// (stored entries <= (1 + max_pad) nnz) and A has at least min_rows rows
// (fp32: also the FP32 values), in two parts: ell_plan sizes and allocates
// the copy on the library stream (false: no copy for this operator),
// ell_fill launches the copy on s
bool ell_plan(const Csr &A, Ell &E, double max_pad, int min_rows, bool fp32);
void ell_fill(const Csr &A, Ell &E, bool fp32, hipStream_t s);

void csr_copy(const Csr &src, Csr &dst);
void csr_to_host(const Csr &A, std::vector<int> &ia, std::vector<int> &ja, std::vector<double> &a);
void csr_from_host(Csr &A, int n, int ncols, const int *ia, const int *ja, const double *a);

// ------------------------------------------------------------------ solve kernels
// y = alpha*A*x + beta*b  (b may alias y; b ignored when beta == 0)
void spmv(const Csr &A, double alpha, const double *x, double beta, const double *b, double *y,
          const char *cls = "spmv");
void spmv32(const Csr &A, float alpha, const float *x, float beta, const float *b, float *y,
            const char *cls = "spmv");
// y = A x and u = (0 + w y)./l1 (the zero-guess l1-Jacobi sweep on y) in one pass
void spmv_zero_relax(const Csr &A, const double *x, double *y, const double *l1, double w,
                     double *u, const char *cls);
void spmv32_zero_relax(const Csr &A, const float *x, float *y, const float *l1, float w, float *u,
                       const char *cls);
// s = A p and part[] = per-block partial sums of s.*p (deterministic)
void spmv_dot(const Csr &A, const double *p, double *s, double *dot_out_dev, const char *cls);
// l1-Jacobi: u_out = u + (w*f - w*A*u)./l1 ; zero_guess: u_out = (w*f)./l1
void relax_l1(const Csr &A, const double *f, const double *u, const double *l1, double w,
              double *u_out, bool zero_guess, const char *cls);
void relax_l1_32(const Csr &A, const float *f, const float *u, const float *l1, float w,
                 float *u_out, bool zero_guess, const char *cls);
// weighted Jacobi (relax type 0): u_out = (1-w) u + w (f - sum_{j!=i} a_ij u_j)/a_ii
void relax_jacobi(const Csr &A, const double *f, const double *u, double w, double *u_out);
void relax_jacobi_32(const Csr &A, const float *f, const float *u, float w, float *u_out);
// hybrid Gauss-Seidel, relax types 3/4/6 (l1 == nullptr: divide by the
// diagonal) and 8/13/14 (l1 = option 4 norms): `parts` row partitions
// (hypre_partition1D), Gauss-Seidel inside one, the pre-sweep u across them.
// Out of place: u is not modified; zero_guess reads u as zeros.
// Backward and symmetric types (4/6/8/14) also need Arev, the row-reversed
// copy of A (csr_reverse_rows; a32 filled for the FP32 variant).
bool is_hybrid_gs(int type);
bool hgs_needs_reverse(int type);
// max_row: the longest row of A (negative: measured here)
int  hgs_parts(int n, int parts, int rows);
int  csr_max_row(const Csr &A);
void relax_hgs(int type, const Csr &A, const Csr *Arev, int max_row, const double *f,
               const double *u, const double *l1, double w, double omega, int parts, double *u_out,
               bool zero_guess, const char *cls);
void relax_hgs_32(int type, const Csr &A, const Csr *Arev, int max_row, const float *f,
                  const float *u, const float *l1, float w, float omega, int parts, float *u_out,
                  bool zero_guess, const char *cls);
// BLAS-1
double dot(int n, const double *x, const double *y);
void   dot_dev(int n, const double *x, const double *y, double *out_dev);
void   dot2_dev(int n, const double *r, const double *s, double *out2_dev); // <r,s>, <r,r>
void   axpy(int n, double alpha, const double *x, double *y);               // y += alpha x
void   scale(int n, double alpha, double *x);
void   fill(int n, double v, double *x);
void   fill32(int n, float v, float *x);
void   copy(int n, const double *x, double *y);
void   pcg_update_xr(int n, const double *alpha_dev, const double *p, const double *s, double *x,
                     double *r);                                            // x += a p; r -= a s
void   pcg_update_p(int n, const double *beta_dev, const double *s, double *p); // p = beta p + s
void   d2f(int n, const double *x, float *y);
void   f2d(int n, const float *x, double *y);
void   scalar_div(const double *num, const double *den, double *out);       // device scalars
// coarsest dense solve, column-major matrix (relax type 9)
void   gselim_solve(int n, const double *Amat_dev, const double *f, double *u);
void   gselim_solve32(int n, const double *Amat_dev, const float *f, float *u);

// ------------------------------------------------------------------ setup kernels
// strength pattern (S.a unused), sequential row order kept
// S of A; optionally also the column counts of S (st_count, n ints), the
// l1 norms of A's rows (l1, n doubles; *l1_zero set to 1 on a zero norm) and
// each row's strength threshold and flags (thr_out, flg_out: n each; bit 0
// negative diagonal, bit 1 all entries weak) for extpi_interp
// column_runs: A's rows hold runs of consecutive columns (the fine-grid
// operator of a stencil in its natural order; a hint for the column counts)
void create_s(const Csr &A, double theta, double max_row_sum, Csr &S, int *st_count = nullptr,
              double *l1 = nullptr, int *l1_zero = nullptr, double *thr_out = nullptr,
              int *flg_out = nullptr, bool column_runs = false);
// l1 norms option 1 of hypre_ParCSRComputeL1Norms
void l1_norms(const Csr &A, double *l1);
// option 4 on one rank (signed diagonal), for the l1 hybrid Gauss-Seidel types
void l1_norms_diag(const Csr &A, double *l1);
// R = A with its rows in reverse order (entries of a row kept in order);
// a32 copied when A has it
void csr_reverse_rows(const Csr &A, Csr &R);
// PMIS: measure (optional input: given measure on device, else Park-Miller
// stream seed 2747), CF out; returns number of rounds
// st_count: the column counts of S when already known (else counted here)
// cf_init 1 (HMIS): CF holds the Ruge-Stueben first pass on entry; cf_init 3
// (second coarsening of aggressive levels): rows without strong connections
// become C points and the first round picks no independent set
int  pmis(const Csr &S, const double *measure_in_dev, int *CF, double *measure_keep_dev,
          const int *st_count = nullptr, int cf_init = 0);
// HMIS (coarsen_type 10): the Ruge-Stueben first pass on the host, then PMIS
// with CF_init 1 on the device; returns the PMIS rounds
int  hmis(const Csr &S, const double *measure_in_dev, int *CF, double *measure_keep_dev,
          const int *st_count = nullptr);
// the first pass alone (host arrays; CF: 1 C, -1 F, -2 Z, -3 SF)
void rs_first_pass_host(int n, const int *S_ia, const int *S_ja, int *CF);
// fine->coarse map; returns n_coarse
int  coarse_map(int n, const int *CF, int *f2c);
// ext+i interpolation + truncation (CF: -3 -> -1 on exit)
// s_thr / s_flg: the strength thresholds and flags create_s produced for S
// (optional; they let the C lists be built in one pass over A)
void extpi_interp(const Csr &A, const Csr &S, int *CF, const int *f2c, int nc,
                  double trunc_factor, int max_elmts, Csr &P, const double *s_thr = nullptr,
                  const int *s_flg = nullptr);
// hypre_BoomerAMGInterpTruncation of P in place (rows of <= 64 entries)
void truncate_rows(Csr &P, double trunc_factor, int max_elmts);
// aggressive coarsening: the second strength matrix over the nc C points of
// CF (f2c their coarse indices), distance one or two in S, first-seen order
void create_2nd_s(const Csr &S, const int *CF, const int *f2c, int nc, Csr &S2);
// C points of CF take the second split CFN (f2c1: coarse index of the first split)
void correct_cf_marker(int n, int *CF, const int *f2c1, const int *CFN);
// modified multipass interpolation (agg_interp_type 8) for nc C points,
// truncated by (trunc_factor, max_elmts) when either is set; CF -3 -> -1
void mod_multipass_interp(const Csr &A, const Csr &S, int *CF, int nc, double trunc_factor,
                          int max_elmts, Csr &P);
void transpose(const Csr &P, Csr &R);
// p_row_bound / a_row_bound: an upper bound of P's / A's row lengths when the
// caller knows one (truncation to P_max_elmts; the previous product's
// ac_max_row), else <= 0; ac_max_row, when given, receives Ac's longest row
void rap(const Csr &R, const Csr &A, const Csr &P, Csr &Ac, int p_row_bound = 0,
         int a_row_bound = 0, int *ac_max_row = nullptr);
// dense column-major copy of a small CSR matrix (n <= 64)
void csr_to_dense_colmajor(const Csr &A, double *M_dev);

// Park-Miller helpers shared with the generators
void pm_uniform(int n, int seed, double *out_dev);   // hypre_Rand() stream

} // namespace hx
