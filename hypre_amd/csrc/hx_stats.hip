// BoomerAMG setup statistics: src/parcsr_ls/par_stats.c:20-1340
// (hypre_BoomerAMGSetupStats), printed after setup with print level 1 or 3
// (par_amg_setup.c:3830-3833), for one rank and the options this library
// runs: PMIS or HMIS coarsening, ext+i interpolation with optional aggressive
// levels (modified multipass), relax_order 0 and no grid relax points.
//
// The per-level figures come from one device pass over each operator and
// interpolation matrix: every row is summed by one thread in storage order
// (the reference's row loop), and the minima and maxima are reduced per block
// and then on the host, keeping the reference's hypre_min / hypre_max operand
// order.  Not on the timed path: nothing here runs unless the print level
// asks for it.
#include "hx_internal.h"
#include "hx_objects.h"

#include <algorithm>
#include <cstdio>
#include <vector>

namespace hx {

namespace {
constexpr int ST_BS = 256;
constexpr int ST_MAXB = 1024;

struct LevelStat
{
   int    min_e, max_e;
   double min_rs, max_rs, min_w, max_w;
};

// hypre_min(a, b) = a < b ? a : b and hypre_max(a, b) = a < b ? b : a, with
// the operands in the order par_stats.c passes them
__device__ __host__ inline double min_rs_step(double rs, double m) { return rs < m ? rs : m; }
__device__ __host__ inline double max_rs_step(double rs, double m) { return rs < m ? m : rs; }

// part[b * 6 + {0..5}]: min entries, max entries, min row sum, max row sum,
// min weight, max weight (weights only when `weights`); a block without rows
// writes +inf / -inf so the host fold skips it
__global__ __launch_bounds__(ST_BS) void k_level_stats(int n, const int *__restrict__ ia,
                                                       const double *__restrict__ a, int weights,
                                                       double w0, double *__restrict__ part)
{
   __shared__ double sh[6][ST_BS];
   const double inf = __builtin_huge_val();
   double mne = inf, mxe = -inf, mnr = inf, mxr = -inf, mnw = w0, mxw = 0.0;
   for (int r = blockIdx.x * ST_BS + threadIdx.x; r < n; r += gridDim.x * ST_BS) {
      const int b = ia[r], e = ia[r + 1];
      double rs = 0.0;
      for (int k = b; k < e; k++) {
         const double v = a[k];
         if (weights) {
            mnw = (mnw < v) ? mnw : v;
            if (v != 1.0) { mxw = (mxw < v) ? v : mxw; }
         }
         rs += v;
      }
      const double ne = (double) (e - b);
      mne = (ne < mne) ? ne : mne;
      mxe = (mxe < ne) ? ne : mxe;
      mnr = min_rs_step(rs, mnr);
      mxr = max_rs_step(rs, mxr);
   }
   const int t = threadIdx.x;
   sh[0][t] = mne; sh[1][t] = mxe; sh[2][t] = mnr; sh[3][t] = mxr; sh[4][t] = mnw; sh[5][t] = mxw;
   __syncthreads();
   for (int off = ST_BS / 2; off > 0; off >>= 1) {
      if (t < off) {
         const int u = t + off;
         sh[0][t] = (sh[0][u] < sh[0][t]) ? sh[0][u] : sh[0][t];
         sh[1][t] = (sh[1][t] < sh[1][u]) ? sh[1][u] : sh[1][t];
         sh[2][t] = min_rs_step(sh[2][u], sh[2][t]);
         sh[3][t] = max_rs_step(sh[3][u], sh[3][t]);
         sh[4][t] = (sh[4][t] < sh[4][u]) ? sh[4][t] : sh[4][u];
         sh[5][t] = (sh[5][t] < sh[5][u]) ? sh[5][u] : sh[5][t];
      }
      __syncthreads();
   }
   if (t < 6) { part[blockIdx.x * 6 + t] = sh[t][0]; }
}

// par_stats.c:587-650 (operators, weights = false) and :851-960
// (interpolation, weights = true)
LevelStat level_stat(const Csr &M, bool weights)
{
   LevelStat s{0, 0, 0.0, 0.0, 1.0, 0.0};
   if (M.nrows == 0) {
      s.min_e = 1000000;
      s.min_rs = 1.0e7;
      s.min_w = 1.0e7;
      return s;
   }
   // the reference starts the smallest weight at P's first stored value
   double w0 = 1.0;
   if (weights && M.nnz > 0) { d2h(&w0, M.a, sizeof(double)); }
   const int nb = std::max(1, std::min(ST_MAXB, (M.nrows + ST_BS - 1) / ST_BS));
   double *part = dalloc_n<double>((size_t) nb * 6);
   k_level_stats<<<nb, ST_BS, 0, stream()>>>(M.nrows, M.ia, M.a, weights ? 1 : 0, w0, part);
   HX_LAUNCH_CHECK();
   std::vector<double> h((size_t) nb * 6);
   d2h(h.data(), part, sizeof(double) * (size_t) nb * 6);
   dfree(part);
   double mne = h[0], mxe = h[1], mnr = h[2], mxr = h[3], mnw = h[4], mxw = h[5];
   for (int b = 1; b < nb; b++) {
      const double *q = &h[(size_t) b * 6];
      mne = (q[0] < mne) ? q[0] : mne;
      mxe = (mxe < q[1]) ? q[1] : mxe;
      mnr = min_rs_step(q[2], mnr);
      mxr = max_rs_step(q[3], mxr);
      mnw = (mnw < q[4]) ? mnw : q[4];
      mxw = (mxw < q[5]) ? q[5] : mxw;
   }
   s.min_e = (int) mne;
   s.max_e = (int) mxe;
   s.min_rs = mnr;
   s.max_rs = mxr;
   if (weights) {
      s.min_w = mnw;
      s.max_w = mxw;
   }
   return s;
}

// printf.c:136 hypre_ndigits_ull
int ndigits(long long v)
{
   int d = 0;
   while (v > 0) {
      v /= 10;
      d++;
   }
   return d;
}
} // namespace

void AMGSolver::print_setup_stats() const
{
   const AMGParams &p = prm;
   const int nl = num_levels;
   std::FILE *o = stdout;
   // par_stats.c:209-325
   std::fprintf(o, "\n\n Num MPI tasks = %d\n", 1);
   std::fprintf(o, " Num OpenMP threads = %d\n", 1);
   std::fprintf(o, " Execution policy = %s\n\n", "Device (HIP)");
   std::fprintf(o, " Max levels = %d\n", p.max_levels);
   std::fprintf(o, " Num levels = %d\n\n", nl);
   std::fprintf(o, " Strength Threshold = %f\n", p.strong_threshold);
   std::fprintf(o, " Interpolation Truncation Factor = %f\n", p.trunc_factor);
   std::fprintf(o, " Maximum Row Sum Threshold for Dependency Weakening = %f\n\n", p.max_row_sum);
   std::fprintf(o, " Coarsening Type = %s \n", p.coarsen_type == 10 ? "HMIS" : "PMIS");
   if (p.agg_num_levels > 0) {
      std::fprintf(o, "\n No. of levels of aggressive coarsening: %d\n\n", p.agg_num_levels);
      std::fprintf(o, " Interpolation on agg. levels = matrix-matrix multipass\n");
   }
   std::fprintf(o, " measures are determined %s\n\n", p.measure_type ? "globally" : "locally");
   std::fprintf(o, " Interpolation = extended+i\n");
   std::fprintf(o, "\nOperator Matrix Information:\n\n");

   // par_stats.c:443-486: column widths from the finest level
   const Csr &A0 = *L[0].A;
   int nd0 = std::max(7, ndigits(A0.nrows));
   const int nd1 = std::max(8, ndigits(A0.nnz));
   int nd2 = 4;
   for (int l = 0; l < nl; l++) {
      const Csr &A = *L[l].A;
      if (A.nrows > 0) { nd2 = std::max(nd2, ndigits(A.nnz / A.nrows)); }
   }
   nd2 += 2;
   const int nd3 = nd0 + nd1 + nd2;
   std::fprintf(o, "%*s", nd0 + 13, "nonzero");
   std::fprintf(o, "%*s", nd1 + 15, "entries/row");
   std::fprintf(o, "%18s\n", "row sums");
   std::fprintf(o, "%s %*s ", "lev", nd0, "rows");
   std::fprintf(o, "%*s", nd1, "entries");
   std::fprintf(o, "%7s %5s %4s", "sparse", "min", "max");
   std::fprintf(o, "%*s %8s %11s\n", nd2 + 2, "avg", "min", "max");
   for (int i = 0; i < 49 + nd3; i++) { std::fputc('=', o); }
   std::fputc('\n', o);

   // par_stats.c:497-700; memory counts each operator once and each P once
   std::vector<double> coeffs(nl), mem(nl), vars(nl);
   for (int l = 0; l < nl; l++) {
      const Csr &A = *L[l].A;
      const double nnz = (double) A.nnz, rows = (double) A.nrows;
      coeffs[l] = nnz;
      mem[l] = nnz;
      vars[l] = rows;
      const LevelStat s = level_stat(A, false);
      std::fprintf(o, "%3d %*d %*.0f  %0.3f  %4d %4d", l, nd0, A.nrows, nd1, nnz, nnz / (rows * rows),
                   s.min_e, s.max_e);
      std::fprintf(o, "  %*.1f  %10.3e  %10.3e\n", nd2, nnz / rows, s.min_rs, s.max_rs);
   }

   // par_stats.c:702-996
   nd0 = 5;
   if (nl - 1) { nd0 = std::max(nd0, ndigits(L[0].P.nrows)); }
   std::fprintf(o, "\n\nInterpolation Matrix Information:\n");
   std::fprintf(o, "%*s ", 2 * nd0 + 21, "entries/row");
   std::fprintf(o, "%10s %10s %19s\n", "min", "max", "row sums");
   std::fprintf(o, "lev %*s x %-*s min  max  avgW", nd0, "rows", nd0, "cols");
   std::fprintf(o, "%11s %11s %9s %11s\n", "weight", "weight", "min", "max");
   for (int i = 0; i < 70 + 2 * nd0; i++) { std::fputc('=', o); }
   std::fputc('\n', o);
   for (int l = 0; l < nl - 1; l++) {
      const Csr &P = L[l].P;
      const double nnz = (double) P.nnz;
      mem[l] += nnz;
      const LevelStat s = level_stat(P, true);
      std::fprintf(o, "%3d %*d x %-*d %3d  %3d", l, nd0, P.nrows, nd0, P.ncols, s.min_e, s.max_e);
      std::fprintf(o, "  %4.1f  %10.3e  %10.3e  %10.3e  %10.3e\n",
                   (nnz - (double) P.ncols) / ((double) (P.nrows - P.ncols)), s.min_w, s.max_w,
                   s.min_rs, s.max_rs);
   }

   // par_stats.c:997-1020
   double grid = 0.0, oper = 0.0, memc = 0.0, total = 0.0;
   for (int l = 0; l < nl; l++) {
      memc += mem[l] / coeffs[0];
      oper += coeffs[l] / coeffs[0];
      total += vars[l];
   }
   if (vars[0] != 0) { grid = total / vars[0]; }
   std::fprintf(o, "\n\n");
   std::fprintf(o, "     Complexity:   grid = %f\n", grid);
   std::fprintf(o, "               operator = %f\n", oper);
   std::fprintf(o, "                 memory = %f\n", memc);
   std::fprintf(o, "\n\n");

   // par_stats.c:1022-1034 and 1246-1322 (V-cycle, relax_order 0, no relax points)
   std::fprintf(o, "\n\nBoomerAMG SOLVER PARAMETERS:\n\n");
   std::fprintf(o, "  Maximum number of cycles:         %d \n", p.max_iter);
   std::fprintf(o, "  Stopping Tolerance:               %e \n", p.tol);
   std::fprintf(o, "  Cycle type (1 = V, 2 = W, etc.):  %d\n\n", p.cycle_type);
   std::fprintf(o, "  Relaxation Parameters:\n");
   std::fprintf(o, "   Visiting Grid:                     down   up  coarse\n");
   std::fprintf(o, "            Number of sweeps:         %4d   %2d  %4d \n", p.num_sweeps[1],
                p.num_sweeps[2], coarse_sweeps);
   std::fprintf(o, "   Type 0=Jac, 3=hGS, 6=hSGS, 9=GE:   %4d   %2d  %4d \n", p.relax_type[1],
                p.relax_type[2], coarse_relax);
   std::fprintf(o, "   Point types, partial sweeps (1=C, -1=F):\n");
   std::fprintf(o, "                  Pre-CG relaxation (down):");
   for (int j = 0; j < p.num_sweeps[1]; j++) { std::fprintf(o, "  %2d", 0); }
   std::fprintf(o, "\n");
   std::fprintf(o, "                   Post-CG relaxation (up):");
   for (int j = 0; j < p.num_sweeps[2]; j++) { std::fprintf(o, "  %2d", 0); }
   std::fprintf(o, "\n");
   std::fprintf(o, "                             Coarsest grid:");
   for (int j = 0; j < coarse_sweeps; j++) { std::fprintf(o, "  %2d", 0); }
   std::fprintf(o, "\n");
   std::fprintf(o, "\n");
   for (int l = 0; l < nl; l++) {
      if (omega(l) != 1) { std::fprintf(o, " Relaxation Weight %f level %d\n", omega(l), l); }
   }
   for (int l = 0; l < nl; l++) {
      if (p.outer_wt != 1) { std::fprintf(o, " Outer relaxation weight %f level %d\n", p.outer_wt, l); }
   }
   std::fflush(o);
}

} // namespace hx
