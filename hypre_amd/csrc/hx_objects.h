// Handle types behind the opaque pointers of the public headers.
#pragma once

#include "hx_internal.h"

#include <string>
#include <vector>

struct hypre_ParCSRMatrix_struct
{
   hx::Csr A;
   bool    owns = true;
   // host mirror used by GetRow / assembled-IJ updates (filled on demand)
   std::vector<int>    h_ia, h_ja;
   std::vector<double> h_a;
   std::vector<int>    row_cols_tmp;
};

struct hypre_ParVector_struct
{
   int     n = 0;
   double *d = nullptr;
   bool    owns = true;
   bool    all_zeros = false;   // hypre_ParVectorAllZeros flag
};

struct hypre_Solver_struct
{
   int kind = 0;                // 1 BoomerAMG, 2 PCG, 3 diagonal scaling
   virtual ~hypre_Solver_struct() {}
};

namespace hx {

struct Level
{
   Csr     Aown;                // owned operator (levels >= 1)
   const Csr *A = nullptr;      // operator of this level
   Csr     P, R, S;             // P, R = P^T for levels < L-1 (S kept on request)
   int    *CF = nullptr;
   double *measure = nullptr;   // kept on request
   double *l1 = nullptr;
   float  *l1f = nullptr;
   int     pmis_rounds = 0;
   // level vectors
   double *f = nullptr, *u = nullptr, *ut = nullptr, *r = nullptr;
   float  *f32 = nullptr, *u32 = nullptr, *ut32 = nullptr, *r32 = nullptr;
   Csr     A32, P32, R32;       // FP32 value copies share index arrays (a32 filled)
   Csr     Arev;                // rows of A reversed, for backward hybrid Gauss-Seidel
   int     max_row = -1;        // longest row of A (hybrid Gauss-Seidel chunking)
   Ell     ell;                 // sliced-ELL copy of A for the solve phase (when built)
   Csr     Av;                  // the operator as the solve phase sees it: *A plus ell
};

struct AMGParams
{
   double strong_threshold = 0.25;
   double max_row_sum = 0.9;
   double trunc_factor = 0.0;
   int    P_max_elmts = 4;
   int    max_levels = 25;
   int    max_coarse_size = 9;
   int    min_coarse_size = 0;
   int    coarsen_type = 8;       // hypre host default is 10; this build runs PMIS
   int    interp_type = 6;
   int    relax_type[4] = {18, 18, 18, 9};
   int    num_sweeps[4] = {1, 1, 1, 1};
   double relax_wt = 1.0;
   double outer_wt = 1.0;
   int    max_iter = 20;
   double tol = 1e-9;
   int    cycle_type = 1;
   int    relax_order = 0;
   int    measure_type = 0;
   int    agg_num_levels = 0;
   int    agg_interp_type = 4;   // hypre's default; this build runs 8 (modified multipass)
   double agg_trunc_factor = 0.0;
   int    agg_P_max_elmts = 0;
   int    num_functions = 1;
   int    print_level = 0;
   int    logging = 0;
   int    keep_transpose = 1;
   int    cycle_fp32 = 0;
   int    keep_setup_data = 0;
   int    hgs_parts = 0;         // hybrid GS partitions on every level (the reference's thread count)
   int    hgs_rows = 4;          // else partitions of about this many rows
   int    solve_format = 1;      // 1: sliced-ELL copies of the level operators where padding allows; 0: CSR
   std::vector<double> level_theta, level_omega;
   std::vector<int>    level_sweeps;
};

struct AMGSolver : public hypre_Solver_struct
{
   AMGParams          prm;
   std::vector<Level> L;
   int     num_levels = 0;
   int     coarse_relax = 9;
   int     coarse_sweeps = 1;
   double *Amat = nullptr;        // coarsest dense matrix (column major)
   int     coarse_n = 0;
   bool    setup_done = false;
   int     num_iterations = 0;
   double  rel_resid = 0.0;
   double  setup_seconds = 0.0;
   AMGSolver() { kind = 1; }
   ~AMGSolver() override;
   void free_hierarchy();
   void setup(const Csr &A);
   // the setup statistics table (par_stats.c:20), print level 1 or 3
   void print_setup_stats() const;
   void cycle(const double *f, double *u, bool u_zero);
   void cycle32(const double *f, double *u);
   int  solve(const Csr &A, hypre_ParVector_struct *b, hypre_ParVector_struct *x);
   double theta(int lev) const;
   double omega(int lev) const;
   int    sweeps(int lev, int k) const;
   void   relax_any(int type, int lev, const Csr &A, const double *f, const double *u,
                    const double *l1, double *uo, bool zero, const char *cls) const;
   void   relax_any32(int type, int lev, const Csr &A, const float *f, const float *u,
                      const float *l1, float *uo, bool zero, const char *cls) const;
   // the solve-phase view of A when A is this hierarchy's fine operator, else A
   const Csr &solve_view(const Csr &A) const;
};

struct PCGSolver : public hypre_Solver_struct
{
   double tol = 1e-6, atol = 0.0, rtol = 0.0, atolf = 0.0, cf_tol = 0.0;
   int    max_iter = 1000, two_norm = 0, rel_change = 0, stop_crit = 0;
   int    recompute_residual = 0, recompute_residual_p = 0, flex = 0, skip_break = 0;
   int    print_level = 0, logging = 0;
   HYPRE_PtrToParSolverFcn precond = nullptr, precond_setup = nullptr;
   HYPRE_Solver precond_solver = nullptr;
   hypre_ParVector_struct *p = nullptr, *s = nullptr, *r = nullptr;
   int    num_iterations = 0, converged = 0;
   double rel_residual_norm = 0.0;
   std::vector<double> history;
   double setup_seconds = 0.0, solve_seconds = 0.0;
   int    n = 0;
   PCGSolver() { kind = 2; }
   ~PCGSolver() override;
   void ensure_vectors(int n_);
};

double wall_seconds();
hypre_ParVector_struct *vec_create(int n);
void                    vec_destroy(hypre_ParVector_struct *v);

} // namespace hx
