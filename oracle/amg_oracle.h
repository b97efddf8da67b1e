/*
 * CPU oracle for the BoomerAMG-PCG hot path (TEST INFRASTRUCTURE ONLY).
 *
 * A plain-C restatement of the host code path that hypre 3.1.0 runs for
 * PMIS coarsening + extended+i interpolation + l1-Jacobi V-cycle + PCG; its
 * independent row loops run under OpenMP with the serial results.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker.  The product library never calls it.
 *
 * Every function cites the reference file:line it follows (paths relative to
 * the hypre source tree).  Floating-point operations are written in the same
 * order as the reference's serial host loops and compiled without FP
 * contraction, so results are bit-reproducible.
 */
#ifndef AMG_ORACLE_H
#define AMG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* CSR matrix; ia[nrows+1], ja[nnz], a[nnz] (a may be NULL: pattern only). */
typedef struct
{
   int     nrows;
   int     ncols;
   int    *ia;
   int    *ja;
   double *a;
} orc_csr;

void orc_csr_free(orc_csr *M);

/* threads of the OpenMP loops (OMP_NUM_THREADS, else the host's cores) */
int orc_num_threads(void);

/* ---------------- Park-Miller RNG (src/utilities/random.c) -------------- */
void   orc_seed_rand(int seed);                 /* random.c hypre_SeedRand */
int    orc_rand_i(void);                        /* random.c hypre_RandI    */
double orc_rand(void);                          /* random.c hypre_Rand     */
/* n consecutive hypre_Rand() draws after hypre_SeedRand(seed) */
void   orc_rand_stream(int seed, int n, double *out);

/* ---------------- problem generators ------------------------------------ */
/* 7-point Laplacian, src/parcsr_ls/par_laplace.c GenerateLaplacian, one rank;
 * row layout: diagonal first, then off-diagonals by ascending column. */
int orc_gen_laplace7(int nx, int ny, int nz, double cx, double cy, double cz,
                     orc_csr *A);
/* 27-point Laplacian, src/parcsr_ls/par_laplace_27pt.c GenerateLaplacian27pt */
int orc_gen_laplace27(int nx, int ny, int nz, orc_csr *A);
/* 7-point diffusion with a piecewise-constant coefficient kappa on cubic
 * cells of `cell` nodes, kappa = 10^(lo + (hi-lo)*U), U drawn from the
 * Park-Miller stream seeded with `seed` in lexicographic cell order; face
 * coefficient = harmonic mean; Dirichlet faces use the node's own kappa. */
int orc_gen_vardiff7(int N, int cell, double log10_lo, double log10_hi,
                     int seed, orc_csr *A);
/* 27-point version of the above (all 26 neighbours, harmonic means). */
int orc_gen_vardiff27(int N, int cell, double log10_lo, double log10_hi,
                      int seed, orc_csr *A);
/* 7-point anisotropic diffusion: diagonal of R_z(theta) diag(1,eps,1) R_z^T
 * (the xy cross term is dropped so the stencil stays 7-point). */
int orc_gen_aniso7(int N, double eps, double theta, orc_csr *A);

/* ---------------- kernels of the setup ---------------------------------- */
/* src/parcsr_ls/par_strength.c:75 hypre_BoomerAMGCreateSHost (num_functions 1) */
int orc_strength(const orc_csr *A, double theta, double max_row_sum, orc_csr *S);
/* PMIS measure init: column counts of S + hypre_Rand() stream.
 * src/parcsr_ls/par_coarsen.c:2094-2330, par_indepset.c:50-75 */
int orc_pmis_measure(const orc_csr *S, int seed, double *measure);
/* PMIS rounds for a given measure (par_coarsen.c:2330-2790).
 * CF out: 1 = C, -1 = F, -3 = isolated F (SF_PT). Returns #rounds. */
int orc_pmis_with_measure(const orc_csr *S, const double *measure_in, int *CF);
/* Full PMIS with the hypre seed (2747 on rank 0, CF_init 0) */
int orc_pmis(const orc_csr *S, int *CF);
/* Ruge-Stueben first pass as HMIS runs it (par_coarsen.c:911 with
 * coarsen_type 10): CF out 1 = C, -1 = F, -2 = Z (nobody depends on it),
 * -3 = SF (no strong connections) */
int orc_rs_first_pass(const orc_csr *S, int *CF);
/* HMIS (par_coarsen.c:2842): the first pass, then PMIS with CF_init 1 for a
 * given measure / the hypre seed.  Returns the PMIS rounds. */
int orc_hmis_with_measure(const orc_csr *S, const double *measure_in, int *CF);
int orc_hmis(const orc_csr *S, int *CF);
/* ---------------- aggressive coarsening (one rank, num_paths 1) -------- */
/* src/parcsr_ls/par_strength.c:1795 hypre_BoomerAMGCreate2ndSHost: strength
 * over the C points of CF (distance one or two in S), pattern only, rows and
 * columns in coarse numbering; C points with an empty row get CF 2.
 * Returns the number of C points. */
int  orc_create_2nd_s(const orc_csr *S, int *CF, orc_csr *S2);
/* PMIS with CF_init 3 (the second coarsening, par_amg_setup.c:1324) */
int  orc_pmis_cf3_with_measure(const orc_csr *S, const double *measure_in, int *CF);
int  orc_pmis_cf3(const orc_csr *S, int *CF);
/* src/parcsr_ls/par_strength.c:3068 hypre_BoomerAMGCorrectCFMarkerHost */
void orc_correct_cf_marker(int n, int *CF, const int *CFN);
/* src/parcsr_mv/par_csr_matop.c:348 hypre_ParMatmul (one rank) */
int  orc_matmul(const orc_csr *A, const orc_csr *B, orc_csr *C);
/* src/parcsr_ls/par_mod_multi_interp.c:17 hypre_BoomerAMGBuildModMultipassHost,
 * interp_type 8; CF -3 -> -1 at the end.  -1 when passes exceed 9. */
int  orc_mod_multipass(const orc_csr *A, const orc_csr *S, int *CF, double trunc_factor,
                       int P_max_elmts, orc_csr *P);

/* src/parcsr_ls/par_lr_interp.c:1024 hypre_BoomerAMGBuildExtPIInterpHost,
 * followed by par_interp.c:2622 truncation when max_elmts > 0 or trunc > 0.
 * Converts CF -3 -> -1 at the end, as the reference does. */
int orc_extpi_interp(const orc_csr *A, const orc_csr *S, int *CF,
                     double trunc_factor, int max_elmts, orc_csr *P);
/* untruncated ext+i (for tests of the truncation step) */
int orc_extpi_interp_raw(const orc_csr *A, const orc_csr *S, const int *CF,
                         orc_csr *P);
/* src/parcsr_mv/par_csr_matrix.c:2943 hypre_ParCSRMatrixTruncate
 * (rescale 1, infinity norm), in place on P */
int orc_truncate(orc_csr *P, double tol, int max_row_elmts);
/* src/utilities/qsort.c:452 hypre_qsort2_abs */
void orc_qsort2_abs(int *v, double *w, int left, int right);
/* src/seq_mv/csr_matop.c:1071 hypre_CSRMatrixTransposeHost (stable) */
int orc_transpose(const orc_csr *P, orc_csr *R);
/* src/parcsr_ls/par_rap.c:30 hypre_BoomerAMGBuildCoarseOperatorKT, R = P^T */
int orc_rap(const orc_csr *R, const orc_csr *A, const orc_csr *P, orc_csr *Ac);
/* src/parcsr_ls/ams.c:517 hypre_ParCSRComputeL1Norms option 1 */
int orc_l1_norms(const orc_csr *A, double *l1);
/* src/parcsr_ls/ams.c:649 option 4 on one rank (signed diagonal) */
int orc_l1_norms_diag(const orc_csr *A, double *l1);

/* ---------------- hybrid Gauss-Seidel (relax 3/4/6/8/13/14) ------------- */
/* src/utilities/utilities.c:99 hypre_partition1D */
void orc_partition1d(int n, int p, int j, int *s, int *e);
int  orc_hgs_parts(int n, int parts, int rows);
/* src/parcsr_ls/par_relax.c:692 hypre_BoomerAMGRelaxHybridGaussSeidel_core
 * with `parts` partitions (the reference's thread count) */
void orc_hybrid_gs(const orc_csr *A, const double *f, const double *l1, double *u,
                   int parts, int direction, int symm, int skip_diag,
                   double relax_weight, double omega);

/* ---------------- solve side -------------------------------------------- */
/* y = alpha*A*x + beta*b (src/seq_mv/csr_matvec.c:22) */
void orc_matvec(double alpha, const orc_csr *A, const double *x, double beta,
                const double *b, double *y);
/* y = alpha*A^T*x + beta*y (src/seq_mv/csr_matvec.c:914) */
void orc_matvecT(double alpha, const orc_csr *A, const double *x, double beta,
                 double *y);
double orc_dot(int n, const double *x, const double *y);

/* ---------------- full AMG ---------------------------------------------- */
typedef struct
{
   double strong_threshold;   /* 0.25 */
   double max_row_sum;        /* 0.9 library default, 1.0 in ij */
   double trunc_factor;       /* 0 */
   int    P_max_elmts;        /* 4 */
   int    max_levels;         /* 25 */
   int    max_coarse_size;    /* 9 */
   int    relax_type;         /* 18 l1-Jacobi, 0 Jacobi, 3/4/6/8/13/14 hybrid GS */
   int    coarse_relax_type;  /* 9 (Gaussian elimination) */
   double relax_weight;       /* 1.0 */
   int    num_sweeps;         /* 1 */
   int    relax_up_type;      /* -1: relax_type on the up cycle too */
   double outer_weight;       /* 1.0 (hypre's omega) */
   int    hgs_parts;          /* hybrid GS partitions; 0: about hgs_rows rows each */
   int    hgs_rows;           /* 4 */
   int    coarsen_type;       /* 8 PMIS, 10 HMIS */
   int    agg_num_levels;     /* 0: levels below this one are coarsened aggressively */
   int    agg_interp_type;    /* 8: modified multipass */
   double agg_trunc_factor;   /* 0 */
   int    agg_P_max_elmts;    /* 0 */
} orc_amg_params;

void orc_amg_default_params(orc_amg_params *p);

typedef struct orc_amg orc_amg;   /* opaque hierarchy */

orc_amg *orc_amg_setup(const orc_csr *A, const orc_amg_params *prm);
void     orc_amg_free(orc_amg *H);
int      orc_amg_num_levels(const orc_amg *H);
/* borrowed pointers into the hierarchy; P/R/S/CF exist for lev < L-1 */
const orc_csr *orc_amg_A(const orc_amg *H, int lev);
const orc_csr *orc_amg_P(const orc_amg *H, int lev);
const orc_csr *orc_amg_R(const orc_amg *H, int lev);
const orc_csr *orc_amg_S(const orc_amg *H, int lev);
const int     *orc_amg_CF(const orc_amg *H, int lev);
const double  *orc_amg_l1(const orc_amg *H, int lev);
int            orc_amg_pmis_rounds(const orc_amg *H, int lev);
int            orc_amg_coarse_relax(const orc_amg *H);
/* the setup statistics text (src/parcsr_ls/par_stats.c:20, print level 1)
 * into buf; returns the length the whole text needs */
int orc_amg_setup_stats(const orc_amg *H, int max_iter, double tol, char *buf, int cap);
/* one V-cycle: u <- V-cycle(f, u); src/parcsr_ls/par_cycle.c:23 */
int orc_amg_cycle(orc_amg *H, const double *f, double *u);
/* AMG as a solver, src/parcsr_ls/par_amg_solve.c:22 (converge_type 0) */
int orc_amg_solve(orc_amg *H, const double *b, double *x, double tol,
                  int max_iter, int *iters, double *rel_res, double *conv);
/* two-norm PCG preconditioned by one V-cycle, src/krylov/pcg.c:313.
 * history[i] = ||r_i||/||b|| for i = 1..iters (may be NULL, size max_iter+1) */
int orc_pcg_solve(orc_amg *H, const double *b, double *x, double tol,
                  int max_iter, int *iters, double *rel_res, double *history);

#ifdef __cplusplus
}
#endif
#endif
