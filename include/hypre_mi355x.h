/*
 * hypre_mi355x.h -- extensions of libhypre_mi355x.so that have no
 * counterpart in hypre's public API.  They serve three purposes:
 *   1. the FP32 V-cycle inside FP64 PCG and the learned-parameter model,
 *   2. inspection of the AMG hierarchy and of single setup kernels, so a
 *      caller can compare them with a CPU implementation of the same steps,
 *   3. kernel timing with HIP events and the problem generators used by the
 *      benchmark workloads.
 * All arrays passed in or out are HOST arrays unless a name says _dev.
 */
#ifndef HYPRE_MI355X_HEADER
#define HYPRE_MI355X_HEADER

#include "HYPRE_utilities.h"
#include "HYPRE_parcsr_mv.h"
#include "HYPRE_parcsr_ls.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- precision of the V-cycle: 0 = FP64 (default), 1 = FP32 operators and
 * level vectors inside the FP64 outer iteration. */
HYPRE_Int HYPRE_MI355X_BoomerAMGSetCyclePrecision(HYPRE_Solver solver, HYPRE_Int fp32);

/* ---- per-level parameter overrides (used by the learned model): strong
 * threshold and relaxation weight of level `level`; sweeps applies to both
 * pre- and post-smoothing on that level. Negative level = all levels. */
HYPRE_Int HYPRE_MI355X_BoomerAMGSetLevelStrongThreshold(HYPRE_Solver solver, HYPRE_Int level,
                                                        HYPRE_Real theta);
HYPRE_Int HYPRE_MI355X_BoomerAMGSetLevelRelaxWt(HYPRE_Solver solver, HYPRE_Int level,
                                                HYPRE_Real omega);
HYPRE_Int HYPRE_MI355X_BoomerAMGSetLevelNumSweeps(HYPRE_Solver solver, HYPRE_Int level,
                                                  HYPRE_Int sweeps);
HYPRE_Int HYPRE_MI355X_BoomerAMGClearLevelParams(HYPRE_Solver solver);
/* number of levels of the set-up hierarchy (hypre keeps it in the solver
 * object, hypre_ParAMGDataNumLevels, but has no public getter) */
HYPRE_Int HYPRE_MI355X_BoomerAMGGetNumLevels(HYPRE_Solver solver, HYPRE_Int *num_levels);
/* row partitions of the hybrid Gauss-Seidel smoothers (relax 3/4/6/8/13/14),
 * which relax each partition in Gauss-Seidel order and couple partitions
 * through the iterate before the sweep, like the reference's OpenMP threads:
 * parts > 0 uses that many partitions on every level (the reference run with
 * `parts` threads); parts == 0 uses partitions of `rows` rows (default 4),
 * fewer (down to 8) on levels with fewer than about 128K partitions. */
HYPRE_Int HYPRE_MI355X_BoomerAMGSetHybridPartitions(HYPRE_Solver solver, HYPRE_Int parts,
                                                    HYPRE_Int rows);
/* kernels of the solve phase.  1 (default): at setup, a sliced-ELL copy of
 * each level operator (64-row slices, entries column-major, padding at most
 * 25% of the entries, levels of at least 4096 rows) for the SpMV and
 * l1-Jacobi kernels and PCG's products with the fine operator, and each
 * coarse level's first l1-Jacobi sweep from a zero guess formed by the
 * restriction that produces its right-hand side.  0: CSR kernels and one
 * launch per sweep, the plain sequence of hypre's cycle.  Both give the same
 * bits.  The copy belongs to the setup: values of A changed in place
 * afterwards need a new setup. */
HYPRE_Int HYPRE_MI355X_BoomerAMGSetSolveFormat(HYPRE_Solver solver, HYPRE_Int format);
/* keep the strength pattern S and the PMIS measure of every level for
 * inspection (off by default; they are freed during setup otherwise) */
HYPRE_Int HYPRE_MI355X_BoomerAMGSetKeepSetupData(HYPRE_Solver solver, HYPRE_Int keep);

/* ---- hierarchy inspection.  which: 'A' level operator, 'P' interpolation,
 * 'R' its stored transpose, 'S' strength pattern (a == NULL allowed).
 * Call with ia == NULL to query sizes; then pass arrays of those sizes. */
HYPRE_Int HYPRE_MI355X_BoomerAMGGetLevelMatrix(HYPRE_Solver solver, HYPRE_Int level, char which,
                                               HYPRE_Int *nrows, HYPRE_Int *ncols,
                                               HYPRE_Int *nnz, HYPRE_Int *ia, HYPRE_Int *ja,
                                               HYPRE_Real *a);
HYPRE_Int HYPRE_MI355X_BoomerAMGGetLevelCF(HYPRE_Solver solver, HYPRE_Int level, HYPRE_Int *cf);
HYPRE_Int HYPRE_MI355X_BoomerAMGGetLevelL1(HYPRE_Solver solver, HYPRE_Int level, HYPRE_Real *l1);
HYPRE_Int HYPRE_MI355X_BoomerAMGGetLevelMeasure(HYPRE_Solver solver, HYPRE_Int level,
                                                HYPRE_Real *measure);
HYPRE_Int HYPRE_MI355X_BoomerAMGGetLevelPMISRounds(HYPRE_Solver solver, HYPRE_Int level,
                                                   HYPRE_Int *rounds);
/* one V-cycle on host vectors of the fine size: u <- V(f, u) */
HYPRE_Int HYPRE_MI355X_BoomerAMGCycleHost(HYPRE_Solver solver, const HYPRE_Real *f,
                                          HYPRE_Real *u);
/* PCG residual history ||r_i||/||b||, i = 1..iters of the last solve */
HYPRE_Int HYPRE_MI355X_PCGGetResidualHistory(HYPRE_Solver solver, HYPRE_Int max_len,
                                             HYPRE_Real *hist, HYPRE_Int *len);
/* wall-clock seconds of the last PCG setup / solve (device-synchronised) */
HYPRE_Int HYPRE_MI355X_PCGGetTimes(HYPRE_Solver solver, HYPRE_Real *setup_s,
                                   HYPRE_Real *solve_s);

/* ---- single setup kernels on host CSR input (ia/ja/a of n rows) -------- */
/* strength pattern: S_ia (n+1) and S_ja (nnz(A) capacity) */
HYPRE_Int HYPRE_MI355X_CreateS(HYPRE_Int n, const HYPRE_Int *ia, const HYPRE_Int *ja,
                               const HYPRE_Real *a, HYPRE_Real theta, HYPRE_Real max_row_sum,
                               HYPRE_Int *S_ia, HYPRE_Int *S_ja);
/* PMIS on a strength pattern; measure == NULL draws the reference's
 * Park-Miller stream (seed 2747), else the given measure is used */
HYPRE_Int HYPRE_MI355X_PMIS(HYPRE_Int n, const HYPRE_Int *S_ia, const HYPRE_Int *S_ja,
                            const HYPRE_Real *measure, HYPRE_Int *cf, HYPRE_Int *rounds);
/* HMIS (coarsen_type 10, src/parcsr_ls/par_coarsen.c:2842): the
 * Ruge-Stueben first pass on the host, then PMIS with CF_init 1 on the device;
 * measure == NULL draws the reference's Park-Miller stream (seed 2747) */
HYPRE_Int HYPRE_MI355X_HMIS(HYPRE_Int n, const HYPRE_Int *S_ia, const HYPRE_Int *S_ja,
                            const HYPRE_Real *measure, HYPRE_Int *cf, HYPRE_Int *rounds);
/* the first pass alone (par_coarsen.c:911 with coarsen_type 10, host only):
 * cf 1 = C, -1 = F, -2 = Z (nobody depends on it), -3 = no strong connections */
HYPRE_Int HYPRE_MI355X_RSFirstPass(HYPRE_Int n, const HYPRE_Int *S_ia, const HYPRE_Int *S_ja,
                                   HYPRE_Int *cf);
/* ext+i interpolation + truncation; P arrays sized by a first call with
 * P_ia == NULL that returns *P_nnz; cf is updated (-3 -> -1) */
HYPRE_Int HYPRE_MI355X_ExtPIInterp(HYPRE_Int n, const HYPRE_Int *ia, const HYPRE_Int *ja,
                                   const HYPRE_Real *a, const HYPRE_Int *S_ia,
                                   const HYPRE_Int *S_ja, HYPRE_Int *cf,
                                   HYPRE_Real trunc_factor, HYPRE_Int max_elmts,
                                   HYPRE_Int *P_ncols, HYPRE_Int *P_nnz, HYPRE_Int *P_ia,
                                   HYPRE_Int *P_ja, HYPRE_Real *P_a);
/* Galerkin product Ac = P^T A P (R = P^T is built internally); two-call
 * pattern as above */
HYPRE_Int HYPRE_MI355X_RAP(HYPRE_Int n, HYPRE_Int nc, const HYPRE_Int *ia, const HYPRE_Int *ja,
                           const HYPRE_Real *a, const HYPRE_Int *P_ia, const HYPRE_Int *P_ja,
                           const HYPRE_Real *P_a, HYPRE_Int *Ac_nnz, HYPRE_Int *Ac_ia,
                           HYPRE_Int *Ac_ja, HYPRE_Real *Ac_a);
/* y = alpha*A*x + beta*y on host vectors, through the HIP SpMV */
HYPRE_Int HYPRE_MI355X_SpMV(HYPRE_Int n, HYPRE_Int ncols, const HYPRE_Int *ia,
                            const HYPRE_Int *ja, const HYPRE_Real *a, HYPRE_Real alpha,
                            const HYPRE_Real *x, HYPRE_Real beta, HYPRE_Real *y);

/* one hybrid Gauss-Seidel relaxation (relax_type 3/4/6/8/13/14; l1 norms
 * required for 8/13/14) with `parts` row partitions, u updated in place;
 * zero_guess != 0 reads u as zeros.  src/parcsr_ls/par_relax.c:692 with the
 * reference's thread count equal to `parts`. */
HYPRE_Int HYPRE_MI355X_HybridGS(HYPRE_Int n, const HYPRE_Int *ia, const HYPRE_Int *ja,
                                const HYPRE_Real *a, const HYPRE_Real *f, const HYPRE_Real *l1,
                                HYPRE_Real *u, HYPRE_Int relax_type, HYPRE_Real relax_weight,
                                HYPRE_Real omega, HYPRE_Int parts, HYPRE_Int zero_guess);

/* ---- ParCSR access ------------------------------------------------------ */
/* borrowed device pointers of the (single-rank) CSR block */
HYPRE_Int HYPRE_MI355X_ParCSRMatrixGetDevice(HYPRE_ParCSRMatrix A, HYPRE_Int *n,
                                             HYPRE_Int *nnz, HYPRE_Int **ia_dev,
                                             HYPRE_Int **ja_dev, HYPRE_Real **a_dev);
HYPRE_Int HYPRE_MI355X_ParCSRMatrixCopyToHost(HYPRE_ParCSRMatrix A, HYPRE_Int *ia,
                                              HYPRE_Int *ja, HYPRE_Real *a);
HYPRE_Int HYPRE_MI355X_ParVectorGetDevice(HYPRE_ParVector v, HYPRE_Int *n, HYPRE_Real **d);
HYPRE_Int HYPRE_MI355X_ParVectorCopyFromHost(HYPRE_ParVector v, const HYPRE_Real *x);
HYPRE_Int HYPRE_MI355X_ParVectorCopyToHost(HYPRE_ParVector v, HYPRE_Real *x);
/* direct construction from host CSR (diagonal must be first in each row) */
HYPRE_ParCSRMatrix HYPRE_MI355X_ParCSRMatrixFromHostCSR(HYPRE_Int n, HYPRE_Int ncols,
                                                        const HYPRE_Int *ia,
                                                        const HYPRE_Int *ja,
                                                        const HYPRE_Real *a);

/* ---- benchmark problem generators (built on the device) ---------------- */
/* 7- or 27-point diffusion on an N^3 grid with a piecewise-constant kappa on
 * cubic cells of `cell` nodes, kappa = 10^(lo + (hi-lo) U), U from the
 * Park-Miller stream seeded with `seed` in lexicographic cell order, harmonic
 * face means, Dirichlet faces using the node's own kappa. */
HYPRE_ParCSRMatrix HYPRE_MI355X_GenerateVarDiffusion(HYPRE_Int N, HYPRE_Int stencil,
                                                     HYPRE_Int cell, HYPRE_Real log10_lo,
                                                     HYPRE_Real log10_hi, HYPRE_Int seed);
/* 7-point anisotropic diffusion, diagonal of R_z(theta) diag(1, eps, 1) R_z^T */
HYPRE_ParCSRMatrix HYPRE_MI355X_GenerateAniso7(HYPRE_Int N, HYPRE_Real eps, HYPRE_Real theta);
/* x_i = 2*U_i - 1 with U the Park-Miller stream from `seed` (seed < 1 -> 1) */
HYPRE_Int HYPRE_MI355X_ParVectorSetRandomPM(HYPRE_ParVector v, HYPRE_Int seed);

/* ---- learned per-level parameters ---------------------------------------
 * A perceptron 8 -> nh1 -> nh2 -> nout (ReLU hidden layers, sigmoid output)
 * on the device, FP32, trained with mean-squared error and Adam.  Parameters
 * are one float array: W1[nin][nh1], b1, W2[nh1][nh2], b2, W3[nh2][nout], b3
 * (row-major, one sample per row of X).  Output y of level l (nout = 3 L):
 * theta = 0.1 + 0.8 y[3l], omega = 0.6 + 0.8 y[3l+1],
 * sweeps = 1 + floor(2 y[3l+2] + 0.5). */
typedef struct hx_param_model_struct *HYPRE_MI355X_ParamModel;
HYPRE_Int HYPRE_MI355X_ParamModelCreate(HYPRE_Int nin, HYPRE_Int nh1, HYPRE_Int nh2, HYPRE_Int nout,
                                        HYPRE_MI355X_ParamModel *model);
HYPRE_Int HYPRE_MI355X_ParamModelDestroy(HYPRE_MI355X_ParamModel model);
HYPRE_Int HYPRE_MI355X_ParamModelNumParams(HYPRE_MI355X_ParamModel model, HYPRE_Int *np);
/* host arrays of NumParams floats; SetParams also resets the Adam state */
HYPRE_Int HYPRE_MI355X_ParamModelSetParams(HYPRE_MI355X_ParamModel model, const float *params);
HYPRE_Int HYPRE_MI355X_ParamModelGetParams(HYPRE_MI355X_ParamModel model, float *params);
/* Y[batch][nout] = sigmoid outputs for X[batch][nin] */
HYPRE_Int HYPRE_MI355X_ParamModelForward(HYPRE_MI355X_ParamModel model, HYPRE_Int batch,
                                         const float *X, float *Y);
/* gradient of mean((Y - T)^2) with respect to the parameters, and the loss */
HYPRE_Int HYPRE_MI355X_ParamModelGradient(HYPRE_MI355X_ParamModel model, HYPRE_Int batch,
                                          const float *X, const float *T, float *grad, float *loss);
/* `steps` full-batch Adam steps (beta1 0.9, beta2 0.999, eps 1e-8); loss
 * before each step into loss_hist[steps] (may be NULL) */
HYPRE_Int HYPRE_MI355X_ParamModelTrain(HYPRE_MI355X_ParamModel model, HYPRE_Int batch,
                                       const float *X, const float *T, HYPRE_Int steps, float lr,
                                       float *loss_hist);
/* the 8 matrix features the model reads: log10(n)/8; the row-length
 * distribution (mean/27, standard deviation/27); the off-diagonal anisotropy
 * ratios (mean min/max |a_ij|, mean 2nd/1st largest |a_ij|); the mean strong
 * fraction at theta 0.25; a coefficient-jump histogram (fractions of rows
 * with a neighbour whose |diagonal| differs by a factor >= 10 and >= 1000) */
HYPRE_Int HYPRE_MI355X_MatrixFeatures(HYPRE_ParCSRMatrix A, HYPRE_Real *features);
/* cost-model reading of a model with nout = 2 ng: output g predicts the
 * time of setting g (theta[g], omega[g], sweeps[g]) on the fine level
 * relative to the defaults, output ng + g the same on the coarse levels;
 * ng = 0 returns to the per-level reading */
HYPRE_Int HYPRE_MI355X_ParamModelSetGrid(HYPRE_MI355X_ParamModel model, HYPRE_Int ng,
                                         const HYPRE_Real *theta, const HYPRE_Real *omega,
                                         const HYPRE_Int *sweeps);
/* prior of the cost-model reading: grid indices of the settings kept on the
 * fine level and on the coarse levels (-1: none) unless another setting is
 * predicted faster by more than `margin` (in relative time); set after the
 * grid, which clears it */
HYPRE_Int HYPRE_MI355X_ParamModelSetPrior(HYPRE_MI355X_ParamModel model, HYPRE_Int fine,
                                          HYPRE_Int coarse, HYPRE_Real margin);
/* features of A -> model -> per-level theta, omega, sweeps of the solver
 * (with a grid: the predicted fastest setting on level 0 and on levels
 * 1..24) */
HYPRE_Int HYPRE_MI355X_BoomerAMGSetParamModel(HYPRE_Solver solver, HYPRE_MI355X_ParamModel model,
                                              HYPRE_ParCSRMatrix A);

/* ---- kernel timing with HIP events --------------------------------------
 * While enabled, every launch of the named kernel class is bracketed by HIP
 * events on the stream it runs on.  Classes: "spmv_A0" (fine-level SpMV of
 * the solve phase), "relax_A0", "spmv", "relax", "setup", "solve". */
HYPRE_Int HYPRE_MI355X_SetKernelTiming(HYPRE_Int on);
/* restrict timing to one class (NULL or "" = all classes) */
HYPRE_Int HYPRE_MI355X_SetKernelTimingClass(const char *name);
HYPRE_Int HYPRE_MI355X_ResetKernelTiming(void);
HYPRE_Int HYPRE_MI355X_GetKernelTiming(const char *name, HYPRE_Real *total_ms,
                                       HYPRE_Int *launches, HYPRE_Real *bytes);

/* ---- profiling ranges ---------------------------------------------------
 * roctx ranges with the names of the reference's GPU profiling ranges
 * (src/utilities/device_markers.c; "PCG-Setup", "PCG-Solve", "AMGsetup",
 * "AMG Level-<l>", "CreateS", "PMIS", "ExtPIInterp", "TripleMat-RAP",
 * "AMGCycle", "Relaxation", "Residual", "Restriction", "Interpolation",
 * "Coarse solve"), for rocprofv3 --marker-trace.  Off by default (also
 * enabled by HYPRE_MI355X_ROCTX=1 in the environment); closing a range waits
 * for the library stream, as the reference does, so timings with ranges on
 * include those waits.  HYPRE_ERROR_GENERIC when the roctx library cannot be
 * loaded. */
HYPRE_Int HYPRE_MI355X_SetProfilingRanges(HYPRE_Int on);
/* ranges opened so far and ranges still open */
HYPRE_Int HYPRE_MI355X_GetProfilingRanges(HYPRE_Int *pushed, HYPRE_Int *open);

/* ---- allocation-free variant (no counterpart in hypre) ------------------
 * For callers that cannot allocate device or pinned memory: code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * HYPRE_ParVectorInnerProd (src/parcsr_mv/HYPRE_parcsr_vector.c) with its
 * per-block partial sums and its result kept in a fixed 16.6 KB device array
 * of the library (one per device, taken at HYPRE_Initialize) instead of
 * memory allocated on first use, and the result copied straight into *prod.
 * Same bits as HYPRE_ParVectorInnerProd.  Any length a HYPRE_Int holds is
 * accepted (a call never needs more than 2048 partial sums); vectors of
 * different lengths return HYPRE_ERROR_ARG.  Calls that overlap on one device
 * share that array and race; calls from one host thread never overlap. */
HYPRE_Int HYPRE_MI355X_ParVectorInnerProdNoAlloc(HYPRE_ParVector x, HYPRE_ParVector y,
                                                 HYPRE_Real *prod);

#ifdef __cplusplus
}
#endif

#endif
