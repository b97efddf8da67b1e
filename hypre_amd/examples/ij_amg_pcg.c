/*
 * ij_amg_pcg.c -- BoomerAMG-preconditioned PCG through hypre's public C API
 * only, the same calls as hypre's ij driver (src/test/ij.c:5555-5700 with
 * -solver 1 -pmis -interptype 6 -Pmx 4 -rlx 18 -ns 1 -rhsrand) and the
 * src/examples/ex5.c tutorial.  Nothing here is specific to the MI355X
 * library: the file compiles unchanged against hypre's headers.
 *
 *   ij_amg_pcg [N] [stencil 7|27] [max_row_sum]
 *
 * The matrix is assembled row by row with HYPRE_IJMatrixSetValues in the
 * column order of hypre's GenerateLaplacian (src/parcsr_ls/par_laplace.c)
 * resp. GenerateLaplacian27pt, so the hierarchy equals that of ij's
 * built-in generators.  RHS as ij -rhsrand: random values from seed 22775,
 * scaled to unit norm (src/test/ij.c:3857-3860).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "HYPRE.h"
#include "HYPRE_IJ_mv.h"
#include "HYPRE_krylov.h"
#include "HYPRE_parcsr_ls.h"
#include "HYPRE_parcsr_mv.h"

static int row_7pt(int N, int i, int *cols, double *vals)
{
   int x = i % N, y = (i / N) % N, z = i / (N * N), c = 0;
   cols[c] = i; vals[c++] = 6.0;
   if (z > 0) { cols[c] = i - N * N; vals[c++] = -1.0; }
   if (y > 0) { cols[c] = i - N; vals[c++] = -1.0; }
   if (x > 0) { cols[c] = i - 1; vals[c++] = -1.0; }
   if (x < N - 1) { cols[c] = i + 1; vals[c++] = -1.0; }
   if (y < N - 1) { cols[c] = i + N; vals[c++] = -1.0; }
   if (z < N - 1) { cols[c] = i + N * N; vals[c++] = -1.0; }
   return c;
}

static int row_27pt(int N, int i, int *cols, double *vals)
{
   int x = i % N, y = (i / N) % N, z = i / (N * N), c = 0;
   cols[c] = i; vals[c++] = 26.0;
   for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
         for (int dx = -1; dx <= 1; dx++) {
            if (!dx && !dy && !dz) { continue; }
            int xx = x + dx, yy = y + dy, zz = z + dz;
            if (xx < 0 || yy < 0 || zz < 0 || xx >= N || yy >= N || zz >= N) { continue; }
            cols[c] = xx + N * (yy + N * zz);
            vals[c++] = -1.0;
         }
   return c;
}

int main(int argc, char **argv)
{
   int N = argc > 1 ? atoi(argv[1]) : 30;
   int stencil = argc > 2 ? atoi(argv[2]) : 7;
   double mrs = argc > 3 ? atof(argv[3]) : 1.0;
   int n = N * N * N, ilower = 0, iupper = n - 1;
   int cols[27], its = 0;
   double vals[27], norm = 0.0, rel = 0.0;

   /* one rank: hypre's sequential headers make MPI_Comm an integer handle
    * (src/utilities/HYPRE_utilities.h:175-177), and any value names the one rank */
   MPI_Comm comm = 0;

   HYPRE_Initialize();

   HYPRE_IJMatrix ijA;
   HYPRE_ParCSRMatrix A;
   HYPRE_IJMatrixCreate(comm, ilower, iupper, ilower, iupper, &ijA);
   HYPRE_IJMatrixSetObjectType(ijA, HYPRE_PARCSR);
   HYPRE_IJMatrixInitialize(ijA);
   for (int i = 0; i < n; i++) {
      int nc = (stencil == 27) ? row_27pt(N, i, cols, vals) : row_7pt(N, i, cols, vals);
      HYPRE_IJMatrixSetValues(ijA, 1, &nc, &i, cols, vals);
   }
   HYPRE_IJMatrixAssemble(ijA);
   HYPRE_IJMatrixGetObject(ijA, (void **) &A);

   HYPRE_IJVector ijb, ijx;
   HYPRE_ParVector b, x;
   HYPRE_IJVectorCreate(comm, ilower, iupper, &ijb);
   HYPRE_IJVectorSetObjectType(ijb, HYPRE_PARCSR);
   HYPRE_IJVectorInitialize(ijb);
   HYPRE_IJVectorAssemble(ijb);
   HYPRE_IJVectorGetObject(ijb, (void **) &b);
   HYPRE_IJVectorCreate(comm, ilower, iupper, &ijx);
   HYPRE_IJVectorSetObjectType(ijx, HYPRE_PARCSR);
   HYPRE_IJVectorInitialize(ijx);
   HYPRE_IJVectorAssemble(ijx);
   HYPRE_IJVectorGetObject(ijx, (void **) &x);
   HYPRE_ParVectorSetRandomValues(b, 22775);
   HYPRE_ParVectorInnerProd(b, b, &norm);
   HYPRE_ParVectorScale(1.0 / sqrt(norm), b);
   HYPRE_ParVectorSetConstantValues(x, 0.0);

   HYPRE_Solver pcg, amg;
   HYPRE_ParCSRPCGCreate(comm, &pcg);
   HYPRE_PCGSetMaxIter(pcg, 1000);
   HYPRE_PCGSetTol(pcg, 1e-8);
   HYPRE_PCGSetTwoNorm(pcg, 1);
   HYPRE_PCGSetPrintLevel(pcg, 0);

   HYPRE_BoomerAMGCreate(&amg);
   HYPRE_BoomerAMGSetCoarsenType(amg, 8);
   HYPRE_BoomerAMGSetInterpType(amg, 6);
   HYPRE_BoomerAMGSetStrongThreshold(amg, 0.25);
   HYPRE_BoomerAMGSetMaxRowSum(amg, mrs);
   HYPRE_BoomerAMGSetPMaxElmts(amg, 4);
   HYPRE_BoomerAMGSetRelaxType(amg, 18);
   HYPRE_BoomerAMGSetNumSweeps(amg, 1);
   HYPRE_BoomerAMGSetMaxIter(amg, 1);
   HYPRE_BoomerAMGSetTol(amg, 0.0);
   HYPRE_ParCSRPCGSetPrecond(pcg, HYPRE_BoomerAMGSolve, HYPRE_BoomerAMGSetup, amg);

   HYPRE_ParCSRPCGSetup(pcg, A, b, x);
   HYPRE_ParCSRPCGSolve(pcg, A, b, x);
   HYPRE_ParCSRPCGGetNumIterations(pcg, &its);
   HYPRE_ParCSRPCGGetFinalRelativeResidualNorm(pcg, &rel);

   /* FP64 true residual r = b - A x */
   HYPRE_IJVector ijr;
   HYPRE_ParVector r;
   double rr = 0.0, bb = 0.0;
   HYPRE_IJVectorCreate(comm, ilower, iupper, &ijr);
   HYPRE_IJVectorSetObjectType(ijr, HYPRE_PARCSR);
   HYPRE_IJVectorInitialize(ijr);
   HYPRE_IJVectorAssemble(ijr);
   HYPRE_IJVectorGetObject(ijr, (void **) &r);
   HYPRE_ParVectorCopy(b, r);
   HYPRE_ParCSRMatrixMatvec(-1.0, A, x, 1.0, r);
   HYPRE_ParVectorInnerProd(r, r, &rr);
   HYPRE_ParVectorInnerProd(b, b, &bb);

   printf("Iterations = %d\n", its);
   printf("Final Relative Residual Norm = %e\n", rel);
   printf("True Relative Residual Norm = %e\n", sqrt(rr / bb));
   printf("error flag = %d\n", HYPRE_GetError());

   HYPRE_IJVectorDestroy(ijr);
   HYPRE_BoomerAMGDestroy(amg);
   HYPRE_ParCSRPCGDestroy(pcg);
   HYPRE_IJVectorDestroy(ijx);
   HYPRE_IJVectorDestroy(ijb);
   HYPRE_IJMatrixDestroy(ijA);
   HYPRE_Finalize();
   return 0;
}
