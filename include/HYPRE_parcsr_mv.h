/*
 * HYPRE_parcsr_mv.h -- ParCSR matrix / vector handles (single rank; the
 * `diag` block is the whole matrix and `offd` is empty).
 * Replaces src/parcsr_mv/HYPRE_parcsr_matrix.c and HYPRE_parcsr_vector.c.
 */
#ifndef HYPRE_PARCSR_MV_HEADER
#define HYPRE_PARCSR_MV_HEADER

#include "HYPRE_utilities.h"

#ifdef __cplusplus
extern "C" {
#endif

struct hypre_ParCSRMatrix_struct;
typedef struct hypre_ParCSRMatrix_struct *HYPRE_ParCSRMatrix;
struct hypre_ParVector_struct;
typedef struct hypre_ParVector_struct *HYPRE_ParVector;

/* src/parcsr_mv/HYPRE_parcsr_matrix.c */
HYPRE_Int HYPRE_ParCSRMatrixDestroy(HYPRE_ParCSRMatrix matrix);
HYPRE_Int HYPRE_ParCSRMatrixGetDims(HYPRE_ParCSRMatrix matrix, HYPRE_BigInt *M, HYPRE_BigInt *N);
HYPRE_Int HYPRE_ParCSRMatrixGetLocalRange(HYPRE_ParCSRMatrix matrix, HYPRE_BigInt *row_start,
                                          HYPRE_BigInt *row_end, HYPRE_BigInt *col_start,
                                          HYPRE_BigInt *col_end);
HYPRE_Int HYPRE_ParCSRMatrixGetRow(HYPRE_ParCSRMatrix matrix, HYPRE_BigInt row,
                                   HYPRE_Int *size, HYPRE_BigInt **col_ind,
                                   HYPRE_Complex **values);
HYPRE_Int HYPRE_ParCSRMatrixRestoreRow(HYPRE_ParCSRMatrix matrix, HYPRE_BigInt row,
                                       HYPRE_Int *size, HYPRE_BigInt **col_ind,
                                       HYPRE_Complex **values);
/* y = alpha*A*x + beta*y  (HYPRE_parcsr_matrix.c HYPRE_ParCSRMatrixMatvec) */
HYPRE_Int HYPRE_ParCSRMatrixMatvec(HYPRE_Complex alpha, HYPRE_ParCSRMatrix A,
                                   HYPRE_ParVector x, HYPRE_Complex beta, HYPRE_ParVector y);
/* y = alpha*A^T*x + beta*y */
HYPRE_Int HYPRE_ParCSRMatrixMatvecT(HYPRE_Complex alpha, HYPRE_ParCSRMatrix A,
                                    HYPRE_ParVector x, HYPRE_Complex beta, HYPRE_ParVector y);

/* src/parcsr_mv/HYPRE_parcsr_vector.c */
HYPRE_Int HYPRE_ParVectorCreate(MPI_Comm comm, HYPRE_BigInt global_size,
                                HYPRE_BigInt *partitioning, HYPRE_ParVector *vector);
HYPRE_Int HYPRE_ParVectorInitialize(HYPRE_ParVector vector);
HYPRE_Int HYPRE_ParVectorDestroy(HYPRE_ParVector vector);
HYPRE_Int HYPRE_ParVectorCopy(HYPRE_ParVector x, HYPRE_ParVector y);
HYPRE_Int HYPRE_ParVectorSetConstantValues(HYPRE_ParVector vector, HYPRE_Complex value);
HYPRE_Int HYPRE_ParVectorSetRandomValues(HYPRE_ParVector vector, HYPRE_Int seed);
HYPRE_Int HYPRE_ParVectorScale(HYPRE_Complex value, HYPRE_ParVector x);
HYPRE_Int HYPRE_ParVectorAxpy(HYPRE_Complex alpha, HYPRE_ParVector x, HYPRE_ParVector y);
HYPRE_Int HYPRE_ParVectorInnerProd(HYPRE_ParVector x, HYPRE_ParVector y, HYPRE_Real *prod);
HYPRE_Int HYPRE_ParVectorGetValues(HYPRE_ParVector vector, HYPRE_Int num_values,
                                   HYPRE_BigInt *indices, HYPRE_Complex *values);

#ifdef __cplusplus
}
#endif

#endif
