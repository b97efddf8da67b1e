/*
 * HYPRE_IJ_mv.h -- the IJ ("index-based") assembly interface, single rank.
 * Entry points replace src/IJ_mv/HYPRE_IJMatrix.c and HYPRE_IJVector.c of
 * hypre 3.1.0 (line numbers below).  Assembled objects live in HBM.
 */
#ifndef HYPRE_IJ_MV_HEADER
#define HYPRE_IJ_MV_HEADER

#include "HYPRE_utilities.h"

#ifdef __cplusplus
extern "C" {
#endif

#ifndef HYPRE_PARCSR
#define HYPRE_PARCSR 5555          /* src/HYPRE.h:33 */
#endif

struct hypre_IJMatrix_struct;
typedef struct hypre_IJMatrix_struct *HYPRE_IJMatrix;
struct hypre_IJVector_struct;
typedef struct hypre_IJVector_struct *HYPRE_IJVector;

/* src/IJ_mv/HYPRE_IJMatrix.c:23 */
HYPRE_Int HYPRE_IJMatrixCreate(MPI_Comm comm, HYPRE_BigInt ilower, HYPRE_BigInt iupper,
                               HYPRE_BigInt jlower, HYPRE_BigInt jupper,
                               HYPRE_IJMatrix *matrix);
/* :177 */
HYPRE_Int HYPRE_IJMatrixDestroy(HYPRE_IJMatrix matrix);
/* :213 */
HYPRE_Int HYPRE_IJMatrixInitialize(HYPRE_IJMatrix matrix);
/* :232 -- DEVICE: SetValues/AddToValues read device pointers */
HYPRE_Int HYPRE_IJMatrixInitialize_v2(HYPRE_IJMatrix matrix,
                                      HYPRE_MemoryLocation memory_location);
/* :271 (overwrite semantics) */
HYPRE_Int HYPRE_IJMatrixSetValues(HYPRE_IJMatrix matrix, HYPRE_Int nrows, HYPRE_Int *ncols,
                                  const HYPRE_BigInt *rows, const HYPRE_BigInt *cols,
                                  const HYPRE_Complex *values);
/* :456 (accumulate semantics) */
HYPRE_Int HYPRE_IJMatrixAddToValues(HYPRE_IJMatrix matrix, HYPRE_Int nrows, HYPRE_Int *ncols,
                                    const HYPRE_BigInt *rows, const HYPRE_BigInt *cols,
                                    const HYPRE_Complex *values);
/* :631 */
HYPRE_Int HYPRE_IJMatrixAssemble(HYPRE_IJMatrix matrix);
/* :890 / :920 */
HYPRE_Int HYPRE_IJMatrixSetObjectType(HYPRE_IJMatrix matrix, HYPRE_Int type);
HYPRE_Int HYPRE_IJMatrixGetObjectType(HYPRE_IJMatrix matrix, HYPRE_Int *type);
/* :1004 -- borrowed HYPRE_ParCSRMatrix owned by the IJ object */
HYPRE_Int HYPRE_IJMatrixGetObject(HYPRE_IJMatrix matrix, void **object);
/* :1024 / :1045 (hints; SetDiagOffdSizes also selects hypre's direct-insert
 * path, whose assembly swaps the diagonal into the first slot) */
HYPRE_Int HYPRE_IJMatrixSetRowSizes(HYPRE_IJMatrix matrix, const HYPRE_Int *sizes);
HYPRE_Int HYPRE_IJMatrixSetDiagOffdSizes(HYPRE_IJMatrix matrix, const HYPRE_Int *diag_sizes,
                                         const HYPRE_Int *offdiag_sizes);
HYPRE_Int HYPRE_IJMatrixSetMaxOffProcElmts(HYPRE_IJMatrix matrix, HYPRE_Int max_off_proc_elmts);
HYPRE_Int HYPRE_IJMatrixSetPrintLevel(HYPRE_IJMatrix matrix, HYPRE_Int print_level);
/* :770 / :695 */
HYPRE_Int HYPRE_IJMatrixGetValues(HYPRE_IJMatrix matrix, HYPRE_Int nrows, HYPRE_Int *ncols,
                                  HYPRE_BigInt *rows, HYPRE_BigInt *cols, HYPRE_Complex *values);
HYPRE_Int HYPRE_IJMatrixGetRowCounts(HYPRE_IJMatrix matrix, HYPRE_Int nrows,
                                     HYPRE_BigInt *rows, HYPRE_Int *ncols);
HYPRE_Int HYPRE_IJMatrixGetLocalRange(HYPRE_IJMatrix matrix, HYPRE_BigInt *ilower,
                                      HYPRE_BigInt *iupper, HYPRE_BigInt *jlower,
                                      HYPRE_BigInt *jupper);
/* :1201 / :1150 -- hypre's text format: "<filename>.00000", a line
 * "ilower iupper jlower jupper", then "<row> <col> <value>" per entry */
HYPRE_Int HYPRE_IJMatrixPrint(HYPRE_IJMatrix matrix, const char *filename);
HYPRE_Int HYPRE_IJMatrixRead(const char *filename, MPI_Comm comm, HYPRE_Int type,
                             HYPRE_IJMatrix *matrix);
/* :1186 -- MatrixMarket coordinate file (real or integer; general or
 * symmetric, the symmetric triangle mirrored); the name is used as given */
HYPRE_Int HYPRE_IJMatrixReadMM(const char *filename, MPI_Comm comm, HYPRE_Int type,
                               HYPRE_IJMatrix *matrix);
/* :1223 / :1168 -- hypre's binary format "<filename>.00000.bin": 11 uint64
 * header words, then row indices, column indices (4 or 8 bytes) and values
 * (4 or 8 bytes) of every entry; written with 4-byte indices and 8-byte values */
HYPRE_Int HYPRE_IJMatrixPrintBinary(HYPRE_IJMatrix matrix, const char *filename);
HYPRE_Int HYPRE_IJMatrixReadBinary(const char *filename, MPI_Comm comm, HYPRE_Int type,
                                   HYPRE_IJMatrix *matrix);

/* src/IJ_mv/HYPRE_IJVector.c:23 */
HYPRE_Int HYPRE_IJVectorCreate(MPI_Comm comm, HYPRE_BigInt jlower, HYPRE_BigInt jupper,
                               HYPRE_IJVector *vector);
/* :150 */
HYPRE_Int HYPRE_IJVectorDestroy(HYPRE_IJVector vector);
/* :280 / :300 */
HYPRE_Int HYPRE_IJVectorInitialize(HYPRE_IJVector vector);
HYPRE_Int HYPRE_IJVectorInitialize_v2(HYPRE_IJVector vector,
                                      HYPRE_MemoryLocation memory_location);
/* :366 (indices may be NULL: consecutive from jlower) */
HYPRE_Int HYPRE_IJVectorSetValues(HYPRE_IJVector vector, HYPRE_Int nvalues,
                                  const HYPRE_BigInt *indices, const HYPRE_Complex *values);
/* :449 */
HYPRE_Int HYPRE_IJVectorAddToValues(HYPRE_IJVector vector, HYPRE_Int nvalues,
                                    const HYPRE_BigInt *indices, const HYPRE_Complex *values);
HYPRE_Int HYPRE_IJVectorAssemble(HYPRE_IJVector vector);
/* :560 */
HYPRE_Int HYPRE_IJVectorGetValues(HYPRE_IJVector vector, HYPRE_Int nvalues,
                                  const HYPRE_BigInt *indices, HYPRE_Complex *values);
HYPRE_Int HYPRE_IJVectorSetObjectType(HYPRE_IJVector vector, HYPRE_Int type);
HYPRE_Int HYPRE_IJVectorGetObjectType(HYPRE_IJVector vector, HYPRE_Int *type);
HYPRE_Int HYPRE_IJVectorGetObject(HYPRE_IJVector vector, void **object);
HYPRE_Int HYPRE_IJVectorGetLocalRange(HYPRE_IJVector vector, HYPRE_BigInt *jlower,
                                      HYPRE_BigInt *jupper);
HYPRE_Int HYPRE_IJVectorSetNumComponents(HYPRE_IJVector vector, HYPRE_Int num_components);
HYPRE_Int HYPRE_IJVectorSetComponent(HYPRE_IJVector vector, HYPRE_Int component);
HYPRE_Int HYPRE_IJVectorSetMaxOffProcElmts(HYPRE_IJVector vector, HYPRE_Int max_off_proc_elmts);
/* src/IJ_mv/HYPRE_IJVector.c:862 / :785 -- "<filename>.00000": "jlower jupper",
 * then "<index> <value>" per entry */
HYPRE_Int HYPRE_IJVectorPrint(HYPRE_IJVector vector, const char *filename);
HYPRE_Int HYPRE_IJVectorRead(const char *filename, MPI_Comm comm, HYPRE_Int type,
                             HYPRE_IJVector *vector);
/* :927 / :849 -- "<filename>.00000.bin": 8 uint64 header words (version,
 * value bytes, partitioning, sizes, components, storage), then the values */
HYPRE_Int HYPRE_IJVectorPrintBinary(HYPRE_IJVector vector, const char *filename);
HYPRE_Int HYPRE_IJVectorReadBinary(const char *filename, MPI_Comm comm, HYPRE_Int type,
                                   HYPRE_IJVector *vector);

#ifdef __cplusplus
}
#endif

#endif
