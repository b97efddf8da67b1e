/*
 * HYPRE_utilities.h -- base types, error flags and library init of the
 * MI355X-native BoomerAMG-PCG library (libhypre_mi355x.so).
 *
 * Mirrors the single-rank, no-MPI build of hypre 3.1.0
 * (src/utilities/HYPRE_utilities.h, src/utilities/mpistubs.h) so that a caller
 * written against hypre compiles unchanged:
 *   HYPRE_Int = int, HYPRE_BigInt = int        (HYPRE_utilities.h:83-84)
 *   HYPRE_Real = HYPRE_Complex = double        (HYPRE_utilities.h:130,156)
 *   MPI_Comm -> hypre_MPI_Comm = HYPRE_Int     (mpistubs.h:34,142,159)
 */
#ifndef HYPRE_UTILITIES_HEADER
#define HYPRE_UTILITIES_HEADER

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int    HYPRE_Int;
typedef int    HYPRE_BigInt;
typedef double HYPRE_Real;
typedef double HYPRE_Complex;

#define HYPRE_MPI_INT    0
#define HYPRE_RELEASE_NAME    "hypre-mi355x"
#define HYPRE_RELEASE_VERSION "3.1.0-mi355x"

/* single-rank MPI stand-ins (src/utilities/mpistubs.h:34-43,142,159) */
#ifndef HYPRE_HAVE_MPI
typedef HYPRE_Int hypre_MPI_Comm;
#define hypre_MPI_COMM_WORLD 0
#define hypre_MPI_COMM_SELF  1
#ifndef MPI_Comm
#define MPI_Comm       hypre_MPI_Comm
#endif
#ifndef MPI_COMM_WORLD
#define MPI_COMM_WORLD hypre_MPI_COMM_WORLD
#endif
#ifndef MPI_COMM_SELF
#define MPI_COMM_SELF  hypre_MPI_COMM_SELF
#endif
#endif

/* error codes (src/utilities/HYPRE_utilities.h:273-277) */
#define HYPRE_ERROR_GENERIC    1
#define HYPRE_ERROR_MEMORY     2
#define HYPRE_ERROR_ARG        4
#define HYPRE_ERROR_CONV     256

/* src/utilities/error.c:154-230 */
HYPRE_Int HYPRE_GetError(void);
HYPRE_Int HYPRE_CheckError(HYPRE_Int hypre_ierr, HYPRE_Int hypre_error_code);
HYPRE_Int HYPRE_GetErrorArg(void);
void      HYPRE_DescribeError(HYPRE_Int hypre_ierr, char *descr);
HYPRE_Int HYPRE_ClearAllErrors(void);
HYPRE_Int HYPRE_ClearError(HYPRE_Int hypre_error_code);

/* src/utilities/general.c:436,485 */
HYPRE_Int HYPRE_Initialize(void);
HYPRE_Int HYPRE_Finalize(void);
HYPRE_Int HYPRE_Initialized(void);
#define HYPRE_Init() HYPRE_Initialize()

/* memory locations / execution policies (src/utilities/HYPRE_utilities.h:460-480).
 * Every object of this library lives in HBM; HOST is accepted for the
 * inputs of the IJ interface, which are always read from host arrays unless
 * HYPRE_IJMatrixInitialize_v2 was given HYPRE_MEMORY_DEVICE. */
typedef enum _HYPRE_MemoryLocation
{
   HYPRE_MEMORY_UNDEFINED = -1,
   HYPRE_MEMORY_HOST,
   HYPRE_MEMORY_DEVICE
} HYPRE_MemoryLocation;

typedef enum _HYPRE_ExecutionPolicy
{
   HYPRE_EXEC_UNDEFINED = -1,
   HYPRE_EXEC_HOST,
   HYPRE_EXEC_DEVICE
} HYPRE_ExecutionPolicy;

HYPRE_Int HYPRE_SetMemoryLocation(HYPRE_MemoryLocation memory_location);
HYPRE_Int HYPRE_GetMemoryLocation(HYPRE_MemoryLocation *memory_location);
HYPRE_Int HYPRE_SetExecutionPolicy(HYPRE_ExecutionPolicy exec_policy);
HYPRE_Int HYPRE_GetExecutionPolicy(HYPRE_ExecutionPolicy *exec_policy);

/* generic object handles (src/utilities/HYPRE_utilities.h:730-753) */
typedef struct hypre_Solver_struct *HYPRE_Solver;
typedef struct hypre_Matrix_struct *HYPRE_Matrix;
typedef struct hypre_Vector_struct *HYPRE_Vector;

typedef HYPRE_Int (*HYPRE_PtrToSolverFcn)(HYPRE_Solver, HYPRE_Matrix,
                                          HYPRE_Vector, HYPRE_Vector);
typedef HYPRE_Int (*HYPRE_PtrToDestroyFcn)(HYPRE_Solver);

#ifdef __cplusplus
}
#endif

#endif
