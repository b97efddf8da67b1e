/*
 * HYPRE.h -- umbrella header of libhypre_mi355x.so (src/HYPRE.h of hypre
 * 3.1.0 plays the same role).
 */
#ifndef HYPRE_HEADER
#define HYPRE_HEADER

#define HYPRE_UNITIALIZED -999
#define HYPRE_PETSC_MAT_PARILUT_SOLVER 222
#define HYPRE_PARILUT                  333
#define HYPRE_STRUCT                  1111
#define HYPRE_SSTRUCT                 3333
#define HYPRE_PARCSR                  5555

#include "HYPRE_utilities.h"
#include "HYPRE_IJ_mv.h"
#include "HYPRE_parcsr_mv.h"
#include "HYPRE_krylov.h"
#include "HYPRE_parcsr_ls.h"

#endif
