/*
 * CPU oracle for the BoomerAMG-PCG hot path -- TEST INFRASTRUCTURE ONLY.
 * Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
 * the checker; never linked into or called by the product library.
 *
 * Restatement of hypre 3.1.0's host path (single rank: every `offd` block is
 * empty).  Loops keep the reference's traversal and floating-point operation
 * order so that patterns and values are reproduced exactly; the file is
 * compiled with -ffp-contract=off.  Citations are file:line into the hypre
 * source tree (src/...).
 *
 * Row loops whose rows are independent run under OpenMP (strength, PMIS
 * measure and rounds, ext+i, truncation, transpose, RAP, l1 norms, SpMV,
 * l1-Jacobi): every row is still computed by one thread with the serial
 * arithmetic, rows that build a CSR matrix are counted first or buffered per
 * thread in row order, and the first-seen markers are private to a thread
 * whose rows come in ascending order, so the results are those of the serial
 * loops bit for bit, for any thread count.  Dot products stay serial.
 */
#include "amg_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int orc_threads(void)
{
#ifdef _OPENMP
   return omp_get_max_threads();
#else
   return 1;
#endif
}

static int orc_thread_id(void)
{
#ifdef _OPENMP
   return omp_get_thread_num();
#else
   return 0;
#endif
}

/* threads of the current team (1 outside a parallel region) */
static int orc_team(void)
{
#ifdef _OPENMP
   return omp_get_num_threads();
#else
   return 1;
#endif
}

/* threads the oracle's parallel loops use */
int orc_num_threads(void) { return orc_threads(); }

/* rows [*r0, *r1) of thread t of nt: contiguous, in thread order */
static void orc_chunk(int n, int nt, int t, int *r0, int *r1)
{
   long q = n / nt, rem = n % nt;
   *r0 = (int) (t * q + (t < rem ? t : rem));
   *r1 = *r0 + (int) q + (t < rem ? 1 : 0);
}

/* in-place exclusive prefix sum: c[0..n] holds counts in c[1..n] -> offsets */
static void orc_prefix(int *c, int n)
{
   int i;
   c[0] = 0;
   for (i = 0; i < n; i++) { c[i + 1] += c[i]; }
}

#define ORC_MAX(a, b) (((a) < (b)) ? (b) : (a))
#define ORC_MIN(a, b) (((a) < (b)) ? (a) : (b))

#define C_PT  1
#define F_PT -1
#define Z_PT -2
#define SF_PT -3

static void *xmalloc(size_t n)
{
   void *p = malloc(n ? n : 1);
   if (!p) { abort(); }
   return p;
}
static void *xcalloc(size_t n, size_t s)
{
   void *p = calloc(n ? n : 1, s ? s : 1);
   if (!p) { abort(); }
   return p;
}

void orc_csr_free(orc_csr *M)
{
   if (!M) { return; }
   free(M->ia); free(M->ja); free(M->a);
   M->ia = NULL; M->ja = NULL; M->a = NULL;
   M->nrows = M->ncols = 0;
}

/* ======================================================================
 * Park-Miller minimal standard generator (src/utilities/random.c:36-103):
 * a = 16807, m = 2^31-1, Schrage factorisation q = 127773, r = 2836.
 * ====================================================================== */
static int orc_Seed = 13579;

void orc_seed_rand(int seed)                   /* random.c hypre_SeedRand */
{
   if (seed < 1) { seed = 1; }
   else if (seed >= 2147483647) { seed = 2147483647 - 1; }
   orc_Seed = seed;
}

int orc_rand_i(void)                           /* random.c hypre_RandI */
{
   int low, high, test;
   high = orc_Seed / 127773;
   low  = orc_Seed % 127773;
   test = 16807 * low - 2836 * high;
   if (test > 0) { orc_Seed = test; }
   else          { orc_Seed = test + 2147483647; }
   return orc_Seed;
}

double orc_rand(void)                          /* random.c hypre_Rand */
{
   return ((double) orc_rand_i()) / ((double) 2147483647);
}

void orc_rand_stream(int seed, int n, double *out)
{
   int i;
   orc_seed_rand(seed);
   for (i = 0; i < n; i++) { out[i] = orc_rand(); }
}

/* ======================================================================
 * Problem generators
 * ====================================================================== */
static int csr_alloc(orc_csr *A, int n, int ncols, long nnz)
{
   A->nrows = n; A->ncols = ncols;
   A->ia = (int *) xmalloc(sizeof(int) * (size_t)(n + 1));
   A->ja = (int *) xmalloc(sizeof(int) * (size_t) nnz);
   A->a  = (double *) xmalloc(sizeof(double) * (size_t) nnz);
   return 0;
}

/* src/parcsr_ls/par_laplace.c GenerateLaplacian (P=Q=R=1), stencil values as
 * in src/test/ij.c:10454-10470 BuildParLaplacian. Row layout at
 * par_laplace.c:158-250: diag, -z, -y, -x, +x, +y, +z. */
int orc_gen_laplace7(int nx, int ny, int nz, double cx, double cy, double cz,
                     orc_csr *A)
{
   double v0 = 0.0, v1 = -cx, v2 = -cy, v3 = -cz;
   long n = (long) nx * ny * nz, nnz = 0;
   int ix, iy, iz, row = 0;
   if (nx > 1) { v0 += 2.0 * cx; }
   if (ny > 1) { v0 += 2.0 * cy; }
   if (nz > 1) { v0 += 2.0 * cz; }
   csr_alloc(A, (int) n, (int) n, 7 * n);
   A->ia[0] = 0;
   for (iz = 0; iz < nz; iz++)
      for (iy = 0; iy < ny; iy++)
         for (ix = 0; ix < nx; ix++)
         {
            A->ja[nnz] = row; A->a[nnz++] = v0;
            if (iz > 0)      { A->ja[nnz] = row - nx * ny; A->a[nnz++] = v3; }
            if (iy > 0)      { A->ja[nnz] = row - nx;      A->a[nnz++] = v2; }
            if (ix > 0)      { A->ja[nnz] = row - 1;       A->a[nnz++] = v1; }
            if (ix + 1 < nx) { A->ja[nnz] = row + 1;       A->a[nnz++] = v1; }
            if (iy + 1 < ny) { A->ja[nnz] = row + nx;      A->a[nnz++] = v2; }
            if (iz + 1 < nz) { A->ja[nnz] = row + nx * ny; A->a[nnz++] = v3; }
            A->ia[++row] = (int) nnz;
         }
   return 0;
}

/* src/parcsr_ls/par_laplace_27pt.c GenerateLaplacian27pt (one rank): diag
 * first, then the 26 neighbours in ascending column order (the z-1 plane, the
 * own plane, the z+1 plane, each row-major). Values: src/test/ij.c:11903-11912 */
int orc_gen_laplace27(int nx, int ny, int nz, orc_csr *A)
{
   double v0 = 26.0, v1 = -1.0;
   long n = (long) nx * ny * nz, nnz = 0;
   int ix, iy, iz, dx, dy, dz, row = 0;
   if (nx == 1 || ny == 1 || nz == 1) { v0 = 8.0; }
   if (nx * ny == 1 || nx * nz == 1 || ny * nz == 1) { v0 = 2.0; }
   csr_alloc(A, (int) n, (int) n, 27 * n);
   A->ia[0] = 0;
   for (iz = 0; iz < nz; iz++)
      for (iy = 0; iy < ny; iy++)
         for (ix = 0; ix < nx; ix++)
         {
            A->ja[nnz] = row; A->a[nnz++] = v0;
            for (dz = -1; dz <= 1; dz++)
               for (dy = -1; dy <= 1; dy++)
                  for (dx = -1; dx <= 1; dx++)
                  {
                     int jx = ix + dx, jy = iy + dy, jz = iz + dz;
                     if (!dx && !dy && !dz) { continue; }
                     if (jx < 0 || jx >= nx || jy < 0 || jy >= ny || jz < 0 || jz >= nz) { continue; }
                     A->ja[nnz] = row + dz * nx * ny + dy * nx + dx;
                     A->a[nnz++] = v1;
                  }
            A->ia[++row] = (int) nnz;
         }
   return 0;
}

static double *cell_kappa(int N, int cell, double lo, double hi, int seed, int *nc_out)
{
   int nc = (N + cell - 1) / cell, c;
   long ncell = (long) nc * nc * nc;
   double *kap = (double *) xmalloc(sizeof(double) * (size_t) ncell);
   orc_seed_rand(seed);
   for (c = 0; c < ncell; c++)
   {
      double u = orc_rand();
      kap[c] = pow(10.0, lo + (hi - lo) * u);
   }
   *nc_out = nc;
   return kap;
}

static double harm(double a, double b) { return 2.0 * a * b / (a + b); }

static int gen_vardiff(int N, int cell, double lo, double hi, int seed, int stencil27,
                       orc_csr *A)
{
   int nc, ix, iy, iz, dx, dy, dz, row = 0;
   long n = (long) N * N * N, nnz = 0;
   double *kap = cell_kappa(N, cell, lo, hi, seed, &nc);
   double *node = (double *) xmalloc(sizeof(double) * (size_t) n);
   for (iz = 0; iz < N; iz++)
      for (iy = 0; iy < N; iy++)
         for (ix = 0; ix < N; ix++)
         {
            node[(long) iz * N * N + (long) iy * N + ix] =
               kap[((long) (iz / cell) * nc + iy / cell) * nc + ix / cell];
         }
   free(kap);
   csr_alloc(A, (int) n, (int) n, (stencil27 ? 27 : 7) * n);
   A->ia[0] = 0;
   for (iz = 0; iz < N; iz++)
      for (iy = 0; iy < N; iy++)
         for (ix = 0; ix < N; ix++)
         {
            double kp = node[row], diag = 0.0;
            long dpos = nnz++;
            A->ja[dpos] = row;
            for (dz = -1; dz <= 1; dz++)
               for (dy = -1; dy <= 1; dy++)
                  for (dx = -1; dx <= 1; dx++)
                  {
                     int jx = ix + dx, jy = iy + dy, jz = iz + dz, col;
                     double k;
                     if (!dx && !dy && !dz) { continue; }
                     if (!stencil27 && (abs(dx) + abs(dy) + abs(dz)) != 1) { continue; }
                     if (jx < 0 || jx >= N || jy < 0 || jy >= N || jz < 0 || jz >= N)
                     {
                        diag += kp;             /* Dirichlet face: own kappa */
                        continue;
                     }
                     col = row + dz * N * N + dy * N + dx;
                     k = harm(kp, node[col]);
                     diag += k;
                     A->ja[nnz] = col; A->a[nnz++] = -k;
                  }
            A->a[dpos] = diag;
            A->ia[++row] = (int) nnz;
         }
   free(node);
   return 0;
}

int orc_gen_vardiff7(int N, int cell, double lo, double hi, int seed, orc_csr *A)
{
   return gen_vardiff(N, cell, lo, hi, seed, 0, A);
}

int orc_gen_vardiff27(int N, int cell, double lo, double hi, int seed, orc_csr *A)
{
   return gen_vardiff(N, cell, lo, hi, seed, 1, A);
}

int orc_gen_aniso7(int N, double eps, double theta, orc_csr *A)
{
   double c = cos(theta), s = sin(theta);
   double kxx = c * c + eps * s * s, kyy = s * s + eps * c * c, kzz = 1.0;
   return orc_gen_laplace7(N, N, N, kxx, kyy, kzz, A);
}

/* ======================================================================
 * Strength matrix: src/parcsr_ls/par_strength.c:75-525 (CreateSHost),
 * num_functions == 1, empty offd.  S keeps A's column order without the
 * (first) diagonal entry.
 * ====================================================================== */
/* one row of S: strong columns in A's order into out (when non-NULL); returns
 * their number */
static int strength_row(const orc_csr *A, int i, double theta, double max_row_sum, int *out)
{
   const int *ia = A->ia, *ja = A->ja;
   const double *a = A->a;
   int jA, cnt = 0;
   double diag = a[ia[i]], row_scale = 0.0, row_sum = diag;
   if (diag < 0)
   {
      for (jA = ia[i] + 1; jA < ia[i + 1]; jA++)
      {
         row_scale = ORC_MAX(row_scale, a[jA]);
         row_sum += a[jA];
      }
   }
   else
   {
      for (jA = ia[i] + 1; jA < ia[i + 1]; jA++)
      {
         row_scale = ORC_MIN(row_scale, a[jA]);
         row_sum += a[jA];
      }
   }
   /* par_strength.c:362 all dependencies weak */
   if ((fabs(row_sum) > fabs(diag) * max_row_sum) && (max_row_sum < 1.0))
   {
      /* no strong entries */
   }
   else if (diag < 0)
   {
      for (jA = ia[i] + 1; jA < ia[i + 1]; jA++)
      {
         if (!(a[jA] <= theta * row_scale)) { if (out) { out[cnt] = ja[jA]; } cnt++; }
      }
   }
   else
   {
      for (jA = ia[i] + 1; jA < ia[i + 1]; jA++)
      {
         if (!(a[jA] >= theta * row_scale)) { if (out) { out[cnt] = ja[jA]; } cnt++; }
      }
   }
   return cnt;
}

int orc_strength(const orc_csr *A, double theta, double max_row_sum, orc_csr *S)
{
   int n = A->nrows, i;
   S->nrows = n; S->ncols = A->ncols;
   S->ia = (int *) xmalloc(sizeof(int) * (size_t)(n + 1));
   S->a = NULL;
#pragma omp parallel for schedule(static)
   for (i = 0; i < n; i++) { S->ia[i + 1] = strength_row(A, i, theta, max_row_sum, NULL); }
   orc_prefix(S->ia, n);
   S->ja = (int *) xmalloc(sizeof(int) * (size_t) (S->ia[n] > 0 ? S->ia[n] : 1));
#pragma omp parallel for schedule(static)
   for (i = 0; i < n; i++) { strength_row(A, i, theta, max_row_sum, S->ja + S->ia[i]); }
   return 0;
}

/* ======================================================================
 * PMIS: src/parcsr_ls/par_coarsen.c:2094-2800 (CoarsenPMISHost, CF_init 0)
 * ====================================================================== */
int orc_pmis_measure(const orc_csr *S, int seed, double *measure)
{
   int n = S->nrows, i;
   /* par_coarsen.c:2248-2262 column counts of S */
   {
      /* integer counts: the same values as adding 1.0 per entry in any order */
      int *cnt = (int *) xcalloc((size_t) (n > 0 ? n : 1), sizeof(int));
      const long nnz = S->ia[n];
      long k;
#pragma omp parallel for schedule(static)
      for (k = 0; k < nnz; k++)
      {
#pragma omp atomic update
         cnt[S->ja[k]]++;
      }
      for (i = 0; i < n; i++) { measure[i] = (double) cnt[i]; }
      free(cnt);
   }
   /* par_indepset.c:50-75 hypre_BoomerAMGIndepSetInit, seed 2747 + rank */
   orc_seed_rand(seed);
   for (i = 0; i < n; i++) { measure[i] += orc_rand(); }
   return 0;
}

/* CF_init 0 or 3: CF is set here; CF_init 1 (HMIS): CF holds the first
 * pass's split on entry and its C points are the first independent set */
static int pmis_core(const orc_csr *S, const double *measure_in, int *CF, int cf_init)
{
   int n = S->nrows, i, j, jS, ig, cnt, graph_size, iter = 0;
   const int *Si = S->ia, *Sj = S->ja;
   double *measure = (double *) xmalloc(sizeof(double) * (size_t) (n > 0 ? n : 1));
   int *graph = (int *) xmalloc(sizeof(int) * (size_t) (n > 0 ? n : 1));
   int *graph2 = (int *) xmalloc(sizeof(int) * (size_t) (n > 0 ? n : 1));
   memcpy(measure, measure_in, sizeof(double) * (size_t) n);

   cnt = 0;
   if (cf_init == 1)
   {
      /* par_coarsen.c:2346-2374 (one rank: no offd entries) */
      for (i = 0; i < n; i++)
      {
         if (CF[i] != SF_PT)
         {
            if (CF[i] == -1) { CF[i] = 0; }
            if (CF[i] == Z_PT)
            {
               if (measure[i] >= 1.0 || Si[i + 1] - Si[i] > 0)
               {
                  CF[i] = 0;
                  graph[cnt++] = i;
               }
               else
               {
                  CF[i] = F_PT;
               }
            }
            else
            {
               graph[cnt++] = i;
            }
         }
         else
         {
            measure[i] = 0;
         }
      }
   }
   else
   {
      /* par_coarsen.c:2386-2407 (CF_init 0, or 3 for the second coarsening
       * of aggressive levels: rows without strong connections become C) */
      for (i = 0; i < n; i++)
      {
         CF[i] = 0;
         if (Si[i + 1] - Si[i] == 0)
         {
            CF[i] = (cf_init == 3) ? C_PT : SF_PT;
            measure[i] = 0;
         }
         else
         {
            graph[cnt++] = i;
         }
      }
   }
   graph_size = cnt;

   while (1)   /* par_coarsen.c:2460 THE INDEPENDENT SET COARSENING LOOP */
   {
      if (graph_size == 0) { break; }
      /* par_coarsen.c:2487: with CF_init 1 the first independent set is the
       * first pass's C points */
      if (!cf_init || iter)
      {
      /* par_coarsen.c:2485-2550 pick the independent set.  Every write of
       * the second loop stores 0, so the outcome does not depend on the
       * order of the rows. */
#pragma omp parallel for private(i) schedule(static)
      for (ig = 0; ig < graph_size; ig++)
      {
         i = graph[ig];
         if (measure[i] > 1) { CF[i] = 1; }
      }
#pragma omp parallel for private(i, j, jS) schedule(static)
      for (ig = 0; ig < graph_size; ig++)
      {
         i = graph[ig];
         if (measure[i] > 1)
         {
            for (jS = Si[i]; jS < Si[i + 1]; jS++)
            {
               j = Sj[jS];
               if (measure[j] > 1)
               {
                  if (measure[i] > measure[j])
                  {
#pragma omp atomic write
                     CF[j] = 0;
                  }
                  else if (measure[j] > measure[i])
                  {
#pragma omp atomic write
                     CF[i] = 0;
                  }
               }
            }
         }
      }
      }
      iter++;
      /* par_coarsen.c:2603-2660 set C and F points.  Which points have
       * CF > 0 does not change in this loop (C points stay 1, the others
       * become -1 or stay 0), so neighbours read the same either way. */
#pragma omp parallel for private(i, j, jS) schedule(static)
      for (ig = 0; ig < graph_size; ig++)
      {
         int c;
         i = graph[ig];
#pragma omp atomic read
         c = CF[i];
         if (measure[i] < 1) { c = F_PT; }
         if (c > 0)
         {
            c = C_PT;
         }
         else
         {
            for (jS = Si[i]; jS < Si[i + 1]; jS++)
            {
               int cj;
               j = Sj[jS];
#pragma omp atomic read
               cj = CF[j];
               if (cj > 0) { c = F_PT; }
            }
         }
#pragma omp atomic write
         CF[i] = c;
      }
      /* par_coarsen.c:2690-2750 update the subgraph */
      cnt = 0;
      for (ig = 0; ig < graph_size; ig++)
      {
         i = graph[ig];
         if (CF[i] != 0) { measure[i] = 0; }
         else            { graph2[cnt++] = i; }
      }
      graph_size = cnt;
      { int *t = graph; graph = graph2; graph2 = t; }
   }
   free(measure); free(graph); free(graph2);
   return iter;
}

int orc_pmis_with_measure(const orc_csr *S, const double *measure_in, int *CF)
{
   return pmis_core(S, measure_in, CF, 0);
}

/* ======================================================================
 * HMIS (coarsen_type 10): src/parcsr_ls/par_coarsen.c:2842-2906
 * hypre_BoomerAMGCoarsenHMIS = the Ruge-Stueben first pass
 * (hypre_BoomerAMGCoarsenRuge, par_coarsen.c:911, called with coarsen_type
 * 10: undecided points that no point depends on become Z_PT, :1119-1123,
 * and the routine returns after the first pass, :1384-1391), then PMIS with
 * CF_init 1 (:2904).  One rank, measure_type 0, cut_factor 0.
 *
 * The first pass keeps undecided points on lists of equal measure ordered
 * by decreasing measure (src/utilities/amg_linklist.c): a point joins the
 * tail of its measure's list, and the next C point is the head of the
 * list of largest measure.  Restated here with that list of lists.
 * ====================================================================== */
#define LIST_HEAD_MARK -1
#define LIST_TAIL_MARK -2

typedef struct lol_elt
{
   int data, head, tail;
   struct lol_elt *next, *prev;
} lol_elt;

typedef struct
{
   lol_elt *head, *tail;
   int *lists, *where;      /* next / previous point of each point on its list */
} lol_t;

static lol_elt *lol_new(int measure, int index, lol_t *L)
{
   lol_elt *e = (lol_elt *) xcalloc(1, sizeof(lol_elt));
   e->data = measure;
   e->head = index;
   e->tail = index;
   L->lists[index] = LIST_TAIL_MARK;
   L->where[index] = LIST_HEAD_MARK;
   return e;
}

/* amg_linklist.c:168 hypre_enter_on_lists */
static void lol_enter(lol_t *L, int measure, int index)
{
   lol_elt *p = L->head, *e;
   if (!p)
   {
      e = lol_new(measure, index, L);
      L->head = L->tail = e;
      return;
   }
   for (; p; p = p->next)
   {
      if (measure > p->data)
      {
         /* a new list in front of p */
         e = lol_new(measure, index, L);
         e->next = p;
         e->prev = p->prev;
         if (p->prev) { p->prev->next = e; } else { L->head = e; }
         p->prev = e;
         return;
      }
      if (measure == p->data)
      {
         /* append at the tail of p */
         L->lists[p->tail] = index;
         L->where[index] = p->tail;
         L->lists[index] = LIST_TAIL_MARK;
         p->tail = index;
         return;
      }
   }
   /* smaller than every list: a new last list */
   e = lol_new(measure, index, L);
   e->prev = L->tail;
   L->tail->next = e;
   L->tail = e;
}

/* amg_linklist.c:40 hypre_remove_point */
static void lol_remove(lol_t *L, int measure, int index)
{
   lol_elt *p;
   for (p = L->head; p; p = p->next)
   {
      if (p->data != measure) { continue; }
      if (p->head == index && p->tail == index)
      {
         /* the point is alone on its list: drop the list */
         if (p->prev) { p->prev->next = p->next; } else { L->head = p->next; }
         if (p->next) { p->next->prev = p->prev; } else { L->tail = p->prev; }
         free(p);
      }
      else if (p->head == index)
      {
         p->head = L->lists[index];
         L->where[L->lists[index]] = LIST_HEAD_MARK;
      }
      else if (p->tail == index)
      {
         p->tail = L->where[index];
         L->lists[L->where[index]] = LIST_TAIL_MARK;
      }
      else
      {
         L->lists[L->where[index]] = L->lists[index];
         L->where[L->lists[index]] = L->where[index];
      }
      return;
   }
   abort();   /* "No such list!" */
}

int orc_rs_first_pass(const orc_csr *S, int *CF)
{
   const int n = S->nrows;
   const int *Si = S->ia, *Sj = S->ja;
   const int nnz = Si[n];
   int *STi = (int *) xcalloc((size_t) n + 1, sizeof(int));
   int *STj = (int *) xmalloc(sizeof(int) * (size_t) (nnz > 0 ? nnz : 1));
   int *measure = (int *) xmalloc(sizeof(int) * (size_t) (n > 0 ? n : 1));
   lol_t L;
   int i, j, k, num_left = 0;
   L.head = L.tail = NULL;
   L.lists = (int *) xcalloc((size_t) (n > 0 ? n : 1), sizeof(int));
   L.where = (int *) xcalloc((size_t) (n > 0 ? n : 1), sizeof(int));

   /* S^T by a stable counting sort (par_coarsen.c:1066-1095) */
   for (i = 0; i < nnz; i++) { STi[Sj[i] + 1]++; }
   for (i = 0; i < n; i++) { STi[i + 1] += STi[i]; }
   {
      int *pos = (int *) xmalloc(sizeof(int) * (size_t) (n > 0 ? n : 1));
      memcpy(pos, STi, sizeof(int) * (size_t) n);
      for (i = 0; i < n; i++)
      {
         for (j = Si[i]; j < Si[i + 1]; j++) { STj[pos[Sj[j]]++] = i; }
      }
      free(pos);
   }
   /* measures: the number of points that depend on i (:1107-1112) */
   for (i = 0; i < n; i++) { measure[i] = STi[i + 1] - STi[i]; }

   /* :1166-1190 */
   for (j = 0; j < n; j++) { CF[j] = 0; }
   for (j = 0; j < n; j++)
   {
      if (Si[j + 1] - Si[j] == 0)
      {
         CF[j] = SF_PT;
         measure[j] = 0;
      }
      else
      {
         num_left++;
      }
   }
   /* :1212-1256 lists, and points nobody depends on become Z_PT */
   for (j = 0; j < n; j++)
   {
      const int m = measure[j];
      if (CF[j] == SF_PT) { continue; }
      if (m > 0)
      {
         lol_enter(&L, m, j);
         continue;
      }
      CF[j] = Z_PT;
      for (k = Si[j]; k < Si[j + 1]; k++)
      {
         const int nb = Sj[k];
         if (CF[nb] == SF_PT) { continue; }
         if (nb < j)
         {
            if (measure[nb] > 0) { lol_remove(&L, measure[nb], nb); }
            lol_enter(&L, ++measure[nb], nb);
         }
         else
         {
            ++measure[nb];
         }
      }
      --num_left;
   }
   /* :1278-1358 the first pass proper */
   while (num_left > 0)
   {
      const int index = L.head->head;
      const int m = measure[index];
      CF[index] = C_PT;
      measure[index] = 0;
      --num_left;
      lol_remove(&L, m, index);
      for (j = STi[index]; j < STi[index + 1]; j++)
      {
         const int nb = STj[j];
         if (CF[nb] != 0) { continue; }
         CF[nb] = F_PT;
         lol_remove(&L, measure[nb], nb);
         --num_left;
         for (k = Si[nb]; k < Si[nb + 1]; k++)
         {
            const int nb2 = Sj[k];
            if (CF[nb2] != 0) { continue; }
            lol_remove(&L, measure[nb2], nb2);
            lol_enter(&L, ++measure[nb2], nb2);
         }
      }
      for (j = Si[index]; j < Si[index + 1]; j++)
      {
         const int nb = Sj[j];
         int mm;
         if (CF[nb] != 0) { continue; }
         mm = measure[nb];
         lol_remove(&L, mm, nb);
         measure[nb] = --mm;
         if (mm > 0)
         {
            lol_enter(&L, mm, nb);
         }
         else
         {
            CF[nb] = F_PT;
            --num_left;
            for (k = Si[nb]; k < Si[nb + 1]; k++)
            {
               const int nb2 = Sj[k];
               if (CF[nb2] != 0) { continue; }
               lol_remove(&L, measure[nb2], nb2);
               lol_enter(&L, ++measure[nb2], nb2);
            }
         }
      }
   }
   while (L.head)
   {
      lol_elt *nx = L.head->next;
      free(L.head);
      L.head = nx;
   }
   free(STi); free(STj); free(measure); free(L.lists); free(L.where);
   return 0;
}

int orc_hmis_with_measure(const orc_csr *S, const double *measure_in, int *CF)
{
   orc_rs_first_pass(S, CF);
   return pmis_core(S, measure_in, CF, 1);
}

int orc_hmis(const orc_csr *S, int *CF)
{
   double *m = (double *) xmalloc(sizeof(double) * (size_t) (S->nrows > 0 ? S->nrows : 1));
   int it;
   orc_pmis_measure(S, 2747, m);
   it = orc_hmis_with_measure(S, m, CF);
   free(m);
   return it;
}

int orc_pmis(const orc_csr *S, int *CF)
{
   double *m = (double *) xmalloc(sizeof(double) * (size_t) S->nrows);
   int it;
   orc_pmis_measure(S, 2747, m);
   it = orc_pmis_with_measure(S, m, CF);
   free(m);
   return it;
}

/* ======================================================================
 * Aggressive coarsening (agg_num_levels), one rank, num_paths 1:
 * src/parcsr_ls/par_amg_setup.c:1311-1352 (second coarsening of the C
 * points over the distance-two strength graph) and :1799-1824 (C/F
 * correction, modified multipass interpolation, agg_interp_type 8).
 * ====================================================================== */

/* src/parcsr_ls/par_strength.c:1795 hypre_BoomerAMGCreate2ndSHost with
 * num_paths 1: row ic (the ic-th C point i1 of CF, row order) holds, in
 * first-seen order, the coarse indices of the C points reached from i1 by one
 * strong connection (:2440-2449) or by two through any point (:2451-2465),
 * except ic itself.  Pattern only.  C points whose row is empty are marked 2
 * in CF (:2700-2722); the second pass writes the same order as the counting
 * pass (:2740-2790). */
int orc_create_2nd_s(const orc_csr *S, int *CF, orc_csr *S2)
{
   const int n = S->nrows;
   int i, nc = 0, ic;
   int *f2c = (int *) xmalloc(sizeof(int) * (size_t) (n > 0 ? n : 1));
   int *c2f = (int *) xmalloc(sizeof(int) * (size_t) (n > 0 ? n : 1));
   int *marker;
   long cap = 1024, cnt = 0;
   for (i = 0; i < n; i++)
   {
      if (CF[i] > 0) { f2c[i] = nc; c2f[nc++] = i; }
      else { f2c[i] = -1; }
   }
   marker = (int *) xmalloc(sizeof(int) * (size_t) (nc > 0 ? nc : 1));
   for (i = 0; i < nc; i++) { marker[i] = -1; }
   S2->nrows = nc; S2->ncols = nc; S2->a = NULL;
   S2->ia = (int *) xmalloc(sizeof(int) * (size_t) (nc + 1));
   S2->ja = (int *) xmalloc(sizeof(int) * (size_t) cap);
   for (ic = 0; ic < nc; ic++)
   {
      const long row_begin = cnt;
      const int i1 = c2f[ic];
      int jj1, jj2;
      S2->ia[ic] = (int) cnt;
      for (jj1 = S->ia[i1]; jj1 < S->ia[i1 + 1]; jj1++)
      {
         const int i2 = S->ja[jj1];
         if (cnt + 1 + (S->ia[i2 + 1] - S->ia[i2]) > cap)
         {
            cap = 2 * cap + (S->ia[i2 + 1] - S->ia[i2]) + 1;
            S2->ja = (int *) realloc(S2->ja, sizeof(int) * (size_t) cap);
         }
         if (CF[i2] > 0)
         {
            const int idx = f2c[i2];
            if (marker[idx] < row_begin) { marker[idx] = (int) cnt; S2->ja[cnt++] = idx; }
         }
         for (jj2 = S->ia[i2]; jj2 < S->ia[i2 + 1]; jj2++)
         {
            const int i3 = S->ja[jj2];
            if (CF[i3] > 0)
            {
               const int idx = f2c[i3];
               if (idx != ic && marker[idx] < row_begin) { marker[idx] = (int) cnt; S2->ja[cnt++] = idx; }
            }
         }
      }
   }
   S2->ia[nc] = (int) cnt;
   for (ic = 0; ic < nc; ic++)
   {
      if (S2->ia[ic + 1] == S2->ia[ic]) { CF[c2f[ic]] = 2; }
   }
   free(f2c); free(c2f); free(marker);
   return nc;
}

/* the second coarsening of the aggressive branch for coarsen_type 8:
 * hypre_BoomerAMGCoarsenPMIS(S2, S2, 3, ...) (par_amg_setup.c:1324),
 * i.e. CF_init 3 (par_coarsen.c:2386-2407: rows without strong connections
 * become C points; :2487: no independent set is picked in the first round) */
int orc_pmis_cf3_with_measure(const orc_csr *S, const double *measure_in, int *CF)
{
   return pmis_core(S, measure_in, CF, 3);
}

int orc_pmis_cf3(const orc_csr *S, int *CF)
{
   double *m = (double *) xmalloc(sizeof(double) * (size_t) (S->nrows > 0 ? S->nrows : 1));
   int it;
   orc_pmis_measure(S, 2747, m);
   it = pmis_core(S, m, CF, 3);
   free(m);
   return it;
}

/* src/parcsr_ls/par_strength.c:3068 hypre_BoomerAMGCorrectCFMarkerHost:
 * the C points of CF (in row order) take the second coarsening's marks;
 * points marked 2 (empty second-strength row) become plain C points. */
void orc_correct_cf_marker(int n, int *CF, const int *CFN)
{
   int i, cnt = 0;
   for (i = 0; i < n; i++)
   {
      if (CF[i] > 0)
      {
         if (CF[i] == 1) { CF[i] = CFN[cnt++]; }
         else { CF[i] = 1; cnt++; }
      }
   }
}

/* src/parcsr_mv/par_csr_matop.c:348 hypre_ParMatmul on one rank: row i of
 * C = A B in first-seen column order; the first product of a column is
 * stored, later ones added in traversal order (:1107-1140).  When C is
 * square, each row first gets its diagonal position with value 0
 * (:476-479, :691-697). */
int orc_matmul(const orc_csr *A, const orc_csr *B, orc_csr *C)
{
   const int n = A->nrows, m = B->ncols;
   const int allsquare = (n == m);
   int i, *marker = (int *) xmalloc(sizeof(int) * (size_t) (m > 0 ? m : 1));
   long cap = 1024, cnt = 0;
   for (i = 0; i < m; i++) { marker[i] = -1; }
   C->nrows = n; C->ncols = m;
   C->ia = (int *) xmalloc(sizeof(int) * (size_t) (n + 1));
   C->ja = (int *) xmalloc(sizeof(int) * (size_t) cap);
   C->a = (double *) xmalloc(sizeof(double) * (size_t) cap);
   for (i = 0; i < n; i++)
   {
      const long row_begin = cnt;
      int jj2, jj3;
      C->ia[i] = (int) cnt;
      if (cnt + 1 > cap)
      {
         cap = 2 * cap + 1;
         C->ja = (int *) realloc(C->ja, sizeof(int) * (size_t) cap);
         C->a = (double *) realloc(C->a, sizeof(double) * (size_t) cap);
      }
      if (allsquare) { marker[i] = (int) cnt; C->ja[cnt] = i; C->a[cnt++] = 0.0; }
      for (jj2 = A->ia[i]; jj2 < A->ia[i + 1]; jj2++)
      {
         const int i2 = A->ja[jj2];
         const double a_entry = A->a[jj2];
         if (cnt + (B->ia[i2 + 1] - B->ia[i2]) > cap)
         {
            cap = 2 * cap + (B->ia[i2 + 1] - B->ia[i2]);
            C->ja = (int *) realloc(C->ja, sizeof(int) * (size_t) cap);
            C->a = (double *) realloc(C->a, sizeof(double) * (size_t) cap);
         }
         for (jj3 = B->ia[i2]; jj3 < B->ia[i2 + 1]; jj3++)
         {
            const int i3 = B->ja[jj3];
            if (marker[i3] < row_begin)
            {
               marker[i3] = (int) cnt;
               C->ja[cnt] = i3;
               C->a[cnt++] = a_entry * B->a[jj3];
            }
            else
            {
               C->a[marker[i3]] += a_entry * B->a[jj3];
            }
         }
      }
   }
   C->ia[n] = (int) cnt;
   free(marker);
   return 0;
}

/* src/parcsr_ls/par_mod_multi_interp.c:530 hypre_GenerateMultipassPi, one
 * rank: for the `num_points` rows of `pass_order` (pass color+1), the
 * entries a_ij of the strong connections j of pass `color` (S order; the
 * entry of A found by a forward search, :717-726), scaled to
 * -a_ij * row_sums[i] / (sum_j a_ij * a_ii) (:729-765; row_sums[i] is
 * divided in place only when the product is nonzero).  Columns number the
 * points of pass `color` in row order. */
static void gen_multipass_pi(const orc_csr *A, const orc_csr *S, const int *pass_order,
                             const int *pass_marker, int num_points, int color,
                             double *row_sums, orc_csr *Pi)
{
   const int n = A->nrows;
   int i, j, nct = 0, cnt = 0;
   int *f2c = (int *) xmalloc(sizeof(int) * (size_t) (n > 0 ? n : 1));
   for (i = 0; i < n; i++) { f2c[i] = (pass_marker[i] == color) ? nct++ : -1; }
   Pi->nrows = num_points; Pi->ncols = nct;
   Pi->ia = (int *) xcalloc((size_t) num_points + 1, sizeof(int));
   for (i = 0; i < num_points; i++)
   {
      const int i1 = pass_order[i];
      for (j = S->ia[i1]; j < S->ia[i1 + 1]; j++)
      {
         if (pass_marker[S->ja[j]] == color) { Pi->ia[i + 1]++; }
      }
   }
   for (i = 0; i < num_points; i++) { Pi->ia[i + 1] += Pi->ia[i]; }
   Pi->ja = (int *) xmalloc(sizeof(int) * (size_t) (Pi->ia[num_points] > 0 ? Pi->ia[num_points] : 1));
   Pi->a = (double *) xmalloc(sizeof(double) * (size_t) (Pi->ia[num_points] > 0 ? Pi->ia[num_points] : 1));
   for (i = 0; i < num_points; i++)
   {
      const int i1 = pass_order[i];
      int j2 = A->ia[i1];
      for (j = S->ia[i1]; j < S->ia[i1 + 1]; j++)
      {
         const int j1 = S->ja[j];
         while (A->ja[j2] != j1) { j2++; }
         if (pass_marker[j1] == color)
         {
            Pi->ja[cnt] = f2c[j1];
            Pi->a[cnt++] = A->a[j2];
         }
      }
   }
   for (i = 0; i < num_points; i++)
   {
      const int i1 = pass_order[i];
      const double diagonal = A->a[A->ia[i1]];
      double row_sum_C = 0.0, value;
      for (j = Pi->ia[i]; j < Pi->ia[i + 1]; j++) { row_sum_C += Pi->a[j]; }
      value = row_sum_C * diagonal;
      if (value != 0) { row_sums[i1] /= value; }
      for (j = Pi->ia[i]; j < Pi->ia[i + 1]; j++) { Pi->a[j] = -Pi->a[j] * row_sums[i1]; }
   }
   free(f2c);
}

/* src/parcsr_ls/par_mod_multi_interp.c:17 hypre_BoomerAMGBuildModMultipassHost
 * with interp_type 8, one rank, num_functions 1.  Points are colored by
 * passes (C points pass 1; a remaining point joins pass p+1 when one of its
 * strong connections is in pass p; :151-219, at most 9 passes), the weights
 * of pass 2 come from gen_multipass_pi and those of pass p+1 are the product
 * of its own gen_multipass_pi matrix with the weights of pass p (:247-259),
 * and P is truncated by (trunc_factor, P_max_elmts) when either is set
 * (:410-419).  CF -3 becomes -1 at the end (:503-506).  Returns 0, or -1 when
 * the passes exceed the reference's limit. */
int orc_mod_multipass(const orc_csr *A, const orc_csr *S, int *CF, double trunc_factor,
                      int P_max_elmts, orc_csr *P)
{
   const int n = A->nrows;
   int i, j, cnt = 0, remaining = 0, current_pass = 1, num_passes = 1, nc = 0, p;
   int *pass_marker = (int *) xcalloc((size_t) (n > 0 ? n : 1), sizeof(int));
   int *pass_order = (int *) xcalloc((size_t) (n > 0 ? n : 1), sizeof(int));
   int *f2c = (int *) xcalloc((size_t) (n > 0 ? n : 1), sizeof(int));
   int *left = (int *) xcalloc((size_t) (n > 0 ? n : 1), sizeof(int));
   int pass_starts[11] = {0};
   double *row_sums;
   orc_csr Pi[10];
   memset(Pi, 0, sizeof(Pi));
   for (i = 0; i < n; i++)
   {
      if (CF[i] == 1) { pass_marker[i] = 1; f2c[i] = cnt; pass_order[cnt++] = i; }
      else { left[remaining++] = i; }
   }
   nc = cnt;
   pass_starts[1] = cnt;
   while (remaining > 0)
   {
      const int old_remaining = remaining;
      int cnt_rem = 0;
      for (i = 0; i < old_remaining; i++)
      {
         const int i1 = left[i];
         int hit = 0;
         for (j = S->ia[i1]; j < S->ia[i1 + 1]; j++)
         {
            if (pass_marker[S->ja[j]] == current_pass) { hit = 1; break; }
         }
         if (hit) { pass_marker[i1] = current_pass + 1; pass_order[cnt++] = i1; remaining--; }
         else { left[cnt_rem++] = i1; }
      }
      remaining = cnt_rem;
      current_pass++;
      num_passes++;
      if (num_passes > 9)
      {
         free(pass_marker); free(pass_order); free(f2c); free(left);
         return -1;
      }
      pass_starts[num_passes] = cnt;
      if (old_remaining == remaining) { break; }
   }
   /* :221-265 row sums of the off-diagonal entries of the F rows */
   row_sums = (double *) xcalloc((size_t) (n > 0 ? n : 1), sizeof(double));
   for (i = 0; i < n; i++)
   {
      if (CF[i] < 0)
      {
         for (j = A->ia[i] + 1; j < A->ia[i + 1]; j++) { row_sums[i] += A->a[j]; }
      }
   }
   if (num_passes > 1)
   {
      gen_multipass_pi(A, S, &pass_order[pass_starts[1]], pass_marker,
                       pass_starts[2] - pass_starts[1], 1, row_sums, &Pi[0]);
   }
   for (i = 1; i < num_passes - 1; i++)
   {
      orc_csr Q;
      gen_multipass_pi(A, S, &pass_order[pass_starts[i + 1]], pass_marker,
                       pass_starts[i + 2] - pass_starts[i + 1], i + 1, row_sums, &Q);
      orc_matmul(&Q, &Pi[i - 1], &Pi[i]);
      orc_csr_free(&Q);
   }
   /* :266-330 assemble P: C rows are unit rows, pass rows copy their Pi row */
   P->nrows = n; P->ncols = nc;
   P->ia = (int *) xcalloc((size_t) n + 1, sizeof(int));
   for (i = 0; i < pass_starts[1]; i++) { P->ia[pass_order[i] + 1] = 1; }
   for (p = 0; p < num_passes - 1; p++)
   {
      int j1 = 0;
      for (i = pass_starts[p + 1]; i < pass_starts[p + 2]; i++, j1++)
      {
         P->ia[pass_order[i] + 1] = Pi[p].ia[j1 + 1] - Pi[p].ia[j1];
      }
   }
   for (i = 0; i < n; i++) { P->ia[i + 1] += P->ia[i]; }
   P->ja = (int *) xmalloc(sizeof(int) * (size_t) (P->ia[n] > 0 ? P->ia[n] : 1));
   P->a = (double *) xmalloc(sizeof(double) * (size_t) (P->ia[n] > 0 ? P->ia[n] : 1));
   for (i = 0; i < pass_starts[1]; i++)
   {
      const int i1 = pass_order[i];
      P->ja[P->ia[i1]] = f2c[i1];
      P->a[P->ia[i1]] = 1.0;
   }
   for (p = 0; p < num_passes - 1; p++)
   {
      int j1 = 0;
      for (i = pass_starts[p + 1]; i < pass_starts[p + 2]; i++, j1++)
      {
         const int i1 = pass_order[i];
         int i2 = Pi[p].ia[j1];
         for (j = P->ia[i1]; j < P->ia[i1 + 1]; j++, i2++)
         {
            P->ja[j] = Pi[p].ja[i2];
            P->a[j] = Pi[p].a[i2];
         }
      }
   }
   for (p = 0; p < num_passes - 1; p++) { orc_csr_free(&Pi[p]); }
   if (trunc_factor != 0.0 || P_max_elmts > 0) { orc_truncate(P, trunc_factor, P_max_elmts); }
   for (i = 0; i < n; i++) { if (CF[i] == SF_PT) { CF[i] = F_PT; } }
   free(pass_marker); free(pass_order); free(f2c); free(left); free(row_sums);
   return 0;
}

/* ======================================================================
 * Extended+i interpolation: src/parcsr_ls/par_lr_interp.c:1024-1900
 * (BuildExtPIInterpHost, one rank, num_functions 1).
 * ====================================================================== */
int orc_extpi_interp_raw(const orc_csr *A, const orc_csr *S, const int *CF, orc_csr *P)
{
   int n = A->nrows, i, nc = 0;
   const int *Ai = A->ia, *Aj = A->ja, *Si = S->ia, *Sj = S->ja;
   const double *Ad = A->a;
   int *f2c = (int *) xmalloc(sizeof(int) * (size_t) (n > 0 ? n : 1));
   long cap;

   /* fine_to_coarse (par_lr_interp.c:1230-1400) */
   for (i = 0; i < n; i++) { f2c[i] = (CF[i] >= 0) ? nc++ : -1; }
   P->nrows = n;
   P->ncols = nc;
   P->ia = (int *) xmalloc(sizeof(int) * (size_t)(n + 1));

   /* first pass: row sizes (par_lr_interp.c:1230-1400).  P_marker holds
    * positions of a running counter; each thread keeps its own marker array
    * and counter and takes its rows in ascending order, which is all the
    * "marker < row start" test needs. */
#pragma omp parallel private(i)
   {
      int *P_marker = (int *) xmalloc(sizeof(int) * (size_t) (n > 0 ? n : 1));
      int jj_counter = 0, jj, kk, i1, k1;
      for (i = 0; i < n; i++) { P_marker[i] = -1; }
#pragma omp for schedule(static)
      for (i = 0; i < n; i++)
      {
         const int row_begin = jj_counter;
         if (CF[i] >= 0)
         {
            jj_counter++;
         }
         else if (CF[i] != SF_PT)
         {
            for (jj = Si[i]; jj < Si[i + 1]; jj++)
            {
               i1 = Sj[jj];
               if (CF[i1] >= 0)
               {
                  if (P_marker[i1] < row_begin) { P_marker[i1] = jj_counter; jj_counter++; }
               }
               else if (CF[i1] != SF_PT)
               {
                  for (kk = Si[i1]; kk < Si[i1 + 1]; kk++)
                  {
                     k1 = Sj[kk];
                     if (CF[k1] >= 0)
                     {
                        if (P_marker[k1] < row_begin) { P_marker[k1] = jj_counter; jj_counter++; }
                     }
                  }
               }
            }
         }
         P->ia[i + 1] = jj_counter - row_begin;
      }
      free(P_marker);
   }
   orc_prefix(P->ia, n);
   cap = P->ia[n];
   P->ja = (int *) xmalloc(sizeof(int) * (size_t) (cap > 0 ? cap : 1));
   P->a = (double *) xmalloc(sizeof(double) * (size_t) (cap > 0 ? cap : 1));

   /* second pass: fill (par_lr_interp.c:1485-1810).  The strong-F tag of a
    * row, -2 - i, is unique to the row, like the reference's decreasing
    * strong_f_marker, and below every position. */
#pragma omp parallel private(i)
   {
      int *P_marker = (int *) xmalloc(sizeof(int) * (size_t) (n > 0 ? n : 1));
      int jj, kk, jj1, i1, i2, k1;
      for (i = 0; i < n; i++) { P_marker[i] = -1; }
#pragma omp for schedule(static)
      for (i = 0; i < n; i++)
      {
         int jj_begin_row = P->ia[i], jj_end_row;
         int jj_counter2 = jj_begin_row;
         const int strong_f_marker = -2 - i;
         if (CF[i] >= 0)
         {
            P->ja[jj_counter2] = f2c[i];
            P->a[jj_counter2] = 1.0;
            jj_counter2++;
         }
         else if (CF[i] != SF_PT)
         {
            double diagonal, sum, distribute;
            int sgn;
            for (jj = Si[i]; jj < Si[i + 1]; jj++)
            {
               i1 = Sj[jj];
               if (CF[i1] >= 0)
               {
                  if (P_marker[i1] < jj_begin_row)
                  {
                     P_marker[i1] = jj_counter2;
                     P->ja[jj_counter2] = f2c[i1];
                     P->a[jj_counter2] = 0.0;
                     jj_counter2++;
                  }
               }
               else if (CF[i1] != SF_PT)
               {
                  P_marker[i1] = strong_f_marker;
                  for (kk = Si[i1]; kk < Si[i1 + 1]; kk++)
                  {
                     k1 = Sj[kk];
                     if (CF[k1] >= 0)
                     {
                        if (P_marker[k1] < jj_begin_row)
                        {
                           P_marker[k1] = jj_counter2;
                           P->ja[jj_counter2] = f2c[k1];
                           P->a[jj_counter2] = 0.0;
                           jj_counter2++;
                        }
                     }
                  }
               }
            }
            jj_end_row = jj_counter2;
            diagonal = Ad[Ai[i]];
            for (jj = Ai[i] + 1; jj < Ai[i + 1]; jj++)
            {
               i1 = Aj[jj];
               if (P_marker[i1] >= jj_begin_row)
               {
                  P->a[P_marker[i1]] += Ad[jj];
               }
               else if (P_marker[i1] == strong_f_marker)
               {
                  sum = 0.0;
                  sgn = 1;
                  if (Ad[Ai[i1]] < 0) { sgn = -1; }
                  for (jj1 = Ai[i1] + 1; jj1 < Ai[i1 + 1]; jj1++)
                  {
                     i2 = Aj[jj1];
                     if ((P_marker[i2] >= jj_begin_row || i2 == i) && (sgn * Ad[jj1]) < 0)
                     {
                        sum += Ad[jj1];
                     }
                  }
                  if (sum != 0)
                  {
                     distribute = Ad[jj] / sum;
                     for (jj1 = Ai[i1] + 1; jj1 < Ai[i1 + 1]; jj1++)
                     {
                        i2 = Aj[jj1];
                        if (P_marker[i2] >= jj_begin_row && (sgn * Ad[jj1]) < 0)
                        {
                           P->a[P_marker[i2]] += distribute * Ad[jj1];
                        }
                        if (i2 == i && (sgn * Ad[jj1]) < 0)
                        {
                           diagonal += distribute * Ad[jj1];
                        }
                     }
                  }
                  else
                  {
                     diagonal += Ad[jj];
                  }
               }
               else if (CF[i1] != SF_PT)
               {
                  diagonal += Ad[jj];
               }
            }
            if (diagonal != 0.0)
            {
               for (jj = jj_begin_row; jj < jj_end_row; jj++)
               {
                  P->a[jj] /= -diagonal;
               }
            }
         }
      }
      free(P_marker);
   }
   free(f2c);
   return 0;
}

/* src/utilities/qsort.c:452-475 (+ hypre_swap2 at :42) */
static void swap2(int *v, double *w, int i, int j)
{
   int t = v[i]; double t2 = w[i];
   v[i] = v[j]; v[j] = t;
   w[i] = w[j]; w[j] = t2;
}

void orc_qsort2_abs(int *v, double *w, int left, int right)
{
   int i, last;
   if (left >= right) { return; }
   swap2(v, w, left, (left + right) / 2);
   last = left;
   for (i = left + 1; i <= right; i++)
   {
      if (fabs(w[i]) > fabs(w[left])) { swap2(v, w, ++last, i); }
   }
   swap2(v, w, left, last);
   orc_qsort2_abs(v, w, left, last - 1);
   orc_qsort2_abs(v, w, last + 1, right);
}

/* src/parcsr_mv/par_csr_matrix.c:2943-3340 hypre_ParCSRMatrixTruncate with
 * rescale = 1, nrm_type = 0 (called from par_interp.c:2622-2652). One thread,
 * empty offd.  Compacts in place. */
int orc_truncate(orc_csr *P, double tol, int max_row_elmts)
{
   int n = P->nrows, i, j;
   int *Pi = P->ia, *Pj = P->ja;
   double *Pd = P->a;
   if (tol <= 0.0 && max_row_elmts == 0) { return 0; }
   if (tol > 0)
   {
      int next_open = 0, now_checking = 0, num_lost = 0;
      for (i = 0; i < n; i++)
      {
         double row_nrm = 0, drop_coeff, row_sum = 0, scale = 0;
         int start_j = Pi[i];
         for (j = Pi[i]; j < Pi[i + 1]; j++)
         {
            row_nrm = (row_nrm < fabs(Pd[j])) ? fabs(Pd[j]) : row_nrm;
         }
         drop_coeff = tol * row_nrm;
         if (num_lost) { Pi[i] -= num_lost; }
         for (j = start_j; j < Pi[i + 1]; j++)
         {
            row_sum += Pd[now_checking];
            if (fabs(Pd[now_checking]) < drop_coeff)
            {
               num_lost++; now_checking++;
            }
            else
            {
               scale += Pd[now_checking];
               Pd[next_open] = Pd[now_checking];
               Pj[next_open] = Pj[now_checking];
               now_checking++; next_open++;
            }
         }
         if (scale != 0.)
         {
            if (scale != row_sum)
            {
               scale = row_sum / scale;
               for (j = Pi[i]; j < (Pi[i + 1] - num_lost); j++) { Pd[j] *= scale; }
            }
         }
      }
      Pi[n] -= num_lost;
   }
   if (max_row_elmts > 0)
   {
      int mx = 0;
      for (i = 0; i < n; i++) { mx = ORC_MAX(mx, Pi[i + 1] - Pi[i]); }
      if (mx > max_row_elmts)
      {
         /* rows keep min(length, max_row_elmts) entries; written to new
          * arrays at offsets counted first (the reference compacts in place
          * with a running cnt_diag, which yields the same arrays) */
         int *ni = (int *) xmalloc(sizeof(int) * (size_t) (n + 1));
         int *nj;
         double *nd;
         for (i = 0; i < n; i++) { ni[i + 1] = ORC_MIN(Pi[i + 1] - Pi[i], max_row_elmts); }
         orc_prefix(ni, n);
         nj = (int *) xmalloc(sizeof(int) * (size_t) (ni[n] > 0 ? ni[n] : 1));
         nd = (double *) xmalloc(sizeof(double) * (size_t) (ni[n] > 0 ? ni[n] : 1));
#pragma omp parallel private(i, j)
         {
            int *aux_j = (int *) xmalloc(sizeof(int) * (size_t) mx);
            double *aux_d = (double *) xmalloc(sizeof(double) * (size_t) mx);
#pragma omp for schedule(static)
            for (i = 0; i < n; i++)
            {
               int last_index = Pi[i + 1];
               int num_elmts = last_index - Pi[i];
               int cnt_diag = ni[i];
               double row_sum = 0, scale;
               if (max_row_elmts < num_elmts)
               {
                  int cnt = 0, row_begin = cnt_diag;
                  for (j = Pi[i]; j < last_index; j++)
                  {
                     aux_j[cnt] = Pj[j];
                     aux_d[cnt++] = Pd[j];
                     row_sum += Pd[j];
                  }
                  orc_qsort2_abs(aux_j, aux_d, 0, cnt - 1);
                  scale = 0;
                  for (j = 0; j < max_row_elmts; j++)
                  {
                     scale += aux_d[j];
                     nj[cnt_diag] = aux_j[j];
                     nd[cnt_diag++] = aux_d[j];
                  }
                  if (scale != 0.)
                  {
                     if (scale != row_sum)
                     {
                        scale = row_sum / scale;
                        for (j = row_begin; j < cnt_diag; j++) { nd[j] *= scale; }
                     }
                  }
               }
               else
               {
                  for (j = Pi[i]; j < last_index; j++)
                  {
                     nj[cnt_diag] = Pj[j];
                     nd[cnt_diag++] = Pd[j];
                  }
               }
            }
            free(aux_j); free(aux_d);
         }
         free(P->ia); free(P->ja); free(P->a);
         P->ia = ni; P->ja = nj; P->a = nd;
      }
   }
   return 0;
}

int orc_extpi_interp(const orc_csr *A, const orc_csr *S, int *CF,
                     double trunc_factor, int max_elmts, orc_csr *P)
{
   int i;
   orc_extpi_interp_raw(A, S, CF, P);
   /* par_lr_interp.c:1870 */
   if (trunc_factor != 0.0 || max_elmts > 0)
   {
      orc_truncate(P, trunc_factor, max_elmts);
   }
   /* par_lr_interp.c:1890-1897 */
   for (i = 0; i < A->nrows; i++)
   {
      if (CF[i] == SF_PT) { CF[i] = F_PT; }
   }
   return 0;
}

/* src/seq_mv/csr_matop.c:1071 hypre_CSRMatrixTransposeHost: stable counting
 * sort, so each row of R lists the fine rows of P in ascending order. */
int orc_transpose(const orc_csr *P, orc_csr *R)
{
   int n = P->nrows, nc = P->ncols, nnz = P->ia[n], c;
   const int nt = orc_threads();
   /* per-thread column counts over contiguous row chunks; offsets ordered by
    * (column, thread) keep each column's rows ascending, as the serial sort */
   int *cnt = (int *) xcalloc((size_t) nt * (size_t) (nc + 1), sizeof(int));
   R->nrows = nc; R->ncols = n;
   R->ia = (int *) xmalloc(sizeof(int) * (size_t)(nc + 1));
   R->ja = (int *) xmalloc(sizeof(int) * (size_t) (nnz > 0 ? nnz : 1));
   R->a = (double *) xmalloc(sizeof(double) * (size_t) (nnz > 0 ? nnz : 1));
#pragma omp parallel num_threads(nt)
   {
      const int t = orc_thread_id(), team = orc_team();
      int r0, r1, i, j;
      int *ct = cnt + (size_t) t * (nc + 1);
      orc_chunk(n, team, t, &r0, &r1);
      for (i = r0; i < r1; i++)
      {
         for (j = P->ia[i]; j < P->ia[i + 1]; j++) { ct[P->ja[j]]++; }
      }
#pragma omp barrier
#pragma omp single
      {
         int run = 0, tt;
         for (c = 0; c < nc; c++)
         {
            R->ia[c] = run;
            for (tt = 0; tt < team; tt++)
            {
               int *cc = cnt + (size_t) tt * (nc + 1);
               const int k = cc[c];
               cc[c] = run;
               run += k;
            }
         }
         R->ia[nc] = run;
      }
      for (i = r0; i < r1; i++)
      {
         for (j = P->ia[i]; j < P->ia[i + 1]; j++)
         {
            const int pos = ct[P->ja[j]]++;
            R->ja[pos] = i;
            R->a[pos] = P->a[j];
         }
      }
   }
   free(cnt);
   return 0;
}

/* src/parcsr_ls/par_rap.c:30-1930 BuildCoarseOperatorKT, one rank: row ic of
 * RA is built by a marker sweep over R(ic,:) x A, then (RA)(ic,:) x P; the
 * diagonal entry is created first (par_rap.c:1612-1620). */
int orc_rap(const orc_csr *R, const orc_csr *A, const orc_csr *P, orc_csr *Ac)
{
   int nc = R->nrows, nf = A->nrows;
   const int nt = orc_threads();
   /* each thread builds the rows of a contiguous chunk into its own buffer
    * (positions counted from 0 for its first row, so the marker tests of the
    * serial loop hold unchanged), then the buffers are concatenated in order */
   int **tj = (int **) xcalloc((size_t) nt, sizeof(int *));
   double **td = (double **) xcalloc((size_t) nt, sizeof(double *));
   long *tcnt = (long *) xcalloc((size_t) nt, sizeof(long));
   int team_size = 1;
   Ac->nrows = nc; Ac->ncols = nc;
   Ac->ia = (int *) xmalloc(sizeof(int) * (size_t)(nc + 1));
#pragma omp parallel num_threads(nt)
   {
      const int t = orc_thread_id(), team = orc_team();
      int ic, jj1, jj2, i1, i2, r0, r1;
      int *A_marker = (int *) xmalloc(sizeof(int) * (size_t) (nf > 0 ? nf : 1));
      int *P_marker = (int *) xmalloc(sizeof(int) * (size_t) (nc > 0 ? nc : 1));
      int *RA_j = (int *) xmalloc(sizeof(int) * (size_t) (nf > 0 ? nf : 1));
      double *RA_d = (double *) xmalloc(sizeof(double) * (size_t) (nf > 0 ? nf : 1));
      long cap = 1024, cnt = 0;
      int ra_nnz = 0;
      int *Acj = (int *) xmalloc(sizeof(int) * (size_t) cap);
      double *Acd = (double *) xmalloc(sizeof(double) * (size_t) cap);
      if (t == 0) { team_size = team; }
      orc_chunk(nc, team, t, &r0, &r1);
      for (i1 = 0; i1 < nf; i1++) { A_marker[i1] = -1; }
      for (i1 = 0; i1 < nc; i1++) { P_marker[i1] = -1; }
      for (ic = r0; ic < r1; ic++)
      {
         long jj_row_begin = cnt;
         int ra_row_begin = ra_nnz;
         int ra_cnt = 0;
         /* RA row: par_rap.c:1712-1800 (A_marker holds absolute RA positions) */
         for (jj1 = R->ia[ic]; jj1 < R->ia[ic + 1]; jj1++)
         {
            double r_entry = R->a[jj1];
            i1 = R->ja[jj1];
            for (jj2 = A->ia[i1]; jj2 < A->ia[i1 + 1]; jj2++)
            {
               double a_entry = A->a[jj2];
               int marker;
               i2 = A->ja[jj2];
               marker = A_marker[i2];
               if (marker < ra_row_begin)
               {
                  A_marker[i2] = ra_nnz;
                  RA_d[ra_nnz - ra_row_begin] = r_entry * a_entry;
                  RA_j[ra_nnz - ra_row_begin] = i2;
                  ra_nnz++;
               }
               else
               {
                  RA_d[marker - ra_row_begin] += r_entry * a_entry;
               }
            }
         }
         ra_cnt = ra_nnz - ra_row_begin;
         /* make room for the row of RAP: at most 1 + sum |P(i1,:)| entries */
         {
            long bound = 1;
            for (jj1 = 0; jj1 < ra_cnt; jj1++)
            {
               i1 = RA_j[jj1];
               bound += P->ia[i1 + 1] - P->ia[i1];
            }
            if (cnt + bound > cap)
            {
               while (cnt + bound > cap) { cap *= 2; }
               Acj = (int *) realloc(Acj, sizeof(int) * (size_t) cap);
               Acd = (double *) realloc(Acd, sizeof(double) * (size_t) cap);
               if (!Acj || !Acd) { abort(); }
            }
         }
         /* diagonal first */
         P_marker[ic] = (int) cnt;
         Acd[cnt] = 0.0;
         Acj[cnt] = ic;
         cnt++;
         /* (RA) x P: par_rap.c:1860-1920 */
         for (jj1 = 0; jj1 < ra_cnt; jj1++)
         {
            double r_a_product = RA_d[jj1];
            i1 = RA_j[jj1];
            for (jj2 = P->ia[i1]; jj2 < P->ia[i1 + 1]; jj2++)
            {
               double p_entry = P->a[jj2];
               int marker;
               i2 = P->ja[jj2];
               marker = P_marker[i2];
               if (marker < jj_row_begin)
               {
                  P_marker[i2] = (int) cnt;
                  Acd[cnt] = r_a_product * p_entry;
                  Acj[cnt] = i2;
                  cnt++;
               }
               else
               {
                  Acd[marker] += r_a_product * p_entry;
               }
            }
         }
         Ac->ia[ic + 1] = (int) (cnt - jj_row_begin);
      }
      tj[t] = Acj; td[t] = Acd; tcnt[t] = cnt;
      free(A_marker); free(P_marker); free(RA_j); free(RA_d);
   }
   {
      long total = 0, off = 0;
      int t;
      for (t = 0; t < team_size; t++) { total += tcnt[t]; }
      orc_prefix(Ac->ia, nc);
      Ac->ja = (int *) xmalloc(sizeof(int) * (size_t) (total > 0 ? total : 1));
      Ac->a = (double *) xmalloc(sizeof(double) * (size_t) (total > 0 ? total : 1));
      for (t = 0; t < team_size; t++)
      {
         memcpy(Ac->ja + off, tj[t], sizeof(int) * (size_t) tcnt[t]);
         memcpy(Ac->a + off, td[t], sizeof(double) * (size_t) tcnt[t]);
         off += tcnt[t];
         free(tj[t]); free(td[t]);
      }
   }
   free(tj); free(td); free(tcnt);
   return 0;
}

/* src/parcsr_ls/ams.c:4523-4600 (option 1) and ams.c:770-797: sum |a_ij|
 * over the row in storage order, negated where the diagonal is negative. */
int orc_l1_norms(const orc_csr *A, double *l1)
{
   int i, j, bad = 0;
#pragma omp parallel for private(j) schedule(static)
   for (i = 0; i < A->nrows; i++)
   {
      l1[i] = 0.0;
      for (j = A->ia[i]; j < A->ia[i + 1]; j++) { l1[i] += fabs(A->a[j]); }
      if (A->a[A->ia[i]] < 0.0) { l1[i] = -l1[i]; }
   }
   for (i = 0; i < A->nrows; i++)
   {
      if (fabs(l1[i]) == 0.0) { bad = 1; break; }
   }
   return bad ? 4 : 0;
}

/* hypre_ParCSRComputeL1Norms option 4 on one rank (src/parcsr_ls/ams.c:649-676
 * and 776-795): |a_ii| (hypre_CSRMatrixExtractDiagonalHost type 1, last
 * entry with j == i, src/seq_mv/csr_matop.c:2091-2117); the truncation of
 * Remark 6.2 leaves it unchanged without an offd block; the sign of a_ii is
 * restored. */
int orc_l1_norms_diag(const orc_csr *A, double *l1)
{
   int i, j, bad = 0;
   for (i = 0; i < A->nrows; i++)
   {
      double d = 0.0, ad = 0.0;
      for (j = A->ia[i]; j < A->ia[i + 1]; j++)
      {
         if (A->ja[j] == i) { d = A->a[j]; ad = fabs(A->a[j]); }
      }
      l1[i] = (d < 0.0) ? -ad : ad;
      if (fabs(l1[i]) == 0.0) { bad = 1; }
   }
   return bad ? 4 : 0;
}

/* hypre_partition1D, src/utilities/utilities.c:99-125 */
void orc_partition1d(int n, int p, int j, int *s, int *e)
{
   int size, rest;
   if (p == 1) { *s = 0; *e = n; return; }
   size = n / p;
   rest = n - size * p;
   if (j < rest) { *s = j * (size + 1); *e = (j + 1) * (size + 1); }
   else { *s = j * size + rest; *e = (j + 1) * size + rest; }
}

/* Number of row partitions of the hybrid Gauss-Seidel smoother on a level of
 * n rows: `parts` when positive (the reference's thread count), otherwise
 * the automatic rule of the library: partitions of `rows` rows, fewer (down
 * to 8) where that would leave fewer than about 128K partitions. */
int orc_hgs_parts(int n, int parts, int rows)
{
   long long want;
   int r;
   if (parts > 0) { return parts; }
   if (rows < 1) { rows = 1; }
   want = ((long long) n + 131071) / 131072;    /* ~128K partitions when possible */
   if (want < 8) { want = 8; }
   r = (rows < want) ? rows : (int) want;
   if (n <= r) { return 1; }
   return (int) (((long long) n + r - 1) / r);
}

/* One hybrid Gauss-Seidel relaxation on a single rank whose rows are split
 * into `parts` contiguous partitions (hypre_partition1D): Gauss-Seidel in
 * row order inside a partition, values of other partitions taken from the
 * copy of u made before the sweep.  src/parcsr_ls/par_relax.c:692-953
 * (hypre_BoomerAMGRelaxHybridGaussSeidel_core; with one partition the
 * unthreaded loops) and the row kernels of src/parcsr_ls/par_relax.h:13-106
 * (NS), 109-222 (NSThreads), 225-331 (scaled), 334-455 (scaled threads).
 * direction 1 forward, -1 backward; symm: forward then backward against the
 * same copy.  l1 == NULL divides by the first entry of the row (the
 * diagonal), as the reference's A_diag_data[A_diag_i[i]]. */
void orc_hybrid_gs(const orc_csr *A, const double *f, const double *l1, double *u,
                   int parts, int direction, int symm, int skip_diag,
                   double relax_weight, double omega)
{
   const int n = A->nrows;
   const int non_scale = (relax_weight == 1.0 && omega == 1.0);
   const double prod = 1.0 - relax_weight * omega;
   const double one_minus_omega = 1.0 - omega;
   const int nsw = symm ? 2 : 1;
   double *v = (double *) xmalloc(sizeof(double) * (size_t) (n > 0 ? n : 1));
   int p, sw;
   memcpy(v, u, sizeof(double) * (size_t) n);
   if (parts < 1) { parts = 1; }
   for (p = 0; p < parts; p++)
   {
      int ns, ne;
      orc_partition1d(n, parts, p, &ns, &ne);
      for (sw = 0; sw < nsw; sw++)
      {
         const int ord = (nsw == 1) ? (direction > 0 ? 1 : -1) : (sw == 0 ? 1 : -1);
         const int ib = ord > 0 ? ns : ne - 1, ie = ord > 0 ? ne : ns - 1;
         int i;
         for (i = ib; i != ie; i += ord)
         {
            const double d = l1 ? l1[i] : A->a[A->ia[i]];
            int jj;
            if (d == 0.0) { continue; }
            if (non_scale)
            {
               double res = f[i];
               for (jj = A->ia[i] + skip_diag; jj < A->ia[i + 1]; jj++)
               {
                  const int c = A->ja[jj];
                  res -= A->a[jj] * ((c >= ns && c < ne) ? u[c] : v[c]);
               }
               if (skip_diag) { u[i] = res / d; }
               else { u[i] += res / d; }
            }
            else
            {
               double res = f[i], res0 = 0.0, res2 = 0.0;
               for (jj = A->ia[i] + skip_diag; jj < A->ia[i + 1]; jj++)
               {
                  const int c = A->ja[jj];
                  if (c >= ns && c < ne)
                  {
                     res0 -= A->a[jj] * u[c];
                     res2 += A->a[jj] * v[c];
                  }
                  else
                  {
                     res -= A->a[jj] * v[c];
                  }
               }
               if (skip_diag) { u[i] *= prod; }
               u[i] += relax_weight * (omega * res + res0 + one_minus_omega * res2) / d;
            }
         }
      }
   }
   free(v);
}

/* ======================================================================
 * Vector kernels
 * ====================================================================== */
void orc_matvec(double alpha, const orc_csr *A, const double *x, double beta,
                const double *b, double *y)
{
   int i, j;
#pragma omp parallel for private(j) schedule(static)
   for (i = 0; i < A->nrows; i++)
   {
      double t = 0.0;
      for (j = A->ia[i]; j < A->ia[i + 1]; j++) { t += A->a[j] * x[A->ja[j]]; }
      if (beta == 0.0) { y[i] = alpha * t; }
      else             { y[i] = alpha * t + beta * b[i]; }
   }
}

void orc_matvecT(double alpha, const orc_csr *A, const double *x, double beta, double *y)
{
   int i, j;
   for (i = 0; i < A->ncols; i++) { y[i] = (beta == 0.0) ? 0.0 : beta * y[i]; }
   for (i = 0; i < A->nrows; i++)
   {
      for (j = A->ia[i]; j < A->ia[i + 1]; j++) { y[A->ja[j]] += alpha * A->a[j] * x[i]; }
   }
}

double orc_dot(int n, const double *x, const double *y)
{
   double s = 0.0;
   int i;
   for (i = 0; i < n; i++) { s += x[i] * y[i]; }
   return s;
}

/* ======================================================================
 * AMG setup: level loop of src/parcsr_ls/par_amg_setup.c:1002-3185
 * ====================================================================== */
#define ORC_MAXLEV 64
struct orc_amg
{
   orc_amg_params prm;
   int      num_levels;
   orc_csr  A[ORC_MAXLEV], P[ORC_MAXLEV], R[ORC_MAXLEV], S[ORC_MAXLEV];
   int     *CF[ORC_MAXLEV];
   double  *l1[ORC_MAXLEV];
   int      rounds[ORC_MAXLEV];
   int      coarse_relax;
   int      coarse_sweeps;
   double  *A_mat;              /* coarsest dense matrix, column major */
   double  *F[ORC_MAXLEV], *U[ORC_MAXLEV], *Vtemp;
};

void orc_amg_default_params(orc_amg_params *p)
{
   p->strong_threshold = 0.25;
   p->max_row_sum = 0.9;
   p->trunc_factor = 0.0;
   p->P_max_elmts = 4;
   p->max_levels = 25;
   p->max_coarse_size = 9;
   p->relax_type = 18;
   p->coarse_relax_type = 9;
   p->relax_weight = 1.0;
   p->num_sweeps = 1;
   p->relax_up_type = -1;
   p->outer_weight = 1.0;
   p->hgs_parts = 0;
   p->hgs_rows = 4;
   p->coarsen_type = 8;
   p->agg_num_levels = 0;
   p->agg_interp_type = 8;
   p->agg_trunc_factor = 0.0;
   p->agg_P_max_elmts = 0;
}

static void csr_copy(const orc_csr *A, orc_csr *B)
{
   int nnz = A->ia[A->nrows];
   B->nrows = A->nrows; B->ncols = A->ncols;
   B->ia = (int *) xmalloc(sizeof(int) * (size_t)(A->nrows + 1));
   B->ja = (int *) xmalloc(sizeof(int) * (size_t) nnz);
   B->a = (double *) xmalloc(sizeof(double) * (size_t) nnz);
   memcpy(B->ia, A->ia, sizeof(int) * (size_t)(A->nrows + 1));
   memcpy(B->ja, A->ja, sizeof(int) * (size_t) nnz);
   memcpy(B->a, A->a, sizeof(double) * (size_t) nnz);
}

static int orc_is_l1_gs(int t) { return t == 8 || t == 13 || t == 14; }

orc_amg *orc_amg_setup(const orc_csr *A0, const orc_amg_params *prm)
{
   orc_amg *H = (orc_amg *) xcalloc(1, sizeof(orc_amg));
   int level = 0, j, n;
   H->prm = *prm;
   csr_copy(A0, &H->A[0]);
   H->coarse_relax = prm->coarse_relax_type;
   H->coarse_sweeps = 1;
   while (1)
   {
      int fine_size = H->A[level].nrows, coarse_size = 0, i;
      orc_csr *A = &H->A[level];
      if (prm->max_levels == 1) { break; }
      const int aggressive = level < prm->agg_num_levels;
      orc_strength(A, prm->strong_threshold, prm->max_row_sum, &H->S[level]);
      H->CF[level] = (int *) xmalloc(sizeof(int) * (size_t) fine_size);
      H->rounds[level] = (prm->coarsen_type == 10) ? orc_hmis(&H->S[level], H->CF[level])
                                                   : orc_pmis(&H->S[level], H->CF[level]);
      if (aggressive)
      {
         /* par_amg_setup.c:1311-1327: the C points are coarsened again over
          * the distance-two strength graph; :1799-1806 the split is
          * corrected before the interpolation */
         orc_csr S2;
         int *CFN;
         const int nc1 = orc_create_2nd_s(&H->S[level], H->CF[level], &S2);
         CFN = (int *) xmalloc(sizeof(int) * (size_t) (nc1 > 0 ? nc1 : 1));
         orc_pmis_cf3(&S2, CFN);
         orc_correct_cf_marker(fine_size, H->CF[level], CFN);
         free(CFN);
         orc_csr_free(&S2);
      }
      for (i = 0; i < fine_size; i++) { if (H->CF[level][i] == 1) { coarse_size++; } }
      /* par_amg_setup.c:1690-1725: no coarse grid -> stop, coarsest solve is
       * one sweep of the down smoother */
      if (coarse_size == 0 || coarse_size == fine_size)
      {
         if (H->coarse_relax == 9) { H->coarse_relax = prm->relax_type; H->coarse_sweeps = 1; }
         orc_csr_free(&H->S[level]);
         if (level > 0) { free(H->CF[level]); H->CF[level] = NULL; }
         break;
      }
      if (aggressive)
      {
         if (orc_mod_multipass(A, &H->S[level], H->CF[level], prm->agg_trunc_factor,
                               prm->agg_P_max_elmts, &H->P[level]) != 0)
         {
            fprintf(stderr, "oracle: multipass interpolation needs more than 9 passes\n");
            abort();
         }
      }
      else
      {
         orc_extpi_interp(A, &H->S[level], H->CF[level], prm->trunc_factor,
                          prm->P_max_elmts, &H->P[level]);
      }
      orc_transpose(&H->P[level], &H->R[level]);
      orc_rap(&H->R[level], A, &H->P[level], &H->A[level + 1]);
      level++;
      /* par_amg_setup.c:3136-3141 (the host-only switch to CLJP at
       * par_amg_setup.c:3128 is not taken by the device execution policy and
       * is not reproduced) */
      if (level == prm->max_levels - 1 || coarse_size <= prm->max_coarse_size) { break; }
   }
   H->num_levels = level + 1;
   /* coarsest solve, par_amg_setup.c:3165-3185 */
   n = H->A[level].nrows;
   if (H->coarse_relax == 9)
   {
      if (n <= prm->max_coarse_size)
      {
         const orc_csr *Ac = &H->A[level];
         H->A_mat = (double *) xcalloc((size_t) n * (size_t) n, sizeof(double));
         for (j = 0; j < n; j++)
         {
            int jj;
            for (jj = Ac->ia[j]; jj < Ac->ia[j + 1]; jj++)
            {
               H->A_mat[j + n * Ac->ja[jj]] = Ac->a[jj];   /* par_gauss_elim.c:312-316 */
            }
         }
      }
      else
      {
         H->coarse_relax = prm->relax_type;
      }
   }
   /* l1 norms, par_amg_setup.c:3297-3365: option 4 for the l1 hybrid
    * Gauss-Seidel types, option 1 for l1-Jacobi (evaluated last there, so it
    * wins when both kinds are used) */
   for (j = 0; j < H->num_levels; j++)
   {
      const int up = prm->relax_up_type >= 0 ? prm->relax_up_type : prm->relax_type;
      int opt = 0;
      if (j < H->num_levels - 1)
      {
         if (prm->relax_type == 18 || up == 18) { opt = 1; }
         else if (orc_is_l1_gs(prm->relax_type) || orc_is_l1_gs(up)) { opt = 4; }
      }
      else
      {
         if (H->coarse_relax == 18) { opt = 1; }
         else if (orc_is_l1_gs(H->coarse_relax)) { opt = 4; }
      }
      H->F[j] = (double *) xcalloc((size_t) H->A[j].nrows, sizeof(double));
      H->U[j] = (double *) xcalloc((size_t) H->A[j].nrows, sizeof(double));
      if (opt)
      {
         H->l1[j] = (double *) xmalloc(sizeof(double) * (size_t) H->A[j].nrows);
         if (opt == 1) { orc_l1_norms(&H->A[j], H->l1[j]); }
         else { orc_l1_norms_diag(&H->A[j], H->l1[j]); }
      }
   }
   H->Vtemp = (double *) xcalloc((size_t) H->A[0].nrows, sizeof(double));
   return H;
}

void orc_amg_free(orc_amg *H)
{
   int j;
   if (!H) { return; }
   for (j = 0; j < H->num_levels; j++)
   {
      orc_csr_free(&H->A[j]); orc_csr_free(&H->P[j]); orc_csr_free(&H->R[j]);
      orc_csr_free(&H->S[j]);
      free(H->CF[j]); free(H->l1[j]); free(H->F[j]); free(H->U[j]);
   }
   free(H->A_mat); free(H->Vtemp);
   free(H);
}

int orc_amg_num_levels(const orc_amg *H) { return H->num_levels; }
const orc_csr *orc_amg_A(const orc_amg *H, int l) { return &H->A[l]; }
const orc_csr *orc_amg_P(const orc_amg *H, int l) { return &H->P[l]; }
const orc_csr *orc_amg_R(const orc_amg *H, int l) { return &H->R[l]; }
const orc_csr *orc_amg_S(const orc_amg *H, int l) { return &H->S[l]; }
const int *orc_amg_CF(const orc_amg *H, int l) { return H->CF[l]; }
const double *orc_amg_l1(const orc_amg *H, int l) { return H->l1[l]; }
int orc_amg_pmis_rounds(const orc_amg *H, int l) { return H->rounds[l]; }
int orc_amg_coarse_relax(const orc_amg *H) { return H->coarse_relax; }

/* relax_type 18: src/parcsr_ls/par_relax.c:1178-1240 Relax7Jacobi
 *   Vtemp = w*(f - A u);  u += Vtemp ./ l1
 * relax_type 0: par_relax.c:180-320 weighted Jacobi (Skip_diag = 1).
 * relax_type 3/4/6 (par_relax.c:946-989, 1160-1171) and 8/13/14
 * (par_relax.c:1266-1383): hybrid Gauss-Seidel, l1 norms for 8/13/14 with
 * Skip_diag = 0 when weight and omega are both 1. */
static void relax(orc_amg *H, int lev, int type, const double *f, double *u)
{
   const orc_csr *A = &H->A[lev];
   int n = A->nrows, i, j;
   double w = H->prm.relax_weight;
   double *v = H->Vtemp;
   if (type == 3 || type == 4 || type == 6 || type == 8 || type == 13 || type == 14)
   {
      const double om = H->prm.outer_weight;
      const int l1t = (type == 8 || type == 13 || type == 14);
      const int skip = l1t ? ((w == 1.0 && om == 1.0) ? 0 : 1) : 1;
      const int dir = (type == 4 || type == 14) ? -1 : 1;
      const int symm = (type == 6 || type == 8);
      orc_hybrid_gs(A, f, l1t ? H->l1[lev] : NULL, u,
                    orc_hgs_parts(n, H->prm.hgs_parts, H->prm.hgs_rows), dir, symm, skip, w, om);
      return;
   }
   if (type == 18)
   {
#pragma omp parallel for private(j) schedule(static)
      for (i = 0; i < n; i++)
      {
         double t = 0.0;
         for (j = A->ia[i]; j < A->ia[i + 1]; j++) { t += A->a[j] * u[A->ja[j]]; }
         v[i] = -w * t + w * f[i];
      }
#pragma omp parallel for schedule(static)
      for (i = 0; i < n; i++) { u[i] += v[i] / H->l1[lev][i]; }
   }
   else   /* Jacobi */
   {
      double omw = 1.0 - w;
      for (i = 0; i < n; i++) { v[i] = u[i]; }
      for (i = 0; i < n; i++)
      {
         double di = A->a[A->ia[i]], res = f[i];
         if (di == 0.0) { continue; }
         for (j = A->ia[i] + 1; j < A->ia[i + 1]; j++) { res -= A->a[j] * v[A->ja[j]]; }
         u[i] *= omw;
         u[i] += w * res / di;
      }
   }
}

/* src/utilities/gselim.h hypre_gselim applied to the column-major matrix
 * (par_gauss_elim.c:640-650, relax type 9) */
static void gselim(double *A, double *x, int n)
{
   int j, k, m;
   if (n == 1)
   {
      if (A[0] != 0.0) { x[0] = x[0] / A[0]; }
      return;
   }
   for (k = 0; k < n - 1; k++)
   {
      if (A[k * n + k] != 0.0)
      {
         double divA = 1.0 / A[k * n + k];
         for (j = k + 1; j < n; j++)
         {
            if (A[j * n + k] != 0.0)
            {
               double factor = A[j * n + k] * divA;
               for (m = k + 1; m < n; m++) { A[j * n + m] -= factor * A[k * n + m]; }
               x[j] -= factor * x[k];
            }
         }
      }
   }
   for (k = n - 1; k > 0; --k)
   {
      if (A[k * n + k] != 0.0)
      {
         x[k] /= A[k * n + k];
         for (j = 0; j < k; j++)
         {
            if (A[j * n + k] != 0.0) { x[j] -= x[k] * A[j * n + k]; }
         }
      }
   }
   if (A[0] != 0.0) { x[0] /= A[0]; }
}

/* src/parcsr_ls/par_cycle.c:23-889, V-cycle, one sweep down/up */
int orc_amg_cycle(orc_amg *H, const double *f, double *u)
{
   int L = H->num_levels, lev, i, s;
   int nsw = H->prm.num_sweeps;
   memcpy(H->F[0], f, sizeof(double) * (size_t) H->A[0].nrows);
   memcpy(H->U[0], u, sizeof(double) * (size_t) H->A[0].nrows);
   for (lev = 0; lev < L - 1; lev++)
   {
      int n = H->A[lev].nrows;
      for (s = 0; s < nsw; s++) { relax(H, lev, H->prm.relax_type, H->F[lev], H->U[lev]); }
      /* residual and restriction, par_cycle.c:755-790 */
      orc_matvec(-1.0, &H->A[lev], H->U[lev], 1.0, H->F[lev], H->Vtemp);
      for (i = 0; i < H->A[lev + 1].nrows; i++) { H->U[lev + 1][i] = 0.0; }
      {
         const orc_csr *R = &H->R[lev];   /* F_c = P^T r via the stored transpose */
         int ic, jj;
         (void) n;
#pragma omp parallel for private(jj) schedule(static)
         for (ic = 0; ic < R->nrows; ic++)
         {
            double t = 0.0;
            for (jj = R->ia[ic]; jj < R->ia[ic + 1]; jj++) { t += R->a[jj] * H->Vtemp[R->ja[jj]]; }
            H->F[lev + 1][ic] = t;
         }
      }
   }
   /* coarsest level */
   lev = L - 1;
   if (H->coarse_relax == 9 && H->A_mat)
   {
      int n = H->A[lev].nrows;
      double *work = (double *) xmalloc(sizeof(double) * (size_t)(n * n));
      double *x = (double *) xmalloc(sizeof(double) * (size_t) n);
      memcpy(work, H->A_mat, sizeof(double) * (size_t)(n * n));
      memcpy(x, H->F[lev], sizeof(double) * (size_t) n);
      gselim(work, x, n);
      memcpy(H->U[lev], x, sizeof(double) * (size_t) n);
      free(work); free(x);
   }
   else
   {
      for (s = 0; s < H->coarse_sweeps; s++) { relax(H, lev, H->coarse_relax, H->F[lev], H->U[lev]); }
   }
   for (lev = L - 2; lev >= 0; lev--)
   {
      /* interpolation u += P e, par_cycle.c:820-840 */
      orc_matvec(1.0, &H->P[lev], H->U[lev + 1], 1.0, H->U[lev], H->U[lev]);
      for (s = 0; s < nsw; s++)
      {
         relax(H, lev, H->prm.relax_up_type >= 0 ? H->prm.relax_up_type : H->prm.relax_type,
               H->F[lev], H->U[lev]);
      }
   }
   memcpy(u, H->U[0], sizeof(double) * (size_t) H->A[0].nrows);
   return 0;
}

/* src/parcsr_ls/par_amg_solve.c:22-424 (tol > 0, converge_type 0) */
int orc_amg_solve(orc_amg *H, const double *b, double *x, double tol, int max_iter,
                  int *iters, double *rel_res, double *conv)
{
   int n = H->A[0].nrows, cycle = 0;
   double *r = (double *) xmalloc(sizeof(double) * (size_t) n);
   double res0, res, rhs, rel;
   orc_matvec(-1.0, &H->A[0], x, 1.0, b, r);
   res0 = sqrt(orc_dot(n, r, r));
   res = res0;
   rhs = sqrt(orc_dot(n, b, b));
   rel = (rhs != 0.0) ? res0 / rhs : res0;
   while (rel >= tol && cycle < max_iter)
   {
      orc_amg_cycle(H, b, x);
      orc_matvec(-1.0, &H->A[0], x, 1.0, b, r);
      res = sqrt(orc_dot(n, r, r));
      rel = (rhs != 0.0) ? res / rhs : res;
      cycle++;
   }
   *iters = cycle;
   *rel_res = rel;
   *conv = (cycle > 0 && res0 != 0.0) ? pow(res / res0, 1.0 / (double) cycle) : 1.0;
   free(r);
   return (cycle == max_iter && tol > 0.) ? 256 : 0;
}

/* src/krylov/pcg.c:313-1013, two_norm = 1, everything else default */
int orc_pcg_solve(orc_amg *H, const double *b, double *x, double tol, int max_iter,
                  int *iters, double *rel_res, double *history)
{
   int n = H->A[0].nrows, i = 0, k, err = 0;
   const orc_csr *A = &H->A[0];
   double *p = (double *) xcalloc((size_t) n, sizeof(double));
   double *s = (double *) xcalloc((size_t) n, sizeof(double));
   double *r = (double *) xcalloc((size_t) n, sizeof(double));
   double bi_prod, eps, gamma, gamma_old, alpha, beta, sdotp, i_prod = 0.0;
   bi_prod = orc_dot(n, b, b);
   if (bi_prod != bi_prod) { free(p); free(s); free(r); return 1; }
   if (bi_prod <= 0.0)
   {
      memcpy(x, b, sizeof(double) * (size_t) n);
      *iters = 0; *rel_res = 0.0;
      free(p); free(s); free(r);
      return 0;
   }
   eps = tol * tol;
   /* r = b - A x */
   orc_matvec(-1.0, A, x, 1.0, b, r);
   memset(p, 0, sizeof(double) * (size_t) n);
   orc_amg_cycle(H, r, p);
   gamma = orc_dot(n, r, p);
   if (gamma != gamma) { free(p); free(s); free(r); return 1; }
   while (i + 1 <= max_iter)
   {
      i++;
      orc_matvec(1.0, A, p, 0.0, NULL, s);
      sdotp = orc_dot(n, s, p);
      if (sdotp == 0.0) { err = 256; break; }
      alpha = gamma / sdotp;
      if (alpha <= 0.0) { err = 256; break; }
      gamma_old = gamma;
#pragma omp parallel for schedule(static)
      for (k = 0; k < n; k++) { x[k] += alpha * p[k]; r[k] += -alpha * s[k]; }
      memset(s, 0, sizeof(double) * (size_t) n);
      orc_amg_cycle(H, r, s);
      gamma = orc_dot(n, r, s);
      i_prod = orc_dot(n, r, r);
      if (history) { history[i] = sqrt(i_prod / bi_prod); }
      if (i_prod / bi_prod < eps) { break; }
      if (gamma <= 0.0) { err = 256; break; }
      beta = gamma / gamma_old;
#pragma omp parallel for schedule(static)
      for (k = 0; k < n; k++) { p[k] = beta * p[k]; p[k] += s[k]; }
   }
   *iters = i;
   *rel_res = sqrt(i_prod / bi_prod);
   if (i >= max_iter && (i_prod / bi_prod) >= eps) { err = 256; }
   free(p); free(s); free(r);
   return err;
}

/* ---------------------------------------------------------------- setup statistics
 * src/parcsr_ls/par_stats.c:20-1340 (hypre_BoomerAMGSetupStats), one rank,
 * for the options of this path: PMIS or HMIS coarsening, ext+i interpolation
 * with optional aggressive levels (modified multipass), relax_order 0 and no
 * grid relax points.  The text is what the reference prints with print level
 * 1 or 3 after setup; it goes to buf (NUL-terminated, at most cap bytes) and
 * the return value is the length the full text needs. */
typedef struct
{
   char  *buf;
   size_t cap, len;
} orc_text;

static void text_add(orc_text *t, const char *fmt, ...)
{
   va_list ap;
   size_t room = (t->len < t->cap) ? t->cap - t->len : 0;
   int n;
   va_start(ap, fmt);
   n = vsnprintf(room ? t->buf + t->len : NULL, room, fmt, ap);
   va_end(ap);
   if (n > 0) { t->len += (size_t) n; }
}

/* printf.c:136 hypre_ndigits_ull */
static int orc_ndigits(long long v)
{
   int d = 0;
   while (v > 0) { v /= 10; d++; }
   return d;
}

/* par_stats.c:587-650 (operators) and :851-960 (interpolation): entries per
 * row, row sums in storage order, and for P the smallest weight and the
 * largest weight other than 1.0; hypre_min / hypre_max keep their operand
 * order so that ties resolve as there */
static void orc_level_stats(const orc_csr *M, int weights, int *min_e, int *max_e,
                            double *min_rs, double *max_rs, double *min_w, double *max_w)
{
   int i, k;
   *min_e = 0; *max_e = 0;
   *min_rs = 0.0; *max_rs = 0.0;
   *min_w = 1.0; *max_w = 0.0;
   if (M->nrows == 0) { *min_e = 1000000; *min_rs = 1.0e7; *min_w = 1.0e7; return; }
   if (weights && M->ia[M->nrows] > 0) { *min_w = M->a[0]; }
   for (i = 0; i < M->nrows; i++)
   {
      int e = M->ia[i + 1] - M->ia[i];
      double rs = 0.0;
      for (k = M->ia[i]; k < M->ia[i + 1]; k++)
      {
         double v = M->a[k];
         if (weights)
         {
            *min_w = (*min_w < v) ? *min_w : v;
            if (v != 1.0) { *max_w = (*max_w < v) ? v : *max_w; }
         }
         rs += v;
      }
      if (i == 0) { *min_e = e; *min_rs = rs; *max_rs = rs; }
      *min_e = (e < *min_e) ? e : *min_e;
      *max_e = (*max_e < e) ? e : *max_e;
      *min_rs = (rs < *min_rs) ? rs : *min_rs;
      *max_rs = (rs < *max_rs) ? *max_rs : rs;
   }
}

int orc_amg_setup_stats(const orc_amg *H, int max_iter, double tol, char *buf, int cap)
{
   const orc_amg_params *p = &H->prm;
   orc_text t = {buf, (size_t)(cap > 0 ? cap : 0), 0};
   int L = H->num_levels, lev, i, nd[4];
   double num_coeffs[ORC_MAXLEV], num_mem[ORC_MAXLEV], num_vars[ORC_MAXLEV];
   double grid = 0.0, oper = 0.0, mem = 0.0, total = 0.0;
   int up = (p->relax_up_type >= 0) ? p->relax_up_type : p->relax_type;

   /* par_stats.c:209-325 */
   text_add(&t, "\n\n Num MPI tasks = %d\n", 1);
   text_add(&t, " Num OpenMP threads = %d\n", 1);
   text_add(&t, " Execution policy = %s\n\n", "Device (HIP)");
   text_add(&t, " Max levels = %d\n", p->max_levels);
   text_add(&t, " Num levels = %d\n\n", L);
   text_add(&t, " Strength Threshold = %f\n", p->strong_threshold);
   text_add(&t, " Interpolation Truncation Factor = %f\n", p->trunc_factor);
   text_add(&t, " Maximum Row Sum Threshold for Dependency Weakening = %f\n\n", p->max_row_sum);
   text_add(&t, " Coarsening Type = %s \n", p->coarsen_type == 10 ? "HMIS" : "PMIS");
   if (p->agg_num_levels > 0)
   {
      text_add(&t, "\n No. of levels of aggressive coarsening: %d\n\n", p->agg_num_levels);
      text_add(&t, " Interpolation on agg. levels = matrix-matrix multipass\n");
   }
   text_add(&t, " measures are determined %s\n\n", "locally");
   text_add(&t, " Interpolation = extended+i\n");
   text_add(&t, "\nOperator Matrix Information:\n\n");

   /* par_stats.c:443-486 */
   nd[0] = orc_ndigits(H->A[0].nrows);
   nd[1] = orc_ndigits(H->A[0].ia[H->A[0].nrows]);
   nd[0] = nd[0] > 7 ? nd[0] : 7;
   nd[1] = nd[1] > 8 ? nd[1] : 8;
   nd[2] = 4;
   for (lev = 0; lev < L; lev++)
   {
      long long nnz = H->A[lev].ia[H->A[lev].nrows];
      int d = orc_ndigits(nnz / H->A[lev].nrows);
      nd[2] = d > nd[2] ? d : nd[2];
   }
   nd[2] += 2;
   nd[3] = nd[0] + nd[1] + nd[2];
   text_add(&t, "%*s", nd[0] + 13, "nonzero");
   text_add(&t, "%*s", nd[1] + 15, "entries/row");
   text_add(&t, "%18s\n", "row sums");
   text_add(&t, "%s %*s ", "lev", nd[0], "rows");
   text_add(&t, "%*s", nd[1], "entries");
   text_add(&t, "%7s %5s %4s", "sparse", "min", "max");
   text_add(&t, "%*s %8s %11s\n", nd[2] + 2, "avg", "min", "max");
   for (i = 0; i < 49 + nd[3]; i++) { text_add(&t, "="); }
   text_add(&t, "\n");

   /* par_stats.c:497-700 */
   for (lev = 0; lev < L; lev++)
   {
      const orc_csr *A = &H->A[lev];
      double nnz = (double) A->ia[A->nrows], rows = (double) A->nrows;
      int mne, mxe;
      double mnr, mxr, mnw, mxw;
      num_coeffs[lev] = nnz;
      num_mem[lev] = nnz;
      num_vars[lev] = rows;
      orc_level_stats(A, 0, &mne, &mxe, &mnr, &mxr, &mnw, &mxw);
      text_add(&t, "%3d %*d %*.0f  %0.3f  %4d %4d", lev, nd[0], A->nrows, nd[1], nnz,
               nnz / (rows * rows), mne, mxe);
      text_add(&t, "  %*.1f  %10.3e  %10.3e\n", nd[2], nnz / rows, mnr, mxr);
   }

   /* par_stats.c:702-996 */
   nd[0] = 5;
   if (L - 1)
   {
      int d = orc_ndigits(H->P[0].nrows);
      nd[0] = d > nd[0] ? d : nd[0];
   }
   text_add(&t, "\n\nInterpolation Matrix Information:\n");
   text_add(&t, "%*s ", 2 * nd[0] + 21, "entries/row");
   text_add(&t, "%10s %10s %19s\n", "min", "max", "row sums");
   text_add(&t, "lev %*s x %-*s min  max  avgW", nd[0], "rows", nd[0], "cols");
   text_add(&t, "%11s %11s %9s %11s\n", "weight", "weight", "min", "max");
   for (i = 0; i < 70 + 2 * nd[0]; i++) { text_add(&t, "="); }
   text_add(&t, "\n");
   for (lev = 0; lev < L - 1; lev++)
   {
      const orc_csr *P = &H->P[lev];
      double nnz = (double) P->ia[P->nrows];
      int mne, mxe;
      double mnr, mxr, mnw, mxw;
      num_mem[lev] += nnz;
      orc_level_stats(P, 1, &mne, &mxe, &mnr, &mxr, &mnw, &mxw);
      text_add(&t, "%3d %*d x %-*d %3d  %3d", lev, nd[0], P->nrows, nd[0], P->ncols, mne, mxe);
      text_add(&t, "  %4.1f  %10.3e  %10.3e  %10.3e  %10.3e\n",
               (nnz - (double) P->ncols) / ((double)(P->nrows - P->ncols)), mnw, mxw, mnr, mxr);
   }

   /* par_stats.c:997-1020 */
   for (lev = 0; lev < L; lev++)
   {
      mem += num_mem[lev] / num_coeffs[0];
      oper += num_coeffs[lev] / num_coeffs[0];
      total += num_vars[lev];
   }
   if (num_vars[0] != 0) { grid = total / num_vars[0]; }
   text_add(&t, "\n\n");
   text_add(&t, "     Complexity:   grid = %f\n", grid);
   text_add(&t, "               operator = %f\n", oper);
   text_add(&t, "                 memory = %f\n", mem);
   text_add(&t, "\n\n");

   /* par_stats.c:1022-1034, 1246-1322 (V-cycle, relax_order 0) */
   text_add(&t, "\n\nBoomerAMG SOLVER PARAMETERS:\n\n");
   text_add(&t, "  Maximum number of cycles:         %d \n", max_iter);
   text_add(&t, "  Stopping Tolerance:               %e \n", tol);
   text_add(&t, "  Cycle type (1 = V, 2 = W, etc.):  %d\n\n", 1);
   text_add(&t, "  Relaxation Parameters:\n");
   text_add(&t, "   Visiting Grid:                     down   up  coarse\n");
   text_add(&t, "            Number of sweeps:         %4d   %2d  %4d \n", p->num_sweeps,
            p->num_sweeps, H->coarse_sweeps);
   text_add(&t, "   Type 0=Jac, 3=hGS, 6=hSGS, 9=GE:   %4d   %2d  %4d \n", p->relax_type, up,
            H->coarse_relax);
   text_add(&t, "   Point types, partial sweeps (1=C, -1=F):\n");
   text_add(&t, "                  Pre-CG relaxation (down):");
   for (i = 0; i < p->num_sweeps; i++) { text_add(&t, "  %2d", 0); }
   text_add(&t, "\n");
   text_add(&t, "                   Post-CG relaxation (up):");
   for (i = 0; i < p->num_sweeps; i++) { text_add(&t, "  %2d", 0); }
   text_add(&t, "\n");
   text_add(&t, "                             Coarsest grid:");
   for (i = 0; i < H->coarse_sweeps; i++) { text_add(&t, "  %2d", 0); }
   text_add(&t, "\n");
   text_add(&t, "\n");
   for (lev = 0; lev < L; lev++)
   {
      if (p->relax_weight != 1) { text_add(&t, " Relaxation Weight %f level %d\n", p->relax_weight, lev); }
   }
   for (lev = 0; lev < L; lev++)
   {
      if (p->outer_weight != 1) { text_add(&t, " Outer relaxation weight %f level %d\n", p->outer_weight, lev); }
   }
   return (int) t.len;
}
