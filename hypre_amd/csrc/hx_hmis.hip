// HMIS coarsening (coarsen_type 10): src/parcsr_ls/par_coarsen.c:2842-2906.
//
// The reference runs the Ruge-Stueben first pass (hypre_BoomerAMGCoarsenRuge,
// par_coarsen.c:911, with coarsen_type 10) on the host even in its device
// builds, because the pass is sequential: each new C point is the undecided
// point of largest measure that entered that measure's list first, and every
// choice changes the measures of the points around it.  This file does the
// same: S is copied to the host, the pass runs there, and its split seeds the
// device PMIS rounds with CF_init 1 (the pass's C points are the first
// independent set).
//
// The pass keeps undecided points in FIFO buckets, one per measure value
// (the reference keeps a list of lists in decreasing measure,
// src/utilities/amg_linklist.c: a point joins the tail of its measure's list
// and the next C point is the head of the list of largest measure, which is
// the same order).
#include "hx_internal.h"

#include <algorithm>
#include <vector>

namespace hx {

namespace {

constexpr int kC = 1, kF = -1, kZ = -2, kSF = -3;

struct MeasureBuckets
{
   std::vector<int> head, tail;   // per measure value
   std::vector<int> next, prev;   // per point
   int top = 0;                   // no bucket above `top` holds a point

   explicit MeasureBuckets(int n) : next(n, -1), prev(n, -1) {}

   void push(int m, int i)
   {
      if (m >= (int) head.size()) {
         head.resize(m + 1, -1);
         tail.resize(m + 1, -1);
      }
      next[i] = -1;
      prev[i] = tail[m];
      if (tail[m] >= 0) { next[tail[m]] = i; } else { head[m] = i; }
      tail[m] = i;
      top = std::max(top, m);
   }

   void erase(int m, int i)
   {
      if (prev[i] >= 0) { next[prev[i]] = next[i]; } else { head[m] = next[i]; }
      if (next[i] >= 0) { prev[next[i]] = prev[i]; } else { tail[m] = prev[i]; }
   }

   // first point of the largest non-empty measure (-1 when all are empty)
   int front()
   {
      while (top > 0 && head[top] < 0) { top--; }
      return top > 0 ? head[top] : -1;
   }
};

} // namespace

// par_coarsen.c:1060-1391 for one rank, measure_type 0, cut_factor 0
void rs_first_pass_host(int n, const int *Si, const int *Sj, int *CF)
{
   const int nnz = Si[n];
   // S^T (stable counting sort): the points that depend on each point
   std::vector<int> STi(n + 1, 0), STj(std::max(nnz, 1));
   for (int k = 0; k < nnz; k++) { STi[Sj[k] + 1]++; }
   for (int i = 0; i < n; i++) { STi[i + 1] += STi[i]; }
   {
      std::vector<int> pos(STi.begin(), STi.end() - 1);
      for (int i = 0; i < n; i++) {
         for (int k = Si[i]; k < Si[i + 1]; k++) { STj[pos[Sj[k]]++] = i; }
      }
   }
   std::vector<int> measure(n);
   for (int i = 0; i < n; i++) { measure[i] = STi[i + 1] - STi[i]; }

   int num_left = 0;
   for (int j = 0; j < n; j++) {
      if (Si[j + 1] == Si[j]) {
         CF[j] = kSF;
         measure[j] = 0;
      } else {
         CF[j] = 0;
         num_left++;
      }
   }
   MeasureBuckets B(n);
   // points nobody depends on become Z; their neighbours' measures grow
   for (int j = 0; j < n; j++) {
      if (CF[j] == kSF) { continue; }
      if (measure[j] > 0) {
         B.push(measure[j], j);
         continue;
      }
      CF[j] = kZ;
      for (int k = Si[j]; k < Si[j + 1]; k++) {
         const int nb = Sj[k];
         if (CF[nb] == kSF) { continue; }
         if (nb < j) {
            if (measure[nb] > 0) { B.erase(measure[nb], nb); }
            B.push(++measure[nb], nb);
         } else {
            ++measure[nb];
         }
      }
      --num_left;
   }
   // a point that becomes F raises the measures of its undecided influences
   auto raise_influences = [&](int p) {
      for (int k = Si[p]; k < Si[p + 1]; k++) {
         const int q = Sj[k];
         if (CF[q] != 0) { continue; }
         B.erase(measure[q], q);
         B.push(++measure[q], q);
      }
   };
   while (num_left > 0) {
      const int c = B.front();
      if (c < 0) { fail(HYPRE_ERROR_GENERIC, "HMIS first pass: no undecided point left"); }
      CF[c] = kC;
      B.erase(measure[c], c);
      measure[c] = 0;
      --num_left;
      for (int k = STi[c]; k < STi[c + 1]; k++) {
         const int nb = STj[k];
         if (CF[nb] != 0) { continue; }
         CF[nb] = kF;
         B.erase(measure[nb], nb);
         --num_left;
         raise_influences(nb);
      }
      for (int k = Si[c]; k < Si[c + 1]; k++) {
         const int nb = Sj[k];
         if (CF[nb] != 0) { continue; }
         B.erase(measure[nb], nb);
         if (--measure[nb] > 0) {
            B.push(measure[nb], nb);
         } else {
            CF[nb] = kF;
            --num_left;
            raise_influences(nb);
         }
      }
   }
}

int hmis(const Csr &S, const double *measure_in_dev, int *CF, double *measure_keep_dev,
         const int *st_count)
{
   const int n = S.nrows;
   if (n == 0) { return 0; }
   TimedLaunch tl("setup_hmis_first_pass", 0.0);
   std::vector<int> ia(n + 1), ja(std::max<long long>(S.nnz, 1)), cf(n);
   d2h(ia.data(), S.ia, sizeof(int) * ((size_t) n + 1));
   if (S.nnz > 0) { d2h(ja.data(), S.ja, sizeof(int) * (size_t) S.nnz); }
   rs_first_pass_host(n, ia.data(), ja.data(), cf.data());
   h2d(CF, cf.data(), sizeof(int) * (size_t) n);
   return pmis(S, measure_in_dev, CF, measure_keep_dev, st_count, 1);
}

} // namespace hx
