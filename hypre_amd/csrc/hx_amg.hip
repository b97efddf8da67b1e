// BoomerAMG (setup level loop, V-cycle, AMG as solver) and PCG on the device.
//
// Host code here only sequences kernels and reads back scalars: level sizes
// during setup, and per PCG iteration the dot products that decide
// convergence.  Level operators, interpolation, its stored transpose, l1 norms
// and level vectors stay in HBM.
#include "hx_objects.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace hx {

double wall_seconds()
{
   using namespace std::chrono;
   return duration_cast<duration<double>>(steady_clock::now().time_since_epoch()).count();
}

hypre_ParVector_struct *vec_create(int n)
{
   auto *v = new hypre_ParVector_struct();
   v->n = n;
   v->d = dalloc_n<double>((size_t) std::max(n, 1));
   v->owns = true;
   return v;
}

void vec_destroy(hypre_ParVector_struct *v)
{
   if (!v) { return; }
   if (v->owns) { dfree(v->d); }
   delete v;
}

// ================================================================== AMG setup
AMGSolver::~AMGSolver() { free_hierarchy(); }

void AMGSolver::free_hierarchy()
{
   side_sync();   // solve-phase copies may still be filling
   for (auto &l : L) {
      l.Aown.free_all();
      l.P.free_all();
      l.R.free_all();
      l.S.free_all();
      l.Arev.free_all();
      l.ell.free_all();
      l.Av = Csr();
      dfree(l.A32.a32); dfree(l.P32.a32); dfree(l.R32.a32);
      dfree(l.CF); dfree(l.measure); dfree(l.l1); dfree(l.l1f);
      dfree(l.f); dfree(l.u); dfree(l.ut); dfree(l.r);
      dfree(l.f32); dfree(l.u32); dfree(l.ut32); dfree(l.r32);
   }
   L.clear();
   dfree(Amat);
   Amat = nullptr;
   num_levels = 0;
   setup_done = false;
}

double AMGSolver::theta(int lev) const
{
   if (lev < (int) prm.level_theta.size() && prm.level_theta[lev] > 0) { return prm.level_theta[lev]; }
   return prm.strong_threshold;
}

double AMGSolver::omega(int lev) const
{
   if (lev < (int) prm.level_omega.size() && prm.level_omega[lev] > 0) { return prm.level_omega[lev]; }
   return prm.relax_wt;
}

int AMGSolver::sweeps(int lev, int k) const
{
   if (lev < (int) prm.level_sweeps.size() && prm.level_sweeps[lev] > 0) { return prm.level_sweeps[lev]; }
   return prm.num_sweeps[k];
}

static void make_fp32_view(const Csr &src, Csr &dst)
{
   dst = src;                 // shares index arrays; owns only a32
   dst.a = nullptr;
   dst.a32 = dalloc_n<float>((size_t) std::max<long long>(src.nnz, 1));
   d2f((int) src.nnz, src.a, dst.a32);
}

// src/parcsr_ls/par_amg_setup.c:1002-3185 (the level loop) for coarsen_type
// 8/9/10, interp_type 6, relax 18/0 with Gaussian elimination on the coarsest
// grid.
// device word that the strength pass of `level` sets on a zero l1 norm
static int *l1_zero_flag(int level) { return device_flags() + 768 + std::min(level, 255); }

void AMGSolver::setup(const Csr &A0)
{
   double t0 = wall_seconds();
   const AMGParams &p = prm;
   if (p.agg_num_levels > 0 && (p.agg_interp_type != 8 || p.coarsen_type != 8)) {
      g_error_arg = 1;
      fail(HYPRE_ERROR_ARG, "aggressive coarsening is built for PMIS (coarsen_type 8) with "
                            "modified multipass interpolation (agg_interp_type 8) only");
   }
   free_hierarchy();
   Range range("AMGsetup");   // par_amg_setup.c:251
   TimedLaunch tl("setup", 0.0);
   L.reserve(std::max(p.max_levels, 1) + 1);
   L.emplace_back();
   L[0].A = &A0;
   // solve-phase copies of the level operators: sliced ELL where the slices'
   // padding stays within 25% of the entries (levels of at least 4096 rows).
   // A level's copy is filled on the side stream while the level's RAP runs
   // (a latency-bound kernel that leaves HBM bandwidth free), the coarsest
   // level's after the loop.
   int ell_started = 0;   // levels whose copy is under way
   auto start_ell = [&](int j) {
      Level &lv = L[j];
      if (p.solve_format == 1 && ell_plan(*lv.A, lv.ell, 0.25, 4096, p.cycle_fp32 != 0)) {
         side_after_main();
         ell_fill(*lv.A, lv.ell, p.cycle_fp32 != 0, side_stream());
      }
      ell_started = j + 1;
   };
   coarse_relax = p.relax_type[3];
   coarse_sweeps = p.num_sweeps[3];
   int level = 0;
   int a_max_row = 0;   // longest row of the current level's operator, once a product gave it
   while (true) {
      Range lrange("AMG Level-%d", level);   // par_amg_setup.c:976, 3103
      Level &lv = L[level];
      const Csr &A = *lv.A;
      const int fine_size = A.nrows;
      if (p.max_levels == 1) { break; }
      // the strength pass also yields the column counts of S (PMIS measure)
      // and, for l1-Jacobi smoothing, the level's l1 norms
      int *stc = dalloc_n<int>((size_t) std::max(fine_size, 1));
      if (p.relax_type[1] == 18 || p.relax_type[2] == 18) {
         lv.l1 = dalloc_n<double>((size_t) std::max(fine_size, 1));
         dzero(l1_zero_flag(level), sizeof(int));
      }
      double *s_thr = dalloc_n<double>((size_t) std::max(fine_size, 1));
      int *s_flg = dalloc_n<int>((size_t) std::max(fine_size, 1));
      {
         Range r("CreateS");   // par_strength.c:538
         create_s(A, theta(level), p.max_row_sum, lv.S, stc, lv.l1, l1_zero_flag(level), s_thr,
                  s_flg, level == 0);
      }
      lv.CF = dalloc_n<int>((size_t) fine_size);
      if (p.keep_setup_data) { lv.measure = dalloc_n<double>((size_t) fine_size); }
      {
         Range r("PMIS");   // par_coarsen.c:2815
         lv.pmis_rounds = (p.coarsen_type == 10) ? hmis(lv.S, nullptr, lv.CF, lv.measure, stc)
                                                 : pmis(lv.S, nullptr, lv.CF, lv.measure, stc);
      }
      dfree(stc);
      int *f2c = dalloc_n<int>((size_t) fine_size + 1);
      const bool aggressive = level < p.agg_num_levels;
      if (aggressive) {
         // par_amg_setup.c:1311-1327: the C points are coarsened again over
         // the distance-two strength graph (PMIS, CF_init 3), and
         // par_amg_setup.c:1799-1806 corrects the split with the result
         const int nc1 = coarse_map(fine_size, lv.CF, f2c);
         Csr S2;
         {
            Range r("Create2ndS");   // par_strength.c:3041
            create_2nd_s(lv.S, lv.CF, f2c, nc1, S2);
         }
         if (nc1 > 0) {
            int *CFN = dalloc_n<int>((size_t) nc1);
            {
               Range r("PMIS");
               pmis(S2, nullptr, CFN, nullptr, nullptr, 3);
            }
            Range r("CorrectCFMarker");   // par_strength.c:3127
            correct_cf_marker(fine_size, lv.CF, f2c, CFN);
            dfree(CFN);
         }
         S2.free_all();
      }
      const int coarse_size = coarse_map(fine_size, lv.CF, f2c);
      // par_amg_setup.c:1690-1725
      if (coarse_size == 0 || coarse_size == fine_size) {
         if (coarse_relax == 9 || coarse_relax == 19 || coarse_relax == 98 || coarse_relax == 99) {
            coarse_relax = p.relax_type[0];
            coarse_sweeps = 1;
         }
         dfree(f2c);
         dfree(s_thr);
         dfree(s_flg);
         if (!p.keep_setup_data) { lv.S.free_all(); }
         if (level > 0) { dfree(lv.CF); lv.CF = nullptr; }
         break;
      }
      if (coarse_size < p.min_coarse_size) {
         dfree(f2c);
         dfree(s_thr);
         dfree(s_flg);
         if (!p.keep_setup_data) { lv.S.free_all(); }
         dfree(lv.CF); lv.CF = nullptr;
         break;
      }
      if (aggressive) {
         mod_multipass_interp(A, lv.S, lv.CF, coarse_size, p.agg_trunc_factor, p.agg_P_max_elmts,
                              lv.P);
      } else {
         Range r("ExtPIInterp");   // par_lr_interp.c:5564
         extpi_interp(A, lv.S, lv.CF, f2c, coarse_size, p.trunc_factor, p.P_max_elmts, lv.P, s_thr,
                      s_flg);
      }
      dfree(s_thr);
      dfree(s_flg);
      dfree(f2c);
      if (!p.keep_setup_data) { lv.S.free_all(); }
      transpose(lv.P, lv.R);
      L.emplace_back();
      Level &nx = L[level + 1];
      Level &cur = L[level];
      start_ell(level);
      // the copy's plan found the level operator's longest row
      if (a_max_row <= 0 && cur.ell.nslices > 0) { a_max_row = cur.ell.wmax; }
      // P's rows: one entry for a C point, at most P_max_elmts otherwise
      const int pmx = aggressive ? p.agg_P_max_elmts : p.P_max_elmts;
      {
         Range r("TripleMat-RAP");   // par_csr_triplemat.c:952
         rap(cur.R, *cur.A, cur.P, nx.Aown, pmx > 0 ? std::max(pmx, 1) : 0, a_max_row,
             &a_max_row);
      }
      nx.A = &nx.Aown;
      level++;
      // par_amg_setup.c:3136-3141 (the host-only switch to CLJP coarsening
      // at par_amg_setup.c:3128 is not part of the device execution path)
      if (level == p.max_levels - 1 || coarse_size <= p.max_coarse_size) { break; }
   }
   num_levels = level + 1;
   // coarsest solve (par_amg_setup.c:3165-3185)
   coarse_n = L[level].A->nrows;
   if (coarse_relax == 9 || coarse_relax == 19 || coarse_relax == 98 || coarse_relax == 99) {
      if (coarse_n <= p.max_coarse_size && coarse_n <= 64) {
         Amat = dalloc_n<double>((size_t) coarse_n * coarse_n);
         csr_to_dense_colmajor(*L[level].A, Amat);
         coarse_relax = 9;
      } else {
         coarse_relax = p.relax_type[1];
      }
   }
   // l1 norms (par_amg_setup.c:3352-3365) and level vectors
   // option 4 for the l1 hybrid Gauss-Seidel types (par_amg_setup.c:3297-3317),
   // option 1 for l1-Jacobi, evaluated after it there and so taking precedence
   auto l1_gs = [](int t) { return t == 8 || t == 13 || t == 14; };
   for (int j = 0; j < num_levels; j++) {
      Range lrange("AMG Level-%d", j);    // par_amg_setup.c:3293-3297
      Range rrange("Relaxation-%d", j);
      Level &lv = L[j];
      const int n = lv.A->nrows;
      int opt = 0;
      if (j < num_levels - 1) {
         if (p.relax_type[1] == 18 || p.relax_type[2] == 18) { opt = 1; }
         else if (l1_gs(p.relax_type[1]) || l1_gs(p.relax_type[2])) { opt = 4; }
      } else {
         if (coarse_relax == 18) { opt = 1; }
         else if (l1_gs(coarse_relax)) { opt = 4; }
      }
      if (opt == 1 && lv.l1) {
         // computed with the level's strength matrix
         int h = 0;
         d2h(&h, l1_zero_flag(j), sizeof(int));
         if (h) { g_error_arg = 1; set_error(HYPRE_ERROR_ARG, "zero l1 norm"); }
      } else if (opt == 1) {
         lv.l1 = dalloc_n<double>((size_t) n);
         l1_norms(*lv.A, lv.l1);
      } else if (opt == 4) {
         if (!lv.l1) { lv.l1 = dalloc_n<double>((size_t) n); }
         l1_norms_diag(*lv.A, lv.l1);
      } else if (lv.l1) {
         dfree(lv.l1);
         lv.l1 = nullptr;
      }
      if (j > 0) {
         lv.f = dalloc_n<double>((size_t) n);
         lv.u = dalloc_n<double>((size_t) n);
      }
      lv.ut = dalloc_n<double>((size_t) n);
      lv.r = dalloc_n<double>((size_t) n);
      if (p.cycle_fp32) {
         make_fp32_view(*lv.A, lv.A32);
         if (j < num_levels - 1) {
            make_fp32_view(lv.P, lv.P32);
            make_fp32_view(lv.R, lv.R32);
         }
         if (lv.l1) {
            lv.l1f = dalloc_n<float>((size_t) n);
            d2f(n, lv.l1, lv.l1f);
         }
         lv.f32 = dalloc_n<float>((size_t) n);
         lv.u32 = dalloc_n<float>((size_t) n);
         lv.ut32 = dalloc_n<float>((size_t) n);
         lv.r32 = dalloc_n<float>((size_t) n);
      }
      const bool hgs = (j < num_levels - 1)
                          ? (is_hybrid_gs(p.relax_type[1]) || is_hybrid_gs(p.relax_type[2]))
                          : (coarse_relax != 9 || !Amat) && is_hybrid_gs(coarse_relax);
      lv.max_row = hgs ? csr_max_row(*lv.A) : -1;
      // row-reversed operator for backward / symmetric hybrid Gauss-Seidel
      const bool rev = (j < num_levels - 1)
                          ? (hgs_needs_reverse(p.relax_type[1]) || hgs_needs_reverse(p.relax_type[2]))
                          : (coarse_relax != 9 || !Amat) && hgs_needs_reverse(coarse_relax);
      if (rev) {
         if (p.cycle_fp32) {
            Csr tmp = *lv.A;          // shares A's arrays; adds the FP32 values
            tmp.a32 = lv.A32.a32;
            csr_reverse_rows(tmp, lv.Arev);
         } else {
            csr_reverse_rows(*lv.A, lv.Arev);
         }
      }
   }
   for (int j = ell_started; j < num_levels; j++) { start_ell(j); }
   // the solve phase runs after the copies are filled
   main_after_side();
   for (int j = 0; j < num_levels; j++) {
      Level &lv = L[j];
      lv.Av = *lv.A;
      lv.Av.ell = nullptr;
      if (lv.ell.stored > 0) {
         lv.Av.ell = &lv.ell;
         if (p.cycle_fp32) { lv.A32.ell = &lv.ell; }
      }
   }
   sync();
   setup_done = true;
   setup_seconds = wall_seconds() - t0;
   // par_amg_setup.c:3830-3833
   if (p.print_level == 1 || p.print_level == 3) { print_setup_stats(); }
}

// ================================================================== V-cycle
// src/parcsr_ls/par_cycle.c:23-889 with cycle_type 1; relax 18 through
// par_relax.c:1178 (Relax7Jacobi), relax 0 through par_relax.c:180, relax
// 3/4/6/8/13/14 through par_relax.c:692 (hybrid Gauss-Seidel).
void AMGSolver::relax_any(int type, int lev, const Csr &A, const double *f, const double *u,
                          const double *l1, double *uo, bool zero, const char *cls) const
{
   const double w = omega(lev);
   if (type == 18) {
      relax_l1(A, f, u, l1, w, uo, zero, cls);
   } else if (type == 0 || type == 7) {
      if (zero) { fill(A.nrows, 0.0, const_cast<double *>(u)); }
      relax_jacobi(A, f, u, w, uo);
   } else if (is_hybrid_gs(type)) {
      relax_hgs(type, A, &L[lev].Arev, L[lev].max_row, f, u, l1, w, prm.outer_wt,
                hgs_parts(A.nrows, prm.hgs_parts, prm.hgs_rows), uo, zero, cls);
   } else {
      fail(HYPRE_ERROR_ARG, "relaxation type not supported");
   }
}

void AMGSolver::relax_any32(int type, int lev, const Csr &A, const float *f, const float *u,
                            const float *l1, float *uo, bool zero, const char *cls) const
{
   const float w = (float) omega(lev);
   if (type == 18) {
      relax_l1_32(A, f, u, l1, w, uo, zero, cls);
   } else if (type == 0 || type == 7) {
      if (zero) { fill32(A.nrows, 0.f, const_cast<float *>(u)); }
      relax_jacobi_32(A, f, u, w, uo);
   } else if (is_hybrid_gs(type)) {
      relax_hgs_32(type, A, &L[lev].Arev, L[lev].max_row, f, u, l1, w, (float) prm.outer_wt,
                   hgs_parts(A.nrows, prm.hgs_parts, prm.hgs_rows), uo, zero, cls);
   } else {
      fail(HYPRE_ERROR_ARG, "relaxation type not supported");
   }
}

void AMGSolver::cycle(const double *f, double *u, bool u_zero)
{
   Range range("AMGCycle");   // par_cycle.c:119
   const int nl = num_levels;
   // level 0 works on the caller's buffers; the result must end in `u`
   double *cur0 = u;
   double *spare0 = L[0].ut;
   std::vector<double *> cur(nl), spare(nl);
   std::vector<const double *> F(nl);
   cur[0] = cur0; spare[0] = spare0; F[0] = f;
   for (int l = 1; l < nl; l++) { cur[l] = L[l].u; spare[l] = L[l].ut; F[l] = L[l].f; }
   bool zero = u_zero;
   bool formed = false;   // this level's first sweep was formed by the restriction
   for (int l = 0; l < nl - 1; l++) {
      Range lrange("AMG Level-%d", l);   // par_cycle.c:302
      const Level &lv = L[l];
      const Csr &A = lv.Av;
      if (formed) {
         std::swap(cur[l], spare[l]);
         zero = false;
      }
      for (int s = formed ? 1 : 0; s < sweeps(l, 1); s++) {
         Range r("Relaxation");   // par_cycle.c:430
         // level-0 sweeps from a zero guess are timed as a class of their own
         const char *cls = (l == 0) ? (zero ? "relaxz_A0" : "relax_A0") : "relax";
         relax_any(prm.relax_type[1], l, A, F[l], cur[l], lv.l1, spare[l], zero, cls);
         std::swap(cur[l], spare[l]);
         zero = false;
      }
      // residual and restriction (par_cycle.c:755-790)
      {
         Range r("Residual");   // par_cycle.c:748
         spmv(A, -1.0, cur[l], 1.0, F[l], lv.r, l == 0 ? "resid_A0" : "spmv");
      }
      Range rr("Restriction");   // par_cycle.c:768
      // a next level that opens with an l1-Jacobi sweep from a zero guess gets
      // that sweep, u = (0 + w f)/l1, from the restriction's epilogue
      formed = prm.solve_format == 1 && (l + 1 < nl - 1) && prm.relax_type[1] == 18 &&
               sweeps(l + 1, 1) >= 1;
      if (formed) {
         spmv_zero_relax(lv.R, lv.r, L[l + 1].f, L[l + 1].l1, omega(l + 1), spare[l + 1],
                         "restrict");
      } else {
         spmv(lv.R, 1.0, lv.r, 0.0, nullptr, L[l + 1].f, "restrict");
      }
      zero = true;
   }
   {
      const int l = nl - 1;
      Range lrange("AMG Level-%d", l);
      Range r("Coarse solve");   // par_cycle.c:381
      const Level &lv = L[l];
      if (coarse_relax == 9 && Amat) {
         gselim_solve(coarse_n, Amat, F[l], cur[l]);
      } else {
         for (int s = 0; s < coarse_sweeps; s++) {
            relax_any(coarse_relax, l, lv.Av, F[l], cur[l], lv.l1, spare[l], zero, "relax");
            std::swap(cur[l], spare[l]);
            zero = false;
         }
      }
   }
   for (int l = nl - 2; l >= 0; l--) {
      Range lrange("AMG Level-%d", l);   // par_cycle.c:806, 860
      const Level &lv = L[l];
      // interpolation u += P e (par_cycle.c:820-840)
      {
         Range r("Interpolation");   // par_cycle.c:827
         spmv(lv.P, 1.0, cur[l + 1], 1.0, cur[l], cur[l], "interp");
      }
      const char *cls = (l == 0) ? "relax_A0" : "relax";
      for (int s = 0; s < sweeps(l, 2); s++) {
         Range r("Relaxation");
         relax_any(prm.relax_type[2], l, lv.Av, F[l], cur[l], lv.l1, spare[l], false, cls);
         std::swap(cur[l], spare[l]);
      }
   }
   if (cur[0] != u) { copy(L[0].A->nrows, cur[0], u); }
   // keep the library's own buffers consistent for the next cycle
   for (int l = 1; l < nl; l++) { L[l].u = cur[l]; L[l].ut = spare[l]; }
   L[0].ut = (cur[0] == u) ? spare[0] : cur[0];
}

// FP32 V-cycle inside the FP64 outer iteration: f and u are FP64 at the
// boundary; everything below runs on FP32 copies; the coarsest dense solve
// computes in FP64.
void AMGSolver::cycle32(const double *f, double *u)
{
   Range range("AMGCycle");
   const int nl = num_levels;
   std::vector<float *> cur(nl), spare(nl);
   for (int l = 0; l < nl; l++) { cur[l] = L[l].u32; spare[l] = L[l].ut32; }
   d2f(L[0].A->nrows, f, L[0].f32);
   bool zero = true;
   bool formed = false;
   for (int l = 0; l < nl - 1; l++) {
      Level &lv = L[l];
      if (formed) {
         std::swap(cur[l], spare[l]);
         zero = false;
      }
      for (int s = formed ? 1 : 0; s < sweeps(l, 1); s++) {
         const char *cls = (l == 0) ? (zero ? "relaxz_A0" : "relax_A0") : "relax";
         relax_any32(prm.relax_type[1], l, lv.A32, lv.f32, cur[l], lv.l1f, spare[l], zero, cls);
         std::swap(cur[l], spare[l]);
         zero = false;
      }
      spmv32(lv.A32, -1.f, cur[l], 1.f, lv.f32, lv.r32, l == 0 ? "resid_A0" : "spmv");
      formed = prm.solve_format == 1 && (l + 1 < nl - 1) && prm.relax_type[1] == 18 &&
               sweeps(l + 1, 1) >= 1;
      if (formed) {
         spmv32_zero_relax(lv.R32, lv.r32, L[l + 1].f32, L[l + 1].l1f, (float) omega(l + 1),
                           spare[l + 1], "restrict");
      } else {
         spmv32(lv.R32, 1.f, lv.r32, 0.f, nullptr, L[l + 1].f32, "restrict");
      }
      zero = true;
   }
   {
      const int l = nl - 1;
      Level &lv = L[l];
      if (coarse_relax == 9 && Amat) {
         gselim_solve32(coarse_n, Amat, lv.f32, cur[l]);
      } else {
         for (int s = 0; s < coarse_sweeps; s++) {
            relax_any32(coarse_relax, l, lv.A32, lv.f32, cur[l], lv.l1f, spare[l], zero, "relax");
            std::swap(cur[l], spare[l]);
            zero = false;
         }
      }
   }
   for (int l = nl - 2; l >= 0; l--) {
      Level &lv = L[l];
      spmv32(lv.P32, 1.f, cur[l + 1], 1.f, cur[l], cur[l], "interp");
      const char *cls = (l == 0) ? "relax_A0" : "relax";
      for (int s = 0; s < sweeps(l, 2); s++) {
         relax_any32(prm.relax_type[2], l, lv.A32, lv.f32, cur[l], lv.l1f, spare[l], false, cls);
         std::swap(cur[l], spare[l]);
      }
   }
   f2d(L[0].A->nrows, cur[0], u);
   for (int l = 0; l < nl; l++) { L[l].u32 = cur[l]; L[l].ut32 = spare[l]; }
}

// src/parcsr_ls/par_amg_solve.c:22-424
const Csr &AMGSolver::solve_view(const Csr &A) const
{
   if (setup_done && !L.empty() && L[0].A == &A && L[0].Av.ell) { return L[0].Av; }
   return A;
}

int AMGSolver::solve(const Csr &A_in, hypre_ParVector_struct *b, hypre_ParVector_struct *x)
{
   const Csr &A = solve_view(A_in);
   const int n = A.nrows;
   const double tol = prm.tol;
   const int max_iter = prm.max_iter;
   double resid_nrm = 0.0, resid_nrm_init = 0.0, rhs_norm = 0.0, relative_resid = 1.0;
   const bool track = (tol > 0.) || prm.print_level > 1 || prm.logging > 1;
   if (track) {
      spmv(A, -1.0, x->d, 1.0, b->d, L[0].r, "spmv");
      resid_nrm = std::sqrt(dot(n, L[0].r, L[0].r));
      if (resid_nrm != resid_nrm) {
         fail(HYPRE_ERROR_GENERIC, "INFs and/or NaNs detected in input");
      }
      resid_nrm_init = resid_nrm;
      rhs_norm = std::sqrt(dot(n, b->d, b->d));
      relative_resid = (rhs_norm != 0.0) ? resid_nrm_init / rhs_norm : resid_nrm_init;
   }
   int cycle_count = 0;
   while ((relative_resid >= tol) && cycle_count < max_iter) {
      bool zero = x->all_zeros;
      if (prm.cycle_fp32) {
         // the FP32 cycle is a correction cycle: x += V(b - A x)
         if (zero) {
            cycle32(b->d, x->d);
         } else {
            double *r = L[0].r;
            double *e = L[0].ut;
            spmv(A, -1.0, x->d, 1.0, b->d, r, "spmv");
            cycle32(r, e);
            axpy(n, 1.0, e, x->d);
         }
      } else {
         cycle(b->d, x->d, zero);
      }
      x->all_zeros = false;
      if (track) {
         spmv(A, -1.0, x->d, 1.0, b->d, L[0].r, "spmv");
         resid_nrm = std::sqrt(dot(n, L[0].r, L[0].r));
         relative_resid = (rhs_norm != 0.0) ? resid_nrm / rhs_norm : resid_nrm;
      }
      ++cycle_count;
      if (!track) { break; }
   }
   num_iterations = cycle_count;
   rel_resid = relative_resid;
   if (cycle_count == max_iter && tol > 0.) { set_error(HYPRE_ERROR_CONV, "AMG did not converge"); }
   if (prm.print_level > 1 && track && cycle_count > 0 && resid_nrm_init != 0.0) {
      double conv = std::pow(resid_nrm / resid_nrm_init, 1.0 / (double) cycle_count);
      printf("\n\n Average Convergence Factor = %f\n", conv);
   }
   return g_error_flag;
}

// ================================================================== PCG
PCGSolver::~PCGSolver()
{
   vec_destroy(p); vec_destroy(s); vec_destroy(r);
}

void PCGSolver::ensure_vectors(int n_)
{
   if (n == n_ && p) { return; }
   vec_destroy(p); vec_destroy(s); vec_destroy(r);
   p = vec_create(n_); s = vec_create(n_); r = vec_create(n_);
   n = n_;
}

} // namespace hx

using namespace hx;

// ------------------------------------------------------------------ helpers
static AMGSolver *as_amg(HYPRE_Solver s)
{
   if (!s || s->kind != 1) { fail(HYPRE_ERROR_ARG, "not a BoomerAMG solver"); }
   return static_cast<AMGSolver *>(s);
}

static PCGSolver *as_pcg(HYPRE_Solver s)
{
   if (!s || s->kind != 2) { fail(HYPRE_ERROR_ARG, "not a PCG solver"); }
   return static_cast<PCGSolver *>(s);
}

static const Csr &mat_of(HYPRE_ParCSRMatrix A)
{
   if (!A) { fail(HYPRE_ERROR_ARG, "NULL matrix"); }
   return A->A;
}

static double dev_scalar_read(const double *d, int k = 1, double *out = nullptr)
{
   double *h = static_cast<double *>(pinned_scratch(64));
   HX_HIP(hipMemcpyAsync(h, d, sizeof(double) * k, hipMemcpyDeviceToHost, stream()));
   sync();
   if (out) { for (int i = 0; i < k; i++) { out[i] = h[i]; } }
   return h[0];
}

// device scalars of the PCG iteration live at byte offset 512 of the status block
static double *scal_dev(int k) { return reinterpret_cast<double *>(device_flags()) + 64 + k; }

extern "C" {

// ================================================================== BoomerAMG API
HYPRE_Int HYPRE_BoomerAMGCreate(HYPRE_Solver *solver)
{
   return guard([&] {
      if (!solver) { g_error_arg = 1; fail(HYPRE_ERROR_ARG, "NULL solver pointer"); }
      ensure_init();
      *solver = new AMGSolver();
   });
}

HYPRE_Int HYPRE_BoomerAMGDestroy(HYPRE_Solver solver)
{
   return guard([&] {
      if (!solver) { return; }
      delete as_amg(solver);
   });
}

HYPRE_Int HYPRE_BoomerAMGSetup(HYPRE_Solver solver, HYPRE_ParCSRMatrix A, HYPRE_ParVector b,
                               HYPRE_ParVector x)
{
   return guard([&] {
      if (!A) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "NULL matrix"); }
      (void) b; (void) x;
      as_amg(solver)->setup(mat_of(A));
   });
}

HYPRE_Int HYPRE_BoomerAMGSolve(HYPRE_Solver solver, HYPRE_ParCSRMatrix A, HYPRE_ParVector b,
                               HYPRE_ParVector x)
{
   return guard([&] {
      AMGSolver *s = as_amg(solver);
      if (!A) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "NULL matrix"); }
      if (!b) { g_error_arg = 3; fail(HYPRE_ERROR_ARG, "NULL rhs"); }
      if (!x) { g_error_arg = 4; fail(HYPRE_ERROR_ARG, "NULL solution"); }
      if (!s->setup_done) { fail(HYPRE_ERROR_GENERIC, "BoomerAMGSolve before setup"); }
      // the reference reads past short vectors here; on the device that is a
      // memory fault, so lengths that do not fit A and the set-up hierarchy
      // are refused
      const Csr &M = mat_of(A);
      if (M.nrows != s->L[0].A->nrows || M.ncols != M.nrows) {
         g_error_arg = 2;
         fail(HYPRE_ERROR_ARG, "matrix size differs from the set-up hierarchy");
      }
      if (b->n != M.nrows) { g_error_arg = 3; fail(HYPRE_ERROR_ARG, "rhs length differs from the matrix"); }
      if (x->n != M.nrows) { g_error_arg = 4; fail(HYPRE_ERROR_ARG, "solution length differs from the matrix"); }
      s->solve(M, b, x);
   });
}

#define HX_AMG_SET(fn, field, type)                                     \
   HYPRE_Int fn(HYPRE_Solver solver, type v)                            \
   {                                                                    \
      return guard([&] { as_amg(solver)->prm.field = v; });            \
   }

HX_AMG_SET(HYPRE_BoomerAMGSetMaxCoarseSize, max_coarse_size, HYPRE_Int)
HX_AMG_SET(HYPRE_BoomerAMGSetMinCoarseSize, min_coarse_size, HYPRE_Int)
HX_AMG_SET(HYPRE_BoomerAMGSetTruncFactor, trunc_factor, HYPRE_Real)
HX_AMG_SET(HYPRE_BoomerAMGSetMaxIter, max_iter, HYPRE_Int)
HX_AMG_SET(HYPRE_BoomerAMGSetTol, tol, HYPRE_Real)
HX_AMG_SET(HYPRE_BoomerAMGSetPrintLevel, print_level, HYPRE_Int)
HX_AMG_SET(HYPRE_BoomerAMGSetLogging, logging, HYPRE_Int)
HX_AMG_SET(HYPRE_BoomerAMGSetRelaxWt, relax_wt, HYPRE_Real)
HX_AMG_SET(HYPRE_BoomerAMGSetOuterWt, outer_wt, HYPRE_Real)
HX_AMG_SET(HYPRE_BoomerAMGSetKeepTranspose, keep_transpose, HYPRE_Int)

HYPRE_Int HYPRE_BoomerAMGSetMaxLevels(HYPRE_Solver solver, HYPRE_Int v)
{
   return guard([&] {
      if (v < 1) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "max_levels < 1"); }
      as_amg(solver)->prm.max_levels = v;
   });
}

HYPRE_Int HYPRE_BoomerAMGGetMaxLevels(HYPRE_Solver solver, HYPRE_Int *v)
{
   return guard([&] { *v = as_amg(solver)->prm.max_levels; });
}

HYPRE_Int HYPRE_BoomerAMGSetStrongThreshold(HYPRE_Solver solver, HYPRE_Real v)
{
   return guard([&] {
      if (v < 0 || v > 1) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "strong threshold out of [0,1]"); }
      as_amg(solver)->prm.strong_threshold = v;
   });
}

HYPRE_Int HYPRE_BoomerAMGGetStrongThreshold(HYPRE_Solver solver, HYPRE_Real *v)
{
   return guard([&] { *v = as_amg(solver)->prm.strong_threshold; });
}

HYPRE_Int HYPRE_BoomerAMGSetMaxRowSum(HYPRE_Solver solver, HYPRE_Real v)
{
   return guard([&] {
      if (v <= 0 || v > 1) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "max_row_sum out of (0,1]"); }
      as_amg(solver)->prm.max_row_sum = v;
   });
}

HYPRE_Int HYPRE_BoomerAMGSetPMaxElmts(HYPRE_Solver solver, HYPRE_Int v)
{
   return guard([&] {
      if (v < 0) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "P_max_elmts < 0"); }
      as_amg(solver)->prm.P_max_elmts = v;
   });
}

HYPRE_Int HYPRE_BoomerAMGSetSeqThreshold(HYPRE_Solver solver, HYPRE_Int v)
{
   return guard([&] {
      (void) as_amg(solver);
      if (v > 0) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "redundant coarse solves are not supported"); }
   });
}

HYPRE_Int HYPRE_BoomerAMGSetRedundant(HYPRE_Solver solver, HYPRE_Int v)
{
   return guard([&] {
      (void) as_amg(solver);
      if (v != 0) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "redundant coarse solves are not supported"); }
   });
}

HYPRE_Int HYPRE_BoomerAMGSetInterpType(HYPRE_Solver solver, HYPRE_Int v)
{
   return guard([&] {
      if (v != 6) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "only extended+i interpolation (6) is supported"); }
      as_amg(solver)->prm.interp_type = v;
   });
}

HYPRE_Int HYPRE_BoomerAMGSetCoarsenType(HYPRE_Solver solver, HYPRE_Int v)
{
   return guard([&] {
      if (v != 8 && v != 9 && v != 10) {
         g_error_arg = 2;
         fail(HYPRE_ERROR_ARG, "only PMIS (8, 9) and HMIS (10) coarsening are supported");
      }
      as_amg(solver)->prm.coarsen_type = v;
   });
}

HYPRE_Int HYPRE_BoomerAMGSetMeasureType(HYPRE_Solver solver, HYPRE_Int v)
{
   return guard([&] {
      if (v != 0) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "only local measures (0) are supported"); }
      as_amg(solver)->prm.measure_type = v;
   });
}

HYPRE_Int HYPRE_BoomerAMGSetCycleType(HYPRE_Solver solver, HYPRE_Int v)
{
   return guard([&] {
      if (v != 1) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "only V-cycles (1) are supported"); }
      as_amg(solver)->prm.cycle_type = v;
   });
}

HYPRE_Int HYPRE_BoomerAMGSetNumSweeps(HYPRE_Solver solver, HYPRE_Int v)
{
   return guard([&] {
      if (v < 1) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "num_sweeps < 1"); }
      AMGSolver *s = as_amg(solver);
      for (int k = 0; k < 3; k++) { s->prm.num_sweeps[k] = v; }
      s->prm.num_sweeps[3] = 1;
   });
}

HYPRE_Int HYPRE_BoomerAMGSetCycleNumSweeps(HYPRE_Solver solver, HYPRE_Int v, HYPRE_Int k)
{
   return guard([&] {
      if (k < 1 || k > 3) { g_error_arg = 3; fail(HYPRE_ERROR_ARG, "cycle index not in 1..3"); }
      if (v < 1) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "num_sweeps < 1"); }
      as_amg(solver)->prm.num_sweeps[k] = v;
   });
}

static bool relax_ok(int t)
{
   return t == 18 || t == 0 || t == 7 || t == 9 || is_hybrid_gs(t);
}

HYPRE_Int HYPRE_BoomerAMGSetRelaxType(HYPRE_Solver solver, HYPRE_Int v)
{
   return guard([&] {
      if (!relax_ok(v) || v == 9) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "relaxation type not supported"); }
      AMGSolver *s = as_amg(solver);
      for (int k = 0; k < 3; k++) { s->prm.relax_type[k] = v; }
      s->prm.relax_type[3] = 9;   // par_amg.c hypre_BoomerAMGSetRelaxType
   });
}

HYPRE_Int HYPRE_BoomerAMGSetCycleRelaxType(HYPRE_Solver solver, HYPRE_Int v, HYPRE_Int k)
{
   return guard([&] {
      if (k < 1 || k > 3) { g_error_arg = 3; fail(HYPRE_ERROR_ARG, "cycle index not in 1..3"); }
      if (!relax_ok(v) || (v == 9 && k != 3)) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "relaxation type not supported"); }
      as_amg(solver)->prm.relax_type[k] = v;
   });
}

HYPRE_Int HYPRE_BoomerAMGSetRelaxOrder(HYPRE_Solver solver, HYPRE_Int v)
{
   return guard([&] {
      if (v != 0) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "only lexicographic relaxation (0) is supported"); }
      as_amg(solver)->prm.relax_order = v;
   });
}

HYPRE_Int HYPRE_BoomerAMGSetDebugFlag(HYPRE_Solver solver, HYPRE_Int v)
{
   return guard([&] { (void) as_amg(solver); (void) v; });
}

HYPRE_Int HYPRE_BoomerAMGSetNumFunctions(HYPRE_Solver solver, HYPRE_Int v)
{
   return guard([&] {
      if (v != 1) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "systems AMG is not supported"); }
      as_amg(solver)->prm.num_functions = v;
   });
}

HYPRE_Int HYPRE_BoomerAMGSetAggNumLevels(HYPRE_Solver solver, HYPRE_Int v)
{
   return guard([&] {
      if (v < 0) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "agg_num_levels must be >= 0"); }
      as_amg(solver)->prm.agg_num_levels = v;
   });
}

// src/parcsr_ls/HYPRE_parcsr_amg.c (hypre_BoomerAMGSetAggInterpType,
// par_amg.c): this build interpolates aggressive levels with the modified
// multipass method (8) only; hypre's default 4 is accepted here and refused
// at setup when aggressive levels are requested
HYPRE_Int HYPRE_BoomerAMGSetAggInterpType(HYPRE_Solver solver, HYPRE_Int v)
{
   return guard([&] {
      if (v < 0 || v > 9) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "agg_interp_type must be 0..9"); }
      as_amg(solver)->prm.agg_interp_type = v;
   });
}

HYPRE_Int HYPRE_BoomerAMGSetAggPMaxElmts(HYPRE_Solver solver, HYPRE_Int v)
{
   return guard([&] {
      if (v < 0) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "agg_P_max_elmts must be >= 0"); }
      as_amg(solver)->prm.agg_P_max_elmts = v;
   });
}

HYPRE_Int HYPRE_BoomerAMGSetAggTruncFactor(HYPRE_Solver solver, HYPRE_Real v)
{
   return guard([&] {
      if (v < 0.0 || v >= 1.0) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "agg_trunc_factor must be in [0, 1)"); }
      as_amg(solver)->prm.agg_trunc_factor = v;
   });
}

// num_paths 1 is the only value of this build (the second strength matrix
// keeps every distance-two connection)
HYPRE_Int HYPRE_BoomerAMGSetNumPaths(HYPRE_Solver solver, HYPRE_Int v)
{
   return guard([&] {
      if (v != 1) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "only num_paths 1 is supported"); }
      (void) as_amg(solver);
   });
}

HYPRE_Int HYPRE_BoomerAMGSetRAP2(HYPRE_Solver solver, HYPRE_Int v)
{
   return guard([&] { (void) as_amg(solver); (void) v; });
}

HYPRE_Int HYPRE_BoomerAMGSetModuleRAP2(HYPRE_Solver solver, HYPRE_Int v)
{
   return guard([&] { (void) as_amg(solver); (void) v; });
}

HYPRE_Int HYPRE_BoomerAMGGetNumIterations(HYPRE_Solver solver, HYPRE_Int *v)
{
   return guard([&] { *v = as_amg(solver)->num_iterations; });
}

HYPRE_Int HYPRE_BoomerAMGGetFinalRelativeResidualNorm(HYPRE_Solver solver, HYPRE_Real *v)
{
   return guard([&] { *v = as_amg(solver)->rel_resid; });
}

HYPRE_Int HYPRE_MI355X_BoomerAMGGetNumLevels(HYPRE_Solver solver, HYPRE_Int *v)
{
   return guard([&] { *v = as_amg(solver)->num_levels; });
}

// ================================================================== PCG API
HYPRE_Int HYPRE_ParCSRPCGCreate(MPI_Comm comm, HYPRE_Solver *solver)
{
   return guard([&] {
      (void) comm;
      if (!solver) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "NULL solver pointer"); }
      ensure_init();
      *solver = new PCGSolver();
   });
}

HYPRE_Int HYPRE_ParCSRPCGDestroy(HYPRE_Solver solver)
{
   return guard([&] {
      if (!solver) { return; }
      delete as_pcg(solver);
   });
}

// src/krylov/pcg.c:198 hypre_PCGSetup: work vectors, then the preconditioner's setup
HYPRE_Int HYPRE_ParCSRPCGSetup(HYPRE_Solver solver, HYPRE_ParCSRMatrix A, HYPRE_ParVector b,
                               HYPRE_ParVector x)
{
   return guard([&] {
      Range range("PCG-Setup");   // pcg.c:215
      PCGSolver *s = as_pcg(solver);
      if (!A) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "NULL matrix"); }
      double t0 = wall_seconds();
      s->ensure_vectors(A->A.nrows);
      // the setup and solve functions are called whether or not they come
      // with a solver object (the diagonal scaling takes none), as in pcg.c
      if (s->precond_setup) { s->precond_setup(s->precond_solver, A, b, x); }
      sync();
      s->setup_seconds = wall_seconds() - t0;
   });
}

static void apply_precond(PCGSolver *s, HYPRE_ParCSRMatrix A, hypre_ParVector_struct *rhs,
                          hypre_ParVector_struct *out)
{
   // ClearVector (par_krylov_func.c:260) then the preconditioner
   fill(out->n, 0.0, out->d);
   out->all_zeros = true;
   if (s->precond) {
      s->precond(s->precond_solver, A, rhs, out);
   } else {
      copy(out->n, rhs->d, out->d);         // identity preconditioner
   }
   out->all_zeros = false;
}

// The operator PCG multiplies with: when the preconditioner is a BoomerAMG
// hierarchy set up on this same matrix, its solve-phase copy (sliced ELL, the
// same products in the same order), else the matrix itself.  Like hypre's
// device SpMV analysis data, the copy belongs to the setup: a matrix whose
// values are changed in place after setup needs a new setup.
static const Csr &pcg_operator(const PCGSolver *s, HYPRE_ParCSRMatrix A)
{
   if (s->precond_solver && s->precond_solver->kind == 1) {
      return static_cast<const AMGSolver *>(s->precond_solver)->solve_view(A->A);
   }
   return A->A;
}

// src/krylov/pcg.c:313-1013 (hypre_PCGSolve)
HYPRE_Int HYPRE_ParCSRPCGSolve(HYPRE_Solver solver, HYPRE_ParCSRMatrix A, HYPRE_ParVector b,
                               HYPRE_ParVector x)
{
   return guard([&] {
      Range range("PCG-Solve");   // pcg.c:379
      PCGSolver *s = as_pcg(solver);
      if (!A) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "NULL matrix"); }
      if (!b) { g_error_arg = 3; fail(HYPRE_ERROR_ARG, "NULL rhs"); }
      if (!x) { g_error_arg = 4; fail(HYPRE_ERROR_ARG, "NULL solution"); }
      const Csr &M = pcg_operator(s, A);
      const int n = M.nrows;
      if (M.ncols != n) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "PCG needs a square matrix"); }
      if (b->n != n) { g_error_arg = 3; fail(HYPRE_ERROR_ARG, "rhs length differs from the matrix"); }
      if (x->n != n) { g_error_arg = 4; fail(HYPRE_ERROR_ARG, "solution length differs from the matrix"); }
      s->ensure_vectors(n);
      double t0 = wall_seconds();
      hypre_ParVector_struct *p = s->p, *sv = s->s, *r = s->r;
      s->history.clear();
      s->converged = 0;
      double bi_prod;
      if (s->two_norm) {
         bi_prod = dot(n, b->d, b->d);
      } else {
         apply_precond(s, A, b, p);
         bi_prod = dot(n, p->d, b->d);
      }
      if (bi_prod != 0. && (bi_prod / bi_prod) != (bi_prod / bi_prod)) {
         fail(HYPRE_ERROR_GENERIC, "INFs and/or NaNs detected in input");
      }
      double eps = s->tol * s->tol;
      if (bi_prod > 0.0) {
         if (s->stop_crit && !s->rel_change && s->atolf <= 0) {
            eps = eps / bi_prod;
         } else if (s->atolf > 0) {
            bi_prod += s->atolf;
         } else {
            eps = std::max(s->tol * s->tol, s->atol * s->atol / bi_prod);
         }
      } else {
         copy(n, b->d, x->d);
         x->all_zeros = false;
         s->num_iterations = 0;
         s->rel_residual_norm = 0.0;
         sync();
         s->solve_seconds = wall_seconds() - t0;
         return;
      }
      // r = b - A x ; p = C r ; gamma = <r,p>
      spmv(M, -1.0, x->d, 1.0, b->d, r->d, "spmv");
      apply_precond(s, A, r, p);
      double gamma = dot(n, r->d, p->d);
      if (gamma != 0. && (gamma / gamma) != (gamma / gamma)) {
         fail(HYPRE_ERROR_GENERIC, "INFs and/or NaNs detected in input");
      }
      double i_prod_0 = s->two_norm ? dot(n, r->d, r->d) : gamma;
      double i_prod = 0.0;
      int i = 0;
      const double REAL_MIN = 2.2250738585072014e-308;
      const double TRUE_MIN = 4.9406564584124654e-324;
      double *d_sdotp = scal_dev(0), *d_alpha = scal_dev(1), *d_gam = scal_dev(2),
             *d_beta = scal_dev(4), *d_gold = scal_dev(5);
      double gamma_old = 0.0;
      if (s->print_level > 1) {
         printf("<b,b>: %e\n\n Iters       ||r||_2     conv.rate  ||r||_2/||b||_2\n", bi_prod);
         printf(" -----    ------------   ---------  ------------\n");
      }
      double prev_norm = std::sqrt(i_prod_0);
      while ((i + 1) <= s->max_iter) {
         i++;
         // s = A p, sdotp = <s,p>
         spmv_dot(M, p->d, sv->d, d_sdotp, "spmv_A0");
         double sdotp = dev_scalar_read(d_sdotp);
         if (sdotp == 0.0) {
            set_error(HYPRE_ERROR_CONV, "Zero sdotp value in PCG");
            if (i == 1) { i_prod = i_prod_0; }
            break;
         }
         double alpha = gamma / sdotp;
         if (alpha <= 0.0) {
            set_error(HYPRE_ERROR_CONV, "Negative or zero alpha value in PCG");
            if (s->skip_break < 3) { if (i == 1) { i_prod = i_prod_0; } break; }
         } else if (!(alpha >= TRUE_MIN)) {
            set_error(HYPRE_ERROR_CONV, "alpha value less than TRUE_MIN in PCG");
            if (s->skip_break < 2) { if (i == 1) { i_prod = i_prod_0; } break; }
         } else if (!(alpha >= REAL_MIN)) {
            set_error(HYPRE_ERROR_CONV, "Subnormal alpha value in PCG");
            if (s->skip_break < 1) { if (i == 1) { i_prod = i_prod_0; } break; }
         }
         gamma_old = gamma;
         // x += alpha p ; r -= alpha s
         double *ha = static_cast<double *>(pinned_scratch(64));
         ha[0] = alpha;
         HX_HIP(hipMemcpyAsync(d_alpha, ha, sizeof(double), hipMemcpyHostToDevice, stream()));
         pcg_update_xr(n, d_alpha, p->d, sv->d, x->d, r->d);
         x->all_zeros = false;
         // s = C r ; gamma = <r,s> ; i_prod = <r,r>
         apply_precond(s, A, r, sv);
         dot2_dev(n, r->d, sv->d, d_gam);
         double g2[2];
         dev_scalar_read(d_gam, 2, g2);
         gamma = g2[0];
         i_prod = s->two_norm ? g2[1] : gamma;
         double nrm = std::sqrt(i_prod);
         s->history.push_back(bi_prod > 0.0 ? std::sqrt(i_prod / bi_prod) : 0.0);
         if (s->print_level > 1) {
            printf("% 5d    %e    %f   %e\n", i, nrm, prev_norm > 0 ? nrm / prev_norm : 0.0,
                   std::sqrt(i_prod / bi_prod));
         }
         prev_norm = nrm;
         bool tentatively_converged = (i_prod / bi_prod < eps);
         if (tentatively_converged && s->recompute_residual) {
            spmv(M, -1.0, x->d, 1.0, b->d, r->d, "spmv");
            if (s->two_norm) {
               i_prod = dot(n, r->d, r->d);
            } else {
               apply_precond(s, A, r, sv);
               i_prod = dot(n, r->d, sv->d);
               gamma = i_prod;
            }
            if (i_prod / bi_prod >= eps) { tentatively_converged = false; }
         }
         if (tentatively_converged && s->rel_change && i_prod > 0.0) {
            double pi_prod = dot(n, p->d, p->d);
            double xi_prod = dot(n, x->d, x->d);
            double ratio = alpha * alpha * pi_prod / xi_prod;
            if (ratio >= eps) { tentatively_converged = false; }
         }
         if (tentatively_converged) {
            s->converged = 1;
            break;
         }
         if (gamma <= 0.0) {
            set_error(HYPRE_ERROR_CONV, "Negative or zero gamma value in PCG");
            if (s->skip_break < 3) { if (i == 1) { i_prod = i_prod_0; } break; }
         } else if (!(gamma >= TRUE_MIN)) {
            set_error(HYPRE_ERROR_CONV, "gamma value less than TRUE_MIN in PCG");
            if (s->skip_break < 2) { if (i == 1) { i_prod = i_prod_0; } break; }
         } else if (!(gamma >= REAL_MIN)) {
            set_error(HYPRE_ERROR_CONV, "Subnormal gamma value in PCG");
            if (s->skip_break < 1) { if (i == 1) { i_prod = i_prod_0; } break; }
         }
         // beta = gamma / gamma_old ; p = s + beta p
         double beta = gamma / gamma_old;
         double *hb = static_cast<double *>(pinned_scratch(64));
         hb[1] = beta;
         HX_HIP(hipMemcpyAsync(d_beta, hb + 1, sizeof(double), hipMemcpyHostToDevice, stream()));
         pcg_update_p(n, d_beta, sv->d, p->d);
      }
      (void) d_gold;
      if (i >= s->max_iter && (i_prod / bi_prod) >= eps && eps > 0) {
         set_error(HYPRE_ERROR_CONV, "Reached max iterations in PCG before convergence");
      }
      s->num_iterations = i;
      s->rel_residual_norm = (bi_prod > 0.0) ? std::sqrt(i_prod / bi_prod) : 0.0;
      sync();
      s->solve_seconds = wall_seconds() - t0;
      if (s->print_level > 0) {
         printf("\n\nIterations = %d\nFinal Relative Residual Norm = %e\n", i, s->rel_residual_norm);
      }
   });
}

#define HX_PCG_SET(fn, field, type)                                     \
   HYPRE_Int fn(HYPRE_Solver solver, type v)                            \
   {                                                                    \
      return guard([&] { as_pcg(solver)->field = v; });                \
   }

HX_PCG_SET(HYPRE_ParCSRPCGSetTol, tol, HYPRE_Real)
HX_PCG_SET(HYPRE_ParCSRPCGSetAbsoluteTol, atol, HYPRE_Real)
HX_PCG_SET(HYPRE_ParCSRPCGSetMaxIter, max_iter, HYPRE_Int)
HX_PCG_SET(HYPRE_ParCSRPCGSetStopCrit, stop_crit, HYPRE_Int)
HX_PCG_SET(HYPRE_ParCSRPCGSetTwoNorm, two_norm, HYPRE_Int)
HX_PCG_SET(HYPRE_ParCSRPCGSetRelChange, rel_change, HYPRE_Int)
HX_PCG_SET(HYPRE_ParCSRPCGSetPrintLevel, print_level, HYPRE_Int)
HX_PCG_SET(HYPRE_ParCSRPCGSetLogging, logging, HYPRE_Int)
HX_PCG_SET(HYPRE_PCGSetTol, tol, HYPRE_Real)
HX_PCG_SET(HYPRE_PCGSetAbsoluteTol, atol, HYPRE_Real)
HX_PCG_SET(HYPRE_PCGSetResidualTol, rtol, HYPRE_Real)
HX_PCG_SET(HYPRE_PCGSetAbsoluteTolFactor, atolf, HYPRE_Real)
HX_PCG_SET(HYPRE_PCGSetConvergenceFactorTol, cf_tol, HYPRE_Real)
HX_PCG_SET(HYPRE_PCGSetStopCrit, stop_crit, HYPRE_Int)
HX_PCG_SET(HYPRE_PCGSetMaxIter, max_iter, HYPRE_Int)
HX_PCG_SET(HYPRE_PCGSetTwoNorm, two_norm, HYPRE_Int)
HX_PCG_SET(HYPRE_PCGSetRelChange, rel_change, HYPRE_Int)
HX_PCG_SET(HYPRE_PCGSetRecomputeResidual, recompute_residual, HYPRE_Int)
HX_PCG_SET(HYPRE_PCGSetRecomputeResidualP, recompute_residual_p, HYPRE_Int)
HX_PCG_SET(HYPRE_PCGSetFlex, flex, HYPRE_Int)
HX_PCG_SET(HYPRE_PCGSetSkipBreak, skip_break, HYPRE_Int)
HX_PCG_SET(HYPRE_PCGSetLogging, logging, HYPRE_Int)
HX_PCG_SET(HYPRE_PCGSetPrintLevel, print_level, HYPRE_Int)

HYPRE_Int HYPRE_ParCSRPCGSetPrecond(HYPRE_Solver solver, HYPRE_PtrToParSolverFcn precond,
                                    HYPRE_PtrToParSolverFcn precond_setup,
                                    HYPRE_Solver precond_solver)
{
   return guard([&] {
      PCGSolver *s = as_pcg(solver);
      s->precond = precond;
      s->precond_setup = precond_setup;
      s->precond_solver = precond_solver;
   });
}

HYPRE_Int HYPRE_PCGSetPrecond(HYPRE_Solver solver, HYPRE_PtrToSolverFcn precond,
                              HYPRE_PtrToSolverFcn precond_setup, HYPRE_Solver precond_solver)
{
   return HYPRE_ParCSRPCGSetPrecond(solver, reinterpret_cast<HYPRE_PtrToParSolverFcn>(precond),
                                    reinterpret_cast<HYPRE_PtrToParSolverFcn>(precond_setup),
                                    precond_solver);
}

HYPRE_Int HYPRE_ParCSRPCGGetPrecond(HYPRE_Solver solver, HYPRE_Solver *precond_data)
{
   return guard([&] { *precond_data = as_pcg(solver)->precond_solver; });
}

HYPRE_Int HYPRE_PCGGetPrecond(HYPRE_Solver solver, HYPRE_Solver *precond_data)
{
   return HYPRE_ParCSRPCGGetPrecond(solver, precond_data);
}

HYPRE_Int HYPRE_ParCSRPCGGetNumIterations(HYPRE_Solver solver, HYPRE_Int *v)
{
   return guard([&] { *v = as_pcg(solver)->num_iterations; });
}

HYPRE_Int HYPRE_PCGGetNumIterations(HYPRE_Solver solver, HYPRE_Int *v)
{
   return HYPRE_ParCSRPCGGetNumIterations(solver, v);
}

HYPRE_Int HYPRE_ParCSRPCGGetFinalRelativeResidualNorm(HYPRE_Solver solver, HYPRE_Real *v)
{
   return guard([&] { *v = as_pcg(solver)->rel_residual_norm; });
}

HYPRE_Int HYPRE_PCGGetFinalRelativeResidualNorm(HYPRE_Solver solver, HYPRE_Real *v)
{
   return HYPRE_ParCSRPCGGetFinalRelativeResidualNorm(solver, v);
}

HYPRE_Int HYPRE_PCGGetConverged(HYPRE_Solver solver, HYPRE_Int *v)
{
   return guard([&] { *v = as_pcg(solver)->converged; });
}

HYPRE_Int HYPRE_PCGGetTol(HYPRE_Solver solver, HYPRE_Real *v)
{
   return guard([&] { *v = as_pcg(solver)->tol; });
}

HYPRE_Int HYPRE_PCGGetMaxIter(HYPRE_Solver solver, HYPRE_Int *v)
{
   return guard([&] { *v = as_pcg(solver)->max_iter; });
}

HYPRE_Int HYPRE_PCGSetup(HYPRE_Solver solver, HYPRE_Matrix A, HYPRE_Vector b, HYPRE_Vector x)
{
   return HYPRE_ParCSRPCGSetup(solver, reinterpret_cast<HYPRE_ParCSRMatrix>(A),
                               reinterpret_cast<HYPRE_ParVector>(b), reinterpret_cast<HYPRE_ParVector>(x));
}

HYPRE_Int HYPRE_PCGSolve(HYPRE_Solver solver, HYPRE_Matrix A, HYPRE_Vector b, HYPRE_Vector x)
{
   return HYPRE_ParCSRPCGSolve(solver, reinterpret_cast<HYPRE_ParCSRMatrix>(A),
                               reinterpret_cast<HYPRE_ParVector>(b), reinterpret_cast<HYPRE_ParVector>(x));
}

// diagonal scaling preconditioner (src/parcsr_ls/HYPRE_parcsr_pcg.c)
__global__ void k_diag_scale(int n, const int *__restrict__ ia, const double *__restrict__ a,
                             const double *__restrict__ y, double *__restrict__ x)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      x[i] = y[i] / a[ia[i]];
   }
}

HYPRE_Int HYPRE_ParCSRDiagScaleSetup(HYPRE_Solver solver, HYPRE_ParCSRMatrix A, HYPRE_ParVector y,
                                     HYPRE_ParVector x)
{
   (void) solver; (void) A; (void) y; (void) x;
   return g_error_flag;
}

HYPRE_Int HYPRE_ParCSRDiagScale(HYPRE_Solver solver, HYPRE_ParCSRMatrix A, HYPRE_ParVector y,
                                HYPRE_ParVector x)
{
   return guard([&] {
      (void) solver;
      const Csr &M = A->A;
      int nb = std::max(1, std::min((M.nrows + 255) / 256, 8192));
      k_diag_scale<<<nb, 256, 0, stream()>>>(M.nrows, M.ia, M.a, y->d, x->d);
      HX_LAUNCH_CHECK();
   });
}

// ================================================================== extensions
HYPRE_Int HYPRE_MI355X_BoomerAMGSetCyclePrecision(HYPRE_Solver solver, HYPRE_Int fp32)
{
   return guard([&] { as_amg(solver)->prm.cycle_fp32 = fp32 ? 1 : 0; });
}

static void set_level_vec(std::vector<double> &v, int level, double x)
{
   if (level < 0) {
      v.assign(64, x);
   } else {
      if ((int) v.size() <= level) { v.resize(level + 1, -1.0); }
      v[level] = x;
   }
}

HYPRE_Int HYPRE_MI355X_BoomerAMGSetLevelStrongThreshold(HYPRE_Solver solver, HYPRE_Int level,
                                                        HYPRE_Real theta)
{
   return guard([&] { set_level_vec(as_amg(solver)->prm.level_theta, level, theta); });
}

HYPRE_Int HYPRE_MI355X_BoomerAMGSetSolveFormat(HYPRE_Solver solver, HYPRE_Int format)
{
   return guard([&] {
      if (format != 0 && format != 1) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "format must be 0 or 1"); }
      as_amg(solver)->prm.solve_format = format;
   });
}

HYPRE_Int HYPRE_MI355X_BoomerAMGSetHybridPartitions(HYPRE_Solver solver, HYPRE_Int parts,
                                                    HYPRE_Int rows)
{
   return guard([&] {
      if (parts < 0) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "parts < 0"); }
      if (parts == 0 && rows < 1) { g_error_arg = 3; fail(HYPRE_ERROR_ARG, "rows < 1"); }
      AMGSolver *s = as_amg(solver);
      s->prm.hgs_parts = parts;
      if (rows > 0) { s->prm.hgs_rows = rows; }
   });
}

HYPRE_Int HYPRE_MI355X_BoomerAMGSetLevelRelaxWt(HYPRE_Solver solver, HYPRE_Int level, HYPRE_Real w)
{
   return guard([&] { set_level_vec(as_amg(solver)->prm.level_omega, level, w); });
}

HYPRE_Int HYPRE_MI355X_BoomerAMGSetLevelNumSweeps(HYPRE_Solver solver, HYPRE_Int level, HYPRE_Int sw)
{
   return guard([&] {
      auto &v = as_amg(solver)->prm.level_sweeps;
      if (level < 0) {
         v.assign(64, sw);
      } else {
         if ((int) v.size() <= level) { v.resize(level + 1, -1); }
         v[level] = sw;
      }
   });
}

HYPRE_Int HYPRE_MI355X_BoomerAMGClearLevelParams(HYPRE_Solver solver)
{
   return guard([&] {
      AMGSolver *s = as_amg(solver);
      s->prm.level_theta.clear();
      s->prm.level_omega.clear();
      s->prm.level_sweeps.clear();
   });
}

HYPRE_Int HYPRE_MI355X_BoomerAMGSetKeepSetupData(HYPRE_Solver solver, HYPRE_Int keep)
{
   return guard([&] { as_amg(solver)->prm.keep_setup_data = keep ? 1 : 0; });
}

HYPRE_Int HYPRE_MI355X_BoomerAMGGetLevelMatrix(HYPRE_Solver solver, HYPRE_Int level, char which,
                                               HYPRE_Int *nrows, HYPRE_Int *ncols, HYPRE_Int *nnz,
                                               HYPRE_Int *ia, HYPRE_Int *ja, HYPRE_Real *a)
{
   return guard([&] {
      AMGSolver *s = as_amg(solver);
      if (level < 0 || level >= s->num_levels) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "bad level"); }
      const Level &lv = s->L[level];
      const Csr *M = nullptr;
      switch (which) {
         case 'A': M = lv.A; break;
         case 'P': M = &lv.P; break;
         case 'R': M = &lv.R; break;
         case 'S': M = &lv.S; break;
         default: g_error_arg = 3; fail(HYPRE_ERROR_ARG, "which must be A, P, R or S");
      }
      if (!M || !M->ia) { fail(HYPRE_ERROR_GENERIC, "matrix not available on this level"); }
      if (nrows) { *nrows = M->nrows; }
      if (ncols) { *ncols = M->ncols; }
      if (nnz) { *nnz = (int) M->nnz; }
      if (ia) {
         d2h(ia, M->ia, sizeof(int) * ((size_t) M->nrows + 1));
         if (ja) { d2h(ja, M->ja, sizeof(int) * (size_t) M->nnz); }
         if (a && M->a) { d2h(a, M->a, sizeof(double) * (size_t) M->nnz); }
      }
   });
}

HYPRE_Int HYPRE_MI355X_BoomerAMGGetLevelCF(HYPRE_Solver solver, HYPRE_Int level, HYPRE_Int *cf)
{
   return guard([&] {
      AMGSolver *s = as_amg(solver);
      if (level < 0 || level >= s->num_levels || !s->L[level].CF) {
         g_error_arg = 2; fail(HYPRE_ERROR_ARG, "no C/F splitting on this level");
      }
      d2h(cf, s->L[level].CF, sizeof(int) * (size_t) s->L[level].A->nrows);
   });
}

HYPRE_Int HYPRE_MI355X_BoomerAMGGetLevelL1(HYPRE_Solver solver, HYPRE_Int level, HYPRE_Real *l1)
{
   return guard([&] {
      AMGSolver *s = as_amg(solver);
      if (level < 0 || level >= s->num_levels || !s->L[level].l1) {
         g_error_arg = 2; fail(HYPRE_ERROR_ARG, "no l1 norms on this level");
      }
      d2h(l1, s->L[level].l1, sizeof(double) * (size_t) s->L[level].A->nrows);
   });
}

HYPRE_Int HYPRE_MI355X_BoomerAMGGetLevelMeasure(HYPRE_Solver solver, HYPRE_Int level, HYPRE_Real *m)
{
   return guard([&] {
      AMGSolver *s = as_amg(solver);
      if (level < 0 || level >= s->num_levels || !s->L[level].measure) {
         g_error_arg = 2; fail(HYPRE_ERROR_ARG, "measure not kept on this level");
      }
      d2h(m, s->L[level].measure, sizeof(double) * (size_t) s->L[level].A->nrows);
   });
}

HYPRE_Int HYPRE_MI355X_BoomerAMGGetLevelPMISRounds(HYPRE_Solver solver, HYPRE_Int level, HYPRE_Int *r)
{
   return guard([&] {
      AMGSolver *s = as_amg(solver);
      if (level < 0 || level >= s->num_levels) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "bad level"); }
      *r = s->L[level].pmis_rounds;
   });
}

HYPRE_Int HYPRE_MI355X_BoomerAMGCycleHost(HYPRE_Solver solver, const HYPRE_Real *f, HYPRE_Real *u)
{
   return guard([&] {
      AMGSolver *s = as_amg(solver);
      if (!s->setup_done) { fail(HYPRE_ERROR_GENERIC, "cycle before setup"); }
      int n = s->L[0].A->nrows;
      double *df = dalloc_n<double>(n), *du = dalloc_n<double>(n);
      h2d(df, f, sizeof(double) * n);
      h2d(du, u, sizeof(double) * n);
      bool zero = true;
      for (int i = 0; i < n; i++) { if (u[i] != 0.0) { zero = false; break; } }
      if (s->prm.cycle_fp32) {
         if (!zero) { fail(HYPRE_ERROR_ARG, "the FP32 cycle takes a zero initial guess"); }
         s->cycle32(df, du);
      } else {
         s->cycle(df, du, zero);
      }
      d2h(u, du, sizeof(double) * n);
      dfree(df);
      dfree(du);
   });
}

HYPRE_Int HYPRE_MI355X_PCGGetResidualHistory(HYPRE_Solver solver, HYPRE_Int max_len, HYPRE_Real *hist,
                                             HYPRE_Int *len)
{
   return guard([&] {
      PCGSolver *s = as_pcg(solver);
      int m = std::min<int>(max_len, (int) s->history.size());
      for (int i = 0; i < m; i++) { hist[i] = s->history[i]; }
      if (len) { *len = (int) s->history.size(); }
   });
}

HYPRE_Int HYPRE_MI355X_PCGGetTimes(HYPRE_Solver solver, HYPRE_Real *setup_s, HYPRE_Real *solve_s)
{
   return guard([&] {
      PCGSolver *s = as_pcg(solver);
      if (setup_s) { *setup_s = s->setup_seconds; }
      if (solve_s) { *solve_s = s->solve_seconds; }
   });
}

} // extern "C"
