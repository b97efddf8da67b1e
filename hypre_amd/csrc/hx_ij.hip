// IJ assembly interface, ParCSR matrix / ParVector handles, library init and
// error API, problem generators, and the single-kernel extension entry points.
//
// IJ assembly follows hypre's host semantics (src/IJ_mv/IJMatrix_parcsr.c):
// SetValues overwrites an existing (row, col) entry or appends it,
// AddToValues accumulates or appends; at Assemble the diagonal moves to the
// front of its row -- keeping the other entries in insertion order (the
// auxiliary-matrix path, IJMatrix_parcsr.c:2784-2940) or by a swap with the
// first entry when the caller pre-sized rows with SetDiagOffdSizes
// (IJMatrix_parcsr.c:2974-2992).  The row order inside the CSR fixes the
// first-seen orders of the setup, so this is part of the parity contract.
#include "hx_objects.h"

#include <algorithm>
#include <cctype>
#include <cstdint>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>

using namespace hx;

struct hypre_IJMatrix_struct
{
   int ilower = 0, iupper = -1, jlower = 0, jupper = -1;
   int object_type = HYPRE_PARCSR;
   bool initialized = false, assembled = false;
   bool direct_insert = false;                    // SetDiagOffdSizes was called
   bool modified = false;                         // values of the assembled copy changed
   HYPRE_MemoryLocation memloc = HYPRE_MEMORY_HOST;
   int print_level = 0;
   // staged operations in call order: row, col, value, op (0 set, 1 add)
   std::vector<int>    st_row, st_col;
   std::vector<double> st_val;
   std::vector<char>   st_op;
   // assembled host copy (for updates after assembly)
   std::vector<int>    h_ia, h_ja;
   std::vector<double> h_a;
   hypre_ParCSRMatrix_struct *par = nullptr;
};

struct hypre_IJVector_struct
{
   int jlower = 0, jupper = -1;
   int object_type = HYPRE_PARCSR;
   bool initialized = false;
   HYPRE_MemoryLocation memloc = HYPRE_MEMORY_HOST;
   std::vector<double> h;                         // host staging
   hypre_ParVector_struct *par = nullptr;
};

extern "C" {

// ================================================================== init / errors
static int g_initialized = 0;
static HYPRE_MemoryLocation g_memloc = HYPRE_MEMORY_DEVICE;

HYPRE_Int HYPRE_Initialize(void)
{
   return guard([&] {
      ensure_init();
      g_initialized = 1;
   });
}

HYPRE_Int HYPRE_Finalize(void)
{
   return guard([&] {
      release_cache();
      g_initialized = 0;
   });
}

HYPRE_Int HYPRE_Initialized(void) { return g_initialized; }

HYPRE_Int HYPRE_GetError(void) { return g_error_flag; }

HYPRE_Int HYPRE_CheckError(HYPRE_Int ierr, HYPRE_Int code) { return ierr & code; }

HYPRE_Int HYPRE_GetErrorArg(void) { return g_error_arg; }

void HYPRE_DescribeError(HYPRE_Int ierr, char *descr)
{
   if (!descr) { return; }
   descr[0] = 0;
   if (ierr == 0) { strcpy(descr, "[No error] "); }
   if (ierr & HYPRE_ERROR_GENERIC) { strcat(descr, "[Generic error] "); }
   if (ierr & HYPRE_ERROR_MEMORY) { strcat(descr, "[Memory error] "); }
   if (ierr & HYPRE_ERROR_ARG) {
      char b[64];
      snprintf(b, sizeof b, "[Error in argument %d] ", g_error_arg);
      strcat(descr, b);
   }
   if (ierr & HYPRE_ERROR_CONV) { strcat(descr, "[Method did not converge] "); }
}

HYPRE_Int HYPRE_ClearAllErrors(void)
{
   g_error_flag = 0;
   g_error_arg = 0;
   return 0;
}

HYPRE_Int HYPRE_ClearError(HYPRE_Int code)
{
   g_error_flag &= ~code;
   return g_error_flag;
}

HYPRE_Int HYPRE_SetMemoryLocation(HYPRE_MemoryLocation m)
{
   g_memloc = m;
   return g_error_flag;
}

HYPRE_Int HYPRE_GetMemoryLocation(HYPRE_MemoryLocation *m)
{
   if (m) { *m = g_memloc; }
   return g_error_flag;
}

HYPRE_Int HYPRE_SetExecutionPolicy(HYPRE_ExecutionPolicy e)
{
   return guard([&] {
      if (e == HYPRE_EXEC_HOST) {
         g_error_arg = 1;
         fail(HYPRE_ERROR_ARG, "this library executes on the GPU only");
      }
   });
}

HYPRE_Int HYPRE_GetExecutionPolicy(HYPRE_ExecutionPolicy *e)
{
   if (e) { *e = HYPRE_EXEC_DEVICE; }
   return g_error_flag;
}

} // extern "C"

// ================================================================== ParCSR helpers
static hypre_ParCSRMatrix_struct *par_from_host(int n, int ncols, const int *ia, const int *ja,
                                                const double *a)
{
   auto *M = new hypre_ParCSRMatrix_struct();
   csr_from_host(M->A, n, ncols, ia, ja, a);
   return M;
}

// host CSR assembly from the staged operations
static void ij_assemble_host(hypre_IJMatrix_struct *m)
{
   const int nrows = m->iupper - m->ilower + 1;
   const size_t nops = m->st_row.size();
   std::vector<int> start((size_t) nrows + 1, 0);
   for (size_t k = 0; k < nops; k++) { start[(size_t) (m->st_row[k] - m->ilower) + 1]++; }
   for (int i = 0; i < nrows; i++) { start[i + 1] += start[i]; }
   std::vector<size_t> order(nops);
   {
      std::vector<int> fillp(start.begin(), start.end() - 1);
      for (size_t k = 0; k < nops; k++) { order[(size_t) fillp[m->st_row[k] - m->ilower]++] = k; }
   }
   const int ncols = m->jupper - m->jlower + 1;
   std::vector<int> marker((size_t) std::max(ncols, 1), -1);
   std::vector<int> ia((size_t) nrows + 1, 0);
   std::vector<int> ja;
   std::vector<double> a;
   ja.reserve(nops);
   a.reserve(nops);
   for (int i = 0; i < nrows; i++) {
      const size_t row_begin = ja.size();
      for (int t = start[i]; t < start[i + 1]; t++) {
         const size_t k = order[(size_t) t];
         const int c = m->st_col[k] - m->jlower;
         if (c < 0 || c >= ncols) {
            fail(HYPRE_ERROR_GENERIC, "column index outside [jlower, jupper]");
         }
         const int pos = marker[(size_t) c];
         if (pos >= (int) row_begin) {
            if (m->st_op[k] == 0) { a[(size_t) pos] = m->st_val[k]; }
            else { a[(size_t) pos] += m->st_val[k]; }
         } else {
            marker[(size_t) c] = (int) ja.size();
            ja.push_back(c);
            a.push_back(m->st_val[k]);
         }
      }
      // diagonal first
      const size_t row_end = ja.size();
      for (size_t q = row_begin; q < row_end; q++) {
         if (ja[q] == i + m->ilower - m->jlower) {
            if (q != row_begin) {
               if (m->direct_insert) {
                  std::swap(ja[q], ja[row_begin]);
                  std::swap(a[q], a[row_begin]);
               } else {
                  const int dj = ja[q];
                  const double da = a[q];
                  for (size_t z = q; z > row_begin; z--) {
                     ja[z] = ja[z - 1];
                     a[z] = a[z - 1];
                  }
                  ja[row_begin] = dj;
                  a[row_begin] = da;
               }
            }
            break;
         }
      }
      for (size_t q = row_begin; q < row_end; q++) { marker[(size_t) ja[q]] = -1; }
      ia[(size_t) i + 1] = (int) ja.size();
   }
   m->h_ia.swap(ia);
   m->h_ja.swap(ja);
   m->h_a.swap(a);
}

template <class T>
static std::vector<T> fetch_input(const T *p, size_t n, HYPRE_MemoryLocation loc)
{
   std::vector<T> v(n);
   if (n == 0 || !p) { return v; }
   if (loc == HYPRE_MEMORY_DEVICE) {
      hipPointerAttribute_t attr;
      if (hipPointerGetAttributes(&attr, p) == hipSuccess && attr.type == hipMemoryTypeDevice) {
         d2h(v.data(), p, n * sizeof(T));
         return v;
      }
      (void) hipGetLastError();
   }
   std::memcpy(v.data(), p, n * sizeof(T));
   return v;
}

static void ij_stage(hypre_IJMatrix_struct *m, HYPRE_Int nrows, HYPRE_Int *ncols, const HYPRE_BigInt *rows,
                     const HYPRE_BigInt *cols, const HYPRE_Complex *values, char op)
{
   if (!m) { g_error_arg = 1; fail(HYPRE_ERROR_ARG, "NULL IJ matrix"); }
   if (nrows < 0) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "nrows < 0"); }
   if (nrows == 0) { return; }
   if (!ncols) { g_error_arg = 3; fail(HYPRE_ERROR_ARG, "NULL ncols"); }
   if (!rows) { g_error_arg = 4; fail(HYPRE_ERROR_ARG, "NULL rows"); }
   if (!cols) { g_error_arg = 5; fail(HYPRE_ERROR_ARG, "NULL cols"); }
   if (!values) { g_error_arg = 6; fail(HYPRE_ERROR_ARG, "NULL values"); }
   std::vector<int> nc = fetch_input<int>(ncols, (size_t) nrows, m->memloc);
   std::vector<int> rw = fetch_input<int>(rows, (size_t) nrows, m->memloc);
   size_t total = 0;
   for (int i = 0; i < nrows; i++) { total += (size_t) nc[(size_t) i]; }
   std::vector<int> cl = fetch_input<int>(cols, total, m->memloc);
   std::vector<double> vl = fetch_input<double>(values, total, m->memloc);
   if (m->assembled) {
      // update existing entries of the assembled matrix (hypre: entries must exist)
      size_t k = 0;
      for (int i = 0; i < nrows; i++) {
         const int r = rw[(size_t) i] - m->ilower;
         for (int t = 0; t < nc[(size_t) i]; t++, k++) {
            if (r < 0 || r > m->iupper - m->ilower) { continue; }
            const int c = cl[k] - m->jlower;
            bool found = false;
            for (int q = m->h_ia[(size_t) r]; q < m->h_ia[(size_t) r + 1]; q++) {
               if (m->h_ja[(size_t) q] == c) {
                  if (op == 0) { m->h_a[(size_t) q] = vl[k]; } else { m->h_a[(size_t) q] += vl[k]; }
                  found = true;
                  break;
               }
            }
            if (!found) { fail(HYPRE_ERROR_GENERIC, "entry not in the assembled sparsity pattern"); }
         }
      }
      m->modified = true;
      return;
   }
   size_t k = 0;
   m->st_row.reserve(m->st_row.size() + total);
   m->st_col.reserve(m->st_col.size() + total);
   m->st_val.reserve(m->st_val.size() + total);
   m->st_op.reserve(m->st_op.size() + total);
   for (int i = 0; i < nrows; i++) {
      const int r = rw[(size_t) i];
      const bool local = (r >= m->ilower && r <= m->iupper);
      for (int t = 0; t < nc[(size_t) i]; t++, k++) {
         if (!local) { continue; }   // single rank: off-process rows do not exist
         m->st_row.push_back(r);
         m->st_col.push_back(cl[k]);
         m->st_val.push_back(vl[k]);
         m->st_op.push_back(op);
      }
   }
}

extern "C" {

// ================================================================== IJ matrix
HYPRE_Int HYPRE_IJMatrixCreate(MPI_Comm comm, HYPRE_BigInt ilower, HYPRE_BigInt iupper,
                               HYPRE_BigInt jlower, HYPRE_BigInt jupper, HYPRE_IJMatrix *matrix)
{
   return guard([&] {
      (void) comm;
      if (!matrix) { g_error_arg = 6; fail(HYPRE_ERROR_ARG, "NULL matrix pointer"); }
      if (ilower > iupper + 1 || ilower < 0) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "bad row range"); }
      if (jlower > jupper + 1 || jlower < 0) { g_error_arg = 4; fail(HYPRE_ERROR_ARG, "bad column range"); }
      auto *m = new hypre_IJMatrix_struct();
      m->ilower = ilower; m->iupper = iupper;
      m->jlower = jlower; m->jupper = jupper;
      *matrix = m;
   });
}

HYPRE_Int HYPRE_IJMatrixDestroy(HYPRE_IJMatrix m)
{
   return guard([&] {
      if (!m) { return; }
      if (m->par) { m->par->A.free_all(); delete m->par; }
      delete m;
   });
}

HYPRE_Int HYPRE_IJMatrixInitialize(HYPRE_IJMatrix m)
{
   return guard([&] {
      if (!m) { g_error_arg = 1; fail(HYPRE_ERROR_ARG, "NULL IJ matrix"); }
      ensure_init();
      m->initialized = true;
   });
}

HYPRE_Int HYPRE_IJMatrixInitialize_v2(HYPRE_IJMatrix m, HYPRE_MemoryLocation loc)
{
   return guard([&] {
      if (!m) { g_error_arg = 1; fail(HYPRE_ERROR_ARG, "NULL IJ matrix"); }
      ensure_init();
      m->memloc = loc;
      m->initialized = true;
   });
}

HYPRE_Int HYPRE_IJMatrixSetValues(HYPRE_IJMatrix m, HYPRE_Int nrows, HYPRE_Int *ncols,
                                  const HYPRE_BigInt *rows, const HYPRE_BigInt *cols,
                                  const HYPRE_Complex *values)
{
   return guard([&] { ij_stage(m, nrows, ncols, rows, cols, values, 0); });
}

HYPRE_Int HYPRE_IJMatrixAddToValues(HYPRE_IJMatrix m, HYPRE_Int nrows, HYPRE_Int *ncols,
                                    const HYPRE_BigInt *rows, const HYPRE_BigInt *cols,
                                    const HYPRE_Complex *values)
{
   return guard([&] { ij_stage(m, nrows, ncols, rows, cols, values, 1); });
}

HYPRE_Int HYPRE_IJMatrixAssemble(HYPRE_IJMatrix m)
{
   return guard([&] {
      if (!m) { g_error_arg = 1; fail(HYPRE_ERROR_ARG, "NULL IJ matrix"); }
      const int nrows = m->iupper - m->ilower + 1;
      const int ncols = m->jupper - m->jlower + 1;
      if (!(m->assembled && m->modified)) { ij_assemble_host(m); }
      m->modified = false;
      m->st_row.clear(); m->st_col.clear(); m->st_val.clear(); m->st_op.clear();
      m->st_row.shrink_to_fit(); m->st_col.shrink_to_fit(); m->st_val.shrink_to_fit(); m->st_op.shrink_to_fit();
      if (m->par) { m->par->A.free_all(); delete m->par; m->par = nullptr; }
      m->par = par_from_host(nrows, ncols, m->h_ia.data(), m->h_ja.data(), m->h_a.data());
      m->assembled = true;
   });
}

HYPRE_Int HYPRE_IJMatrixSetObjectType(HYPRE_IJMatrix m, HYPRE_Int type)
{
   return guard([&] {
      if (!m) { g_error_arg = 1; fail(HYPRE_ERROR_ARG, "NULL IJ matrix"); }
      if (type != HYPRE_PARCSR) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "only HYPRE_PARCSR objects"); }
      m->object_type = type;
   });
}

HYPRE_Int HYPRE_IJMatrixGetObjectType(HYPRE_IJMatrix m, HYPRE_Int *type)
{
   return guard([&] { *type = m->object_type; });
}

HYPRE_Int HYPRE_IJMatrixGetObject(HYPRE_IJMatrix m, void **object)
{
   return guard([&] {
      if (!m) { g_error_arg = 1; fail(HYPRE_ERROR_ARG, "NULL IJ matrix"); }
      if (!m->assembled || !m->par) { fail(HYPRE_ERROR_GENERIC, "IJ matrix not assembled"); }
      *object = m->par;
   });
}

HYPRE_Int HYPRE_IJMatrixSetRowSizes(HYPRE_IJMatrix m, const HYPRE_Int *sizes)
{
   return guard([&] {
      if (!m) { g_error_arg = 1; fail(HYPRE_ERROR_ARG, "NULL IJ matrix"); }
      (void) sizes;
   });
}

HYPRE_Int HYPRE_IJMatrixSetDiagOffdSizes(HYPRE_IJMatrix m, const HYPRE_Int *diag_sizes,
                                         const HYPRE_Int *offdiag_sizes)
{
   return guard([&] {
      if (!m) { g_error_arg = 1; fail(HYPRE_ERROR_ARG, "NULL IJ matrix"); }
      (void) diag_sizes; (void) offdiag_sizes;
      m->direct_insert = true;
   });
}

HYPRE_Int HYPRE_IJMatrixSetMaxOffProcElmts(HYPRE_IJMatrix m, HYPRE_Int v)
{
   (void) m; (void) v;
   return g_error_flag;
}

HYPRE_Int HYPRE_IJMatrixSetPrintLevel(HYPRE_IJMatrix m, HYPRE_Int v)
{
   return guard([&] { m->print_level = v; });
}

HYPRE_Int HYPRE_IJMatrixGetLocalRange(HYPRE_IJMatrix m, HYPRE_BigInt *il, HYPRE_BigInt *iu,
                                      HYPRE_BigInt *jl, HYPRE_BigInt *ju)
{
   return guard([&] {
      *il = m->ilower; *iu = m->iupper; *jl = m->jlower; *ju = m->jupper;
   });
}

HYPRE_Int HYPRE_IJMatrixGetRowCounts(HYPRE_IJMatrix m, HYPRE_Int nrows, HYPRE_BigInt *rows,
                                     HYPRE_Int *ncols)
{
   return guard([&] {
      if (!m->assembled) { fail(HYPRE_ERROR_GENERIC, "IJ matrix not assembled"); }
      for (int i = 0; i < nrows; i++) {
         int r = rows[i] - m->ilower;
         ncols[i] = m->h_ia[(size_t) r + 1] - m->h_ia[(size_t) r];
      }
   });
}

HYPRE_Int HYPRE_IJMatrixGetValues(HYPRE_IJMatrix m, HYPRE_Int nrows, HYPRE_Int *ncols,
                                  HYPRE_BigInt *rows, HYPRE_BigInt *cols, HYPRE_Complex *values)
{
   return guard([&] {
      if (!m->assembled) { fail(HYPRE_ERROR_GENERIC, "IJ matrix not assembled"); }
      size_t k = 0;
      for (int i = 0; i < nrows; i++) {
         int r = rows[i] - m->ilower;
         for (int t = 0; t < ncols[i]; t++, k++) {
            int c = cols[k] - m->jlower;
            double v = 0.0;
            for (int q = m->h_ia[(size_t) r]; q < m->h_ia[(size_t) r + 1]; q++) {
               if (m->h_ja[(size_t) q] == c) { v = m->h_a[(size_t) q]; break; }
            }
            values[k] = v;
         }
      }
   });
}

// ================================================================== IJ matrix files
// One rank: the reference's per-rank files carry the rank suffix .00000
// (src/IJ_mv/IJMatrix.c:126-134, par_csr_matrix.c:1000, 1194).
static std::string rank_file(const char *name, const char *ext)
{
   return std::string(name) + ".00000" + ext;
}

struct FileCloser
{
   FILE *fp;
   ~FileCloser() { if (fp) { fclose(fp); } }
};

// MatrixMarket banner and size line (the checks of src/utilities/mmio.c as
// the reference applies them in src/IJ_mv/IJMatrix.c:136-170): a sparse
// real or integer coordinate matrix, general or symmetric.
static void mm_header(FILE *fp, int &nrow, int &ncol, int &nnz, bool &sym)
{
   char line[1025];
   if (!fgets(line, sizeof line, fp)) { fail(HYPRE_ERROR_GENERIC, "Could not process Matrix Market banner."); }
   char banner[64] = {0}, object[64] = {0}, format[64] = {0}, field[64] = {0}, symm[64] = {0};
   if (sscanf(line, "%63s %63s %63s %63s %63s", banner, object, format, field, symm) != 5 ||
       std::strcmp(banner, "%%MatrixMarket") != 0) {
      fail(HYPRE_ERROR_GENERIC, "Could not process Matrix Market banner.");
   }
   auto lower = [](char *t) { for (; *t; t++) { *t = (char) std::tolower((unsigned char) *t); } };
   lower(object); lower(format); lower(field); lower(symm);
   if (std::strcmp(object, "matrix") != 0) { fail(HYPRE_ERROR_GENERIC, "Invalid Matrix Market file."); }
   const bool real = !std::strcmp(field, "real") || !std::strcmp(field, "integer");
   if (!real || std::strcmp(format, "coordinate") != 0) {
      fail(HYPRE_ERROR_GENERIC, "Only sparse real-valued/integer coordinate matrices are supported");
   }
   if (!std::strcmp(symm, "symmetric")) {
      sym = true;
   } else if (!std::strcmp(symm, "general")) {
      sym = false;
   } else {
      fail(HYPRE_ERROR_GENERIC, "Only general and symmetric Matrix Market matrices are supported");
   }
   // comment lines, then "rows cols entries"
   while (true) {
      if (!fgets(line, sizeof line, fp)) { fail(HYPRE_ERROR_GENERIC, "MM read size error !"); }
      if (line[0] == '%') { continue; }
      if (sscanf(line, "%d %d %d", &nrow, &ncol, &nnz) == 3) { break; }
      bool blank = true;
      for (char *t = line; *t; t++) { if (!std::isspace((unsigned char) *t)) { blank = false; } }
      if (!blank) { fail(HYPRE_ERROR_GENERIC, "MM read size error !"); }
   }
}

// src/IJ_mv/IJMatrix.c:109-237 hypre_IJMatrixRead (IJ text or MatrixMarket)
static void ij_matrix_read(const char *filename, MPI_Comm comm, HYPRE_Int type,
                           HYPRE_IJMatrix *matrix, bool mm)
{
   if (!filename) { g_error_arg = 1; fail(HYPRE_ERROR_ARG, "NULL file name"); }
   if (!matrix) { g_error_arg = 4; fail(HYPRE_ERROR_ARG, "NULL matrix pointer"); }
   const std::string path = mm ? std::string(filename) : rank_file(filename, "");
   FileCloser fc{fopen(path.c_str(), "r")};
   if (!fc.fp) { g_error_arg = 1; fail(HYPRE_ERROR_ARG, "cannot open " + path); }
   int il, iu, jl, ju;
   bool sym = false;
   if (mm) {
      int nrow, ncol, nnz;
      mm_header(fc.fp, nrow, ncol, nnz, sym);
      il = 0; iu = nrow - 1; jl = 0; ju = ncol - 1;
   } else if (fscanf(fc.fp, "%d %d %d %d", &il, &iu, &jl, &ju) != 4) {
      fail(HYPRE_ERROR_GENERIC, "Error in IJ matrix input file.");
   }
   std::vector<int> rows, cols;
   std::vector<double> vals;
   int I, J, ret;
   double v;
   // whitespace must follow the column index (IJMatrix.c:184-186)
   while ((ret = fscanf(fc.fp, "%d %d%*[ \t]%le", &I, &J, &v)) != EOF) {
      if (ret != 3) { fail(HYPRE_ERROR_GENERIC, "Error in IJ matrix input file."); }
      if (mm) { I--; J--; }
      rows.push_back(I); cols.push_back(J); vals.push_back(v);
      if (sym && I != J) { rows.push_back(J); cols.push_back(I); vals.push_back(v); }
   }
   HYPRE_IJMatrix m = nullptr;
   HYPRE_IJMatrixCreate(comm, il, iu, jl, ju, &m);
   if (g_error_flag) { fail(HYPRE_ERROR_GENERIC, "IJ matrix create failed"); }
   HYPRE_IJMatrixSetObjectType(m, type);
   HYPRE_IJMatrixInitialize_v2(m, HYPRE_MEMORY_HOST);
   if (!rows.empty()) {
      std::vector<int> ones(rows.size(), 1);
      ij_stage(m, (int) rows.size(), ones.data(), rows.data(), cols.data(), vals.data(), 0);
   }
   HYPRE_IJMatrixAssemble(m);
   *matrix = m;
}

HYPRE_Int HYPRE_IJMatrixRead(const char *filename, MPI_Comm comm, HYPRE_Int type, HYPRE_IJMatrix *matrix)
{
   return guard([&] { ij_matrix_read(filename, comm, type, matrix, false); });
}

HYPRE_Int HYPRE_IJMatrixReadMM(const char *filename, MPI_Comm comm, HYPRE_Int type,
                               HYPRE_IJMatrix *matrix)
{
   return guard([&] { ij_matrix_read(filename, comm, type, matrix, true); });
}

// src/parcsr_mv/par_csr_matrix.c:990-1060 hypre_ParCSRMatrixPrintIJ
HYPRE_Int HYPRE_IJMatrixPrint(HYPRE_IJMatrix m, const char *filename)
{
   return guard([&] {
      if (!m || m->object_type != HYPRE_PARCSR) { g_error_arg = 1; fail(HYPRE_ERROR_ARG, "bad IJ matrix"); }
      if (!m->assembled) { fail(HYPRE_ERROR_GENERIC, "IJ matrix not assembled"); }
      const std::string path = rank_file(filename, "");
      FileCloser fc{fopen(path.c_str(), "w")};
      if (!fc.fp) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "cannot open " + path); }
      fprintf(fc.fp, "%d %d %d %d\n", m->ilower, m->iupper, m->jlower, m->jupper);
      const int nrows = m->iupper - m->ilower + 1;
      for (int i = 0; i < nrows; i++) {
         for (int q = m->h_ia[(size_t) i]; q < m->h_ia[(size_t) i + 1]; q++) {
            fprintf(fc.fp, "%d %d %.14e\n", i + m->ilower, m->h_ja[(size_t) q] + m->jlower,
                    m->h_a[(size_t) q]);
         }
      }
   });
}

static void read_exact_n(FILE *fp, void *dst, size_t elem, size_t n, const char *what)
{
   if (n && fread(dst, elem, n, fp) != n) { fail(HYPRE_ERROR_GENERIC, what); }
}

// src/IJ_mv/IJMatrix.c:252-500 hypre_IJMatrixReadBinary: header of 11 uint64
// (version 1, index bytes, value bytes, global rows, columns and nonzeros,
// local nonzeros, ilower, iupper, jlower, jupper), then the row index of
// every entry, its column index and its value
HYPRE_Int HYPRE_IJMatrixReadBinary(const char *filename, MPI_Comm comm, HYPRE_Int type,
                                   HYPRE_IJMatrix *matrix)
{
   return guard([&] {
      if (!filename) { g_error_arg = 1; fail(HYPRE_ERROR_ARG, "NULL file name"); }
      if (!matrix) { g_error_arg = 4; fail(HYPRE_ERROR_ARG, "NULL matrix pointer"); }
      const std::string path = rank_file(filename, ".bin");
      FileCloser fc{fopen(path.c_str(), "rb")};
      if (!fc.fp) { fail(HYPRE_ERROR_GENERIC, "Could not open input file " + path); }
      uint64_t h[11];
      read_exact_n(fc.fp, h, sizeof(*h), 11, "Could not read header entries");
      if (h[0] != 1) { fail(HYPRE_ERROR_GENERIC, "Unsupported header version"); }
      if (h[6] > (uint64_t) INT32_MAX) { fail(HYPRE_ERROR_GENERIC, "Detected integer overflow at 7th header entry"); }
      if (h[1] != 4 && h[1] != 8) { fail(HYPRE_ERROR_GENERIC, "Unsupported data type for row/column indices"); }
      if (h[2] != 4 && h[2] != 8) { fail(HYPRE_ERROR_GENERIC, "Unsupported data type for matrix coefficients"); }
      const size_t nnz = (size_t) h[6];
      for (int k = 7; k < 11; k++) {
         if (h[k] > (uint64_t) INT32_MAX) { fail(HYPRE_ERROR_GENERIC, "row or column range beyond 32-bit indices"); }
      }
      auto read_idx = [&](std::vector<int> &out, const char *what) {
         out.resize(nnz);
         if (h[1] == 4) {
            std::vector<uint32_t> b(nnz);
            read_exact_n(fc.fp, b.data(), sizeof(*b.data()), nnz, what);
            for (size_t k = 0; k < nnz; k++) { out[k] = (int) b[k]; }
         } else {
            std::vector<uint64_t> b(nnz);
            read_exact_n(fc.fp, b.data(), sizeof(*b.data()), nnz, what);
            for (size_t k = 0; k < nnz; k++) {
               if (b[k] > (uint64_t) INT32_MAX) { fail(HYPRE_ERROR_GENERIC, "index beyond 32 bits"); }
               out[k] = (int) b[k];
            }
         }
      };
      std::vector<int> rows, cols;
      read_idx(rows, "Could not read all row indices");
      read_idx(cols, "Could not read all column indices");
      std::vector<double> vals(nnz);
      if (h[2] == 4) {
         std::vector<float> b(nnz);
         read_exact_n(fc.fp, b.data(), sizeof(*b.data()), nnz, "Could not read all matrix coefficients");
         for (size_t k = 0; k < nnz; k++) { vals[k] = (double) b[k]; }
      } else {
         read_exact_n(fc.fp, vals.data(), sizeof(*vals.data()), nnz, "Could not read all matrix coefficients");
      }
      HYPRE_IJMatrix m = nullptr;
      HYPRE_IJMatrixCreate(comm, (int) h[7], (int) h[8], (int) h[9], (int) h[10], &m);
      if (g_error_flag) { fail(HYPRE_ERROR_GENERIC, "IJ matrix create failed"); }
      HYPRE_IJMatrixSetObjectType(m, type);
      HYPRE_IJMatrixInitialize_v2(m, HYPRE_MEMORY_HOST);
      if (nnz) {
         std::vector<int> ones(nnz, 1);
         ij_stage(m, (int) nnz, ones.data(), rows.data(), cols.data(), vals.data(), 0);
      }
      HYPRE_IJMatrixAssemble(m);
      *matrix = m;
   });
}

// src/parcsr_mv/par_csr_matrix.c:1120-1370 hypre_ParCSRMatrixPrintBinaryIJ
HYPRE_Int HYPRE_IJMatrixPrintBinary(HYPRE_IJMatrix m, const char *filename)
{
   return guard([&] {
      if (!m || m->object_type != HYPRE_PARCSR) { g_error_arg = 1; fail(HYPRE_ERROR_ARG, "bad IJ matrix"); }
      if (!m->assembled) { fail(HYPRE_ERROR_GENERIC, "IJ matrix not assembled"); }
      const std::string path = rank_file(filename, ".bin");
      FileCloser fc{fopen(path.c_str(), "wb")};
      if (!fc.fp) { fail(HYPRE_ERROR_GENERIC, "Could not open output file " + path); }
      const int nrows = m->iupper - m->ilower + 1;
      const size_t nnz = (size_t) m->h_ia[(size_t) nrows];
      const uint64_t h[11] = {1, sizeof(HYPRE_BigInt), sizeof(HYPRE_Complex), (uint64_t) nrows,
                              (uint64_t) (m->jupper - m->jlower + 1), (uint64_t) nnz, (uint64_t) nnz,
                              (uint64_t) m->ilower, (uint64_t) m->iupper, (uint64_t) m->jlower,
                              (uint64_t) m->jupper};
      std::vector<HYPRE_BigInt> ri(nnz), ci(nnz);
      for (int i = 0; i < nrows; i++) {
         for (int q = m->h_ia[(size_t) i]; q < m->h_ia[(size_t) i + 1]; q++) {
            ri[(size_t) q] = i + m->ilower;
            ci[(size_t) q] = m->h_ja[(size_t) q] + m->jlower;
         }
      }
      if (fwrite(h, sizeof(uint64_t), 11, fc.fp) != 11 ||
          fwrite(ri.data(), sizeof(HYPRE_BigInt), nnz, fc.fp) != nnz ||
          fwrite(ci.data(), sizeof(HYPRE_BigInt), nnz, fc.fp) != nnz ||
          fwrite(m->h_a.data(), sizeof(HYPRE_Complex), nnz, fc.fp) != nnz) {
         fail(HYPRE_ERROR_GENERIC, "Could not write the binary IJ matrix");
      }
   });
}

// ================================================================== IJ vector
HYPRE_Int HYPRE_IJVectorCreate(MPI_Comm comm, HYPRE_BigInt jlower, HYPRE_BigInt jupper,
                               HYPRE_IJVector *vector)
{
   return guard([&] {
      (void) comm;
      if (!vector) { g_error_arg = 4; fail(HYPRE_ERROR_ARG, "NULL vector pointer"); }
      if (jlower > jupper + 1 || jlower < 0) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "bad range"); }
      auto *v = new hypre_IJVector_struct();
      v->jlower = jlower;
      v->jupper = jupper;
      *vector = v;
   });
}

HYPRE_Int HYPRE_IJVectorDestroy(HYPRE_IJVector v)
{
   return guard([&] {
      if (!v) { return; }
      vec_destroy(v->par);
      delete v;
   });
}

static void ijv_init(hypre_IJVector_struct *v)
{
   ensure_init();
   const int n = v->jupper - v->jlower + 1;
   v->h.assign((size_t) n, 0.0);
   if (!v->par) { v->par = vec_create(n); }
   fill(n, 0.0, v->par->d);
   v->par->all_zeros = true;
   v->initialized = true;
}

HYPRE_Int HYPRE_IJVectorInitialize(HYPRE_IJVector v)
{
   return guard([&] {
      if (!v) { g_error_arg = 1; fail(HYPRE_ERROR_ARG, "NULL IJ vector"); }
      ijv_init(v);
   });
}

HYPRE_Int HYPRE_IJVectorInitialize_v2(HYPRE_IJVector v, HYPRE_MemoryLocation loc)
{
   return guard([&] {
      if (!v) { g_error_arg = 1; fail(HYPRE_ERROR_ARG, "NULL IJ vector"); }
      v->memloc = loc;
      ijv_init(v);
   });
}

static void ijv_set(hypre_IJVector_struct *v, HYPRE_Int nvalues, const HYPRE_BigInt *indices,
                    const HYPRE_Complex *values, bool add)
{
   if (!v) { g_error_arg = 1; fail(HYPRE_ERROR_ARG, "NULL IJ vector"); }
   if (!v->initialized) { fail(HYPRE_ERROR_GENERIC, "IJ vector not initialized"); }
   if (nvalues <= 0) { return; }
   if (!values) { g_error_arg = 4; fail(HYPRE_ERROR_ARG, "NULL values"); }
   std::vector<double> vals = fetch_input<double>(values, (size_t) nvalues, v->memloc);
   const int n = v->jupper - v->jlower + 1;
   if (!indices) {
      for (int k = 0; k < nvalues && k < n; k++) {
         v->h[(size_t) k] = add ? v->h[(size_t) k] + vals[(size_t) k] : vals[(size_t) k];
      }
   } else {
      std::vector<int> idx = fetch_input<int>(indices, (size_t) nvalues, v->memloc);
      for (int k = 0; k < nvalues; k++) {
         const int i = idx[(size_t) k] - v->jlower;
         if (i < 0 || i >= n) { continue; }
         v->h[(size_t) i] = add ? v->h[(size_t) i] + vals[(size_t) k] : vals[(size_t) k];
      }
   }
}

HYPRE_Int HYPRE_IJVectorSetValues(HYPRE_IJVector v, HYPRE_Int nvalues, const HYPRE_BigInt *indices,
                                  const HYPRE_Complex *values)
{
   return guard([&] { ijv_set(v, nvalues, indices, values, false); });
}

HYPRE_Int HYPRE_IJVectorAddToValues(HYPRE_IJVector v, HYPRE_Int nvalues, const HYPRE_BigInt *indices,
                                    const HYPRE_Complex *values)
{
   return guard([&] { ijv_set(v, nvalues, indices, values, true); });
}

HYPRE_Int HYPRE_IJVectorAssemble(HYPRE_IJVector v)
{
   return guard([&] {
      if (!v) { g_error_arg = 1; fail(HYPRE_ERROR_ARG, "NULL IJ vector"); }
      if (!v->initialized) { ijv_init(v); }
      h2d(v->par->d, v->h.data(), sizeof(double) * v->h.size());
      bool z = true;
      for (double x : v->h) { if (x != 0.0) { z = false; break; } }
      v->par->all_zeros = z;
   });
}

HYPRE_Int HYPRE_IJVectorGetValues(HYPRE_IJVector v, HYPRE_Int nvalues, const HYPRE_BigInt *indices,
                                  HYPRE_Complex *values)
{
   return guard([&] {
      if (!v || !v->par) { g_error_arg = 1; fail(HYPRE_ERROR_ARG, "vector not initialized"); }
      const int n = v->jupper - v->jlower + 1;
      std::vector<double> h((size_t) n);
      d2h(h.data(), v->par->d, sizeof(double) * (size_t) n);
      for (int k = 0; k < nvalues; k++) {
         int i = indices ? indices[k] - v->jlower : k;
         values[k] = (i >= 0 && i < n) ? h[(size_t) i] : 0.0;
      }
   });
}

HYPRE_Int HYPRE_IJVectorSetObjectType(HYPRE_IJVector v, HYPRE_Int type)
{
   return guard([&] {
      if (type != HYPRE_PARCSR) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "only HYPRE_PARCSR objects"); }
      v->object_type = type;
   });
}

HYPRE_Int HYPRE_IJVectorGetObjectType(HYPRE_IJVector v, HYPRE_Int *type)
{
   return guard([&] { *type = v->object_type; });
}

HYPRE_Int HYPRE_IJVectorGetObject(HYPRE_IJVector v, void **object)
{
   return guard([&] {
      if (!v || !v->par) { g_error_arg = 1; fail(HYPRE_ERROR_ARG, "vector not initialized"); }
      *object = v->par;
   });
}

HYPRE_Int HYPRE_IJVectorGetLocalRange(HYPRE_IJVector v, HYPRE_BigInt *jl, HYPRE_BigInt *ju)
{
   return guard([&] { *jl = v->jlower; *ju = v->jupper; });
}

HYPRE_Int HYPRE_IJVectorSetNumComponents(HYPRE_IJVector v, HYPRE_Int nc)
{
   return guard([&] {
      (void) v;
      if (nc != 1) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "multi-component vectors are not supported"); }
   });
}

HYPRE_Int HYPRE_IJVectorSetComponent(HYPRE_IJVector v, HYPRE_Int c)
{
   return guard([&] {
      (void) v;
      if (c != 0) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "component out of range"); }
   });
}

HYPRE_Int HYPRE_IJVectorSetMaxOffProcElmts(HYPRE_IJVector v, HYPRE_Int m)
{
   (void) v; (void) m;
   return g_error_flag;
}

// src/IJ_mv/HYPRE_IJVector.c:862-925 (text) and par_vector.c:1360-1440 (binary)
static std::vector<double> ijv_values(hypre_IJVector_struct *v)
{
   if (!v) { g_error_arg = 1; fail(HYPRE_ERROR_ARG, "NULL IJ vector"); }
   if (!v->par) { fail(HYPRE_ERROR_GENERIC, "IJ vector not initialized"); }
   const int n = v->jupper - v->jlower + 1;
   std::vector<double> h((size_t) std::max(n, 0));
   if (n > 0) { d2h(h.data(), v->par->d, sizeof(double) * (size_t) n); }
   return h;
}

HYPRE_Int HYPRE_IJVectorPrint(HYPRE_IJVector v, const char *filename)
{
   return guard([&] {
      const std::vector<double> h = ijv_values(v);
      const std::string path = rank_file(filename, "");
      FileCloser fc{fopen(path.c_str(), "w")};
      if (!fc.fp) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "cannot open " + path); }
      fprintf(fc.fp, "%d %d\n", v->jlower, v->jupper);
      for (size_t i = 0; i < h.size(); i++) { fprintf(fc.fp, "%d %.14e\n", (int) i + v->jlower, h[i]); }
   });
}

HYPRE_Int HYPRE_IJVectorPrintBinary(HYPRE_IJVector v, const char *filename)
{
   return guard([&] {
      const std::vector<double> h = ijv_values(v);
      const std::string path = rank_file(filename, ".bin");
      FileCloser fc{fopen(path.c_str(), "wb")};
      if (!fc.fp) { fail(HYPRE_ERROR_GENERIC, "Could not open output file " + path); }
      // version, value bytes, partitioning [jlower, jupper+1), global size,
      // local size, components, storage method
      const uint64_t hd[8] = {1, sizeof(HYPRE_Complex), (uint64_t) v->jlower,
                              (uint64_t) (v->jupper + 1), (uint64_t) h.size(), (uint64_t) h.size(), 1,
                              0};
      if (fwrite(hd, sizeof(uint64_t), 8, fc.fp) != 8 ||
          fwrite(h.data(), sizeof(double), h.size(), fc.fp) != h.size()) {
         fail(HYPRE_ERROR_GENERIC, "Could not write the binary IJ vector");
      }
   });
}

// src/IJ_mv/HYPRE_IJVector.c:785-842 HYPRE_IJVectorRead
HYPRE_Int HYPRE_IJVectorRead(const char *filename, MPI_Comm comm, HYPRE_Int type,
                             HYPRE_IJVector *vector)
{
   return guard([&] {
      if (!filename) { g_error_arg = 1; fail(HYPRE_ERROR_ARG, "NULL file name"); }
      if (!vector) { g_error_arg = 4; fail(HYPRE_ERROR_ARG, "NULL vector pointer"); }
      const std::string path = rank_file(filename, "");
      FileCloser fc{fopen(path.c_str(), "r")};
      if (!fc.fp) { g_error_arg = 1; fail(HYPRE_ERROR_ARG, "cannot open " + path); }
      int jl, ju;
      if (fscanf(fc.fp, "%d %d", &jl, &ju) != 2) { fail(HYPRE_ERROR_GENERIC, "Error in IJ vector input file."); }
      std::vector<int> idx;
      std::vector<double> val;
      int j, ret;
      double x;
      while ((ret = fscanf(fc.fp, "%d%*[ \t]%le", &j, &x)) != EOF) {
         if (ret != 2) { fail(HYPRE_ERROR_GENERIC, "Error in IJ vector input file."); }
         idx.push_back(j);
         val.push_back(x);
      }
      HYPRE_IJVector vec = nullptr;
      HYPRE_IJVectorCreate(comm, jl, ju, &vec);
      if (g_error_flag) { fail(HYPRE_ERROR_GENERIC, "IJ vector create failed"); }
      HYPRE_IJVectorSetObjectType(vec, type);
      HYPRE_IJVectorInitialize_v2(vec, HYPRE_MEMORY_HOST);
      if (!idx.empty()) { ijv_set(vec, (int) idx.size(), idx.data(), val.data(), false); }
      HYPRE_IJVectorAssemble(vec);
      *vector = vec;
   });
}

// src/IJ_mv/IJVector.c:87-245 hypre_IJVectorReadBinary
HYPRE_Int HYPRE_IJVectorReadBinary(const char *filename, MPI_Comm comm, HYPRE_Int type,
                                   HYPRE_IJVector *vector)
{
   return guard([&] {
      if (!filename) { g_error_arg = 1; fail(HYPRE_ERROR_ARG, "NULL file name"); }
      if (!vector) { g_error_arg = 4; fail(HYPRE_ERROR_ARG, "NULL vector pointer"); }
      const std::string path = rank_file(filename, ".bin");
      FileCloser fc{fopen(path.c_str(), "rb")};
      if (!fc.fp) { fail(HYPRE_ERROR_GENERIC, "Could not open input file " + path); }
      uint64_t h[8];
      read_exact_n(fc.fp, h, sizeof(*h), 8, "Could not read header entries");
      if (h[0] != 1) { fail(HYPRE_ERROR_GENERIC, "Unsupported header version"); }
      if (h[7] == 1) { fail(HYPRE_ERROR_GENERIC, "row-wise ordering is not supported!"); }
      if (h[5] > h[4]) { fail(HYPRE_ERROR_GENERIC, "Invalid vector size!"); }
      if (h[6] != 1) { fail(HYPRE_ERROR_GENERIC, "multi-component vectors are not supported"); }
      if (h[3] > (uint64_t) INT32_MAX + 1 || h[2] > h[3]) { fail(HYPRE_ERROR_GENERIC, "bad partitioning"); }
      const size_t n = (size_t) h[5];
      std::vector<double> val(n);
      if (h[1] == 4) {
         std::vector<float> b(n);
         read_exact_n(fc.fp, b.data(), sizeof(*b.data()), n, "Could not read all vector coefficients");
         for (size_t k = 0; k < n; k++) { val[k] = (double) b[k]; }
      } else if (h[1] == 8) {
         read_exact_n(fc.fp, val.data(), sizeof(*val.data()), n, "Could not read all vector coefficients");
      } else {
         fail(HYPRE_ERROR_GENERIC, "Unsupported data type for vector entries");
      }
      HYPRE_IJVector vec = nullptr;
      HYPRE_IJVectorCreate(comm, (int) h[2], (int) (h[3] - 1), &vec);
      if (g_error_flag) { fail(HYPRE_ERROR_GENERIC, "IJ vector create failed"); }
      HYPRE_IJVectorSetObjectType(vec, type);
      HYPRE_IJVectorInitialize_v2(vec, HYPRE_MEMORY_HOST);
      if (n) { ijv_set(vec, (int) n, nullptr, val.data(), false); }
      HYPRE_IJVectorAssemble(vec);
      *vector = vec;
   });
}

// ================================================================== ParCSR matrix
HYPRE_Int HYPRE_ParCSRMatrixDestroy(HYPRE_ParCSRMatrix M)
{
   return guard([&] {
      if (!M) { return; }
      if (M->owns) { M->A.free_all(); }
      delete M;
   });
}

HYPRE_Int HYPRE_ParCSRMatrixGetDims(HYPRE_ParCSRMatrix M, HYPRE_BigInt *m, HYPRE_BigInt *n)
{
   return guard([&] { *m = M->A.nrows; *n = M->A.ncols; });
}

HYPRE_Int HYPRE_ParCSRMatrixGetLocalRange(HYPRE_ParCSRMatrix M, HYPRE_BigInt *rs, HYPRE_BigInt *re,
                                          HYPRE_BigInt *cs, HYPRE_BigInt *ce)
{
   return guard([&] { *rs = 0; *re = M->A.nrows - 1; *cs = 0; *ce = M->A.ncols - 1; });
}

HYPRE_Int HYPRE_ParCSRMatrixGetRow(HYPRE_ParCSRMatrix M, HYPRE_BigInt row, HYPRE_Int *size,
                                   HYPRE_BigInt **col_ind, HYPRE_Complex **values)
{
   return guard([&] {
      if (M->h_ia.empty()) { csr_to_host(M->A, M->h_ia, M->h_ja, M->h_a); }
      if (row < 0 || row >= M->A.nrows) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "row out of range"); }
      int b = M->h_ia[(size_t) row], e = M->h_ia[(size_t) row + 1];
      *size = e - b;
      if (col_ind) { *col_ind = M->h_ja.data() + b; }
      if (values) { *values = M->h_a.data() + b; }
   });
}

HYPRE_Int HYPRE_ParCSRMatrixRestoreRow(HYPRE_ParCSRMatrix M, HYPRE_BigInt row, HYPRE_Int *size,
                                       HYPRE_BigInt **col_ind, HYPRE_Complex **values)
{
   (void) M; (void) row; (void) size; (void) col_ind; (void) values;
   return g_error_flag;
}

// Size checks of src/parcsr_mv/par_csr_matvec.c:69-82 (y = A x) and
// :596-609 (y = A^T x; the reference compares x with A's columns there too):
// 11 when x's length is off, 12 when y's is, 13 when both are, returned as the
// call's value without touching the global error flag.  The reference still
// multiplies (the code is informational); here a product whose operands do not
// fit is not launched, since it would read and write past the vectors.
static int matvec_size_code(const Csr &A, int xn, int yn)
{
   const bool bx = A.ncols != xn, by = A.nrows != yn;
   return (bx && by) ? 13 : bx ? 11 : by ? 12 : 0;
}

HYPRE_Int HYPRE_ParCSRMatrixMatvec(HYPRE_Complex alpha, HYPRE_ParCSRMatrix A, HYPRE_ParVector x,
                                   HYPRE_Complex beta, HYPRE_ParVector y)
{
   int code = 0;
   const int flag = guard([&] {
      if (!A || !x || !y) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "NULL argument"); }
      code = matvec_size_code(A->A, x->n, y->n);
      if (code) { return; }
      spmv(A->A, alpha, x->d, beta, y->d, y->d, "spmv");
      y->all_zeros = false;
   });
   return code ? code : flag;
}

HYPRE_Int HYPRE_ParCSRMatrixMatvecT(HYPRE_Complex alpha, HYPRE_ParCSRMatrix A, HYPRE_ParVector x,
                                    HYPRE_Complex beta, HYPRE_ParVector y)
{
   int code = 0;
   const int flag = guard([&] {
      if (!A || !x || !y) { g_error_arg = 2; fail(HYPRE_ERROR_ARG, "NULL argument"); }
      code = matvec_size_code(A->A, x->n, y->n);
      if (x->n != A->A.nrows || y->n != A->A.ncols) {
         if (!code) { code = 13; }
         return;
      }
      // explicit transpose (a stored transpose is what the V-cycle uses)
      Csr At;
      Csr tmpP = A->A;
      transpose(tmpP, At);
      spmv(At, alpha, x->d, beta, y->d, y->d, "spmv");
      At.free_all();
      y->all_zeros = false;
   });
   return code ? code : flag;
}

// ================================================================== ParVector
HYPRE_Int HYPRE_ParVectorCreate(MPI_Comm comm, HYPRE_BigInt global_size, HYPRE_BigInt *partitioning,
                                HYPRE_ParVector *vector)
{
   return guard([&] {
      (void) comm; (void) partitioning;
      if (!vector) { g_error_arg = 4; fail(HYPRE_ERROR_ARG, "NULL vector pointer"); }
      ensure_init();
      auto *v = new hypre_ParVector_struct();
      v->n = global_size;
      v->d = nullptr;
      *vector = v;
   });
}

HYPRE_Int HYPRE_ParVectorInitialize(HYPRE_ParVector v)
{
   return guard([&] {
      if (!v->d) {
         v->d = dalloc_n<double>((size_t) std::max(v->n, 1));
         v->owns = true;
      }
      fill(v->n, 0.0, v->d);
      v->all_zeros = true;
   });
}

HYPRE_Int HYPRE_ParVectorDestroy(HYPRE_ParVector v)
{
   return guard([&] { vec_destroy(v); });
}

HYPRE_Int HYPRE_ParVectorCopy(HYPRE_ParVector x, HYPRE_ParVector y)
{
   return guard([&] {
      copy(x->n, x->d, y->d);
      y->all_zeros = x->all_zeros;
   });
}

HYPRE_Int HYPRE_ParVectorSetConstantValues(HYPRE_ParVector v, HYPRE_Complex value)
{
   return guard([&] {
      fill(v->n, value, v->d);
      v->all_zeros = false;   // only SetZeros raises the flag (par_vector.c:428)
   });
}

__global__ void k_two_u_minus_one(int n, double *x)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      x[i] = 2.0 * x[i] - 1.0;
   }
}

// src/parcsr_mv/par_vector.c hypre_ParVectorSetRandomValues (rank 0: seed * 1)
// + src/seq_mv/vector.c hypre_SeqVectorSetRandomValues: 2*hypre_Rand()-1
HYPRE_Int HYPRE_ParVectorSetRandomValues(HYPRE_ParVector v, HYPRE_Int seed)
{
   return guard([&] {
      pm_uniform(v->n, seed, v->d);
      int nb = std::max(1, std::min((v->n + 255) / 256, 8192));
      k_two_u_minus_one<<<nb, 256, 0, stream()>>>(v->n, v->d);
      HX_LAUNCH_CHECK();
      v->all_zeros = false;
   });
}

HYPRE_Int HYPRE_MI355X_ParVectorSetRandomPM(HYPRE_ParVector v, HYPRE_Int seed)
{
   return HYPRE_ParVectorSetRandomValues(v, seed);
}

HYPRE_Int HYPRE_ParVectorScale(HYPRE_Complex value, HYPRE_ParVector x)
{
   return guard([&] { scale(x->n, value, x->d); });
}

HYPRE_Int HYPRE_ParVectorAxpy(HYPRE_Complex alpha, HYPRE_ParVector x, HYPRE_ParVector y)
{
   return guard([&] {
      axpy(x->n, alpha, x->d, y->d);
      y->all_zeros = false;
   });
}

HYPRE_Int HYPRE_ParVectorInnerProd(HYPRE_ParVector x, HYPRE_ParVector y, HYPRE_Real *prod)
{
   return guard([&] { *prod = dot(x->n, x->d, y->d); });
}

HYPRE_Int HYPRE_ParVectorGetValues(HYPRE_ParVector v, HYPRE_Int num_values, HYPRE_BigInt *indices,
                                   HYPRE_Complex *values)
{
   return guard([&] {
      std::vector<double> h((size_t) v->n);
      d2h(h.data(), v->d, sizeof(double) * (size_t) v->n);
      for (int k = 0; k < num_values; k++) {
         int i = indices ? indices[k] : k;
         values[k] = (i >= 0 && i < v->n) ? h[(size_t) i] : 0.0;
      }
   });
}

// ================================================================== device access
HYPRE_Int HYPRE_MI355X_ParCSRMatrixGetDevice(HYPRE_ParCSRMatrix A, HYPRE_Int *n, HYPRE_Int *nnz,
                                             HYPRE_Int **ia_dev, HYPRE_Int **ja_dev, HYPRE_Real **a_dev)
{
   return guard([&] {
      if (n) { *n = A->A.nrows; }
      if (nnz) { *nnz = (int) A->A.nnz; }
      if (ia_dev) { *ia_dev = A->A.ia; }
      if (ja_dev) { *ja_dev = A->A.ja; }
      if (a_dev) { *a_dev = A->A.a; }
   });
}

HYPRE_Int HYPRE_MI355X_ParCSRMatrixCopyToHost(HYPRE_ParCSRMatrix A, HYPRE_Int *ia, HYPRE_Int *ja,
                                              HYPRE_Real *a)
{
   return guard([&] {
      d2h(ia, A->A.ia, sizeof(int) * ((size_t) A->A.nrows + 1));
      if (ja) { d2h(ja, A->A.ja, sizeof(int) * (size_t) A->A.nnz); }
      if (a) { d2h(a, A->A.a, sizeof(double) * (size_t) A->A.nnz); }
   });
}

HYPRE_Int HYPRE_MI355X_ParVectorGetDevice(HYPRE_ParVector v, HYPRE_Int *n, HYPRE_Real **d)
{
   return guard([&] {
      if (n) { *n = v->n; }
      if (d) { *d = v->d; }
   });
}

HYPRE_Int HYPRE_MI355X_ParVectorCopyFromHost(HYPRE_ParVector v, const HYPRE_Real *x)
{
   return guard([&] {
      h2d(v->d, x, sizeof(double) * (size_t) v->n);
      bool z = true;
      for (int i = 0; i < v->n; i++) { if (x[i] != 0.0) { z = false; break; } }
      v->all_zeros = z;
   });
}

HYPRE_Int HYPRE_MI355X_ParVectorCopyToHost(HYPRE_ParVector v, HYPRE_Real *x)
{
   return guard([&] { d2h(x, v->d, sizeof(double) * (size_t) v->n); });
}

HYPRE_ParCSRMatrix HYPRE_MI355X_ParCSRMatrixFromHostCSR(HYPRE_Int n, HYPRE_Int ncols, const HYPRE_Int *ia,
                                                        const HYPRE_Int *ja, const HYPRE_Real *a)
{
   hypre_ParCSRMatrix_struct *M = nullptr;
   guard([&] {
      ensure_init();
      M = par_from_host(n, ncols, ia, ja, a);
   });
   return M;
}

} // extern "C"

// ================================================================== generators
// 7-point (GenerateLaplacian, src/parcsr_ls/par_laplace.c, row layout
// par_laplace.c:158-250: diag, -z, -y, -x, +x, +y, +z) and 27-point
// (GenerateLaplacian27pt, src/parcsr_ls/par_laplace_27pt.c: diag, then the
// neighbours by ascending column) stencils, built on the device.
__device__ __forceinline__ void stencil_coords(long long row, int nx, int ny, int &ix, int &iy, int &iz)
{
   ix = (int) (row % nx);
   long long t = row / nx;
   iy = (int) (t % ny);
   iz = (int) (t / ny);
}

__global__ void k_gen_count(int nx, int ny, int nz, int pts, int *__restrict__ cnt)
{
   long long n = (long long) nx * ny * nz;
   for (long long r = (long long) blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (long long) gridDim.x * blockDim.x) {
      int ix, iy, iz;
      stencil_coords(r, nx, ny, ix, iy, iz);
      int c = 1;
      if (pts == 7) {
         c += (iz > 0) + (iy > 0) + (ix > 0) + (ix + 1 < nx) + (iy + 1 < ny) + (iz + 1 < nz);
      } else {
         int cx = 1 + (ix > 0) + (ix + 1 < nx);
         int cy = 1 + (iy > 0) + (iy + 1 < ny);
         int cz = 1 + (iz > 0) + (iz + 1 < nz);
         c = cx * cy * cz;
      }
      cnt[r] = c;
   }
}

// values: 7pt: v[0] diag, v[1] x, v[2] y, v[3] z; 27pt: v[0] diag, v[1] off
// kappa (optional): node coefficient for the variable-coefficient problems;
// then off-diagonals are -harmonic means and the diagonal their sum plus the
// node's own kappa for every missing (Dirichlet) neighbour.
__global__ void k_gen_fill(int nx, int ny, int nz, int pts, double v0, double v1, double v2, double v3,
                           const double *__restrict__ kappa, const int *__restrict__ ia,
                           int *__restrict__ ja, double *__restrict__ a)
{
   long long n = (long long) nx * ny * nz;
   const long long nxy = (long long) nx * ny;
   for (long long r = (long long) blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (long long) gridDim.x * blockDim.x) {
      int ix, iy, iz;
      stencil_coords(r, nx, ny, ix, iy, iz);
      int o = ia[r];
      const int d = o++;
      ja[d] = (int) r;
      double kp = kappa ? kappa[r] : 0.0;
      double diag = 0.0;
      for (int dz = -1; dz <= 1; dz++) {
         for (int dy = -1; dy <= 1; dy++) {
            for (int dx = -1; dx <= 1; dx++) {
               if (!dx && !dy && !dz) { continue; }
               const int man = abs(dx) + abs(dy) + abs(dz);
               if (pts == 7 && man != 1) { continue; }
               const int jx = ix + dx, jy = iy + dy, jz = iz + dz;
               const bool inside = jx >= 0 && jx < nx && jy >= 0 && jy < ny && jz >= 0 && jz < nz;
               if (!inside) {
                  if (kappa) { diag += kp; }
                  continue;
               }
               const long long col = r + dz * nxy + (long long) dy * nx + dx;
               double val;
               if (kappa) {
                  double kq = kappa[col];
                  double k = 2.0 * kp * kq / (kp + kq);
                  diag += k;
                  val = -k;
               } else if (pts == 7) {
                  val = dx ? v1 : (dy ? v2 : v3);
               } else {
                  val = v1;
               }
               ja[o] = (int) col;
               a[o] = val;
               o++;
            }
         }
      }
      a[d] = kappa ? diag : v0;
   }
}

static hypre_ParCSRMatrix_struct *gen_stencil(int nx, int ny, int nz, int pts, double v0, double v1,
                                              double v2, double v3, const double *kappa)
{
   ensure_init();
   long long n = (long long) nx * ny * nz;
   if (n > 0x7fffffffLL) { fail(HYPRE_ERROR_ARG, "grid too large for 32-bit indices"); }
   auto *M = new hypre_ParCSRMatrix_struct();
   Csr &A = M->A;
   A.nrows = A.ncols = (int) n;
   A.ia = dalloc_n<int>((size_t) n + 1);
   int nb = (int) std::max<long long>(1, std::min<long long>((n + 255) / 256, 16384));
   k_gen_count<<<nb, 256, 0, stream()>>>(nx, ny, nz, pts, A.ia);
   HX_LAUNCH_CHECK();
   A.nnz = exclusive_scan(A.ia, A.ia, (int) n);
   A.ja = dalloc_n<int>((size_t) A.nnz);
   A.a = dalloc_n<double>((size_t) A.nnz);
   k_gen_fill<<<nb, 256, 0, stream()>>>(nx, ny, nz, pts, v0, v1, v2, v3, kappa, A.ia, A.ja, A.a);
   HX_LAUNCH_CHECK();
   sync();
   return M;
}

__global__ void k_node_kappa(int N, int cell, int nc, const double *__restrict__ u, double lo, double hi,
                             double *__restrict__ kappa)
{
   long long n = (long long) N * N * N;
   for (long long r = (long long) blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (long long) gridDim.x * blockDim.x) {
      int ix, iy, iz;
      stencil_coords(r, N, N, ix, iy, iz);
      long long c = ((long long) (iz / cell) * nc + iy / cell) * nc + ix / cell;
      kappa[r] = pow(10.0, lo + (hi - lo) * u[c]);
   }
}

extern "C" {

HYPRE_ParCSRMatrix GenerateLaplacian(MPI_Comm comm, HYPRE_BigInt nx, HYPRE_BigInt ny, HYPRE_BigInt nz,
                                     HYPRE_Int P, HYPRE_Int Q, HYPRE_Int R, HYPRE_Int p, HYPRE_Int q,
                                     HYPRE_Int r, HYPRE_Real *value)
{
   hypre_ParCSRMatrix_struct *M = nullptr;
   guard([&] {
      (void) comm;
      if (P != 1 || Q != 1 || R != 1 || p || q || r) {
         fail(HYPRE_ERROR_ARG, "single-rank build: processor grid must be 1 x 1 x 1");
      }
      M = gen_stencil(nx, ny, nz, 7, value[0], value[1], value[2], value[3], nullptr);
   });
   return M;
}

HYPRE_ParCSRMatrix GenerateLaplacian27pt(MPI_Comm comm, HYPRE_BigInt nx, HYPRE_BigInt ny, HYPRE_BigInt nz,
                                         HYPRE_Int P, HYPRE_Int Q, HYPRE_Int R, HYPRE_Int p, HYPRE_Int q,
                                         HYPRE_Int r, HYPRE_Real *value)
{
   hypre_ParCSRMatrix_struct *M = nullptr;
   guard([&] {
      (void) comm;
      if (P != 1 || Q != 1 || R != 1 || p || q || r) {
         fail(HYPRE_ERROR_ARG, "single-rank build: processor grid must be 1 x 1 x 1");
      }
      M = gen_stencil(nx, ny, nz, 27, value[0], value[1], 0.0, 0.0, nullptr);
   });
   return M;
}

HYPRE_ParCSRMatrix HYPRE_MI355X_GenerateVarDiffusion(HYPRE_Int N, HYPRE_Int stencil, HYPRE_Int cell,
                                                     HYPRE_Real lo, HYPRE_Real hi, HYPRE_Int seed)
{
   hypre_ParCSRMatrix_struct *M = nullptr;
   guard([&] {
      ensure_init();
      if (stencil != 7 && stencil != 27) { fail(HYPRE_ERROR_ARG, "stencil must be 7 or 27"); }
      int nc = (N + cell - 1) / cell;
      long long ncell = (long long) nc * nc * nc;
      double *u = dalloc_n<double>((size_t) ncell);
      pm_uniform((int) ncell, seed, u);
      long long n = (long long) N * N * N;
      double *kappa = dalloc_n<double>((size_t) n);
      int nb = (int) std::max<long long>(1, std::min<long long>((n + 255) / 256, 16384));
      k_node_kappa<<<nb, 256, 0, stream()>>>(N, cell, nc, u, lo, hi, kappa);
      HX_LAUNCH_CHECK();
      M = gen_stencil(N, N, N, stencil, 0.0, 0.0, 0.0, 0.0, kappa);
      dfree(u);
      dfree(kappa);
   });
   return M;
}

HYPRE_ParCSRMatrix HYPRE_MI355X_GenerateAniso7(HYPRE_Int N, HYPRE_Real eps, HYPRE_Real theta)
{
   hypre_ParCSRMatrix_struct *M = nullptr;
   guard([&] {
      double c = std::cos(theta), s = std::sin(theta);
      double kxx = c * c + eps * s * s, kyy = s * s + eps * c * c, kzz = 1.0;
      double v0 = 0.0;
      if (N > 1) { v0 += 2.0 * kxx; v0 += 2.0 * kyy; v0 += 2.0 * kzz; }
      M = gen_stencil(N, N, N, 7, v0, -kxx, -kyy, -kzz, nullptr);
   });
   return M;
}

// ================================================================== single kernels
HYPRE_Int HYPRE_MI355X_CreateS(HYPRE_Int n, const HYPRE_Int *ia, const HYPRE_Int *ja, const HYPRE_Real *a,
                               HYPRE_Real theta, HYPRE_Real max_row_sum, HYPRE_Int *S_ia, HYPRE_Int *S_ja)
{
   return guard([&] {
      ensure_init();
      Csr A, S;
      csr_from_host(A, n, n, ia, ja, a);
      create_s(A, theta, max_row_sum, S);
      d2h(S_ia, S.ia, sizeof(int) * ((size_t) n + 1));
      if (S.nnz) { d2h(S_ja, S.ja, sizeof(int) * (size_t) S.nnz); }
      A.free_all();
      S.free_all();
   });
}

HYPRE_Int HYPRE_MI355X_PMIS(HYPRE_Int n, const HYPRE_Int *S_ia, const HYPRE_Int *S_ja,
                            const HYPRE_Real *measure, HYPRE_Int *cf, HYPRE_Int *rounds)
{
   return guard([&] {
      ensure_init();
      Csr S;
      csr_from_host(S, n, n, S_ia, S_ja, nullptr);
      double *m = nullptr;
      if (measure) {
         m = dalloc_n<double>((size_t) n);
         h2d(m, measure, sizeof(double) * (size_t) n);
      }
      int *dcf = dalloc_n<int>((size_t) n);
      int r = pmis(S, m, dcf, nullptr);
      d2h(cf, dcf, sizeof(int) * (size_t) n);
      if (rounds) { *rounds = r; }
      dfree(m);
      dfree(dcf);
      S.free_all();
   });
}

HYPRE_Int HYPRE_MI355X_HMIS(HYPRE_Int n, const HYPRE_Int *S_ia, const HYPRE_Int *S_ja,
                            const HYPRE_Real *measure, HYPRE_Int *cf, HYPRE_Int *rounds)
{
   return guard([&] {
      ensure_init();
      Csr S;
      csr_from_host(S, n, n, S_ia, S_ja, nullptr);
      double *m = nullptr;
      if (measure) {
         m = dalloc_n<double>((size_t) n);
         h2d(m, measure, sizeof(double) * (size_t) n);
      }
      int *dcf = dalloc_n<int>((size_t) n);
      int r = hmis(S, m, dcf, nullptr);
      d2h(cf, dcf, sizeof(int) * (size_t) n);
      if (rounds) { *rounds = r; }
      dfree(m);
      dfree(dcf);
      S.free_all();
   });
}

HYPRE_Int HYPRE_MI355X_RSFirstPass(HYPRE_Int n, const HYPRE_Int *S_ia, const HYPRE_Int *S_ja,
                                   HYPRE_Int *cf)
{
   return guard([&] {
      if (n < 0 || !S_ia || !cf || (n > 0 && S_ia[n] > 0 && !S_ja)) {
         g_error_arg = 1;
         fail(HYPRE_ERROR_ARG, "bad strength pattern");
      }
      rs_first_pass_host(n, S_ia, S_ja, cf);
   });
}

HYPRE_Int HYPRE_MI355X_ExtPIInterp(HYPRE_Int n, const HYPRE_Int *ia, const HYPRE_Int *ja, const HYPRE_Real *a,
                                   const HYPRE_Int *S_ia, const HYPRE_Int *S_ja, HYPRE_Int *cf,
                                   HYPRE_Real trunc_factor, HYPRE_Int max_elmts, HYPRE_Int *P_ncols,
                                   HYPRE_Int *P_nnz, HYPRE_Int *P_ia, HYPRE_Int *P_ja, HYPRE_Real *P_a)
{
   return guard([&] {
      ensure_init();
      Csr A, S, P;
      csr_from_host(A, n, n, ia, ja, a);
      csr_from_host(S, n, n, S_ia, S_ja, nullptr);
      int *dcf = dalloc_n<int>((size_t) n);
      h2d(dcf, cf, sizeof(int) * (size_t) n);
      int *f2c = dalloc_n<int>((size_t) n + 1);
      int nc = coarse_map(n, dcf, f2c);
      extpi_interp(A, S, dcf, f2c, nc, trunc_factor, max_elmts, P);
      if (P_ncols) { *P_ncols = nc; }
      if (P_nnz) { *P_nnz = (int) P.nnz; }
      if (P_ia) {
         d2h(P_ia, P.ia, sizeof(int) * ((size_t) n + 1));
         if (P_ja) { d2h(P_ja, P.ja, sizeof(int) * (size_t) P.nnz); }
         if (P_a) { d2h(P_a, P.a, sizeof(double) * (size_t) P.nnz); }
         d2h(cf, dcf, sizeof(int) * (size_t) n);
      }
      dfree(dcf);
      dfree(f2c);
      A.free_all(); S.free_all(); P.free_all();
   });
}

HYPRE_Int HYPRE_MI355X_RAP(HYPRE_Int n, HYPRE_Int nc, const HYPRE_Int *ia, const HYPRE_Int *ja, const HYPRE_Real *a,
                           const HYPRE_Int *P_ia, const HYPRE_Int *P_ja, const HYPRE_Real *P_a,
                           HYPRE_Int *Ac_nnz, HYPRE_Int *Ac_ia, HYPRE_Int *Ac_ja, HYPRE_Real *Ac_a)
{
   return guard([&] {
      ensure_init();
      Csr A, P, R, Ac;
      csr_from_host(A, n, n, ia, ja, a);
      csr_from_host(P, n, nc, P_ia, P_ja, P_a);
      transpose(P, R);
      rap(R, A, P, Ac);
      if (Ac_nnz) { *Ac_nnz = (int) Ac.nnz; }
      if (Ac_ia) {
         d2h(Ac_ia, Ac.ia, sizeof(int) * ((size_t) nc + 1));
         if (Ac_ja) { d2h(Ac_ja, Ac.ja, sizeof(int) * (size_t) Ac.nnz); }
         if (Ac_a) { d2h(Ac_a, Ac.a, sizeof(double) * (size_t) Ac.nnz); }
      }
      A.free_all(); P.free_all(); R.free_all(); Ac.free_all();
   });
}

HYPRE_Int HYPRE_MI355X_SpMV(HYPRE_Int n, HYPRE_Int ncols, const HYPRE_Int *ia, const HYPRE_Int *ja,
                            const HYPRE_Real *a, HYPRE_Real alpha, const HYPRE_Real *x, HYPRE_Real beta,
                            HYPRE_Real *y)
{
   return guard([&] {
      ensure_init();
      Csr A;
      csr_from_host(A, n, ncols, ia, ja, a);
      double *dx = dalloc_n<double>((size_t) ncols), *dy = dalloc_n<double>((size_t) n);
      h2d(dx, x, sizeof(double) * (size_t) ncols);
      h2d(dy, y, sizeof(double) * (size_t) n);
      spmv(A, alpha, dx, beta, dy, dy, "spmv");
      d2h(y, dy, sizeof(double) * (size_t) n);
      dfree(dx);
      dfree(dy);
      A.free_all();
   });
}

HYPRE_Int HYPRE_MI355X_HybridGS(HYPRE_Int n, const HYPRE_Int *ia, const HYPRE_Int *ja,
                                const HYPRE_Real *a, const HYPRE_Real *f, const HYPRE_Real *l1,
                                HYPRE_Real *u, HYPRE_Int relax_type, HYPRE_Real relax_weight,
                                HYPRE_Real omega, HYPRE_Int parts, HYPRE_Int zero_guess)
{
   return guard([&] {
      ensure_init();
      if (!is_hybrid_gs(relax_type)) {
         g_error_arg = 8;
         fail(HYPRE_ERROR_ARG, "not a hybrid Gauss-Seidel relaxation type");
      }
      const bool l1t = relax_type == 8 || relax_type == 13 || relax_type == 14;
      if (l1t && !l1) { g_error_arg = 6; fail(HYPRE_ERROR_ARG, "l1 norms required"); }
      if (parts < 1) { g_error_arg = 11; fail(HYPRE_ERROR_ARG, "parts < 1"); }
      Csr A;
      csr_from_host(A, n, n, ia, ja, a);
      const size_t nn = (size_t) std::max(n, 1);
      double *df = dalloc_n<double>(nn), *du = dalloc_n<double>(nn), *dout = dalloc_n<double>(nn);
      double *dl1 = nullptr;
      h2d(df, f, sizeof(double) * (size_t) n);
      if (!zero_guess) { h2d(du, u, sizeof(double) * (size_t) n); }
      if (l1t) {
         dl1 = dalloc_n<double>(nn);
         h2d(dl1, l1, sizeof(double) * (size_t) n);
      }
      Csr Arev;
      if (hgs_needs_reverse(relax_type)) { csr_reverse_rows(A, Arev); }
      relax_hgs(relax_type, A, &Arev, -1, df, du, dl1, relax_weight, omega, parts, dout,
                zero_guess != 0, "relax");
      Arev.free_all();
      d2h(u, dout, sizeof(double) * (size_t) n);
      dfree(df);
      dfree(du);
      dfree(dout);
      if (dl1) { dfree(dl1); }
      A.free_all();
   });
}

} // extern "C"
