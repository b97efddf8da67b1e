// Learned per-level AMG parameters (no counterpart in hypre).
//
// A 3-layer perceptron maps 8 features of a matrix to per-level strong
// thresholds, relaxation weights and sweep counts:
//   Z1 = X W1 + b1, H1 = relu(Z1); Z2 = H1 W2 + b2, H2 = relu(Z2);
//   Z3 = H2 W3 + b3, Y = sigmoid(Z3)        (row-major, one sample per row)
// trained with a mean-squared error against targets in [0, 1] and Adam.  All
// products run on the matrix cores (v_mfma_f32_16x16x4f32, one wave per
// 16x16 output tile); the bias/column sums and the loss are reduced in a
// fixed order, so a training run is bit-reproducible.
//
// Two readings of the output, chosen by the caller:
// * per-level settings (nout = 3 L): theta = 0.1 + 0.8 y[3l],
//   omega = 0.6 + 0.8 y[3l+1], sweeps = 1 + floor(2 y[3l+2] + 0.5);
// * a cost model over a grid of G settings (nout = 2 G, grid set with
//   HYPRE_MI355X_ParamModelSetGrid): y[g] predicts the time of setting g on
//   the fine level relative to the defaults, y[G + g] the same for the coarse
//   levels; the predicted fastest setting of each group is applied.
#include "hx_internal.h"
#include "hx_objects.h"

#include <algorithm>
#include <cmath>
#include <vector>

struct hx_param_model_struct
{
   int    nin = 0, nh1 = 0, nh2 = 0, nout = 0;
   int    np = 0;
   int    t = 0;                 // Adam step counter
   float *p = nullptr;           // parameters, device
   float *g = nullptr;           // gradient
   float *m = nullptr, *v = nullptr;
   // activations for the last batch
   int    cap = 0;
   float *X = nullptr, *Z1 = nullptr, *H1 = nullptr, *Z2 = nullptr, *H2 = nullptr, *Z3 = nullptr;
   float *Y = nullptr, *T = nullptr, *D3 = nullptr, *D2 = nullptr, *D1 = nullptr, *DH = nullptr;
   float *loss = nullptr;        // device scalar
   int    oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, oW3 = 0, ob3 = 0;
   // prior of the cost-model reading: grid settings kept on the fine and
   // coarse levels unless another is predicted faster by more than the margin
   int    prior_f = -1, prior_c = -1;
   double prior_margin = 0.0;
   // settings of the cost-model reading (empty: per-level reading)
   std::vector<double> g_theta, g_omega;
   std::vector<int>    g_sweeps;
};

namespace hx {

constexpr int kModelLevels = 25;   // levels a cost model sets (hypre's max_levels default)

typedef float floatx4 __attribute__((ext_vector_type(4)));

// C[M x N] = op(A) op(B) (+ bias[col]); optional H = relu(C).  op(A)(i, k) is
// A[i*lda + k] or, transposed, A[k*lda + i]; likewise for B.
// Lane l of the wave feeds A(i0 + l%16, k + l/16) and B(k + l/16, j0 + l%16)
// and owns C(i0 + 4*(l/16) + r, j0 + l%16), r = 0..3.
__global__ __launch_bounds__(64) void k_gemm16(int M, int N, int K, const float *__restrict__ A,
                                               int lda, int ta, const float *__restrict__ B, int ldb,
                                               int tb, const float *__restrict__ bias,
                                               float *__restrict__ C, int ldc,
                                               float *__restrict__ H)
{
   const int lane = threadIdx.x;
   const int tiles_n = (N + 15) / 16;
   const int i0 = (blockIdx.x / tiles_n) * 16, j0 = (blockIdx.x % tiles_n) * 16;
   const int ia = i0 + (lane & 15), jb = j0 + (lane & 15), kl = lane >> 4;
   floatx4 acc = {0.f, 0.f, 0.f, 0.f};
   for (int k0 = 0; k0 < K; k0 += 4) {
      const int k = k0 + kl;
      float a = 0.f, b = 0.f;
      if (ia < M && k < K) { a = ta ? A[(size_t) k * lda + ia] : A[(size_t) ia * lda + k]; }
      if (jb < N && k < K) { b = tb ? B[(size_t) jb * ldb + k] : B[(size_t) k * ldb + jb]; }
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
   }
   for (int r = 0; r < 4; r++) {
      const int i = i0 + 4 * kl + r, j = jb;
      if (i < M && j < N) {
         float c = acc[r];
         if (bias) { c += bias[j]; }
         C[(size_t) i * ldc + j] = c;
         if (H) { H[(size_t) i * ldc + j] = c > 0.f ? c : 0.f; }
      }
   }
}

static void gemm(int M, int N, int K, const float *A, int lda, bool ta, const float *B, int ldb,
                 bool tb, const float *bias, float *C, int ldc, float *H)
{
   if (M <= 0 || N <= 0) { return; }
   const int tiles = ((M + 15) / 16) * ((N + 15) / 16);
   k_gemm16<<<tiles, 64, 0, stream()>>>(M, N, K, A, lda, ta ? 1 : 0, B, ldb, tb ? 1 : 0, bias, C,
                                        ldc, H);
   HX_LAUNCH_CHECK();
}

__device__ __forceinline__ float sigmoidf(float z) { return 1.f / (1.f + expf(-z)); }

__global__ void k_sigmoid(int n, const float *__restrict__ z, float *__restrict__ y)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      y[i] = sigmoidf(z[i]);
   }
}

// Y = sigmoid(Z3); D3 = dL/dZ3 for L = mean((Y - T)^2); loss in one block,
// partial sums in thread order then a fixed tree.
__global__ __launch_bounds__(256) void k_mse_head(int n, const float *__restrict__ z,
                                                  const float *__restrict__ t,
                                                  float *__restrict__ y, float *__restrict__ d,
                                                  float *__restrict__ loss)
{
   __shared__ float sh[256];
   const float inv = 1.f / (float) n;
   float s = 0.f;
   for (int i = threadIdx.x; i < n; i += 256) {
      const float yi = sigmoidf(z[i]);
      const float e = yi - t[i];
      y[i] = yi;
      d[i] = 2.f * e * inv * yi * (1.f - yi);
      s += e * e;
   }
   sh[threadIdx.x] = s;
   __syncthreads();
   for (int off = 128; off > 0; off >>= 1) {
      if (threadIdx.x < off) { sh[threadIdx.x] += sh[threadIdx.x + off]; }
      __syncthreads();
   }
   if (threadIdx.x == 0) { *loss = sh[0] * inv; }
}

// dZ = dH * (Z > 0)
__global__ void k_relu_back(int n, const float *__restrict__ dh, const float *__restrict__ z,
                            float *__restrict__ dz)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      dz[i] = z[i] > 0.f ? dh[i] : 0.f;
   }
}

// db[j] = sum_i D[i][j], rows in order
__global__ void k_colsum(int rows, int cols, const float *__restrict__ D, float *__restrict__ db)
{
   for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < cols; j += gridDim.x * blockDim.x) {
      float s = 0.f;
      for (int i = 0; i < rows; i++) { s += D[(size_t) i * cols + j]; }
      db[j] = s;
   }
}

__global__ void k_adam(int n, float *__restrict__ p, const float *__restrict__ g,
                       float *__restrict__ m, float *__restrict__ v, float lr, float b1, float b2,
                       float eps, float c1, float c2)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      const float gi = g[i];
      const float mi = b1 * m[i] + (1.f - b1) * gi;
      const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
      m[i] = mi;
      v[i] = vi;
      p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
   }
}

static int grid1(int n) { return std::max(1, std::min((n + 255) / 256, 1024)); }

static void ensure_batch(hx_param_model_struct *m, int B)
{
   if (B <= m->cap) { return; }
   float *old[] = {m->X, m->Z1, m->H1, m->Z2, m->H2, m->Z3, m->Y, m->T, m->D3, m->D2, m->D1, m->DH};
   for (float *q : old) { dfree(q); }
   const int wh = std::max(m->nh1, m->nh2);
   m->X = dalloc_n<float>((size_t) B * m->nin);
   m->Z1 = dalloc_n<float>((size_t) B * m->nh1);
   m->H1 = dalloc_n<float>((size_t) B * m->nh1);
   m->Z2 = dalloc_n<float>((size_t) B * m->nh2);
   m->H2 = dalloc_n<float>((size_t) B * m->nh2);
   m->Z3 = dalloc_n<float>((size_t) B * m->nout);
   m->Y = dalloc_n<float>((size_t) B * m->nout);
   m->T = dalloc_n<float>((size_t) B * m->nout);
   m->D3 = dalloc_n<float>((size_t) B * m->nout);
   m->D2 = dalloc_n<float>((size_t) B * wh);
   m->D1 = dalloc_n<float>((size_t) B * wh);
   m->DH = dalloc_n<float>((size_t) B * wh);
   m->cap = B;
}

static void forward(hx_param_model_struct *m, int B)
{
   const float *p = m->p;
   gemm(B, m->nh1, m->nin, m->X, m->nin, false, p + m->oW1, m->nh1, false, p + m->ob1, m->Z1, m->nh1, m->H1);
   gemm(B, m->nh2, m->nh1, m->H1, m->nh1, false, p + m->oW2, m->nh2, false, p + m->ob2, m->Z2, m->nh2, m->H2);
   gemm(B, m->nout, m->nh2, m->H2, m->nh2, false, p + m->oW3, m->nout, false, p + m->ob3, m->Z3, m->nout, nullptr);
}

// gradient of the mean-squared error into m->g, loss into m->loss
static void backward(hx_param_model_struct *m, int B)
{
   float *g = m->g;
   const float *p = m->p;
   const int n3 = B * m->nout;
   k_mse_head<<<1, 256, 0, stream()>>>(n3, m->Z3, m->T, m->Y, m->D3, m->loss);
   HX_LAUNCH_CHECK();
   // layer 3
   gemm(m->nh2, m->nout, B, m->H2, m->nh2, true, m->D3, m->nout, false, nullptr, g + m->oW3, m->nout, nullptr);
   k_colsum<<<grid1(m->nout), 256, 0, stream()>>>(B, m->nout, m->D3, g + m->ob3);
   gemm(B, m->nh2, m->nout, m->D3, m->nout, false, p + m->oW3, m->nout, true, nullptr, m->DH, m->nh2, nullptr);
   k_relu_back<<<grid1(B * m->nh2), 256, 0, stream()>>>(B * m->nh2, m->DH, m->Z2, m->D2);
   // layer 2
   gemm(m->nh1, m->nh2, B, m->H1, m->nh1, true, m->D2, m->nh2, false, nullptr, g + m->oW2, m->nh2, nullptr);
   k_colsum<<<grid1(m->nh2), 256, 0, stream()>>>(B, m->nh2, m->D2, g + m->ob2);
   gemm(B, m->nh1, m->nh2, m->D2, m->nh2, false, p + m->oW2, m->nh2, true, nullptr, m->DH, m->nh1, nullptr);
   k_relu_back<<<grid1(B * m->nh1), 256, 0, stream()>>>(B * m->nh1, m->DH, m->Z1, m->D1);
   // layer 1
   gemm(m->nin, m->nh1, B, m->X, m->nin, true, m->D1, m->nh1, false, nullptr, g + m->oW1, m->nh1, nullptr);
   k_colsum<<<grid1(m->nh1), 256, 0, stream()>>>(B, m->nh1, m->D1, g + m->ob1);
   HX_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ matrix features
// Per row i (off-diagonal entries j):
//   q0 = max|a_ij|/|a_ii|, q1 = min|a_ij|/max|a_ij|, q2 = 2nd largest/largest,
//   q3 = fraction of strong entries at theta 0.25 (par_strength.c rule),
//   q4 = |sum_j a_ij| / |a_ii|, q5 = log10|a_ii|, q6 = (log10|a_ii|)^2, q7 = row length.
// Block partial sums (rows in order per thread, fixed tree), summed on the host.
constexpr int NQ = 8;

// Per-row sums of the features (one thread per row, block partials).  The
// coefficient jump of row i is tested against its neighbours' diagonals by
// products (|a_jj| >= 10 |a_ii| or |a_ii| >= 10 |a_jj|), which every IEEE
// implementation rounds alike, instead of comparing logarithms.
__global__ __launch_bounds__(256) void k_row_features(int n, const int *__restrict__ ia,
                                                      const int *__restrict__ ja,
                                                      const double *__restrict__ a,
                                                      double *__restrict__ part)
{
   __shared__ double sh[NQ][256];
   double q[NQ] = {0, 0, 0, 0, 0, 0, 0, 0};
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      const int s = ia[i], e = ia[i + 1];
      const double d = a[s];
      const double ad = fabs(d) > 0 ? fabs(d) : 1.0;
      double mx = 0.0, mx2 = 0.0, mn = 1e300, scale = 0.0;
      bool j1 = false, j3 = false;
      for (int k = s + 1; k < e; k++) {
         const double v = a[k], av = fabs(v);
         if (av > mx) { mx2 = mx; mx = av; } else if (av > mx2) { mx2 = av; }
         mn = av < mn ? av : mn;
         scale = (d < 0) ? ((scale < v) ? v : scale) : ((scale < v) ? scale : v);
         const double dj = fabs(a[ia[ja[k]]]);
         const double aj = dj > 0 ? dj : 1.0;
         j1 = j1 || aj >= 10.0 * ad || ad >= 10.0 * aj;
         j3 = j3 || aj >= 1000.0 * ad || ad >= 1000.0 * aj;
      }
      int strong = 0;
      const double thr = 0.25 * scale;
      for (int k = s + 1; k < e; k++) {
         const double v = a[k];
         strong += (d < 0) ? !(v <= thr) : !(v >= thr);
      }
      const int no = e - s - 1;
      const double len = (double) (e - s);
      q[0] += len;
      q[1] += len * len;
      q[2] += (no > 0 && mx > 0) ? mn / mx : 0.0;
      q[3] += (no > 1 && mx > 0) ? mx2 / mx : 0.0;
      q[4] += no > 0 ? (double) strong / no : 0.0;
      q[5] += j1 ? 1.0 : 0.0;
      q[6] += j3 ? 1.0 : 0.0;
   }
   for (int c = 0; c < NQ; c++) { sh[c][threadIdx.x] = q[c]; }
   __syncthreads();
   for (int off = 128; off > 0; off >>= 1) {
      if (threadIdx.x < off) {
         for (int c = 0; c < NQ; c++) { sh[c][threadIdx.x] += sh[c][threadIdx.x + off]; }
      }
      __syncthreads();
   }
   if (threadIdx.x < NQ) { part[(size_t) blockIdx.x * NQ + threadIdx.x] = sh[threadIdx.x][0]; }
}

// f = [log10(n)/8, mean row length/27, std(row length)/27, mean min/max
// |a_ij|, mean 2nd/1st largest |a_ij| (off-diagonal anisotropy), mean strong
// fraction at theta 0.25, fraction of rows with a coefficient jump of at
// least one decade to a neighbour, the same for three decades]
void matrix_features(const Csr &A, double *f)
{
   const int n = A.nrows;
   const int nb = std::max(1, std::min((n + 255) / 256, 1024));
   double *part = dalloc_n<double>((size_t) nb * NQ);
   k_row_features<<<nb, 256, 0, stream()>>>(n, A.ia, A.ja, A.a, part);
   HX_LAUNCH_CHECK();
   std::vector<double> h((size_t) nb * NQ);
   d2h(h.data(), part, sizeof(double) * h.size());
   dfree(part);
   double q[NQ] = {0, 0, 0, 0, 0, 0, 0, 0};
   for (int b = 0; b < nb; b++) {
      for (int c = 0; c < NQ; c++) { q[c] += h[(size_t) b * NQ + c]; }
   }
   const double inv = n > 0 ? 1.0 / n : 0.0;
   const double mean_len = q[0] * inv;
   const double var_len = std::max(0.0, q[1] * inv - mean_len * mean_len);
   f[0] = n > 0 ? log10((double) n) / 8.0 : 0.0;
   f[1] = mean_len / 27.0;
   f[2] = std::sqrt(var_len) / 27.0;
   f[3] = q[2] * inv;
   f[4] = q[3] * inv;
   f[5] = q[4] * inv;
   f[6] = q[5] * inv;
   f[7] = q[6] * inv;
}

} // namespace hx

using namespace hx;

static hx_param_model_struct *as_model(HYPRE_MI355X_ParamModel m)
{
   if (!m) { fail(HYPRE_ERROR_ARG, "null parameter model"); }
   return reinterpret_cast<hx_param_model_struct *>(m);
}

extern "C" {

HYPRE_Int HYPRE_MI355X_ParamModelCreate(HYPRE_Int nin, HYPRE_Int nh1, HYPRE_Int nh2, HYPRE_Int nout,
                                        HYPRE_MI355X_ParamModel *model)
{
   return guard([&] {
      ensure_init();
      if (!model || nin <= 0 || nh1 <= 0 || nh2 <= 0 || nout <= 0) {
         fail(HYPRE_ERROR_ARG, "bad parameter model sizes");
      }
      auto *m = new hx_param_model_struct();
      m->nin = nin; m->nh1 = nh1; m->nh2 = nh2; m->nout = nout;
      m->oW1 = 0;
      m->ob1 = m->oW1 + nin * nh1;
      m->oW2 = m->ob1 + nh1;
      m->ob2 = m->oW2 + nh1 * nh2;
      m->oW3 = m->ob2 + nh2;
      m->ob3 = m->oW3 + nh2 * nout;
      m->np = m->ob3 + nout;
      m->p = dalloc_n<float>(m->np);
      m->g = dalloc_n<float>(m->np);
      m->m = dalloc_n<float>(m->np);
      m->v = dalloc_n<float>(m->np);
      m->loss = dalloc_n<float>(4);
      dzero(m->p, sizeof(float) * m->np);
      dzero(m->g, sizeof(float) * m->np);
      dzero(m->m, sizeof(float) * m->np);
      dzero(m->v, sizeof(float) * m->np);
      *model = reinterpret_cast<HYPRE_MI355X_ParamModel>(m);
   });
}

HYPRE_Int HYPRE_MI355X_ParamModelDestroy(HYPRE_MI355X_ParamModel model)
{
   return guard([&] {
      if (!model) { return; }
      auto *m = as_model(model);
      float *bufs[] = {m->p, m->g, m->m, m->v, m->loss, m->X, m->Z1, m->H1, m->Z2, m->H2, m->Z3,
                       m->Y, m->T, m->D3, m->D2, m->D1, m->DH};
      for (float *q : bufs) { dfree(q); }
      delete m;
   });
}

HYPRE_Int HYPRE_MI355X_ParamModelNumParams(HYPRE_MI355X_ParamModel model, HYPRE_Int *np)
{
   return guard([&] { *np = as_model(model)->np; });
}

HYPRE_Int HYPRE_MI355X_ParamModelSetParams(HYPRE_MI355X_ParamModel model, const float *params)
{
   return guard([&] {
      auto *m = as_model(model);
      h2d(m->p, params, sizeof(float) * m->np);
      dzero(m->m, sizeof(float) * m->np);
      dzero(m->v, sizeof(float) * m->np);
      m->t = 0;
   });
}

HYPRE_Int HYPRE_MI355X_ParamModelGetParams(HYPRE_MI355X_ParamModel model, float *params)
{
   return guard([&] {
      auto *m = as_model(model);
      d2h(params, m->p, sizeof(float) * m->np);
   });
}

HYPRE_Int HYPRE_MI355X_ParamModelForward(HYPRE_MI355X_ParamModel model, HYPRE_Int batch,
                                         const float *X, float *Y)
{
   return guard([&] {
      auto *m = as_model(model);
      if (batch <= 0) { fail(HYPRE_ERROR_ARG, "batch must be positive"); }
      ensure_batch(m, batch);
      h2d(m->X, X, sizeof(float) * (size_t) batch * m->nin);
      forward(m, batch);
      const int n3 = batch * m->nout;
      k_sigmoid<<<grid1(n3), 256, 0, stream()>>>(n3, m->Z3, m->Y);
      HX_LAUNCH_CHECK();
      d2h(Y, m->Y, sizeof(float) * n3);
   });
}

HYPRE_Int HYPRE_MI355X_ParamModelGradient(HYPRE_MI355X_ParamModel model, HYPRE_Int batch,
                                          const float *X, const float *T, float *grad, float *loss)
{
   return guard([&] {
      auto *m = as_model(model);
      if (batch <= 0) { fail(HYPRE_ERROR_ARG, "batch must be positive"); }
      ensure_batch(m, batch);
      h2d(m->X, X, sizeof(float) * (size_t) batch * m->nin);
      h2d(m->T, T, sizeof(float) * (size_t) batch * m->nout);
      forward(m, batch);
      backward(m, batch);
      if (grad) { d2h(grad, m->g, sizeof(float) * m->np); }
      if (loss) { d2h(loss, m->loss, sizeof(float)); }
   });
}

HYPRE_Int HYPRE_MI355X_ParamModelTrain(HYPRE_MI355X_ParamModel model, HYPRE_Int batch,
                                       const float *X, const float *T, HYPRE_Int steps, float lr,
                                       float *loss_hist)
{
   return guard([&] {
      auto *m = as_model(model);
      if (batch <= 0 || steps < 0) { fail(HYPRE_ERROR_ARG, "bad batch or step count"); }
      ensure_batch(m, batch);
      h2d(m->X, X, sizeof(float) * (size_t) batch * m->nin);
      h2d(m->T, T, sizeof(float) * (size_t) batch * m->nout);
      const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
      float *hist = steps > 0 ? dalloc_n<float>(steps) : nullptr;
      for (int s = 0; s < steps; s++) {
         forward(m, batch);
         backward(m, batch);
         m->t++;
         const float c1 = 1.f - powf(b1, (float) m->t), c2 = 1.f - powf(b2, (float) m->t);
         k_adam<<<grid1(m->np), 256, 0, stream()>>>(m->np, m->p, m->g, m->m, m->v, lr, b1, b2, eps,
                                                   c1, c2);
         HX_LAUNCH_CHECK();
         d2d(hist + s, m->loss, sizeof(float));
      }
      if (steps > 0 && loss_hist) { d2h(loss_hist, hist, sizeof(float) * steps); }
      dfree(hist);
   });
}

HYPRE_Int HYPRE_MI355X_MatrixFeatures(HYPRE_ParCSRMatrix A, HYPRE_Real *features)
{
   return guard([&] {
      if (!A || !features) { fail(HYPRE_ERROR_ARG, "null argument"); }
      matrix_features(A->A,features);
   });
}

HYPRE_Int HYPRE_MI355X_ParamModelSetGrid(HYPRE_MI355X_ParamModel model, HYPRE_Int ng,
                                         const HYPRE_Real *theta, const HYPRE_Real *omega,
                                         const HYPRE_Int *sweeps)
{
   return guard([&] {
      auto *m = as_model(model);
      if (ng == 0) {
         m->g_theta.clear();
         m->g_omega.clear();
         m->g_sweeps.clear();
         return;
      }
      if (ng < 0 || 2 * ng != m->nout || !theta || !omega || !sweeps) {
         fail(HYPRE_ERROR_ARG, "grid size must be half the model's outputs");
      }
      for (int g = 0; g < ng; g++) {
         if (!(theta[g] >= 0.0 && theta[g] < 1.0) || !(omega[g] > 0.0) || sweeps[g] < 1) {
            fail(HYPRE_ERROR_ARG, "grid setting out of range");
         }
      }
      m->g_theta.assign(theta, theta + ng);
      m->g_omega.assign(omega, omega + ng);
      m->g_sweeps.assign(sweeps, sweeps + ng);
      m->prior_f = m->prior_c = -1;
      m->prior_margin = 0.0;
   });
}

HYPRE_Int HYPRE_MI355X_ParamModelSetPrior(HYPRE_MI355X_ParamModel model, HYPRE_Int fine,
                                          HYPRE_Int coarse, HYPRE_Real margin)
{
   return guard([&] {
      auto *m = as_model(model);
      const int G = (int) m->g_theta.size();
      if (G == 0) { fail(HYPRE_ERROR_ARG, "the prior needs a grid (ParamModelSetGrid)"); }
      if (fine < -1 || fine >= G || coarse < -1 || coarse >= G || !(margin >= 0.0)) {
         fail(HYPRE_ERROR_ARG, "prior index outside the grid or negative margin");
      }
      m->prior_f = fine;
      m->prior_c = coarse;
      m->prior_margin = margin;
   });
}

HYPRE_Int HYPRE_MI355X_BoomerAMGSetParamModel(HYPRE_Solver solver, HYPRE_MI355X_ParamModel model,
                                              HYPRE_ParCSRMatrix A)
{
   return guard([&] {
      auto *m = as_model(model);
      if (!solver || !A) { fail(HYPRE_ERROR_ARG, "null argument"); }
      if (m->nin != 8 || (m->g_theta.empty() && m->nout % 3 != 0)) {
         fail(HYPRE_ERROR_ARG, "model shape is not 8 -> 3L or 8 -> 2G with a grid");
      }
      double f[8];
      matrix_features(A->A,f);
      float xf[8];
      for (int c = 0; c < 8; c++) { xf[c] = (float) f[c]; }
      std::vector<float> y(m->nout);
      ensure_batch(m, 1);
      h2d(m->X, xf, sizeof(xf));
      forward(m, 1);
      k_sigmoid<<<1, 256, 0, stream()>>>(m->nout, m->Z3, m->Y);
      HX_LAUNCH_CHECK();
      d2h(y.data(), m->Y, sizeof(float) * m->nout);
      if (!m->g_theta.empty()) {
         // cost model: the first predicted minimum of each group, unless the
         // group's prior setting is predicted within the margin of it
         const int G = (int) m->g_theta.size();
         int gf = (int) (std::min_element(y.begin(), y.begin() + G) - y.begin());
         int gc = (int) (std::min_element(y.begin() + G, y.begin() + 2 * G) - y.begin()) - G;
         if (m->prior_f >= 0 && (double) y[gf] > (double) y[m->prior_f] - m->prior_margin) {
            gf = m->prior_f;
         }
         if (m->prior_c >= 0 &&
             (double) y[G + gc] > (double) y[G + m->prior_c] - m->prior_margin) {
            gc = m->prior_c;
         }
         for (int l = 0; l < kModelLevels; l++) {
            const int g = (l == 0) ? gf : gc;
            HYPRE_MI355X_BoomerAMGSetLevelStrongThreshold(solver, l, m->g_theta[g]);
            HYPRE_MI355X_BoomerAMGSetLevelRelaxWt(solver, l, m->g_omega[g]);
            HYPRE_MI355X_BoomerAMGSetLevelNumSweeps(solver, l, m->g_sweeps[g]);
         }
         return;
      }
      const int L = m->nout / 3;
      for (int l = 0; l < L; l++) {
         const double theta = 0.1 + 0.8 * (double) y[3 * l];
         const double omega = 0.6 + 0.8 * (double) y[3 * l + 1];
         const int sweeps = 1 + (int) std::floor(2.0 * (double) y[3 * l + 2] + 0.5);
         HYPRE_MI355X_BoomerAMGSetLevelStrongThreshold(solver, l, theta);
         HYPRE_MI355X_BoomerAMGSetLevelRelaxWt(solver, l, omega);
         HYPRE_MI355X_BoomerAMGSetLevelNumSweeps(solver, l, sweeps);
      }
   });
}

} // extern "C"
