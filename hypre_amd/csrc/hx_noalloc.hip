// HYPRE_MI355X_ParVectorInnerProdNoAlloc: HYPRE_ParVectorInnerProd for
// callers that cannot allocate (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).
//
// HYPRE_ParVectorInnerProd keeps its per-block partial sums in a buffer the
// library allocates on first use and reads the result back through pinned
// host memory it also allocates.  This variant keeps the partial sums and the
// result in `scratch`, a fixed device array of this file, and copies the
// result straight into the caller's variable, so a call allocates nothing.
// The arithmetic is the same as HYPRE_ParVectorInnerProd's (the same grid,
// the same per-block tree and the same final fold), so both return the same
// bits.
//
// The array's address on a device comes from hipGetSymbolAddress, taken once
// per device when the library is initialised on it (HYPRE_Initialize) and
// kept per device; host code never passes the array's name itself.  All
// calls on one device share its array: calls that overlap on one device race
// (the library runs its work in order on one stream, so calls made from one
// host thread never overlap).  Every partial sum a call reads is written by
// that call, so nothing has to be cleared between calls.
#include "hx_internal.h"
#include "hx_objects.h"

#include <algorithm>

namespace hx {

namespace {
constexpr int NA_BS = 256;
constexpr int NA_PARTS = 2048;            // most blocks one call launches
constexpr int NA_MAX_DEV = 64;
} // namespace

// NA_PARTS partial sums, then the result: 16.6 KB per device, held until the
// process ends
static __device__ __align__(256) double scratch[NA_PARTS + 32];

// the array's address on each device (hipGetSymbolAddress)
static double *g_scratch_dev[NA_MAX_DEV] = {nullptr};

namespace {

__global__ __launch_bounds__(NA_BS) void k_na_dot_part(int n, const double *__restrict__ x,
                                                       const double *__restrict__ y,
                                                       double *__restrict__ part)
{
   __shared__ double sh[NA_BS];
   double s = 0.0;
   for (long long i = (long long) blockIdx.x * NA_BS + threadIdx.x; i < n;
        i += (long long) gridDim.x * NA_BS) {
      s += x[i] * y[i];
   }
   sh[threadIdx.x] = s;
   __syncthreads();
   for (int off = NA_BS / 2; off > 0; off >>= 1) {
      if (threadIdx.x < off) { sh[threadIdx.x] += sh[threadIdx.x + off]; }
      __syncthreads();
   }
   if (threadIdx.x == 0) { part[blockIdx.x] = sh[0]; }
}

__global__ __launch_bounds__(NA_BS) void k_na_reduce(const double *__restrict__ part, int np,
                                                     double *__restrict__ out)
{
   __shared__ double sh[NA_BS];
   double s = 0.0;
   for (int i = threadIdx.x; i < np; i += NA_BS) { s += part[i]; }
   sh[threadIdx.x] = s;
   __syncthreads();
   for (int off = NA_BS / 2; off > 0; off >>= 1) {
      if (threadIdx.x < off) { sh[threadIdx.x] += sh[threadIdx.x + off]; }
      __syncthreads();
   }
   if (threadIdx.x == 0) { *out = sh[0]; }
}

} // namespace

// called by the library's initialisation with the device current
void noalloc_attach_device()
{
   int dev = 0;
   HX_HIP(hipGetDevice(&dev));
   if (dev < 0 || dev >= NA_MAX_DEV) { fail(HYPRE_ERROR_GENERIC, "device index out of range"); }
   if (g_scratch_dev[dev]) { return; }
   void *p = nullptr;
   HX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
   g_scratch_dev[dev] = static_cast<double *>(p);
}

} // namespace hx

using namespace hx;

HYPRE_Int HYPRE_MI355X_ParVectorInnerProdNoAlloc(HYPRE_ParVector x, HYPRE_ParVector y,
                                                 HYPRE_Real *prod)
{
   return guard([&] {
      if (!x || !y || !prod) { g_error_arg = 1; fail(HYPRE_ERROR_ARG, "null argument"); }
      if (x->n != y->n) {
         g_error_arg = 2;
         fail(HYPRE_ERROR_ARG, "vectors of different lengths");
      }
      if (x->n < 0) { g_error_arg = 1; fail(HYPRE_ERROR_ARG, "negative vector length"); }
      int dev = 0;
      HX_HIP(hipGetDevice(&dev));
      double *sp = (dev >= 0 && dev < NA_MAX_DEV) ? g_scratch_dev[dev] : nullptr;
      if (!sp) {
         fail(HYPRE_ERROR_GENERIC, "HYPRE_Initialize has not run on the current device");
      }
      const int n = x->n;
      // the grid of HYPRE_ParVectorInnerProd: at most NA_PARTS blocks, so the
      // partial sums always fit the array (the bound is checked regardless)
      const int nb = (int) std::max<long long>(
         1, std::min<long long>(((long long) n + NA_BS * 4 - 1) / (NA_BS * 4), NA_PARTS));
      if (nb > NA_PARTS) { fail(HYPRE_ERROR_GENERIC, "partial sums exceed the fixed scratch"); }
      hipStream_t st = stream();
      k_na_dot_part<<<nb, NA_BS, 0, st>>>(n, x->d, y->d, sp);
      k_na_reduce<<<1, NA_BS, 0, st>>>(sp, nb, sp + NA_PARTS);
      HX_LAUNCH_CHECK();
      HX_HIP(hipMemcpyAsync(prod, sp + NA_PARTS, sizeof(double), hipMemcpyDeviceToHost, st));
      HX_HIP(hipStreamSynchronize(st));
   });
}
