// Runtime of libhypre_mi355x.so: error flag, library stream, caching device
// allocator, prefix scan / reductions, HIP-event kernel timing, CSR helpers.
#include "hx_internal.h"

#include <dlfcn.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>

namespace hx {

int g_error_flag = 0;
int g_error_arg  = 0;

void set_error(int code, const char *msg)
{
   g_error_flag |= code;
   if (msg && getenv("HYPRE_MI355X_VERBOSE")) { fprintf(stderr, "[hypre-mi355x] %s\n", msg); }
}

void fail(int code, const std::string &msg) { throw Error(code, msg); }

// ------------------------------------------------------------------ stream
static hipStream_t g_stream = nullptr;
static bool        g_inited = false;
static int        *g_flags = nullptr;
static void       *g_pinned = nullptr;
static size_t      g_pinned_bytes = 0;

void ensure_init()
{
   if (g_inited) { return; }
   int ndev = 0;
   HX_HIP(hipGetDeviceCount(&ndev));
   if (ndev < 1) { fail(HYPRE_ERROR_GENERIC, "no HIP device"); }
   HX_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
   HX_HIP(hipMalloc(&g_flags, 1024 * sizeof(int)));
   HX_HIP(hipMemset(g_flags, 0, 1024 * sizeof(int)));
   noalloc_attach_device();
   g_inited = true;
}

hipStream_t stream()
{
   ensure_init();
   return g_stream;
}

void sync() { HX_HIP(hipStreamSynchronize(stream())); }

// A second stream for setup work that runs beside the library stream (the
// solve-phase copies of the level operators).  Work is handed over in both
// directions with events: side_after_main() makes later side-stream work wait
// for everything enqueued on the library stream so far (so no block the
// allocator recycled is still in use there), main_after_side() the reverse.
// Both streams have the default priority.
static hipStream_t g_side = nullptr;
static hipEvent_t  g_ev_main = nullptr, g_ev_side = nullptr;

hipStream_t side_stream()
{
   ensure_init();
   if (!g_side) {
      HX_HIP(hipStreamCreateWithFlags(&g_side, hipStreamNonBlocking));
      HX_HIP(hipEventCreateWithFlags(&g_ev_main, hipEventDisableTiming));
      HX_HIP(hipEventCreateWithFlags(&g_ev_side, hipEventDisableTiming));
   }
   return g_side;
}

void side_after_main()
{
   hipStream_t s = side_stream();
   HX_HIP(hipEventRecord(g_ev_main, stream()));
   HX_HIP(hipStreamWaitEvent(s, g_ev_main, 0));
}

void main_after_side()
{
   if (!g_side) { return; }
   HX_HIP(hipEventRecord(g_ev_side, g_side));
   HX_HIP(hipStreamWaitEvent(stream(), g_ev_side, 0));
}

void side_sync()
{
   if (g_side) { HX_HIP(hipStreamSynchronize(g_side)); }
}

// ------------------------------------------------------------------ profiling ranges
// roctx is loaded when ranges are first enabled (the rocprofiler-sdk roctx
// library that rocprofv3 --marker-trace records, else the older roctx64).
static int (*g_roctx_push)(const char *) = nullptr;
static int (*g_roctx_pop)() = nullptr;
static int g_ranges = -1;          // -1: environment not read yet
static long long g_range_pushed = 0, g_range_open = 0;

static bool roctx_load()
{
   if (g_roctx_push) { return true; }
   static const char *libs[] = {"librocprofiler-sdk-roctx.so.1", "/opt/rocm/lib/librocprofiler-sdk-roctx.so.1",
                                "libroctx64.so.4", "/opt/rocm/lib/libroctx64.so.4"};
   for (const char *name : libs) {
      void *h = dlopen(name, RTLD_NOW | RTLD_GLOBAL);
      if (!h) { continue; }
      auto push = reinterpret_cast<int (*)(const char *)>(dlsym(h, "roctxRangePushA"));
      auto pop = reinterpret_cast<int (*)()>(dlsym(h, "roctxRangePop"));
      if (push && pop) {
         g_roctx_push = push;
         g_roctx_pop = pop;
         return true;
      }
   }
   return false;
}

bool ranges_on()
{
   if (g_ranges < 0) {
      const char *e = std::getenv("HYPRE_MI355X_ROCTX");
      g_ranges = (e && std::atoi(e) != 0 && roctx_load()) ? 1 : 0;
   }
   return g_ranges == 1;
}

void range_push(const char *name)
{
   g_roctx_push(name);
   g_range_pushed++;
   g_range_open++;
}

void range_pop()
{
   // src/utilities/device_markers.c:117-131: the stream sync inside a range
   g_roctx_push("StreamSync0");
   (void) hipStreamSynchronize(stream());
   g_roctx_pop();
   g_roctx_pop();
   g_range_open--;
}

Range::Range(const char *fmt, int v) : on(ranges_on())
{
   if (on) {
      char name[64];
      std::snprintf(name, sizeof(name), fmt, v);
      range_push(name);
   }
}

int *device_flags()
{
   ensure_init();
   return g_flags;
}

void *pinned_scratch(size_t bytes)
{
   ensure_init();
   if (bytes > g_pinned_bytes) {
      if (g_pinned) { HX_HIP(hipHostFree(g_pinned)); }
      size_t nb = std::max<size_t>(bytes, 1 << 16);
      HX_HIP(hipHostMalloc(&g_pinned, nb, hipHostMallocDefault));
      g_pinned_bytes = nb;
   }
   return g_pinned;
}

// ------------------------------------------------------------------ allocator
// Size classes: powers of two up to 64 MiB, then multiples of 64 MiB.  Freed
// blocks are kept for reuse (stream-ordered: every kernel runs on the single
// library stream, so a block reused by a later launch is never in flight).
static std::multimap<size_t, void *>     g_free;
static std::unordered_map<void *, size_t> g_live;
static size_t g_cached_bytes = 0, g_live_bytes = 0;

static size_t size_class(size_t b)
{
   const size_t big = size_t(64) << 20;
   if (b <= 256) { return 256; }
   if (b <= big) {
      size_t c = 256;
      while (c < b) { c <<= 1; }
      return c;
   }
   return (b + big - 1) / big * big;
}

void *dalloc(size_t bytes)
{
   ensure_init();
   size_t c = size_class(bytes);
   auto it = g_free.find(c);
   void *p = nullptr;
   if (it != g_free.end()) {
      p = it->second;
      g_free.erase(it);
      g_cached_bytes -= c;
   } else {
      hipError_t e = hipMalloc(&p, c);
      if (e != hipSuccess) {
         (void) hipGetLastError();
         release_cache();
         e = hipMalloc(&p, c);
         if (e != hipSuccess) {
            (void) hipGetLastError();
            fail(HYPRE_ERROR_MEMORY, "hipMalloc of " + std::to_string(c) + " bytes failed");
         }
      }
   }
   g_live[p] = c;
   g_live_bytes += c;
   return p;
}

void dfree(void *p)
{
   if (!p) { return; }
   auto it = g_live.find(p);
   if (it == g_live.end()) { return; }
   g_free.emplace(it->second, p);
   g_cached_bytes += it->second;
   g_live_bytes -= it->second;
   g_live.erase(it);
}

void release_cache()
{
   if (!g_inited) { return; }
   (void) hipStreamSynchronize(g_stream);
   if (g_side) { (void) hipStreamSynchronize(g_side); }
   for (auto &kv : g_free) { (void) hipFree(kv.second); }
   g_free.clear();
   g_cached_bytes = 0;
}

void h2d(void *d, const void *h, size_t bytes)
{
   if (!bytes) { return; }
   HX_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, stream()));
   sync();
}

void d2h(void *h, const void *d, size_t bytes)
{
   if (!bytes) { return; }
   HX_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, stream()));
   sync();
}

void d2d(void *dst, const void *src, size_t bytes)
{
   if (!bytes) { return; }
   HX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream()));
}

__global__ void k_zero_bytes(unsigned int *p, size_t nwords)
{
   size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
   size_t stride = (size_t) gridDim.x * blockDim.x;
   for (; i < nwords; i += stride) { p[i] = 0u; }
}

void dzero(void *d, size_t bytes)
{
   if (!bytes) { return; }
   // a fill kernel instead of hipMemsetAsync keeps the stream capturable
   size_t nwords = (bytes + 3) / 4;
   int blocks = (int) std::min<size_t>((nwords + 255) / 256, 8192);
   k_zero_bytes<<<blocks, 256, 0, stream()>>>(static_cast<unsigned int *>(d), nwords);
   HX_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ scan
// Three passes: per-block totals and maxima (2048 items / block), a
// single-block scan of the totals (64-bit, overflow-checked) and reduction of
// the maxima, then the per-block exclusive scan.
constexpr int SCAN_BS = 256;
constexpr int SCAN_IT = 8;
constexpr int SCAN_TILE = SCAN_BS * SCAN_IT;

__device__ __forceinline__ long long block_excl_scan_ll(long long v, long long *sh, long long *total)
{
   // Hillis-Steele in shared memory over 256 threads
   int t = threadIdx.x;
   sh[t] = v;
   __syncthreads();
   for (int off = 1; off < SCAN_BS; off <<= 1) {
      long long add = (t >= off) ? sh[t - off] : 0;
      __syncthreads();
      sh[t] += add;
      __syncthreads();
   }
   long long incl = sh[t];
   *total = sh[SCAN_BS - 1];
   __syncthreads();
   return incl - v;
}

__global__ __launch_bounds__(SCAN_BS) void k_scan_tiles(const int *__restrict__ in, int n,
                                                        long long *__restrict__ tile_sum,
                                                        int *__restrict__ tile_max)
{
   __shared__ long long sh[SCAN_BS];
   __shared__ int wmax[SCAN_BS / 64];
   long long base = (long long) blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_IT;
   long long s = 0;
   int m = INT_MIN;
#pragma unroll
   for (int k = 0; k < SCAN_IT; k++) {
      long long i = base + k;
      if (i < n) {
         const int v = in[i];
         s += v;
         m = max(m, v);
      }
   }
   for (int o = 32; o > 0; o >>= 1) { m = max(m, __shfl_xor(m, o, 64)); }
   if ((threadIdx.x & 63) == 0) { wmax[threadIdx.x >> 6] = m; }
   long long tot;
   block_excl_scan_ll(s, sh, &tot);   // its barriers order the wmax stores
   if (threadIdx.x == 0) {
      tile_sum[blockIdx.x] = tot;
      int bm = wmax[0];
      for (int w = 1; w < SCAN_BS / 64; w++) { bm = max(bm, wmax[w]); }
      tile_max[blockIdx.x] = bm;
   }
}

__global__ __launch_bounds__(SCAN_BS) void k_scan_sums(long long *__restrict__ tile_sum,
                                                       const int *__restrict__ tile_max, int ntiles,
                                                       int *__restrict__ out_total_slot,
                                                       int *__restrict__ overflow)
{
   __shared__ long long sh[SCAN_BS];
   __shared__ int wmax[SCAN_BS / 64];
   int per = (ntiles + SCAN_BS - 1) / SCAN_BS;
   int b0 = threadIdx.x * per;
   long long s = 0;
   int m = INT_MIN;
   for (int k = 0; k < per; k++) {
      int i = b0 + k;
      if (i < ntiles) {
         s += tile_sum[i];
         m = max(m, tile_max[i]);
      }
   }
   for (int o = 32; o > 0; o >>= 1) { m = max(m, __shfl_xor(m, o, 64)); }
   if ((threadIdx.x & 63) == 0) { wmax[threadIdx.x >> 6] = m; }
   long long tot;
   long long ex = block_excl_scan_ll(s, sh, &tot);
   long long run = ex;
   for (int k = 0; k < per; k++) {
      int i = b0 + k;
      if (i < ntiles) {
         long long v = tile_sum[i];
         tile_sum[i] = run;
         run += v;
      }
   }
   if (threadIdx.x == 0) {
      *out_total_slot = (tot > 0x7fffffffLL) ? 0x7fffffff : (int) tot;
      overflow[0] = (tot > 0x7fffffffLL) ? 1 : 0;
      int bm = wmax[0];
      for (int w = 1; w < SCAN_BS / 64; w++) { bm = max(bm, wmax[w]); }
      overflow[1] = bm;
      reinterpret_cast<long long *>(overflow + 2)[0] = tot;
   }
}

__global__ __launch_bounds__(SCAN_BS) void k_scan_apply(const int *in, int n, int *out,
                                                        const long long *__restrict__ tile_off)
{
   __shared__ long long sh[SCAN_BS];
   long long base = (long long) blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_IT;
   int v[SCAN_IT];
   long long s = 0;
#pragma unroll
   for (int k = 0; k < SCAN_IT; k++) {
      long long i = base + k;
      v[k] = (i < n) ? in[i] : 0;
      s += v[k];
   }
   long long tot;
   long long ex = block_excl_scan_ll(s, sh, &tot) + tile_off[blockIdx.x];
   __syncthreads();
#pragma unroll
   for (int k = 0; k < SCAN_IT; k++) {
      long long i = base + k;
      if (i < n) { out[i] = (int) ex; }
      ex += v[k];
   }
}

// the three passes on the library stream; the total's 32-bit overflow flag,
// the largest input and the 64-bit total land at flags + 8, 9 and 10..11
static void scan_launch(const int *in, int *out, int n)
{
   int ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
   long long *tiles = dalloc_n<long long>(ntiles);
   int *tmax = dalloc_n<int>(ntiles);
   int *flags = device_flags();
   k_scan_tiles<<<ntiles, SCAN_BS, 0, stream()>>>(in, n, tiles, tmax);
   HX_LAUNCH_CHECK();
   k_scan_sums<<<1, SCAN_BS, 0, stream()>>>(tiles, tmax, ntiles, out + n, flags + 8);
   HX_LAUNCH_CHECK();
   k_scan_apply<<<ntiles, SCAN_BS, 0, stream()>>>(in, n, out, tiles);
   HX_LAUNCH_CHECK();
   // stream-ordered reuse: later work on this stream runs after the passes
   dfree(tiles);
   dfree(tmax);
}

long long exclusive_scan(const int *in, int *out, int n, int *max_in)
{
   if (n <= 0) {
      int z = 0;
      h2d(out, &z, sizeof(int));
      if (max_in) { *max_in = 0; }
      return 0;
   }
   scan_launch(in, out, n);
   int *h = static_cast<int *>(pinned_scratch(64));
   HX_HIP(hipMemcpyAsync(h, device_flags() + 8, 4 * sizeof(int), hipMemcpyDeviceToHost, stream()));
   sync();
   long long tot;
   std::memcpy(&tot, h + 2, sizeof(long long));
   if (h[0]) { fail(HYPRE_ERROR_MEMORY, "prefix sum exceeds 32-bit index range"); }
   if (max_in) { *max_in = h[1]; }
   return tot;
}

void exclusive_scan_nowait(const int *in, int *out, int n)
{
   if (n <= 0) {
      dzero(out, sizeof(int));
      return;
   }
   scan_launch(in, out, n);
}

__global__ void k_reduce_int(const int *__restrict__ in, int n, unsigned long long *sum, int *mx)
{
   __shared__ long long ss[256];
   __shared__ int sm[256];
   long long s = 0;
   int m = INT_MIN;
   for (long long i = (long long) blockIdx.x * 256 + threadIdx.x; i < n; i += (long long) gridDim.x * 256) {
      int v = in[i];
      s += v;
      m = v > m ? v : m;
   }
   ss[threadIdx.x] = s;
   sm[threadIdx.x] = m;
   __syncthreads();
   for (int off = 128; off > 0; off >>= 1) {
      if (threadIdx.x < off) {
         ss[threadIdx.x] += ss[threadIdx.x + off];
         sm[threadIdx.x] = max(sm[threadIdx.x], sm[threadIdx.x + off]);
      }
      __syncthreads();
   }
   if (threadIdx.x == 0) {
      atomicAdd(sum, (unsigned long long) ss[0]);
      atomicMax(mx, sm[0]);
   }
}

__global__ void k_reduce_init(unsigned long long *sum, int *mx)
{
   *sum = 0ULL;
   *mx = INT_MIN;
}

static void reduce_int(const int *in, int n, long long *sum, int *mx)
{
   int *flags = device_flags();
   unsigned long long *dsum = reinterpret_cast<unsigned long long *>(flags + 16);
   int *dmx = flags + 18;
   k_reduce_init<<<1, 1, 0, stream()>>>(dsum, dmx);
   if (n > 0) {
      int blocks = std::min((n + 255) / 256, 2048);
      k_reduce_int<<<blocks, 256, 0, stream()>>>(in, n, dsum, dmx);
      HX_LAUNCH_CHECK();
   }
   int *h = static_cast<int *>(pinned_scratch(64));
   HX_HIP(hipMemcpyAsync(h, flags + 16, 3 * sizeof(int), hipMemcpyDeviceToHost, stream()));
   sync();
   unsigned long long s;
   std::memcpy(&s, h, sizeof(s));
   *sum = (long long) s;
   *mx = h[2];
}

long long reduce_sum_int(const int *in, int n)
{
   long long s; int m;
   reduce_int(in, n, &s, &m);
   return s;
}

int reduce_max_int(const int *in, int n)
{
   long long s; int m;
   reduce_int(in, n, &s, &m);
   return n > 0 ? m : 0;
}

// ------------------------------------------------------------------ timing
struct TimingRec
{
   std::string name;
   hipEvent_t  e0, e1;
   double      bytes;
};
static bool                    g_timing = false;
static std::string             g_timing_only;      // empty: every class
static std::vector<TimingRec>  g_trecs;
static std::vector<hipEvent_t> g_event_pool;

bool timing_on() { return g_timing; }

static hipEvent_t get_event()
{
   if (!g_event_pool.empty()) {
      hipEvent_t e = g_event_pool.back();
      g_event_pool.pop_back();
      return e;
   }
   hipEvent_t e;
   HX_HIP(hipEventCreate(&e));
   return e;
}

TimedLaunch::TimedLaunch(const char *cls, double bytes)
{
   if (!g_timing || !cls) { return; }
   if (!g_timing_only.empty() && g_timing_only != cls) { return; }
   TimingRec r;
   r.name = cls;
   r.e0 = get_event();
   r.e1 = get_event();
   r.bytes = bytes;
   HX_HIP(hipEventRecord(r.e0, stream()));
   g_trecs.push_back(r);
   slot = (int) g_trecs.size() - 1;
}

TimedLaunch::~TimedLaunch()
{
   if (slot >= 0 && slot < (int) g_trecs.size()) {
      (void) hipEventRecord(g_trecs[slot].e1, stream());
   }
}

static void timing_reset()
{
   for (auto &r : g_trecs) {
      g_event_pool.push_back(r.e0);
      g_event_pool.push_back(r.e1);
   }
   g_trecs.clear();
}

// ------------------------------------------------------------------ CSR
void Csr::free_all()
{
   dfree(ia); dfree(ja); dfree(a); dfree(a32);
   ia = ja = nullptr; a = nullptr; a32 = nullptr;
   nrows = ncols = 0; nnz = 0;
   ell = nullptr;
}

void Ell::free_all()
{
   dfree(sw); dfree(j); dfree(j16); dfree(j8); dfree(jb); dfree(mode); dfree(a); dfree(a32);
   sw = j = nullptr; j16 = nullptr; j8 = nullptr; jb = mode = nullptr; a = nullptr; a32 = nullptr;
   nslices = 0; stored = 0; wmax = 0;
}

void Csr::alloc(int n, int nc, long long nz, bool values)
{
   nrows = n; ncols = nc; nnz = nz;
   ia = dalloc_n<int>((size_t) n + 1);
   ja = dalloc_n<int>((size_t) std::max<long long>(nz, 1));
   a = values ? dalloc_n<double>((size_t) std::max<long long>(nz, 1)) : nullptr;
   a32 = nullptr;
}

void csr_copy(const Csr &s, Csr &d)
{
   d.alloc(s.nrows, s.ncols, s.nnz, s.a != nullptr);
   d2d(d.ia, s.ia, sizeof(int) * ((size_t) s.nrows + 1));
   d2d(d.ja, s.ja, sizeof(int) * (size_t) s.nnz);
   if (s.a) { d2d(d.a, s.a, sizeof(double) * (size_t) s.nnz); }
}

void csr_to_host(const Csr &A, std::vector<int> &ia, std::vector<int> &ja, std::vector<double> &a)
{
   ia.resize((size_t) A.nrows + 1);
   ja.resize((size_t) A.nnz);
   d2h(ia.data(), A.ia, sizeof(int) * ia.size());
   d2h(ja.data(), A.ja, sizeof(int) * ja.size());
   if (A.a) {
      a.resize((size_t) A.nnz);
      d2h(a.data(), A.a, sizeof(double) * a.size());
   } else {
      a.clear();
   }
}

void csr_from_host(Csr &A, int n, int ncols, const int *ia, const int *ja, const double *a)
{
   long long nnz = ia[n];
   A.alloc(n, ncols, nnz, a != nullptr);
   h2d(A.ia, ia, sizeof(int) * ((size_t) n + 1));
   h2d(A.ja, ja, sizeof(int) * (size_t) nnz);
   if (a) { h2d(A.a, a, sizeof(double) * (size_t) nnz); }
}

} // namespace hx

// ------------------------------------------------------------------ C API
extern "C" {

HYPRE_Int HYPRE_MI355X_SetProfilingRanges(HYPRE_Int on)
{
   if (on && !hx::roctx_load()) {
      hx::set_error(HYPRE_ERROR_GENERIC, "roctx library not found");
      hx::g_ranges = 0;
      return hx::g_error_flag;
   }
   hx::g_ranges = on ? 1 : 0;
   return hx::g_error_flag;
}

HYPRE_Int HYPRE_MI355X_GetProfilingRanges(HYPRE_Int *pushed, HYPRE_Int *open)
{
   if (pushed) { *pushed = (HYPRE_Int) hx::g_range_pushed; }
   if (open) { *open = (HYPRE_Int) hx::g_range_open; }
   return hx::g_error_flag;
}

HYPRE_Int HYPRE_MI355X_SetKernelTiming(HYPRE_Int on)
{
   hx::g_timing = (on != 0);
   return hx::g_error_flag;
}

HYPRE_Int HYPRE_MI355X_SetKernelTimingClass(const char *name)
{
   hx::g_timing_only = name ? name : "";
   return hx::g_error_flag;
}

HYPRE_Int HYPRE_MI355X_ResetKernelTiming(void)
{
   return hx::guard([&] {
      hx::sync();
      hx::timing_reset();
   });
}

HYPRE_Int HYPRE_MI355X_GetKernelTiming(const char *name, HYPRE_Real *total_ms, HYPRE_Int *launches,
                                       HYPRE_Real *bytes)
{
   return hx::guard([&] {
      hx::sync();
      double t = 0.0, b = 0.0;
      int cnt = 0;
      for (auto &r : hx::g_trecs) {
         if (r.name != name) { continue; }
         float ms = 0.f;
         HX_HIP(hipEventElapsedTime(&ms, r.e0, r.e1));
         t += ms;
         b += r.bytes;
         cnt++;
      }
      if (total_ms) { *total_ms = t; }
      if (launches) { *launches = cnt; }
      if (bytes) { *bytes = b; }
   });
}

} // extern "C"
