// Internal helpers shared by the libssim translation units (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "../../include/ssim.h"

namespace ssim {

// Thread-local message of the last failure, returned by ssim_last_error().
void set_error(const char* fmt, ...);
void clear_error();

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

// Round a byte count up so consecutive workspace slices stay 256-B aligned.
inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// All-pairs force (allpairs.hip); shared by the integrators.
size_t allpairs_work_bytes(int dim, int ns, int nt);
int allpairs(int dim, int prec, int ns, const double* sx, const double* sy,
             const double* sz, const double* sm, int nt, const double* tx,
             const double* ty, const double* tz, double* ax, double* ay,
             double* az, double G, double eps2, void* work, size_t work_bytes,
             hipStream_t st);

// Elementwise integrator kernels (integrate.hip).
int kick(int dim, int n, double* vx, double* vy, double* vz, const double* ax,
         const double* ay, const double* az, double h, hipStream_t st);
int drift(int dim, int n, double* x, double* y, double* z, const double* vx,
          const double* vy, const double* vz, double h, hipStream_t st);

// Yoshida 4th-order composition weights w1, w0 (nbody_kernel.py:64-65).
double yoshida_w1();
double yoshida_w0();

// Exclusive scan of n ints into out[0..n] (out[n] = total), sort.hip;
// `work` holds scan_work_bytes(n) (unused for n <= 65536: one workgroup).
// every > 0 (dividing n) also writes sel[j] = out[every j], j = 0..n/every.
size_t scan_work_bytes(int n);
int scan_exclusive(int n, const int* in, int* out, void* work, hipStream_t st,
                   int every = 0, int* sel = nullptr);

// Launch timing (abi.hip): SSIM_TIMED wraps a launch statement.
bool timing_on();
void* timing_start(hipStream_t st);
void timing_stop(const char* name, void* start, hipStream_t st);

// Stable LSD radix sort of (keys, vals) in place on keys in [0, 2^bits).
// first_counted: the caller's key kernel has already written the first
// pass's per-tile digit counts (8 low key bits) into radix_first_counts(work,
// n), in k_radix_hist's layout: tiles of RADIX_TILE keys, RADIX_T threads
// each taking keys base + r RADIX_T + t, r < RADIX_TILE / RADIX_T, counts
// digit-major (row stride *hrow, the tile count rounded up to 4, zero past
// the last tile).  vals_identity: the input values are 0..n-1 (not read;
// vals is output only).
constexpr int RADIX_T = 256, RADIX_TILE = 1024;
size_t radix_sort_work_bytes(int n);
int* radix_first_counts(void* work, int n, int* hrow);
int radix_sort_pairs(int n, int bits, int* keys, int* vals, void* work,
                     size_t work_bytes, hipStream_t st, bool first_counted = false,
                     bool vals_identity = false);

}  // namespace ssim

// Record a HIP failure and return SSIM_EHIP from the enclosing function.
#define SSIM_HIP(call)                                                       \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess) {                                                  \
      ssim::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                      __FILE__, __LINE__);                                   \
      return SSIM_EHIP;                                                      \
    }                                                                        \
  } while (0)

// Check the launch that was just enqueued.
#define SSIM_LAUNCHED(name)                                                   \
  do {                                                                        \
    hipError_t e_ = hipGetLastError();                                        \
    if (e_ != hipSuccess) {                                                   \
      ssim::set_error("launch of %s failed: %s", name, hipGetErrorString(e_)); \
      return SSIM_EHIP;                                                       \
    }                                                                         \
  } while (0)

// Reject an argument: record the message and return SSIM_EINVAL.
#define SSIM_REQUIRE(cond, ...)   \
  do {                            \
    if (!(cond)) {                \
      ssim::set_error(__VA_ARGS__); \
      return SSIM_EINVAL;         \
    }                             \
  } while (0)

// Enqueue `launch` (a hipLaunchKernelGGL statement) on stream st; while
// timing is enabled it is bracketed by events recorded under `name`.
#define SSIM_TIMED(name, st, launch)                                \
  do {                                                              \
    void* t0_ = ssim::timing_on() ? ssim::timing_start(st) : nullptr; \
    launch;                                                         \
    if (t0_) ssim::timing_stop(name, t0_, st);                      \
  } while (0)

// Propagate a non-zero status from a nested call.
#define SSIM_TRY(call)       \
  do {                       \
    int rc_ = (call);        \
    if (rc_ != SSIM_OK) return rc_; \
  } while (0)

// ---------------------------------------------------------------------
// Device arithmetic helpers
// ---------------------------------------------------------------------

// 1/sqrt(x) to ~1 ulp in FP64: the hardware v_rsq_f64 estimate (relative
// error about 2^-23) refined by one third-order Householder step,
//   y1 = y0 * (1 + e/2 + 3 e^2/8),  e = 1 - x*y0^2,
// whose truncation error 5/16 e^3 is ~2^-69, below FP64 rounding.  The
// library is built with -ffp-contract=off, so every fma here is explicit.
__device__ __forceinline__ double ssim_rsqrt_f64(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double t = x * y;
  const double e = fma(-t, y, 1.0);
  const double p = fma(e, 0.375, 0.5);
  const double q = e * p;
  return fma(y, q, y);
}

// 1/sqrt(x) in FP32: v_rsq_f32 (about 1 ulp) plus one Newton step.
__device__ __forceinline__ float ssim_rsqrt_f32(float x) {
  const float y = __builtin_amdgcn_rsqf(x);
  const float t = x * y;
  const float e = fmaf(-t, y, 1.0f);
  return fmaf(y * 0.5f, e, y);
}

// Split function of the particle-mesh + near-field force, shared by the
// mesh kernel (pm.hip) and the near-field table (pp.hip):
//   g(r) = erfc(r/2rs) + r/(rs sqrt(pi)) exp(-r^2/4rs^2)   (Gaussian split)
// tapered to zero over the last rs before rcut with the C2 step
// S(u) = 1 - u^3 (10 - 15u + 6u^2), u = (r - (rcut - rs)) / rs, so the
// near part F g S vanishes smoothly at rcut and the mesh part F (1 - g S)
// is continuous (a hard cut leaves a jump of F g(rcut) in the mesh kernel,
// sampled unevenly by the mesh nodes).  rcut <= 0: no taper.
__host__ __device__ inline double ssim_split_g(double r, double rs, double rcut) {
  const double x = r / (2.0 * rs);
  double g = erfc(x) + r / (rs * 1.7724538509055160273) * exp(-x * x);
  if (rcut > 0.0) {
    if (r >= rcut) return 0.0;
    const double u = (r - (rcut - rs)) / rs;
    if (u > 0.0) g *= fmax(1.0 - u * u * u * (10.0 - 15.0 * u + 6.0 * u * u), 0.0);
  }
  return g;
}

// XCD-aware block order.  Workgroups are dispatched round-robin over the 8
// XCDs (block b runs on XCD b % 8), each with its own L2; a kernel whose
// neighbouring blocks share data (a mesh stencil, a source list, a cache
// line) renumbers them so that XCD x takes the contiguous logical range
// [start(x), start(x) + count(x)) in dispatch order.  A bijection on
// [0, nblocks).
constexpr int SSIM_XCDS = 8;
__device__ __forceinline__ int ssim_xcd_block(int b, int nblocks) {
  const int per = nblocks / SSIM_XCDS, rem = nblocks % SSIM_XCDS;
  const int x = b % SSIM_XCDS, i = b / SSIM_XCDS;
  return (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + i;
}

// Bin key of mesh node (i, j, k) in ssim_split_bin's blocked layout:
// blocks of B^3 nodes, nb blocks per axis, block-major then node-major,
//   ((i/B nb + j/B) nb + k/B) B^3 + ((i%B) B + j%B) B + k%B.
__host__ __device__ inline int ssim_blocked_key(int i, int j, int k, int B, int nb) {
  const int bi = i / B, bj = j / B, bk = k / B;
  return ((bi * nb + bj) * nb + bk) * (B * B * B) +
         ((i - bi * B) * B + (j - bj * B)) * B + (k - bk * B);
}

// Exclusive prefix of v over a 256-thread workgroup (wave shuffles and one
// LDS exchange of the 4 wave totals, `red` = 4 ints of LDS); *total = the
// workgroup's sum.  Every thread of the workgroup must call it.
__device__ __forceinline__ int ssim_block256_excl(int v, int* red, int* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int y = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int z = __shfl_up(y, off, 64);
    if (lane >= off) y += z;
  }
  if (lane == 63) red[w] = y;
  __syncthreads();
  int pre = 0, tot = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int r = red[q];
    pre += q < w ? r : 0;
    tot += r;
  }
  *total = tot;
  __syncthreads();  // red is free again
  return pre + y - v;
}

// The offset of workgroup g among ng workgroup totals bsum[0..ng) (the sum
// of those before it), by all 256 threads; `red` = 4 ints of LDS.
__device__ __forceinline__ int ssim_block256_offset(const int* bsum, int g, int* red) {
  int v = 0;
  for (int i = threadIdx.x; i < g; i += 256) v += bsum[i];
  int tot;
  ssim_block256_excl(v, red, &tot);
  return tot;
}

// Knuth TwoSum: s + err == a + b exactly (6 flops, no ordering needed).
template <typename F>
__device__ __forceinline__ void ssim_two_sum(F a, F b, F& s, F& err) {
  s = a + b;
  const F bb = s - a;
  err = (a - (s - bb)) + (b - bb);
}

// Accumulate b into the compensated pair (hi, lo).
template <typename F>
__device__ __forceinline__ void ssim_comp_add(F& hi, F& lo, F b) {
  F s, e;
  ssim_two_sum(hi, b, s, e);
  hi = s;
  lo += e;
}
