// Kick / drift kernels and the self-gravity step driver.
//
// The step driver replaces the cooperative `selfgrav_kernel`
// (selfgrav_kernel.py:262-420).  That kernel runs one persistent grid with
// grid-wide barriers between phases because it also couples several GPUs
// through host memory; on one MI355X each phase is an ordinary launch on
// one stream (launch cost ~µs against a force evaluation of ms), and the
// tracers ride along as extra targets of the same all-pairs launch instead
// of a second tile loop: they see the masses at the same new positions the
// masses see, which is exactly the reference's order (masses drift, force,
// then the tracers' full KDK against the new mass positions).
//
// Arithmetic follows the reference line by line and is built without FMA
// contraction: v += a * (0.5*dt), x += v * dt.
#include <cmath>

#include "ssim_internal.h"

namespace {

constexpr int EW_BLOCK = 256;

template <int D>
__global__ __launch_bounds__(EW_BLOCK) void k_kick(
    int n, double* __restrict__ vx, double* __restrict__ vy,
    double* __restrict__ vz, const double* __restrict__ ax,
    const double* __restrict__ ay, const double* __restrict__ az, double h) {
  const int i = blockIdx.x * EW_BLOCK + threadIdx.x;
  if (i >= n) return;
  vx[i] = vx[i] + ax[i] * h;
  vy[i] = vy[i] + ay[i] * h;
  if (D == 3) vz[i] = vz[i] + az[i] * h;
}

template <int D>
__global__ __launch_bounds__(EW_BLOCK) void k_drift(
    int n, double* __restrict__ x, double* __restrict__ y,
    double* __restrict__ z, const double* __restrict__ vx,
    const double* __restrict__ vy, const double* __restrict__ vz, double h) {
  const int i = blockIdx.x * EW_BLOCK + threadIdx.x;
  if (i >= n) return;
  x[i] = x[i] + vx[i] * h;
  y[i] = y[i] + vy[i] * h;
  if (D == 3) z[i] = z[i] + vz[i] * h;
}

// First half of a KDK step, fused: v += a*half; x += v*full.
template <int D>
__global__ __launch_bounds__(EW_BLOCK) void k_kick_drift(
    int n, double* __restrict__ x, double* __restrict__ y,
    double* __restrict__ z, double* __restrict__ vx, double* __restrict__ vy,
    double* __restrict__ vz, const double* __restrict__ ax,
    const double* __restrict__ ay, const double* __restrict__ az, double half,
    double full) {
  const int i = blockIdx.x * EW_BLOCK + threadIdx.x;
  if (i >= n) return;
  const double nvx = vx[i] + ax[i] * half;
  const double nvy = vy[i] + ay[i] * half;
  vx[i] = nvx;
  vy[i] = nvy;
  x[i] = x[i] + nvx * full;
  y[i] = y[i] + nvy * full;
  if (D == 3) {
    const double nvz = vz[i] + az[i] * half;
    vz[i] = nvz;
    z[i] = z[i] + nvz * full;
  }
}

// Second half-kick, optionally followed by the f32 snapshot of the step:
// masses [x|y|(z)|vx|vy|(vz)] (n each), then tracers [tx|ty|(tz)] (t each).
template <int D>
__global__ __launch_bounds__(EW_BLOCK) void k_kick_snap(
    int n, int t, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, double* __restrict__ vx,
    double* __restrict__ vy, double* __restrict__ vz,
    const double* __restrict__ ax, const double* __restrict__ ay,
    const double* __restrict__ az, double half, float* __restrict__ snap) {
  const int i = blockIdx.x * EW_BLOCK + threadIdx.x;
  if (i >= n + t) return;
  const double nvx = vx[i] + ax[i] * half;
  const double nvy = vy[i] + ay[i] * half;
  vx[i] = nvx;
  vy[i] = nvy;
  double nvz = 0.0;
  if (D == 3) {
    nvz = vz[i] + az[i] * half;
    vz[i] = nvz;
  }
  if (!snap) return;
  if (i < n) {
    snap[i] = (float)x[i];
    snap[(size_t)n + i] = (float)y[i];
    if (D == 3) snap[(size_t)2 * n + i] = (float)z[i];
    snap[(size_t)D * n + i] = (float)nvx;
    snap[(size_t)(D + 1) * n + i] = (float)nvy;
    if (D == 3) snap[(size_t)5 * n + i] = (float)nvz;
  } else {
    float* ts = snap + (size_t)2 * D * n;
    const int k = i - n;
    ts[k] = (float)x[i];
    ts[(size_t)t + k] = (float)y[i];
    if (D == 3) ts[(size_t)2 * t + k] = (float)z[i];
  }
}

// Drift in a periodic box: x += v*h, then x -= L floor((x - o) / L) on each
// axis (exactly x while it stays inside [o, o + L)).
template <int D>
__global__ __launch_bounds__(EW_BLOCK) void k_drift_wrap(
    int n, double* __restrict__ x, double* __restrict__ y, double* __restrict__ z,
    const double* __restrict__ vx, const double* __restrict__ vy,
    const double* __restrict__ vz, double h, double o0, double o1, double o2,
    double L) {
  const int i = blockIdx.x * EW_BLOCK + threadIdx.x;
  if (i >= n) return;
  const double nx = x[i] + vx[i] * h;
  const double ny = y[i] + vy[i] * h;
  x[i] = nx - L * floor((nx - o0) / L);
  y[i] = ny - L * floor((ny - o1) / L);
  if (D == 3) {
    const double nz = z[i] + vz[i] * h;
    z[i] = nz - L * floor((nz - o2) / L);
  }
}

__global__ __launch_bounds__(EW_BLOCK) void k_fill_f32(long long n,
                                                       float* __restrict__ d,
                                                       float v) {
  const long long i = (long long)blockIdx.x * EW_BLOCK + threadIdx.x;
  if (i < n) d[i] = v;
}

// dst[idx[k]] = vals[k] (add == 0) or dst[idx[k]] += vals[k]; the caller
// guarantees unique indices when adding.
__global__ __launch_bounds__(EW_BLOCK) void k_scatter_f64(
    int n, const long long* __restrict__ idx, const double* __restrict__ vals,
    double* __restrict__ dst, int add) {
  const int k = blockIdx.x * EW_BLOCK + threadIdx.x;
  if (k >= n) return;
  const long long i = idx[k];
  dst[i] = add ? dst[i] + vals[k] : vals[k];
}

template <int D>
int run_selfgrav(int prec, int scheme, int n, int t, double* x, double* y,
                 double* z, double* vx, double* vy, double* vz, double* ax,
                 double* ay, double* az, const double* m, double G,
                 double eps2, double dt, int steps, int init_acc, float* snap,
                 long long snap_stride, void* work, size_t work_bytes,
                 hipStream_t st) {
  const int nb = n + t;
  const dim3 grid(ssim::ceil_div(nb, EW_BLOCK));
  if (init_acc)
    SSIM_TRY(ssim::allpairs(D, prec, n, x, y, z, m, nb, x, y, z, ax, ay, az, G,
                            eps2, work, work_bytes, st));
  double w[3];
  int nsub;
  if (scheme == SSIM_SCHEME_KDK) {
    w[0] = 1.0;
    nsub = 1;
  } else {
    w[0] = ssim::yoshida_w1();
    w[1] = ssim::yoshida_w0();
    w[2] = ssim::yoshida_w1();
    nsub = 3;
  }
  for (int s = 0; s < steps; ++s) {
    for (int k = 0; k < nsub; ++k) {
      const double full = w[k] * dt;
      const double half = 0.5 * full;
      hipLaunchKernelGGL((k_kick_drift<D>), grid, dim3(EW_BLOCK), 0, st, nb, x,
                         y, z, vx, vy, vz, ax, ay, az, half, full);
      SSIM_LAUNCHED("k_kick_drift");
      SSIM_TRY(ssim::allpairs(D, prec, n, x, y, z, m, nb, x, y, z, ax, ay, az,
                              G, eps2, work, work_bytes, st));
      float* sp = (snap && k == nsub - 1) ? snap + (size_t)s * snap_stride
                                          : nullptr;
      hipLaunchKernelGGL((k_kick_snap<D>), grid, dim3(EW_BLOCK), 0, st, n, t, x,
                         y, z, vx, vy, vz, ax, ay, az, half, sp);
      SSIM_LAUNCHED("k_kick_snap");
    }
  }
  return SSIM_OK;
}

}  // namespace

namespace ssim {

// Same expression as the reference (2 ** (1/3) through libm pow).
double yoshida_w1() { return 1.0 / (2.0 - std::pow(2.0, 1.0 / 3.0)); }
double yoshida_w0() {
  return -std::pow(2.0, 1.0 / 3.0) / (2.0 - std::pow(2.0, 1.0 / 3.0));
}

int kick(int dim, int n, double* vx, double* vy, double* vz, const double* ax,
         const double* ay, const double* az, double h, hipStream_t st) {
  SSIM_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3, got %d", dim);
  SSIM_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return SSIM_OK;
  SSIM_REQUIRE(vx && vy && ax && ay && (dim == 2 || (vz && az)),
               "null pointer");
  const dim3 grid(ceil_div(n, EW_BLOCK));
  if (dim == 2)
    hipLaunchKernelGGL(k_kick<2>, grid, dim3(EW_BLOCK), 0, st, n, vx, vy, vz,
                       ax, ay, az, h);
  else
    hipLaunchKernelGGL(k_kick<3>, grid, dim3(EW_BLOCK), 0, st, n, vx, vy, vz,
                       ax, ay, az, h);
  SSIM_LAUNCHED("k_kick");
  return SSIM_OK;
}

int drift(int dim, int n, double* x, double* y, double* z, const double* vx,
          const double* vy, const double* vz, double h, hipStream_t st) {
  SSIM_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3, got %d", dim);
  SSIM_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return SSIM_OK;
  SSIM_REQUIRE(x && y && vx && vy && (dim == 2 || (z && vz)), "null pointer");
  const dim3 grid(ceil_div(n, EW_BLOCK));
  if (dim == 2)
    hipLaunchKernelGGL(k_drift<2>, grid, dim3(EW_BLOCK), 0, st, n, x, y, z, vx,
                       vy, vz, h);
  else
    hipLaunchKernelGGL(k_drift<3>, grid, dim3(EW_BLOCK), 0, st, n, x, y, z, vx,
                       vy, vz, h);
  SSIM_LAUNCHED("k_drift");
  return SSIM_OK;
}

}  // namespace ssim

extern "C" {

size_t ssim_selfgrav_work_bytes(int dim, int prec, int n, int t) {
  (void)prec;
  return ssim::allpairs_work_bytes(dim, n, n + t);
}

int ssim_selfgrav_run(int dim, int prec, int scheme, int n, int t, double* x,
                      double* y, double* z, double* vx, double* vy, double* vz,
                      double* ax, double* ay, double* az, const double* m,
                      double G, double eps2, double dt, int steps, int init_acc,
                      float* snap, long long snap_stride, void* work,
                      size_t work_bytes, void* stream) {
  SSIM_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3, got %d", dim);
  SSIM_REQUIRE(scheme == SSIM_SCHEME_KDK || scheme == SSIM_SCHEME_YOSHIDA4,
               "unknown scheme %d", scheme);
  SSIM_REQUIRE(n >= 1, "empty state (n=%d)", n);
  SSIM_REQUIRE(t >= 0, "tracer count must be >= 0");
  SSIM_REQUIRE(steps >= 1, "steps must be >= 1, got %d", steps);
  SSIM_REQUIRE(eps2 > 0.0, "softening eps2 must be > 0, got %g", eps2);
  SSIM_REQUIRE(x && y && vx && vy && ax && ay && m, "null state pointer");
  SSIM_REQUIRE(dim == 2 || (z && vz && az), "dim 3 needs z pointers");
  SSIM_REQUIRE(!snap || snap_stride >= (long long)2 * dim * n + (long long)dim * t,
               "snapshot stride %lld too small", snap_stride);
  hipStream_t st = ssim::as_stream(stream);
  if (dim == 2)
    return run_selfgrav<2>(prec, scheme, n, t, x, y, z, vx, vy, vz, ax, ay, az,
                           m, G, eps2, dt, steps, init_acc, snap, snap_stride,
                           work, work_bytes, st);
  return run_selfgrav<3>(prec, scheme, n, t, x, y, z, vx, vy, vz, ax, ay, az, m,
                         G, eps2, dt, steps, init_acc, snap, snap_stride, work,
                         work_bytes, st);
}

int ssim_fill_f32(long long n, float* dst, float value, void* stream) {
  SSIM_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return SSIM_OK;
  SSIM_REQUIRE(dst, "null pointer");
  hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)ssim::ceil_div(n, EW_BLOCK)),
                     dim3(EW_BLOCK), 0, ssim::as_stream(stream), n, dst, value);
  SSIM_LAUNCHED("k_fill_f32");
  return SSIM_OK;
}

int ssim_scatter_f64(int n, const long long* idx, const double* vals,
                     double* dst, int add, void* stream) {
  SSIM_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return SSIM_OK;
  SSIM_REQUIRE(idx && vals && dst, "null pointer");
  hipLaunchKernelGGL(k_scatter_f64, dim3(ssim::ceil_div(n, EW_BLOCK)),
                     dim3(EW_BLOCK), 0, ssim::as_stream(stream), n, idx, vals,
                     dst, add);
  SSIM_LAUNCHED("k_scatter_f64");
  return SSIM_OK;
}

int ssim_kick_snap(int dim, int n, int t, const double* x, const double* y,
                   const double* z, double* vx, double* vy, double* vz,
                   const double* ax, const double* ay, const double* az,
                   double h, float* snap, void* stream) {
  SSIM_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3, got %d", dim);
  SSIM_REQUIRE(n >= 0 && t >= 0, "negative size");
  if (n + t == 0) return SSIM_OK;
  SSIM_REQUIRE(x && y && vx && vy && ax && ay && (dim == 2 || (z && vz && az)),
               "null pointer");
  hipStream_t st = ssim::as_stream(stream);
  const dim3 grid(ssim::ceil_div(n + t, EW_BLOCK));
  if (dim == 2)
    hipLaunchKernelGGL(k_kick_snap<2>, grid, dim3(EW_BLOCK), 0, st, n, t, x, y,
                       z, vx, vy, vz, ax, ay, az, h, snap);
  else
    hipLaunchKernelGGL(k_kick_snap<3>, grid, dim3(EW_BLOCK), 0, st, n, t, x, y,
                       z, vx, vy, vz, ax, ay, az, h, snap);
  SSIM_LAUNCHED("k_kick_snap");
  return SSIM_OK;
}

int ssim_kick(int dim, int n, double* vx, double* vy, double* vz,
              const double* ax, const double* ay, const double* az, double h,
              void* stream) {
  return ssim::kick(dim, n, vx, vy, vz, ax, ay, az, h, ssim::as_stream(stream));
}

int ssim_drift(int dim, int n, double* x, double* y, double* z,
               const double* vx, const double* vy, const double* vz, double h,
               void* stream) {
  return ssim::drift(dim, n, x, y, z, vx, vy, vz, h, ssim::as_stream(stream));
}

int ssim_drift_periodic(int dim, int n, double* x, double* y, double* z,
                        const double* vx, const double* vy, const double* vz, double h,
                        const double* origin, double box, void* stream) {
  SSIM_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3, got %d", dim);
  SSIM_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return SSIM_OK;
  SSIM_REQUIRE(x && y && vx && vy && origin && (dim == 2 || (z && vz)), "null pointer");
  SSIM_REQUIRE(box > 0.0, "box edge must be > 0, got %g", box);
  hipStream_t st = ssim::as_stream(stream);
  const dim3 grid(ssim::ceil_div(n, EW_BLOCK));
  const double o2 = dim == 3 ? origin[2] : 0.0;
  if (dim == 2)
    hipLaunchKernelGGL(k_drift_wrap<2>, grid, dim3(EW_BLOCK), 0, st, n, x, y, z, vx,
                       vy, vz, h, origin[0], origin[1], o2, box);
  else
    hipLaunchKernelGGL(k_drift_wrap<3>, grid, dim3(EW_BLOCK), 0, st, n, x, y, z, vx,
                       vy, vz, h, origin[0], origin[1], o2, box);
  SSIM_LAUNCHED("k_drift_wrap");
  return SSIM_OK;
}

int ssim_kick_drift(int dim, int n, double* x, double* y, double* z, double* vx,
                    double* vy, double* vz, const double* ax, const double* ay,
                    const double* az, double half, double full, void* stream) {
  SSIM_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3, got %d", dim);
  SSIM_REQUIRE(n >= 0, "negative size");
  if (n == 0) return SSIM_OK;
  SSIM_REQUIRE(x && y && vx && vy && ax && ay && (dim == 2 || (z && vz && az)),
               "null pointer");
  hipStream_t st = ssim::as_stream(stream);
  const dim3 grid(ssim::ceil_div(n, EW_BLOCK));
  if (dim == 2)
    hipLaunchKernelGGL(k_kick_drift<2>, grid, dim3(EW_BLOCK), 0, st, n, x, y, z, vx, vy,
                       vz, ax, ay, az, half, full);
  else
    hipLaunchKernelGGL(k_kick_drift<3>, grid, dim3(EW_BLOCK), 0, st, n, x, y, z, vx, vy,
                       vz, ax, ay, az, half, full);
  SSIM_LAUNCHED("k_kick_drift");
  return SSIM_OK;
}

}  // extern "C"
