// FP64 energy diagnostic with compensated, fixed-order accumulation.
//
// E_kin = 1/2 sum_i m_i |v_i|^2
// E_pot = -1/2 G sum_i m_i sum_{j != i} m_j / sqrt(|x_j - x_i|^2 + eps2)
//
// This is the reference's host-side `energie` (test_selfgrav.py:54-63): the
// same Plummer potential with the i == j term excluded (the reference adds
// the diagonal back; here it is never added, the weight of j == i is 0).
// The pair loop reuses the tiling of the force kernel.  Each workgroup
// writes (kin_hi, kin_lo, pot_hi, pot_lo) and one final workgroup sums the
// partials in block order, so the result does not depend on scheduling.
#include "ssim_internal.h"

#include <mutex>

namespace {

constexpr int EN_BLOCK = 256;

template <int D>
__global__ __launch_bounds__(EN_BLOCK) void k_energy_partial(
    int n, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ vx,
    const double* __restrict__ vy, const double* __restrict__ vz,
    const double* __restrict__ m, double eps2, double* __restrict__ part) {
  __shared__ double4 tile[EN_BLOCK];
  __shared__ double red[4][EN_BLOCK];
  const int i = blockIdx.x * EN_BLOCK + threadIdx.x;
  const bool ok = i < n;
  const double px = ok ? x[i] : 0.0;
  const double py = ok ? y[i] : 0.0;
  const double pz = (D == 3 && ok) ? z[i] : 0.0;

  double phi_hi = 0.0, phi_lo = 0.0;
  for (int j0 = 0; j0 < n; j0 += EN_BLOCK) {
    const int j = j0 + threadIdx.x;
    double4 s = make_double4(0.0, 0.0, 0.0, 0.0);
    if (j < n) {
      s.x = x[j];
      s.y = y[j];
      s.z = (D == 3) ? z[j] : 0.0;
      s.w = m[j];
    }
    tile[threadIdx.x] = s;
    __syncthreads();
    double acc = 0.0;
#pragma unroll 4
    for (int q = 0; q < EN_BLOCK; ++q) {
      const double4 src = tile[q];
      const double dx = src.x - px;
      const double dy = src.y - py;
      double r2 = fma(dx, dx, eps2);
      r2 = fma(dy, dy, r2);
      if (D == 3) {
        const double dz = src.z - pz;
        r2 = fma(dz, dz, r2);
      }
      const double wgt = (j0 + q == i) ? 0.0 : src.w;
      acc = fma(wgt, ssim_rsqrt_f64(r2), acc);
    }
    ssim_comp_add(phi_hi, phi_lo, acc);
    __syncthreads();
  }

  double kin = 0.0, pot_hi = 0.0, pot_lo = 0.0;
  if (ok) {
    const double mi = m[i];
    double v2 = vx[i] * vx[i] + vy[i] * vy[i];
    if (D == 3) v2 = v2 + vz[i] * vz[i];
    kin = 0.5 * mi * v2;
    // m_i * (phi_hi + phi_lo) as a double-double product
    const double p = mi * phi_hi;
    pot_lo = fma(mi, phi_hi, -p) + mi * phi_lo;
    pot_hi = p;
  }
  // fixed-order tree reduction of (kin, 0) and (pot_hi, pot_lo)
  red[0][threadIdx.x] = kin;
  red[1][threadIdx.x] = 0.0;
  red[2][threadIdx.x] = pot_hi;
  red[3][threadIdx.x] = pot_lo;
  __syncthreads();
  for (int w = EN_BLOCK / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      const int o = threadIdx.x + w;
      for (int c = 0; c < 4; c += 2) {
        double s, e;
        ssim_two_sum(red[c][threadIdx.x], red[c][o], s, e);
        red[c][threadIdx.x] = s;
        red[c + 1][threadIdx.x] += red[c + 1][o] + e;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) part[(size_t)blockIdx.x * 4 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(64) void k_energy_final(int nblk, double G,
                                                      const double* __restrict__ part,
                                                      double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  double kh = 0.0, kl = 0.0, ph = 0.0, pl = 0.0;
  for (int b = 0; b < nblk; ++b) {
    ssim_comp_add(kh, kl, part[4 * b + 0]);
    kl += part[4 * b + 1];
    ssim_comp_add(ph, pl, part[4 * b + 2]);
    pl += part[4 * b + 3];
  }
  const double ekin = kh + kl;
  const double epot = -0.5 * G * (ph + pl);
  out[0] = ekin;
  out[1] = epot;
  out[2] = ekin + epot;
}

}  // namespace

extern "C" {

size_t ssim_energy_work_bytes(int dim, int n) {
  (void)dim;
  if (n < 1) return 0;
  return ssim::align256((size_t)ssim::ceil_div(n, EN_BLOCK) * 4 * sizeof(double));
}

}  // extern "C"

namespace {

int energy_args_ok(int dim, int n, const double* x, const double* y, const double* z,
                   const double* vx, const double* vy, const double* vz, const double* m,
                   double eps2, const double* out) {
  SSIM_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3, got %d", dim);
  SSIM_REQUIRE(n >= 1, "empty state (n=%d)", n);
  SSIM_REQUIRE(eps2 > 0.0, "softening eps2 must be > 0, got %g", eps2);
  SSIM_REQUIRE(x && y && vx && vy && m && out, "null pointer");
  SSIM_REQUIRE(dim == 2 || (z && vz), "dim 3 needs z pointers");
  return SSIM_OK;
}

// The two launches of one energy evaluation; part holds
// ceil(n / EN_BLOCK) * 4 doubles, every one of them written by the first
// launch before the second reads it (no initial contents needed).
int energy_launch(int dim, int n, const double* x, const double* y, const double* z,
                  const double* vx, const double* vy, const double* vz, const double* m,
                  double G, double eps2, double* out, double* part, hipStream_t st) {
  const int nblk = ssim::ceil_div(n, EN_BLOCK);
  if (dim == 2)
    hipLaunchKernelGGL(k_energy_partial<2>, dim3(nblk), dim3(EN_BLOCK), 0, st,
                       n, x, y, z, vx, vy, vz, m, eps2, part);
  else
    hipLaunchKernelGGL(k_energy_partial<3>, dim3(nblk), dim3(EN_BLOCK), 0, st,
                       n, x, y, z, vx, vy, vz, m, eps2, part);
  SSIM_LAUNCHED("k_energy_partial");
  hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(64), 0, st, nblk, G, part,
                     out);
  SSIM_LAUNCHED("k_energy_final");
  return SSIM_OK;
}

// Scratch of ssim_energy_noalloc: the per-workgroup partials of up to
// SSIM_ENERGY_NOALLOC_MAX_N bodies (2 MiB), one copy per GPU for the life of the
// process.  Host code never passes the array's name to a kernel (on the
// host it is only a stand-in for the symbol): the device address comes from
// hipGetSymbolAddress, once per device, and is kept in g_scratch_dev.
constexpr int NOALLOC_MAX_BLK = SSIM_ENERGY_NOALLOC_MAX_N / EN_BLOCK;
__device__ __align__(256) double scratch[NOALLOC_MAX_BLK * 4];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[MAX_DEVICES] = {};

// The device address of this GPU's scratch array (looked up on the first
// call on each device, with that device current).
int scratch_for_current_device(double** out) {
  int dev = -1;
  SSIM_HIP(hipGetDevice(&dev));
  SSIM_REQUIRE(dev >= 0 && dev < MAX_DEVICES, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    SSIM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    SSIM_REQUIRE(p != nullptr, "no device address for the energy scratch array");
    g_scratch_dev[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_dev[dev];
  return SSIM_OK;
}

}  // namespace

extern "C" {

int ssim_energy(int dim, int n, const double* x, const double* y,
                const double* z, const double* vx, const double* vy,
                const double* vz, const double* m, double G, double eps2,
                double* out, void* work, size_t work_bytes, void* stream) {
  SSIM_TRY(energy_args_ok(dim, n, x, y, z, vx, vy, vz, m, eps2, out));
  const size_t need = ssim_energy_work_bytes(dim, n);
  SSIM_REQUIRE(work && work_bytes >= need, "workspace too small: %zu < %zu",
               work_bytes, need);
  return energy_launch(dim, n, x, y, z, vx, vy, vz, m, G, eps2, out,
                       static_cast<double*>(work), ssim::as_stream(stream));
}

int ssim_energy_noalloc(int dim, int n, const double* x, const double* y,
                        const double* z, const double* vx, const double* vy,
                        const double* vz, const double* m, double G, double eps2,
                        double* out, void* stream) {
  SSIM_TRY(energy_args_ok(dim, n, x, y, z, vx, vy, vz, m, eps2, out));
  SSIM_REQUIRE(n <= SSIM_ENERGY_NOALLOC_MAX_N,
               "ssim_energy_noalloc: n = %d exceeds its scratch limit of %d bodies "
               "(use ssim_energy with a workspace)", n, SSIM_ENERGY_NOALLOC_MAX_N);
  double* part = nullptr;
  SSIM_TRY(scratch_for_current_device(&part));
  return energy_launch(dim, n, x, y, z, vx, vy, vz, m, G, eps2, out, part,
                       ssim::as_stream(stream));
}

}  // extern "C"
