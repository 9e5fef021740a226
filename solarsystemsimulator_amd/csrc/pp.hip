// Near-field (particle-particle) part of the split particle-mesh force.
//
// The Plummer pair force F(d) = G m d / (|d|^2 + eps2)^(3/2) is split with
// the Gaussian filter of the long-range mesh part,
//   g(r) = erfc(r / 2rs) + r / (rs sqrt(pi)) exp(-r^2 / 4rs^2),
//   F = F g(r)  [near, summed here over r < rcut]  +  F (1 - g(r))  [mesh].
// Sources come from a cell list (cells.hip) with cells no smaller than rcut,
// so the 27 cells around a target's cell hold every source within rcut.
// Terms are FP32 on origin-relative coordinates (error ~1e-7 of each term,
// far below the method's 1e-3 budget); the per-target sum is ADDED to the
// FP64 accelerations.
#include "ssim_internal.h"

namespace {

constexpr int PB = 256;

__global__ __launch_bounds__(PB) void k_pp_short(
    int nt, const double* __restrict__ tx, const double* __restrict__ ty,
    const double* __restrict__ tz, double ox, double oy, double oz,
    double inv_cell, int nc0, int nc1, int nc2,
    const int* __restrict__ cell_start, const float4* __restrict__ src,
    float eps2, float inv_2rs, float c_lin, float rcut2, double G,
    double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az) {
  const int i = blockIdx.x * PB + threadIdx.x;
  if (i >= nt) return;
  const double rx = tx[i] - ox, ry = ty[i] - oy, rz = tz[i] - oz;
  const float px = (float)rx, py = (float)ry, pz = (float)rz;
  const int a = min(max((int)floor(rx * inv_cell), 0), nc0 - 1);
  const int b = min(max((int)floor(ry * inv_cell), 0), nc1 - 1);
  const int c = min(max((int)floor(rz * inv_cell), 0), nc2 - 1);
  float sx = 0.f, sy = 0.f, sz = 0.f;
  for (int da = max(a - 1, 0); da <= min(a + 1, nc0 - 1); ++da)
    for (int db = max(b - 1, 0); db <= min(b + 1, nc1 - 1); ++db) {
      const int row = (da * nc1 + db) * nc2;
      const int c0 = max(c - 1, 0), c1 = min(c + 1, nc2 - 1);
      // the cells c0..c1 of one row are contiguous in the sorted order
      const int s0 = cell_start[row + c0], s1 = cell_start[row + c1 + 1];
      for (int s = s0; s < s1; ++s) {
        const float4 q = src[s];
        const float dx = q.x - px, dy = q.y - py, dz = q.z - pz;
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        if (r2 >= rcut2) continue;
        const float r = sqrtf(r2);
        const float x = r * inv_2rs;
        const float g = erfcf(x) + c_lin * r * __expf(-x * x);
        const float inv = ssim_rsqrt_f32(r2 + eps2);
        const float f = q.w * (inv * inv) * inv * g;
        sx = fmaf(f, dx, sx);
        sy = fmaf(f, dy, sy);
        sz = fmaf(f, dz, sz);
      }
    }
  ax[i] += G * (double)sx;
  ay[i] += G * (double)sy;
  az[i] += G * (double)sz;
}

}  // namespace

extern "C" {

int ssim_pp_short(int nt, const double* tx, const double* ty, const double* tz,
                  const double* origin, double cell, const int* nc,
                  const int* cell_start, const void* packed, double G,
                  double eps2, double rs, double rcut, double* ax, double* ay,
                  double* az, void* stream) {
  SSIM_REQUIRE(nt >= 0, "nt must be >= 0");
  if (nt == 0) return SSIM_OK;
  SSIM_REQUIRE(tx && ty && tz && origin && nc && cell_start && packed && ax &&
                   ay && az,
               "null pointer");
  SSIM_REQUIRE(rs > 0.0 && rcut > 0.0 && rcut <= cell,
               "need rs > 0 and 0 < rcut <= cell (rs=%g rcut=%g cell=%g)", rs,
               rcut, cell);
  SSIM_REQUIRE(eps2 > 0.0, "softening eps2 must be > 0");
  hipStream_t st = ssim::as_stream(stream);
  const float c_lin = (float)(1.0 / (rs * 1.7724538509055160273));
  hipLaunchKernelGGL(k_pp_short, dim3(ssim::ceil_div(nt, PB)), dim3(PB), 0, st,
                     nt, tx, ty, tz, origin[0], origin[1], origin[2],
                     1.0 / cell, nc[0], nc[1], nc[2], cell_start,
                     static_cast<const float4*>(packed), (float)eps2,
                     (float)(0.5 / rs), c_lin, (float)(rcut * rcut), G, ax, ay,
                     az);
  SSIM_LAUNCHED("k_pp_short");
  return SSIM_OK;
}

}  // extern "C"
