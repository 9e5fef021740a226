// All-pairs softened gravity on gfx950.
//
// a_i = G * sum_j m_j (x_j - x_i) / (|x_j - x_i|^2 + eps2)^(3/2)
//
// Follows the tiled sum of `beschleunigung` (selfgrav_kernel.py:222-260):
// every thread owns targets, the workgroup stages sources tile by tile,
// the tile overhang carries zero mass so the trip count is uniform, and the
// self term vanishes because dx = 0.  What is different here:
//  * a thread owns T targets (register blocking) so one LDS broadcast read of
//    a source feeds T interactions;
//  * the source range can be split over gridDim.y chunks when there are too
//    few targets to fill 256 CUs; chunk partials are summed afterwards in a
//    fixed order, so results are bitwise reproducible;
//  * accumulation is blocked: each 256-source tile is summed plainly, the
//    tile sums enter a TwoSum-compensated (hi, lo) total.  The summation
//    error is then ~tile-size * eps instead of ~N * eps;
//  * FP64 terms use v_rsq_f64 + one Householder step (ssim_rsqrt_f64),
//    accurate to ~1 ulp without a divide or a full-precision sqrt;
//  * the FP32 path splits every FP64 coordinate into hi + lo floats and forms
//    dx = (xj_hi - xi_hi) + (xj_lo - xi_lo): the hi difference of nearby
//    bodies is exact (Sterbenz), so dx keeps ~1e-7 relative accuracy even
//    when |dx| << |x|.  Sub-blocks of 32 terms are summed plainly in FP32 and
//    merged into an FP32 TwoSum (hi, lo) pair.
#include "ssim_internal.h"

namespace {

constexpr int AP_BLOCK = 256;  // threads per workgroup = sources per tile
constexpr int AP_T = 2;        // targets per thread
constexpr int AP_SUB32 = 32;   // FP32 plain-sum sub-block

// Number of source chunks for nt targets / ns sources: enough workgroups to
// cover the 256 CUs several times over, but at least 4 tiles per chunk.
int source_chunks(int ns, int nt) {
  const int bx = ssim::ceil_div(nt, AP_BLOCK * AP_T);
  int s = ssim::ceil_div(2048, bx);
  const int max_s = ns / (AP_BLOCK * 4);
  if (s > max_s) s = max_s;
  if (s < 1) s = 1;
  return s;
}

int chunk_len(int ns, int s) {
  // multiple of the tile so only the final chunk has an overhang
  int c = ssim::ceil_div(ns, s);
  return ssim::ceil_div(c, AP_BLOCK) * AP_BLOCK;
}

template <int D>
__device__ __forceinline__ void store_result(int i, int nt, int comp_stride,
                                             const double (&hi)[3],
                                             const double (&lo)[3], double G,
                                             double* ax, double* ay, double* az,
                                             double* part, bool direct) {
  if (direct) {
    ax[i] = G * (hi[0] + lo[0]);
    ay[i] = G * (hi[1] + lo[1]);
    if (D == 3) az[i] = G * (hi[2] + lo[2]);
  } else {
    // part layout [chunk][D][2][nt]
    double* p = part + (size_t)blockIdx.y * comp_stride;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      p[(size_t)(2 * c) * nt + i] = hi[c];
      p[(size_t)(2 * c + 1) * nt + i] = lo[c];
    }
  }
}

template <int D, int T>
__global__ __launch_bounds__(AP_BLOCK) void k_allpairs_f64(
    const double* __restrict__ sx, const double* __restrict__ sy,
    const double* __restrict__ sz, const double* __restrict__ sm, int ns,
    int chunk, const double* __restrict__ tx, const double* __restrict__ ty,
    const double* __restrict__ tz, int nt, double* __restrict__ ax,
    double* __restrict__ ay, double* __restrict__ az,
    double* __restrict__ part, double G, double eps2) {
  __shared__ double4 tile[AP_BLOCK];
  const int base = blockIdx.x * (AP_BLOCK * T) + threadIdx.x;

  double px[T], py[T], pz[T];
  double hi[T][3], lo[T][3];
#pragma unroll
  for (int k = 0; k < T; ++k) {
    const int i = base + k * AP_BLOCK;
    const bool ok = i < nt;
    px[k] = ok ? tx[i] : 0.0;
    py[k] = ok ? ty[i] : 0.0;
    pz[k] = (D == 3 && ok) ? tz[i] : 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      hi[k][c] = 0.0;
      lo[k][c] = 0.0;
    }
  }

  const int s0 = blockIdx.y * chunk;
  const int s1 = min(ns, s0 + chunk);
  for (int j0 = s0; j0 < s1; j0 += AP_BLOCK) {
    const int j = j0 + threadIdx.x;
    double4 s;
    if (j < s1) {
      s.x = sx[j];
      s.y = sy[j];
      s.z = (D == 3) ? sz[j] : 0.0;
      s.w = sm[j];
    } else {  // zero-mass padding contributes exactly nothing
      s.x = 0.0;
      s.y = 0.0;
      s.z = 0.0;
      s.w = 0.0;
    }
    tile[threadIdx.x] = s;
    __syncthreads();

    double acc[T][3];
#pragma unroll
    for (int k = 0; k < T; ++k)
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[k][c] = 0.0;

#pragma unroll 4
    for (int q = 0; q < AP_BLOCK; ++q) {
      const double4 src = tile[q];
#pragma unroll
      for (int k = 0; k < T; ++k) {
        const double dx = src.x - px[k];
        const double dy = src.y - py[k];
        double r2 = fma(dx, dx, eps2);
        r2 = fma(dy, dy, r2);
        double dz = 0.0;
        if (D == 3) {
          dz = src.z - pz[k];
          r2 = fma(dz, dz, r2);
        }
        const double inv = ssim_rsqrt_f64(r2);
        const double inv2 = inv * inv;
        const double f = (src.w * inv) * inv2;
        acc[k][0] = fma(f, dx, acc[k][0]);
        acc[k][1] = fma(f, dy, acc[k][1]);
        if (D == 3) acc[k][2] = fma(f, dz, acc[k][2]);
      }
    }
#pragma unroll
    for (int k = 0; k < T; ++k)
#pragma unroll
      for (int c = 0; c < D; ++c) ssim_comp_add(hi[k][c], lo[k][c], acc[k][c]);
    __syncthreads();
  }

  const bool direct = gridDim.y == 1;
  const int comp_stride = D * 2 * nt;
#pragma unroll
  for (int k = 0; k < T; ++k) {
    const int i = base + k * AP_BLOCK;
    if (i < nt)
      store_result<D>(i, nt, comp_stride, hi[k], lo[k], G, ax, ay, az, part,
                      direct);
  }
}

template <int D, int T>
__global__ __launch_bounds__(AP_BLOCK) void k_allpairs_f32c(
    const double* __restrict__ sx, const double* __restrict__ sy,
    const double* __restrict__ sz, const double* __restrict__ sm, int ns,
    int chunk, const double* __restrict__ tx, const double* __restrict__ ty,
    const double* __restrict__ tz, int nt, double* __restrict__ ax,
    double* __restrict__ ay, double* __restrict__ az,
    double* __restrict__ part, double G, double eps2) {
  __shared__ float4 tile_hi[AP_BLOCK];  // x, y, z, m
  __shared__ float4 tile_lo[AP_BLOCK];  // x, y, z, 0
  const int base = blockIdx.x * (AP_BLOCK * T) + threadIdx.x;
  const float eps2f = (float)eps2;

  float ph[T][3], pl[T][3];
  float hi[T][3], lo[T][3];
#pragma unroll
  for (int k = 0; k < T; ++k) {
    const int i = base + k * AP_BLOCK;
    const bool ok = i < nt;
    const double p[3] = {ok ? tx[i] : 0.0, ok ? ty[i] : 0.0,
                         (D == 3 && ok) ? tz[i] : 0.0};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      ph[k][c] = (float)p[c];
      pl[k][c] = (float)(p[c] - (double)ph[k][c]);
      hi[k][c] = 0.0f;
      lo[k][c] = 0.0f;
    }
  }

  const int s0 = blockIdx.y * chunk;
  const int s1 = min(ns, s0 + chunk);
  for (int j0 = s0; j0 < s1; j0 += AP_BLOCK) {
    const int j = j0 + threadIdx.x;
    float4 h = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 l = make_float4(0.f, 0.f, 0.f, 0.f);
    if (j < s1) {
      const double x = sx[j], y = sy[j], z = (D == 3) ? sz[j] : 0.0;
      h.x = (float)x;
      h.y = (float)y;
      h.z = (float)z;
      h.w = (float)sm[j];
      l.x = (float)(x - (double)h.x);
      l.y = (float)(y - (double)h.y);
      l.z = (float)(z - (double)h.z);
    }
    tile_hi[threadIdx.x] = h;
    tile_lo[threadIdx.x] = l;
    __syncthreads();

    for (int q0 = 0; q0 < AP_BLOCK; q0 += AP_SUB32) {
      float acc[T][3];
#pragma unroll
      for (int k = 0; k < T; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[k][c] = 0.0f;
#pragma unroll 8
      for (int q = q0; q < q0 + AP_SUB32; ++q) {
        const float4 sh = tile_hi[q];
        const float4 sl = tile_lo[q];
#pragma unroll
        for (int k = 0; k < T; ++k) {
          const float dx = (sh.x - ph[k][0]) + (sl.x - pl[k][0]);
          const float dy = (sh.y - ph[k][1]) + (sl.y - pl[k][1]);
          float r2 = fmaf(dx, dx, eps2f);
          r2 = fmaf(dy, dy, r2);
          float dz = 0.0f;
          if (D == 3) {
            dz = (sh.z - ph[k][2]) + (sl.z - pl[k][2]);
            r2 = fmaf(dz, dz, r2);
          }
          const float inv = ssim_rsqrt_f32(r2);
          const float f = (sh.w * inv) * (inv * inv);
          acc[k][0] = fmaf(f, dx, acc[k][0]);
          acc[k][1] = fmaf(f, dy, acc[k][1]);
          if (D == 3) acc[k][2] = fmaf(f, dz, acc[k][2]);
        }
      }
#pragma unroll
      for (int k = 0; k < T; ++k)
#pragma unroll
        for (int c = 0; c < D; ++c) ssim_comp_add(hi[k][c], lo[k][c], acc[k][c]);
    }
    __syncthreads();
  }

  const bool direct = gridDim.y == 1;
  const int comp_stride = D * 2 * nt;
#pragma unroll
  for (int k = 0; k < T; ++k) {
    const int i = base + k * AP_BLOCK;
    if (i < nt) {
      double h64[3], l64[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        h64[c] = (double)hi[k][c];
        l64[c] = (double)lo[k][c];
      }
      store_result<D>(i, nt, comp_stride, h64, l64, G, ax, ay, az, part,
                      direct);
    }
  }
}

// Sum chunk partials [S][D][2][nt] in chunk order (double-double).
template <int D>
__global__ __launch_bounds__(256) void k_allpairs_reduce(
    const double* __restrict__ part, int S, int nt, double G,
    double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nt) return;
  double* outs[3] = {ax, ay, az};
#pragma unroll
  for (int c = 0; c < D; ++c) {
    double hi = 0.0, lo = 0.0;
    for (int s = 0; s < S; ++s) {
      const double* p = part + (size_t)s * D * 2 * nt + (size_t)(2 * c) * nt;
      ssim_comp_add(hi, lo, p[i]);
      lo += p[nt + i];
    }
    outs[c][i] = G * (hi + lo);
  }
}

template <int D>
int launch_allpairs(int prec, int ns, const double* sx, const double* sy,
                    const double* sz, const double* sm, int nt,
                    const double* tx, const double* ty, const double* tz,
                    double* ax, double* ay, double* az, double G, double eps2,
                    void* work, hipStream_t st) {
  const int S = source_chunks(ns, nt);
  const int chunk = chunk_len(ns, S);
  const int S_eff = ssim::ceil_div(ns, chunk);  // chunks that hold sources
  dim3 grid(ssim::ceil_div(nt, AP_BLOCK * AP_T), S_eff);
  double* part = static_cast<double*>(work);
  if (prec == SSIM_PREC_F64)
    SSIM_TIMED("k_allpairs_f64", st,
               hipLaunchKernelGGL((k_allpairs_f64<D, AP_T>), grid, dim3(AP_BLOCK),
                                  0, st, sx, sy, sz, sm, ns, chunk, tx, ty, tz, nt,
                                  ax, ay, az, part, G, eps2));
  else
    hipLaunchKernelGGL((k_allpairs_f32c<D, AP_T>), grid, dim3(AP_BLOCK), 0, st,
                       sx, sy, sz, sm, ns, chunk, tx, ty, tz, nt, ax, ay, az,
                       part, G, eps2);
  SSIM_LAUNCHED("k_allpairs");
  if (S_eff > 1) {
    hipLaunchKernelGGL((k_allpairs_reduce<D>), dim3(ssim::ceil_div(nt, 256)),
                       dim3(256), 0, st, part, S_eff, nt, G, ax, ay, az);
    SSIM_LAUNCHED("k_allpairs_reduce");
  }
  return SSIM_OK;
}

}  // namespace

namespace ssim {

size_t allpairs_work_bytes(int dim, int ns, int nt) {
  if (ns < 1 || nt < 1) return 0;
  const int S = source_chunks(ns, nt);
  const int S_eff = ceil_div(ns, chunk_len(ns, S));
  if (S_eff <= 1) return 0;
  return align256((size_t)S_eff * dim * 2 * nt * sizeof(double));
}

int allpairs(int dim, int prec, int ns, const double* sx, const double* sy,
             const double* sz, const double* sm, int nt, const double* tx,
             const double* ty, const double* tz, double* ax, double* ay,
             double* az, double G, double eps2, void* work, size_t work_bytes,
             hipStream_t st) {
  SSIM_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3, got %d", dim);
  SSIM_REQUIRE(prec == SSIM_PREC_F64 || prec == SSIM_PREC_F32C,
               "unknown precision %d", prec);
  SSIM_REQUIRE(ns >= 1 && nt >= 1, "need ns >= 1 and nt >= 1 (ns=%d nt=%d)", ns,
               nt);
  SSIM_REQUIRE(eps2 > 0.0, "softening eps2 must be > 0, got %g", eps2);
  SSIM_REQUIRE(sx && sy && sm && tx && ty && ax && ay,
               "null source/target/output pointer");
  SSIM_REQUIRE(dim == 2 || (sz && tz && az), "dim 3 needs z pointers");
  const size_t need = allpairs_work_bytes(dim, ns, nt);
  SSIM_REQUIRE(work_bytes >= need && (need == 0 || work),
               "workspace too small: %zu < %zu bytes", work_bytes, need);
  if (dim == 2)
    return launch_allpairs<2>(prec, ns, sx, sy, sz, sm, nt, tx, ty, tz, ax, ay,
                              az, G, eps2, work, st);
  return launch_allpairs<3>(prec, ns, sx, sy, sz, sm, nt, tx, ty, tz, ax, ay,
                            az, G, eps2, work, st);
}

}  // namespace ssim

extern "C" {

size_t ssim_allpairs_work_bytes(int dim, int prec, int ns, int nt) {
  (void)prec;
  return ssim::allpairs_work_bytes(dim, ns, nt);
}

int ssim_allpairs(int dim, int prec, int ns, const double* sx, const double* sy,
                  const double* sz, const double* sm, int nt, const double* tx,
                  const double* ty, const double* tz, double* ax, double* ay,
                  double* az, double G, double eps2, void* work,
                  size_t work_bytes, void* stream) {
  return ssim::allpairs(dim, prec, ns, sx, sy, sz, sm, nt, tx, ty, tz, ax, ay,
                        az, G, eps2, work, work_bytes, ssim::as_stream(stream));
}

}  // extern "C"
