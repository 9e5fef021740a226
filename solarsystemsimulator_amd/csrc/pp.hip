// Near-field (particle-particle) part of the split particle-mesh force.
//
// The Plummer pair force F(d) = G m d / (|d|^2 + eps2)^(3/2) is split with
// the Gaussian filter of the long-range mesh part, tapered to zero at rcut
// (ssim_split_g),
//   g(r) = erfc(r / 2rs) + r / (rs sqrt(pi)) exp(-r^2 / 4rs^2),
//   F = F gS(r)  [near, summed here over r < rcut]  +  F (1 - gS(r))  [mesh].
//
// Sources and targets are binned on the same grid of cells no smaller than
// rcut (cells.hip), so the 27 cells around a target's cell hold every source
// within rcut.  Work is cut into chunks of <= 64 targets of one cell; one
// wave takes a chunk, and all its lanes walk the same source list (the
// three contiguous cells of each of the 9 neighbour rows, in the sources'
// sorted order), so the source loads are wave-uniform (scalar loads) and
// the loop has no divergence.  A source that lies beyond rcut of every
// target of the chunk is skipped after the distance test (a wave-uniform
// branch); the targets of a cell come grouped by octant (ssim_cells_bin),
// so a cell's partial last chunk is compact and skips most of the
// 27-cell block.  Each target's sum is made by one lane in the fixed source
// order, so results do not depend on scheduling.
//
// Terms are FP32 on origin-relative coordinates (error ~1e-7 of each term,
// far below the method's 1e-3 budget); the per-target sum is ADDED to the
// FP64 accelerations.  gS(r) comes from a table over u = r^2 in [0, rcut^2]
// of GT + 1 values (computed in FP64 per call) in LDS, linearly
// interpolated in u: at 4096 intervals the error is below 7e-6 of the pair
// force (largest at r ~ rs/20, where g - 1 ~ r^3 is least smooth in u) and
// ~1e-7 elsewhere, and a tested pair then costs one table read and one
// hardware reciprocal square root (v_rsq_f32, 1 ulp) instead of erfc, exp
// and two square roots.  The hi/lo near field (ssim_pp_near_hilo) does the
// pair arithmetic in FP64 on positions given as FP32 hi + lo pairs and
// interpolates the same table quadratically (table entries GT + 1 and
// GT + 2 are zero).
#include "ssim_internal.h"

namespace {

constexpr int PB = 512;            // 8 waves share one LDS table
constexpr int WAVES = PB / 64;
constexpr int CHUNK = 64;          // targets per wave
constexpr int SG = 4;              // sources loaded and tested per group
constexpr int GT = 4096;           // intervals of the gS(sqrt(u)) table

// tab[i] = (gS(r_i), gS(r_{i+1}) - gS(r_i)), r_i = rcut sqrt(i / gt), for
// i = 0..gt + 1; tab[gt] = tab[gt + 1] = (0, 0): gS vanishes from rcut on
__global__ __launch_bounds__(PB) void k_gtab(double rs, double rcut, int gt,
                                             float2* __restrict__ tab) {
  const int i = blockIdx.x * PB + threadIdx.x;
  if (i > gt + 1) return;
  const double r0 = rcut * sqrt((double)i / gt), r1 = rcut * sqrt((double)(i + 1) / gt);
  const double g0 = ssim_split_g(r0, rs, rcut);
  tab[i] = make_float2((float)g0, (float)(ssim_split_g(r1, rs, rcut) - g0));
}

// The table of the filtered near field, in s = sqrt(r^2 + eps2) rather
// than u = r^2: gS is smooth in s down to r = 0 (in u it goes as 1 - c u^1.5,
// whose derivatives diverge at 0, so u-tables lose most accuracy on the
// closest pairs), and s = (r^2 + eps2) (r^2 + eps2)^(-1/2) costs one multiply
// next to the reciprocal square root the pair term needs anyway.
// tab[i] = (gS(r_i), gS(r_{i+1}) - gS(r_i)) at s_i = eps + i (s_max - eps) /
// gt, r_i = sqrt(s_i^2 - eps2), s_max = sqrt(rcut^2 + eps2), i = 0..gt + 1;
// tab[gt] = tab[gt + 1] = (0, 0).
__global__ __launch_bounds__(PB) void k_gtab_s(double rs, double rcut, double eps2, int gt,
                                               float2* __restrict__ tab) {
  const int i = blockIdx.x * PB + threadIdx.x;
  if (i > gt + 1) return;
  const double e = sqrt(eps2), smax = sqrt(rcut * rcut + eps2);
  const double ds = (smax - e) / gt;
  const double s0 = e + i * ds, s1 = e + (i + 1) * ds;
  const double r0 = sqrt(fmax(s0 * s0 - eps2, 0.0)), r1 = sqrt(fmax(s1 * s1 - eps2, 0.0));
  const double g0 = i >= gt ? 0.0 : ssim_split_g(r0, rs, rcut);
  const double g1 = i + 1 >= gt ? 0.0 : ssim_split_g(r1, rs, rcut);
  tab[i] = make_float2((float)g0, (float)(g1 - g0));
}

// The chunk plan by two fully parallel launches (ncells <= 1024 x 1024):
// k_chunk_sum writes each cell's chunk count and each 256-thread
// workgroup's total over its 1024 cells, k_chunk_scan adds the totals of
// the workgroups before its own to its scan and writes cstart and the chunk
// -> cell map.  Same plan as k_chunk_plan.
constexpr int CP_T = 256, CP_BLOCKS = 1024;

__global__ __launch_bounds__(CP_T) void k_chunk_sum(int ncells,
                                                    const int* __restrict__ tstart,
                                                    int* __restrict__ nchunk,
                                                    int* __restrict__ bsum) {
  __shared__ int red[4];
  const int c0 = (blockIdx.x * CP_T + threadIdx.x) * 4;
  int v = 0;
#pragma unroll
  for (int e = 0; e < 4; ++e)
    if (c0 + e < ncells) {
      const int k = (tstart[c0 + e + 1] - tstart[c0 + e] + CHUNK - 1) / CHUNK;
      nchunk[c0 + e] = k;
      v += k;
    }
  int tot;
  ssim_block256_excl(v, red, &tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(CP_T) void k_chunk_scan(int ncells,
                                                     const int* __restrict__ nchunk,
                                                     const int* __restrict__ bsum,
                                                     int* __restrict__ cstart,
                                                     int* __restrict__ map) {
  __shared__ int red[4];
  const int off = ssim_block256_offset(bsum, blockIdx.x, red);
  const int c0 = (blockIdx.x * CP_T + threadIdx.x) * 4;
  int v[4], s = 0;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    v[e] = c0 + e < ncells ? nchunk[c0 + e] : 0;
    s += v[e];
  }
  int tot;
  int run = off + ssim_block256_excl(s, red, &tot);
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    if (c0 + e < ncells) {
      cstart[c0 + e] = run;
      for (int k = 0; k < v[e]; ++k) map[run + k] = c0 + e;
    }
    run += v[e];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) cstart[ncells] = off + tot;
}

__global__ __launch_bounds__(PB) void k_chunk_count(int ncells,
                                                    const int* __restrict__ tstart,
                                                    int* __restrict__ nchunk) {
  const int c = blockIdx.x * PB + threadIdx.x;
  if (c >= ncells) return;
  nchunk[c] = (tstart[c + 1] - tstart[c] + CHUNK - 1) / CHUNK;
}

__global__ __launch_bounds__(PB) void k_chunk_map(int ncells,
                                                  const int* __restrict__ cstart,
                                                  int* __restrict__ map) {
  const int c = blockIdx.x * PB + threadIdx.x;
  if (c >= ncells) return;
  for (int k = cstart[c]; k < cstart[c + 1]; ++k) map[k] = c;
}

// s += m d gS(r) / (r^2 + eps2)^(3/2); gS from the LDS table, linear in
// u = r^2.  t = u GT / rcut^2 is clamped to GT, whose entry is (0, 0), so a
// source at or beyond rcut adds exactly nothing without a select, and the
// interpolation weight t - floor(t) is one v_fract_f32.
__device__ __forceinline__ void pp_term(float m, float dx, float dy, float dz, float r2,
                                        const float2* tab, float tab_scale, float eps2,
                                        float& sx, float& sy, float& sz) {
  const float t = fminf(r2 * tab_scale, (float)GT);
  const float2 e = tab[(int)t];
  const float g = fmaf(__builtin_amdgcn_fractf(t), e.y, e.x);
  const float inv = __builtin_amdgcn_rsqf(r2 + eps2);
  const float f = m * (inv * inv) * inv * g;
  sx = fmaf(f, dx, sx);
  sy = fmaf(f, dy, sy);
  sz = fmaf(f, dz, sz);
}

// The term in FP64 for the hi/lo kernel: d and r^2 in FP64, gS from the
// table in FP32 but interpolated quadratically through the nodes i, i + 1,
// i + 2 (g_i + t d_i + t (t - 1) / 2 (d_{i+1} - d_i), d = the stored
// differences: the table's error drops from ~1e-6 to ~1e-8 of g), and
// (r^2 + eps2)^(-1/2) from the FP32 estimate and one Newton step in FP64
// (relative error ~1e-13).  With packed FP32 off, an FP64 FMA issues at the
// rate of an FP32 one.
__device__ __forceinline__ void pp_term_d(float m, double dx, double dy, double dz,
                                          double r2, const float2* tab, float tab_scale,
                                          double eps2, double& sx, double& sy, double& sz) {
  const float t = fminf((float)r2 * tab_scale, (float)GT);
  const int i = (int)t;
  const float2 e0 = tab[i], e1 = tab[i + 1];
  const float fr = __builtin_amdgcn_fractf(t);
  const float c = fmaf((fr - 1.0f) * (e1.y - e0.y), 0.5f, e0.y);
  const float g = fmaf(fr, c, e0.x);
  const double x = r2 + eps2;
  double y = (double)__builtin_amdgcn_rsqf((float)x);
  const double res = fma(-(x * y), y, 1.0);
  y = fma(0.5 * y, res, y);
  const double f = (double)m * (double)g * (y * y * y);
  sx = fma(f, dx, sx);
  sy = fma(f, dy, sy);
  sz = fma(f, dz, sz);
}

// COUNT: instead of forces, count the pairs tested and the pairs inside
// rcut (excluding a body with itself), for the work accounting of a run.
template <bool COUNT>
__global__ __launch_bounds__(PB) void k_pp_cells(
    const double* __restrict__ tx, const double* __restrict__ ty,
    const double* __restrict__ tz, double ox, double oy, double oz, int nc0,
    int nc1, int nc2, const int* __restrict__ cell_start,
    const float4* __restrict__ src, const int* __restrict__ tstart,
    const int* __restrict__ torder, const int* __restrict__ cstart,
    const int* __restrict__ cmap, int ncells, float eps2,
    const float2* __restrict__ gtab, float tab_scale, float rcut2, double G,
    double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az,
    int accumulate, unsigned long long* __restrict__ counts) {
  // the grid is sized for the most chunks possible: a workgroup past the
  // last chunk leaves before it loads the table (uniform per workgroup)
  if (blockIdx.x * WAVES >= cstart[ncells]) return;
  __shared__ float2 tab[GT + 2];
  for (int i = threadIdx.x; i <= GT + 1; i += PB) tab[i] = gtab[i];
  __syncthreads();  // before any wave leaves
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(blockIdx.x * WAVES + (threadIdx.x >> 6));
  if (w >= cstart[ncells]) return;
  const int cell = __builtin_amdgcn_readfirstlane(cmap[w]);
  const int first = __builtin_amdgcn_readfirstlane(
      tstart[cell] + (w - cstart[cell]) * CHUNK);
  const int count = __builtin_amdgcn_readfirstlane(min(CHUNK, tstart[cell + 1] - first));
  const bool live = lane < count;
  const int tgt = live ? torder[first + lane] : 0;
  // idle lanes sit far outside the grid, so no source is ever within rcut
  const float px = live ? (float)(tx[tgt] - ox) : 1e18f;
  const float py = live ? (float)(ty[tgt] - oy) : 1e18f;
  const float pz = live ? (float)(tz[tgt] - oz) : 1e18f;
  const int a = cell / (nc1 * nc2), b = (cell / nc2) % nc1, c = cell % nc2;
  const int c0 = max(c - 1, 0), c1 = min(c + 1, nc2 - 1);
  float sx = 0.f, sy = 0.f, sz = 0.f;
  unsigned int n_test = 0, n_in = 0;
  for (int da = max(a - 1, 0); da <= min(a + 1, nc0 - 1); ++da)
    for (int db = max(b - 1, 0); db <= min(b + 1, nc1 - 1); ++db) {
      const int row = (da * nc1 + db) * nc2;
      const int s0 = cell_start[row + c0], s1 = cell_start[row + c1 + 1];
      // groups of SG sources: their (scalar) loads are issued together, and
      // a group beyond rcut of every lane adds exactly zero and is skipped
      int s = s0;
      for (; s + SG <= s1; s += SG) {
        float4 q[SG];
#pragma unroll
        for (int j = 0; j < SG; ++j) q[j] = src[s + j];
        float dx[SG], dy[SG], dz[SG], r2[SG];
        unsigned long long any = 0;  // lanes with a source of the group inside
#pragma unroll
        for (int j = 0; j < SG; ++j) {
          dx[j] = q[j].x - px;
          dy[j] = q[j].y - py;
          dz[j] = q[j].z - pz;
          r2[j] = fmaf(dx[j], dx[j], fmaf(dy[j], dy[j], dz[j] * dz[j]));
          any |= __ballot(r2[j] < rcut2);
        }
        if (COUNT) {
#pragma unroll
          for (int j = 0; j < SG; ++j) {
            n_test += 1;
            n_in += (r2[j] < rcut2 && r2[j] > 0.f) ? 1 : 0;
          }
          continue;
        }
        if (any == 0) continue;
#pragma unroll
        for (int j = 0; j < SG; ++j)
          pp_term(q[j].w, dx[j], dy[j], dz[j], r2[j], tab, tab_scale, eps2, sx, sy, sz);
      }
      for (; s < s1; ++s) {
        const float4 q = src[s];
        const float dx = q.x - px, dy = q.y - py, dz = q.z - pz;
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        const bool in = r2 < rcut2;
        if (COUNT) {
          n_test += 1;
          n_in += (in && r2 > 0.f) ? 1 : 0;
          continue;
        }
        if (in) pp_term(q.w, dx, dy, dz, r2, tab, tab_scale, eps2, sx, sy, sz);
      }
    }
  if (COUNT) {
    if (live) {
      atomicAdd(&counts[0], (unsigned long long)n_test);
      atomicAdd(&counts[1], (unsigned long long)n_in);
    }
    return;
  }
  if (live) {
    if (accumulate) {
      ax[tgt] += G * (double)sx;
      ay[tgt] += G * (double)sy;
      az[tgt] += G * (double)sz;
    } else {
      ax[tgt] = G * (double)sx;
      ay[tgt] = G * (double)sy;
      az[tgt] = G * (double)sz;
    }
  }
}


// The near field on block-relative coordinates, for masses binned by
// ssim_split_bin (which also writes `near`: each mass relative to the
// corner of its own block, i.e. its cell).  Same chunks, cell walk and
// wave-uniform source groups as k_pp_cells, but the walk goes cell by cell:
// for cell k of the 27 the target is re-expressed relative to that cell's
// corner (t_k = t_0 - offset_k, exact to ~1e-9 of the unit), so both
// operands of every difference are < 2 cells and keep their low bits
// (origin-relative FP32 coordinates lose ~6e-8 absolute, ~1e-5 of a pair
// force at r ~ rcut/2).  The FP32 terms of one cell are summed in FP32 and
// each cell's partial is added to FP64 sums (blocked compensated
// summation).  Fixed source order per target: deterministic.
template <bool HILO>
__global__ __launch_bounds__(PB) void k_pp_near(
    const double* __restrict__ tx, const double* __restrict__ ty,
    const double* __restrict__ tz, double ox, double oy, double oz, double cellsz,
    int nc0, int nc1, int nc2, const int* __restrict__ cell_start,
    const float4* __restrict__ near, const float4* __restrict__ near_lo,
    const int* __restrict__ tstart,
    const int* __restrict__ torder, const int* __restrict__ cstart,
    const int* __restrict__ cmap, int ncells, double eps2,
    const float2* __restrict__ gtab, float tab_scale, float rcut2, double G,
    double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az,
    int accumulate) {
  if (blockIdx.x * WAVES >= cstart[ncells]) return;
  __shared__ float2 tab[GT + 2];
  for (int i = threadIdx.x; i <= GT + 1; i += PB) tab[i] = gtab[i];
  __syncthreads();  // before any wave leaves
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(blockIdx.x * WAVES + (threadIdx.x >> 6));
  if (w >= cstart[ncells]) return;
  const int cell = __builtin_amdgcn_readfirstlane(cmap[w]);
  const int first = __builtin_amdgcn_readfirstlane(
      tstart[cell] + (w - cstart[cell]) * CHUNK);
  const int count = __builtin_amdgcn_readfirstlane(min(CHUNK, tstart[cell + 1] - first));
  const bool live = lane < count;
  const int tgt = live ? torder[first + lane] : 0;
  const int a = cell / (nc1 * nc2), b = (cell / nc2) % nc1, c = cell % nc2;
  const float csz = (float)cellsz, eps2f = (float)eps2;
  const float pre2 = rcut2 * 1.0001f;  // HILO's FP32 distance test
  // the target relative to its cell's corner; idle lanes sit far away, so
  // no source is ever within rcut of them
  const double d0x = live ? tx[tgt] - (ox + a * cellsz) : 1e18;
  const double d0y = live ? ty[tgt] - (oy + b * cellsz) : 1e18;
  const double d0z = live ? tz[tgt] - (oz + c * cellsz) : 1e18;
  const float t0x = (float)d0x, t0y = (float)d0y, t0z = (float)d0z;
  double hx = 0.0, hy = 0.0, hz = 0.0;
  const int c0 = max(c - 1, 0), c1 = min(c + 1, nc2 - 1);
  for (int da = max(a - 1, 0); da <= min(a + 1, nc0 - 1); ++da)
    for (int db = max(b - 1, 0); db <= min(b + 1, nc1 - 1); ++db) {
      const int row = (da * nc1 + db) * nc2;
      // HILO: the target in this cell's frame in FP64
      const double qx = d0x - (double)(da - a) * cellsz, qy = d0y - (double)(db - b) * cellsz;
      const float px = t0x - (float)(da - a) * csz;
      const float py = t0y - (float)(db - b) * csz;
      for (int cc = c0; cc <= c1; ++cc) {
        const double qz = d0z - (double)(cc - c) * cellsz;
        const float pz = t0z - (float)(cc - c) * csz;
        const int s1 = cell_start[row + cc + 1];
        int s = cell_start[row + cc];
        if (HILO) {
          // the distance test in FP32 on the hi parts with a margin (their
          // r^2 is within ~1e-6 of the exact one), then (hi - q) + lo and
          // the term in FP64 for the sources some lane has inside; a term
          // beyond rcut adds exactly zero (table clamp), so the test only
          // decides what is computed, never what is summed
          for (; s + SG <= s1; s += SG) {
            float4 q[SG], ql[SG];
#pragma unroll
            for (int j = 0; j < SG; ++j) {
              q[j] = near[s + j];
              ql[j] = near_lo[s + j];
            }
            unsigned long long need[SG];
#pragma unroll
            for (int j = 0; j < SG; ++j) {
              const float ex = q[j].x - px, ey = q[j].y - py, ez = q[j].z - pz;
              need[j] = __ballot(fmaf(ex, ex, fmaf(ey, ey, ez * ez)) < pre2);
            }
#pragma unroll
            for (int j = 0; j < SG; ++j)
              if (need[j]) {
                const double dx = ((double)q[j].x - qx) + (double)ql[j].x;
                const double dy = ((double)q[j].y - qy) + (double)ql[j].y;
                const double dz = ((double)q[j].z - qz) + (double)ql[j].z;
                pp_term_d(q[j].w, dx, dy, dz, fma(dx, dx, fma(dy, dy, dz * dz)), tab,
                          tab_scale, eps2, hx, hy, hz);
              }
          }
          for (; s < s1; ++s) {
            const float4 q = near[s], l = near_lo[s];
            const double dx = ((double)q.x - qx) + (double)l.x;
            const double dy = ((double)q.y - qy) + (double)l.y;
            const double dz = ((double)q.z - qz) + (double)l.z;
            const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
            if (r2 < (double)pre2)  // as the groups: every term with t < GT
              pp_term_d(q.w, dx, dy, dz, r2, tab, tab_scale, eps2, hx, hy, hz);
          }
          continue;
        }
        float sx = 0.f, sy = 0.f, sz = 0.f;
        for (; s + SG <= s1; s += SG) {
          float4 q[SG];
#pragma unroll
          for (int j = 0; j < SG; ++j) q[j] = near[s + j];
          float dx[SG], dy[SG], dz[SG], r2[SG];
          unsigned long long need[SG];
#pragma unroll
          for (int j = 0; j < SG; ++j) {
            dx[j] = q[j].x - px;
            dy[j] = q[j].y - py;
            dz[j] = q[j].z - pz;
            r2[j] = fmaf(dx[j], dx[j], fmaf(dy[j], dy[j], dz[j] * dz[j]));
            need[j] = __ballot(r2[j] < rcut2);
          }
          // a source beyond rcut of every lane costs its distance only
#pragma unroll
          for (int j = 0; j < SG; ++j)
            if (need[j])
              pp_term(q[j].w, dx[j], dy[j], dz[j], r2[j], tab, tab_scale, eps2f, sx, sy, sz);
        }
        for (; s < s1; ++s) {
          const float4 q = near[s];
          const float dx = q.x - px, dy = q.y - py, dz = q.z - pz;
          const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
          if (r2 < rcut2) pp_term(q.w, dx, dy, dz, r2, tab, tab_scale, eps2f, sx, sy, sz);
        }
        hx += (double)sx;
        hy += (double)sy;
        hz += (double)sz;
      }
    }
  if (live) {
    if (accumulate) {
      ax[tgt] += G * hx;
      ay[tgt] += G * hy;
      az[tgt] += G * hz;
    } else {
      ax[tgt] = G * hx;
      ay[tgt] = G * hy;
      az[tgt] = G * hz;
    }
  }
}

// ---------------------------------------------------------------------
// The near field as a filter and a sparse sum (k_pp_bits).
//
// One wave takes a chunk of <= 64 targets of one cell, as above, and its
// sources are the 27 neighbour cells' masses: 9 rows (da, db) of up to 3
// cells adjacent in z, each row one contiguous range of the binned masses.
// Work proceeds in batches of <= CAP sources; batch b of B takes every
// B-th source of each row (positions b, b + B, ...), so every target's
// pairs inside rcut spread evenly over the batches.  Per batch:
//
// 1. staging: the batch's sources relative to the target cell's corner
//    (block-relative hi + lo plus the cell offset, in FP64) with their
//    masses into the wave's LDS records (FP32 float4; PRECISE: FP64), and
//    each source's filter row (x', y', z', |s'|^2 in FP16, cell units about
//    the cell centre) kept in the staging lane's registers;
// 2. filter: r^2 - thr for 64 targets x 32 sources per step by two
//    v_mfma_f32_32x32x8_f16 (K = 5 used: s'.(-2p') + |s'|^2 + |p'|^2 - thr),
//    whose sign bits v_alignbit packs into one 32-bit mask per target and
//    32-source tile (1 VALU per 32 pairs); the FP16 rounding moves r^2 by
//    < 0.005 cell^2 and thr = (rcut / cell)^2 + 0.02, so no pair inside
//    rcut is lost, and the few extra pairs add exactly zero (table clamp);
// 3. sum: each lane walks the set bits of its own masks (v_ffbl) and adds
//    the pair term from the LDS record, so only pairs inside rcut (plus the
//    filter's margin) cost a term, against every pair of a source that any
//    lane of the wave needs in k_pp_near.
//
// A target's terms are summed in the order batch, tile, bit, a function of
// its cell alone: results do not depend on the other targets of its wave
// or on scheduling.  The factor gS comes from a table of GTB intervals in
// s = sqrt(r^2 + eps2) (k_gtab_s).  PRECISE (FP64 records): quadratic
// interpolation (error ~1e-9 of gS), (r^2 + eps2)^(-1/2) refined by a Newton
// step in FP64 and FP64 sums, as pp_term_d; otherwise linear interpolation
// (~2e-7 of gS) and FP32 pair terms (v_rsq) on positions rounded to FP32
// relative to the target cell's corner, summed in FP32 over at most 8 terms
// and then in FP64.
constexpr int BW = 4;              // waves per workgroup
constexpr int BT = BW * 64;
constexpr int GTB = 2048;          // intervals of the filtered kernel's table
constexpr float FILTER_MARGIN = 0.02f;  // cell units^2, > 4x the FP16 error bound
#ifndef BITS_CAP
#define BITS_CAP 192                // staged sources per batch, FP32 records
#endif

struct __align__(16) SrcRec {
  double x, y;
  double z;
  float m, pad;
};

typedef _Float16 h4_t __attribute__((ext_vector_type(4)));
typedef float f16v_t __attribute__((ext_vector_type(16)));

// Position within a 32-source tile (= bit of the tile mask) of MFMA row R:
// lane l of v_mfma_f32_32x32x8 holds rows (r & 3) + 8 (r >> 2) + 4 (l >> 5)
// in its registers r = 0..15; alignbit leaves register r in bit 15 - r, and
// the lane halves l >> 5 = 0, 1 give the low and the high 16 bits.
__device__ __forceinline__ int tile_pos(int R) {
  const int h = (R >> 2) & 1, r = (R & 3) + 4 * (R >> 3);
  return 16 * h + 15 - r;
}

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool PRECISE>
struct BitsRec {
  using T = SrcRec;
};
template <>
struct BitsRec<false> {
  using T = float4;
};

template <bool PRECISE, int CAP>
__global__ __launch_bounds__(BT) void k_pp_bits(
    const double* __restrict__ tx, const double* __restrict__ ty,
    const double* __restrict__ tz, double ox, double oy, double oz, double cellsz,
    int nc0, int nc1, int nc2, const int* __restrict__ cell_start,
    const float4* __restrict__ near, const float4* __restrict__ near_lo,
    const int* __restrict__ cstart, const int* __restrict__ cmap,
    const int* __restrict__ tstart, const int* __restrict__ torder, int ncells,
    double eps2, const float2* __restrict__ gtab, float tab_scale, float tab_off, float thr,
    double G, double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az,
    int accumulate) {
  using Rec = typename BitsRec<PRECISE>::T;
  constexpr int NT = CAP / 32;   // tiles per batch
  constexpr int NS = CAP / 64;   // staging slots per lane
  const int nchunks = cstart[ncells];
  const int nblk = (nchunks + BW - 1) / BW;
  if ((int)blockIdx.x >= nblk) return;
  __shared__ float2 tab[GTB + 2];
  __shared__ Rec rec[BW][CAP];
  __shared__ unsigned int msk[BW][NT][64];
  __shared__ int rowtab[BW][9][8];
  for (int i = threadIdx.x; i <= GTB + 1; i += BT) tab[i] = gtab[i];
  __syncthreads();  // the only workgroup barrier: before any wave leaves
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  // XCD-contiguous chunk ranges: neighbouring cells share their sources
  const int lb = ssim_xcd_block(blockIdx.x, nblk);
  const int w = __builtin_amdgcn_readfirstlane(lb * BW + wv);
  if (w >= nchunks) return;
  const int cell = __builtin_amdgcn_readfirstlane(cmap[w]);
  const int first = __builtin_amdgcn_readfirstlane(tstart[cell] + (w - cstart[cell]) * CHUNK);
  const int count = __builtin_amdgcn_readfirstlane(min(CHUNK, tstart[cell + 1] - first));
  const bool live = lane < count;
  const int tgt = live ? torder[first + lane] : 0;
  const int a = cell / (nc1 * nc2), b = (cell / nc2) % nc1, c = cell % nc2;
  const double inv_cs = 1.0 / cellsz;
  // the target relative to its cell's corner (FP64), and in cell units
  // about the cell centre for the filter
  const double px = live ? tx[tgt] - (ox + a * cellsz) : 0.0;
  const double py = live ? ty[tgt] - (oy + b * cellsz) : 0.0;
  const double pz = live ? tz[tgt] - (oz + c * cellsz) : 0.0;
  const float px32 = (float)px, py32 = (float)py, pz32 = (float)pz;
  const float fx = (float)(px * inv_cs) - 0.5f, fy = (float)(py * inv_cs) - 0.5f,
              fz = (float)(pz * inv_cs) - 0.5f;
  // B operands (k x target): lanes 0-31 k 0..3 = (-2p', 1), lanes 32-63
  // k 4..7 = (|p'|^2 - thr, 0, 0, 0); an idle column gets +1000 (no bits)
  const float pq = live ? fmaf(fx, fx, fmaf(fy, fy, fz * fz)) - thr : 1000.0f;
  const float ox_ = __shfl_xor(fx, 32), oy_ = __shfl_xor(fy, 32), oz_ = __shfl_xor(fz, 32);
  const float oq = __shfl_xor(pq, 32);
  h4_t B1, B2;
  if (lane < 32) {
    // B1: target lane (own), B2: target lane + 32 (partner)
    B1 = h4_t{(_Float16)(-2.0f * fx), (_Float16)(-2.0f * fy), (_Float16)(-2.0f * fz),
              (_Float16)1.0f};
    B2 = h4_t{(_Float16)(-2.0f * ox_), (_Float16)(-2.0f * oy_), (_Float16)(-2.0f * oz_),
              (_Float16)1.0f};
  } else {
    // B1: |p'|^2 - thr of target lane - 32 (partner), B2: of target lane (own)
    B1 = h4_t{(_Float16)oq, (_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f};
    B2 = h4_t{(_Float16)pq, (_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f};
  }
  const h4_t A_far = h4_t{(_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f, (_Float16)1000.0f};
  const h4_t A_hi = h4_t{(_Float16)1.0f, (_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f};

  // the 9 rows: lane r < 9 describes row (da, db) = (r / 3 - 1, r % 3 - 1)
  int (*rt)[8] = rowtab[wv];
  int len = 0;
  if (lane < 9) {
    const int da = lane / 3 - 1, db = lane % 3 - 1;
    const int A = a + da, Bq = b + db;
    const int c0 = max(c - 1, 0), c1 = min(c + 1, nc2 - 1);
    int r0 = 0, e1 = 0, e2 = 0;
    if (A >= 0 && A < nc0 && Bq >= 0 && Bq < nc1) {
      const int row = (A * nc1 + Bq) * nc2;
      r0 = cell_start[row + c0];
      const int end = cell_start[row + c1 + 1];
      e1 = c0 + 1 <= c1 ? cell_start[row + c0 + 1] : end;
      e2 = c0 + 2 <= c1 ? cell_start[row + c0 + 2] : end;
      len = end - r0;
    }
    rt[lane][0] = r0;
    rt[lane][1] = e1;
    rt[lane][2] = e2;
    rt[lane][3] = c0 - c;
    rt[lane][4] = da;
    rt[lane][5] = db;
  }
  int total = len;
#pragma unroll
  for (int off = 1; off < 16; off <<= 1) total += __shfl_xor(total, off);
  total = __builtin_amdgcn_readfirstlane(total);
  // batches: sum over rows of ceil(len_r / B) <= total / B + 9 <= CAP
  const int nbatch = max(1, (total + (CAP - 9) - 1) / (CAP - 9));

  double hx = 0.0, hy = 0.0, hz = 0.0;
  const float eps2f = (float)eps2;
  // the table offset in a VGPR once: the pair loop's fma already takes
  // tab_scale from an SGPR, and a VOP3 reads at most one
  float toff;
  asm volatile("v_mov_b32 %0, %1" : "=v"(toff) : "s"(tab_off));
  for (int bt = 0; bt < nbatch; ++bt) {
    // slots of this batch per row and their exclusive prefix
    const int nr = (lane < 9 && len > bt) ? (len - bt + nbatch - 1) / nbatch : 0;
    int incl = nr;
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
      const int v = __shfl_up(incl, off, 64);
      if (lane >= off) incl += v;
    }
    const int nslots = __builtin_amdgcn_readlane(incl, 8);
    if (lane < 9) rt[lane][6] = incl - nr;
    int sb[9];
#pragma unroll
    for (int r = 1; r < 9; ++r) sb[r] = __builtin_amdgcn_readlane(incl - nr, r);
    wave_sync_lds();
    // 1. staging: slot S = lane + 64 k is tile t = S / 32, filter row
    //    R = lane & 31, record position 32 t + tile_pos(R) (loads first)
    float4 hv[NS], lv[NS];
    int srow[NS], sidx[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      const int S = lane + 64 * k;
      int row = 0;
#pragma unroll
      for (int r = 1; r < 9; ++r) row += (S >= sb[r]) ? 1 : 0;
      srow[k] = row;
      sidx[k] = -1;
      if (S < nslots) {
        const int s = rt[row][0] + bt + (S - rt[row][6]) * nbatch;
        sidx[k] = s;
        hv[k] = near[s];
        lv[k] = near_lo ? near_lo[s] : make_float4(0.f, 0.f, 0.f, 0.f);
      }
    }
    h4_t arow[NS];
    const int rpos = tile_pos(lane & 31) + (lane & 32);  // record position in k's 64
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      const int s = sidx[k];
      arow[k] = A_far;
      if (s >= 0) {
        const int* q = rt[srow[k]];
        const int cc = q[3] + (s >= q[1] ? 1 : 0) + (s >= q[2] ? 1 : 0);
        const double x = ((double)hv[k].x + (double)lv[k].x) + (double)q[4] * cellsz;
        const double y = ((double)hv[k].y + (double)lv[k].y) + (double)q[5] * cellsz;
        const double z = ((double)hv[k].z + (double)lv[k].z) + (double)cc * cellsz;
        if constexpr (PRECISE) {
          SrcRec r;
          r.x = x;
          r.y = y;
          r.z = z;
          r.m = hv[k].w;
          r.pad = 0.0f;
          rec[wv][64 * k + rpos] = r;
        } else {
          rec[wv][64 * k + rpos] = make_float4((float)x, (float)y, (float)z, hv[k].w);
        }
        const float sx = (float)(x * inv_cs) - 0.5f, sy = (float)(y * inv_cs) - 0.5f,
                    sz = (float)(z * inv_cs) - 0.5f;
        arow[k] = h4_t{(_Float16)sx, (_Float16)sy, (_Float16)sz,
                       (_Float16)fmaf(sx, sx, fmaf(sy, sy, sz * sz))};
      }
    }
    wave_sync_lds();
    // 2. filter: one 32-bit mask per target and tile; tile 2k's rows are
    //    lanes 0-31's arow[k], tile 2k + 1's those of lanes 32-63
    const int ntiles = (nslots + 31) >> 5;
    unsigned int occ = 0;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      if (t < ntiles) {
        h4_t A = arow[t >> 1];
        if (t & 1) {
          const int2 v = *reinterpret_cast<const int2*>(&A);
          int2 u;
          u.x = __shfl_xor(v.x, 32);
          u.y = __shfl_xor(v.y, 32);
          A = *reinterpret_cast<const h4_t*>(&u);
        }
        if (lane >= 32) A = A_hi;
        const f16v_t zero = {};
        const f16v_t d1 = __builtin_amdgcn_mfma_f32_32x32x8f16(A, B1, zero, 0, 0, 0);
        const f16v_t d2 = __builtin_amdgcn_mfma_f32_32x32x8f16(A, B2, zero, 0, 0, 0);
        unsigned int m1 = 0, m2 = 0;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          m1 = __builtin_amdgcn_alignbit(m1, __float_as_uint(d1[r]), 31);
          m2 = __builtin_amdgcn_alignbit(m2, __float_as_uint(d2[r]), 31);
        }
        // lane < 32 owns target lane (m1 halves), lane >= 32 target lane (m2)
        const unsigned int recv = __shfl_xor(lane < 32 ? m2 : m1, 32);
        const unsigned int mk = lane < 32 ? (m1 | (recv << 16)) : (recv | (m2 << 16));
        msk[wv][t][lane] = mk;
        occ |= (mk != 0u ? 1u : 0u) << t;
      }
    }
    wave_sync_lds();
    // 3. sum over the set bits
    unsigned int mcur = 0;
    const Rec* tile = rec[wv];
    float sx = 0.f, sy = 0.f, sz = 0.f;
    int it = 0;
    while ((mcur | occ) != 0u) {
      if (mcur == 0u) {
        const int t = __builtin_ctz(occ);
        occ &= occ - 1u;
        mcur = msk[wv][t][lane];
        tile = rec[wv] + 32 * t;
      }
      const int j = __builtin_ctz(mcur);
      mcur &= mcur - 1u;
      const Rec r = tile[j];
      if constexpr (PRECISE) {
        const double dx = r.x - px, dy = r.y - py, dz = r.z - pz;
        const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
        const double xx = r2 + eps2;
        double yy = (double)__builtin_amdgcn_rsqf((float)xx);
        const double res = fma(-(xx * yy), yy, 1.0);
        yy = fma(0.5 * yy, res, yy);
        // table in s = sqrt(r^2 + eps2) = xx yy, quadratic interpolation
        const float tt = __builtin_amdgcn_fmed3f(fmaf((float)(xx * yy), tab_scale, tab_off),
                                                 0.0f, (float)GTB);
        const int ti = (int)tt;
        const float fr = __builtin_amdgcn_fractf(tt);
        const float2 e0 = tab[ti], e1 = tab[ti + 1];
        const float cq = fmaf((fr - 1.0f) * (e1.y - e0.y), 0.5f, e0.y);
        const float g = fmaf(fr, cq, e0.x);
        const double f = ((double)r.m * (double)g) * (yy * yy * yy);
        hx = fma(f, dx, hx);
        hy = fma(f, dy, hy);
        hz = fma(f, dz, hz);
      } else {
        const float dx = r.x - px32, dy = r.y - py32, dz = r.z - pz32;
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        const float xx = r2 + eps2f;
        const float inv = __builtin_amdgcn_rsqf(xx);
        // table in s = sqrt(r^2 + eps2) = xx inv, linear interpolation
        const float tt = __builtin_amdgcn_fmed3f(fmaf(xx * inv, tab_scale, toff), 0.0f,
                                                 (float)GTB);
        const float2 e = tab[(int)tt];
        const float g = fmaf(__builtin_amdgcn_fractf(tt), e.y, e.x);
        const float f = r.w * (inv * inv) * inv * g;
        sx = fmaf(f, dx, sx);
        sy = fmaf(f, dy, sy);
        sz = fmaf(f, dz, sz);
        if ((++it & 7) == 0) {
          hx += (double)sx;
          hy += (double)sy;
          hz += (double)sz;
          sx = sy = sz = 0.f;
        }
      }
    }
    if constexpr (!PRECISE) {
      hx += (double)sx;
      hy += (double)sy;
      hz += (double)sz;
    }
    wave_sync_lds();  // the records and masks are rewritten by the next batch
  }
  if (live) {
    if (accumulate) {
      ax[tgt] += G * hx;
      ay[tgt] += G * hy;
      az[tgt] += G * hz;
    } else {
      ax[tgt] = G * hx;
      ay[tgt] = G * hy;
      az[tgt] = G * hz;
    }
  }
}

// The table scale s = GT / rcut^2 rounded up until rcut2 * s >= GT in
// FP32 (rcut2 = rcut^2 in FP32): r2 >= rcut2 then gives t = min(r2 s, GT) =
// GT exactly, whose entries are zero, so a source at or beyond rcut adds
// exactly nothing whether or not another lane of the wave needs it, and a
// target's sum does not depend on which targets share its wave.
float tab_scale_for(double rcut, int gt = GT) {
  const float rcut2 = (float)(rcut * rcut);
  float s = (float)(gt / (rcut * rcut));
  while (rcut2 * s < (float)gt) s = nextafterf(s, INFINITY);
  return s;
}

}  // namespace

extern "C" {

size_t ssim_pp_work_bytes(int nt, int ncells) {
  const size_t nc = (size_t)(ncells > 0 ? ncells : 1);
  const size_t chunks = (size_t)(nt > 0 ? nt : 1) / CHUNK + nc + 1;
  return ssim::align256(nc * sizeof(int)) + ssim::align256((nc + 1) * sizeof(int)) +
         ssim::align256(chunks * sizeof(int)) + ssim::align256((GT + 2) * sizeof(float2)) +
         ssim::scan_work_bytes((int)nc);
}

}  // extern "C"

namespace {

__global__ void k_zero_u64(unsigned long long* c) {
  if (threadIdx.x < 2) c[threadIdx.x] = 0;
}

// The table and the chunk plan in `work` (ssim_pp_work_bytes): gS table,
// chunk starts per cell and the chunk -> cell map.
int pp_plan(int nt, int ncells, const int* tcell_start, double rs, double rcut,
            void* work, int** cstart_out, int** cmap_out, float2** gtab_out,
            hipStream_t st, int gt = GT, double eps2_s = 0.0) {
  unsigned char* w = static_cast<unsigned char*>(work);
  int* nchunk = reinterpret_cast<int*>(w);
  w += ssim::align256((size_t)ncells * sizeof(int));
  int* cstart = reinterpret_cast<int*>(w);
  w += ssim::align256(((size_t)ncells + 1) * sizeof(int));
  int* cmap = reinterpret_cast<int*>(w);
  const size_t max_chunks = (size_t)nt / CHUNK + ncells + 1;
  w += ssim::align256(max_chunks * sizeof(int));
  float2* gtab = reinterpret_cast<float2*>(w);
  w += ssim::align256((GT + 2) * sizeof(float2));
  // eps2_s > 0: the table in s = sqrt(r^2 + eps2_s) of k_pp_bits
  if (eps2_s > 0.0) {
    hipLaunchKernelGGL(k_gtab_s, dim3(ssim::ceil_div(gt + 2, PB)), dim3(PB), 0, st, rs,
                       rcut, eps2_s, gt, gtab);
    SSIM_LAUNCHED("k_gtab_s");
  } else {
    hipLaunchKernelGGL(k_gtab, dim3(ssim::ceil_div(gt + 2, PB)), dim3(PB), 0, st, rs,
                       rcut, gt, gtab);
    SSIM_LAUNCHED("k_gtab");
  }
  const int cblocks = ssim::ceil_div(ncells, 4 * CP_T);
  if (cblocks <= CP_BLOCKS) {
    // w: the scan workspace, >= ceil(ncells / 1024) + 1 ints
    int* bsum = reinterpret_cast<int*>(w);
    hipLaunchKernelGGL(k_chunk_sum, dim3(cblocks), dim3(CP_T), 0, st, ncells, tcell_start,
                       nchunk, bsum);
    SSIM_LAUNCHED("k_chunk_sum");
    hipLaunchKernelGGL(k_chunk_scan, dim3(cblocks), dim3(CP_T), 0, st, ncells, nchunk, bsum,
                       cstart, cmap);
    SSIM_LAUNCHED("k_chunk_scan");
  } else {
    const dim3 gc(ssim::ceil_div(ncells, PB));
    hipLaunchKernelGGL(k_chunk_count, gc, dim3(PB), 0, st, ncells, tcell_start, nchunk);
    SSIM_LAUNCHED("k_chunk_count");
    SSIM_TRY(ssim::scan_exclusive(ncells, nchunk, cstart, w, st));
    hipLaunchKernelGGL(k_chunk_map, gc, dim3(PB), 0, st, ncells, cstart, cmap);
    SSIM_LAUNCHED("k_chunk_map");
  }
  *cstart_out = cstart;
  *cmap_out = cmap;
  *gtab_out = gtab;
  return SSIM_OK;
}

int pp_impl(int nt, const double* tx, const double* ty, const double* tz,
            const double* origin, double cell, const int* nc,
            const int* cell_start, const void* packed, const int* tcell_start,
            const int* torder, double G, double eps2, double rs, double rcut,
            double* ax, double* ay, double* az, int accumulate,
            unsigned long long* counts, void* work, size_t work_bytes, void* stream) {
  SSIM_REQUIRE(nt >= 0, "nt must be >= 0");
  if (nt == 0) return SSIM_OK;
  SSIM_REQUIRE(tx && ty && tz && origin && nc && cell_start && packed &&
                   tcell_start && torder && work && (counts || (ax && ay && az)),
               "null pointer");
  SSIM_REQUIRE(nc[0] >= 1 && nc[1] >= 1 && nc[2] >= 1, "bad cell counts");
  const long long ncl = (long long)nc[0] * nc[1] * nc[2];
  SSIM_REQUIRE(ncl <= (1 << 26), "too many cells: %lld", ncl);
  const int ncells = (int)ncl;
  SSIM_REQUIRE(rs > 0.0 && rcut > 0.0 && rcut <= cell,
               "need rs > 0 and 0 < rcut <= cell (rs=%g rcut=%g cell=%g)", rs,
               rcut, cell);
  SSIM_REQUIRE(eps2 > 0.0, "softening eps2 must be > 0");
  SSIM_REQUIRE(work_bytes >= ssim_pp_work_bytes(nt, ncells), "workspace too small");
  hipStream_t st = ssim::as_stream(stream);
  int *cstart, *cmap;
  float2* gtab;
  SSIM_TRY(pp_plan(nt, ncells, tcell_start, rs, rcut, work, &cstart, &cmap, &gtab, st));
  const size_t max_chunks = (size_t)nt / CHUNK + ncells + 1;
  // chunks <= nt / 64 + ncells: launch that bound, surplus waves exit
  const dim3 grid(ssim::ceil_div((long long)max_chunks, WAVES));
  if (counts)
    hipLaunchKernelGGL(k_pp_cells<true>, grid, dim3(PB), 0, st, tx, ty, tz,
                       origin[0], origin[1], origin[2], nc[0], nc[1], nc[2],
                       cell_start, static_cast<const float4*>(packed), tcell_start,
                       torder, cstart, cmap, ncells, (float)eps2, gtab,
                       tab_scale_for(rcut), (float)(rcut * rcut), G, ax, ay, az,
                       accumulate, counts);
  else
    SSIM_TIMED("k_pp_cells", st,
               hipLaunchKernelGGL(k_pp_cells<false>, grid, dim3(PB), 0, st, tx, ty,
                                  tz, origin[0], origin[1], origin[2], nc[0], nc[1],
                                  nc[2], cell_start,
                                  static_cast<const float4*>(packed), tcell_start,
                                  torder, cstart, cmap, ncells, (float)eps2, gtab,
                                  tab_scale_for(rcut), (float)(rcut * rcut), G, ax,
                                  ay, az, accumulate, nullptr));
  SSIM_LAUNCHED("k_pp_cells");
  return SSIM_OK;
}

}  // namespace

extern "C" {

static int pp_near_impl(int nt, const double* tx, const double* ty, const double* tz,
                 const double* origin, double cell, const int* nc, const int* cell_start,
                 const void* near, const void* near_lo, const int* tcell_start, const int* torder, double G,
                 double eps2, double rs, double rcut, double* ax, double* ay, double* az,
                 int accumulate, void* work, size_t work_bytes, void* stream) {
  SSIM_REQUIRE(nt >= 0, "nt must be >= 0");
  if (nt == 0) return SSIM_OK;
  SSIM_REQUIRE(tx && ty && tz && origin && nc && cell_start && near &&
                   tcell_start && torder && work && ax && ay && az,
               "null pointer");
  SSIM_REQUIRE(nc[0] >= 1 && nc[1] >= 1 && nc[2] >= 1, "bad cell counts");
  const long long ncl = (long long)nc[0] * nc[1] * nc[2];
  SSIM_REQUIRE(ncl <= (1 << 26), "too many cells: %lld", ncl);
  const int ncells = (int)ncl;
  SSIM_REQUIRE(rs > 0.0 && rcut > 0.0 && rcut <= cell,
               "need rs > 0 and 0 < rcut <= cell (rs=%g rcut=%g cell=%g)", rs, rcut, cell);
  SSIM_REQUIRE(eps2 > 0.0, "softening eps2 must be > 0");
  SSIM_REQUIRE(work_bytes >= ssim_pp_work_bytes(nt, ncells), "workspace too small");
  hipStream_t st = ssim::as_stream(stream);
  int *cstart, *cmap;
  float2* gtab;
  SSIM_TRY(pp_plan(nt, ncells, tcell_start, rs, rcut, work, &cstart, &cmap, &gtab, st));
  const int max_chunks = nt / CHUNK + ncells + 1;
  const dim3 grid(ssim::ceil_div(max_chunks, WAVES));
  const float4* nl4 = static_cast<const float4*>(near_lo);
  if (near_lo) {
    SSIM_TIMED("k_pp_near_hilo", st,
               hipLaunchKernelGGL(k_pp_near<true>, grid, dim3(PB), 0, st, tx, ty, tz,
                                  origin[0], origin[1], origin[2], cell, nc[0], nc[1], nc[2],
                                  cell_start, static_cast<const float4*>(near), nl4,
                                  tcell_start, torder, cstart, cmap, ncells, eps2,
                                  gtab, tab_scale_for(rcut), (float)(rcut * rcut), G,
                                  ax, ay, az, accumulate));
  } else {
    SSIM_TIMED("k_pp_near", st,
               hipLaunchKernelGGL(k_pp_near<false>, grid, dim3(PB), 0, st, tx, ty, tz,
                                  origin[0], origin[1], origin[2], cell, nc[0], nc[1], nc[2],
                                  cell_start, static_cast<const float4*>(near), nl4,
                                  tcell_start, torder, cstart, cmap, ncells, eps2,
                                  gtab, tab_scale_for(rcut), (float)(rcut * rcut), G,
                                  ax, ay, az, accumulate));
  }
  SSIM_LAUNCHED("k_pp_near");
  return SSIM_OK;
}

int ssim_pp_near(int nt, const double* tx, const double* ty, const double* tz,
                 const double* origin, double cell, const int* nc, const int* cell_start,
                 const void* near, const int* tcell_start, const int* torder, double G,
                 double eps2, double rs, double rcut, double* ax, double* ay, double* az,
                 int accumulate, void* work, size_t work_bytes, void* stream) {
  return pp_near_impl(nt, tx, ty, tz, origin, cell, nc, cell_start, near, nullptr,
                      tcell_start, torder, G, eps2, rs, rcut, ax, ay, az, accumulate, work,
                      work_bytes, stream);
}

int ssim_pp_near_hilo(int nt, const double* tx, const double* ty, const double* tz,
                      const double* origin, double cell, const int* nc,
                      const int* cell_start, const void* near, const void* near_lo,
                      const int* tcell_start, const int* torder, double G, double eps2,
                      double rs, double rcut, double* ax, double* ay, double* az,
                      int accumulate, void* work, size_t work_bytes, void* stream) {
  SSIM_REQUIRE(near_lo || nt == 0, "null near_lo");
  return pp_near_impl(nt, tx, ty, tz, origin, cell, nc, cell_start, near, near_lo,
                      tcell_start, torder, G, eps2, rs, rcut, ax, ay, az, accumulate, work,
                      work_bytes, stream);
}

int ssim_pp_near_bits(int nt, const double* tx, const double* ty, const double* tz,
                      const double* origin, double cell, const int* nc,
                      const int* cell_start, const void* near, const void* near_lo,
                      const int* tcell_start, const int* torder, double G, double eps2,
                      double rs, double rcut, double* ax, double* ay, double* az,
                      int accumulate, int precise, void* work, size_t work_bytes,
                      void* stream) {
  SSIM_REQUIRE(nt >= 0, "nt must be >= 0");
  if (nt == 0) return SSIM_OK;
  SSIM_REQUIRE(tx && ty && tz && origin && nc && cell_start && near && tcell_start &&
                   torder && work && ax && ay && az,
               "null pointer");
  SSIM_REQUIRE(nc[0] >= 1 && nc[1] >= 1 && nc[2] >= 1, "bad cell counts");
  const long long ncl = (long long)nc[0] * nc[1] * nc[2];
  SSIM_REQUIRE(ncl <= (1 << 26), "too many cells: %lld", ncl);
  const int ncells = (int)ncl;
  SSIM_REQUIRE(rs > 0.0 && rcut > 0.0 && rcut <= cell,
               "need rs > 0 and 0 < rcut <= cell (rs=%g rcut=%g cell=%g)", rs, rcut, cell);
  SSIM_REQUIRE(eps2 > 0.0, "softening eps2 must be > 0");
  SSIM_REQUIRE(work_bytes >= ssim_pp_work_bytes(nt, ncells), "workspace too small");
  hipStream_t st = ssim::as_stream(stream);
  int *cstart, *cmap;
  float2* gtab;
  SSIM_TRY(pp_plan(nt, ncells, tcell_start, rs, rcut, work, &cstart, &cmap, &gtab, st,
                   GTB, eps2));
  const int max_chunks = nt / CHUNK + ncells + 1;
  const dim3 grid(ssim::ceil_div(max_chunks, BW));
  const float rc = (float)(rcut / cell);
  const float thr = rc * rc + FILTER_MARGIN;
  const float4* n4 = static_cast<const float4*>(near);
  const float4* nl4 = static_cast<const float4*>(near_lo);
  // t = (s - eps) GTB / (s_max - eps), s = sqrt(r^2 + eps2)
  const double e = sqrt(eps2), smax = sqrt(rcut * rcut + eps2);
  const float scale = (float)(GTB / (smax - e)), off = (float)(-e * (GTB / (smax - e)));
#define SSIM_BITS_LAUNCH(MODE, CAP)                                                   \
  SSIM_TIMED(MODE ? "k_pp_bits_precise" : "k_pp_bits", st,                            \
             hipLaunchKernelGGL((k_pp_bits<MODE, CAP>), grid, dim3(BT), 0, st, tx, ty, tz, \
                                origin[0], origin[1], origin[2], cell, nc[0], nc[1], nc[2], \
                                cell_start, n4, nl4, cstart, cmap, tcell_start, torder,    \
                                ncells, eps2, gtab, scale, off, thr, G, ax, ay, az,        \
                                accumulate))
  if (precise)
    SSIM_BITS_LAUNCH(true, 128);
  else
    SSIM_BITS_LAUNCH(false, BITS_CAP);
#undef SSIM_BITS_LAUNCH
  SSIM_LAUNCHED("k_pp_bits");
  return SSIM_OK;
}

int ssim_pp_short(int nt, const double* tx, const double* ty, const double* tz,
                  const double* origin, double cell, const int* nc,
                  const int* cell_start, const void* packed,
                  const int* tcell_start, const int* torder, double G,
                  double eps2, double rs, double rcut, double* ax, double* ay,
                  double* az, int accumulate, void* work, size_t work_bytes,
                  void* stream) {
  return pp_impl(nt, tx, ty, tz, origin, cell, nc, cell_start, packed,
                 tcell_start, torder, G, eps2, rs, rcut, ax, ay, az, accumulate,
                 nullptr, work, work_bytes, stream);
}

int ssim_pp_count(int nt, const double* tx, const double* ty, const double* tz,
                  const double* origin, double cell, const int* nc,
                  const int* cell_start, const void* packed,
                  const int* tcell_start, const int* torder, double eps2,
                  double rs, double rcut, unsigned long long* counts, void* work,
                  size_t work_bytes, void* stream) {
  SSIM_REQUIRE(counts, "null pointer");
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, ssim::as_stream(stream),
                     counts);
  SSIM_LAUNCHED("k_zero_u64");
  return pp_impl(nt, tx, ty, tz, origin, cell, nc, cell_start, packed,
                 tcell_start, torder, 1.0, eps2, rs, rcut, nullptr, nullptr,
                 nullptr, 1, counts, work, work_bytes, stream);
}

}  // extern "C"
