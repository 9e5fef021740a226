// Few massive bodies + many asteroids: the raster-batch integrator of
// `frame_kernel` (nbody_kernel.py:97-560) on one MI355X.
//
// Physics (same as the reference): Yoshida 4th order as three weighted
// velocity-Verlet steps; the coarse sub-step dtH follows the closest
// massive-massive encounter (tEnc/20, floor maxSubDt/1000); asteroids whose
// encounter time with a massive body is below dtH are "hot" and integrate
// [0, dtH] in a private adaptive fine loop against linearly interpolated
// massive positions, landing exactly on the sub-sample grid j*dtRaster/mSub
// and reporting the first contact; every visible asteroid pulls back on
// the massive bodies (Newton's third law); frozen bodies (vis = 0) do not
// move.  Outputs per raster: f32 snapshots of asteroids and massive bodies,
// sub-samples of hot asteroids, first-contact time/partner.
//
// Structure (MI355X-first, one GPU):
//  * one cooperative launch per batch; every workgroup keeps its own copy of
//    the <= 64 massive bodies in LDS and advances it with exactly the same
//    instructions, so all copies stay bit-identical without broadcasts, and
//    the sub-step control (dtH, remaining, guard) is computed per
//    workgroup from that copy — uniform control flow with no global flags;
//  * the only grid-wide dependency is the asteroids' back-reaction on the
//    massive bodies: each wave reduces its lanes with shuffles, each
//    workgroup writes one partial per massive body, and after one grid
//    barrier every workgroup sums the partials in workgroup order
//    (deterministic; the reference uses float atomics).  One grid barrier
//    per Verlet step instead of the reference's four.
//  * FP64 throughout, built without FMA contraction (operation order of the
//    reference's formulas).
#include <hip/hip_cooperative_groups.h>

#include "ssim_internal.h"

namespace cg = cooperative_groups;

namespace {

constexpr int FB = 256;
constexpr int MMAX = 64;
constexpr int NWAVE = FB / 64;

// D = 2 (the reference's plane) or 3 (the same scheme in space).  Every
// vector quantity is an array of D components; |d|^2 is summed x, y, (z)
// in that order, so the D = 2 instance performs exactly the reference's
// operations.
template <int D>
struct FrameArgs {
  double* ap[D];    // asteroid positions, velocities, accelerations
  double* av[D];
  double* aacc[D];
  const double* am;
  const unsigned char* aVis;
  const double* aRad;
  double* mp[D];    // massive bodies
  double* mv[D];
  double* macc[D];
  const double* mm;
  const unsigned char* mVis;
  const double* mRad;
  unsigned char* hot;
  double* part;  // [2][gridDim][D][MMAX]
  float *snap, *snapM, *subPos, *hitT;
  unsigned char* subN;
  int* hitM;
  int nAst, M, mSub, K, maxSubSteps;
  double G, soft, dtRaster, maxSubDt, w1, w0;
};

template <int D>
struct Massive {
  double x[D][MMAX], v[D][MMAX], acc[D][MMAX], x0[D][MMAX];
  double m[MMAX], rad[MMAX];
  unsigned char vis[MMAX];
};

// d[0]^2 + d[1]^2 (+ d[2]^2), left to right
template <int D>
__device__ __forceinline__ double sq_sum(const double (&d)[D]) {
  double r = d[0] * d[0];
#pragma unroll
  for (int c = 1; c < D; ++c) r = r + d[c] * d[c];
  return r;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Asteroid accelerations against the massive bodies in LDS and the
// back-reaction partial of this workgroup.  kick: 0 = only store the new
// acceleration (initial), 1 = quiet asteroids also get the velocity kick
// v += 0.5 (a_old + a_new) dt (phase B, nbody_kernel.py:336-368).
template <int D>
__device__ void accel_phase(const FrameArgs<D>& A, const Massive<D>& S,
                            double (*s_wave)[D][MMAX], int kick, double dt,
                            double* part_out) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int M = A.M;
  for (int k = threadIdx.x; k < NWAVE * D * MMAX; k += FB)
    (&s_wave[0][0][0])[k] = 0.0;
  __syncthreads();
  const int stride = gridDim.x * FB;
  for (int i0 = blockIdx.x * FB + wave * 64; i0 < A.nAst; i0 += stride) {
    const int i = i0 + lane;
    const bool on = i < A.nAst && A.aVis[i];
    double p[D], pm = 0.0;
#pragma unroll
    for (int c = 0; c < D; ++c) p[c] = on ? A.ap[c][i] : 0.0;
    if (on) pm = A.am[i];
    double ac[D];
#pragma unroll
    for (int c = 0; c < D; ++c) ac[c] = 0.0;
    for (int kk = 0; kk < M; ++kk) {
      if (!S.vis[kk]) continue;
      double bk[D];
#pragma unroll
      for (int c = 0; c < D; ++c) bk[c] = 0.0;
      if (on) {
        double d[D];
#pragma unroll
        for (int c = 0; c < D; ++c) d[c] = S.x[c][kk] - p[c];
        const double r2 = sq_sum<D>(d) + A.soft;
        const double f = A.G / (r2 * sqrt(r2));
#pragma unroll
        for (int c = 0; c < D; ++c) {
          ac[c] += f * S.m[kk] * d[c];
          bk[c] = -f * pm * d[c];
        }
      }
#pragma unroll
      for (int c = 0; c < D; ++c) bk[c] = wave_sum(bk[c]);
      if (lane == 0) {
#pragma unroll
        for (int c = 0; c < D; ++c) s_wave[wave][c][kk] += bk[c];
      }
    }
    if (i < A.nAst) {
      if (!kick) {
#pragma unroll
        for (int c = 0; c < D; ++c) A.aacc[c][i] = ac[c];
      } else if (on && !A.hot[i]) {
#pragma unroll
        for (int c = 0; c < D; ++c) {
          A.av[c][i] += 0.5 * (A.aacc[c][i] + ac[c]) * dt;
          A.aacc[c][i] = ac[c];
        }
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < D * MMAX) {
    const int c = threadIdx.x / MMAX, kk = threadIdx.x % MMAX;
    double s = 0.0;
    for (int w = 0; w < NWAVE; ++w) s += s_wave[w][c][kk];
    part_out[(size_t)blockIdx.x * D * MMAX + c * MMAX + kk] = s;
  }
}

// Massive accelerations (back-reaction summed in workgroup order + massive
// pairs); kick != 0 applies v += 0.5 (a_old + a_new) dt to visible bodies
// (phase C, nbody_kernel.py:376-399).  Threads < M, result in LDS.
template <int D>
__device__ void massive_phase(const FrameArgs<D>& A, Massive<D>& S,
                              const double* part, int kick, double dt) {
  const int i = threadIdx.x;
  if (i >= A.M) return;
  double ac[D];
#pragma unroll
  for (int c = 0; c < D; ++c) ac[c] = 0.0;
  for (int b = 0; b < (int)gridDim.x; ++b)
#pragma unroll
    for (int c = 0; c < D; ++c) ac[c] += part[(size_t)b * D * MMAX + c * MMAX + i];
  if (S.vis[i]) {
    for (int j = 0; j < A.M; ++j) {
      if (j == i || !S.vis[j]) continue;
      double d[D];
#pragma unroll
      for (int c = 0; c < D; ++c) d[c] = S.x[c][j] - S.x[c][i];
      const double r2 = sq_sum<D>(d) + A.soft;
      const double f = A.G / (r2 * sqrt(r2));
#pragma unroll
      for (int c = 0; c < D; ++c) ac[c] += f * S.m[j] * d[c];
    }
    if (kick)
#pragma unroll
      for (int c = 0; c < D; ++c) S.v[c][i] += 0.5 * (S.acc[c][i] + ac[c]) * dt;
  }
#pragma unroll
  for (int c = 0; c < D; ++c) S.acc[c][i] = ac[c];
}

// acceleration of a point p from the massive bodies at the interpolated
// positions x0 + (x - x0) al (the fine loop's moving sources)
template <int D>
__device__ __forceinline__ void interp_accel(const FrameArgs<D>& A, const Massive<D>& S,
                                             const double (&p)[D], double al,
                                             double (&ac)[D]) {
#pragma unroll
  for (int c = 0; c < D; ++c) ac[c] = 0.0;
  for (int kk = 0; kk < A.M; ++kk) {
    if (!S.vis[kk]) continue;
    double d[D];
#pragma unroll
    for (int c = 0; c < D; ++c)
      d[c] = (S.x0[c][kk] + (S.x[c][kk] - S.x0[c][kk]) * al) - p[c];
    const double r2 = sq_sum<D>(d) + A.soft;
    const double f = A.G / (r2 * sqrt(r2));
#pragma unroll
    for (int c = 0; c < D; ++c) ac[c] += f * S.m[kk] * d[c];
  }
}

template <int D>
__global__ __launch_bounds__(FB) void k_nbody_frame(FrameArgs<D> A) {
  cg::grid_group grid = cg::this_grid();
  __shared__ Massive<D> S;
  __shared__ double s_wave[NWAVE][D][MMAX];
  __shared__ double s_remaining, s_dtH;
  __shared__ int s_done, s_guard;
  const int M = A.M;
  const int tid = blockIdx.x * FB + threadIdx.x;
  const int stride = gridDim.x * FB;
  const size_t pstride = (size_t)gridDim.x * D * MMAX;

  if (threadIdx.x < M) {
    const int i = threadIdx.x;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      S.x[c][i] = A.mp[c][i];
      S.v[c][i] = A.mv[c][i];
    }
    S.m[i] = A.mm[i];
    S.vis[i] = A.mVis[i];
    S.rad[i] = A.mRad[i];
  }
  if (threadIdx.x == 0) s_guard = 0;
  __syncthreads();

  // ---- initial accelerations (nbody_kernel.py:151-216) ----
  int pb = 0;
  accel_phase<D>(A, S, s_wave, 0, 0.0, A.part + pb * pstride);
  grid.sync();
  massive_phase<D>(A, S, A.part + pb * pstride, 0, 0.0);
  pb ^= 1;
  __syncthreads();

  const double wd[3] = {A.w1, A.w0, A.w1};
  const double subDtRaster = A.mSub > 0 ? A.dtRaster / A.mSub : 0.0;
  for (int ks = 0; ks < A.K; ++ks) {
    if (threadIdx.x == 0) s_remaining = A.dtRaster;
    if (A.mSub > 0)
      for (int i = tid; i < A.nAst; i += stride) A.subN[(size_t)ks * A.nAst + i] = 0;
    __syncthreads();
    for (;;) {
      // coarse step from the massive pairs (nbody_kernel.py:245-279)
      if (threadIdx.x == 0) {
        if (s_remaining <= 1e-12 || s_guard++ >= A.maxSubSteps) {
          s_done = 1;
        } else {
          double dtH = A.maxSubDt;
          for (int i = 0; i < M; ++i) {
            if (!S.vis[i]) continue;
            for (int j = i + 1; j < M; ++j) {
              if (!S.vis[j]) continue;
              double d[D], dv[D];
#pragma unroll
              for (int c = 0; c < D; ++c) {
                d[c] = S.x[c][j] - S.x[c][i];
                dv[c] = S.v[c][j] - S.v[c][i];
              }
              double dist = sqrt(sq_sum<D>(d));
              if (dist < 1e-12) dist = 1e-12;
              const double vrel = sqrt(sq_sum<D>(dv));
              if (vrel > 1e-9) {
                const double tEnc = dist / vrel / 20.0;
                if (tEnc < dtH) dtH = tEnc;
              }
            }
          }
          if (dtH < A.maxSubDt / 1000.0) dtH = A.maxSubDt / 1000.0;
          if (dtH > s_remaining) dtH = s_remaining;
          s_dtH = dtH;
          s_remaining -= dtH;
          s_done = 0;
        }
      }
      __syncthreads();
      if (s_done) break;
      const double dtH = s_dtH;

      // hot classification (nbody_kernel.py:283-301)
      for (int i = tid; i < A.nAst; i += stride) {
        unsigned char h = 0;
        if (A.aVis[i]) {
          double p[D], pv[D];
#pragma unroll
          for (int c = 0; c < D; ++c) {
            p[c] = A.ap[c][i];
            pv[c] = A.av[c][i];
          }
          for (int kk = 0; kk < M; ++kk) {
            if (!S.vis[kk]) continue;
            double d[D], dv[D];
#pragma unroll
            for (int c = 0; c < D; ++c) {
              d[c] = S.x[c][kk] - p[c];
              dv[c] = S.v[c][kk] - pv[c];
            }
            double dist = sqrt(sq_sum<D>(d));
            if (dist < 1e-12) dist = 1e-12;
            const double vrel = sqrt(sq_sum<D>(dv));
            if (vrel > 1e-9 && dist / vrel / 20.0 < dtH) {
              h = 1;
              break;
            }
          }
        }
        A.hot[i] = h;
      }
      if (threadIdx.x < M)
#pragma unroll
        for (int c = 0; c < D; ++c) S.x0[c][threadIdx.x] = S.x[c][threadIdx.x];
      __syncthreads();

      // three Yoshida-weighted Verlet steps (nbody_kernel.py:314-401)
      for (int w = 0; w < 3; ++w) {
        const double dt = wd[w] * dtH;
        for (int i = tid; i < A.nAst; i += stride) {  // phase A
          if (!A.aVis[i] || A.hot[i]) continue;
#pragma unroll
          for (int c = 0; c < D; ++c)
            A.ap[c][i] += A.av[c][i] * dt + 0.5 * A.aacc[c][i] * dt * dt;
        }
        if (threadIdx.x < M && S.vis[threadIdx.x]) {
          const int i = threadIdx.x;
#pragma unroll
          for (int c = 0; c < D; ++c)
            S.x[c][i] += S.v[c][i] * dt + 0.5 * S.acc[c][i] * dt * dt;
        }
        __syncthreads();
        accel_phase<D>(A, S, s_wave, 1, dt, A.part + pb * pstride);  // phase B
        grid.sync();
        massive_phase<D>(A, S, A.part + pb * pstride, 1, dt);         // phase C
        pb ^= 1;
        __syncthreads();
      }

      // private fine loop of the hot asteroids (nbody_kernel.py:403-535)
      const double tSeg0 = A.dtRaster - s_remaining - dtH;
      for (int i = tid; i < A.nAst; i += stride) {
        if (!A.aVis[i] || !A.hot[i]) continue;
        double p[D], pv[D];
#pragma unroll
        for (int c = 0; c < D; ++c) {
          p[c] = A.ap[c][i];
          pv[c] = A.av[c][i];
        }
        const double sgn = dtH >= 0.0 ? 1.0 : -1.0;
        const double span = fabs(dtH);
        double tau = 0.0;
        int guardF = 0;
        while (tau < span - 1e-15 && guardF++ < 8000) {
          const double al = tau / span;
          double dtF = A.maxSubDt;
          for (int kk = 0; kk < M; ++kk) {
            if (!S.vis[kk]) continue;
            double d[D], dv[D];
#pragma unroll
            for (int c = 0; c < D; ++c) {
              d[c] = (S.x0[c][kk] + (S.x[c][kk] - S.x0[c][kk]) * al) - p[c];
              dv[c] = S.v[c][kk] - pv[c];
            }
            double dist = sqrt(sq_sum<D>(d));
            if (dist < 1e-12) dist = 1e-12;
            if (A.hitT[i] < 0.0f && dist <= S.rad[kk] + A.aRad[i]) {
              A.hitT[i] = (float)(ks * A.dtRaster + tSeg0 + tau);
              A.hitM[i] = kk;
            }
            const double vrel = sqrt(sq_sum<D>(dv));
            if (vrel > 1e-9) {
              const double tE = dist / vrel / 20.0;
              if (tE < dtF) dtF = tE;
            }
          }
          if (dtF < A.maxSubDt / 1000.0) dtF = A.maxSubDt / 1000.0;
          if (dtF > span - tau) dtF = span - tau;
          if (A.mSub > 0) {  // land exactly on the next sub-sample time
            const double tAbs = tSeg0 + tau;
            const int jn = (int)floor(tAbs / subDtRaster + 1e-9) + 1;
            if (jn <= A.mSub) {
              const double rest = jn * subDtRaster - tAbs;
              if (rest > 0.0 && dtF > rest) dtF = rest;
            }
          }
          double tSub = tau;
          for (int wf = 0; wf < 3; ++wf) {
            const double df = wd[wf] * dtF * sgn;
            double ac[D], nc[D];
            interp_accel<D>(A, S, p, tSub / span, ac);
#pragma unroll
            for (int c = 0; c < D; ++c) p[c] += pv[c] * df + 0.5 * ac[c] * df * df;
            tSub += wd[wf] * dtF;
            interp_accel<D>(A, S, p, tSub / span, nc);
#pragma unroll
            for (int c = 0; c < D; ++c) pv[c] += 0.5 * (ac[c] + nc[c]) * df;
          }
          tau += dtF;
          if (A.mSub > 0) {
            const double q = (tSeg0 + tau) / subDtRaster;
            const int j = (int)floor(q + 0.5);
            if (j >= 1 && j <= A.mSub && fabs(q - j) < 1e-6) {
              const size_t b = ((size_t)ks * A.mSub + (j - 1)) * D * A.nAst;
#pragma unroll
              for (int c = 0; c < D; ++c) A.subPos[b + (size_t)c * A.nAst + i] = (float)p[c];
              A.subN[(size_t)ks * A.nAst + i]++;
            }
          }
        }
        double ac[D];
#pragma unroll
        for (int c = 0; c < D; ++c) {
          A.ap[c][i] = p[c];
          A.av[c][i] = pv[c];
          ac[c] = 0.0;
        }
        for (int kk = 0; kk < M; ++kk) {
          if (!S.vis[kk]) continue;
          double d[D];
#pragma unroll
          for (int c = 0; c < D; ++c) d[c] = S.x[c][kk] - p[c];
          const double r2 = sq_sum<D>(d) + A.soft;
          const double f = A.G / (r2 * sqrt(r2));
#pragma unroll
          for (int c = 0; c < D; ++c) ac[c] += f * S.m[kk] * d[c];
        }
#pragma unroll
        for (int c = 0; c < D; ++c) A.aacc[c][i] = ac[c];
      }
      __syncthreads();
    }

    // f32 snapshot of this raster (nbody_kernel.py:539-557): positions then
    // velocities, component by component
    const size_t b = (size_t)ks * 2 * D * A.nAst;
    for (int i = tid; i < A.nAst; i += stride)
#pragma unroll
      for (int c = 0; c < D; ++c) {
        A.snap[b + (size_t)c * A.nAst + i] = (float)A.ap[c][i];
        A.snap[b + (size_t)(D + c) * A.nAst + i] = (float)A.av[c][i];
      }
    if (blockIdx.x == 0 && threadIdx.x < M) {
      const size_t bm = (size_t)ks * 2 * D * M;
      const int j = threadIdx.x;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        A.snapM[bm + c * M + j] = (float)S.x[c][j];
        A.snapM[bm + (D + c) * M + j] = (float)S.v[c][j];
      }
    }
  }
  if (blockIdx.x == 0 && threadIdx.x < M) {
    const int i = threadIdx.x;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      A.mp[c][i] = S.x[c][i];
      A.mv[c][i] = S.v[c][i];
      A.macc[c][i] = S.acc[c][i];
    }
  }
}

template <int D>
int frame_blocks(int nAst) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_nbody_frame<D>, FB,
                                                   0) != hipSuccess ||
      per_cu < 1) {
    (void)hipGetLastError();
    per_cu = 1;
  }
  int dev = 0, cus = 1;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
      hipSuccess) {
    (void)hipGetLastError();
    cus = 1;
  }
  int want = ssim::ceil_div(nAst > 0 ? nAst : 1, FB);
  const int cap = per_cu * cus;
  return want < cap ? want : cap;
}

// workspace layout (hot flags, back-reaction partials) and the cooperative
// launch, shared by both dimensions
template <int D>
int launch_frame(FrameArgs<D>& A, void* work, void* stream) {
  unsigned char* w = static_cast<unsigned char*>(work);
  A.hot = w;
  A.part = reinterpret_cast<double*>(w + ssim::align256((size_t)(A.nAst > 0 ? A.nAst : 1)));
  const int blocks = frame_blocks<D>(A.nAst);
  void* params[] = {&A};
  SSIM_HIP(hipLaunchCooperativeKernel((const void*)k_nbody_frame<D>, dim3(blocks),
                                      dim3(FB), params, 0,
                                      ssim::as_stream(stream)));
  SSIM_LAUNCHED("k_nbody_frame");
  return SSIM_OK;
}

template <int D>
void set_scalars(FrameArgs<D>& A, int n_ast, const double* am,
                 const unsigned char* avis, const double* arad, int m,
                 const double* mm, const unsigned char* mvis, const double* mrad,
                 double G, double soft, double dt_raster, int k, double max_sub_dt,
                 int max_sub_steps, double w1, double w0, int m_sub, float* snap,
                 float* snap_m, float* sub_pos, unsigned char* sub_n, float* hit_t,
                 int* hit_m) {
  A.am = am; A.aVis = avis; A.aRad = arad;
  A.mm = mm; A.mVis = mvis; A.mRad = mrad;
  A.snap = snap; A.snapM = snap_m; A.subPos = sub_pos; A.hitT = hit_t;
  A.subN = sub_n; A.hitM = hit_m;
  A.nAst = n_ast; A.M = m; A.mSub = m_sub; A.K = k; A.maxSubSteps = max_sub_steps;
  A.G = G; A.soft = soft; A.dtRaster = dt_raster; A.maxSubDt = max_sub_dt;
  A.w1 = w1; A.w0 = w0;
}

}  // namespace

extern "C" {

size_t ssim_nbody_work_bytes(int n_ast) {
  // hot flags + back-reaction partials (3 components) for the largest
  // possible grid
  const int nb = ssim::ceil_div(n_ast > 0 ? n_ast : 1, FB);
  const size_t blocks = nb < 65536 ? (size_t)nb : 65536;
  return ssim::align256((size_t)(n_ast > 0 ? n_ast : 1)) +
         ssim::align256(2 * blocks * 3 * MMAX * sizeof(double));
}

#define FRAME_CHECKS(POS_OK, MPOS_OK)                                               \
  SSIM_REQUIRE(n_ast >= 0, "n_ast must be >= 0");                                  \
  SSIM_REQUIRE(m >= 0 && m <= MMAX, "zu viele massive Koerper: %d > %d", m, MMAX); \
  SSIM_REQUIRE(k >= 1, "k must be >= 1, got %d", k);                               \
  SSIM_REQUIRE(m_sub >= 0, "m_sub must be >= 0");                                  \
  SSIM_REQUIRE(n_ast == 0 || ((POS_OK) && am && avis && arad && snap && hit_t &&   \
                              hit_m),                                              \
               "null asteroid pointer");                                           \
  SSIM_REQUIRE(m == 0 || ((MPOS_OK) && mm && mvis && mrad && snap_m),              \
               "null massive pointer");                                            \
  SSIM_REQUIRE(m_sub == 0 || n_ast == 0 || (sub_pos && sub_n),                     \
               "sub-samples need sub_pos and sub_n");                              \
  SSIM_REQUIRE(work && work_bytes >= ssim_nbody_work_bytes(n_ast),                 \
               "workspace too small")

int ssim_nbody_frame(int n_ast, double* ax, double* ay, double* avx,
                     double* avy, double* aacc_x, double* aacc_y,
                     const double* am, const unsigned char* avis,
                     const double* arad, int m, double* mx, double* my,
                     double* mvx, double* mvy, double* macc_x, double* macc_y,
                     const double* mm, const unsigned char* mvis,
                     const double* mrad, double G, double soft,
                     double dt_raster, int k, double max_sub_dt,
                     int max_sub_steps, double w1, double w0, int m_sub,
                     float* snap, float* snap_m, float* sub_pos,
                     unsigned char* sub_n, float* hit_t, int* hit_m,
                     void* work, size_t work_bytes, void* stream) {
  FRAME_CHECKS(ax && ay && avx && avy && aacc_x && aacc_y,
               mx && my && mvx && mvy && macc_x && macc_y);
  FrameArgs<2> A;
  A.ap[0] = ax; A.ap[1] = ay; A.av[0] = avx; A.av[1] = avy;
  A.aacc[0] = aacc_x; A.aacc[1] = aacc_y;
  A.mp[0] = mx; A.mp[1] = my; A.mv[0] = mvx; A.mv[1] = mvy;
  A.macc[0] = macc_x; A.macc[1] = macc_y;
  set_scalars<2>(A, n_ast, am, avis, arad, m, mm, mvis, mrad, G, soft, dt_raster,
                 k, max_sub_dt, max_sub_steps, w1, w0, m_sub, snap, snap_m,
                 sub_pos, sub_n, hit_t, hit_m);
  return launch_frame<2>(A, work, stream);
}

int ssim_nbody_frame3d(int n_ast, double* ax, double* ay, double* az,
                       double* avx, double* avy, double* avz, double* aacc_x,
                       double* aacc_y, double* aacc_z, const double* am,
                       const unsigned char* avis, const double* arad, int m,
                       double* mx, double* my, double* mz, double* mvx,
                       double* mvy, double* mvz, double* macc_x, double* macc_y,
                       double* macc_z, const double* mm,
                       const unsigned char* mvis, const double* mrad, double G,
                       double soft, double dt_raster, int k, double max_sub_dt,
                       int max_sub_steps, double w1, double w0, int m_sub,
                       float* snap, float* snap_m, float* sub_pos,
                       unsigned char* sub_n, float* hit_t, int* hit_m,
                       void* work, size_t work_bytes, void* stream) {
  FRAME_CHECKS(ax && ay && az && avx && avy && avz && aacc_x && aacc_y && aacc_z,
               mx && my && mz && mvx && mvy && mvz && macc_x && macc_y && macc_z);
  FrameArgs<3> A;
  A.ap[0] = ax; A.ap[1] = ay; A.ap[2] = az;
  A.av[0] = avx; A.av[1] = avy; A.av[2] = avz;
  A.aacc[0] = aacc_x; A.aacc[1] = aacc_y; A.aacc[2] = aacc_z;
  A.mp[0] = mx; A.mp[1] = my; A.mp[2] = mz;
  A.mv[0] = mvx; A.mv[1] = mvy; A.mv[2] = mvz;
  A.macc[0] = macc_x; A.macc[1] = macc_y; A.macc[2] = macc_z;
  set_scalars<3>(A, n_ast, am, avis, arad, m, mm, mvis, mrad, G, soft, dt_raster,
                 k, max_sub_dt, max_sub_steps, w1, w0, m_sub, snap, snap_m,
                 sub_pos, sub_n, hit_t, hit_m);
  return launch_frame<3>(A, work, stream);
}

}  // extern "C"
