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

struct FrameArgs {
  double *ax, *ay, *avx, *avy, *aAccX, *aAccY;
  const double* am;
  const unsigned char* aVis;
  const double* aRad;
  double *mx, *my, *mvx, *mvy, *mAccX, *mAccY;
  const double* mm;
  const unsigned char* mVis;
  const double* mRad;
  unsigned char* hot;
  double* part;  // [2][gridDim][2][MMAX]
  float *snap, *snapM, *subPos, *hitT;
  unsigned char* subN;
  int* hitM;
  int nAst, M, mSub, K, maxSubSteps;
  double G, soft, dtRaster, maxSubDt, w1, w0;
};

struct Massive {
  double x[MMAX], y[MMAX], vx[MMAX], vy[MMAX], m[MMAX], accx[MMAX], accy[MMAX];
  double x0[MMAX], y0[MMAX], rad[MMAX];
  unsigned char vis[MMAX];
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Asteroid accelerations against the massive bodies in LDS and the
// back-reaction partial of this workgroup.  kick: 0 = only store the new
// acceleration (initial), 1 = quiet asteroids also get the velocity kick
// v += 0.5 (a_old + a_new) dt (phase B, nbody_kernel.py:336-368).
__device__ void accel_phase(const FrameArgs& A, const Massive& S,
                            double (*s_wave)[2][MMAX], int kick, double dt,
                            double* part_out) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int M = A.M;
  for (int k = threadIdx.x; k < NWAVE * 2 * MMAX; k += FB)
    (&s_wave[0][0][0])[k] = 0.0;
  __syncthreads();
  const int stride = gridDim.x * FB;
  for (int i0 = blockIdx.x * FB + wave * 64; i0 < A.nAst; i0 += stride) {
    const int i = i0 + lane;
    const bool on = i < A.nAst && A.aVis[i];
    double px = 0.0, py = 0.0, pm = 0.0;
    if (on) {
      px = A.ax[i];
      py = A.ay[i];
      pm = A.am[i];
    }
    double acx = 0.0, acy = 0.0;
    for (int kk = 0; kk < M; ++kk) {
      if (!S.vis[kk]) continue;
      double bx = 0.0, by = 0.0;
      if (on) {
        const double dx = S.x[kk] - px, dy = S.y[kk] - py;
        const double r2 = dx * dx + dy * dy + A.soft;
        const double f = A.G / (r2 * sqrt(r2));
        acx += f * S.m[kk] * dx;
        acy += f * S.m[kk] * dy;
        bx = -f * pm * dx;
        by = -f * pm * dy;
      }
      bx = wave_sum(bx);
      by = wave_sum(by);
      if (lane == 0) {
        s_wave[wave][0][kk] += bx;
        s_wave[wave][1][kk] += by;
      }
    }
    if (i < A.nAst) {
      if (!kick) {
        A.aAccX[i] = acx;
        A.aAccY[i] = acy;
      } else if (on && !A.hot[i]) {
        A.avx[i] += 0.5 * (A.aAccX[i] + acx) * dt;
        A.avy[i] += 0.5 * (A.aAccY[i] + acy) * dt;
        A.aAccX[i] = acx;
        A.aAccY[i] = acy;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < 2 * MMAX) {
    const int c = threadIdx.x / MMAX, kk = threadIdx.x % MMAX;
    double s = 0.0;
    for (int w = 0; w < NWAVE; ++w) s += s_wave[w][c][kk];
    part_out[(size_t)blockIdx.x * 2 * MMAX + c * MMAX + kk] = s;
  }
}

// Massive accelerations (back-reaction summed in workgroup order + massive
// pairs); kick != 0 applies v += 0.5 (a_old + a_new) dt to visible bodies
// (phase C, nbody_kernel.py:376-399).  Threads < M, result in LDS.
__device__ void massive_phase(const FrameArgs& A, Massive& S, const double* part,
                              int kick, double dt) {
  const int i = threadIdx.x;
  if (i >= A.M) return;
  double acx = 0.0, acy = 0.0;
  for (int b = 0; b < (int)gridDim.x; ++b) {
    acx += part[(size_t)b * 2 * MMAX + i];
    acy += part[(size_t)b * 2 * MMAX + MMAX + i];
  }
  if (S.vis[i]) {
    for (int j = 0; j < A.M; ++j) {
      if (j == i || !S.vis[j]) continue;
      const double dx = S.x[j] - S.x[i], dy = S.y[j] - S.y[i];
      const double r2 = dx * dx + dy * dy + A.soft;
      const double f = A.G / (r2 * sqrt(r2));
      acx += f * S.m[j] * dx;
      acy += f * S.m[j] * dy;
    }
    if (kick) {
      S.vx[i] += 0.5 * (S.accx[i] + acx) * dt;
      S.vy[i] += 0.5 * (S.accy[i] + acy) * dt;
    }
  }
  S.accx[i] = acx;
  S.accy[i] = acy;
}

__global__ __launch_bounds__(FB) void k_nbody_frame(FrameArgs A) {
  cg::grid_group grid = cg::this_grid();
  __shared__ Massive S;
  __shared__ double s_wave[NWAVE][2][MMAX];
  __shared__ double s_remaining, s_dtH;
  __shared__ int s_done, s_guard;
  const int M = A.M;
  const int tid = blockIdx.x * FB + threadIdx.x;
  const int stride = gridDim.x * FB;
  const size_t pstride = (size_t)gridDim.x * 2 * MMAX;

  if (threadIdx.x < M) {
    const int i = threadIdx.x;
    S.x[i] = A.mx[i];
    S.y[i] = A.my[i];
    S.vx[i] = A.mvx[i];
    S.vy[i] = A.mvy[i];
    S.m[i] = A.mm[i];
    S.vis[i] = A.mVis[i];
    S.rad[i] = A.mRad[i];
  }
  if (threadIdx.x == 0) s_guard = 0;
  __syncthreads();

  // ---- initial accelerations (nbody_kernel.py:151-216) ----
  int pb = 0;
  accel_phase(A, S, s_wave, 0, 0.0, A.part + pb * pstride);
  grid.sync();
  massive_phase(A, S, A.part + pb * pstride, 0, 0.0);
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
              const double dx = S.x[j] - S.x[i], dy = S.y[j] - S.y[i];
              double dist = sqrt(dx * dx + dy * dy);
              if (dist < 1e-12) dist = 1e-12;
              const double dvx = S.vx[j] - S.vx[i], dvy = S.vy[j] - S.vy[i];
              const double vrel = sqrt(dvx * dvx + dvy * dvy);
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
          const double px = A.ax[i], py = A.ay[i];
          const double pvx = A.avx[i], pvy = A.avy[i];
          for (int kk = 0; kk < M; ++kk) {
            if (!S.vis[kk]) continue;
            const double dx = S.x[kk] - px, dy = S.y[kk] - py;
            double dist = sqrt(dx * dx + dy * dy);
            if (dist < 1e-12) dist = 1e-12;
            const double dvx = S.vx[kk] - pvx, dvy = S.vy[kk] - pvy;
            const double vrel = sqrt(dvx * dvx + dvy * dvy);
            if (vrel > 1e-9 && dist / vrel / 20.0 < dtH) {
              h = 1;
              break;
            }
          }
        }
        A.hot[i] = h;
      }
      if (threadIdx.x < M) {
        S.x0[threadIdx.x] = S.x[threadIdx.x];
        S.y0[threadIdx.x] = S.y[threadIdx.x];
      }
      __syncthreads();

      // three Yoshida-weighted Verlet steps (nbody_kernel.py:314-401)
      for (int w = 0; w < 3; ++w) {
        const double dt = wd[w] * dtH;
        for (int i = tid; i < A.nAst; i += stride) {  // phase A
          if (!A.aVis[i] || A.hot[i]) continue;
          A.ax[i] += A.avx[i] * dt + 0.5 * A.aAccX[i] * dt * dt;
          A.ay[i] += A.avy[i] * dt + 0.5 * A.aAccY[i] * dt * dt;
        }
        if (threadIdx.x < M && S.vis[threadIdx.x]) {
          const int i = threadIdx.x;
          S.x[i] += S.vx[i] * dt + 0.5 * S.accx[i] * dt * dt;
          S.y[i] += S.vy[i] * dt + 0.5 * S.accy[i] * dt * dt;
        }
        __syncthreads();
        accel_phase(A, S, s_wave, 1, dt, A.part + pb * pstride);  // phase B
        grid.sync();
        massive_phase(A, S, A.part + pb * pstride, 1, dt);         // phase C
        pb ^= 1;
        __syncthreads();
      }

      // private fine loop of the hot asteroids (nbody_kernel.py:403-535)
      const double tSeg0 = A.dtRaster - s_remaining - dtH;
      for (int i = tid; i < A.nAst; i += stride) {
        if (!A.aVis[i] || !A.hot[i]) continue;
        double px = A.ax[i], py = A.ay[i];
        double pvx = A.avx[i], pvy = A.avy[i];
        const double sgn = dtH >= 0.0 ? 1.0 : -1.0;
        const double span = fabs(dtH);
        double tau = 0.0;
        int guardF = 0;
        while (tau < span - 1e-15 && guardF++ < 8000) {
          const double al = tau / span;
          double dtF = A.maxSubDt;
          for (int kk = 0; kk < M; ++kk) {
            if (!S.vis[kk]) continue;
            const double mxi = S.x0[kk] + (S.x[kk] - S.x0[kk]) * al;
            const double myi = S.y0[kk] + (S.y[kk] - S.y0[kk]) * al;
            const double dx = mxi - px, dy = myi - py;
            double dist = sqrt(dx * dx + dy * dy);
            if (dist < 1e-12) dist = 1e-12;
            if (A.hitT[i] < 0.0f && dist <= S.rad[kk] + A.aRad[i]) {
              A.hitT[i] = (float)(ks * A.dtRaster + tSeg0 + tau);
              A.hitM[i] = kk;
            }
            const double dvx = S.vx[kk] - pvx, dvy = S.vy[kk] - pvy;
            const double vrel = sqrt(dvx * dvx + dvy * dvy);
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
            const double a0 = tSub / span;
            double acx = 0.0, acy = 0.0;
            for (int kk = 0; kk < M; ++kk) {
              if (!S.vis[kk]) continue;
              const double mxi = S.x0[kk] + (S.x[kk] - S.x0[kk]) * a0;
              const double myi = S.y0[kk] + (S.y[kk] - S.y0[kk]) * a0;
              const double dx = mxi - px, dy = myi - py;
              const double r2 = dx * dx + dy * dy + A.soft;
              const double f = A.G / (r2 * sqrt(r2));
              acx += f * S.m[kk] * dx;
              acy += f * S.m[kk] * dy;
            }
            px += pvx * df + 0.5 * acx * df * df;
            py += pvy * df + 0.5 * acy * df * df;
            tSub += wd[wf] * dtF;
            const double a1 = tSub / span;
            double ncx = 0.0, ncy = 0.0;
            for (int kk = 0; kk < M; ++kk) {
              if (!S.vis[kk]) continue;
              const double mxi = S.x0[kk] + (S.x[kk] - S.x0[kk]) * a1;
              const double myi = S.y0[kk] + (S.y[kk] - S.y0[kk]) * a1;
              const double dx = mxi - px, dy = myi - py;
              const double r2 = dx * dx + dy * dy + A.soft;
              const double f = A.G / (r2 * sqrt(r2));
              ncx += f * S.m[kk] * dx;
              ncy += f * S.m[kk] * dy;
            }
            pvx += 0.5 * (acx + ncx) * df;
            pvy += 0.5 * (acy + ncy) * df;
          }
          tau += dtF;
          if (A.mSub > 0) {
            const double q = (tSeg0 + tau) / subDtRaster;
            const int j = (int)floor(q + 0.5);
            if (j >= 1 && j <= A.mSub && fabs(q - j) < 1e-6) {
              const size_t b = ((size_t)ks * A.mSub + (j - 1)) * 2 * A.nAst;
              A.subPos[b + i] = (float)px;
              A.subPos[b + A.nAst + i] = (float)py;
              A.subN[(size_t)ks * A.nAst + i]++;
            }
          }
        }
        A.ax[i] = px;
        A.ay[i] = py;
        A.avx[i] = pvx;
        A.avy[i] = pvy;
        double acx = 0.0, acy = 0.0;
        for (int kk = 0; kk < M; ++kk) {
          if (!S.vis[kk]) continue;
          const double dx = S.x[kk] - px, dy = S.y[kk] - py;
          const double r2 = dx * dx + dy * dy + A.soft;
          const double f = A.G / (r2 * sqrt(r2));
          acx += f * S.m[kk] * dx;
          acy += f * S.m[kk] * dy;
        }
        A.aAccX[i] = acx;
        A.aAccY[i] = acy;
      }
      __syncthreads();
    }

    // f32 snapshot of this raster (nbody_kernel.py:539-557)
    const size_t b = (size_t)ks * 4 * A.nAst;
    for (int i = tid; i < A.nAst; i += stride) {
      A.snap[b + i] = (float)A.ax[i];
      A.snap[b + A.nAst + i] = (float)A.ay[i];
      A.snap[b + 2 * (size_t)A.nAst + i] = (float)A.avx[i];
      A.snap[b + 3 * (size_t)A.nAst + i] = (float)A.avy[i];
    }
    if (blockIdx.x == 0 && threadIdx.x < M) {
      const size_t bm = (size_t)ks * 4 * M;
      const int j = threadIdx.x;
      A.snapM[bm + j] = (float)S.x[j];
      A.snapM[bm + M + j] = (float)S.y[j];
      A.snapM[bm + 2 * M + j] = (float)S.vx[j];
      A.snapM[bm + 3 * M + j] = (float)S.vy[j];
    }
  }
  if (blockIdx.x == 0 && threadIdx.x < M) {
    const int i = threadIdx.x;
    A.mx[i] = S.x[i];
    A.my[i] = S.y[i];
    A.mvx[i] = S.vx[i];
    A.mvy[i] = S.vy[i];
    A.mAccX[i] = S.accx[i];
    A.mAccY[i] = S.accy[i];
  }
}

int frame_blocks(int nAst) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_nbody_frame, FB,
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

}  // namespace

extern "C" {

size_t ssim_nbody_work_bytes(int n_ast) {
  // hot flags + back-reaction partials for the largest possible grid
  const int nb = ssim::ceil_div(n_ast > 0 ? n_ast : 1, FB);
  const size_t blocks = nb < 65536 ? (size_t)nb : 65536;
  return ssim::align256((size_t)(n_ast > 0 ? n_ast : 1)) +
         ssim::align256(2 * blocks * 2 * MMAX * sizeof(double));
}

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
  SSIM_REQUIRE(n_ast >= 0, "n_ast must be >= 0");
  SSIM_REQUIRE(m >= 0 && m <= MMAX, "zu viele massive Koerper: %d > %d", m, MMAX);
  SSIM_REQUIRE(k >= 1, "k must be >= 1, got %d", k);
  SSIM_REQUIRE(m_sub >= 0, "m_sub must be >= 0");
  SSIM_REQUIRE(n_ast == 0 || (ax && ay && avx && avy && aacc_x && aacc_y && am &&
                              avis && arad && snap && hit_t && hit_m),
               "null asteroid pointer");
  SSIM_REQUIRE(m == 0 || (mx && my && mvx && mvy && macc_x && macc_y && mm &&
                          mvis && mrad && snap_m),
               "null massive pointer");
  SSIM_REQUIRE(m_sub == 0 || n_ast == 0 || (sub_pos && sub_n),
               "sub-samples need sub_pos and sub_n");
  SSIM_REQUIRE(work && work_bytes >= ssim_nbody_work_bytes(n_ast),
               "workspace too small");
  FrameArgs A;
  A.ax = ax; A.ay = ay; A.avx = avx; A.avy = avy; A.aAccX = aacc_x; A.aAccY = aacc_y;
  A.am = am; A.aVis = avis; A.aRad = arad;
  A.mx = mx; A.my = my; A.mvx = mvx; A.mvy = mvy; A.mAccX = macc_x; A.mAccY = macc_y;
  A.mm = mm; A.mVis = mvis; A.mRad = mrad;
  unsigned char* w = static_cast<unsigned char*>(work);
  A.hot = w;
  A.part = reinterpret_cast<double*>(w + ssim::align256((size_t)(n_ast > 0 ? n_ast : 1)));
  A.snap = snap; A.snapM = snap_m; A.subPos = sub_pos; A.hitT = hit_t;
  A.subN = sub_n; A.hitM = hit_m;
  A.nAst = n_ast; A.M = m; A.mSub = m_sub; A.K = k; A.maxSubSteps = max_sub_steps;
  A.G = G; A.soft = soft; A.dtRaster = dt_raster; A.maxSubDt = max_sub_dt;
  A.w1 = w1; A.w0 = w0;
  const int blocks = frame_blocks(n_ast);
  void* params[] = {&A};
  SSIM_HIP(hipLaunchCooperativeKernel((const void*)k_nbody_frame, dim3(blocks),
                                      dim3(FB), params, 0,
                                      ssim::as_stream(stream)));
  SSIM_LAUNCHED("k_nbody_frame");
  return SSIM_OK;
}

}  // extern "C"
