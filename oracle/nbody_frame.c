/*
 * CPU oracle of the few-massive + many-asteroid raster integrator —
 * TEST INFRASTRUCTURE ONLY (see ssim_oracle.c for the rules).
 *
 * Serial, plain-double restatement of `frame_kernel` (nbody_kernel.py:
 * 97-560) for one device: initial accelerations with back-reaction
 * (151-216), per raster the coarse step from massive pairs (245-279), hot
 * classification (283-301), three Yoshida-weighted Verlets with phases A/B/C
 * (314-401), the private fine loop with interpolated massive positions,
 * contact time and sub-sample landing (403-535), and the f32 snapshots
 * (539-557).  Back-reaction sums run over asteroids in index order.
 *
 * Pinned by tests/test_nbody_oracle.py against the reference's NumPy
 * `reference_advance` (test_kernel.py:16-81) within the reference's own
 * 5e-6 tolerance, its diver test (test_kernel.py:190-219), batch == single
 * steps, frozen bodies, and a closed-form circular orbit.
 *
 * The same scheme in space (D = 3, `oracle_nbody_frame3d`) adds a z
 * component to every vector (|d|^2 summed x, y, z); the plane case keeps
 * the reference's operations exactly.  The 3-D form is pinned by
 * tests/test_nbody3d_oracle.py: a planar scene (z = vz = 0) gives the 2-D
 * results bit for bit, a planar scene rotated into a tilted plane gives the
 * rotated 2-D results (a dropped or misplaced z term in forces, encounter
 * times, contact distances or sub-samples breaks this), a circular inclined
 * orbit follows its closed form, and the 3-D form of the reference's NumPy
 * `reference_advance` agrees within the reference's 5e-6.
 */
#include <math.h>
#include <stddef.h>

#define MMAX 64

/* |d|^2 summed x, y, (z) left to right */
static double sq_sum(int D, const double* d) {
  double r = d[0] * d[0];
  for (int c = 1; c < D; ++c) r = r + d[c] * d[c];
  return r;
}

/* One device's frame_kernel for D = 2 or 3.  a/av/aac: D component arrays
 * of the asteroids' positions, velocities, accelerations; mp/mv/mac the
 * massive bodies'.  snap [K][2D][nAst], snapM [K][2D][M] (positions then
 * velocities), subPos [K][mSub][D][nAst]. */
static void frame(int D, int nAst, double* const* a, double* const* av,
                  double* const* aac, const double* am, const unsigned char* aVis,
                  const double* aRad, int M, double* const* mp, double* const* mv,
                  double* const* mac, const double* mm, const unsigned char* mVis,
                  const double* mRad, double G, double soft, double dtRaster,
                  int K, double maxSubDt, int maxSubSteps, double w1, double w0,
                  int mSub, float* snap, float* snapM, float* subPos,
                  unsigned char* subN, float* hitT, int* hitM, unsigned char* hot) {
  double back[3][MMAX], m0[3][MMAX];
  int guard = 0;
  /* initial accelerations (nbody_kernel.py:151-216) */
  for (int c = 0; c < D; ++c)
    for (int j = 0; j < M; ++j) back[c][j] = 0.0;
  for (int i = 0; i < nAst; ++i) {
    double ac[3] = {0.0, 0.0, 0.0};
    if (aVis[i]) {
      for (int kk = 0; kk < M; ++kk) {
        if (!mVis[kk]) continue;
        double d[3];
        for (int c = 0; c < D; ++c) d[c] = mp[c][kk] - a[c][i];
        const double r2 = sq_sum(D, d) + soft;
        const double f = G / (r2 * sqrt(r2));
        for (int c = 0; c < D; ++c) {
          ac[c] += f * mm[kk] * d[c];
          back[c][kk] += -f * am[i] * d[c];
        }
      }
    }
    for (int c = 0; c < D; ++c) aac[c][i] = ac[c];
  }
  for (int i = 0; i < M; ++i) {
    double ac[3];
    for (int c = 0; c < D; ++c) ac[c] = back[c][i];
    if (mVis[i])
      for (int j = 0; j < M; ++j) {
        if (j == i || !mVis[j]) continue;
        double d[3];
        for (int c = 0; c < D; ++c) d[c] = mp[c][j] - mp[c][i];
        const double r2 = sq_sum(D, d) + soft;
        const double f = G / (r2 * sqrt(r2));
        for (int c = 0; c < D; ++c) ac[c] += f * mm[j] * d[c];
      }
    for (int c = 0; c < D; ++c) mac[c][i] = ac[c];
  }
  const double wd[3] = {w1, w0, w1};
  const double subDt = mSub > 0 ? dtRaster / mSub : 0.0;
  for (int ks = 0; ks < K; ++ks) {
    double remaining = dtRaster;
    if (mSub > 0)
      for (int i = 0; i < nAst; ++i) subN[(size_t)ks * nAst + i] = 0;
    for (;;) {
      /* coarse step from the massive pairs (245-279) */
      if (remaining <= 1e-12 || guard++ >= maxSubSteps) break;
      double dtH = maxSubDt;
      for (int i = 0; i < M; ++i) {
        if (!mVis[i]) continue;
        for (int j = i + 1; j < M; ++j) {
          if (!mVis[j]) continue;
          double d[3], dv[3];
          for (int c = 0; c < D; ++c) {
            d[c] = mp[c][j] - mp[c][i];
            dv[c] = mv[c][j] - mv[c][i];
          }
          double dist = sqrt(sq_sum(D, d));
          if (dist < 1e-12) dist = 1e-12;
          const double vrel = sqrt(sq_sum(D, dv));
          if (vrel > 1e-9) {
            const double tEnc = dist / vrel / 20.0;
            if (tEnc < dtH) dtH = tEnc;
          }
        }
      }
      if (dtH < maxSubDt / 1000.0) dtH = maxSubDt / 1000.0;
      if (dtH > remaining) dtH = remaining;
      remaining -= dtH;

      /* hot classification (283-301) */
      for (int i = 0; i < nAst; ++i) {
        unsigned char h = 0;
        if (aVis[i])
          for (int kk = 0; kk < M; ++kk) {
            if (!mVis[kk]) continue;
            double d[3], dv[3];
            for (int c = 0; c < D; ++c) {
              d[c] = mp[c][kk] - a[c][i];
              dv[c] = mv[c][kk] - av[c][i];
            }
            double dist = sqrt(sq_sum(D, d));
            if (dist < 1e-12) dist = 1e-12;
            const double vrel = sqrt(sq_sum(D, dv));
            if (vrel > 1e-9 && dist / vrel / 20.0 < dtH) {
              h = 1;
              break;
            }
          }
        hot[i] = h;
      }
      for (int c = 0; c < D; ++c)
        for (int j = 0; j < M; ++j) m0[c][j] = mp[c][j];
      /* three Yoshida-weighted Verlet steps (314-401) */
      for (int w = 0; w < 3; ++w) {
        const double dt = wd[w] * dtH;
        for (int i = 0; i < nAst; ++i) {
          if (!aVis[i] || hot[i]) continue;
          for (int c = 0; c < D; ++c)
            a[c][i] += av[c][i] * dt + 0.5 * aac[c][i] * dt * dt;
        }
        for (int i = 0; i < M; ++i) {
          if (!mVis[i]) continue;
          for (int c = 0; c < D; ++c)
            mp[c][i] += mv[c][i] * dt + 0.5 * mac[c][i] * dt * dt;
        }
        for (int c = 0; c < D; ++c)
          for (int j = 0; j < M; ++j) back[c][j] = 0.0;
        for (int i = 0; i < nAst; ++i) {
          if (!aVis[i]) continue;
          double ac[3] = {0.0, 0.0, 0.0};
          for (int kk = 0; kk < M; ++kk) {
            if (!mVis[kk]) continue;
            double d[3];
            for (int c = 0; c < D; ++c) d[c] = mp[c][kk] - a[c][i];
            const double r2 = sq_sum(D, d) + soft;
            const double f = G / (r2 * sqrt(r2));
            for (int c = 0; c < D; ++c) {
              ac[c] += f * mm[kk] * d[c];
              back[c][kk] += -f * am[i] * d[c];
            }
          }
          if (!hot[i])
            for (int c = 0; c < D; ++c) {
              av[c][i] += 0.5 * (aac[c][i] + ac[c]) * dt;
              aac[c][i] = ac[c];
            }
        }
        for (int i = 0; i < M; ++i) {
          double ac[3];
          for (int c = 0; c < D; ++c) ac[c] = back[c][i];
          if (mVis[i]) {
            for (int j = 0; j < M; ++j) {
              if (j == i || !mVis[j]) continue;
              double d[3];
              for (int c = 0; c < D; ++c) d[c] = mp[c][j] - mp[c][i];
              const double r2 = sq_sum(D, d) + soft;
              const double f = G / (r2 * sqrt(r2));
              for (int c = 0; c < D; ++c) ac[c] += f * mm[j] * d[c];
            }
            for (int c = 0; c < D; ++c) mv[c][i] += 0.5 * (mac[c][i] + ac[c]) * dt;
          }
          for (int c = 0; c < D; ++c) mac[c][i] = ac[c];
        }
      }
      /* private fine loop of the hot asteroids (403-535) */
      const double tSeg0 = dtRaster - remaining - dtH;
      for (int i = 0; i < nAst; ++i) {
        if (!aVis[i] || !hot[i]) continue;
        double p[3], pv[3];
        for (int c = 0; c < D; ++c) {
          p[c] = a[c][i];
          pv[c] = av[c][i];
        }
        const double sgn = dtH >= 0.0 ? 1.0 : -1.0;
        const double span = fabs(dtH);
        double tau = 0.0;
        int guardF = 0;
        while (tau < span - 1e-15 && guardF++ < 8000) {
          const double al = tau / span;
          double dtF = maxSubDt;
          for (int kk = 0; kk < M; ++kk) {
            if (!mVis[kk]) continue;
            double d[3], dv[3];
            for (int c = 0; c < D; ++c) {
              const double mi = m0[c][kk] + (mp[c][kk] - m0[c][kk]) * al;
              d[c] = mi - p[c];
              dv[c] = mv[c][kk] - pv[c];
            }
            double dist = sqrt(sq_sum(D, d));
            if (dist < 1e-12) dist = 1e-12;
            if (hitT[i] < 0.0f && dist <= mRad[kk] + aRad[i]) {
              hitT[i] = (float)(ks * dtRaster + tSeg0 + tau);
              hitM[i] = kk;
            }
            const double vrel = sqrt(sq_sum(D, dv));
            if (vrel > 1e-9) {
              const double tE = dist / vrel / 20.0;
              if (tE < dtF) dtF = tE;
            }
          }
          if (dtF < maxSubDt / 1000.0) dtF = maxSubDt / 1000.0;
          if (dtF > span - tau) dtF = span - tau;
          if (mSub > 0) {
            const double tAbs = tSeg0 + tau;
            const int jn = (int)floor(tAbs / subDt + 1e-9) + 1;
            if (jn <= mSub) {
              const double rest = jn * subDt - tAbs;
              if (rest > 0.0 && dtF > rest) dtF = rest;
            }
          }
          double tSub = tau;
          for (int wf = 0; wf < 3; ++wf) {
            const double df = wd[wf] * dtF * sgn;
            double ac[3] = {0.0, 0.0, 0.0}, nc[3] = {0.0, 0.0, 0.0};
            for (int pass = 0; pass < 2; ++pass) {
              /* pass 0: a at tSub before the drift; pass 1: after it */
              const double al2 = tSub / span;
              double* acc = pass == 0 ? ac : nc;
              for (int kk = 0; kk < M; ++kk) {
                if (!mVis[kk]) continue;
                double d[3];
                for (int c = 0; c < D; ++c) {
                  const double mi = m0[c][kk] + (mp[c][kk] - m0[c][kk]) * al2;
                  d[c] = mi - p[c];
                }
                const double r2 = sq_sum(D, d) + soft;
                const double f = G / (r2 * sqrt(r2));
                for (int c = 0; c < D; ++c) acc[c] += f * mm[kk] * d[c];
              }
              if (pass == 0) {
                for (int c = 0; c < D; ++c) p[c] += pv[c] * df + 0.5 * ac[c] * df * df;
                tSub += wd[wf] * dtF;
              }
            }
            for (int c = 0; c < D; ++c) pv[c] += 0.5 * (ac[c] + nc[c]) * df;
          }
          tau += dtF;
          if (mSub > 0) {
            const double q = (tSeg0 + tau) / subDt;
            const int j = (int)floor(q + 0.5);
            if (j >= 1 && j <= mSub && fabs(q - j) < 1e-6) {
              const size_t b = ((size_t)ks * mSub + (j - 1)) * D * nAst;
              for (int c = 0; c < D; ++c) subPos[b + (size_t)c * nAst + i] = (float)p[c];
              subN[(size_t)ks * nAst + i]++;
            }
          }
        }
        double ac[3] = {0.0, 0.0, 0.0};
        for (int c = 0; c < D; ++c) {
          a[c][i] = p[c];
          av[c][i] = pv[c];
        }
        for (int kk = 0; kk < M; ++kk) {
          if (!mVis[kk]) continue;
          double d[3];
          for (int c = 0; c < D; ++c) d[c] = mp[c][kk] - p[c];
          const double r2 = sq_sum(D, d) + soft;
          const double f = G / (r2 * sqrt(r2));
          for (int c = 0; c < D; ++c) ac[c] += f * mm[kk] * d[c];
        }
        for (int c = 0; c < D; ++c) aac[c][i] = ac[c];
      }
    }
    /* f32 snapshots (539-557) */
    const size_t b = (size_t)ks * 2 * D * nAst;
    for (int i = 0; i < nAst; ++i)
      for (int c = 0; c < D; ++c) {
        snap[b + (size_t)c * nAst + i] = (float)a[c][i];
        snap[b + (size_t)(D + c) * nAst + i] = (float)av[c][i];
      }
    const size_t bm = (size_t)ks * 2 * D * M;
    for (int j = 0; j < M; ++j)
      for (int c = 0; c < D; ++c) {
        snapM[bm + (size_t)c * M + j] = (float)mp[c][j];
        snapM[bm + (size_t)(D + c) * M + j] = (float)mv[c][j];
      }
  }
}

void oracle_nbody_frame(int nAst, double* ax, double* ay, double* avx,
                        double* avy, double* aAccX, double* aAccY,
                        const double* am, const unsigned char* aVis,
                        const double* aRad, int M, double* mx, double* my,
                        double* mvx, double* mvy, double* mAccX, double* mAccY,
                        const double* mm, const unsigned char* mVis,
                        const double* mRad, double G, double soft,
                        double dtRaster, int K, double maxSubDt,
                        int maxSubSteps, double w1, double w0, int mSub,
                        float* snap, float* snapM, float* subPos,
                        unsigned char* subN, float* hitT, int* hitM,
                        unsigned char* hot) {
  double* a[2] = {ax, ay};
  double* av[2] = {avx, avy};
  double* aac[2] = {aAccX, aAccY};
  double* mp[2] = {mx, my};
  double* mv[2] = {mvx, mvy};
  double* mac[2] = {mAccX, mAccY};
  frame(2, nAst, a, av, aac, am, aVis, aRad, M, mp, mv, mac, mm, mVis, mRad, G,
        soft, dtRaster, K, maxSubDt, maxSubSteps, w1, w0, mSub, snap, snapM,
        subPos, subN, hitT, hitM, hot);
}

void oracle_nbody_frame3d(int nAst, double* ax, double* ay, double* az,
                          double* avx, double* avy, double* avz, double* aAccX,
                          double* aAccY, double* aAccZ, const double* am,
                          const unsigned char* aVis, const double* aRad, int M,
                          double* mx, double* my, double* mz, double* mvx,
                          double* mvy, double* mvz, double* mAccX, double* mAccY,
                          double* mAccZ, const double* mm,
                          const unsigned char* mVis, const double* mRad,
                          double G, double soft, double dtRaster, int K,
                          double maxSubDt, int maxSubSteps, double w1, double w0,
                          int mSub, float* snap, float* snapM, float* subPos,
                          unsigned char* subN, float* hitT, int* hitM,
                          unsigned char* hot) {
  double* a[3] = {ax, ay, az};
  double* av[3] = {avx, avy, avz};
  double* aac[3] = {aAccX, aAccY, aAccZ};
  double* mp[3] = {mx, my, mz};
  double* mv[3] = {mvx, mvy, mvz};
  double* mac[3] = {mAccX, mAccY, mAccZ};
  frame(3, nAst, a, av, aac, am, aVis, aRad, M, mp, mv, mac, mm, mVis, mRad, G,
        soft, dtRaster, K, maxSubDt, maxSubSteps, w1, w0, mSub, snap, snapM,
        subPos, subN, hitT, hitM, hot);
}
