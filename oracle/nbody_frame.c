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
 */
#include <math.h>
#include <stddef.h>

#define MMAX 64

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
  double backX[MMAX], backY[MMAX], mx0[MMAX], my0[MMAX];
  int guard = 0;
  /* initial accelerations */
  for (int j = 0; j < M; ++j) backX[j] = backY[j] = 0.0;
  for (int i = 0; i < nAst; ++i) {
    double acx = 0.0, acy = 0.0;
    if (aVis[i]) {
      for (int kk = 0; kk < M; ++kk) {
        if (!mVis[kk]) continue;
        const double dx = mx[kk] - ax[i], dy = my[kk] - ay[i];
        const double r2 = dx * dx + dy * dy + soft;
        const double f = G / (r2 * sqrt(r2));
        acx += f * mm[kk] * dx;
        acy += f * mm[kk] * dy;
        backX[kk] += -f * am[i] * dx;
        backY[kk] += -f * am[i] * dy;
      }
    }
    aAccX[i] = acx;
    aAccY[i] = acy;
  }
  for (int i = 0; i < M; ++i) {
    double acx = backX[i], acy = backY[i];
    if (mVis[i])
      for (int j = 0; j < M; ++j) {
        if (j == i || !mVis[j]) continue;
        const double dx = mx[j] - mx[i], dy = my[j] - my[i];
        const double r2 = dx * dx + dy * dy + soft;
        const double f = G / (r2 * sqrt(r2));
        acx += f * mm[j] * dx;
        acy += f * mm[j] * dy;
      }
    mAccX[i] = acx;
    mAccY[i] = acy;
  }
  const double wd[3] = {w1, w0, w1};
  const double subDt = mSub > 0 ? dtRaster / mSub : 0.0;
  for (int ks = 0; ks < K; ++ks) {
    double remaining = dtRaster;
    if (mSub > 0)
      for (int i = 0; i < nAst; ++i) subN[(size_t)ks * nAst + i] = 0;
    for (;;) {
      if (remaining <= 1e-12 || guard++ >= maxSubSteps) break;
      double dtH = maxSubDt;
      for (int i = 0; i < M; ++i) {
        if (!mVis[i]) continue;
        for (int j = i + 1; j < M; ++j) {
          if (!mVis[j]) continue;
          const double dx = mx[j] - mx[i], dy = my[j] - my[i];
          double dist = sqrt(dx * dx + dy * dy);
          if (dist < 1e-12) dist = 1e-12;
          const double dvx = mvx[j] - mvx[i], dvy = mvy[j] - mvy[i];
          const double vrel = sqrt(dvx * dvx + dvy * dvy);
          if (vrel > 1e-9) {
            const double tEnc = dist / vrel / 20.0;
            if (tEnc < dtH) dtH = tEnc;
          }
        }
      }
      if (dtH < maxSubDt / 1000.0) dtH = maxSubDt / 1000.0;
      if (dtH > remaining) dtH = remaining;
      remaining -= dtH;

      for (int i = 0; i < nAst; ++i) {
        unsigned char h = 0;
        if (aVis[i])
          for (int kk = 0; kk < M; ++kk) {
            if (!mVis[kk]) continue;
            const double dx = mx[kk] - ax[i], dy = my[kk] - ay[i];
            double dist = sqrt(dx * dx + dy * dy);
            if (dist < 1e-12) dist = 1e-12;
            const double dvx = mvx[kk] - avx[i], dvy = mvy[kk] - avy[i];
            const double vrel = sqrt(dvx * dvx + dvy * dvy);
            if (vrel > 1e-9 && dist / vrel / 20.0 < dtH) {
              h = 1;
              break;
            }
          }
        hot[i] = h;
      }
      for (int j = 0; j < M; ++j) {
        mx0[j] = mx[j];
        my0[j] = my[j];
      }
      for (int w = 0; w < 3; ++w) {
        const double dt = wd[w] * dtH;
        for (int i = 0; i < nAst; ++i) {
          if (!aVis[i] || hot[i]) continue;
          ax[i] += avx[i] * dt + 0.5 * aAccX[i] * dt * dt;
          ay[i] += avy[i] * dt + 0.5 * aAccY[i] * dt * dt;
        }
        for (int i = 0; i < M; ++i) {
          if (!mVis[i]) continue;
          mx[i] += mvx[i] * dt + 0.5 * mAccX[i] * dt * dt;
          my[i] += mvy[i] * dt + 0.5 * mAccY[i] * dt * dt;
        }
        for (int j = 0; j < M; ++j) backX[j] = backY[j] = 0.0;
        for (int i = 0; i < nAst; ++i) {
          if (!aVis[i]) continue;
          double acx = 0.0, acy = 0.0;
          for (int kk = 0; kk < M; ++kk) {
            if (!mVis[kk]) continue;
            const double dx = mx[kk] - ax[i], dy = my[kk] - ay[i];
            const double r2 = dx * dx + dy * dy + soft;
            const double f = G / (r2 * sqrt(r2));
            acx += f * mm[kk] * dx;
            acy += f * mm[kk] * dy;
            backX[kk] += -f * am[i] * dx;
            backY[kk] += -f * am[i] * dy;
          }
          if (!hot[i]) {
            avx[i] += 0.5 * (aAccX[i] + acx) * dt;
            avy[i] += 0.5 * (aAccY[i] + acy) * dt;
            aAccX[i] = acx;
            aAccY[i] = acy;
          }
        }
        for (int i = 0; i < M; ++i) {
          double acx = backX[i], acy = backY[i];
          if (mVis[i]) {
            for (int j = 0; j < M; ++j) {
              if (j == i || !mVis[j]) continue;
              const double dx = mx[j] - mx[i], dy = my[j] - my[i];
              const double r2 = dx * dx + dy * dy + soft;
              const double f = G / (r2 * sqrt(r2));
              acx += f * mm[j] * dx;
              acy += f * mm[j] * dy;
            }
            mvx[i] += 0.5 * (mAccX[i] + acx) * dt;
            mvy[i] += 0.5 * (mAccY[i] + acy) * dt;
          }
          mAccX[i] = acx;
          mAccY[i] = acy;
        }
      }
      const double tSeg0 = dtRaster - remaining - dtH;
      for (int i = 0; i < nAst; ++i) {
        if (!aVis[i] || !hot[i]) continue;
        double px = ax[i], py = ay[i], pvx = avx[i], pvy = avy[i];
        const double sgn = dtH >= 0.0 ? 1.0 : -1.0;
        const double span = fabs(dtH);
        double tau = 0.0;
        int guardF = 0;
        while (tau < span - 1e-15 && guardF++ < 8000) {
          const double al = tau / span;
          double dtF = maxSubDt;
          for (int kk = 0; kk < M; ++kk) {
            if (!mVis[kk]) continue;
            const double mxi = mx0[kk] + (mx[kk] - mx0[kk]) * al;
            const double myi = my0[kk] + (my[kk] - my0[kk]) * al;
            const double dx = mxi - px, dy = myi - py;
            double dist = sqrt(dx * dx + dy * dy);
            if (dist < 1e-12) dist = 1e-12;
            if (hitT[i] < 0.0f && dist <= mRad[kk] + aRad[i]) {
              hitT[i] = (float)(ks * dtRaster + tSeg0 + tau);
              hitM[i] = kk;
            }
            const double dvx = mvx[kk] - pvx, dvy = mvy[kk] - pvy;
            const double vrel = sqrt(dvx * dvx + dvy * dvy);
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
            const double a0 = tSub / span;
            double acx = 0.0, acy = 0.0;
            for (int kk = 0; kk < M; ++kk) {
              if (!mVis[kk]) continue;
              const double mxi = mx0[kk] + (mx[kk] - mx0[kk]) * a0;
              const double myi = my0[kk] + (my[kk] - my0[kk]) * a0;
              const double dx = mxi - px, dy = myi - py;
              const double r2 = dx * dx + dy * dy + soft;
              const double f = G / (r2 * sqrt(r2));
              acx += f * mm[kk] * dx;
              acy += f * mm[kk] * dy;
            }
            px += pvx * df + 0.5 * acx * df * df;
            py += pvy * df + 0.5 * acy * df * df;
            tSub += wd[wf] * dtF;
            const double a1 = tSub / span;
            double ncx = 0.0, ncy = 0.0;
            for (int kk = 0; kk < M; ++kk) {
              if (!mVis[kk]) continue;
              const double mxi = mx0[kk] + (mx[kk] - mx0[kk]) * a1;
              const double myi = my0[kk] + (my[kk] - my0[kk]) * a1;
              const double dx = mxi - px, dy = myi - py;
              const double r2 = dx * dx + dy * dy + soft;
              const double f = G / (r2 * sqrt(r2));
              ncx += f * mm[kk] * dx;
              ncy += f * mm[kk] * dy;
            }
            pvx += 0.5 * (acx + ncx) * df;
            pvy += 0.5 * (acy + ncy) * df;
          }
          tau += dtF;
          if (mSub > 0) {
            const double q = (tSeg0 + tau) / subDt;
            const int j = (int)floor(q + 0.5);
            if (j >= 1 && j <= mSub && fabs(q - j) < 1e-6) {
              const size_t b = ((size_t)ks * mSub + (j - 1)) * 2 * nAst;
              subPos[b + i] = (float)px;
              subPos[b + nAst + i] = (float)py;
              subN[(size_t)ks * nAst + i]++;
            }
          }
        }
        ax[i] = px;
        ay[i] = py;
        avx[i] = pvx;
        avy[i] = pvy;
        double acx = 0.0, acy = 0.0;
        for (int kk = 0; kk < M; ++kk) {
          if (!mVis[kk]) continue;
          const double dx = mx[kk] - px, dy = my[kk] - py;
          const double r2 = dx * dx + dy * dy + soft;
          const double f = G / (r2 * sqrt(r2));
          acx += f * mm[kk] * dx;
          acy += f * mm[kk] * dy;
        }
        aAccX[i] = acx;
        aAccY[i] = acy;
      }
    }
    const size_t b = (size_t)ks * 4 * nAst;
    for (int i = 0; i < nAst; ++i) {
      snap[b + i] = (float)ax[i];
      snap[b + nAst + i] = (float)ay[i];
      snap[b + 2 * (size_t)nAst + i] = (float)avx[i];
      snap[b + 3 * (size_t)nAst + i] = (float)avy[i];
    }
    const size_t bm = (size_t)ks * 4 * M;
    for (int j = 0; j < M; ++j) {
      snapM[bm + j] = (float)mx[j];
      snapM[bm + M + j] = (float)my[j];
      snapM[bm + 2 * M + j] = (float)mvx[j];
      snapM[bm + 3 * M + j] = (float)mvy[j];
    }
  }
}
