/*
 * CPU oracle of the GPU physics backend — TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker or the timed CPU baseline.
 * The product (solarsystemsimulator_amd) never calls it.
 *
 * Plain C restatement of the reference's FP64 algorithms:
 *  - oracle_allpairs: the softened all-pairs sum of `beschleunigung`
 *    (selfgrav_kernel.py:222-260) and `allpairs_accel` (test_pm.py:32-41),
 *    a_i = G sum_j m_j (x_j - x_i) / (|x_j - x_i|^2 + eps2)^(3/2).
 *    Every term is formed in long double (x87, 64-bit mantissa) and summed
 *    with Neumaier compensation in long double, so the result is the
 *    correctly rounded double up to a few units in the 19th digit — a
 *    reference for the GPU's 1e-12 FP64 gate, not a copy of its arithmetic.
 *  - oracle_energy: `energie` (test_selfgrav.py:54-63), same precision.
 *  - oracle_selfgrav_run: the leapfrog KDK of `referenz`
 *    (test_selfgrav.py:66-88) and of `selfgrav_kernel` (selfgrav_kernel.py:
 *    333-417), with masses first and tracers after them, tracers feeling the
 *    masses only; scheme 1 composes three KDK sub-steps with the Yoshida
 *    weights w1, w0, w1 (nbody_kernel.py:64-65).  Velocity/position updates
 *    are plain double `v = v + a*h` (no FMA; built with -ffp-contract=off).
 *  - oracle_allpairs_f64_plain: the same sum in plain double arithmetic with
 *    left-to-right accumulation (what a naive port would compute); used as
 *    the timed CPU baseline ("port"), never as a reference.
 *
 * Pinned by tests/test_oracle.py against closed forms (two-body force and
 * energy, symmetric configurations, Kepler orbit closure, 4th-order error
 * scaling of the Yoshida weights), against exact Decimal brute force on
 * small N, against a NumPy restatement of the reference's `referenz`, and
 * against the reference's own test thresholds (test_selfgrav.py A, B, F).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>

typedef long double ld;

static void neu_add(ld* s, ld* c, ld x) {
  ld t = *s + x;
  if (fabsl(*s) >= fabsl(x))
    *c += (*s - t) + x;
  else
    *c += (x - t) + *s;
  *s = t;
}

void oracle_allpairs(int dim, int ns, const double* sx, const double* sy,
                     const double* sz, const double* sm, int nt,
                     const double* tx, const double* ty, const double* tz,
                     double* ax, double* ay, double* az, double G,
                     double eps2) {
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < nt; ++i) {
    ld s[3] = {0, 0, 0}, c[3] = {0, 0, 0};
    const ld px = tx[i], py = ty[i], pz = dim == 3 ? (ld)tz[i] : 0.0L;
    for (int j = 0; j < ns; ++j) {
      const ld dx = (ld)sx[j] - px;
      const ld dy = (ld)sy[j] - py;
      const ld dz = dim == 3 ? (ld)sz[j] - pz : 0.0L;
      const ld r2 = dx * dx + dy * dy + dz * dz + (ld)eps2;
      const ld f = (ld)sm[j] / (r2 * sqrtl(r2));
      neu_add(&s[0], &c[0], f * dx);
      neu_add(&s[1], &c[1], f * dy);
      if (dim == 3) neu_add(&s[2], &c[2], f * dz);
    }
    ax[i] = (double)((ld)G * (s[0] + c[0]));
    ay[i] = (double)((ld)G * (s[1] + c[1]));
    if (dim == 3) az[i] = (double)((ld)G * (s[2] + c[2]));
  }
}

void oracle_allpairs_f64_plain(int dim, int ns, const double* sx,
                               const double* sy, const double* sz,
                               const double* sm, int nt, const double* tx,
                               const double* ty, const double* tz, double* ax,
                               double* ay, double* az, double G, double eps2) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nt; ++i) {
    double s0 = 0, s1 = 0, s2 = 0;
    const double px = tx[i], py = ty[i], pz = dim == 3 ? tz[i] : 0.0;
    for (int j = 0; j < ns; ++j) {
      const double dx = sx[j] - px, dy = sy[j] - py;
      const double dz = dim == 3 ? sz[j] - pz : 0.0;
      const double r2 = dx * dx + dy * dy + dz * dz + eps2;
      const double f = G * sm[j] / (r2 * sqrt(r2));
      s0 += f * dx;
      s1 += f * dy;
      s2 += f * dz;
    }
    ax[i] = s0;
    ay[i] = s1;
    if (dim == 3) az[i] = s2;
  }
}

/* out[0] = E_kin, out[1] = E_pot, out[2] = E over the n masses. */
void oracle_energy(int dim, int n, const double* x, const double* y,
                   const double* z, const double* vx, const double* vy,
                   const double* vz, const double* m, double G, double eps2,
                   double* out) {
  ld ks = 0, kc = 0, ps = 0, pc = 0;
  for (int i = 0; i < n; ++i) {
    ld v2 = (ld)vx[i] * vx[i] + (ld)vy[i] * vy[i];
    if (dim == 3) v2 += (ld)vz[i] * vz[i];
    neu_add(&ks, &kc, 0.5L * (ld)m[i] * v2);
  }
#pragma omp parallel
  {
    ld lps = 0, lpc = 0;
#pragma omp for schedule(dynamic, 16)
    for (int i = 0; i < n; ++i) {
      ld s = 0, c = 0;
      for (int j = 0; j < n; ++j) {
        if (j == i) continue;
        const ld dx = (ld)x[j] - x[i], dy = (ld)y[j] - y[i];
        const ld dz = dim == 3 ? (ld)z[j] - z[i] : 0.0L;
        neu_add(&s, &c, (ld)m[j] / sqrtl(dx * dx + dy * dy + dz * dz + (ld)eps2));
      }
      neu_add(&lps, &lpc, (ld)m[i] * (s + c));
    }
#pragma omp critical
    {
      neu_add(&ps, &pc, lps);
      pc += lpc;
    }
  }
  const ld ek = ks + kc;
  const ld ep = -0.5L * (ld)G * (ps + pc);
  out[0] = (double)ek;
  out[1] = (double)ep;
  out[2] = (double)(ek + ep);
}

static void accel_all(int dim, int n, int t, const double* x, const double* y,
                      const double* z, const double* m, double* ax, double* ay,
                      double* az, double G, double eps2) {
  oracle_allpairs(dim, n, x, y, z, m, n + t, x, y, z, ax, ay, az, G, eps2);
}

void oracle_selfgrav_run(int dim, int scheme, int n, int t, double* x,
                         double* y, double* z, double* vx, double* vy,
                         double* vz, double* ax, double* ay, double* az,
                         const double* m, double G, double eps2, double dt,
                         int steps, int init_acc) {
  const int nb = n + t;
  if (init_acc) accel_all(dim, n, t, x, y, z, m, ax, ay, az, G, eps2);
  double w[3];
  int nsub;
  if (scheme == 0) {
    w[0] = 1.0;
    nsub = 1;
  } else {
    const double c = pow(2.0, 1.0 / 3.0);
    w[0] = 1.0 / (2.0 - c);
    w[1] = -c / (2.0 - c);
    w[2] = w[0];
    nsub = 3;
  }
  for (int s = 0; s < steps; ++s) {
    for (int k = 0; k < nsub; ++k) {
      const double full = w[k] * dt;
      const double half = 0.5 * full;
      for (int i = 0; i < nb; ++i) {
        vx[i] = vx[i] + ax[i] * half;
        vy[i] = vy[i] + ay[i] * half;
        x[i] = x[i] + vx[i] * full;
        y[i] = y[i] + vy[i] * full;
        if (dim == 3) {
          vz[i] = vz[i] + az[i] * half;
          z[i] = z[i] + vz[i] * full;
        }
      }
      accel_all(dim, n, t, x, y, z, m, ax, ay, az, G, eps2);
      for (int i = 0; i < nb; ++i) {
        vx[i] = vx[i] + ax[i] * half;
        vy[i] = vy[i] + ay[i] * half;
        if (dim == 3) vz[i] = vz[i] + az[i] * half;
      }
    }
  }
}
