/*
 * ssim.h — C-ABI of libssim.so, the MI355X (gfx950) GPU physics backend of
 * the solar-system / self-gravity simulator.
 *
 * Conventions shared by every entry point
 * ---------------------------------------
 *  - All array arguments are DEVICE pointers (HBM), structure-of-arrays,
 *    one array per axis.  `dim` is 2 or 3; for dim == 2 every z pointer may
 *    be NULL and is never touched.
 *  - Nothing here allocates device memory.  Entry points that need scratch
 *    take a `work` pointer plus its size in bytes; the matching
 *    `*_work_bytes` query returns the size required.
 *  - Work is enqueued on `stream` (a hipStream_t passed as void*; NULL is
 *    the legacy default stream) and is NOT synchronised: the caller orders
 *    its reads on the same stream.
 *  - Return value: 0 on success, SSIM_EINVAL for a rejected argument,
 *    SSIM_EHIP for a HIP runtime error.  ssim_last_error() returns the
 *    message of the last failure on the calling thread.
 *  - Entry points are reentrant for distinct buffers; calls that share a
 *    buffer must be ordered by the caller (same stream).
 *
 * Reference citations are path:line in the Peuqui/SolarSystemSimulator
 * sources (backend/...).  The reference has no C-ABI: each function below
 * replaces a kernel or CuPy operation that its Python classes launch.
 */
#ifndef SSIM_H
#define SSIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SSIM_OK 0
#define SSIM_EINVAL (-1)
#define SSIM_EHIP (-2)

/* Precision of the pairwise force sum. */
#define SSIM_PREC_F64 0   /* FP64 terms, blocked + two-sum accumulation     */
#define SSIM_PREC_F32C 1  /* FP32 terms on hi/lo split positions, two-sum   */

/* Integrator schemes. */
#define SSIM_SCHEME_KDK 0      /* leapfrog kick-drift-kick                  */
#define SSIM_SCHEME_YOSHIDA4 1 /* 3 weighted KDK sub-steps (w1, w0, w1)     */

/* Message of the last failed call on this thread ("" if none). */
const char* ssim_last_error(void);
/* Library version string and the gfx target it was built for. */
const char* ssim_version(void);
/* Number of HIP devices visible (0 on a host without GPU, never fails). */
int ssim_device_count(void);

/* ------------------------------------------------------------------ */
/* All-pairs softened gravity                                          */
/* ------------------------------------------------------------------ */

/* a_i = G * sum_j m_j (x_j - x_i) / (|x_j - x_i|^2 + eps2)^(3/2)
 * for every target i against every source j (the self term is 0 because
 * dx = 0; eps2 must be > 0).  Sources (sx,sy,sz,sm)[ns], targets
 * (tx,ty,tz)[nt], output (ax,ay,az)[nt] (overwritten).
 * Replaces `beschleunigung` (selfgrav_kernel.py:222-260) and
 * `allpairs_accel` (test_pm.py:32-41). */
size_t ssim_allpairs_work_bytes(int dim, int prec, int ns, int nt);
int ssim_allpairs(int dim, int prec, int ns, const double* sx, const double* sy,
                  const double* sz, const double* sm, int nt, const double* tx,
                  const double* ty, const double* tz, double* ax, double* ay,
                  double* az, double G, double eps2, void* work,
                  size_t work_bytes, void* stream);

/* ------------------------------------------------------------------ */
/* Self-gravity integrator (masses + massless tracers)                 */
/* ------------------------------------------------------------------ */

/* State layout: every per-body array holds n masses followed by t tracers
 * (length n + t).  Only the first n entries of m are read (tracers have no
 * mass and are no source).  a (ax,ay,az) holds the accelerations that
 * belong to the current positions; with init_acc != 0 they are computed
 * first (what every launch of the reference does, selfgrav_kernel.py:
 * 305-328), with init_acc == 0 the caller asserts they are current.
 * Per step: scheme KDK = v += a*(dt/2); x += v*dt; a = F(x); v += a*(dt/2)
 * (selfgrav_kernel.py:333-417); scheme YOSHIDA4 = the same with dt scaled
 * by w1, w0, w1 (nbody_kernel.py:64-65,314-401).  If snap != NULL, step s
 * writes f32 [x|y|(z)|vx|vy|(vz)] of the masses (2*dim*n values) followed
 * by [tx|ty|(tz)] of the tracers (dim*t values) at snap + s*snap_stride;
 * the dim == 2 layout is the reference's (steps, 4N+2T) row
 * (selfgrav_kernel.py:560-627). */
size_t ssim_selfgrav_work_bytes(int dim, int prec, int n, int t);
int ssim_selfgrav_run(int dim, int prec, int scheme, int n, int t, double* x,
                      double* y, double* z, double* vx, double* vy, double* vz,
                      double* ax, double* ay, double* az, const double* m,
                      double G, double eps2, double dt, int steps, int init_acc,
                      float* snap, long long snap_stride, void* work,
                      size_t work_bytes, void* stream);

/* Elementwise kick v += a*h and drift x += v*h on n bodies (no FMA
 * contraction: the same rounding as `vx += ax * h` in NumPy,
 * test_selfgrav.py:81-87). */
int ssim_kick(int dim, int n, double* vx, double* vy, double* vz,
              const double* ax, const double* ay, const double* az, double h,
              void* stream);
int ssim_drift(int dim, int n, double* x, double* y, double* z,
               const double* vx, const double* vy, const double* vz, double h,
               void* stream);

/* ------------------------------------------------------------------ */
/* FP64 energy diagnostic                                              */
/* ------------------------------------------------------------------ */

/* E_kin = 1/2 sum m v^2, E_pot = -1/2 G sum_{i != j} m_i m_j /
 * sqrt(r^2 + eps2) over the n masses, accumulated with two-sum in a fixed
 * order (bitwise reproducible).  out[0] = E_kin, out[1] = E_pot,
 * out[2] = E_kin + E_pot (device pointer, 3 doubles).  The formula is the
 * reference's `energie` (test_selfgrav.py:54-63), which the reference only
 * evaluates on the host. */
size_t ssim_energy_work_bytes(int dim, int n);
int ssim_energy(int dim, int n, const double* x, const double* y,
                const double* z, const double* vx, const double* vy,
                const double* vz, const double* m, double G, double eps2,
                double* out, void* work, size_t work_bytes, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* SSIM_H */
