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
/* Launch timing for measurement runs: ssim_timing(1) clears the records and
 * brackets every later launch of the library's main kernels (k_allpairs_f64,
 * k_fft_pass, k_conv_x3, k_pp_cells, k_tsc_deposit, k_tsc_gather, k_cic_deposit,
 * k_cic_gather, k_mlp_features, k_mlp_forward) with HIP events on the
 * launch's stream; ssim_timing(0) stops recording.  ssim_timing_report
 * waits for the recorded launches of `kernel` and returns their number and
 * summed duration.  Not thread-safe against concurrent launches. */
int ssim_timing(int enable);
int ssim_timing_report(const char* kernel, double* total_ms, long long* launches);

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
/* Drift in the periodic box [o, o + box)^dim: x += v*h, then every
 * coordinate is wrapped as x - box floor((x - o) / box) (unchanged while
 * inside).  The periodic particle-mesh integrator's drift; no reference
 * counterpart (the reference's PM is isolated, pm_kernel.py:166-293). */
int ssim_drift_periodic(int dim, int n, double* x, double* y, double* z,
                        const double* vx, const double* vy, const double* vz, double h,
                        const double* origin, double box, void* stream);
/* v += a*half then x += v*full (the opening half of a KDK step) in one
 * pass, with the rounding of ssim_kick followed by ssim_drift. */
int ssim_kick_drift(int dim, int n, double* x, double* y, double* z, double* vx,
                    double* vy, double* vz, const double* ax, const double* ay,
                    const double* az, double half, double full, void* stream);
/* Half-kick v += a*h of n masses + t tracers followed (if snap != NULL) by
 * the f32 snapshot row described at ssim_selfgrav_run; used by integrators
 * whose force comes from elsewhere (particle mesh, NBodyPM.step_batch,
 * pm_kernel.py:254-281). */
int ssim_kick_snap(int dim, int n, int t, const double* x, const double* y,
                   const double* z, double* vx, double* vy, double* vz,
                   const double* ax, const double* ay, const double* az,
                   double h, float* snap, void* stream);

/* dst[0..n) = value (a kernel, not hipMemset, so it can be graph-captured);
 * e.g. the per-batch reset of the contact times (nbody_kernel.py:811). */
int ssim_fill_f32(long long n, float* dst, float value, void* stream);
/* dst[idx[k]] = vals[k], or += when add != 0 (indices unique): the point
 * updates of the resident state (nbody_kernel.py:943-1040). */
int ssim_scatter_f64(int n, const long long* idx, const double* vals,
                     double* dst, int add, void* stream);

/* ------------------------------------------------------------------ */
/* Few massive bodies + many asteroids (raster batches)                */
/* ------------------------------------------------------------------ */

/* K raster samples of dt_raster in one launch (frame_kernel,
 * nbody_kernel.py:97-560, single device: its gpuId/nGpus/GSync arguments
 * are gone).  Asteroid arrays have n_ast entries, massive arrays m <= 64;
 * aacc_x/y and macc_x/y are scratch (recomputed at the start, as in the
 * reference).
 * soft is added to r^2.  Outputs: snap f32 [k][4][n_ast] (x|y|vx|vy),
 * snap_m f32 [k][4][m], sub_pos f32 [k][m_sub][2][n_ast] and sub_n u8
 * [k][n_ast] (hot sub-samples; complete when sub_n == m_sub), hit_t f32
 * [n_ast] (batch time of first contact; the caller resets it to -1 per
 * batch) and hit_m i32 [n_ast] (massive partner).  Cooperative launch
 * (grid barriers), deterministic back-reaction sums. */
size_t ssim_nbody_work_bytes(int n_ast);
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
                     void* work, size_t work_bytes, void* stream);

/* The same raster integrator in space (D = 3): every formula of
 * ssim_nbody_frame with a z component added (|d|^2 = dx^2 + dy^2 + dz^2,
 * hot classification, encounter times and contact distances in 3-D); the
 * reference runs only the plane (nbody_kernel.py:97-560 has x, y), so this
 * entry is its 3-D generalisation for scenes with inclinations.  Layouts:
 * snap f32 [k][6][n_ast] and snap_m f32 [k][6][m] as x, y, z, vx, vy, vz;
 * sub_pos f32 [k][m_sub][3][n_ast]; work: ssim_nbody_work_bytes(n_ast). */
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
                       void* work, size_t work_bytes, void* stream);

/* Film ring slots of k raster samples, written on the device from the
 * outputs of ssim_nbody_frame: replaces the host side of
 * film_producer.py:81-123 (`slot_bytes`, `schreibe_slot`) applied to
 * nbody_kernel.py:836-859 (scatter to original order) and 894-934 (complete
 * fine tracks, subN == m_sub).  Slot f at slots + f * slot_bytes:
 *   x[n] f32 | y[n] f32 | u32 nh | u32 idx[sub_max] | f32 track[sub_max][m_sub][2]
 * nh = min(totals[f], sub_max); idx ascending original indices (the lowest
 * kept on overflow, totals[f] > sub_max); entries past nh untouched.
 * a_idx (u32 [n_ast], strictly increasing) and m_idx (u32 [m]) map the
 * integrator's slots to original indices, n = n_ast + m.  m_sub == 0 or
 * sub_max == 0 writes nh = 0.  work: ssim_ring_work_bytes(k, n_ast). */
long long ssim_ring_slot_bytes(int n, int m_sub, int sub_max);
size_t ssim_ring_work_bytes(int k, int n_ast);
int ssim_ring_slots(int k, int n, int n_ast, int m, int m_sub, int sub_max,
                    const float* snap, const float* snap_m, const float* sub_pos,
                    const unsigned char* sub_n, const unsigned* a_idx,
                    const unsigned* m_idx, unsigned char* slots, int* totals,
                    void* work, size_t work_bytes, void* stream);

/* The k raster samples of ssim_nbody_frame as rows f32 [k][4][n]
 * (x, y, vx, vy) in original body order: the array step_batch returns
 * (nbody_kernel.py:836-856 scatters it on the host), kept on the device
 * for the producer's detection.  a_idx / m_idx as for ssim_ring_slots. */
int ssim_frame_rows(int k, int n, int na, int m, const float* snap,
                    const float* snap_m, const unsigned* a_idx,
                    const unsigned* m_idx, float* rows, void* stream);

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

/* ssim_energy without a workspace and without any allocation: no
 * counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops that report the
 * energy every step, harnesses that forbid hipMalloc).  The per-workgroup
 * partials live in a fixed device array inside the library (2 MiB per GPU,
 * held until the process ends), so n is limited to
 * SSIM_ENERGY_NOALLOC_MAX_N bodies (larger n: SSIM_EINVAL, use
 * ssim_energy).  Results are bit-identical to ssim_energy's.  Every call
 * on one GPU shares that array: calls that overlap on one GPU (different
 * streams, or host threads without ordering) race and must be serialised
 * by the caller. */
#define SSIM_ENERGY_NOALLOC_MAX_N (1 << 24)
int ssim_energy_noalloc(int dim, int n, const double* x, const double* y,
                        const double* z, const double* vx, const double* vy,
                        const double* vz, const double* m, double G, double eps2,
                        double* out, void* stream);

/* ------------------------------------------------------------------ */
/* FFT (real <-> half-complex, NumPy layout and normalisation)         */
/* ------------------------------------------------------------------ */

/* dims: ndim (1..3) powers of two in [2, 4096], last axis contiguous.
 * The complex half spectrum has shape (dims[0], .., dims[ndim-1]/2 + 1),
 * interleaved (re, im) in FP64 (prec 0) or FP32 (prec 1).  `work` holds
 * the twiddle tables: size from ssim_fft_work_bytes, filled once by
 * ssim_fft_prepare.  r2c = numpy.fft.rfftn, c2r = numpy.fft.irfftn (scaled
 * by 1/prod(dims)); c2r overwrites its complex input.
 * Replaces cp.fft.rfft2 / irfft2 (pm_kernel.py:60,144-146). */
size_t ssim_fft_work_bytes(int ndim, const int* dims, int prec);
int ssim_fft_prepare(int ndim, const int* dims, int prec, void* work,
                     size_t work_bytes, void* stream);
int ssim_fft_r2c(int ndim, const int* dims, int prec, const void* in,
                 void* out, const void* work, size_t work_bytes, void* stream);
int ssim_fft_c2r(int ndim, const int* dims, int prec, void* in, void* out,
                 const void* work, size_t work_bytes, void* stream);

/* ------------------------------------------------------------------ */
/* Particle-mesh building blocks                                       */
/* ------------------------------------------------------------------ */

/* Cloud-in-cell deposit of masses m[np] at (x,y,z) onto a mesh of
 * grid_n[dim] nodes (C order, x slowest) with origin[dim] and cell size h,
 * ADDED to `mesh` (FP64 or FP32 per mesh_prec; the caller zeroes it).
 * Isolated mesh (periodic = 0): cell = clip(floor((x-x0)/h), 0, n-2) as in
 * `_cic_deposit` (pm_kernel.py:63-90); periodic: floor wrapped mod n.  If
 * cells != NULL the integer cell indices are written as cells[dim*p + c].
 * The summation order of the atomic adds is not fixed (like cp.bincount). */
int ssim_cic_deposit(int dim, int periodic, int np, const double* x,
                     const double* y, const double* z, const double* m,
                     const int* grid_n, const double* origin, double h,
                     int mesh_prec, void* mesh, int* cells, void* stream);
/* The same CIC deposit, reproducible and without floating-point
 * atomics: the mesh is OVERWRITTEN (every node written once, no zeroing
 * needed), bit-identical from run to run.  The masses are filed by tiles of
 * 8^3 (3-D) / 32^2 (2-D) nodes (one integer count per distinct tile of a
 * wave, a scan, a ranked scatter), and one workgroup per tile sums its
 * nodes in LDS as int64 fixed point (each of the reference's FP64 corner
 * products rounded to 2^(e-62), 2^e >= masses the tile scans x max |m|)
 * before writing them.  Same cell rule and `cells` output as
 * ssim_cic_deposit.  work: ssim_cic_deposit_tiles_work_bytes.  Replaces
 * the reference's `cp.bincount` (pm_kernel.py:89). */
size_t ssim_cic_deposit_tiles_work_bytes(int dim, int np, const int* grid_n);
int ssim_cic_deposit_tiles(int dim, int periodic, int np, const double* x,
                           const double* y, const double* z, const double* m,
                           const int* grid_n, const double* origin, double h,
                           int mesh_prec, void* mesh, int* cells, void* work,
                           size_t work_bytes, void* stream);
/* CIC interpolation of nfields (1..3) mesh fields at np positions, same
 * stencil as the deposit (`_cic_gather`, pm_kernel.py:93-107). */
int ssim_cic_gather(int dim, int periodic, int np, const double* x,
                    const double* y, const double* z, const int* grid_n,
                    const double* origin, double h, int mesh_prec, int nfields,
                    const void* f0, const void* f1, const void* f2,
                    double* out0, double* out1, double* out2, void* stream);
/* out[2c] = min, out[2c+1] = max of array c (narrays 1..3), on the device
 * (the reductions of `grid_fuer`, pm_kernel.py:117-118). */
size_t ssim_minmax_work_bytes(int n);
int ssim_minmax(int n, int narrays, const double* a0, const double* a1,
                const double* a2, double* out, void* work, size_t work_bytes,
                void* stream);
/* out[2c], out[2c+1] = min, max over the n bodies of the positions that
 * the next half kick (v += a half) and drift (x += v full) will produce,
 * x_c + (v_c + a_c half) full, with the kernels' own operations, so the
 * box is exact (device output, workspace ssim_minmax_work_bytes(n)).  The
 * split-force integrator places its mesh from it before the step, so a
 * step needs no host round trip.  New: the reference computes the box of
 * the current positions (`grid_fuer`, pm_kernel.py:117-118). */
int ssim_box_after_kdk(int dim, int n, const double* x, const double* y,
                       const double* z, const double* vx, const double* vy,
                       const double* vz, const double* ax, const double* ay,
                       const double* az, double half, double full, double* out,
                       void* work, size_t work_bytes, void* stream);
/* The closing half kick of a step, v += a half (as ssim_kick), fused with
 * ssim_box_after_kdk for the next step (out = the box of x + (v + a half)
 * full with the kicked v): the kick kernel writes per-workgroup extrema to
 * work (ssim_kick_box_work_bytes(n) bytes, no initial contents needed; the
 * first 256 bytes are left untouched) and a one-workgroup launch folds them
 * into out.  Calls sharing a work buffer must be ordered.  No reference
 * counterpart. */
size_t ssim_kick_box_work_bytes(int n);
int ssim_kick_box(int dim, int n, const double* x, const double* y, const double* z,
                  double* vx, double* vy, double* vz, const double* ax,
                  const double* ay, const double* az, double half, double full,
                  double* out, void* work, size_t work_bytes, void* stream);
/* The closing half kick of a step and the next step's opening half kick
 * and drift in one pass: v += a half, v += a half, x += v full (the
 * operations, in order, of ssim_kick then ssim_kick_drift).  For a caller
 * that computes the next box beforehand (ssim_box_after_kdk on the same
 * state gives the box of the positions stored here).  No reference
 * counterpart. */
int ssim_kicks_drift(int dim, int n, double* x, double* y, double* z, double* vx,
                     double* vy, double* vz, const double* ax, const double* ay,
                     const double* az, double half, double full, void* stream);
/* ssim_kick_box followed by the next step's opening kick and drift, in one
 * pass: v += a half, v += a half, x += v full (the operations, in order, of
 * ssim_kick_box then ssim_kick_drift, so the state is bit-identical), and
 * out = the box of the new positions.  For a step that another step with
 * the same accelerations' kick follows; work as ssim_kick_box.  No
 * reference counterpart. */
int ssim_kick_drift_box(int dim, int n, double* x, double* y, double* z, double* vx,
                        double* vy, double* vz, const double* ax, const double* ay,
                        const double* az, double half, double full, double* out,
                        void* work, size_t work_bytes, void* stream);
/* Real-space force Green's function component `comp` on the zero-padded
 * (2n)^dim mesh: -G d_comp / (|d|^2 + eps2)^(3/2) with wrapped offsets
 * (build_force_kernels, pm_kernel.py:34-60, generalised to 3-D).  rs > 0
 * keeps only the long-range part: times 1 - g(r) with the Gaussian split
 * g(r) = erfc(r/2rs) + r/(rs sqrt(pi)) exp(-r^2/4rs^2) for r < rcut (all
 * r if rcut <= 0) and the full kernel beyond rcut, where the near-field sum
 * stops; it is zero where an offset is -n (never used by two nodes of the
 * n-mesh), which makes it exactly odd.  rs = 0 is the reference's full
 * kernel (rcut ignored). */
int ssim_pm_force_kernel(int dim, int n, int comp, double h, double eps2,
                         double G, double rs, double rcut, int mesh_prec,
                         void* out, void* stream);
/* Divide a half spectrum on the padded mesh (padded_n^dim) by the squared
 * assignment window prod_c sinc^(2 order)(pi f_c / padded_n), order 2 for
 * CIC and 3 for TSC (no reference counterpart: used by the split mesh +
 * near-field force). */
int ssim_pm_deconvolve(int dim, int padded_n, int prec, int order,
                       void* spectrum, void* stream);
/* Triangular-shaped-cloud deposit of np masses on an isolated grid_n^3
 * mesh (ADDS to mesh): weights (1/2 - d)^2/2, 3/4 - d^2, (1/2 + d)^2/2 on
 * the nodes c-1, c, c+1 of each axis around the nearest node c, d the
 * offset from it in cells; nodes outside the mesh are skipped.  No
 * reference counterpart (the split force's assignment). */
int ssim_tsc_deposit(int np, const double* x, const double* y, const double* z,
                     const double* m, int grid_n, const double* origin, double h,
                     int mesh_prec, void* mesh, void* stream);
/* The TSC deposit without atomics (OVERWRITES mesh) of np masses (work:
 * ssim_tsc_deposit_work_bytes(np), their packed copy in bin order): the
 * masses are binned
 * by nearest node (ssim_cells_build with origin - h/2, cell h and
 * grid_n^3 cells: bin_start, bin_order) and every node sums the bodies of
 * the 27 bins around it in a fixed order, so the mesh is bit-reproducible.
 * Same weights as ssim_tsc_deposit. */
size_t ssim_tsc_deposit_work_bytes(int np);
int ssim_tsc_deposit_binned(int np, const double* x, const double* y,
                            const double* z, const double* m, int grid_n,
                            const double* origin, double h, const int* bin_start,
                            const int* bin_order, int mesh_prec, void* mesh,
                            void* work, size_t work_bytes, void* stream);
/* The same atomic-free deposit (OVERWRITES mesh) over the blocked bins of
 * ssim_split_bin (bin_start, and body = its double4 copy of the masses in
 * bin order) with the same mesh and block edge (>= 2); bit-reproducible.
 * The masses must lie inside the mesh.  No reference counterpart. */
int ssim_tsc_deposit_blocked(int grid_n, const double* origin, double h, int block,
                             const int* bin_start, const void* body, int mesh_prec,
                             void* mesh, void* stream);
/* The same deposit (OVERWRITES mesh) from ssim_split_bin's block starts
 * (cell_start) and body, for block edge 5: one workgroup per tile of 2^3
 * blocks sums its nodes in LDS as int64 fixed point (contributions
 * m w_x w_y w_z 2^s, rounded to nearest), so the sums are exact and do not
 * depend on the order of the additions: bit-reproducible.  mass_bound >=
 * sum |m| fixes s = 62 - ceil(log2 mass_bound) (no sum can overflow; each
 * contribution is exact to 2^-(s+1)).  Only the block order of the masses
 * is used.  No reference counterpart. */
int ssim_tsc_deposit_tiles(int grid_n, const double* origin, double h, int block,
                           const int* cell_start, const void* body, double mass_bound,
                           int mesh_prec, void* mesh, void* stream);
/* TSC interpolation of three meshes at np points: o_c = sum w f_c
 * (overwrites o_c, FP64, or adds to it when accumulate != 0); thread i handles point order[i] (order = NULL:
 * point i) — a cell-sorted order makes neighbouring threads read
 * neighbouring nodes.  f1 = f2 = NULL: f0 holds the three FP32 fields
 * interleaved as float4 [grid_n^3] (ssim_pm_conv_isolated3's layout). */
int ssim_tsc_gather(int np, const int* order, const double* x, const double* y,
                    const double* z, int grid_n, const double* origin, double h,
                    int mesh_prec, const void* f0, const void* f1, const void* f2,
                    double* o0, double* o1, double* o2, int accumulate,
                    void* stream);
/* ssim_tsc_gather of the interleaved FP32 fields (accumulate as there)
 * with the learned correction folded in: corr[np][3] are the network's
 * outputs (ssim_mlp_forward with out != NULL and no ax), s the gathered
 * split force of a body, u = scale |s| (relative != 0) or scale, d = u corr
 * without its component along (v0, v1, v2) for bodies below n_proj, and
 * s + d is stored — what ssim_mlp_forward would add afterwards, with the
 * same operations, in one pass over the bodies.  No reference counterpart. */
int ssim_tsc_gather_corrected(int np, const int* order, const double* x,
                              const double* y, const double* z, int grid_n,
                              const double* origin, double h, const void* f4,
                              double* o0, double* o1, double* o2, int accumulate,
                              const float* corr, float scale, int relative,
                              const double* v0, const double* v1, const double* v2,
                              int n_proj, void* stream);
/* ssim_tsc_gather of the interleaved FP32 fields (corr == NULL) or
 * ssim_tsc_gather_corrected (corr != NULL; n_proj and the velocities as
 * there), the same results, and out[0..6) = the box (min, max per axis) of
 * the positions x + ((v + a half) + a half) full, a the stored sum: where
 * the bodies go with this step's closing half kick and the next step's
 * opening half kick and drift (ssim_kick_drift_box, ssim_kicks_drift).  v0,
 * v1, v2: the velocities of all np points.  The gather writes
 * per-workgroup extrema to work (ssim_tsc_gather_box_work_bytes(np)) and a
 * one-workgroup launch folds them into out, which may be pinned host
 * memory.  np >= 1.  No reference counterpart. */
size_t ssim_tsc_gather_box_work_bytes(int np);
int ssim_tsc_gather_box(int np, const int* order, const double* x, const double* y,
                        const double* z, int grid_n, const double* origin, double h,
                        const void* f4, double* o0, double* o1, double* o2, int accumulate,
                        const float* corr, float scale, int relative, const double* v0,
                        const double* v1, const double* v2, int n_proj, double half,
                        double full, double* out, void* work, size_t work_bytes,
                        void* stream);
/* out = a * b over ncomplex interleaved complex values (the spectral
 * products of pm_force_field, pm_kernel.py:144-146). */
int ssim_spectrum_mul(long long ncomplex, int prec, const void* a,
                      const void* b, void* out, void* stream);
/* out[i] = Im a[i] over ncomplex values (the split kernel's spectrum is
 * imaginary; ssim_pm_conv_isolated3 keeps only that part). */
int ssim_spectrum_imag(long long ncomplex, int prec, const void* a, void* out,
                       void* stream);
/* Isolated 3-D convolution of the n^3 mass mesh rho with three kernels
 * given by their imaginary spectra kim_* on the padded (2n)^3 mesh
 * (rfftn half-spectrum layout, real values): field_c = corner n^3 of
 * irfftn(rfftn(pad(rho)) * i kim_c), the result of ssim_pm_pad +
 * ssim_fft_r2c + ssim_spectrum_mul + ssim_fft_c2r + crop, computed without
 * the padding or the discarded part (pruned passes).  With fy = fz = NULL
 * the three fields are written interleaved into fx as [n^3][4] (component
 * c at index 4 i + c; the 4th slot is either left as it is or set to 0).  spec and tmp hold
 * (2n)^2 (n+1) complex values each; fft_work is prepared for dims
 * (2n, 2n, 2n).  No reference counterpart (split force). */
int ssim_pm_conv_isolated3(int n, int prec, const void* rho, const void* kim_x,
                           const void* kim_y, const void* kim_z, void* fx,
                           void* fy, void* fz, void* spec, void* tmp,
                           const void* fft_work, size_t fft_work_bytes,
                           void* stream);
/* Periodic Poisson solve in place on an rfftn half spectrum of MASS per
 * node: phi_k = -4 pi G (m_k / h^dim) / k^2, phi_0 = 0. */
int ssim_pm_poisson(int dim, const int* grid_n, double h, double G, int prec,
                    void* spectrum, void* stream);
/* The periodic potential of a 3-D mesh of mass per node in one call:
 * rfftn, the ssim_pm_poisson factor, irfftn, with the x axis's forward
 * transform, factor and inverse transform fused in one pass (the two
 * spectrum round trips of that axis saved); the result equals
 * ssim_fft_r2c + ssim_pm_poisson + ssim_fft_c2r bit for bit.  The
 * reference's PM is 2-D and isolated (pm_kernel.py:127-147); this serves
 * the periodic 3-D plan (TODO.md:225-253).  rho, phi: real grid_n^3;
 * spec: the half spectrum (grid_n[2]/2 + 1 in the last axis) as scratch;
 * fft_work prepared for grid_n. */
int ssim_pm_potential_periodic(const int* grid_n, int prec, const void* rho, void* spec,
                               void* phi, double h, double G, const void* fft_work,
                               size_t fft_work_bytes, void* stream);
/* Periodic central difference g_c = -(phi[+1] - phi[-1]) / (2h). */
int ssim_pm_gradient(int dim, const int* grid_n, double h, int prec,
                     const void* phi, void* g0, void* g1, void* g2,
                     void* stream);
/* crop == 0: copy an n^dim block into the corner of a (2n)^dim block (dst
 * must be zeroed); crop == 1: copy the corner back out
 * (pm_force_field's padding and cropping, pm_kernel.py:141-146). */
int ssim_pm_pad(int dim, int n, int prec, const void* src, void* dst, int crop,
                void* stream);

/* ------------------------------------------------------------------ */
/* Cell list and near-field (particle-particle) force                  */
/* ------------------------------------------------------------------ */

/* Bin n bodies into nc[0] x nc[1] x nc[2] cubic cells of edge `cell` from
 * `origin` (outside coordinates clamp to border cells).  cell_start gets
 * ncells + 1 offsets, order the body indices sorted by cell (ascending
 * index within a cell), packed (optional, float4 per body) the
 * origin-relative FP32 position and mass in that order.  The uniform-grid
 * hashing of the reference's collision search (film_producer.py:235-240)
 * and the near-field source lists use it.  No hipMemset inside. */
size_t ssim_cells_work_bytes(int n, int ncells);
int ssim_cells_build(int n, const double* x, const double* y, const double* z,
                     const double* m, const double* origin, double cell,
                     const int* nc, int* cell_start, int* order, void* packed,
                     void* work, size_t work_bytes, void* stream);
/* The same binning with the bodies of a cell grouped by the octant of the
 * cell they lie in (x, then y, then z halves), so any run of consecutive
 * bodies of a cell is spatially compact; ascending body index inside an
 * octant (a stable radix sort of cell * 8 + octant, so the lists, and what
 * is computed from them, do not change from run to run).  For the
 * near-field targets.  No hipMemset inside. */
size_t ssim_cells_bin_work_bytes(int n, int ncells);
int ssim_cells_bin(int n, const double* x, const double* y, const double* z,
                   const double* origin, double cell, const int* nc,
                   int* cell_start, int* order, void* work, size_t work_bytes,
                   void* stream);
/* One stable binning of the n masses (x, y, z, m) for both halves of the
 * split force on a grid_n^3 mesh (origin, h): every mass goes to the TSC
 * node its deposit stencil is centred on, r = floor((x - o)/h + 1/2) per
 * axis, and the nodes are grouped in blocks of block^3 (nb = ceil(grid_n /
 * block) per axis) that serve as the near-field cells, edge block * h
 * (>= rcut), grid origin o - h/2.  Bins are ordered block-major
 * (ssim_tsc_deposit_blocked's key), bodies inside a bin by index.
 * Outputs: bin_start[nb^3 block^3 + 1], cell_start[nb^3 + 1] (the block
 * starts, for ssim_pp_short), order[n] (mass indices in bin order),
 * packed[n] float4 (x, y, z relative to o - h/2, m) for ssim_pp_short and
 * body[n] double4 ((x - o)/h per axis, m) for ssim_tsc_deposit_blocked and
 * ssim_tsc_deposit_tiles, and (near != NULL) near[n] float4 (x, y, z
 * relative to the corner of the mass's own block, m), the copy
 * ssim_pp_near's pair terms read, and (near_lo != NULL) near_lo[n] float4
 * (the remainders of those three coordinates, 0): hi + lo for
 * ssim_pp_near_hilo.  With bin_start = NULL only the blocks are binned
 * (key = block index, stable order, order inside a block by index): all
 * that ssim_pp_short and ssim_tsc_deposit_tiles read.  It replaces the two
 * cell lists (nodes and near-field cells) of ssim_cells_build.  No
 * hipMemset inside.  No reference counterpart. */
size_t ssim_split_bin_work_bytes(int n, int grid_n, int block);
int ssim_split_bin(int n, const double* x, const double* y, const double* z,
                   const double* m, int grid_n, const double* origin, double h,
                   int block, int* bin_start, int* cell_start, int* order, void* packed,
                   void* body, void* near, void* near_lo, void* work, size_t work_bytes,
                   void* stream);
/* One binning of all n_all bodies (the n_mass masses first, then the
 * tracers) on ssim_split_bin's blocks, for the split force's near field
 * and tiled deposit: every body gets the block its TSC stencil is centred
 * on (ssim_split_bin's rule), one stable sort orders them by block, index
 * order inside a block.  Outputs: tcell_start[nb^3 + 1] and torder[n_all]
 * (the target lists of ssim_pp_near, which may be any binning of the
 * targets on this grid) and, for the masses, exactly ssim_split_bin's
 * outputs with bin_start = NULL: cell_start[nb^3 + 1], order[n_mass],
 * packed, body, near, near_lo (the masses lead every block's segment of
 * torder, so their order is its restriction to indices < n_mass).  It
 * replaces ssim_split_bin + ssim_cells_bin (two sorts) by one.  No hipMemset
 * inside.  No reference counterpart. */
size_t ssim_split_bin_all_work_bytes(int n_all, int grid_n, int block);
int ssim_split_bin_all(int n_all, int n_mass, const double* x, const double* y,
                       const double* z, const double* m, int grid_n, const double* origin,
                       double h, int block, int* cell_start, int* order, void* packed,
                       void* body, void* near, void* near_lo, int* tcell_start,
                       int* torder, void* work, size_t work_bytes, void* stream);
/* ADD (accumulate != 0) or STORE the near-field force G m d g(r) /
 * (r^2 + eps2)^(3/2), r < rcut, of the binned sources (cell_start +
 * packed, from ssim_cells_build or ssim_split_bin) to / into nt targets'
 * accelerations (3-D); every target is written once.  The targets are binned on the same grid
 * (tcell_start, torder from ssim_cells_bin or ssim_cells_build) and
 * rcut <= cell.  Each target's sum runs over the sources in their sorted
 * order, so the result is deterministic.  New: the reference has no split
 * force. */
size_t ssim_pp_work_bytes(int nt, int ncells);
int ssim_pp_short(int nt, const double* tx, const double* ty, const double* tz,
                  const double* origin, double cell, const int* nc,
                  const int* cell_start, const void* packed,
                  const int* tcell_start, const int* torder, double G,
                  double eps2, double rs, double rcut, double* ax, double* ay,
                  double* az, int accumulate, void* work, size_t work_bytes,
                  void* stream);
/* ssim_pp_short for masses binned by ssim_split_bin, on their
 * block-relative copy `near` (cell_start, near from ssim_split_bin; targets
 * from ssim_cells_bin on the same grid): each of the 27 cells is walked
 * with the target re-expressed relative to that cell's corner, so the FP32
 * pair terms keep the positions' low bits; the terms of a cell are summed
 * in FP32 and the cells' partials in FP64.  Deterministic.  Work and the
 * STORE / ADD semantics as ssim_pp_short.  New: the reference has no split
 * force. */
int ssim_pp_near(int nt, const double* tx, const double* ty, const double* tz,
                 const double* origin, double cell, const int* nc, const int* cell_start,
                 const void* near, const int* tcell_start, const int* torder, double G,
                 double eps2, double rs, double rcut, double* ax, double* ay, double* az,
                 int accumulate, void* work, size_t work_bytes, void* stream);
/* ssim_pp_near with each mass position as an FP32 pair: near (hi) from
 * ssim_split_bin and near_lo (lo, its near_lo output), and each target as
 * FP64 relative to the cell's corner, so that the pair difference (t - hi)
 * - lo keeps about 48 bits; the pair factor is interpolated quadratically
 * in a table of GT + 2 entries.  Same order of summation, same semantics;
 * the force error against the FP64 sum drops from p99 ~2e-6 to < 1e-6.
 * New: the reference has no split force. */
int ssim_pp_near_hilo(int nt, const double* tx, const double* ty, const double* tz,
                      const double* origin, double cell, const int* nc,
                      const int* cell_start, const void* near, const void* near_lo,
                      const int* tcell_start, const int* torder, double G, double eps2,
                      double rs, double rcut, double* ax, double* ay, double* az,
                      int accumulate, void* work, size_t work_bytes, void* stream);
/* The near field of ssim_pp_near / ssim_pp_near_hilo (same inputs, same
 * STORE / ADD semantics; near_lo may be NULL: lo = 0) as a filter and a
 * sparse sum: per chunk of <= 64 targets, the 27 cells' masses are staged
 * in LDS relative to the target cell's corner in FP64 (hi + lo + cell
 * offset), in batches of <= 256 that take every B-th mass of each row; an
 * FP16 MFMA evaluates r^2 - thr for 64 x 32 pairs at a time and packs the
 * signs into per-target bit masks (thr = (rcut/cell)^2 + 0.02 cell^2, more
 * than 4x the FP16 error bound, so no pair inside rcut is missed); each
 * lane then sums the terms of its own set bits only, in FP64 positions and
 * sums.  precise != 0: quadratic table interpolation and an FP64-refined
 * (r^2 + eps2)^(-1/2), as ssim_pp_near_hilo; 0: linear interpolation and
 * the FP32 v_rsq.  Deterministic; a target's result depends only on its
 * position and its cell's masses.  Replaces the reference's tile loop over
 * all sources (backend/selfgrav_kernel.py:231-259) inside rcut.  New: the
 * reference has no split force. */
int ssim_pp_near_bits(int nt, const double* tx, const double* ty, const double* tz,
                      const double* origin, double cell, const int* nc,
                      const int* cell_start, const void* near, const void* near_lo,
                      const int* tcell_start, const int* torder, double G, double eps2,
                      double rs, double rcut, double* ax, double* ay, double* az,
                      int accumulate, int precise, void* work, size_t work_bytes,
                      void* stream);
/* Work accounting of ssim_pp_short with the same arguments: counts[0] =
 * source-target pairs tested, counts[1] = pairs closer than rcut (a body
 * with itself excluded); counts is a device array of 2, no forces. */
int ssim_pp_count(int nt, const double* tx, const double* ty, const double* tz,
                  const double* origin, double cell, const int* nc,
                  const int* cell_start, const void* packed,
                  const int* tcell_start, const int* torder, double eps2,
                  double rs, double rcut, unsigned long long* counts, void* work,
                  size_t work_bytes, void* stream);

/* ------------------------------------------------------------------ */
/* Uniform-grid collision search (bit-exact ordered pair list)         */
/* ------------------------------------------------------------------ */

/* out[0] (device) = the q-th percentile of the speeds |(gvx, gvy)| of
 * the alive bodies (is_ast and vis != 0), NumPy's 'linear' rule on FP32
 * data: the speeds sorted, virtual index (m - 1) (q / 100) in FP32 and the
 * two-sided lerp of numpy's percentile; NaN if no body is alive.  The cell
 * size of the collision search comes from its 95th percentile
 * (`cp.percentile(cp.hypot(gvx, gvy), 95)`, film_producer.py:357).  Each
 * speed is sqrt of the FP64 sum of squares rounded to FP32. */
size_t ssim_speed_percentile_work_bytes(int n);
int ssim_speed_percentile(int n, const float* gvx, const float* gvy,
                          const unsigned char* is_ast, const unsigned char* vis, float q,
                          float* out, void* work, size_t work_bytes, void* stream);
/* Candidates per batch of ssim_collide_hits for the following calls (a
 * multiple of 32 up to 2^25, the default; the workspace holds one batch's
 * hit bits).  The pair list does not depend on it.  A knob for tests of
 * the batching and for memory-tight callers. */
int ssim_collide_batch(long long candidates);
/* Asteroid-asteroid contact pairs of one sample, exactly the list of
 * `Erkennungskarte._hits` (film_producer.py:223-317) for the stripe
 * [lo, hi) with halo (single device: lo = -inf, hi = +inf): same alive
 * rule, XOR cell hash of floor(f64(x)/h), stable key order, neighbour
 * offsets (0,0),(1,0),(0,1),(1,1),(1,-1), FP32 closest-approach prefilter
 * over [-dt_y, dt_y] with radius sum + 1e-6, ownership min(x) < hi and the
 * FP64 recheck, in the reference's order.  Inputs are the f32 sample
 * (gx, gy, gvx, gvy), f32 contact radii rr, u8 flags is_ast and vis.
 * counts[0] = candidate pairs, counts[1] = FP32 hits, counts[2] = final
 * pairs (or -1 when f32_hit_capacity was too small: retry with
 * counts[1]).  Pairs are written to out_i/out_j (int64) up to
 * pair_capacity; the workspace depends on n and f32_hit_capacity.
 * Candidate pairs must stay below 2^31 (SSIM_EINVAL otherwise).  One
 * prefilter evaluation per candidate (load-balanced runs of 32 over the
 * flattened candidate list, batched), no atomics: deterministic.
 * Synchronises the stream (counts are returned on the host). */
size_t ssim_collide_work_bytes(int n, long long f32_hit_capacity);
int ssim_collide_hits(int n, const float* gx, const float* gy, const float* gvx,
                      const float* gvy, const float* rr,
                      const unsigned char* is_ast, const unsigned char* vis,
                      float lo, float hi, float halo, double dt_y, double h,
                      long long f32_hit_capacity, long long pair_capacity,
                      long long* out_i, long long* out_j, long long* counts,
                      void* work, size_t work_bytes, void* stream);

/* Massive-body contacts and the runaway guard of one sample
 * (`analyze_merge`, film_producer.py:707-779), FP32 like the reference:
 * for each of the nm alive massive bodies m_alive[r] and every visible body
 * j != m_alive[r], does the segment prev -> cur (frac = tt - 1) or
 * cur -> cur + v dt (frac = tt) pass within rr[m] + rr[j]?  Hits come out
 * row-major (row r, column j ascending) as (out_row, out_col, out_frac) up
 * to `capacity`; counts[0] = number of hits.  With do_runaway, the indices
 * of asteroids with v^2 > 9 (two_gm / max(|x - b|, 1e-6)) (b = (bx, by),
 * the massive barycentre) go to `runaway` in ascending order, counts[1] =
 * their number.  Synchronises the stream. */
size_t ssim_merge_work_bytes(int n, int nm);
int ssim_merge_scan(int n, const float* gx, const float* gy, const float* gvx,
                    const float* gvy, const float* px, const float* py,
                    const float* rr, const unsigned char* vis,
                    const unsigned char* is_ast, int nm, const int* m_alive,
                    float dt, int do_runaway, float bx, float by, float two_gm,
                    long long capacity, int* out_row, int* out_col,
                    float* out_frac, long long runaway_capacity,
                    long long* runaway, long long* counts, void* work,
                    size_t work_bytes, void* stream);

/* Bounce response of a contact list (`bounce_deltas`, film_producer.py:
 * 916-1019, which the reference runs on the host): the greedy
 * one-contact-per-body filter in list order, the entry time of
 * |dp + dv t| = r_i + r_j clipped to [-dt_y, dt_y], validity, approach
 * test, the impulse -(1+e) v_rel / (1/m_i + 1/m_j) and the mass-weighted
 * 1.1x overlap push, in FP64 on the f32 sample.  Dense outputs ddx, ddy,
 * ddvx, ddvy [n] (zero for untouched bodies); act_pairs = indices into the
 * list of the pairs that bounced (counts[0] of them, list order);
 * affected = bodies with a delta, ascending (counts[1]).  Synchronises. */
size_t ssim_bounce_work_bytes(int n, int npairs);
int ssim_bounce_deltas(int n, int npairs, const long long* hi, const long long* hj,
                       const float* x, const float* y, const float* vx,
                       const float* vy, const double* real_r, const double* mass,
                       double dt_y, double restitution, double* ddx, double* ddy,
                       double* ddvx, double* ddvy, long long* act_pairs,
                       long long* affected, long long* counts, void* work,
                       size_t work_bytes, void* stream);

/* Bounce deltas onto the resident state: body aff[q] (original index)
 * gets x, y, vx, vy += ddx/ddy/ddvx/ddvy[aff[q]] (the f64 length-n outputs
 * of ssim_bounce_deltas) in its integrator slot; the device form of the
 * host's apply_deltas after bounce_deltas (film_producer.py:1021-1026,
 * nbody_kernel.py:952-986).  aff: distinct indices; inv_kind /
 * inv_pos / massive / asteroids as for ssim_apply_merges. */
int ssim_add_body_deltas(int naff, const long long* aff, const double* ddx,
                         const double* ddy, const double* ddvx, const double* ddvy,
                         const int* inv_kind, const int* inv_pos, void* const* massive,
                         void* const* asteroids, void* stream);

/* The results of the massive-body contact scan applied to the state, on the
 * device (the host closure `apply_merges`, film_producer.py:781-861), in
 * list order — a later pair sees the masses and flags the earlier ones
 * left.  Per pair (pair_i, pair_j, pair_frac): a kernel contact time
 * contact_t[j] (NaN: none) replaces frac when |contact_t[j] - (k_ev + 1)|
 * <= 1 and is consumed (set to NaN) either way; with shatter_check, two
 * visible massive bodies, neither a star / black hole, meeting at
 * |v_j - v_i| >= 1.5 sqrt(2 G (m_i + m_j) / max(1e-12, r_i + r_j)) stop the
 * whole call (result[2] = i, result[3] = j, shatter_t = t0_days + (k_ev + 1
 * + frac) raster_days; the runaway list is not applied); otherwise, when
 * both are visible and m_i > 0, the lighter body merges into the heavier:
 * mass-weighted position and velocity (f32 sample, FP64 arithmetic), mass
 * sum, radius (r_a^3 + r_b^3)^(1/3), the lighter one deactivated (mass 0,
 * not visible).  Then every visible body of `runaway` is deactivated.
 * sample: f32 [x|y|vx|vy] rows of n bodies; mass, vis, real_r: the mirrors
 * by original index (updated); inv_kind / inv_pos map an original index to
 * the massive (-1) or asteroid (0) arrays and a position in them;
 * massive / asteroids: 7 device pointers each (x, y, vx, vy, m as double,
 * visible as uint8, radius as double).  events receives one 32-byte record
 * per merge or removal in the reference's ring format ("<dIIfIff": time in
 * days, a, b, new mass, kind 0 = merge / 2 = runaway with a = 0xFFFFFFFF,
 * x, y), at most npairs + nrun; result[0] = records, result[1] = merges.
 * One thread; device memory for every array. */
int ssim_apply_merges(int n, const float* sample, int npairs, const int* pair_i,
                      const int* pair_j, const double* pair_frac, int nrun,
                      const int* runaway, double* mass, unsigned char* vis,
                      double* real_r, const unsigned char* is_ast,
                      const unsigned char* is_star_bh, double* contact_t, int k_ev,
                      int shatter_check, double G, double t0_days, double raster_days,
                      const int* inv_kind, const int* inv_pos, void* const* massive,
                      void* const* asteroids, unsigned char* events, int* result,
                      double* shatter_t, void* stream);

/* ------------------------------------------------------------------ */
/* Learned near-field correction (MLP, BF16 MFMA, FP32 accumulation)   */
/* ------------------------------------------------------------------ */
/* None of these has a reference counterpart: the reference has no learned
 * component.  Network 32 -> 64 -> 64 -> 64 -> 3 with SiLU; parameters are
 * one FP32 block of ssim_mlp_param_count() floats (W1[64][32] b1 W2[64][64]
 * b2 W3[64][64] b3 W4[16][64] b4[16], rows 3..15 of W4/b4 zero) plus its
 * BF16 copy `wbf` (same offsets) used by the matrix cores.  Features are
 * BF16 [nb][32]. */
int ssim_mlp_param_count(void);
/* wbf = BF16 (round to nearest even) of params. */
int ssim_mlp_to_bf16(const float* params, unsigned short* wbf, void* stream);
/* Features of nb bodies (the first n_mass are masses, the rest tracers;
 * thread i fills body order[i]'s row, order = NULL: body i)
 * from the CIC mass mesh (mesh_n^3, C order, FP64 or FP32 per mesh_prec)
 * with origin[3] and cell h: the 27 node masses of the 3x3x3 stencil around
 * the nearest node divided by mass_unit, twice the offset from that node in
 * cells, 1 for a mass / 0 for a tracer, and a zero. */
int ssim_mlp_features(int nb, int n_mass, const int* order, const double* x,
                      const double* y, const double* z, int mesh_n,
                      const double* origin, double h,
                      int mesh_prec, const void* rho, double mass_unit,
                      unsigned short* feat, void* stream);
/* Forward pass.  out (optional) receives FP32 [nb][3]; if ax != NULL,
 * d = u * out is ADDED to (ax, ay, az) with u = scale, or, if relative != 0,
 * u = scale * |a| of the acceleration being corrected; for the first n_proj
 * bodies (the masses) the component of d along the velocity (vx, vy, vz) is
 * removed first, so the correction does no work on them.  For training pass
 * zs (FP32 [3][nb][64] pre-activations) and hs (BF16 [3][nb][64]
 * activations). */
int ssim_mlp_forward(int nb, const unsigned short* feat,
                     const unsigned short* wbf, const float* params,
                     float scale, int relative, float* out, double* ax,
                     double* ay, double* az, const double* vx, const double* vy,
                     const double* vz, int n_proj, float* zs, unsigned short* hs,
                     void* stream);
/* Training target: out[i][c] = (ref_c[i] - mod_c[i]) / u (FP32), the
 * residual of the model force against the exact direct sum, with u = a_unit
 * or, if relative != 0, u = a_unit * |mod[i]|; clipped to [-clip, clip]
 * when clip > 0. */
int ssim_mlp_residual(int nb, const double* ref_x, const double* ref_y,
                      const double* ref_z, const double* mod_x,
                      const double* mod_y, const double* mod_z, double a_unit,
                      int relative, double clip, float* out, void* stream);
/* gout = 2 (out - target) / nb; loss_part[block] = partial sums of the
 * squared error over blocks of 256 bodies. */
int ssim_mlp_mse_grad(int nb, const float* out, const float* target,
                      float* gout, float* loss_part, void* stream);
/* Backward pass: parameter gradient `grad` (FP32, param block layout) of
 * the loss whose output gradient is gout [nb][3], from the activations
 * recorded by ssim_mlp_forward.  Sums are in a fixed order. */
size_t ssim_mlp_train_work_bytes(int nb);
int ssim_mlp_backward(int nb, const unsigned short* feat,
                      const unsigned short* wbf, const float* zs,
                      const unsigned short* hs, const float* gout, float* grad,
                      void* work, size_t work_bytes, void* stream);
/* One Adam step (bias-corrected, step >= 1) on params; refreshes wbf. */
int ssim_mlp_adam(float* params, const float* grad, float* m, float* v,
                  float lr, float beta1, float beta2, float eps, int step,
                  unsigned short* wbf, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* SSIM_H */
