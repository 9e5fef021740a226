// Particle-mesh building blocks on gfx950: cloud-in-cell deposit and
// gather, min/max reduction for the adaptive grid, Green's functions in real
// and Fourier space, spectral products and the central-difference gradient.
//
// Mesh layout: C order with the x cell index slowest, flat index
// ((i * n1) + j) * n2 + k — the reference's `i * grid_n + j` in 2-D
// (pm_kernel.py:77-82).  Mesh values are MASS per node (what
// `cp.bincount` returns there); the Poisson Green's function divides by the
// cell volume.
//
// CIC cell indices are computed exactly like the reference: fx = (x - x0)/h
// in IEEE FP64 (the library is built without FMA contraction), i =
// floor(fx), clipped to [0, n-2] for an isolated mesh (pm_kernel.py:66-75)
// or wrapped mod n for a periodic one; the weights use the clipped index
// (tx = fx - i).  The products keep the reference's operand order
// (m * (1-tx)) * (1-ty) ... and the gather sums the corners in its order.
#include <cmath>

#include "ssim_internal.h"

namespace {

constexpr int PM_BLOCK = 256;

struct Grid {
  int n[3];       // n0 (x), n1 (y), n2 (z; 1 in 2-D)
  double o[3];    // origin
  double h;
  int periodic;
};

__device__ __forceinline__ void cell_of(double x, double o, double h, int n,
                                        int periodic, int& i, int& i1,
                                        double& t) {
  const double f = (x - o) / h;
  int c = (int)floor(f);
  if (periodic) {
    t = f - (double)c;
    c %= n;
    if (c < 0) c += n;
    i = c;
    i1 = (c + 1 == n) ? 0 : c + 1;
  } else {
    c = c < 0 ? 0 : (c > n - 2 ? n - 2 : c);
    t = f - (double)c;
    i = c;
    i1 = c + 1;
  }
}

template <typename M>
__device__ __forceinline__ void atomic_add(M* p, M v) {
  unsafeAtomicAdd(p, v);
}

template <int D, typename M>
__global__ __launch_bounds__(PM_BLOCK) void k_cic_deposit(
    int np, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ m, Grid g,
    M* __restrict__ mesh, int* __restrict__ cells) {
  const int p = blockIdx.x * PM_BLOCK + threadIdx.x;
  if (p >= np) return;
  int i, i1, j, j1, k = 0, k1 = 0;
  double tx, ty, tz = 0.0;
  cell_of(x[p], g.o[0], g.h, g.n[0], g.periodic, i, i1, tx);
  cell_of(y[p], g.o[1], g.h, g.n[1], g.periodic, j, j1, ty);
  if (D == 3) cell_of(z[p], g.o[2], g.h, g.n[2], g.periodic, k, k1, tz);
  if (cells) {
    cells[(size_t)D * p] = i;
    cells[(size_t)D * p + 1] = j;
    if (D == 3) cells[(size_t)D * p + 2] = k;
  }
  const double w = m[p];
  const double ux = 1.0 - tx, uy = 1.0 - ty;
  const long long n1 = g.n[1], n2 = g.n[2];
  if (D == 2) {
    atomic_add(&mesh[i * n1 + j], (M)(w * ux * uy));
    atomic_add(&mesh[i1 * n1 + j], (M)(w * tx * uy));
    atomic_add(&mesh[i * n1 + j1], (M)(w * ux * ty));
    atomic_add(&mesh[i1 * n1 + j1], (M)(w * tx * ty));
  } else {
    const double uz = 1.0 - tz;
    const double a00 = w * ux * uy, a10 = w * tx * uy;
    const double a01 = w * ux * ty, a11 = w * tx * ty;
    atomic_add(&mesh[(i * n1 + j) * n2 + k], (M)(a00 * uz));
    atomic_add(&mesh[(i1 * n1 + j) * n2 + k], (M)(a10 * uz));
    atomic_add(&mesh[(i * n1 + j1) * n2 + k], (M)(a01 * uz));
    atomic_add(&mesh[(i1 * n1 + j1) * n2 + k], (M)(a11 * uz));
    atomic_add(&mesh[(i * n1 + j) * n2 + k1], (M)(a00 * tz));
    atomic_add(&mesh[(i1 * n1 + j) * n2 + k1], (M)(a10 * tz));
    atomic_add(&mesh[(i * n1 + j1) * n2 + k1], (M)(a01 * tz));
    atomic_add(&mesh[(i1 * n1 + j1) * n2 + k1], (M)(a11 * tz));
  }
}

// ---------------------------------------------------------------------
// Reproducible CIC deposit without floating-point atomics
// (ssim_cic_deposit_tiles).  The mesh is cut into tiles of CE^D nodes
// (CE = 8 in 3-D, 32 in 2-D).  Each mass is filed under the tile of its
// cell (cell_of, the reference's cell rule) and a face group: bit a set
// when the cell's upper node on axis a lies in the next tile (or wraps),
// i.e. when the mass also reaches that neighbour.  A key kernel counts the
// (tile, group) bins with one integer atomic per distinct bin of a wave
// (masses arrive in runs when they are spatially ordered), a scan gives the
// bin starts and a ranked scatter lays the masses out in bin order with
// their cells and weights.  One wave per tile then takes the masses of its
// own tile and, from each tile below it on the axes of an offset e, only
// the groups that contain e (the masses whose upper nodes cross into this
// tile: ~1.4 masses read per mass, against 8 when every mass of the 2^D
// source tiles is scanned), adds every contribution that lands in the tile
// to an LDS accumulator as int64 fixed point, and writes each node once.
// Integer sums are exact in any order, so the mesh is bit-identical from
// run to run; each contribution is the reference's FP64 product ((m ux) uy)
// uz rounded to 2^(e - 62), with 2^e >= (masses the tile reads) x max |m|
// (over all masses: the key kernel's workgroup maxima, folded by the
// scatter), and each node is the nearest double of its exact fixed-point
// sum.
// ---------------------------------------------------------------------
template <int D>
struct CicTile {
  static constexpr int E = D == 3 ? 8 : 32;           // nodes per tile edge
  static constexpr int NN = D == 3 ? 512 : 1024;      // nodes per tile
  static constexpr int G = 1 << D;                    // face groups per tile
  static constexpr int R = D == 3 ? 20 : 6;           // source ranges per tile
};
constexpr int CT_THREADS = 64;   // one wave per tile

__device__ __forceinline__ int tiles_of(int n, int E) { return (n + E - 1) / E; }

template <int D>
__device__ __forceinline__ void cic_cells(const double* __restrict__ x,
                                          const double* __restrict__ y,
                                          const double* __restrict__ z, int p, const Grid& g,
                                          int (&c)[3], int (&c1)[3], double (&t)[3]) {
  cell_of(x[p], g.o[0], g.h, g.n[0], g.periodic, c[0], c1[0], t[0]);
  cell_of(y[p], g.o[1], g.h, g.n[1], g.periodic, c[1], c1[1], t[1]);
  c[2] = c1[2] = 0;
  t[2] = 0.0;
  if (D == 3) cell_of(z[p], g.o[2], g.h, g.n[2], g.periodic, c[2], c1[2], t[2]);
}

template <int D>
__global__ __launch_bounds__(PM_BLOCK) void k_cic_tile_key(
    int np, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ m, Grid g,
    int* __restrict__ key, int* __restrict__ rank, int* __restrict__ count,
    int* __restrict__ cells, double* __restrict__ part_max, int kbits) {
  constexpr int E = CicTile<D>::E, G = CicTile<D>::G;
  const int p = blockIdx.x * PM_BLOCK + threadIdx.x;
  const bool on = p < np;
  int k = 0;
  if (on) {
    int c[3], c1[3];
    double t[3];
    cic_cells<D>(x, y, z, p, g, c, c1, t);
    if (cells) {
      cells[(size_t)D * p] = c[0];
      cells[(size_t)D * p + 1] = c[1];
      if (D == 3) cells[(size_t)D * p + 2] = c[2];
    }
    const int t1 = tiles_of(g.n[1], E), t2 = D == 3 ? tiles_of(g.n[2], E) : 1;
    int grp = 0;
#pragma unroll
    for (int a = 0; a < D; ++a) grp |= (c1[a] / E != c[a] / E ? 1 : 0) << a;
    k = (((c[0] / E) * t1 + c[1] / E) * t2 + (D == 3 ? c[2] / E : 0)) * G + grp;
    key[p] = k;
  }
  // max |m| of this workgroup's masses (the tiles' fixed-point scale; the
  // scatter's first workgroup folds the workgroups' maxima)
  __shared__ double wmax[PM_BLOCK / 64];
  const int lane = threadIdx.x & 63;
  double am = on ? fabs(m[p]) : 0.0;
  for (int off = 32; off > 0; off >>= 1) am = fmax(am, __shfl_xor(am, off, 64));
  if (lane == 0) wmax[threadIdx.x >> 6] = am;
  __syncthreads();
  if (threadIdx.x == 0) {
    double bm = wmax[0];
    for (int q = 1; q < PM_BLOCK / 64; ++q) bm = fmax(bm, wmax[q]);
    part_max[blockIdx.x] = bm;
  }
  // the lanes with the same key (one ballot per key bit), one atomic per
  // distinct key of the wave by its first lane, all in one instruction (a
  // single round trip), and the ranks (slots below in the same bin)
  unsigned long long peers = __ballot(on);
  for (int b = 0; b < kbits; ++b) {
    const bool bit = (k >> b) & 1;
    const unsigned long long bal = __ballot(bit);
    peers &= bit ? bal : ~bal;
  }
  const int leader = on ? __ffsll((long long)peers) - 1 : lane;
  int base = 0;
  if (on && leader == lane) base = atomicAdd(&count[k], __popcll(peers));
  base = __shfl(base, leader, 64);
  if (on) rank[p] = base + __popcll(peers & ((1ull << lane) - 1ull));
}

// count[0..n) = 0 and the 64-bit maximum's two words (mx) = 0
__global__ __launch_bounds__(PM_BLOCK) void k_zero_count(int n, int* __restrict__ c,
                                                         int* __restrict__ mx) {
  const int i = blockIdx.x * PM_BLOCK + threadIdx.x;
  if (i < n) c[i] = 0;
  else if (i < n + 2) mx[i - n] = 0;
}

// the masses' records in bin order, slot = start[key] + rank: the weights
// and the mass (double4) and the cell (int4), computed again from the
// positions; workgroup 0 also folds the key kernel's nparts workgroup
// maxima of |m| into *mmax
template <int D>
__global__ __launch_bounds__(PM_BLOCK) void k_cic_tile_scatter(
    int np, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ m, Grid g,
    const int* __restrict__ key, const int* __restrict__ rank,
    const int* __restrict__ start, double4* __restrict__ trec, int4* __restrict__ treci,
    int nparts, const double* __restrict__ part_max, double* __restrict__ mmax) {
  if (blockIdx.x == 0) {
    __shared__ double wm[PM_BLOCK / 64];
    double bm = 0.0;
    for (int q = threadIdx.x; q < nparts; q += PM_BLOCK) bm = fmax(bm, part_max[q]);
    for (int off = 32; off > 0; off >>= 1) bm = fmax(bm, __shfl_xor(bm, off, 64));
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = bm;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int q = 1; q < PM_BLOCK / 64; ++q) bm = fmax(bm, wm[q]);
      *mmax = fmax(wm[0], bm);
    }
  }
  const int p = blockIdx.x * PM_BLOCK + threadIdx.x;
  if (p >= np) return;
  int c[3], c1[3];
  double t[3];
  cic_cells<D>(x, y, z, p, g, c, c1, t);
  const int slot = start[key[p]] + rank[p];
  trec[slot] = make_double4(t[0], t[1], t[2], m[p]);
  treci[slot] = make_int4(c[0], c[1], c[2], 0);
}

template <int D, typename M>
__global__ __launch_bounds__(CT_THREADS) void k_cic_deposit_tiles(
    Grid g, const int* __restrict__ start, const double4* __restrict__ trec,
    const int4* __restrict__ treci, const double* __restrict__ mmax,
    M* __restrict__ mesh) {
  constexpr int E = CicTile<D>::E, NN = CicTile<D>::NN, G = CicTile<D>::G,
                R = CicTile<D>::R;
  __shared__ unsigned long long acc[NN];
  const int tid = threadIdx.x;
  const int nt[3] = {tiles_of(g.n[0], E), tiles_of(g.n[1], E),
                     D == 3 ? tiles_of(g.n[2], E) : 1};
  const int T[3] = {(int)blockIdx.x / (nt[1] * nt[2]), ((int)blockIdx.x / nt[2]) % nt[1],
                    (int)blockIdx.x % nt[2]};
  const int lo[3] = {T[0] * E, T[1] * E, T[2] * E};
  // the source ranges: the whole of this tile, and from the tile below on
  // the axes of each offset e (wrapped when periodic, none below tile 0
  // otherwise, none on an axis of a single tile) the groups g that contain
  // e; r0[q] - pre[q] maps a virtual index v in [pre[q], pre[q + 1]) to
  // its record
  int rbase[R], pre[R + 1];
  int nr = 0, tot = 0;
#pragma unroll
  for (int e = 0; e < G; ++e) {
    int s3[3] = {T[0], T[1], T[2]};
    bool ok = true;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      if (!((e >> a) & 1)) continue;
      if (nt[a] == 1) ok = false;
      else if (T[a] > 0) s3[a] = T[a] - 1;
      else if (g.periodic) s3[a] = nt[a] - 1;
      else ok = false;
    }
    const int S = ((s3[0] * nt[1] + s3[1]) * nt[2] + s3[2]) * G;
    if (e == 0) {
      const int b0 = start[S], b1 = start[S + G];
      rbase[nr] = b0 - tot;
      pre[nr++] = tot;
      tot += b1 - b0;
      continue;
    }
#pragma unroll
    for (int gg = 0; gg < G; ++gg) {
      if ((gg & e) != e) continue;
      int b0 = 0, b1 = 0;
      if (ok) {
        b0 = start[S + gg];
        b1 = start[S + gg + 1];
      }
      rbase[nr] = b0 - tot;
      pre[nr++] = tot;
      tot += b1 - b0;
    }
  }
  pre[R] = tot;
  // fixed-point scale of this tile: 2^e >= (masses read) x max |m| (the
  // global maximum, from the key kernel)
  for (int q = tid; q < NN; q += CT_THREADS) acc[q] = 0ull;
  const double mx = *mmax;
  int ex = 0;
  frexp(mx * (double)(tot > 0 ? tot : 1), &ex);
  ex = ex < -900 ? -900 : ex;
  __syncthreads();
  const double scale = ldexp(1.0, 62 - ex), inv_scale = ldexp(1.0, ex - 62);
  for (int v = tid; v < tot; v += CT_THREADS) {
    int rb = rbase[0];
#pragma unroll
    for (int q = 1; q < R; ++q) rb = v >= pre[q] ? rbase[q] : rb;
    const int sidx = v + rb;
    const int4 ci = treci[sidx];
    const double4 r = trec[sidx];
    const int i = ci.x, j = ci.y, c = ci.z;
    const int i1 = (g.periodic && i + 1 == g.n[0]) ? 0 : i + 1;
    const int j1 = (g.periodic && j + 1 == g.n[1]) ? 0 : j + 1;
    const double tx = r.x, ty = r.y, tz = r.z;
    const double ux = 1.0 - tx, uy = 1.0 - ty;
    const int ni[2] = {i - lo[0], i1 - lo[0]}, nj[2] = {j - lo[1], j1 - lo[1]};
    if (D == 2) {
      const double w2[2][2] = {{r.w * ux * uy, r.w * ux * ty}, {r.w * tx * uy, r.w * tx * ty}};
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int e = 0; e < 2; ++e)
          if (ni[a] >= 0 && ni[a] < E && nj[e] >= 0 && nj[e] < E)
            atomicAdd(&acc[ni[a] * E + nj[e]],
                      (unsigned long long)__double2ll_rn(w2[a][e] * scale));
    } else {
      const int c1 = (g.periodic && c + 1 == g.n[2]) ? 0 : c + 1;
      const double uz = 1.0 - tz;
      const int nk[2] = {c - lo[2], c1 - lo[2]};
      const double a2[2][2] = {{r.w * ux * uy, r.w * ux * ty}, {r.w * tx * uy, r.w * tx * ty}};
      const double wz[2] = {uz, tz};
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int e = 0; e < 2; ++e)
#pragma unroll
          for (int f = 0; f < 2; ++f)
            if (ni[a] >= 0 && ni[a] < E && nj[e] >= 0 && nj[e] < E && nk[f] >= 0 &&
                nk[f] < E)
              atomicAdd(&acc[(ni[a] * E + nj[e]) * E + nk[f]],
                        (unsigned long long)__double2ll_rn(a2[a][e] * wz[f] * scale));
    }
  }
  __syncthreads();
  for (int q = tid; q < NN; q += CT_THREADS) {
    const int gi = lo[0] + (D == 3 ? q / (E * E) : q / E);
    const int gj = lo[1] + (D == 3 ? (q / E) % E : q % E);
    const int gk = D == 3 ? lo[2] + q % E : 0;
    if (gi < g.n[0] && gj < g.n[1] && (D == 2 || gk < g.n[2]))
      mesh[((long long)gi * g.n[1] + gj) * g.n[2] + gk] =
          (M)((double)(long long)acc[q] * inv_scale);
  }
}

// out_c[p] = sum over the CIC corners of field_c * weight, c < nf.
template <int D, typename M>
__global__ __launch_bounds__(PM_BLOCK) void k_cic_gather(
    int np, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, Grid g, int nf, const M* __restrict__ f0,
    const M* __restrict__ f1, const M* __restrict__ f2, double* __restrict__ o0,
    double* __restrict__ o1, double* __restrict__ o2) {
  const int p = blockIdx.x * PM_BLOCK + threadIdx.x;
  if (p >= np) return;
  int i, i1, j, j1, k = 0, k1 = 0;
  double tx, ty, tz = 0.0;
  cell_of(x[p], g.o[0], g.h, g.n[0], g.periodic, i, i1, tx);
  cell_of(y[p], g.o[1], g.h, g.n[1], g.periodic, j, j1, ty);
  if (D == 3) cell_of(z[p], g.o[2], g.h, g.n[2], g.periodic, k, k1, tz);
  const double ux = 1.0 - tx, uy = 1.0 - ty;
  const long long n1 = g.n[1], n2 = g.n[2];
  const M* fs[3] = {f0, f1, f2};
  double* os[3] = {o0, o1, o2};
  for (int c = 0; c < nf; ++c) {
    const M* f = fs[c];
    double v;
    if (D == 2) {
      v = (double)f[i * n1 + j] * ux * uy + (double)f[i1 * n1 + j] * tx * uy +
          (double)f[i * n1 + j1] * ux * ty + (double)f[i1 * n1 + j1] * tx * ty;
    } else {
      const double uz = 1.0 - tz;
      v = (double)f[(i * n1 + j) * n2 + k] * ux * uy * uz +
          (double)f[(i1 * n1 + j) * n2 + k] * tx * uy * uz +
          (double)f[(i * n1 + j1) * n2 + k] * ux * ty * uz +
          (double)f[(i1 * n1 + j1) * n2 + k] * tx * ty * uz +
          (double)f[(i * n1 + j) * n2 + k1] * ux * uy * tz +
          (double)f[(i1 * n1 + j) * n2 + k1] * tx * uy * tz +
          (double)f[(i * n1 + j1) * n2 + k1] * ux * ty * tz +
          (double)f[(i1 * n1 + j1) * n2 + k1] * tx * ty * tz;
    }
    os[c][p] = v;
  }
}

// Per-block min/max of up to 3 arrays -> part[block][2*c + {0,1}].
__global__ __launch_bounds__(PM_BLOCK) void k_minmax_partial(
    int n, int na, const double* __restrict__ a0, const double* __restrict__ a1,
    const double* __restrict__ a2, double* __restrict__ part) {
  __shared__ double smin[3][PM_BLOCK], smax[3][PM_BLOCK];
  const double* as[3] = {a0, a1, a2};
  for (int c = 0; c < na; ++c) {
    double lo = INFINITY, hi = -INFINITY;
    for (int i = blockIdx.x * PM_BLOCK + threadIdx.x; i < n;
         i += gridDim.x * PM_BLOCK) {
      const double v = as[c][i];
      lo = fmin(lo, v);
      hi = fmax(hi, v);
    }
    smin[c][threadIdx.x] = lo;
    smax[c][threadIdx.x] = hi;
  }
  __syncthreads();
  for (int w = PM_BLOCK / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int c = 0; c < na; ++c) {
        smin[c][threadIdx.x] = fmin(smin[c][threadIdx.x], smin[c][threadIdx.x + w]);
        smax[c][threadIdx.x] = fmax(smax[c][threadIdx.x], smax[c][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int c = 0; c < na; ++c) {
      part[(size_t)blockIdx.x * 6 + 2 * c] = smin[c][0];
      part[(size_t)blockIdx.x * 6 + 2 * c + 1] = smax[c][0];
    }
}

// Per-block extrema of the positions the next kick(half) + drift(full)
// produces, p = x + (v + a half) full, with the same operations (no fused
// multiply-add) as k_kick_snap and k_drift, so the box is exact.
__global__ __launch_bounds__(PM_BLOCK) void k_box_kdk_partial(
    int n, int na, const double* __restrict__ x0, const double* __restrict__ x1,
    const double* __restrict__ x2, const double* __restrict__ v0,
    const double* __restrict__ v1, const double* __restrict__ v2,
    const double* __restrict__ a0, const double* __restrict__ a1,
    const double* __restrict__ a2, double half, double full,
    double* __restrict__ part) {
  __shared__ double smin[3][PM_BLOCK], smax[3][PM_BLOCK];
  const double* xs[3] = {x0, x1, x2};
  const double* vs[3] = {v0, v1, v2};
  const double* as[3] = {a0, a1, a2};
  for (int c = 0; c < na; ++c) {
    double lo = INFINITY, hi = -INFINITY;
    for (int i = blockIdx.x * PM_BLOCK + threadIdx.x; i < n;
         i += gridDim.x * PM_BLOCK) {
      const double nv = vs[c][i] + as[c][i] * half;
      const double p = xs[c][i] + nv * full;
      lo = fmin(lo, p);
      hi = fmax(hi, p);
    }
    smin[c][threadIdx.x] = lo;
    smax[c][threadIdx.x] = hi;
  }
  __syncthreads();
  for (int w = PM_BLOCK / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int c = 0; c < na; ++c) {
        smin[c][threadIdx.x] = fmin(smin[c][threadIdx.x], smin[c][threadIdx.x + w]);
        smax[c][threadIdx.x] = fmax(smax[c][threadIdx.x], smax[c][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int c = 0; c < na; ++c) {
      part[(size_t)blockIdx.x * 6 + 2 * c] = smin[c][0];
      part[(size_t)blockIdx.x * 6 + 2 * c + 1] = smax[c][0];
    }
}

// ssim_kick_box: bodies per thread, all three components of each loaded at
// once (the box partial pass uses the same grid)
#ifndef SSIM_KB_ITEMS
#define SSIM_KB_ITEMS 4
#endif
constexpr int KB_ITEMS = SSIM_KB_ITEMS;

// The closing half kick of a step (v += a half, stored) fused with the box
// of the positions the next step's kick + drift will produce (as
// k_box_kdk_partial): every workgroup writes its extrema (wave shuffles,
// then the workgroup's waves) to part, which k_minmax_final folds into out.
// DRIFT: the next step's opening half kick and drift are stored too (v +=
// a half twice, x += v full: the operations of k_kick_box followed by
// k_kick_drift), and the box is that of the stored positions.  BOX false:
// no box (part is not written).
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

template <bool DRIFT, bool BOX>
__global__ __launch_bounds__(PM_BLOCK) void k_kick_box(
    int n, int na, double* __restrict__ x0, double* __restrict__ x1,
    double* __restrict__ x2, double* __restrict__ v0, double* __restrict__ v1,
    double* __restrict__ v2, const double* __restrict__ a0,
    const double* __restrict__ a1, const double* __restrict__ a2, double half,
    double full, double* __restrict__ part) {
  constexpr int NW = PM_BLOCK / 64;
  __shared__ double wmin[3][NW], wmax[3][NW];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  // a tile of KB_ITEMS bodies per thread (stride PM_BLOCK): every load of
  // every component is issued before the arithmetic and the stores
  const int base = blockIdx.x * PM_BLOCK * KB_ITEMS + threadIdx.x;
  double av[3][KB_ITEMS], vv[3][KB_ITEMS], xv[3][KB_ITEMS];
#pragma unroll
  for (int q = 0; q < KB_ITEMS; ++q) {
    const int i = min(base + q * PM_BLOCK, n - 1);
    av[0][q] = a0[i];
    vv[0][q] = v0[i];
    xv[0][q] = x0[i];
    if (na > 1) {
      av[1][q] = a1[i];
      vv[1][q] = v1[i];
      xv[1][q] = x1[i];
    }
    if (na > 2) {
      av[2][q] = a2[i];
      vv[2][q] = v2[i];
      xv[2][q] = x2[i];
    }
  }
  double* vs[3] = {v0, v1, v2};
  double* xs[3] = {x0, x1, x2};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (c >= na) break;
    double lo = INFINITY, hi = -INFINITY;
#pragma unroll
    for (int q = 0; q < KB_ITEMS; ++q) {
      const int i = base + q * PM_BLOCK;
      if (i < n) {
        const double v = vv[c][q] + av[c][q] * half;  // this step's closing kick
        const double nv = v + av[c][q] * half;        // the next step's opening kick
        const double p = xv[c][q] + nv * full;        // and its drift
        if (DRIFT) {
          vs[c][i] = nv;
          xs[c][i] = p;
        } else {
          vs[c][i] = v;
        }
        lo = fmin(lo, p);
        hi = fmax(hi, p);
      }
    }
    if (!BOX) continue;
    lo = wave_min(lo);
    hi = wave_max(hi);
    if (lane == 0) {
      wmin[c][wv] = lo;
      wmax[c][wv] = hi;
    }
  }
  if (!BOX) return;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int c = 0; c < na; ++c) {
      double lo = wmin[c][0], hi = wmax[c][0];
      for (int w = 1; w < NW; ++w) {
        lo = fmin(lo, wmin[c][w]);
        hi = fmax(hi, wmax[c][w]);
      }
      part[(size_t)blockIdx.x * 6 + 2 * c] = lo;
      part[(size_t)blockIdx.x * 6 + 2 * c + 1] = hi;
    }
  }
}

// Triangular-shaped-cloud (TSC) assignment on an isolated 3-D mesh, for the
// split force: a body at f = (x - o)/h (node units) spreads over the 3 nodes
// c-1, c, c+1 of each axis around its nearest node c = floor(f + 1/2), with
// d = f - c in [-1/2, 1/2) and weights
//   w(-1) = (1/2 - d)^2 / 2,  w(0) = 3/4 - d^2,  w(+1) = (1/2 + d)^2 / 2,
// which sum to 1 and reproduce linear functions exactly (sum w(k) k = d).
// Its window is sinc^3 per axis against sinc^2 for CIC, so aliasing of the
// mesh force drops several-fold.  Nodes outside the mesh are skipped (the
// split force keeps two empty cells around the bodies).
__device__ __forceinline__ void tsc_axis(double x, double o, double h, int& c,
                                         double w[3]) {
  const double f = (x - o) / h;
  const double r = floor(f + 0.5);
  const double d = f - r;
  c = (int)r;
  const double a = 0.5 - d, b = 0.5 + d;
  w[0] = 0.5 * a * a;
  w[1] = 0.75 - d * d;
  w[2] = 0.5 * b * b;
}

// TSC weight of the body at x on node `node` of its axis (0 beyond the
// 3-node stencil); the same arithmetic as tsc_axis.
__device__ __forceinline__ double tsc_node_weight(double x, double o, double h,
                                                  int node) {
  const double f = (x - o) / h;
  const double r = floor(f + 0.5);
  const double d = f - r;
  const int off = node - (int)r;
  if (off == 0) return 0.75 - d * d;
  if (off == -1) {
    const double a = 0.5 - d;
    return 0.5 * a * a;
  }
  if (off == 1) {
    const double b = 0.5 + d;
    return 0.5 * b * b;
  }
  return 0.0;
}

template <typename M>
__global__ __launch_bounds__(PM_BLOCK) void k_tsc_deposit(
    int np, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ m, Grid g,
    M* __restrict__ mesh) {
  const int p = blockIdx.x * PM_BLOCK + threadIdx.x;
  if (p >= np) return;
  int c[3];
  double w[3][3];
  tsc_axis(x[p], g.o[0], g.h, c[0], w[0]);
  tsc_axis(y[p], g.o[1], g.h, c[1], w[1]);
  tsc_axis(z[p], g.o[2], g.h, c[2], w[2]);
  const double mp = m[p];
  for (int a = 0; a < 3; ++a) {
    const int i = c[0] + a - 1;
    if (i < 0 || i >= g.n[0]) continue;
    for (int b = 0; b < 3; ++b) {
      const int j = c[1] + b - 1;
      if (j < 0 || j >= g.n[1]) continue;
      const double wab = mp * w[0][a] * w[1][b];
      const long long row = ((long long)i * g.n[1] + j) * g.n[2];
      for (int e = 0; e < 3; ++e) {
        const int k = c[2] + e - 1;
        if (k < 0 || k >= g.n[2]) continue;
        atomic_add(&mesh[row + k], (M)(wab * w[2][e]));
      }
    }
  }
}

// Tiled TSC deposit (ssim_tsc_deposit_tiles): one workgroup per tile of
// DT^3 blocks (E = DT B nodes a side) accumulates its nodes in LDS as
// int64 fixed point (contributions m w_x w_y w_z 2^s, rounded): integer
// addition is associative, so the LDS atomics give the same sums in any
// order and the mesh is bit-reproducible.  The masses come from the
// (DT + 2)^3 blocks around the tile (16 contiguous runs of the block-sorted
// order); those whose stencil reaches the tile are compacted into an LDS
// list (wave ballots), and every listed mass adds its contributions to the
// tile's nodes, one mass a lane.  Each node is written once, as acc 2^-s.
constexpr int DT = 2;          // blocks per tile edge
constexpr int DEP_T = 256;     // threads of a tile workgroup
constexpr int DEP_LCAP = 2048; // listed masses per round

template <typename M, int B>
__global__ __launch_bounds__(DEP_T) void k_tsc_deposit_tiles(
    const double4* __restrict__ body, const int* __restrict__ cell_start, int nb,
    int n, double scale, double inv_scale, M* __restrict__ mesh) {
  constexpr int E = DT * B, NN = E * E * E;
  __shared__ unsigned long long acc[NN];
  __shared__ int list[DEP_LCAP];
  __shared__ int nlist;
  const int t = threadIdx.x, lane = t & 63;
  const int ntile = (nb + DT - 1) / DT;
  const int tz = blockIdx.x % ntile, ty = (blockIdx.x / ntile) % ntile,
            tx = blockIdx.x / (ntile * ntile);
  const int lo0 = tx * E, lo1 = ty * E, lo2 = tz * E;
  for (int q = t; q < NN; q += DEP_T) acc[q] = 0ull;
  if (t == 0) nlist = 0;
  __syncthreads();
  // the 16 rows of blocks around the tile: z-runs of up to DT + 2 blocks
  const int bz0 = max(tz * DT - 1, 0), bz1 = min(tz * DT + DT, nb - 1);
  // one list entry: add a listed mass's contributions to the tile
  auto add_mass = [&](int sidx) {
    const double4 p = body[sidx];
    const double f[3] = {p.x, p.y, p.z};
    const int lo[3] = {lo0, lo1, lo2};
    double w[3][3];
    int r[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double rr = floor(f[c] + 0.5);
      const double d = f[c] - rr;
      r[c] = (int)rr - lo[c];  // centre node, tile-relative
      const double a = 0.5 - d, b = 0.5 + d;
      w[c][0] = 0.5 * a * a;
      w[c][1] = 0.75 - d * d;
      w[c][2] = 0.5 * b * b;
    }
    const double ms = p.w * scale;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int i = r[0] + a - 1;
      if (i < 0 || i >= E) continue;
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const int j = r[1] + b - 1;
        if (j < 0 || j >= E) continue;
        const double wab = ms * w[0][a] * w[1][b];
#pragma unroll
        for (int e = 0; e < 3; ++e) {
          const int k = r[2] + e - 1;
          if (k < 0 || k >= E) continue;
          atomicAdd(&acc[(i * E + j) * E + k],
                    (unsigned long long)__double2ll_rn(wab * w[2][e]));
        }
      }
    }
  };
  for (int row = 0; row < (DT + 2) * (DT + 2); ++row) {
    const int bx = tx * DT - 1 + row / (DT + 2), by = ty * DT - 1 + row % (DT + 2);
    if (bx < 0 || bx >= nb || by < 0 || by >= nb) continue;  // uniform
    const int cb = (bx * nb + by) * nb;
    const int s0 = cell_start[cb + bz0], s1 = cell_start[cb + bz1 + 1];
    for (int base = s0; base < s1; base += DEP_T) {
      const int sidx = base + t;
      bool rel = false;
      if (sidx < s1) {
        const double4 p = body[sidx];
        const int r0 = (int)floor(p.x + 0.5) - lo0, r1 = (int)floor(p.y + 0.5) - lo1,
                  r2 = (int)floor(p.z + 0.5) - lo2;
        rel = r0 >= -1 && r0 <= E && r1 >= -1 && r1 <= E && r2 >= -1 && r2 <= E;
      }
      const unsigned long long mask = __ballot(rel);
      int wbase = 0;
      if (lane == 0 && mask) wbase = atomicAdd(&nlist, __popcll(mask));
      wbase = __shfl(wbase, 0, 64);
      if (rel) list[wbase + __popcll(mask & ((1ull << lane) - 1ull))] = sidx;
      __syncthreads();
      const int cnt = nlist;
      __syncthreads();
      if (cnt > DEP_LCAP - DEP_T) {  // flush before the next round could overflow
        for (int q = t; q < cnt; q += DEP_T) add_mass(list[q]);
        __syncthreads();
        if (t == 0) nlist = 0;
        __syncthreads();
      }
    }
  }
  const int cnt = nlist;
  for (int q = t; q < cnt; q += DEP_T) add_mass(list[q]);
  __syncthreads();
  for (int q = t; q < NN; q += DEP_T) {
    const int i = q / (E * E), j = (q / E) % E, k = q % E;
    const int gi = lo0 + i, gj = lo1 + j, gk = lo2 + k;
    if (gi < n && gj < n && gk < n)
      mesh[((long long)gi * n + gj) * n + gk] = (M)((double)(long long)acc[q] * inv_scale);
  }
}

// The same deposit without atomics: one thread per node sums, in a fixed
// order, the bodies binned at the 27 nodes around it (bins = nearest node,
// stable order inside a bin), so the mesh is bit-reproducible.  A body binned
// one node off by rounding at a bin edge has weight ~0 on the nodes it is
// then missed by (the TSC weights vanish at |d| = 1/2 on the far side).
// (x, y, z, m) of the binned bodies in bin order: one 32-byte load per
// visit in the node loop below instead of four scattered 8-byte loads.
__global__ __launch_bounds__(PM_BLOCK) void k_pack_binned(
    int np, const int* __restrict__ bin_order, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z,
    const double* __restrict__ m, double4* __restrict__ packed) {
  const int s = blockIdx.x * PM_BLOCK + threadIdx.x;
  if (s >= np) return;
  const int p = bin_order[s];
  packed[s] = make_double4(x[p], y[p], z[p], m[p]);
}

template <typename M>
__global__ __launch_bounds__(PM_BLOCK) void k_tsc_deposit_binned(
    const double4* __restrict__ body, Grid g, const int* __restrict__ bin_start,
    M* __restrict__ mesh) {
  const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
  const long long node = (long long)blockIdx.x * PM_BLOCK + threadIdx.x;
  if (node >= (long long)n0 * n1 * n2) return;
  const int k = (int)(node % n2), j = (int)((node / n2) % n1),
            i = (int)(node / ((long long)n1 * n2));
  double acc = 0.0;
  for (int a = -1; a <= 1; ++a) {
    const int bi = i + a;
    if (bi < 0 || bi >= n0) continue;
    for (int b = -1; b <= 1; ++b) {
      const int bj = j + b;
      if (bj < 0 || bj >= n1) continue;
      const long long row = ((long long)bi * n1 + bj) * n2;
      const int k0 = max(k - 1, 0), k1 = min(k + 1, n2 - 1);
      // bins k0..k1 of one row are contiguous in the sorted order
      const int s0 = bin_start[row + k0], s1 = bin_start[row + k1 + 1];
      for (int s = s0; s < s1; ++s) {
        const double4 q = body[s];
        const double wx = tsc_node_weight(q.x, g.o[0], g.h, i);
        const double wy = tsc_node_weight(q.y, g.o[1], g.h, j);
        const double wz = tsc_node_weight(q.z, g.o[2], g.h, k);
        acc += q.w * wx * wy * wz;
      }
    }
  }
  mesh[node] = (M)acc;
}

// TSC weight on node `node` of a body at node-unit coordinate f = (x-o)/h
// (the value tsc_node_weight computes first).
__device__ __forceinline__ double tsc_weight_f(double f, int node) {
  const double r = floor(f + 0.5);
  const double d = f - r;
  const int off = node - (int)r;
  if (off == 0) return 0.75 - d * d;
  if (off < -1 || off > 1) return 0.0;
  const double a = off < 0 ? 0.5 - d : 0.5 + d;
  return 0.5 * a * a;
}

// The atomic-free TSC deposit over blocked bins: node (i, j, k) sums the
// bodies of the 27 bins around it (a bin holds the bodies whose stencil is
// centred on its node, so every body of them touches the node).  The bin
// key is a sum of per-axis terms, key = X(i) + Y(j) + Z(k) with
//   X(i) = (i/B) nb^2 B^3 + (i%B) B^2,  Y(j) = (j/B) nb B^3 + (j%B) B,
//   Z(k) = (k/B) B^3 + k%B,
// so each node needs nine divisions.  The z bins k-1..k+1 of a row are one
// run of the sorted order, or two when they straddle a block boundary
// (B >= 2: never more); both are walked as one index range so every lane
// of a wave runs one loop.  body[] holds (f_x, f_y, f_z, m) with
// f = (x - o)/h computed by ssim_split_bin.
// BC > 0: the block edge as a compile-time constant (its divisions become
// multiply-shifts); BC == 0: the runtime value Brt.
template <typename M, int BC>
__global__ __launch_bounds__(PM_BLOCK) void k_tsc_deposit_blocked(
    const double4* __restrict__ body, Grid g, int Brt, int nb,
    const int* __restrict__ bin_start, M* __restrict__ mesh) {
  const int B = BC > 0 ? BC : Brt;
  const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
  const long long node = (long long)blockIdx.x * PM_BLOCK + threadIdx.x;
  if (node >= (long long)n0 * n1 * n2) return;
  const int k = (int)(node % n2), j = (int)((node / n2) % n1),
            i = (int)(node / ((long long)n1 * n2));
  const int B3 = B * B * B;
  int X[3], Y[3], Z[3], zb[3];
  for (int q = 0; q < 3; ++q) {
    const int ii = max(i + q - 1, 0), jj = max(j + q - 1, 0), kk = max(k + q - 1, 0);
    const int bi = ii / B, bj = jj / B, bk = kk / B;
    X[q] = bi * nb * nb * B3 + (ii - bi * B) * B * B;
    Y[q] = bj * nb * B3 + (jj - bj * B) * B;
    Z[q] = bk * B3 + (kk - bk * B);
    zb[q] = bk;
  }
  const int q0 = k > 0 ? 0 : 1, q1 = k + 1 < n2 ? 2 : 1;
  // first z bin of the second run (q1 + 1: one run)
  int qs = q1 + 1;
  for (int q = q0 + 1; q <= q1; ++q)
    if (zb[q] != zb[q0] && qs > q1) qs = q;
  // every row's bin range first (rows outside the mesh: empty), so the
  // bin_start loads are in flight together; then the bodies
  const bool split = qs <= q1;
  const int ze = split ? Z[qs - 1] : Z[q1];
  // second run: [Z[qs], Z[q1]] when split, else an empty range
  const int zb0 = split ? Z[qs] : Z[q0], zb1 = split ? Z[q1] + 1 : Z[q0];
  int s0[9], la[9], t0[9], len[9];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const int r = 3 * a + b;
      const bool in = i + a - 1 >= 0 && i + a - 1 < n0 && j + b - 1 >= 0 && j + b - 1 < n1;
      const int xy = in ? X[a] + Y[b] : 0;
      const int a0 = bin_start[xy + Z[q0]], a1 = bin_start[xy + ze + 1];
      const int b0 = bin_start[xy + zb0], b1 = bin_start[xy + zb1];
      s0[r] = a0;
      la[r] = in ? a1 - a0 : 0;
      t0[r] = b0;
      len[r] = in ? la[r] + (b1 - b0) : 0;
    }
  double acc = 0.0;
#pragma unroll
  for (int r = 0; r < 9; ++r)
    for (int u = 0; u < len[r]; ++u) {
      const double4 p = body[u < la[r] ? s0[r] + u : t0[r] + (u - la[r])];
      acc += p.w * tsc_weight_f(p.x, i) * tsc_weight_f(p.y, j) * tsc_weight_f(p.z, k);
    }
  mesh[node] = (M)acc;
}

template <typename M>
__global__ __launch_bounds__(PM_BLOCK) void k_tsc_gather(
    int np, const int* __restrict__ order, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, Grid g, const M* __restrict__ f0,
    const M* __restrict__ f1, const M* __restrict__ f2, double* __restrict__ o0,
    double* __restrict__ o1, double* __restrict__ o2, int accumulate) {
  const int i_ = blockIdx.x * PM_BLOCK + threadIdx.x;
  if (i_ >= np) return;
  const int p = order ? order[i_] : i_;
  int c[3];
  double w[3][3];
  tsc_axis(x[p], g.o[0], g.h, c[0], w[0]);
  tsc_axis(y[p], g.o[1], g.h, c[1], w[1]);
  tsc_axis(z[p], g.o[2], g.h, c[2], w[2]);
  double v0 = 0.0, v1 = 0.0, v2 = 0.0;
  for (int a = 0; a < 3; ++a) {
    const int i = c[0] + a - 1;
    if (i < 0 || i >= g.n[0]) continue;
    for (int b = 0; b < 3; ++b) {
      const int j = c[1] + b - 1;
      if (j < 0 || j >= g.n[1]) continue;
      const double wab = w[0][a] * w[1][b];
      const long long row = ((long long)i * g.n[1] + j) * g.n[2];
      for (int e = 0; e < 3; ++e) {
        const int k = c[2] + e - 1;
        if (k < 0 || k >= g.n[2]) continue;
        const double wt = wab * w[2][e];
        v0 += wt * (double)f0[row + k];
        v1 += wt * (double)f1[row + k];
        v2 += wt * (double)f2[row + k];
      }
    }
  }
  if (accumulate) {
    o0[p] += v0;
    o1[p] += v1;
    o2[p] += v2;
  } else {
    o0[p] = v0;
    o1[p] = v1;
    o2[p] = v2;
  }
}

// The same gather from three FP32 fields interleaved as float4 [n^3]
// (one 16-byte load per node instead of three 4-byte loads).
// Learned correction folded into the gather (k_tsc_gather4<true>): corr
// holds the network's three outputs per body; with the split force s at the
// body (after the accumulate), u = scale |s| (relative) or scale, d = u corr,
// bodies below n_proj lose the component of d along their velocity, and
// s + d is stored: the operations of ssim_mlp_forward's own update.
struct F3 {
  float x, y, z;
};
typedef float f3_t __attribute__((ext_vector_type(3)));

struct Corr {
  const float* out;
  float scale;
  int relative;
  const double *v0, *v1, *v2;
  int n_proj;
};

// BOX (k_tsc_gather4<., ., true>): with the stored acceleration a of each
// body, the position the next closing and opening half kicks and drift
// give, x + ((v + a half) + a half) full (the operations of
// ssim_kick_drift_box), is folded into per-workgroup extrema written to
// bx.part (the layout k_minmax_final folds).
struct BoxOut {
  const double *v0, *v1, *v2;
  double half, full;
  double* part;
};

template <bool CORR, bool BUF, bool BOX>
__global__ __launch_bounds__(PM_BLOCK) void k_tsc_gather4(
    int np, const int* __restrict__ order, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, Grid g, const float4* __restrict__ f,
    double* __restrict__ o0, double* __restrict__ o1, double* __restrict__ o2,
    int accumulate, Corr cr, BoxOut bx) {
  // XCD-contiguous ranges of the visited bodies (in cell order, or in
  // index order when that is spatially coherent): an XCD's L2 then holds
  // the part of the mesh its bodies touch
  const int i_ = ssim_xcd_block(blockIdx.x, gridDim.x) * PM_BLOCK + threadIdx.x;
  if (!BOX && i_ >= np) return;
  double blo[3] = {INFINITY, INFINITY, INFINITY}, bhi[3] = {-INFINITY, -INFINITY, -INFINITY};
  if (i_ < np) {
  const int p = order ? order[i_] : i_;
  // every load that does not depend on the stencil is issued first
  // (uniform branches), so its latency hides behind the mesh loads
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  if (accumulate) {
    a0 = o0[p];
    a1 = o1[p];
    a2 = o2[p];
  }
  float r0 = 0.f, r1 = 0.f, r2 = 0.f;
  double w0 = 0.0, w1 = 0.0, w2 = 0.0;
  if (CORR) {
    r0 = cr.out[(size_t)p * 3 + 0];
    r1 = cr.out[(size_t)p * 3 + 1];
    r2 = cr.out[(size_t)p * 3 + 2];
    if (cr.n_proj > 0) {
      const int pv = p < cr.n_proj ? p : 0;
      w0 = cr.v0[pv];
      w1 = cr.v1[pv];
      w2 = cr.v2[pv];
    }
  }
  double bv0 = 0.0, bv1 = 0.0, bv2 = 0.0;
  if (BOX) {
    bv0 = bx.v0[p];
    bv1 = bx.v1[p];
    bv2 = bx.v2[p];
  }
  const double px = x[p], py = y[p], pz = z[p];
  int c[3];
  double w[3][3];
  tsc_axis(px, g.o[0], g.h, c[0], w[0]);
  tsc_axis(py, g.o[1], g.h, c[1], w[1]);
  tsc_axis(pz, g.o[2], g.h, c[2], w[2]);
  // branch-free stencil: nodes outside the mesh are clamped into it with
  // weight 0 (an exact zero term), so all 27 loads can be in flight at once
  int ix[3][3];
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int i = c[d] + a - 1;
      const bool in = i >= 0 && i < g.n[d];
      ix[d][a] = in ? i : 0;
      w[d][a] = in ? w[d][a] : 0.0;
    }
  // the three fields of a node (12 of its 16 bytes: one dwordx3 load); a
  // mesh under 2 GiB is read with buffer loads (32-bit offsets from one
  // resource, no 64-bit address registers, so all 27 stay in flight)
  F3 q[3][3][3];
  if (BUF) {
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void*)f, (short)0, (int)((long long)g.n[0] * g.n[1] * g.n[2] * 16), 0x00020000);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const unsigned row = (unsigned)((ix[0][a] * g.n[1] + ix[1][b]) * g.n[2]);
#pragma unroll
        for (int e = 0; e < 3; ++e) {
          const f3_t v = __builtin_amdgcn_raw_buffer_load_b96(
              rs, (row + (unsigned)ix[2][e]) * 16u, 0, 0);
          q[a][b][e] = F3{v.x, v.y, v.z};
        }
      }
    __builtin_amdgcn_sched_barrier(0);  // all 27 loads issued before any use
  } else {
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const long long row = ((long long)ix[0][a] * g.n[1] + ix[1][b]) * g.n[2];
#pragma unroll
        for (int e = 0; e < 3; ++e)
          q[a][b][e] = *reinterpret_cast<const F3*>(&f[row + ix[2][e]]);
      }
  }
  double v0 = 0.0, v1 = 0.0, v2 = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const double wab = w[0][a] * w[1][b];
#pragma unroll
      for (int e = 0; e < 3; ++e) {
        const double wt = wab * w[2][e];
        v0 += wt * (double)q[a][b][e].x;
        v1 += wt * (double)q[a][b][e].y;
        v2 += wt * (double)q[a][b][e].z;
      }
    }
  double s0 = accumulate ? a0 + v0 : v0;
  double s1 = accumulate ? a1 + v1 : v1;
  double s2 = accumulate ? a2 + v2 : v2;
  if (CORR) {
    const double u = cr.relative ? (double)cr.scale * sqrt(s0 * s0 + s1 * s1 + s2 * s2)
                                 : (double)cr.scale;
    double d0 = u * (double)r0;
    double d1 = u * (double)r1;
    double d2 = u * (double)r2;
    if (p < cr.n_proj) {
      const double vv = w0 * w0 + w1 * w1 + w2 * w2;
      if (vv > 0.0) {
        const double cc = (d0 * w0 + d1 * w1 + d2 * w2) / vv;
        d0 -= cc * w0;
        d1 -= cc * w1;
        d2 -= cc * w2;
      }
    }
    s0 += d0;
    s1 += d1;
    s2 += d2;
  }
  o0[p] = s0;
  o1[p] = s1;
  o2[p] = s2;
  if (BOX) {
    const double sv[3] = {s0, s1, s2}, vv[3] = {bv0, bv1, bv2}, xv[3] = {px, py, pz};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double nv = (vv[q] + sv[q] * bx.half) + sv[q] * bx.half;
      blo[q] = bhi[q] = xv[q] + nv * bx.full;
    }
  }
  }
  if (BOX) {
    constexpr int NW = PM_BLOCK / 64;
    __shared__ double wmin[3][NW], wmax[3][NW];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double lo = wave_min(blo[q]), hi = wave_max(bhi[q]);
      if (lane == 0) {
        wmin[q][wv] = lo;
        wmax[q][wv] = hi;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0)
      for (int q = 0; q < 3; ++q) {
        double lo = wmin[q][0], hi = wmax[q][0];
        for (int w2 = 1; w2 < NW; ++w2) {
          lo = fmin(lo, wmin[q][w2]);
          hi = fmax(hi, wmax[q][w2]);
        }
        bx.part[(size_t)blockIdx.x * 6 + 2 * q] = lo;
        bx.part[(size_t)blockIdx.x * 6 + 2 * q + 1] = hi;
      }
  }
}

// One workgroup folds the per-block extrema (min and max are exact, so the
// order of the fold does not matter).
__global__ __launch_bounds__(PM_BLOCK) void k_minmax_final(
    int nblk, int na, const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double smin[PM_BLOCK], smax[PM_BLOCK];
  for (int c = 0; c < na; ++c) {
    double lo = INFINITY, hi = -INFINITY;
    for (int b = threadIdx.x; b < nblk; b += PM_BLOCK) {
      lo = fmin(lo, part[(size_t)b * 6 + 2 * c]);
      hi = fmax(hi, part[(size_t)b * 6 + 2 * c + 1]);
    }
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    __syncthreads();
    for (int w = PM_BLOCK / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w) {
        smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + w]);
        smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      out[2 * c] = smin[0];
      out[2 * c + 1] = smax[0];
    }
    __syncthreads();
  }
}

// Real-space force kernel on a zero-padded (2N)^D mesh: K_c(d) =
// -G d_c / (|d|^2 + eps2)^(3/2), d = off * h with off = idx < N ? idx :
// idx - 2N (build_force_kernels, pm_kernel.py:42-59).  Component c of D.
// With rs > 0 only the long-range part is kept: K_c * (1 - gS(|d|)), gS the
// tapered split function (ssim_split_g; zero beyond rcut, so the mesh has
// the full K_c there); the complement K_c gS is summed directly by the
// near-field kernel.
template <typename M>
__global__ __launch_bounds__(PM_BLOCK) void k_force_kernel(
    int D, int N, int comp, double h, double eps2, double G, double rs,
    double rcut, M* __restrict__ out) {
  const long long P = 2LL * N;
  const long long total = (D == 2) ? P * P : P * P * P;
  const long long t = (long long)blockIdx.x * PM_BLOCK + threadIdx.x;
  if (t >= total) return;
  long long idx[3];
  if (D == 2) {
    idx[0] = t / P;
    idx[1] = t % P;
    idx[2] = 0;
  } else {
    idx[0] = t / (P * P);
    idx[1] = (t / P) % P;
    idx[2] = t % P;
  }
  double d[3];
  for (int c = 0; c < 3; ++c) {
    const double off = (double)(idx[c] < N ? idx[c] : idx[c] - P);
    d[c] = (c < D) ? off * h : 0.0;
  }
  double r2 = d[0] * d[0] + d[1] * d[1];
  if (D == 3) r2 = r2 + d[2] * d[2];
  const double rr = r2;
  r2 = r2 + eps2;
  const double inv_r3 = 1.0 / (r2 * sqrt(r2));
  double k = ((-G) * d[comp]) * inv_r3;
  if (rs > 0.0) {
    // the near-field sum takes F g S; the mesh keeps the rest, which is the
    // whole force beyond rcut, so no part of any pair is dropped
    k = k * (1.0 - ssim_split_g(sqrt(rr), rs, rcut));
    // offset -N pairs no two nodes of the N-mesh; zero there the kernel is
    // exactly odd, so its spectrum is imaginary
    for (int c = 0; c < D; ++c)
      if (idx[c] == N) k = 0.0;
  }
  out[t] = (M)k;
}

// Divide a half spectrum on the padded P^D mesh by the squared assignment
// window, prod_c sinc^(2 order)(pi f_c / P) (order 2: CIC, 3: TSC), undoing
// assignment + interpolation smoothing.
template <typename M>
__global__ __launch_bounds__(PM_BLOCK) void k_deconvolve(int D, int P, int order,
                                                         M* __restrict__ spec) {
  const int W = P / 2 + 1;
  const long long total = (D == 2) ? (long long)P * W : (long long)P * P * W;
  const long long t = (long long)blockIdx.x * PM_BLOCK + threadIdx.x;
  if (t >= total) return;
  int f[3];
  f[D - 1] = (int)(t % W);
  if (D == 2) {
    const int a = (int)(t / W);
    f[0] = a <= P / 2 ? a : a - P;
  } else {
    const int b = (int)((t / W) % P);
    const int a = (int)(t / ((long long)W * P));
    f[0] = a <= P / 2 ? a : a - P;
    f[1] = b <= P / 2 ? b : b - P;
  }
  double w = 1.0;
  for (int c = 0; c < D; ++c) {
    if (f[c] == 0) continue;
    const double x = M_PI * (double)f[c] / (double)P;
    const double s = sin(x) / x;
    for (int q = 0; q < 2 * order; ++q) w *= s;
  }
  spec[2 * t] = (M)(spec[2 * t] / w);
  spec[2 * t + 1] = (M)(spec[2 * t + 1] / w);
}

// out = a * b on nc complex values (in place allowed).
template <typename M>
__global__ __launch_bounds__(PM_BLOCK) void k_cmul(long long nc,
                                                   const M* __restrict__ a,
                                                   const M* __restrict__ b,
                                                   M* __restrict__ out) {
  const long long t = (long long)blockIdx.x * PM_BLOCK + threadIdx.x;
  if (t >= nc) return;
  const M ar = a[2 * t], ai = a[2 * t + 1];
  const M br = b[2 * t], bi = b[2 * t + 1];
  out[2 * t] = ar * br - ai * bi;
  out[2 * t + 1] = ar * bi + ai * br;
}

template <typename M>
__global__ __launch_bounds__(PM_BLOCK) void k_imag(long long nc, const M* __restrict__ a,
                                                   M* __restrict__ out) {
  const long long t = (long long)blockIdx.x * PM_BLOCK + threadIdx.x;
  if (t < nc) out[t] = a[2 * t + 1];
}

// Periodic Poisson in Fourier space on the rfftn half spectrum:
// phi_k = -4 pi G rho_k / k^2 with rho = mass / h^D, k = 2 pi f / (n h),
// f the signed integer frequency; the k = 0 mode is set to 0.
template <typename M>
__global__ __launch_bounds__(PM_BLOCK) void k_poisson(int D, int n0, int n1,
                                                      int n2, double h,
                                                      double G, M* __restrict__ spec) {
  const int nl = (D == 2) ? n1 : n2;    // last (halved) axis length
  const int W = nl / 2 + 1;
  const long long total = (D == 2) ? (long long)n0 * W : (long long)n0 * n1 * W;
  const long long t = (long long)blockIdx.x * PM_BLOCK + threadIdx.x;
  if (t >= total) return;
  int f[3];
  const int c = (int)(t % W);
  if (D == 2) {
    const int a = (int)(t / W);
    f[0] = a <= n0 / 2 ? a : a - n0;
    f[1] = c;
    f[2] = 0;
  } else {
    const int b = (int)((t / W) % n1);
    const int a = (int)(t / ((long long)W * n1));
    f[0] = a <= n0 / 2 ? a : a - n0;
    f[1] = b <= n1 / 2 ? b : b - n1;
    f[2] = c;
  }
  const int ns[3] = {n0, (D == 2) ? nl : n1, n2};
  double k2 = 0.0;
  for (int q = 0; q < D; ++q) {
    const double k = 2.0 * M_PI * (double)f[q] / ((double)ns[q] * h);
    k2 += k * k;
  }
  double vol = h * h;
  if (D == 3) vol *= h;
  double g = (k2 > 0.0) ? (-4.0 * M_PI * G) / (k2 * vol) : 0.0;
  spec[2 * t] = (M)(spec[2 * t] * g);
  spec[2 * t + 1] = (M)(spec[2 * t + 1] * g);
}

// Central difference a_c = -(phi[+1] - phi[-1]) / (2h), periodic wrap.
template <typename M>
__global__ __launch_bounds__(PM_BLOCK) void k_gradient(
    int D, int n0, int n1, int n2, double h, const M* __restrict__ phi,
    M* __restrict__ g0, M* __restrict__ g1, M* __restrict__ g2) {
  const long long N = (long long)n0 * n1 * (D == 3 ? n2 : 1);
  const long long t = (long long)blockIdx.x * PM_BLOCK + threadIdx.x;
  if (t >= N) return;
  const int nz = (D == 3) ? n2 : 1;
  const int k = (int)(t % nz);
  const int j = (int)((t / nz) % n1);
  const int i = (int)(t / ((long long)nz * n1));
  auto at = [&](int a, int b, int c) -> double {
    return (double)phi[((long long)a * n1 + b) * nz + c];
  };
  const double s = -1.0 / (2.0 * h);
  const int ip = i + 1 == n0 ? 0 : i + 1, im = i == 0 ? n0 - 1 : i - 1;
  const int jp = j + 1 == n1 ? 0 : j + 1, jm = j == 0 ? n1 - 1 : j - 1;
  g0[t] = (M)((at(ip, j, k) - at(im, j, k)) * s);
  g1[t] = (M)((at(i, jp, k) - at(i, jm, k)) * s);
  if (D == 3) {
    const int kp = k + 1 == nz ? 0 : k + 1, km = k == 0 ? nz - 1 : k - 1;
    g2[t] = (M)((at(i, j, kp) - at(i, j, km)) * s);
  }
}

// Copy an n^D block into the corner of a zeroed (2n)^D block (and back).
template <typename M>
__global__ __launch_bounds__(PM_BLOCK) void k_pad(int D, int n, const M* __restrict__ src,
                                                  M* __restrict__ dst, int crop) {
  const long long total = (D == 2) ? (long long)n * n : (long long)n * n * n;
  const long long t = (long long)blockIdx.x * PM_BLOCK + threadIdx.x;
  if (t >= total) return;
  const long long P = 2LL * n;
  long long big;
  if (D == 2)
    big = (t / n) * P + (t % n);
  else
    big = ((t / ((long long)n * n)) * P + (t / n) % n) * P + (t % n);
  if (crop)
    dst[t] = src[big];
  else
    dst[big] = src[t];
}

Grid make_grid(int dim, const int* n, const double* origin, double h,
               int periodic) {
  Grid g{};
  for (int c = 0; c < 3; ++c) {
    g.n[c] = (c < dim) ? n[c] : 1;
    g.o[c] = (c < dim) ? origin[c] : 0.0;
  }
  g.h = h;
  g.periodic = periodic;
  return g;
}

int check_grid(int dim, const int* n, double h, int periodic) {
  SSIM_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3, got %d", dim);
  SSIM_REQUIRE(n != nullptr, "null grid dims");
  for (int c = 0; c < dim; ++c)
    SSIM_REQUIRE(n[c] >= (periodic ? 1 : 2), "grid dimension %d too small", c);
  SSIM_REQUIRE(h > 0.0, "cell size must be > 0, got %g", h);
  return SSIM_OK;
}

}  // namespace

extern "C" {

int ssim_cic_deposit(int dim, int periodic, int np, const double* x,
                     const double* y, const double* z, const double* m,
                     const int* grid_n, const double* origin, double h,
                     int mesh_prec, void* mesh, int* cells, void* stream) {
  SSIM_TRY(check_grid(dim, grid_n, h, periodic));
  SSIM_REQUIRE(np >= 0, "np must be >= 0");
  if (np == 0) return SSIM_OK;
  SSIM_REQUIRE(x && y && m && mesh && origin && (dim == 2 || z), "null pointer");
  const Grid g = make_grid(dim, grid_n, origin, h, periodic);
  hipStream_t st = ssim::as_stream(stream);
  const dim3 grid(ssim::ceil_div(np, PM_BLOCK));
#define LAUNCH_DEP(D, M)                                                     \
  SSIM_TIMED("k_cic_deposit", st,                                             \
             hipLaunchKernelGGL((k_cic_deposit<D, M>), grid, dim3(PM_BLOCK), 0, \
                                st, np, x, y, z, m, g, static_cast<M*>(mesh),    \
                                cells))
  if (mesh_prec == SSIM_PREC_F64) {
    if (dim == 2) LAUNCH_DEP(2, double); else LAUNCH_DEP(3, double);
  } else {
    if (dim == 2) LAUNCH_DEP(2, float); else LAUNCH_DEP(3, float);
  }
#undef LAUNCH_DEP
  SSIM_LAUNCHED("k_cic_deposit");
  return SSIM_OK;
}

size_t ssim_cic_deposit_tiles_work_bytes(int dim, int np, const int* grid_n) {
  if ((dim != 2 && dim != 3) || !grid_n) return 0;
  const int E = dim == 3 ? CicTile<3>::E : CicTile<2>::E;
  long long nt = 1;
  for (int c = 0; c < dim; ++c) nt *= (grid_n[c] + E - 1) / E;
  const long long nbins = nt << dim;
  const size_t n = (size_t)(np > 0 ? np : 1);
  return 2 * ssim::align256(n * sizeof(int)) + ssim::align256((size_t)nbins * sizeof(int)) +
         ssim::align256(((size_t)nbins + 1) * sizeof(int)) +
         ssim::scan_work_bytes((int)nbins) + ssim::align256(n * sizeof(double4)) +
         ssim::align256(n * sizeof(int4)) + ssim::align256(sizeof(double)) +
         ssim::align256((size_t)((n + PM_BLOCK - 1) / PM_BLOCK) * sizeof(double));
}

int ssim_cic_deposit_tiles(int dim, int periodic, int np, const double* x,
                           const double* y, const double* z, const double* m,
                           const int* grid_n, const double* origin, double h,
                           int mesh_prec, void* mesh, int* cells, void* work,
                           size_t work_bytes, void* stream) {
  SSIM_TRY(check_grid(dim, grid_n, h, periodic));
  SSIM_REQUIRE(np >= 0, "np must be >= 0");
  SSIM_REQUIRE(mesh && origin && (np == 0 || (x && y && m && (dim == 2 || z))),
               "null pointer");
  SSIM_REQUIRE(work && work_bytes >= ssim_cic_deposit_tiles_work_bytes(dim, np, grid_n),
               "workspace too small");
  const Grid g = make_grid(dim, grid_n, origin, h, periodic);
  const int E = dim == 3 ? CicTile<3>::E : CicTile<2>::E;
  long long ntl = 1;
  for (int c = 0; c < dim; ++c) ntl *= (grid_n[c] + E - 1) / E;
  SSIM_REQUIRE(ntl <= (1LL << 26), "too many tiles: %lld", ntl);
  const int nt = (int)ntl, nbins = nt << dim;
  hipStream_t st = ssim::as_stream(stream);
  const size_t n = (size_t)(np > 0 ? np : 1);
  unsigned char* w = static_cast<unsigned char*>(work);
  int* key = reinterpret_cast<int*>(w);
  w += ssim::align256(n * sizeof(int));
  int* rank = reinterpret_cast<int*>(w);
  w += ssim::align256(n * sizeof(int));
  int* count = reinterpret_cast<int*>(w);
  w += ssim::align256((size_t)nbins * sizeof(int));
  int* start = reinterpret_cast<int*>(w);
  w += ssim::align256(((size_t)nbins + 1) * sizeof(int));
  void* scan_work = w;
  w += ssim::scan_work_bytes(nbins);
  double4* trec = reinterpret_cast<double4*>(w);
  w += ssim::align256(n * sizeof(double4));
  int4* treci = reinterpret_cast<int4*>(w);
  w += ssim::align256(n * sizeof(int4));
  double* mmax = reinterpret_cast<double*>(w);
  w += ssim::align256(sizeof(double));
  double* part_max = reinterpret_cast<double*>(w);
  // counts zeroed by a kernel (graph-capturable), then keys, starts, order
  hipLaunchKernelGGL(k_zero_count, dim3(ssim::ceil_div(nbins + 2, PM_BLOCK)), dim3(PM_BLOCK),
                     0, st, nbins, count, reinterpret_cast<int*>(mmax));  // np = 0: mmax = 0
  SSIM_LAUNCHED("k_zero_count");
  const dim3 gp(ssim::ceil_div(np, PM_BLOCK));
  const int nparts = ssim::ceil_div(np, PM_BLOCK);
  int kbits = 0;  // bits of the largest bin index
  while ((1LL << kbits) < nbins) ++kbits;
  if (np > 0) {
    if (dim == 2)
      hipLaunchKernelGGL(k_cic_tile_key<2>, gp, dim3(PM_BLOCK), 0, st, np, x, y, z, m, g,
                         key, rank, count, cells, part_max, kbits);
    else
      hipLaunchKernelGGL(k_cic_tile_key<3>, gp, dim3(PM_BLOCK), 0, st, np, x, y, z, m, g,
                         key, rank, count, cells, part_max, kbits);
    SSIM_LAUNCHED("k_cic_tile_key");
  }
  SSIM_TRY(ssim::scan_exclusive(nbins, count, start, scan_work, st));
  if (np > 0) {
    if (dim == 2)
      hipLaunchKernelGGL(k_cic_tile_scatter<2>, gp, dim3(PM_BLOCK), 0, st, np, x, y, z, m, g,
                         key, rank, start, trec, treci, nparts, part_max, mmax);
    else
      hipLaunchKernelGGL(k_cic_tile_scatter<3>, gp, dim3(PM_BLOCK), 0, st, np, x, y, z, m, g,
                         key, rank, start, trec, treci, nparts, part_max, mmax);
    SSIM_LAUNCHED("k_cic_tile_scatter");
  }
#define LAUNCH_TILES(D, M)                                                         \
  SSIM_TIMED("k_cic_deposit", st,                                                   \
             hipLaunchKernelGGL((k_cic_deposit_tiles<D, M>), dim3(nt), dim3(CT_THREADS), \
                                0, st, g, start, trec, treci, mmax, static_cast<M*>(mesh)))
  if (mesh_prec == SSIM_PREC_F64) {
    if (dim == 2) LAUNCH_TILES(2, double); else LAUNCH_TILES(3, double);
  } else {
    if (dim == 2) LAUNCH_TILES(2, float); else LAUNCH_TILES(3, float);
  }
#undef LAUNCH_TILES
  SSIM_LAUNCHED("k_cic_deposit_tiles");
  return SSIM_OK;
}

int ssim_cic_gather(int dim, int periodic, int np, const double* x,
                    const double* y, const double* z, const int* grid_n,
                    const double* origin, double h, int mesh_prec, int nfields,
                    const void* f0, const void* f1, const void* f2,
                    double* out0, double* out1, double* out2, void* stream) {
  SSIM_TRY(check_grid(dim, grid_n, h, periodic));
  SSIM_REQUIRE(np >= 0 && nfields >= 1 && nfields <= 3, "bad np / nfields");
  if (np == 0) return SSIM_OK;
  SSIM_REQUIRE(x && y && origin && (dim == 2 || z) && f0 && out0 &&
                   (nfields < 2 || (f1 && out1)) && (nfields < 3 || (f2 && out2)),
               "null pointer");
  const Grid g = make_grid(dim, grid_n, origin, h, periodic);
  hipStream_t st = ssim::as_stream(stream);
  const dim3 grid(ssim::ceil_div(np, PM_BLOCK));
#define LAUNCH_GAT(D, M)                                                      \
  SSIM_TIMED("k_cic_gather", st,                                              \
             hipLaunchKernelGGL((k_cic_gather<D, M>), grid, dim3(PM_BLOCK), 0,  \
                                st, np, x, y, z, g, nfields,                    \
                                static_cast<const M*>(f0),                      \
                                static_cast<const M*>(f1),                      \
                                static_cast<const M*>(f2), out0, out1, out2))
  if (mesh_prec == SSIM_PREC_F64) {
    if (dim == 2) LAUNCH_GAT(2, double); else LAUNCH_GAT(3, double);
  } else {
    if (dim == 2) LAUNCH_GAT(2, float); else LAUNCH_GAT(3, float);
  }
#undef LAUNCH_GAT
  SSIM_LAUNCHED("k_cic_gather");
  return SSIM_OK;
}

size_t ssim_minmax_work_bytes(int n) {
  return ssim::align256((size_t)ssim::ceil_div(n < 1 ? 1 : n, PM_BLOCK * 8) * 6 *
                        sizeof(double));
}

int ssim_minmax(int n, int narrays, const double* a0, const double* a1,
                const double* a2, double* out, void* work, size_t work_bytes,
                void* stream) {
  SSIM_REQUIRE(n >= 1, "minmax of an empty array");
  SSIM_REQUIRE(narrays >= 1 && narrays <= 3, "narrays must be 1..3");
  SSIM_REQUIRE(a0 && out && work && (narrays < 2 || a1) && (narrays < 3 || a2),
               "null pointer");
  SSIM_REQUIRE(work_bytes >= ssim_minmax_work_bytes(n), "workspace too small");
  hipStream_t st = ssim::as_stream(stream);
  int nblk = ssim::ceil_div(n, PM_BLOCK * 8);
  double* part = static_cast<double*>(work);
  hipLaunchKernelGGL(k_minmax_partial, dim3(nblk), dim3(PM_BLOCK), 0, st, n,
                     narrays, a0, a1, a2, part);
  SSIM_LAUNCHED("k_minmax_partial");
  hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(PM_BLOCK), 0, st, nblk, narrays,
                     part, out);
  SSIM_LAUNCHED("k_minmax_final");
  return SSIM_OK;
}

int ssim_box_after_kdk(int dim, int n, const double* x, const double* y,
                       const double* z, const double* vx, const double* vy,
                       const double* vz, const double* ax, const double* ay,
                       const double* az, double half, double full, double* out,
                       void* work, size_t work_bytes, void* stream) {
  SSIM_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
  SSIM_REQUIRE(n >= 1, "box of an empty set");
  SSIM_REQUIRE(x && y && vx && vy && ax && ay && out && work &&
                   (dim == 2 || (z && vz && az)),
               "null pointer");
  SSIM_REQUIRE(work_bytes >= ssim_minmax_work_bytes(n), "workspace too small");
  hipStream_t st = ssim::as_stream(stream);
  int nblk = ssim::ceil_div(n, PM_BLOCK * 8);
  double* part = static_cast<double*>(work);
  hipLaunchKernelGGL(k_box_kdk_partial, dim3(nblk), dim3(PM_BLOCK), 0, st, n, dim, x,
                     y, z, vx, vy, vz, ax, ay, az, half, full, part);
  SSIM_LAUNCHED("k_box_kdk_partial");
  hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(PM_BLOCK), 0, st, nblk, dim, part,
                     out);
  SSIM_LAUNCHED("k_minmax_final");
  return SSIM_OK;
}


size_t ssim_kick_box_work_bytes(int n) {
  const int nblk = ssim::ceil_div(n < 1 ? 1 : n, PM_BLOCK * KB_ITEMS);
  return 256 + ssim::align256((size_t)nblk * 6 * sizeof(double));
}

static int kick_box_impl(bool drift, int dim, int n, double* x, double* y, double* z,
                         double* vx, double* vy, double* vz, const double* ax,
                         const double* ay, const double* az, double half, double full,
                         double* out, void* work, size_t work_bytes, void* stream) {
  SSIM_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
  SSIM_REQUIRE(n >= 1, "box of an empty set");
  SSIM_REQUIRE(x && y && vx && vy && ax && ay && out && work &&
                   (dim == 2 || (z && vz && az)),
               "null pointer");
  SSIM_REQUIRE(work_bytes >= ssim_kick_box_work_bytes(n), "workspace too small");
  hipStream_t st = ssim::as_stream(stream);
  const int nblk = ssim::ceil_div(n, PM_BLOCK * KB_ITEMS);
  // the workgroups' extrema are folded by a second launch: a completion
  // counter in the first would need an agent-scope release per workgroup,
  // i.e. an L2 write-back each
  double* part = reinterpret_cast<double*>(static_cast<unsigned char*>(work) + 256);
  if (drift)
    hipLaunchKernelGGL((k_kick_box<true, true>), dim3(nblk), dim3(PM_BLOCK), 0, st, n, dim,
                       x, y, z, vx, vy, vz, ax, ay, az, half, full, part);
  else
    hipLaunchKernelGGL((k_kick_box<false, true>), dim3(nblk), dim3(PM_BLOCK), 0, st, n,
                       dim, x, y, z, vx, vy, vz, ax, ay, az, half, full, part);
  SSIM_LAUNCHED("k_kick_box");
  hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(PM_BLOCK), 0, st, nblk, dim, part, out);
  SSIM_LAUNCHED("k_minmax_final");
  return SSIM_OK;
}

int ssim_kick_box(int dim, int n, const double* x, const double* y, const double* z,
                  double* vx, double* vy, double* vz, const double* ax,
                  const double* ay, const double* az, double half, double full,
                  double* out, void* work, size_t work_bytes, void* stream) {
  // the positions are only read without the drift
  return kick_box_impl(false, dim, n, const_cast<double*>(x), const_cast<double*>(y),
                       const_cast<double*>(z), vx, vy, vz, ax, ay, az, half, full, out,
                       work, work_bytes, stream);
}

int ssim_kicks_drift(int dim, int n, double* x, double* y, double* z, double* vx,
                     double* vy, double* vz, const double* ax, const double* ay,
                     const double* az, double half, double full, void* stream) {
  SSIM_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
  SSIM_REQUIRE(n >= 0, "n must be >= 0");
  if (n == 0) return SSIM_OK;
  SSIM_REQUIRE(x && y && vx && vy && ax && ay && (dim == 2 || (z && vz && az)),
               "null pointer");
  hipStream_t st = ssim::as_stream(stream);
  hipLaunchKernelGGL((k_kick_box<true, false>), dim3(ssim::ceil_div(n, PM_BLOCK * KB_ITEMS)),
                     dim3(PM_BLOCK), 0, st, n, dim, x, y, z, vx, vy, vz, ax, ay, az, half,
                     full, nullptr);
  SSIM_LAUNCHED("k_kick_box");
  return SSIM_OK;
}

int ssim_kick_drift_box(int dim, int n, double* x, double* y, double* z, double* vx,
                        double* vy, double* vz, const double* ax, const double* ay,
                        const double* az, double half, double full, double* out,
                        void* work, size_t work_bytes, void* stream) {
  return kick_box_impl(true, dim, n, x, y, z, vx, vy, vz, ax, ay, az, half, full, out,
                       work, work_bytes, stream);
}

int ssim_pm_force_kernel(int dim, int n, int comp, double h, double eps2,
                         double G, double rs, double rcut, int mesh_prec,
                         void* out, void* stream) {
  SSIM_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
  SSIM_REQUIRE(n >= 1 && comp >= 0 && comp < dim, "bad n / component");
  SSIM_REQUIRE(h > 0.0 && eps2 >= 0.0 && rs >= 0.0,
               "need h > 0, eps2 >= 0, rs >= 0");
  SSIM_REQUIRE(out, "null pointer");
  const long long P = 2LL * n;
  const long long total = dim == 2 ? P * P : P * P * P;
  hipStream_t st = ssim::as_stream(stream);
  const dim3 grid((unsigned)ssim::ceil_div(total, PM_BLOCK));
  if (mesh_prec == SSIM_PREC_F64)
    hipLaunchKernelGGL(k_force_kernel<double>, grid, dim3(PM_BLOCK), 0, st, dim,
                       n, comp, h, eps2, G, rs, rcut, static_cast<double*>(out));
  else
    hipLaunchKernelGGL(k_force_kernel<float>, grid, dim3(PM_BLOCK), 0, st, dim,
                       n, comp, h, eps2, G, rs, rcut, static_cast<float*>(out));
  SSIM_LAUNCHED("k_force_kernel");
  return SSIM_OK;
}

int ssim_pm_deconvolve(int dim, int padded_n, int prec, int order,
                       void* spectrum, void* stream) {
  SSIM_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
  SSIM_REQUIRE(order == 2 || order == 3, "order must be 2 (CIC) or 3 (TSC)");
  SSIM_REQUIRE(padded_n >= 2 && spectrum, "bad arguments");
  const long long W = padded_n / 2 + 1;
  const long long total = dim == 2 ? padded_n * W : (long long)padded_n * padded_n * W;
  hipStream_t st = ssim::as_stream(stream);
  const dim3 grid((unsigned)ssim::ceil_div(total, PM_BLOCK));
  if (prec == SSIM_PREC_F64)
    hipLaunchKernelGGL(k_deconvolve<double>, grid, dim3(PM_BLOCK), 0, st, dim,
                       padded_n, order, static_cast<double*>(spectrum));
  else
    hipLaunchKernelGGL(k_deconvolve<float>, grid, dim3(PM_BLOCK), 0, st, dim,
                       padded_n, order, static_cast<float*>(spectrum));
  SSIM_LAUNCHED("k_deconvolve");
  return SSIM_OK;
}

int ssim_spectrum_mul(long long ncomplex, int prec, const void* a,
                      const void* b, void* out, void* stream) {
  SSIM_REQUIRE(ncomplex >= 0, "bad size");
  if (ncomplex == 0) return SSIM_OK;
  SSIM_REQUIRE(a && b && out, "null pointer");
  hipStream_t st = ssim::as_stream(stream);
  const dim3 grid((unsigned)ssim::ceil_div(ncomplex, PM_BLOCK));
  if (prec == SSIM_PREC_F64)
    hipLaunchKernelGGL(k_cmul<double>, grid, dim3(PM_BLOCK), 0, st, ncomplex,
                       static_cast<const double*>(a), static_cast<const double*>(b),
                       static_cast<double*>(out));
  else
    hipLaunchKernelGGL(k_cmul<float>, grid, dim3(PM_BLOCK), 0, st, ncomplex,
                       static_cast<const float*>(a), static_cast<const float*>(b),
                       static_cast<float*>(out));
  SSIM_LAUNCHED("k_cmul");
  return SSIM_OK;
}

int ssim_tsc_deposit(int np, const double* x, const double* y, const double* z,
                     const double* m, int grid_n, const double* origin, double h,
                     int mesh_prec, void* mesh, void* stream) {
  SSIM_REQUIRE(np >= 0 && grid_n >= 1 && h > 0.0, "bad TSC arguments");
  if (np == 0) return SSIM_OK;
  SSIM_REQUIRE(x && y && z && m && origin && mesh, "null pointer");
  const int ns[3] = {grid_n, grid_n, grid_n};
  const Grid g = make_grid(3, ns, origin, h, 0);
  hipStream_t st = ssim::as_stream(stream);
  const dim3 grid(ssim::ceil_div(np, PM_BLOCK));
  if (mesh_prec == SSIM_PREC_F64)
    SSIM_TIMED("k_tsc_deposit", st,
               hipLaunchKernelGGL(k_tsc_deposit<double>, grid, dim3(PM_BLOCK), 0,
                                  st, np, x, y, z, m, g, static_cast<double*>(mesh)));
  else
    SSIM_TIMED("k_tsc_deposit", st,
               hipLaunchKernelGGL(k_tsc_deposit<float>, grid, dim3(PM_BLOCK), 0,
                                  st, np, x, y, z, m, g, static_cast<float*>(mesh)));
  SSIM_LAUNCHED("k_tsc_deposit");
  return SSIM_OK;
}

size_t ssim_tsc_deposit_work_bytes(int np) {
  return ssim::align256((size_t)(np > 0 ? np : 1) * sizeof(double4));
}

int ssim_tsc_deposit_binned(int np, const double* x, const double* y,
                            const double* z, const double* m, int grid_n,
                            const double* origin, double h, const int* bin_start,
                            const int* bin_order, int mesh_prec, void* mesh,
                            void* work, size_t work_bytes, void* stream) {
  SSIM_REQUIRE(np >= 0 && grid_n >= 1 && h > 0.0, "bad TSC arguments");
  SSIM_REQUIRE(x && y && z && m && origin && bin_start && bin_order && mesh && work,
               "null pointer");
  SSIM_REQUIRE(work_bytes >= ssim_tsc_deposit_work_bytes(np), "workspace too small");
  const int ns[3] = {grid_n, grid_n, grid_n};
  const Grid g = make_grid(3, ns, origin, h, 0);
  hipStream_t st = ssim::as_stream(stream);
  double4* body = static_cast<double4*>(work);
  if (np > 0) {
    hipLaunchKernelGGL(k_pack_binned, dim3(ssim::ceil_div(np, PM_BLOCK)),
                       dim3(PM_BLOCK), 0, st, np, bin_order, x, y, z, m, body);
    SSIM_LAUNCHED("k_pack_binned");
  }
  const long long nodes = (long long)grid_n * grid_n * grid_n;
  const dim3 grid((unsigned)ssim::ceil_div(nodes, PM_BLOCK));
  if (mesh_prec == SSIM_PREC_F64)
    SSIM_TIMED("k_tsc_deposit", st,
               hipLaunchKernelGGL(k_tsc_deposit_binned<double>, grid, dim3(PM_BLOCK),
                                  0, st, body, g, bin_start,
                                  static_cast<double*>(mesh)));
  else
    SSIM_TIMED("k_tsc_deposit", st,
               hipLaunchKernelGGL(k_tsc_deposit_binned<float>, grid, dim3(PM_BLOCK),
                                  0, st, body, g, bin_start,
                                  static_cast<float*>(mesh)));
  SSIM_LAUNCHED("k_tsc_deposit_binned");
  return SSIM_OK;
}

int ssim_tsc_deposit_blocked(int grid_n, const double* origin, double h, int block,
                             const int* bin_start, const void* body, int mesh_prec,
                             void* mesh, void* stream) {
  SSIM_REQUIRE(grid_n >= 1 && h > 0.0 && block >= 2, "bad TSC arguments (block >= 2)");
  SSIM_REQUIRE(origin && bin_start && body && mesh, "null pointer");
  const int ns[3] = {grid_n, grid_n, grid_n};
  const Grid g = make_grid(3, ns, origin, h, 0);
  const int nb = (grid_n + block - 1) / block;
  hipStream_t st = ssim::as_stream(stream);
  const long long nodes = (long long)grid_n * grid_n * grid_n;
  const dim3 grid((unsigned)ssim::ceil_div(nodes, PM_BLOCK));
  const double4* bp = static_cast<const double4*>(body);
#define LAUNCH_DEPB(M, BC)                                                          \
  SSIM_TIMED("k_tsc_deposit", st,                                                   \
             hipLaunchKernelGGL((k_tsc_deposit_blocked<M, BC>), grid, dim3(PM_BLOCK), 0, \
                                st, bp, g, block, nb, bin_start, static_cast<M*>(mesh)))
#define LAUNCH_DEPB_B(M)                       \
  switch (block) {                             \
    case 2: LAUNCH_DEPB(M, 2); break;          \
    case 3: LAUNCH_DEPB(M, 3); break;          \
    case 4: LAUNCH_DEPB(M, 4); break;          \
    case 5: LAUNCH_DEPB(M, 5); break;          \
    case 6: LAUNCH_DEPB(M, 6); break;          \
    default: LAUNCH_DEPB(M, 0); break;         \
  }
  if (mesh_prec == SSIM_PREC_F64) {
    LAUNCH_DEPB_B(double)
  } else {
    LAUNCH_DEPB_B(float)
  }
#undef LAUNCH_DEPB_B
#undef LAUNCH_DEPB
  SSIM_LAUNCHED("k_tsc_deposit_blocked");
  return SSIM_OK;
}

int ssim_tsc_deposit_tiles(int grid_n, const double* origin, double h, int block,
                           const int* cell_start, const void* body, double mass_bound,
                           int mesh_prec, void* mesh, void* stream) {
  SSIM_REQUIRE(grid_n >= 1 && h > 0.0, "bad TSC arguments");
  SSIM_REQUIRE(block == 5, "ssim_tsc_deposit_tiles: block edge 5 only (got %d)", block);
  SSIM_REQUIRE(mass_bound > 0.0 && std::isfinite(mass_bound), "mass_bound must be > 0");
  SSIM_REQUIRE(origin && cell_start && body && mesh, "null pointer");
  const int nb = (grid_n + block - 1) / block;
  const int ntile = (nb + DT - 1) / DT;
  int e = 0;
  std::frexp(mass_bound, &e);  // mass_bound <= 2^e: no sum exceeds 2^62
  e = e < -900 ? -900 : e;     // keeps 2^(62 - e) finite
  const double scale = std::ldexp(1.0, 62 - e), inv_scale = std::ldexp(1.0, e - 62);
  hipStream_t st = ssim::as_stream(stream);
  const dim3 grid((unsigned)(ntile * ntile * ntile));
  const double4* bp = static_cast<const double4*>(body);
  if (mesh_prec == SSIM_PREC_F64)
    SSIM_TIMED("k_tsc_deposit", st,
               hipLaunchKernelGGL((k_tsc_deposit_tiles<double, 5>), grid, dim3(DEP_T), 0,
                                  st, bp, cell_start, nb, grid_n, scale, inv_scale,
                                  static_cast<double*>(mesh)));
  else
    SSIM_TIMED("k_tsc_deposit", st,
               hipLaunchKernelGGL((k_tsc_deposit_tiles<float, 5>), grid, dim3(DEP_T), 0,
                                  st, bp, cell_start, nb, grid_n, scale, inv_scale,
                                  static_cast<float*>(mesh)));
  SSIM_LAUNCHED("k_tsc_deposit_tiles");
  return SSIM_OK;
}

// float4 meshes under 2 GiB are gathered with buffer loads
static bool buffer_mesh(int grid_n) {
  return (long long)grid_n * grid_n * grid_n * 16 <= 0x7fffffffLL;
}

int ssim_tsc_gather(int np, const int* order, const double* x, const double* y,
                    const double* z, int grid_n, const double* origin, double h,
                    int mesh_prec, const void* f0, const void* f1, const void* f2,
                    double* o0, double* o1, double* o2, int accumulate,
                    void* stream) {
  SSIM_REQUIRE(np >= 0 && grid_n >= 1 && h > 0.0, "bad TSC arguments");
  if (np == 0) return SSIM_OK;
  SSIM_REQUIRE(x && y && z && origin && f0 && o0 && o1 && o2, "null pointer");
  SSIM_REQUIRE((f1 == nullptr) == (f2 == nullptr),
               "f1 and f2 are both given or both NULL (interleaved)");
  if (!f1) {
    SSIM_REQUIRE(mesh_prec == SSIM_PREC_F32C,
                 "interleaved fields are FP32 float4 [n^3]");
    const int ns[3] = {grid_n, grid_n, grid_n};
    const Grid g = make_grid(3, ns, origin, h, 0);
    hipStream_t st = ssim::as_stream(stream);
    if (buffer_mesh(grid_n))
      SSIM_TIMED("k_tsc_gather", st,
                 hipLaunchKernelGGL((k_tsc_gather4<false, true, false>),
                                    dim3(ssim::ceil_div(np, PM_BLOCK)), dim3(PM_BLOCK), 0,
                                    st, np, order, x, y, z, g,
                                    static_cast<const float4*>(f0), o0, o1, o2,
                                    accumulate, Corr{}, BoxOut{}));
    else
      SSIM_TIMED("k_tsc_gather", st,
                 hipLaunchKernelGGL((k_tsc_gather4<false, false, false>),
                                    dim3(ssim::ceil_div(np, PM_BLOCK)), dim3(PM_BLOCK), 0,
                                    st, np, order, x, y, z, g,
                                    static_cast<const float4*>(f0), o0, o1, o2,
                                    accumulate, Corr{}, BoxOut{}));
    SSIM_LAUNCHED("k_tsc_gather4");
    return SSIM_OK;
  }
  const int ns[3] = {grid_n, grid_n, grid_n};
  const Grid g = make_grid(3, ns, origin, h, 0);
  hipStream_t st = ssim::as_stream(stream);
  const dim3 grid(ssim::ceil_div(np, PM_BLOCK));
  if (mesh_prec == SSIM_PREC_F64)
    SSIM_TIMED("k_tsc_gather", st,
               hipLaunchKernelGGL(k_tsc_gather<double>, grid, dim3(PM_BLOCK), 0, st,
                                  np, order, x, y, z, g, static_cast<const double*>(f0),
                                  static_cast<const double*>(f1),
                                  static_cast<const double*>(f2), o0, o1, o2,
                                  accumulate));
  else
    SSIM_TIMED("k_tsc_gather", st,
               hipLaunchKernelGGL(k_tsc_gather<float>, grid, dim3(PM_BLOCK), 0, st,
                                  np, order, x, y, z, g, static_cast<const float*>(f0),
                                  static_cast<const float*>(f1),
                                  static_cast<const float*>(f2), o0, o1, o2,
                                  accumulate));
  SSIM_LAUNCHED("k_tsc_gather");
  return SSIM_OK;
}

int ssim_tsc_gather_corrected(int np, const int* order, const double* x,
                              const double* y, const double* z, int grid_n,
                              const double* origin, double h, const void* f4,
                              double* o0, double* o1, double* o2, int accumulate,
                              const float* corr, float scale, int relative,
                              const double* v0, const double* v1, const double* v2,
                              int n_proj, void* stream) {
  SSIM_REQUIRE(np >= 0 && grid_n >= 1 && h > 0.0, "bad TSC arguments");
  if (np == 0) return SSIM_OK;
  SSIM_REQUIRE(x && y && z && origin && f4 && o0 && o1 && o2 && corr, "null pointer");
  SSIM_REQUIRE(n_proj <= 0 || (v0 && v1 && v2), "n_proj > 0 needs the velocities");
  const int ns[3] = {grid_n, grid_n, grid_n};
  const Grid g = make_grid(3, ns, origin, h, 0);
  hipStream_t st = ssim::as_stream(stream);
  const Corr cr{corr, scale, relative, v0, v1, v2, n_proj};
  if (buffer_mesh(grid_n))
    SSIM_TIMED("k_tsc_gather", st,
               hipLaunchKernelGGL((k_tsc_gather4<true, true, false>),
                                  dim3(ssim::ceil_div(np, PM_BLOCK)), dim3(PM_BLOCK), 0, st,
                                  np, order, x, y, z, g, static_cast<const float4*>(f4), o0,
                                  o1, o2, accumulate, cr, BoxOut{}));
  else
    SSIM_TIMED("k_tsc_gather", st,
               hipLaunchKernelGGL((k_tsc_gather4<true, false, false>),
                                  dim3(ssim::ceil_div(np, PM_BLOCK)), dim3(PM_BLOCK), 0, st,
                                  np, order, x, y, z, g, static_cast<const float4*>(f4), o0,
                                  o1, o2, accumulate, cr, BoxOut{}));
  SSIM_LAUNCHED("k_tsc_gather4");
  return SSIM_OK;
}

size_t ssim_tsc_gather_box_work_bytes(int np) {
  return ssim::align256((size_t)ssim::ceil_div(np > 0 ? np : 1, PM_BLOCK) * 6 *
                        sizeof(double));
}

int ssim_tsc_gather_box(int np, const int* order, const double* x, const double* y,
                        const double* z, int grid_n, const double* origin, double h,
                        const void* f4, double* o0, double* o1, double* o2, int accumulate,
                        const float* corr, float scale, int relative, const double* v0,
                        const double* v1, const double* v2, int n_proj, double half,
                        double full, double* out, void* work, size_t work_bytes,
                        void* stream) {
  SSIM_REQUIRE(np >= 1 && grid_n >= 1 && h > 0.0, "bad TSC arguments (np >= 1 for a box)");
  SSIM_REQUIRE(x && y && z && origin && f4 && o0 && o1 && o2 && v0 && v1 && v2 && out &&
                   work,
               "null pointer");
  SSIM_REQUIRE(work_bytes >= ssim_tsc_gather_box_work_bytes(np), "workspace too small");
  const int ns[3] = {grid_n, grid_n, grid_n};
  const Grid g = make_grid(3, ns, origin, h, 0);
  hipStream_t st = ssim::as_stream(stream);
  const int nblk = (int)ssim::ceil_div(np, PM_BLOCK);
  double* part = static_cast<double*>(work);
  const Corr cr{corr, scale, relative, v0, v1, v2, corr ? n_proj : 0};
  const BoxOut bo{v0, v1, v2, half, full, part};
  const float4* fp = static_cast<const float4*>(f4);
#define SSIM_GB_LAUNCH(C, B)                                                             \
  SSIM_TIMED("k_tsc_gather", st,                                                        \
             hipLaunchKernelGGL((k_tsc_gather4<C, B, true>), dim3(nblk), dim3(PM_BLOCK), 0, \
                                st, np, order, x, y, z, g, fp, o0, o1, o2, accumulate, cr, \
                                bo))
  const bool buf = buffer_mesh(grid_n);
  if (corr && buf) SSIM_GB_LAUNCH(true, true);
  else if (corr) SSIM_GB_LAUNCH(true, false);
  else if (buf) SSIM_GB_LAUNCH(false, true);
  else SSIM_GB_LAUNCH(false, false);
#undef SSIM_GB_LAUNCH
  SSIM_LAUNCHED("k_tsc_gather4");
  hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(PM_BLOCK), 0, st, nblk, 3, part, out);
  SSIM_LAUNCHED("k_minmax_final");
  return SSIM_OK;
}

int ssim_spectrum_imag(long long ncomplex, int prec, const void* a, void* out,
                       void* stream) {
  SSIM_REQUIRE(ncomplex >= 0, "bad size");
  if (ncomplex == 0) return SSIM_OK;
  SSIM_REQUIRE(a && out, "null pointer");
  hipStream_t st = ssim::as_stream(stream);
  const dim3 grid((unsigned)ssim::ceil_div(ncomplex, PM_BLOCK));
  if (prec == SSIM_PREC_F64)
    hipLaunchKernelGGL(k_imag<double>, grid, dim3(PM_BLOCK), 0, st, ncomplex,
                       static_cast<const double*>(a), static_cast<double*>(out));
  else
    hipLaunchKernelGGL(k_imag<float>, grid, dim3(PM_BLOCK), 0, st, ncomplex,
                       static_cast<const float*>(a), static_cast<float*>(out));
  SSIM_LAUNCHED("k_imag");
  return SSIM_OK;
}

int ssim_pm_poisson(int dim, const int* grid_n, double h, double G, int prec,
                    void* spectrum, void* stream) {
  SSIM_TRY(check_grid(dim, grid_n, h, 1));
  SSIM_REQUIRE(spectrum, "null pointer");
  const int nl = grid_n[dim - 1];
  const long long W = nl / 2 + 1;
  const long long total = dim == 2 ? grid_n[0] * W : (long long)grid_n[0] * grid_n[1] * W;
  hipStream_t st = ssim::as_stream(stream);
  const dim3 grid((unsigned)ssim::ceil_div(total, PM_BLOCK));
  const int n2 = dim == 3 ? grid_n[2] : 1;
  if (prec == SSIM_PREC_F64)
    hipLaunchKernelGGL(k_poisson<double>, grid, dim3(PM_BLOCK), 0, st, dim,
                       grid_n[0], grid_n[1], n2, h, G, static_cast<double*>(spectrum));
  else
    hipLaunchKernelGGL(k_poisson<float>, grid, dim3(PM_BLOCK), 0, st, dim,
                       grid_n[0], grid_n[1], n2, h, G, static_cast<float*>(spectrum));
  SSIM_LAUNCHED("k_poisson");
  return SSIM_OK;
}

int ssim_pm_gradient(int dim, const int* grid_n, double h, int prec,
                     const void* phi, void* g0, void* g1, void* g2, void* stream) {
  SSIM_TRY(check_grid(dim, grid_n, h, 1));
  SSIM_REQUIRE(phi && g0 && g1 && (dim == 2 || g2), "null pointer");
  const int n2 = dim == 3 ? grid_n[2] : 1;
  const long long total = (long long)grid_n[0] * grid_n[1] * n2;
  hipStream_t st = ssim::as_stream(stream);
  const dim3 grid((unsigned)ssim::ceil_div(total, PM_BLOCK));
  if (prec == SSIM_PREC_F64)
    hipLaunchKernelGGL(k_gradient<double>, grid, dim3(PM_BLOCK), 0, st, dim,
                       grid_n[0], grid_n[1], n2, h, static_cast<const double*>(phi),
                       static_cast<double*>(g0), static_cast<double*>(g1),
                       static_cast<double*>(g2));
  else
    hipLaunchKernelGGL(k_gradient<float>, grid, dim3(PM_BLOCK), 0, st, dim,
                       grid_n[0], grid_n[1], n2, h, static_cast<const float*>(phi),
                       static_cast<float*>(g0), static_cast<float*>(g1),
                       static_cast<float*>(g2));
  SSIM_LAUNCHED("k_gradient");
  return SSIM_OK;
}

int ssim_pm_pad(int dim, int n, int prec, const void* src, void* dst, int crop,
                void* stream) {
  SSIM_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
  SSIM_REQUIRE(n >= 1 && src && dst, "bad arguments");
  const long long total = dim == 2 ? (long long)n * n : (long long)n * n * n;
  hipStream_t st = ssim::as_stream(stream);
  const dim3 grid((unsigned)ssim::ceil_div(total, PM_BLOCK));
  if (prec == SSIM_PREC_F64)
    hipLaunchKernelGGL(k_pad<double>, grid, dim3(PM_BLOCK), 0, st, dim, n,
                       static_cast<const double*>(src), static_cast<double*>(dst), crop);
  else
    hipLaunchKernelGGL(k_pad<float>, grid, dim3(PM_BLOCK), 0, st, dim, n,
                       static_cast<const float*>(src), static_cast<float*>(dst), crop);
  SSIM_LAUNCHED("k_pad");
  return SSIM_OK;
}

}  // extern "C"
