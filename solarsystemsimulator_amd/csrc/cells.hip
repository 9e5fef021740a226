// Uniform cell list (counting sort) for near-field work on gfx950.
//
// Bodies are binned into an nc[0] x nc[1] x nc[2] grid of cubic cells of
// edge `cell` starting at `origin` (coordinates outside are clamped into the
// border cells, so every body is binned).  Output:
//   cell_start[ncells + 1]  exclusive prefix of the per-cell counts,
//   order[n]                body indices sorted by cell, ascending body
//                           index inside a cell (deterministic),
//   packed[n]               float4 (x - ox, y - oy, z - oz, m) in that order.
// Steps: key + histogram (atomics on the counts), a stable radix sort of
// (key, index) (sort.hip) that leaves ascending body indices inside a cell,
// an exclusive scan of the counts, and the packed copy.
#include "ssim_internal.h"

namespace {

constexpr int CB = 256;

__global__ __launch_bounds__(CB) void k_zero_i32(int n, int* __restrict__ a) {
  const int i = blockIdx.x * CB + threadIdx.x;
  if (i < n) a[i] = 0;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

__global__ __launch_bounds__(CB) void k_cell_key(
    int n, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, double ox, double oy, double oz,
    double inv_cell, int nc0, int nc1, int nc2, int* __restrict__ key,
    int* __restrict__ order, int* __restrict__ count) {
  const int i = blockIdx.x * CB + threadIdx.x;
  if (i >= n) return;
  const int a = clampi((int)floor((x[i] - ox) * inv_cell), 0, nc0 - 1);
  const int b = clampi((int)floor((y[i] - oy) * inv_cell), 0, nc1 - 1);
  const int c = clampi((int)floor((z[i] - oz) * inv_cell), 0, nc2 - 1);
  const int k = (a * nc1 + b) * nc2 + c;
  key[i] = k;
  order[i] = i;
  atomicAdd(&count[k], 1);
}

// Key of the target binning: cell * 8 + octant of the body inside its cell
// (bit 2: upper half in x, bit 1: in y, bit 0: in z); a stable sort of
// these keys leaves each cell's targets grouped by octant, ascending body
// index inside an octant.
__device__ __forceinline__ int octant_bit(double f, int a) {
  return clampi((int)floor(2.0 * f) - 2 * a, 0, 1);
}

__global__ __launch_bounds__(CB) void k_cell_key8(
    int n, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, double ox, double oy, double oz,
    double inv_cell, int nc0, int nc1, int nc2, int* __restrict__ key,
    int* __restrict__ order) {
  const int i = blockIdx.x * CB + threadIdx.x;
  if (i >= n) return;
  const double fa = (x[i] - ox) * inv_cell, fb = (y[i] - oy) * inv_cell,
               fc = (z[i] - oz) * inv_cell;
  const int a = clampi((int)floor(fa), 0, nc0 - 1);
  const int b = clampi((int)floor(fb), 0, nc1 - 1);
  const int c = clampi((int)floor(fc), 0, nc2 - 1);
  key[i] = ((a * nc1 + b) * nc2 + c) * 8 + octant_bit(fa, a) * 4 +
           octant_bit(fb, b) * 2 + octant_bit(fc, c);
  order[i] = i;
}

// cell_start[c] = sub_start[8 c], c in [0, ncells]
__global__ __launch_bounds__(CB) void k_every8(int ncells, const int* __restrict__ sub,
                                               int* __restrict__ cell_start) {
  const int c = blockIdx.x * CB + threadIdx.x;
  if (c <= ncells) cell_start[c] = sub[8 * c];
}

// Split-force mass bins: the TSC stencil centre r = floor((x - o)/h + 1/2)
// of each axis (the operations of the deposit's weights, so a body's bin is
// always its stencil's centre), clamped to the mesh, as a blocked key.
__global__ __launch_bounds__(CB) void k_split_key(
    int n, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, double ox, double oy, double oz, double h, int gn,
    int B, int nb, int nodes, int* __restrict__ key, int* __restrict__ order,
    int* __restrict__ count) {
  const int i = blockIdx.x * CB + threadIdx.x;
  if (i >= n) return;
  const double top = gn - 1;
  const int a = (int)fmin(fmax(floor((x[i] - ox) / h + 0.5), 0.0), top);
  const int b = (int)fmin(fmax(floor((y[i] - oy) / h + 0.5), 0.0), top);
  const int c = (int)fmin(fmax(floor((z[i] - oz) / h + 0.5), 0.0), top);
  // nodes: the node bin in its block; else the block alone
  const int k = nodes ? ssim_blocked_key(a, b, c, B, nb) : ((a / B) * nb + b / B) * nb + c / B;
  key[i] = k;
  order[i] = i;
  if (count) atomicAdd(&count[k], 1);  // node bins only (blocks: k_run_starts)
}

// The block keys of k_split_key (nodes = 0) fused with the radix sort's
// first-pass counts: workgroup b takes the keys of tile b as
// radix_sort_pairs's counting kernel does (RADIX_T threads, keys base +
// r RADIX_T + t), writes them and its tile's counts of the 8 low key bits
// (digit-major, row stride hrow); workgroups past the last tile write the
// rows' zero padding.  The values are the identity (radix_sort_pairs with
// vals_identity), so no order array is written here.  zero[0..nzero) is
// cleared on the way (the mass counts k_run_starts adds to).
__global__ __launch_bounds__(ssim::RADIX_T) void k_split_key_count(
    int n, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, double ox, double oy, double oz, double h, int gn,
    int B, int nb, int hrow, int* __restrict__ key, int* __restrict__ hist,
    int nzero, int* __restrict__ zero) {
  constexpr int T = ssim::RADIX_T, R = ssim::RADIX_TILE / ssim::RADIX_T;
  __shared__ int cnt[256];
  const int t = threadIdx.x;
  // zero[0..nzero): counts that a later launch accumulates into
  for (int q = blockIdx.x * T + t; q < nzero; q += gridDim.x * T) zero[q] = 0;
  cnt[t] = 0;
  __syncthreads();
  const int base = blockIdx.x * ssim::RADIX_TILE;
  const double top = gn - 1;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int i = base + r * T + t;
    if (i < n) {
      const int a = (int)fmin(fmax(floor((x[i] - ox) / h + 0.5), 0.0), top);
      const int b = (int)fmin(fmax(floor((y[i] - oy) / h + 0.5), 0.0), top);
      const int c = (int)fmin(fmax(floor((z[i] - oz) / h + 0.5), 0.0), top);
      const int k = ((a / B) * nb + b / B) * nb + c / B;
      key[i] = k;
      atomicAdd(&cnt[k & 255], 1);
    }
  }
  __syncthreads();
  hist[(size_t)t * hrow + blockIdx.x] = cnt[t];
}

// start[c] = the number of sorted keys below c, c in [0, nbins]: bin c's first
// slot, what the exclusive scan of the bin counts gives.  One launch, two
// independent parts.  The first nrun workgroups: wave w looks at the slots p
// in [64w, 64w + 64) below n; the bins c in (K(64w - 1), K(its last slot)]
// start among them, at 64w + #{p : K(p) < c} (K(-1) taken as the first
// key), and these ranges tile (first key, last key].  A bin range the wave
// fills 64 bins per round, each by a 6-step search over the lanes' sorted
// keys.  The other workgroups, one thread per bin: the bins up to the first
// key start at 0 and those past the last key at n (a mesh margin leaves
// thousands of empty bins at both ends).
__global__ __launch_bounds__(CB) void k_run_starts(int n, const int* __restrict__ key,
                                                   int nbins, int nrun,
                                                   int* __restrict__ start,
                                                   const int* __restrict__ vals, int nv,
                                                   int* __restrict__ vcount) {
  if ((int)blockIdx.x >= nrun) {
    const int c = (blockIdx.x - nrun) * CB + threadIdx.x;
    if (c > nbins) return;
    if (n == 0 || c <= key[0]) start[c] = 0;
    else if (c > key[n - 1]) start[c] = n;
    return;
  }
  const int lane = threadIdx.x & 63;
  const int p0 = (int)(((size_t)blockIdx.x * CB + threadIdx.x) >> 6) * 64;
  if (p0 >= n) return;  // whole waves
  const int p = p0 + lane;
  const int kp = p < n ? key[p] : nbins;
  const int lo = p0 > 0 ? key[p0 - 1] : key[0];
  const int hi = __shfl(kp, min(63, n - 1 - p0), 64);
  for (int c0 = lo + 1; c0 <= hi; c0 += 64) {
    const int c = c0 + lane;
    int r = 0;  // lanes with a key below c (<= 63: the last slot's key is hi)
#pragma unroll
    for (int st = 32; st >= 1; st >>= 1)
      if (__shfl(kp, r + st - 1, 64) < c) r += st;
    if (c <= hi) start[c] = p0 + r;
  }
  if (vcount) {
    // vcount[k] += the slots of key k whose value is below nv: the first
    // lane of each run of equal keys in the wave adds the run's count
    const bool below = p < n && vals[p] < nv;
    const int prev = __shfl_up(kp, 1, 64);
    const bool head = p < n && (lane == 0 || prev != kp);
    const unsigned long long heads = __ballot(head), hits = __ballot(below);
    if (head) {
      const unsigned long long later = heads & ~((2ull << lane) - 1ull);
      const unsigned long long upto = later ? (1ull << __builtin_ctzll(later)) - 1ull : ~0ull;
      const int cnt = __popcll(hits & upto & ~((1ull << lane) - 1ull));
      if (cnt) atomicAdd(&vcount[kp], cnt);
    }
  }
}

// vals / nv / vcount (optional, vcount zeroed beforehand): vcount[k] = the
// number of slots of key k whose value is below nv.
int run_starts(int n, const int* key, int nbins, int* start, hipStream_t st,
               const int* vals = nullptr, int nv = 0, int* vcount = nullptr) {
  const int nrun = n > 0 ? (int)ssim::ceil_div(ssim::ceil_div(n, 64) * 64, CB) : 0;
  const int nfill = (int)ssim::ceil_div(nbins + 1, CB);
  hipLaunchKernelGGL(k_run_starts, dim3(nrun + nfill), dim3(CB), 0, st, n, key, nbins, nrun,
                     start, vals, nv, vcount);
  SSIM_LAUNCHED("k_run_starts");
  return SSIM_OK;
}

// Both packed copies in bin order: float4 (x - cx, y - cy, z - cz, m) for
// the near field (relative to the cell grid's origin) and double4
// ((x - ox)/h, (y - oy)/h, (z - oz)/h, m) for the deposit (the node-unit
// coordinates its weights start from, computed with the same operations).
// Pack the body p into slot s: float4 (x - cx, y - cy, z - cz, m) for the
// near field on origin-relative coordinates, double4 ((x - ox)/h, (y - oy)/h,
// (z - oz)/h, m) for the deposit (the node-unit coordinates its weights
// start from, computed with the same operations) and, with `near`, float4
// (x, y, z relative to the corner of block blk, m) with their FP32
// remainders in near_lo.
__device__ __forceinline__ void split_pack_one(
    int s, int p, int blk, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ m, double ox, double oy,
    double oz, double h, double cx, double cy, double cz, float4* __restrict__ packed,
    double4* __restrict__ body, int nb, double cell, float4* __restrict__ near,
    float4* __restrict__ near_lo) {
  const double px = x[p], py = y[p], pz = z[p], pm = m[p];
  packed[s] = make_float4((float)(px - cx), (float)(py - cy), (float)(pz - cz),
                          (float)pm);
  body[s] = make_double4((px - ox) / h, (py - oy) / h, (pz - oz) / h, pm);
  if (near) {
    // relative to the corner of the block the body is binned in, so the
    // near field's differences lose no bits to the magnitude of the
    // coordinates
    const int A = blk / (nb * nb), B = (blk / nb) % nb, C = blk % nb;
    const double rx = px - (cx + A * cell), ry = py - (cy + B * cell),
                 rz = pz - (cz + C * cell);
    const float hx = (float)rx, hy = (float)ry, hz = (float)rz;
    near[s] = make_float4(hx, hy, hz, (float)pm);
    // the remainders: hi + lo is the block-relative position to ~2^-48
    if (near_lo)
      near_lo[s] = make_float4((float)(rx - (double)hx), (float)(ry - (double)hy),
                               (float)(rz - (double)hz), 0.0f);
  }
}

// Both packed copies in bin order (slot s holds order[s], binned under
// key[s], whose block is key[s] / key_div).
__global__ __launch_bounds__(CB) void k_split_pack(
    int n, const int* __restrict__ order, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z,
    const double* __restrict__ m, double ox, double oy, double oz, double h,
    double cx, double cy, double cz, float4* __restrict__ packed,
    double4* __restrict__ body, const int* __restrict__ key, int key_div, int nb,
    double cell, float4* __restrict__ near, float4* __restrict__ near_lo) {
  const int s = blockIdx.x * CB + threadIdx.x;
  if (s >= n) return;
  split_pack_one(s, order[s], key[s] / key_div, x, y, z, m, ox, oy, oz, h, cx, cy, cz,
                 packed, body, nb, cell, near, near_lo);
}

// All bodies binned by block (tstart, torder: index order inside a block,
// masses = indices below n_mass first): mcount[c] = the masses of block c,
// by a binary search for n_mass in its sorted segment.
__global__ __launch_bounds__(CB) void k_mass_counts(int ncells, const int* __restrict__ tstart,
                                                    const int* __restrict__ torder,
                                                    int n_mass, int* __restrict__ mcount) {
  const int c = blockIdx.x * CB + threadIdx.x;
  if (c >= ncells) return;
  const int s0 = tstart[c];
  int lo = s0, hi = tstart[c + 1];  // first slot with an index >= n_mass
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (torder[mid] < n_mass) lo = mid + 1;
    else hi = mid;
  }
  mcount[c] = lo - s0;
}

// Mass counts and their exclusive scan by two fully parallel launches
// (ncells <= CB * SCAN_BLOCKS): k_mass_sum writes mcount[c] (the binary
// search of k_mass_counts, one cell per thread: one search's latency) and
// each workgroup's total of CB cells; k_cells_scan adds the totals of the
// workgroups before its own to the workgroup's scan.  Same result as
// k_mass_counts + ssim::scan_exclusive.
constexpr int SCAN_BLOCKS = 1024;

__global__ __launch_bounds__(CB) void k_mass_sum(int ncells, const int* __restrict__ tstart,
                                                 const int* __restrict__ torder, int n_mass,
                                                 int* __restrict__ mcount,
                                                 int* __restrict__ bsum) {
  __shared__ int red[4];
  const int c = blockIdx.x * CB + threadIdx.x;
  int v = 0;
  if (c < ncells) {
    const int s0 = tstart[c];
    int lo = s0, hi = tstart[c + 1];  // first slot with an index >= n_mass
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (torder[mid] < n_mass) lo = mid + 1;
      else hi = mid;
    }
    v = lo - s0;
    mcount[c] = v;
  }
  int tot;
  ssim_block256_excl(v, red, &tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(CB) void k_cells_scan(int ncells, const int* __restrict__ cnt,
                                                   const int* __restrict__ bsum,
                                                   int* __restrict__ start) {
  __shared__ int red[4];
  const int off = ssim_block256_offset(bsum, blockIdx.x, red);
  const int c = blockIdx.x * CB + threadIdx.x;
  const int v = c < ncells ? cnt[c] : 0;
  int tot;
  const int run = off + ssim_block256_excl(v, red, &tot);
  if (c < ncells) start[c] = run;
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) start[ncells] = off + tot;
}

// The masses' starts and their packed copies in one launch (ncells <=
// DIRECT_SCAN_CELLS): workgroup b takes the PK_CELLS blocks from c0 =
// PK_CELLS b, sums the mass counts before them itself (16-byte loads, eight
// in flight per thread), scans its own counts in one wave and writes their
// starts (start[ncells] = the total), then packs its blocks' masses: the
// w-th of them is the k-th mass of block c, at sorted position tstart[c] + k
// (the masses lead every block's segment), slot start[c] + k — the slots and
// records of k_split_pack_all, with a thread for each mass.
constexpr int DIRECT_SCAN_CELLS = 32768;
constexpr int PK_CELLS = 64;
__global__ __launch_bounds__(CB) void k_cells_scan_pack(
    int ncells, int n_mass, const int* __restrict__ cnt, const int* __restrict__ tstart,
    const int* __restrict__ torder, int* __restrict__ start, int* __restrict__ order,
    const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ m, double ox, double oy,
    double oz, double h, double cx, double cy, double cz, float4* __restrict__ packed,
    double4* __restrict__ body, int nb, double cell, float4* __restrict__ near,
    float4* __restrict__ near_lo) {
  static_assert(PK_CELLS == 64, "one wave scans the workgroup's counts");
  __shared__ int red[4];
  __shared__ int lpre[PK_CELLS + 1], lts[PK_CELLS];
  const int t = threadIdx.x;
  const int c0 = blockIdx.x * PK_CELLS;
  const int4* c4 = reinterpret_cast<const int4*>(cnt);
  const int n4 = c0 / 4;
  int pre = 0;
#pragma unroll 8
  for (int i = t; i < n4; i += CB) {
    const int4 q = c4[i];
    pre += (q.x + q.y) + (q.z + q.w);
  }
  int off;
  ssim_block256_excl(pre, red, &off);
  if (t < PK_CELLS) {
    const int c = c0 + t;
    const int v = c < ncells ? cnt[c] : 0;
    int run = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int u = __shfl_up(run, d, 64);
      if (t >= d) run += u;
    }
    lpre[t] = run - v;
    if (t == PK_CELLS - 1) lpre[PK_CELLS] = run;
    lts[t] = c < ncells ? tstart[c] : 0;
    if (c < ncells) start[c] = off + run - v;
    if (c == ncells - 1) start[ncells] = off + run;
  }
  __syncthreads();
  const int total = lpre[PK_CELLS];
  for (int w = t; w < total; w += CB) {
    // the block holding mass w: the last j with lpre[j] <= w (empty blocks
    // repeat the prefix, so the last one is the block that has masses)
    int j = 0;
#pragma unroll
    for (int st = 32; st >= 1; st >>= 1)
      if (lpre[j + st] <= w) j += st;
    const int i = torder[lts[j] + (w - lpre[j])];
    const int s = off + w;
    order[s] = i;
    split_pack_one(s, i, c0 + j, x, y, z, m, ox, oy, oz, h, cx, cy, cz, packed, body, nb,
                   cell, near, near_lo);
  }
}

// The masses' slots from the all-body order: the mass at sorted position p
// of block c goes to mstart[c] + (p - tstart[c]) (the masses lead every
// block's segment), with order and both packed copies.
__global__ __launch_bounds__(CB) void k_split_pack_all(
    int n_all, int n_mass, const int* __restrict__ torder, const int* __restrict__ key,
    const int* __restrict__ tstart, const int* __restrict__ mstart,
    int* __restrict__ order, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ m, double ox, double oy,
    double oz, double h, double cx, double cy, double cz, float4* __restrict__ packed,
    double4* __restrict__ body, int nb, double cell, float4* __restrict__ near,
    float4* __restrict__ near_lo) {
  const int p = blockIdx.x * CB + threadIdx.x;
  if (p >= n_all) return;
  const int i = torder[p];
  if (i >= n_mass) return;
  const int c = key[p];
  const int s = mstart[c] + (p - tstart[c]);
  order[s] = i;
  split_pack_one(s, i, c, x, y, z, m, ox, oy, oz, h, cx, cy, cz, packed, body, nb, cell,
                 near, near_lo);
}

__global__ __launch_bounds__(CB) void k_pack(
    int n, const int* __restrict__ order, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z,
    const double* __restrict__ m, double ox, double oy, double oz,
    float4* __restrict__ packed) {
  const int s = blockIdx.x * CB + threadIdx.x;
  if (s >= n) return;
  const int p = order[s];
  packed[s] = make_float4((float)(x[p] - ox), (float)(y[p] - oy),
                          (float)(z[p] - oz), m ? (float)m[p] : 0.0f);
}

}  // namespace

extern "C" {

size_t ssim_cells_work_bytes(int n, int ncells) {
  return ssim::align256((size_t)(n > 0 ? n : 1) * sizeof(int)) +
         ssim::align256((size_t)(ncells > 0 ? ncells : 1) * sizeof(int)) +
         ssim::scan_work_bytes(ncells) + ssim::radix_sort_work_bytes(n);
}

int ssim_cells_build(int n, const double* x, const double* y, const double* z,
                     const double* m, const double* origin, double cell,
                     const int* nc, int* cell_start, int* order, void* packed,
                     void* work, size_t work_bytes, void* stream) {
  SSIM_REQUIRE(n >= 0, "n must be >= 0");
  SSIM_REQUIRE(origin && nc && cell > 0.0, "bad cell grid");
  SSIM_REQUIRE(nc[0] >= 1 && nc[1] >= 1 && nc[2] >= 1, "bad cell counts");
  const long long ncl = (long long)nc[0] * nc[1] * nc[2];
  SSIM_REQUIRE(ncl <= (1 << 26), "too many cells: %lld", ncl);
  const int ncells = (int)ncl;
  SSIM_REQUIRE(cell_start && order && (n == 0 || (x && y && z)), "null pointer");
  SSIM_REQUIRE(work && work_bytes >= ssim_cells_work_bytes(n, ncells),
               "workspace too small");
  hipStream_t st = ssim::as_stream(stream);
  unsigned char* w = static_cast<unsigned char*>(work);
  int* key = reinterpret_cast<int*>(w);
  w += ssim::align256((size_t)(n > 0 ? n : 1) * sizeof(int));
  int* count = reinterpret_cast<int*>(w);
  w += ssim::align256((size_t)ncells * sizeof(int));
  void* scan_work = w;
  w += ssim::scan_work_bytes(ncells);
  // zeroed by a kernel (not hipMemsetAsync) so the sequence can be captured
  // in a HIP graph
  hipLaunchKernelGGL(k_zero_i32, dim3(ssim::ceil_div(ncells, CB)), dim3(CB), 0,
                     st, ncells, count);
  SSIM_LAUNCHED("k_zero_i32");
  if (n > 0) {
    hipLaunchKernelGGL(k_cell_key, dim3(ssim::ceil_div(n, CB)), dim3(CB), 0, st,
                       n, x, y, z, origin[0], origin[1], origin[2], 1.0 / cell,
                       nc[0], nc[1], nc[2], key, order, count);
    SSIM_LAUNCHED("k_cell_key");
    int bits = 1;
    while (bits < 31 && (1LL << bits) < ncl) ++bits;
    SSIM_TRY(ssim::radix_sort_pairs(n, bits, key, order, w,
                                    ssim::radix_sort_work_bytes(n), st));
  }
  SSIM_TRY(ssim::scan_exclusive(ncells, count, cell_start, scan_work, st));
  if (n > 0 && packed) {
    hipLaunchKernelGGL(k_pack, dim3(ssim::ceil_div(n, CB)), dim3(CB), 0, st, n,
                       order, x, y, z, m, origin[0], origin[1], origin[2],
                       static_cast<float4*>(packed));
    SSIM_LAUNCHED("k_pack");
  }
  return SSIM_OK;
}

size_t ssim_cells_bin_work_bytes(int n, int ncells) {
  const size_t nsub = 8 * (size_t)(ncells > 0 ? ncells : 1);
  return ssim::align256((size_t)(n > 0 ? n : 1) * sizeof(int)) +
         ssim::align256((nsub + 1) * sizeof(int)) + ssim::radix_sort_work_bytes(n);
}

int ssim_cells_bin(int n, const double* x, const double* y, const double* z,
                   const double* origin, double cell, const int* nc,
                   int* cell_start, int* order, void* work, size_t work_bytes,
                   void* stream) {
  SSIM_REQUIRE(n >= 0, "n must be >= 0");
  SSIM_REQUIRE(origin && nc && cell > 0.0, "bad cell grid");
  SSIM_REQUIRE(nc[0] >= 1 && nc[1] >= 1 && nc[2] >= 1, "bad cell counts");
  const long long ncl = (long long)nc[0] * nc[1] * nc[2];
  SSIM_REQUIRE(ncl <= (1 << 26), "too many cells: %lld", ncl);
  const int ncells = (int)ncl;
  SSIM_REQUIRE(cell_start && order && (n == 0 || (x && y && z)), "null pointer");
  SSIM_REQUIRE(work && work_bytes >= ssim_cells_bin_work_bytes(n, ncells),
               "workspace too small");
  hipStream_t st = ssim::as_stream(stream);
  const int nsub = 8 * ncells;
  unsigned char* w = static_cast<unsigned char*>(work);
  int* key = reinterpret_cast<int*>(w);
  w += ssim::align256((size_t)(n > 0 ? n : 1) * sizeof(int));
  int* sub_start = reinterpret_cast<int*>(w);
  w += ssim::align256(((size_t)nsub + 1) * sizeof(int));
  if (n > 0) {
    hipLaunchKernelGGL(k_cell_key8, dim3(ssim::ceil_div(n, CB)), dim3(CB), 0, st,
                       n, x, y, z, origin[0], origin[1], origin[2], 1.0 / cell,
                       nc[0], nc[1], nc[2], key, order);
    SSIM_LAUNCHED("k_cell_key8");
    int bits = 1;
    while (bits < 31 && (1LL << bits) < nsub) ++bits;
    SSIM_TRY(ssim::radix_sort_pairs(n, bits, key, order, w,
                                    ssim::radix_sort_work_bytes(n), st));
  }
  // octant-bin starts from the sorted keys (n = 0: all zero), every 8th one
  // is a cell start
  SSIM_TRY(run_starts(n, key, nsub, sub_start, st));
  hipLaunchKernelGGL(k_every8, dim3(ssim::ceil_div(ncells + 1, CB)), dim3(CB), 0, st,
                     ncells, sub_start, cell_start);
  SSIM_LAUNCHED("k_every8");
  return SSIM_OK;
}

static long long split_bins(int grid_n, int block) {
  const long long nb = (grid_n + block - 1) / block;
  return nb * nb * nb * block * block * block;
}

size_t ssim_split_bin_work_bytes(int n, int grid_n, int block) {
  if (grid_n < 1 || block < 1) return 0;
  const long long K = split_bins(grid_n, block);
  if (K > (1LL << 30)) return 0;
  return ssim::align256((size_t)(n > 0 ? n : 1) * sizeof(int)) +
         ssim::align256((size_t)K * sizeof(int)) + ssim::scan_work_bytes((int)K) +
         ssim::radix_sort_work_bytes(n);
}

int ssim_split_bin(int n, const double* x, const double* y, const double* z,
                   const double* m, int grid_n, const double* origin, double h,
                   int block, int* bin_start, int* cell_start, int* order, void* packed,
                   void* body, void* near, void* near_lo, void* work, size_t work_bytes,
                   void* stream) {
  SSIM_REQUIRE(n >= 0, "n must be >= 0");
  SSIM_REQUIRE(grid_n >= 1 && block >= 1 && h > 0.0 && origin, "bad mesh");
  const long long K = split_bins(grid_n, block);
  SSIM_REQUIRE(K <= (1LL << 30), "too many bins: %lld", K);
  SSIM_REQUIRE(cell_start && order && packed && body && (n == 0 || (x && y && z && m)),
               "null pointer");
  SSIM_REQUIRE(!near_lo || near, "near_lo needs near");
  SSIM_REQUIRE(work && work_bytes >= ssim_split_bin_work_bytes(n, grid_n, block),
               "workspace too small");
  hipStream_t st = ssim::as_stream(stream);
  const int nb = (grid_n + block - 1) / block;
  const int ncells = nb * nb * nb;
  const int B3 = block * block * block;
  const bool nodes = bin_start != nullptr;  // node bins, or the blocks alone
  const long long nbins = nodes ? K : ncells;
  unsigned char* w = static_cast<unsigned char*>(work);
  int* key = reinterpret_cast<int*>(w);
  w += ssim::align256((size_t)(n > 0 ? n : 1) * sizeof(int));
  int* count = reinterpret_cast<int*>(w);
  w += ssim::align256((size_t)K * sizeof(int));
  void* scan_work = w;
  w += ssim::scan_work_bytes((int)K);
  if (nodes) {
    hipLaunchKernelGGL(k_zero_i32, dim3(ssim::ceil_div(nbins, CB)), dim3(CB), 0, st,
                       (int)nbins, count);
    SSIM_LAUNCHED("k_zero_i32");
  }
  if (n > 0) {
    hipLaunchKernelGGL(k_split_key, dim3(ssim::ceil_div(n, CB)), dim3(CB), 0, st, n, x,
                       y, z, origin[0], origin[1], origin[2], h, grid_n, block, nb,
                       nodes ? 1 : 0, key, order, nodes ? count : nullptr);
    SSIM_LAUNCHED("k_split_key");
    int bits = 1;
    while (bits < 31 && (1LL << bits) < nbins) ++bits;
    SSIM_TRY(ssim::radix_sort_pairs(n, bits, key, order, w,
                                    ssim::radix_sort_work_bytes(n), st));
  }
  if (nodes) {
    SSIM_TRY(ssim::scan_exclusive((int)K, count, bin_start, scan_work, st, B3, cell_start));
  } else {
    // the block starts from the sorted keys (n = 0: all zero)
    SSIM_TRY(run_starts(n, key, ncells, cell_start, st));
  }
  if (n > 0) {
    const double hh = 0.5 * h;
    hipLaunchKernelGGL(k_split_pack, dim3(ssim::ceil_div(n, CB)), dim3(CB), 0, st, n,
                       order, x, y, z, m, origin[0], origin[1], origin[2], h,
                       origin[0] - hh, origin[1] - hh, origin[2] - hh,
                       static_cast<float4*>(packed), static_cast<double4*>(body), key,
                       nodes ? B3 : 1, nb, block * h, static_cast<float4*>(near),
                       static_cast<float4*>(near_lo));
    SSIM_LAUNCHED("k_split_pack");
  }
  return SSIM_OK;
}


size_t ssim_split_bin_all_work_bytes(int n_all, int grid_n, int block) {
  if (grid_n < 1 || block < 1) return 0;
  const long long nb = (grid_n + block - 1) / block;
  const long long ncells = nb * nb * nb;
  if (ncells > (1LL << 30)) return 0;
  return ssim::align256((size_t)(n_all > 0 ? n_all : 1) * sizeof(int)) +
         ssim::align256((size_t)ncells * sizeof(int)) + ssim::scan_work_bytes((int)ncells) +
         ssim::radix_sort_work_bytes(n_all);
}

int ssim_split_bin_all(int n_all, int n_mass, const double* x, const double* y,
                       const double* z, const double* m, int grid_n, const double* origin,
                       double h, int block, int* cell_start, int* order, void* packed,
                       void* body, void* near, void* near_lo, int* tcell_start,
                       int* torder, void* work, size_t work_bytes, void* stream) {
  SSIM_REQUIRE(n_all >= 0 && n_mass >= 0 && n_mass <= n_all, "need 0 <= n_mass <= n_all");
  SSIM_REQUIRE(grid_n >= 1 && block >= 1 && h > 0.0 && origin, "bad mesh");
  const long long nbl = (grid_n + block - 1) / block;
  SSIM_REQUIRE(nbl * nbl * nbl <= (1LL << 30), "too many blocks: %lld", nbl * nbl * nbl);
  SSIM_REQUIRE(cell_start && order && packed && body && tcell_start && torder &&
               (n_all == 0 || (x && y && z)) && (n_mass == 0 || m), "null pointer");
  SSIM_REQUIRE(!near_lo || near, "near_lo needs near");
  SSIM_REQUIRE(work && work_bytes >= ssim_split_bin_all_work_bytes(n_all, grid_n, block),
               "workspace too small");
  hipStream_t st = ssim::as_stream(stream);
  const int nb = (int)nbl;
  const int ncells = nb * nb * nb;
  unsigned char* w = static_cast<unsigned char*>(work);
  int* key = reinterpret_cast<int*>(w);
  w += ssim::align256((size_t)(n_all > 0 ? n_all : 1) * sizeof(int));
  int* mcount = reinterpret_cast<int*>(w);
  w += ssim::align256((size_t)ncells * sizeof(int));
  void* scan_work = w;
  w += ssim::scan_work_bytes(ncells);
  int bits = 1;
  while (bits < 31 && (1LL << bits) < ncells) ++bits;
  // the common case: keys, the sort's first counts and the zeroed mass
  // counts in one launch (identity values), the mass counts with the run
  // starts, and a one-launch scan of them
  const bool fused = n_all >= 2 && ncells <= DIRECT_SCAN_CELLS;
  if (n_all >= 2) {
    int hrow = 0;
    int* hist = ssim::radix_first_counts(w, n_all, &hrow);
    hipLaunchKernelGGL(k_split_key_count, dim3(hrow), dim3(ssim::RADIX_T), 0, st, n_all, x,
                       y, z, origin[0], origin[1], origin[2], h, grid_n, block, nb, hrow, key,
                       hist, fused ? ncells : 0, mcount);
    SSIM_LAUNCHED("k_split_key_count");
    SSIM_TRY(ssim::radix_sort_pairs(n_all, bits, key, torder, w,
                                    ssim::radix_sort_work_bytes(n_all), st, true, true));
  } else if (n_all == 1) {
    hipLaunchKernelGGL(k_split_key, dim3(1), dim3(CB), 0, st, n_all, x, y, z, origin[0],
                       origin[1], origin[2], h, grid_n, block, nb, 0, key, torder, nullptr);
    SSIM_LAUNCHED("k_split_key");
  }
  const int sblocks = ssim::ceil_div(ncells, CB);
  const double hh = 0.5 * h;
  if (fused) {
    SSIM_TRY(run_starts(n_all, key, ncells, tcell_start, st, torder, n_mass, mcount));
    hipLaunchKernelGGL(k_cells_scan_pack, dim3(ssim::ceil_div(ncells, PK_CELLS)), dim3(CB), 0,
                       st, ncells, n_mass, mcount, tcell_start, torder, cell_start, order, x,
                       y, z, m, origin[0], origin[1], origin[2], h, origin[0] - hh,
                       origin[1] - hh, origin[2] - hh, static_cast<float4*>(packed),
                       static_cast<double4*>(body), nb, block * h, static_cast<float4*>(near),
                       static_cast<float4*>(near_lo));
    SSIM_LAUNCHED("k_cells_scan_pack");
    return SSIM_OK;
  } else {
    SSIM_TRY(run_starts(n_all, key, ncells, tcell_start, st));
    if (sblocks <= SCAN_BLOCKS &&
        ssim::scan_work_bytes(ncells) >= (sblocks + 1) * sizeof(int)) {
      // the workgroup totals in scan_work
      int* bsum = static_cast<int*>(scan_work);
      hipLaunchKernelGGL(k_mass_sum, dim3(sblocks), dim3(CB), 0, st, ncells, tcell_start,
                         torder, n_mass, mcount, bsum);
      SSIM_LAUNCHED("k_mass_sum");
      hipLaunchKernelGGL(k_cells_scan, dim3(sblocks), dim3(CB), 0, st, ncells, mcount,
                         bsum, cell_start);
      SSIM_LAUNCHED("k_cells_scan");
    } else {
      hipLaunchKernelGGL(k_mass_counts, dim3(ssim::ceil_div(ncells, CB)), dim3(CB), 0, st,
                         ncells, tcell_start, torder, n_mass, mcount);
      SSIM_LAUNCHED("k_mass_counts");
      SSIM_TRY(ssim::scan_exclusive(ncells, mcount, cell_start, scan_work, st));
    }
  }
  if (n_mass > 0) {
    hipLaunchKernelGGL(k_split_pack_all, dim3(ssim::ceil_div(n_all, CB)), dim3(CB), 0, st,
                       n_all, n_mass, torder, key, tcell_start, cell_start, order, x, y, z,
                       m, origin[0], origin[1], origin[2], h, origin[0] - hh,
                       origin[1] - hh, origin[2] - hh, static_cast<float4*>(packed),
                       static_cast<double4*>(body), nb, block * h,
                       static_cast<float4*>(near), static_cast<float4*>(near_lo));
    SSIM_LAUNCHED("k_split_pack_all");
  }
  return SSIM_OK;
}

}  // extern "C"
