// Stable LSD radix sort of (int key, int value) pairs and an exclusive scan,
// shared by the cell lists (cells.hip) and the near-field chunk map (pp.hip).
//
// The sort takes 8 key bits per pass over tiles of RTILE = 1024 elements
// (4 rounds of one element per thread of a 256-thread workgroup):
//   k_radix_hist    — per-tile digit counts of the pass in LDS (digit-major,
//                     hist[d * hrow + b], rows padded to a multiple of 4);
//   k_radix_scatter — the tile's first slot per digit (its own exclusive
//                     prefix over the tile counts, or k_radix_rows's when
//                     there are many tiles), then per round the rank of an
//                     element among the lanes of its wave with the same
//                     digit (nine ballots), wave totals folded in wave
//                     order, and the store.
// Tiles, rounds, waves and lanes are taken in index order, so equal keys
// keep their input order: the sort is stable, and the result does not
// depend on scheduling.  A pass is 2 launches (3 over more than FUSE_TILES
// tiles, with k_radix_rows before the scatter), and the passes alternate
// between two scratch pairs so the last one writes the caller's arrays.
// (Counting the next pass's digits inside the scatter, by destination tile,
// costs global atomics; the extra pass over the keys counts in LDS.)
#include "ssim_internal.h"

namespace {

constexpr int RT = ssim::RADIX_T;  // threads per workgroup (one per digit)
constexpr int RIPT = 4;            // rounds per tile
constexpr int RTILE = RT * RIPT;   // elements per tile
static_assert(RTILE == ssim::RADIX_TILE, "tile of the exported layout");
constexpr int FUSE_TILES = 64;     // up to this many tiles a scatter block
                                   // computes its own offsets (its row loop
                                   // is latency-bound; beyond, k_radix_rows)
constexpr int SCAN_T = 1024;       // threads of the single-workgroup scan

// launched with hrow workgroups: the ones past the last tile write the
// zero padding of the rows
__global__ __launch_bounds__(RT) void k_radix_hist(int n, const int* __restrict__ keys,
                                                   int shift, int hrow,
                                                   int* __restrict__ hist) {
  __shared__ int cnt[256];
  const int t = threadIdx.x;
  cnt[t] = 0;
  __syncthreads();
  const int base = blockIdx.x * RTILE;
  int k[RIPT];
#pragma unroll
  for (int r = 0; r < RIPT; ++r) {
    const int i = base + r * RT + t;
    k[r] = i < n ? keys[i] : 0;
  }
#pragma unroll
  for (int r = 0; r < RIPT; ++r)
    if (base + r * RT + t < n) atomicAdd(&cnt[(k[r] >> shift) & 255], 1);
  __syncthreads();
  hist[(size_t)t * hrow + blockIdx.x] = cnt[t];
}

// Exclusive prefix of v over the 256 threads of the workgroup: inclusive
// scans of the four waves by shuffles, then the totals of the waves before
// (s: 4 ints of LDS, free again on return).
__device__ int excl_scan256(int v, int* s) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int y = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int z = __shfl_up(y, off, 64);
    if (lane >= off) y += z;
  }
  if (lane == 63) s[w] = y;
  __syncthreads();
  int pre = 0;
#pragma unroll
  for (int q = 0; q < RT / 64; ++q) pre += q < w ? s[q] : 0;
  __syncthreads();
  return pre + y - v;
}

// Counts are digit-major, hist[d * hrow + b], rows padded to hrow (a
// multiple of 4, zero beyond the tiles) so a thread sums its digit's row
// with 16-byte loads.  Returns the row total; *pre = the tiles before b.
__device__ __forceinline__ int row_sums(const int* __restrict__ row, int hrow, int b,
                                        int* pre) {
  const int4* r4 = reinterpret_cast<const int4*>(row);
  int tot = 0, p = 0;
#pragma unroll 4
  for (int q = 0; q < hrow / 4; ++q) {
    const int4 c = r4[q];
    const int j = 4 * q;
    tot += c.x + c.y + c.z + c.w;
    p += (j < b ? c.x : 0) + (j + 1 < b ? c.y : 0) + (j + 2 < b ? c.z : 0) +
         (j + 3 < b ? c.w : 0);
  }
  *pre = p;
  return tot;
}

// One workgroup per digit d: offs[d * hrow + b] = the digit's count in the
// tiles before b, tot[d] = its count over all tiles.
__global__ __launch_bounds__(RT) void k_radix_rows(const int* __restrict__ hist, int hrow,
                                                   int* __restrict__ offs,
                                                   int* __restrict__ tot) {
  __shared__ int s[256];
  __shared__ int sum;
  const int t = threadIdx.x, d = blockIdx.x;
  const int4* r4 = reinterpret_cast<const int4*>(hist + (size_t)d * hrow);
  int4* o4 = reinterpret_cast<int4*>(offs + (size_t)d * hrow);
  int carry = 0;
  for (int q0 = 0; q0 < hrow / 4; q0 += RT) {
    const int q = q0 + t;
    const int4 c = q < hrow / 4 ? r4[q] : make_int4(0, 0, 0, 0);
    const int v = c.x + c.y + c.z + c.w;
    const int ex = carry + excl_scan256(v, s);
    if (t == RT - 1) sum = ex + v;
    if (q < hrow / 4) o4[q] = make_int4(ex, ex + c.x, ex + c.x + c.y, ex + c.x + c.y + c.z);
    __syncthreads();
    carry = sum;
    __syncthreads();
  }
  if (t == 0) tot[d] = carry;
}

__global__ __launch_bounds__(RT) void k_radix_scatter(
    int n, const int* __restrict__ keys, const int* __restrict__ vals, int shift,
    int hrow, const int* __restrict__ hist, const int* __restrict__ offs,
    int* __restrict__ keys_out, int* __restrict__ vals_out) {
  constexpr int W = RT / 64;
  __shared__ int s[256];
  __shared__ int run[256];
  __shared__ int wc[W][257];  // per-wave digit counts (257: idle lanes)
  __shared__ int wb[W][256];  // first slot of (wave, digit) in this round
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int base = blockIdx.x * RTILE;
  int k[RIPT], v[RIPT];
#pragma unroll
  for (int r = 0; r < RIPT; ++r) {
    const int i = base + r * RT + t;
    k[r] = i < n ? keys[i] : 0;
    v[r] = vals ? (i < n ? vals[i] : 0) : i;  // vals == NULL: the identity
  }
  if (offs) {
    run[t] = excl_scan256(offs[(size_t)256 * hrow + t], s) + offs[(size_t)t * hrow + blockIdx.x];
  } else {
    int pre;
    const int tot = row_sums(hist + (size_t)t * hrow, hrow, blockIdx.x, &pre);
    run[t] = excl_scan256(tot, s) + pre;
  }
  for (int q = t; q < W * 257; q += RT) (&wc[0][0])[q] = 0;
  __syncthreads();
  const unsigned long long lt = (1ull << lane) - 1ull;
#pragma unroll
  for (int r = 0; r < RIPT; ++r) {
    const bool valid = base + r * RT + t < n;
    const int d = valid ? (k[r] >> shift) & 255 : 256;
    unsigned long long peers = __ballot(1);
#pragma unroll
    for (int b = 0; b < 9; ++b) {
      const unsigned long long bal = __ballot((d >> b) & 1);
      peers &= ((d >> b) & 1) ? bal : ~bal;
    }
    const int rank = __popcll(peers & lt);
    if (rank == 0) wc[w][d] = __popcll(peers);
    __syncthreads();
    int r0 = run[t];
#pragma unroll
    for (int q = 0; q < W; ++q) {
      wb[q][t] = r0;
      r0 += wc[q][t];
      wc[q][t] = 0;
    }
    run[t] = r0;
    __syncthreads();
    if (valid) {
      const int slot = wb[w][d] + rank;
      keys_out[slot] = k[r];
      vals_out[slot] = v[r];
    }
  }
}

__global__ __launch_bounds__(RT) void k_copy_pairs(int n, const int* __restrict__ ka,
                                                   const int* __restrict__ va,
                                                   int* __restrict__ kb,
                                                   int* __restrict__ vb) {
  const int i = blockIdx.x * RT + threadIdx.x;
  if (i >= n) return;
  kb[i] = ka[i];
  vb[i] = va[i];
}

// Exclusive scan of in[0..n) into out[0..n] (out[n] = total) in one
// workgroup: the whole scan for n <= SCAN1_MAX, the block totals of longer
// ones (any n; a round per 32768 values).  Rounds of S1_TILES tiles of 4 SCAN_T
// values: every thread loads its 4 consecutive values of each tile at once
// (coalesced across the wave, one memory latency per round), scans its
// tile sums inside its wave by shuffles, and the first wave scans the
// S1_TILES x 16 wave totals (tile-major); a carry joins the rounds.
constexpr int S1_TILES = 8;
constexpr int SCAN1_MAX = 1 << 16;

__global__ __launch_bounds__(SCAN_T) void k_scan1(int n, const int* __restrict__ in,
                                                  int* __restrict__ out, int every,
                                                  int* __restrict__ sel) {
  constexpr int NW = SCAN_T / 64;
  static_assert(S1_TILES * NW == 128, "two wave totals per lane of the first wave");
  __shared__ int wsum[S1_TILES * NW];
  __shared__ int woff[S1_TILES * NW];
  __shared__ int rtotal;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  int carry = 0;
  for (int base = 0; base < n; base += S1_TILES * 4 * SCAN_T) {
    int v[S1_TILES][4], x[S1_TILES], s[S1_TILES];
#pragma unroll
    for (int k = 0; k < S1_TILES; ++k) {
      const int i0 = base + k * 4 * SCAN_T + 4 * t;
#pragma unroll
      for (int e = 0; e < 4; ++e) v[k][e] = i0 + e < n ? in[i0 + e] : 0;
    }
#pragma unroll
    for (int k = 0; k < S1_TILES; ++k) {
      s[k] = v[k][0] + v[k][1] + v[k][2] + v[k][3];
      int y = s[k];
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int z = __shfl_up(y, off, 64);
        if (lane >= off) y += z;
      }
      x[k] = y;
      if (lane == 63) wsum[k * NW + w] = y;
    }
    __syncthreads();
    if (w == 0) {
      const int a = wsum[2 * lane], b = wsum[2 * lane + 1];
      int y = a + b;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int z = __shfl_up(y, off, 64);
        if (lane >= off) y += z;
      }
      woff[2 * lane] = y - a - b;
      woff[2 * lane + 1] = y - b;
      if (lane == 63) rtotal = y;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < S1_TILES; ++k) {
      const int i0 = base + k * 4 * SCAN_T + 4 * t;
      int e = carry + woff[k * NW + w] + x[k] - s[k];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (i0 + q < n) {
          out[i0 + q] = e;
          if (every && (i0 + q) % every == 0) sel[(i0 + q) / every] = e;
        }
        e += v[k][q];
      }
    }
    carry += rtotal;
    __syncthreads();  // wsum, woff and rtotal are rewritten by the next round
  }
  if (t == 0) {
    out[n] = carry;
    if (every) sel[n / every] = carry;
  }
}

// Long arrays: blocks of SB = 4 * RT elements.  k_block_sum writes each
// block's total, k_scan1 turns them into block offsets, k_block_scan scans a
// block in LDS and adds its offset.
constexpr int SB = 4 * RT;

__device__ __forceinline__ int block_reduce_sum(int v, int* red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int w = RT / 2; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  return red[0];
}

__global__ __launch_bounds__(RT) void k_block_sum(int n, const int* __restrict__ in,
                                                  int* __restrict__ bsum) {
  __shared__ int red[RT];
  const int base = blockIdx.x * SB + 4 * threadIdx.x;
  int v = 0;
  for (int k = 0; k < 4; ++k)
    if (base + k < n) v += in[base + k];
  const int tot = block_reduce_sum(v, red);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(RT) void k_block_scan(int n, const int* __restrict__ in,
                                                   const int* __restrict__ boff,
                                                   int nblocks, int* __restrict__ out,
                                                   int every, int* __restrict__ sel) {
  __shared__ int part[RT];
  const int t = threadIdx.x;
  const int base = blockIdx.x * SB + 4 * t;
  int v[4], s = 0;
  for (int k = 0; k < 4; ++k) {
    v[k] = base + k < n ? in[base + k] : 0;
    s += v[k];
  }
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < RT; off <<= 1) {
    const int add = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += add;
    __syncthreads();
  }
  int run = boff[blockIdx.x] + (t == 0 ? 0 : part[t - 1]);
  for (int k = 0; k < 4; ++k) {
    if (base + k < n) {
      out[base + k] = run;
      if (every && (base + k) % every == 0) sel[(base + k) / every] = run;
    }
    run += v[k];
  }
  if (blockIdx.x == 0 && t == 0) {
    out[n] = boff[nblocks];
    if (every) sel[n / every] = boff[nblocks];
  }
}

// Mid-size arrays (SCAN1_FAST < n <= SB x SB2_MAX): k_block_sum writes each
// block's total, then every block of k_block_scan2 sums the totals before
// its own (all threads, <= SB2_MAX / RT loads each) and scans its block:
// two fully parallel launches instead of a single-workgroup pass.
constexpr int SCAN1_FAST = 4096;
constexpr int SB2_MAX = 1024;

__global__ __launch_bounds__(RT) void k_block_scan2(int n, const int* __restrict__ in,
                                                    const int* __restrict__ bsum,
                                                    int* __restrict__ out, int every,
                                                    int* __restrict__ sel) {
  __shared__ int red[4];
  static_assert(RT == 256, "ssim_block256 helpers");
  const int off = ssim_block256_offset(bsum, blockIdx.x, red);
  const int base = blockIdx.x * SB + 4 * threadIdx.x;
  int v[4], s = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = base + k < n ? in[base + k] : 0;
    s += v[k];
  }
  int tot;
  int run = off + ssim_block256_excl(s, red, &tot);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (base + k < n) {
      out[base + k] = run;
      if (every && (base + k) % every == 0) sel[(base + k) / every] = run;
    }
    run += v[k];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
    out[n] = off + tot;
    if (every) sel[n / every] = off + tot;
  }
}

}  // namespace

namespace ssim {

size_t scan_work_bytes(int n) {
  return 2 * align256(((size_t)ceil_div(n > 0 ? n : 1, SB) + 1) * sizeof(int));
}

int scan_exclusive(int n, const int* in, int* out, void* work, hipStream_t st, int every,
                   int* sel) {
  SSIM_REQUIRE(every == 0 || (every > 0 && n % every == 0 && sel),
               "scan: every must divide n");
  const int nb = ceil_div(n, SB);
  if (n <= SCAN1_FAST || (n <= SCAN1_MAX && !work)) {
    hipLaunchKernelGGL(k_scan1, dim3(1), dim3(SCAN_T), 0, st, n, in, out, every, sel);
    SSIM_LAUNCHED("k_scan1");
    return SSIM_OK;
  }
  SSIM_REQUIRE(work, "scan: workspace missing");
  if (nb <= SB2_MAX) {
    int* bsum = static_cast<int*>(work);
    hipLaunchKernelGGL(k_block_sum, dim3(nb), dim3(RT), 0, st, n, in, bsum);
    SSIM_LAUNCHED("k_block_sum");
    hipLaunchKernelGGL(k_block_scan2, dim3(nb), dim3(RT), 0, st, n, in, bsum, out, every,
                       sel);
    SSIM_LAUNCHED("k_block_scan2");
    return SSIM_OK;
  }
  int* bsum = static_cast<int*>(work);
  int* boff = bsum + align256(((size_t)nb + 1) * sizeof(int)) / sizeof(int);
  hipLaunchKernelGGL(k_block_sum, dim3(nb), dim3(RT), 0, st, n, in, bsum);
  SSIM_LAUNCHED("k_block_sum");
  hipLaunchKernelGGL(k_scan1, dim3(1), dim3(SCAN_T), 0, st, nb, bsum, boff, 0, nullptr);
  SSIM_LAUNCHED("k_scan1");
  hipLaunchKernelGGL(k_block_scan, dim3(nb), dim3(RT), 0, st, n, in, boff, nb, out, every,
                     sel);
  SSIM_LAUNCHED("k_block_scan");
  return SSIM_OK;
}

size_t radix_sort_work_bytes(int n) {
  const size_t nn = (size_t)(n > 0 ? n : 1);
  const size_t tiles = (nn + RTILE - 1) / RTILE;
  const size_t hrow = (tiles + 3) & ~size_t(3);
  return 4 * align256(nn * sizeof(int)) + align256(4 * 256 * hrow * sizeof(int)) +
         align256((256 * hrow + 256) * sizeof(int));
}

int* radix_first_counts(void* work, int n, int* hrow) {
  const int tiles = ceil_div(n > 0 ? n : 1, RTILE);
  *hrow = (tiles + 3) & ~3;
  return reinterpret_cast<int*>(static_cast<unsigned char*>(work) +
                                4 * align256((size_t)(n > 0 ? n : 1) * sizeof(int)));
}

int radix_sort_pairs(int n, int bits, int* keys, int* vals, void* work,
                     size_t work_bytes, hipStream_t st, bool first_counted,
                     bool vals_identity) {
  SSIM_REQUIRE(!(first_counted || vals_identity) || n >= 2,
               "radix sort: counted keys / identity values need n >= 2");
  if (n <= 1) return SSIM_OK;
  SSIM_REQUIRE(bits >= 1 && bits <= 31, "radix sort: bad key width %d", bits);
  SSIM_REQUIRE(work && work_bytes >= radix_sort_work_bytes(n),
               "radix sort: workspace too small");
  const int tiles = ceil_div(n, RTILE);
  const int hrow = (tiles + 3) & ~3;  // padded row of one digit's tile counts
  const int passes = (bits + 7) / 8;
  const size_t ht = 256 * (size_t)hrow;
  unsigned char* w = static_cast<unsigned char*>(work);
  int* pk[2];
  int* pv[2];
  for (int q = 0; q < 2; ++q) {
    pk[q] = reinterpret_cast<int*>(w);
    w += align256((size_t)n * sizeof(int));
    pv[q] = reinterpret_cast<int*>(w);
    w += align256((size_t)n * sizeof(int));
  }
  int* hist = reinterpret_cast<int*>(w);
  w += align256(4 * ht * sizeof(int));
  int* offs = reinterpret_cast<int*>(w);
  const int* sk = keys;
  const int* sv = vals_identity ? nullptr : vals;
  for (int p = 0; p < passes; ++p) {
    const bool last = p == passes - 1 && passes >= 2;
    int* dk = last ? keys : pk[p & 1];
    int* dv = last ? vals : pv[p & 1];
    const int* h = hist;
    if (p > 0 || !first_counted) {
      hipLaunchKernelGGL(k_radix_hist, dim3(hrow), dim3(RT), 0, st, n, sk, 8 * p, hrow,
                         hist);
      SSIM_LAUNCHED("k_radix_hist");
    }
    const int* o = nullptr;
    if (tiles > FUSE_TILES) {
      hipLaunchKernelGGL(k_radix_rows, dim3(256), dim3(RT), 0, st, h, hrow, offs,
                         offs + ht);
      SSIM_LAUNCHED("k_radix_rows");
      o = offs;
    }
    hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(RT), 0, st, n, sk, sv, 8 * p,
                       hrow, h, o, dk, dv);
    SSIM_LAUNCHED("k_radix_scatter");
    sk = dk;
    sv = dv;
  }
  if (passes == 1) {
    hipLaunchKernelGGL(k_copy_pairs, dim3(ceil_div(n, RT)), dim3(RT), 0, st, n, sk, sv,
                       keys, vals);
    SSIM_LAUNCHED("k_copy_pairs");
  }
  return SSIM_OK;
}

}  // namespace ssim
