// Stable LSD radix sort of (int key, int value) pairs and an exclusive scan,
// shared by the cell list (cells.hip).
//
// Each pass sorts on 8 bits of the key:
//   1. k_radix_hist   — every 256-element tile counts its digits in LDS and
//                       writes the counts digit-major: hist[d * tiles + b];
//   2. scan_exclusive — the exclusive prefix of hist is the first output
//                       slot of digit d coming from tile b;
//   3. k_radix_scatter— an element's slot is that prefix plus the number of
//                       elements before it in its tile with the same digit.
// Tiles are taken in index order and ranks inside a tile follow the index,
// so equal keys keep their input order: the sort is stable, and the result
// does not depend on scheduling.
#include "ssim_internal.h"

namespace {

constexpr int RT = 256;       // tile = threads per workgroup
constexpr int SCAN_T = 1024;  // threads of the single-workgroup scan

__global__ __launch_bounds__(RT) void k_radix_hist(int n, const int* __restrict__ keys,
                                                   int shift, int tiles,
                                                   int* __restrict__ hist) {
  __shared__ int cnt[256];
  const int t = threadIdx.x;
  const int i = blockIdx.x * RT + t;
  cnt[t] = 0;
  __syncthreads();
  if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 255], 1);
  __syncthreads();
  hist[(size_t)t * tiles + blockIdx.x] = cnt[t];
}

__global__ __launch_bounds__(RT) void k_radix_scatter(
    int n, const int* __restrict__ keys, const int* __restrict__ vals, int shift,
    int tiles, const int* __restrict__ offs, int* __restrict__ keys_out,
    int* __restrict__ vals_out) {
  __shared__ int dig[RT];
  const int t = threadIdx.x;
  const int i = blockIdx.x * RT + t;
  const int k = i < n ? keys[i] : 0;
  const int d = i < n ? (k >> shift) & 255 : -1;
  dig[t] = d;
  __syncthreads();
  if (i >= n) return;
  // rank among the earlier elements of the tile with the same digit; every
  // thread reads the same dig[j] in step j (an LDS broadcast)
  int rank = 0;
  for (int j = 0; j < RT; ++j) rank += (j < t) & (dig[j] == d);
  const int slot = offs[(size_t)d * tiles + blockIdx.x] + rank;
  keys_out[slot] = k;
  vals_out[slot] = vals[i];
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

// Exclusive scan of in[0..n) into out[0..n], one workgroup: every thread
// sums a contiguous chunk, a Hillis-Steele scan over the chunk sums gives
// each chunk's base, and the chunks are written out.  out[n] = total.
// Used directly for short arrays and for the block sums of long ones.
__global__ __launch_bounds__(SCAN_T) void k_scan(int n, const int* __restrict__ in,
                                                 int* __restrict__ out) {
  __shared__ int part[SCAN_T];
  const int t = threadIdx.x;
  const int per = (n + SCAN_T - 1) / SCAN_T;
  const int b0 = min(n, t * per), b1 = min(n, b0 + per);
  int s = 0;
  for (int i = b0; i < b1; ++i) s += in[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < SCAN_T; off <<= 1) {
    const int v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = t == 0 ? 0 : part[t - 1];
  for (int i = b0; i < b1; ++i) {
    const int c = in[i];
    out[i] = run;
    run += c;
  }
  if (t == SCAN_T - 1) out[n] = part[SCAN_T - 1];
}

// Long arrays: blocks of SB = 4 * RT elements.  k_block_sum writes each
// block's total, k_scan turns them into block offsets, k_block_scan scans a
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
                                                   int nblocks, int* __restrict__ out) {
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
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (blockIdx.x == 0 && t == 0) out[n] = boff[nblocks];
}

}  // namespace

namespace ssim {

size_t scan_work_bytes(int n) {
  return 2 * align256(((size_t)ceil_div(n > 0 ? n : 1, SB) + 1) * sizeof(int));
}

int scan_exclusive(int n, const int* in, int* out, void* work, hipStream_t st) {
  if (n <= 4 * SB) {
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(SCAN_T), 0, st, n, in, out);
    SSIM_LAUNCHED("k_scan");
    return SSIM_OK;
  }
  SSIM_REQUIRE(work, "scan: workspace missing");
  const int nb = ceil_div(n, SB);
  int* bsum = static_cast<int*>(work);
  int* boff = bsum + align256(((size_t)nb + 1) * sizeof(int)) / sizeof(int);
  hipLaunchKernelGGL(k_block_sum, dim3(nb), dim3(RT), 0, st, n, in, bsum);
  SSIM_LAUNCHED("k_block_sum");
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(SCAN_T), 0, st, nb, bsum, boff);
  SSIM_LAUNCHED("k_scan");
  hipLaunchKernelGGL(k_block_scan, dim3(nb), dim3(RT), 0, st, n, in, boff, nb, out);
  SSIM_LAUNCHED("k_block_scan");
  return SSIM_OK;
}

size_t radix_sort_work_bytes(int n) {
  const size_t nn = (size_t)(n > 0 ? n : 1);
  const size_t tiles = (nn + RT - 1) / RT;
  return 2 * align256(nn * sizeof(int)) + 2 * align256((256 * tiles + 1) * sizeof(int)) +
         scan_work_bytes((int)(256 * tiles));
}

int radix_sort_pairs(int n, int bits, int* keys, int* vals, void* work,
                     size_t work_bytes, hipStream_t st) {
  if (n <= 1) return SSIM_OK;
  SSIM_REQUIRE(bits >= 1 && bits <= 31, "radix sort: bad key width %d", bits);
  SSIM_REQUIRE(work && work_bytes >= radix_sort_work_bytes(n),
               "radix sort: workspace too small");
  const int tiles = ceil_div(n, RT);
  unsigned char* w = static_cast<unsigned char*>(work);
  int* kt = reinterpret_cast<int*>(w);
  w += align256((size_t)n * sizeof(int));
  int* vt = reinterpret_cast<int*>(w);
  w += align256((size_t)n * sizeof(int));
  int* hist = reinterpret_cast<int*>(w);
  w += align256((256 * (size_t)tiles + 1) * sizeof(int));
  int* offs = reinterpret_cast<int*>(w);
  w += align256((256 * (size_t)tiles + 1) * sizeof(int));
  void* scan_work = w;
  int *ka = keys, *va = vals, *kb = kt, *vb = vt;
  const int passes = (bits + 7) / 8;
  for (int p = 0; p < passes; ++p) {
    const int shift = 8 * p;
    hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(RT), 0, st, n, ka, shift,
                       tiles, hist);
    SSIM_LAUNCHED("k_radix_hist");
    SSIM_TRY(scan_exclusive(256 * tiles, hist, offs, scan_work, st));
    hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(RT), 0, st, n, ka, va,
                       shift, tiles, offs, kb, vb);
    SSIM_LAUNCHED("k_radix_scatter");
    int* tk = ka; ka = kb; kb = tk;
    int* tv = va; va = vb; vb = tv;
  }
  if (ka != keys) {
    hipLaunchKernelGGL(k_copy_pairs, dim3(tiles), dim3(RT), 0, st, n, ka, va,
                       keys, vals);
    SSIM_LAUNCHED("k_copy_pairs");
  }
  return SSIM_OK;
}

}  // namespace ssim
