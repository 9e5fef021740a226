// Uniform-grid asteroid-asteroid collision search with the reference's
// exact pair list (`Erkennungskarte._hits`, film_producer.py:223-317).
//
// Order and arithmetic follow the reference so that the list is bit-exact:
//  * alive = asteroid & visible & x in [lo, hi + halo) (FP32 compares);
//    alive bodies are enumerated in ascending index (the "rows");
//  * cell ix = floor(f64(x) / h) as int64, key = ix*73856093 XOR
//    iy*19349663 (int64 wrap-around, `_zellschluessel` line 320-321);
//  * the rows sorted by key, ties by row (a stable argsort);
//  * for every neighbour offset in (0,0),(1,0),(0,1),(1,1),(1,-1) and every
//    row, the candidates are the sorted entries whose key equals the
//    offset cell's key (hash collisions included, like searchsorted on the
//    keys); pair (a, b) is kept if a < b for (0,0) and a != b otherwise;
//  * FP32 closest-approach prefilter over [-dt, dt] with the radius sum
//    widened by 1e-6 AU, every operation rounded separately (the reference
//    runs one CuPy elementwise kernel per operation; built here without FMA
//    contraction);
//  * ownership min(x_a, x_b) < hi, then the FP64 recheck on the FP32 inputs.
// The output is ordered by (offset, row, sorted position).  The rows are
// sorted by key with a stable LSD radix sort (three passes of the 31-bit
// sorter over the 64-bit key); each (offset, row) has one contiguous range
// of sorted candidates, so the candidate list is the concatenation of those
// ranges in output order.  Flattened that way, the candidates are cut into
// runs of 32 per thread (load-balanced however the cells are filled): one
// pass evaluates the FP32 prefilter once per candidate into a bitmask, a
// scan of the per-thread hit counts gives every hit its place, and a second
// pass writes the hits — the same list as the reference's, in its order,
// without atomics.  The candidates go in batches of 2^25 so the workspace
// does not depend on their number.
#include <algorithm>
#include <climits>

#include "ssim_internal.h"

namespace {

constexpr int CB = 256;
constexpr long long HX = 73856093LL, HY = 19349663LL;

__global__ __launch_bounds__(CB) void k_alive_flags(
    int n, const float* __restrict__ gx, const unsigned char* __restrict__ ast,
    const unsigned char* __restrict__ vis, float lo, float hi_halo,
    int* __restrict__ flag) {
  const int i = blockIdx.x * CB + threadIdx.x;
  if (i >= n) return;
  flag[i] = (ast[i] && vis[i] != 0 && gx[i] >= lo && gx[i] < hi_halo) ? 1 : 0;
}

// Block-wise exclusive scan with per-block totals (3-phase scan).
__global__ __launch_bounds__(1024) void k_scan_block(int n, const int* __restrict__ in,
                                                     int* __restrict__ out,
                                                     int* __restrict__ totals) {
  __shared__ int s[1024];
  const int i = blockIdx.x * 1024 + threadIdx.x;
  const int v = i < n ? in[i] : 0;
  s[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int t = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < n) out[i] = s[threadIdx.x] - v;
  if (threadIdx.x == 1023) totals[blockIdx.x] = s[1023];
}

__global__ void k_scan_totals(int nb, int* __restrict__ totals, int* __restrict__ sum) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int run = 0;
  for (int b = 0; b < nb; ++b) {
    const int t = totals[b];
    totals[b] = run;
    run += t;
  }
  *sum = run;
}

__global__ __launch_bounds__(1024) void k_scan_add(int n, int* __restrict__ out,
                                                   const int* __restrict__ totals) {
  const int i = blockIdx.x * 1024 + threadIdx.x;
  if (i < n) out[i] += totals[blockIdx.x];
}

__global__ __launch_bounds__(CB) void k_compact_alive(int n, const int* __restrict__ flag,
                                                      const int* __restrict__ pos,
                                                      int* __restrict__ ai) {
  const int i = blockIdx.x * CB + threadIdx.x;
  if (i < n && flag[i]) ai[pos[i]] = i;
}

__device__ __forceinline__ long long cell_key(long long ix, long long iy) {
  // int64 products wrap like NumPy/CuPy int64 arithmetic
  const unsigned long long a = (unsigned long long)ix * (unsigned long long)HX;
  const unsigned long long b = (unsigned long long)iy * (unsigned long long)HY;
  return (long long)(a ^ b);
}

// Cell and key of every alive row; the key's sign bit is flipped so that
// unsigned order is the reference's signed order.
__global__ __launch_bounds__(CB) void k_keys(int na, const int* __restrict__ ai,
                                             const float* __restrict__ gx,
                                             const float* __restrict__ gy, double h,
                                             long long* __restrict__ ix,
                                             long long* __restrict__ iy,
                                             unsigned long long* __restrict__ ukey,
                                             int* __restrict__ rkey,
                                             int* __restrict__ srow) {
  const int p = blockIdx.x * CB + threadIdx.x;
  if (p >= na) return;
  const int b = ai[p];
  const long long cx = (long long)floor((double)gx[b] / h);
  const long long cy = (long long)floor((double)gy[b] / h);
  ix[p] = cx;
  iy[p] = cy;
  const unsigned long long k = (unsigned long long)cell_key(cx, cy) ^ 0x8000000000000000ULL;
  ukey[p] = k;
  rkey[p] = (int)(k & 0x7fffffffULL);  // bits 0..30: the first radix pass
  srow[p] = p;
}

// The next 31-bit (or last 2-bit) slice of the key, in the current order.
__global__ __launch_bounds__(CB) void k_key_slice(int na, int shift, unsigned mask,
                                                  const unsigned long long* __restrict__ ukey,
                                                  const int* __restrict__ srow,
                                                  int* __restrict__ rkey) {
  const int c = blockIdx.x * CB + threadIdx.x;
  if (c >= na) return;
  rkey[c] = (int)((ukey[srow[c]] >> shift) & mask);
}

// Sorted keys and the candidates' data in sorted order: (x, y, vx, vy),
// radius and body index of sorted entry c; and the same data of the rows
// (alive bodies in index order), ra4 / rar.
__global__ __launch_bounds__(CB) void k_sorted_pack(
    int na, const int* __restrict__ ai, const int* __restrict__ srow,
    const unsigned long long* __restrict__ ukey, const float* __restrict__ gx,
    const float* __restrict__ gy, const float* __restrict__ gvx,
    const float* __restrict__ gvy, const float* __restrict__ rr,
    unsigned long long* __restrict__ skey, float4* __restrict__ sb,
    float* __restrict__ srr, int* __restrict__ sidx, float4* __restrict__ ra4,
    float* __restrict__ rar) {
  const int c = blockIdx.x * CB + threadIdx.x;
  if (c >= na) return;
  const int r = srow[c];
  const int b = ai[r];
  skey[c] = ukey[r];
  sb[c] = make_float4(gx[b], gy[b], gvx[b], gvy[b]);
  srr[c] = rr[b];
  sidx[c] = b;
  const int a = ai[c];
  ra4[c] = make_float4(gx[a], gy[a], gvx[a], gvy[a]);
  rar[c] = rr[a];
}

__device__ __forceinline__ int lower_bound(const unsigned long long* a, int n,
                                           unsigned long long v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ int upper_bound(const unsigned long long* a, int n,
                                           unsigned long long v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] <= v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

struct Pf {
  const float *gx, *gy, *gvx, *gvy, *rr;
  float dt, tol;
};

// FP32 prefilter of one candidate pair (a: row body, b: candidate), one
// rounding per operation.
__device__ __forceinline__ bool prefilter_v(float4 A, float ra, float4 B, float rb, float dt,
                                           float tol) {
  const float dpx = B.x - A.x;
  const float dpy = B.y - A.y;
  const float dvx = B.z - A.z;
  const float dvy = B.w - A.w;
  const float dv2 = dvx * dvx + dvy * dvy;
  const float num = -(dpx * dvx + dpy * dvy);
  const float den = dv2 > 0.0f ? dv2 : 1.0f;
  float t = num / den;
  t = t < -dt ? -dt : (t > dt ? dt : t);
  const float cx = dpx + dvx * t;
  const float cy = dpy + dvy * t;
  const float rs = ra + rb + tol;
  return cx * cx + cy * cy <= rs * rs;
}

__device__ __forceinline__ bool prefilter(const Pf& P, int a, int b) {
  return prefilter_v(make_float4(P.gx[a], P.gy[a], P.gvx[a], P.gvy[a]), P.rr[a],
                     make_float4(P.gx[b], P.gy[b], P.gvx[b], P.gvy[b]), P.rr[b], P.dt,
                     P.tol);
}

// Candidate range of every (offset, row) t = o na + p: the sorted entries
// whose key is the offset cell's key (start, length).
__global__ __launch_bounds__(CB) void k_segments(int na, const long long* __restrict__ ix,
                                                 const long long* __restrict__ iy,
                                                 const unsigned long long* __restrict__ skey,
                                                 int* __restrict__ sstart,
                                                 int* __restrict__ slen) {
  const int t = blockIdx.x * CB + threadIdx.x;
  if (t >= 5 * na) return;
  const int o = t / na, p = t % na;
  const int ox[5] = {0, 1, 0, 1, 1}, oy[5] = {0, 0, 1, 1, -1};
  const unsigned long long k2 =
      (unsigned long long)cell_key(ix[p] + ox[o], iy[p] + oy[o]) ^ 0x8000000000000000ULL;
  const int lo = lower_bound(skey, na, k2);
  sstart[t] = lo;
  slen[t] = upper_bound(skey, na, k2) - lo;
}

// A wave takes 64 cpt consecutive candidates of the flattened list, in cpt
// rounds of 64 consecutive ones (lane l, round q: candidate gw + 64 q + l,
// so the sorted candidates' loads are coalesced); each lane follows the
// segment (offset, row) of its candidate forward through the segment
// offsets soff (5 na + 1 of them).  pass 0 stores each round's hit ballot
// and the wave's hit count; pass 1 writes the hits in list order: place =
// hits of earlier waves (wpos) + earlier rounds + earlier lanes.  cpt (1 ..
// CPT) is chosen per batch so that a launch has at least WAVES_MIN waves:
// few candidates take few rounds each instead of filling few waves.
constexpr int CPT = 32;
constexpr long long WAVES_MIN = 8192;
constexpr long long CBATCH = 1LL << 25;  // candidates per batch (workspace size)
long long g_batch = CBATCH;              // the batch in use (ssim_collide_batch)

__global__ __launch_bounds__(CB) void k_cand_waves(
    int na, long long base, long long batch, long long total, const int* __restrict__ soff,
    const int* __restrict__ sstart, const int* __restrict__ ai,
    const float4* __restrict__ ra4, const float* __restrict__ rar,
    const float4* __restrict__ sb, const float* __restrict__ srr,
    const int* __restrict__ sidx, float dt, float tol, int pass, int cpt,
    unsigned long long* __restrict__ bits, int* __restrict__ wcnt,
    const int* __restrict__ wpos, int hoff, int* __restrict__ outa, int* __restrict__ outb) {
  const int lane = threadIdx.x & 63;
  const long long wave = ((long long)blockIdx.x * CB + threadIdx.x) >> 6;
  const long long gw = base + wave * 64 * cpt;
  const long long gend = min(total, base + batch);
  if (gw >= gend) return;  // whole waves
  const int nseg = 5 * na;
  // the segment of this lane's first candidate (or of the last one)
  const long long g0 = min(gw + lane, gend - 1);
  int lo = 0, hi = nseg;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if ((long long)soff[mid] <= g0) lo = mid; else hi = mid;
  }
  int t = lo;
  long long sbeg = soff[t], send = soff[t + 1];
  int o = t / na, p = t % na;
  int a = ai[p];
  float4 A = ra4[p];
  float ra = rar[p];
  const unsigned long long lt = (1ull << lane) - 1ull;
  unsigned long long* wb = bits + wave * cpt;
  int cnt = 0;
  int w = pass ? hoff + wpos[wave] : 0;
  for (int q = 0; q < cpt; ++q) {
    const long long g = gw + (long long)q * 64 + lane;
    const bool in = g < gend;
    bool hit = false;
    int b = 0;
    if (in) {
      while (g >= send) {  // the next non-empty segment
        ++t;
        sbeg = send;
        send = soff[t + 1];
        o = t / na;
        p = t % na;
        a = ai[p];
        A = ra4[p];
        ra = rar[p];
      }
      const int c = sstart[t] + (int)(g - sbeg);
      b = sidx[c];
      if (pass == 0) {
        const bool keep = (o == 0) ? (a < b) : (a != b);
        hit = keep && prefilter_v(A, ra, sb[c], srr[c], dt, tol);
      }
    }
    if (pass == 0) {
      const unsigned long long bal = __ballot(hit);
      if (lane == 0) wb[q] = bal;
      cnt += __popcll(bal);
    } else {
      const unsigned long long bal = wb[q];
      if ((bal >> lane) & 1ull) {
        const int at = w + __popcll(bal & lt);
        outa[at] = a;
        outb[at] = b;
      }
      w += __popcll(bal);
    }
  }
  if (pass == 0 && lane == 0) wcnt[wave] = cnt;
}

// Ownership rule + FP64 recheck on the FP32 inputs -> flag per f32 hit.
__global__ __launch_bounds__(CB) void k_recheck(int nh, const int* __restrict__ ha,
                                                const int* __restrict__ hb, Pf P,
                                                double dt, float hi,
                                                int* __restrict__ flag) {
  const int t = blockIdx.x * CB + threadIdx.x;
  if (t >= nh) return;
  const int a = ha[t], b = hb[t];
  int ok = 0;
  if (fminf(P.gx[a], P.gx[b]) < hi) {
    const double dpx = (double)P.gx[b] - (double)P.gx[a];
    const double dpy = (double)P.gy[b] - (double)P.gy[a];
    const double dvx = (double)P.gvx[b] - (double)P.gvx[a];
    const double dvy = (double)P.gvy[b] - (double)P.gvy[a];
    const double dv2 = dvx * dvx + dvy * dvy;
    double tm = -(dpx * dvx + dpy * dvy) / (dv2 > 0.0 ? dv2 : 1.0);
    tm = tm < -dt ? -dt : (tm > dt ? dt : tm);
    const double cx = dpx + dvx * tm;
    const double cy = dpy + dvy * tm;
    const double rs = (double)P.rr[a] + (double)P.rr[b];
    ok = cx * cx + cy * cy <= rs * rs;
  }
  flag[t] = ok;
}

__global__ __launch_bounds__(CB) void k_compact_pairs(int nh, const int* __restrict__ flag,
                                                      const int* __restrict__ pos,
                                                      const int* __restrict__ ha,
                                                      const int* __restrict__ hb,
                                                      long long* __restrict__ oi,
                                                      long long* __restrict__ oj,
                                                      long long cap) {
  const int t = blockIdx.x * CB + threadIdx.x;
  if (t >= nh || !flag[t]) return;
  const long long q = pos[t];
  if (q < cap) {
    oi[q] = ha[t];
    oj[q] = hb[t];
  }
}

// *out (zeroed first) += the sum of a[0..n): 16 values per thread, one
// 64-bit atomic per workgroup (integer adds: the order does not matter)
__global__ void k_zero_u64_1(unsigned long long* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *out = 0ull;
}

__global__ __launch_bounds__(CB) void k_sum_i32_u64(int n, const int* __restrict__ a,
                                                    unsigned long long* __restrict__ out) {
  __shared__ unsigned long long s[CB / 64];
  const int base = blockIdx.x * CB * 16 + threadIdx.x;
  unsigned long long v = 0;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int i = base + k * CB;
    if (i < n) v += (unsigned long long)a[i];
  }
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int q = 0; q < CB / 64; ++q) t += s[q];
    atomicAdd(out, t);
  }
}


// Exclusive scan of n ints (any n) with the 3-phase kernels.
int scan(int n, const int* in, int* out, int* totals, int* sum, hipStream_t st) {
  const int nb = ssim::ceil_div(n > 0 ? n : 1, 1024);
  hipLaunchKernelGGL(k_scan_block, dim3(nb), dim3(1024), 0, st, n, in, out, totals);
  hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(64), 0, st, nb, totals, sum);
  hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(1024), 0, st, n, out, totals);
  SSIM_LAUNCHED("scan");
  return SSIM_OK;
}

struct Layout {
  size_t flag, pos, ai, ix, iy, ukey, rkey, srow, skey, sb, srr, sidx, ra4, rar, sstart,
      slen, soff, bits, wcnt, wpos, sums, scan, radix, ha, hb, hflag, hpos, total;
};

Layout layout(int n, long long hcap) {
  Layout L{};
  const size_t nn = (size_t)(n > 0 ? n : 1);
  const size_t nw = (size_t)(CBATCH / 64);     // waves of a batch at cpt = 1
  const size_t hc = (size_t)(hcap > 0 ? hcap : 1);
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o += ssim::align256(bytes);
    return at;
  };
  L.flag = take(nn * 4);
  L.pos = take((nn + 1) * 4);
  L.ai = take(nn * 4);
  L.ix = take(nn * 8);
  L.iy = take(nn * 8);
  L.ukey = take(nn * 8);
  L.rkey = take(nn * 4);
  L.srow = take(nn * 4);
  L.skey = take(nn * 8);
  L.sb = take(nn * 16);
  L.srr = take(nn * 4);
  L.sidx = take(nn * 4);
  L.ra4 = take(nn * 16);
  L.rar = take(nn * 4);
  L.sstart = take(5 * nn * 4);
  L.slen = take(5 * nn * 4);
  L.soff = take((5 * nn + 1) * 4);
  L.bits = take(nw * 8);
  L.wcnt = take(nw * 4);
  L.wpos = take((nw + 1) * 4);
  L.sums = take(64);
  // the largest scan: 5 rows per body, a batch's waves or the FP32 hits
  size_t scan_max = 5 * nn > nw ? 5 * nn : nw;
  scan_max = scan_max > hc ? scan_max : hc;
  L.scan = take(ssim::scan_work_bytes((int)scan_max));
  L.radix = take(ssim::radix_sort_work_bytes((int)nn));
  L.ha = take(hc * 4);
  L.hb = take(hc * 4);
  L.hflag = take(hc * 4);
  L.hpos = take((hc + 1) * 4);
  L.total = o;
  return L;
}

// ---- the cell size's speed percentile (film_producer.py:357) ----

// alive speeds as float bits (>= 0, so their integer order is their order):
// sqrt of the FP64 sum of squares, rounded once to FP32
__global__ __launch_bounds__(CB) void k_alive_speed(int n, const float* __restrict__ gvx,
                                                    const float* __restrict__ gvy,
                                                    const unsigned char* __restrict__ ast,
                                                    const unsigned char* __restrict__ vis,
                                                    const int* __restrict__ pos,
                                                    int* __restrict__ bits,
                                                    int* __restrict__ idx) {
  const int i = blockIdx.x * CB + threadIdx.x;
  if (i >= n || !(ast[i] && vis[i] != 0)) return;
  const double vx = gvx[i], vy = gvy[i];
  const float sp = (float)sqrt(vx * vx + vy * vy);
  bits[pos[i]] = __float_as_int(sp);
  idx[pos[i]] = pos[i];
}

// slots m..n-1 (m = the alive count, on the device) get the largest key
__global__ __launch_bounds__(CB) void k_pad_bits(int n, const int* __restrict__ mcount,
                                                 int* __restrict__ bits, int* __restrict__ idx) {
  const int i = blockIdx.x * CB + threadIdx.x;
  if (i < n && i >= *mcount) {
    bits[i] = 0x7fffffff;
    idx[i] = i;
  }
}

__global__ __launch_bounds__(CB) void k_flag_alive(int n, const unsigned char* __restrict__ ast,
                                                   const unsigned char* __restrict__ vis,
                                                   int* __restrict__ flag) {
  const int i = blockIdx.x * CB + threadIdx.x;
  if (i < n) flag[i] = (ast[i] && vis[i] != 0) ? 1 : 0;
}

// NumPy's 'linear' percentile of m sorted FP32 values, in FP32 as NumPy
// computes it for an FP32 array: virtual index (m - 1) * (q / 100), gamma
// its fraction, a + (b - a) gamma (b - (b - a)(1 - gamma) for gamma >= 1/2)
__global__ void k_percentile_lerp(const int* __restrict__ sorted_bits,
                                  const int* __restrict__ mcount, float qf,
                                  float* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int m = *mcount;
  if (m < 1) {
    out[0] = __int_as_float(0x7fc00000);
    return;
  }
  const float vi = (float)(m - 1) * qf;
  const float fl = floorf(vi);
  int lo = (int)fl;
  lo = lo < 0 ? 0 : (lo > m - 1 ? m - 1 : lo);
  const int hi = lo + 1 < m ? lo + 1 : m - 1;
  const float g = vi - fl;
  const float a = __int_as_float(sorted_bits[lo]), b = __int_as_float(sorted_bits[hi]);
  const float d = b - a;
  out[0] = g >= 0.5f ? b - d * (1.0f - g) : a + d * g;
}

// ---- massive-body merges and runaway guard (`analyze_merge`,
// film_producer.py:707-779), FP32 like the reference's CuPy arrays ----

// Does the segment p0 -> p1 pass within rsum of c?  tt = parameter of the
// closest point in [0, 1] (seg_hit, film_producer.py:730-746).
__device__ __forceinline__ bool seg_hit(float p0x, float p0y, float p1x,
                                        float p1y, float cx, float cy,
                                        float rsum2, float& tt) {
  const float sx = p1x - p0x, sy = p1y - p0y;
  const float seg2 = sx * sx + sy * sy;
  float t = ((cx - p0x) * sx + (cy - p0y) * sy) / (seg2 > 0.0f ? seg2 : 1.0f);
  t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
  tt = t;
  const float dx = p0x + t * sx - cx;
  const float dy = p0y + t * sy - cy;
  return dx * dx + dy * dy <= rsum2;
}

struct MergeIn {
  const float *gx, *gy, *gvx, *gvy, *px, *py, *rr;
  const unsigned char *vis, *ast;
  const int* midx;
  int n, nm;
  float dt;
};

// pass 0: per (row) hit counts; pass 1: ordered (row, col, frac) output.
__global__ __launch_bounds__(CB) void k_merge_rows(MergeIn I, int pass,
                                                   int* __restrict__ cnt,
                                                   const int* __restrict__ off,
                                                   int* __restrict__ orow,
                                                   int* __restrict__ ocol,
                                                   float* __restrict__ ofrac) {
  // one workgroup per (row, block of columns); columns in ascending order
  const int row = blockIdx.y;
  const int j = blockIdx.x * CB + threadIdx.x;
  __shared__ int s_n[CB];
  const int g = I.midx[row];
  bool hit = false;
  float frac = 0.0f;
  if (j < I.n && j != g && I.vis[j] != 0) {
    const float cx = I.gx[g], cy = I.gy[g];
    const float rs = I.rr[g] + I.rr[j];
    const float rs2 = rs * rs;
    float ta, tb;
    const bool ha = seg_hit(I.px[j], I.py[j], I.gx[j], I.gy[j], cx, cy, rs2, ta);
    const float q1x = I.gx[j] + I.gvx[j] * I.dt;
    const float q1y = I.gy[j] + I.gvy[j] * I.dt;
    const bool hb = seg_hit(I.gx[j], I.gy[j], q1x, q1y, cx, cy, rs2, tb);
    hit = ha || hb;
    frac = ha ? ta - 1.0f : tb;
  }
  // ordered position of this hit inside the workgroup (prefix count)
  s_n[threadIdx.x] = hit ? 1 : 0;
  __syncthreads();
  for (int o = 1; o < CB; o <<= 1) {
    const int v = threadIdx.x >= o ? s_n[threadIdx.x - o] : 0;
    __syncthreads();
    s_n[threadIdx.x] += v;
    __syncthreads();
  }
  const int slot = (int)blockIdx.y * gridDim.x + blockIdx.x;
  if (pass == 0) {
    if (threadIdx.x == CB - 1) cnt[slot] = s_n[CB - 1];
  } else if (hit) {
    const int w = off[slot] + s_n[threadIdx.x] - 1;
    orow[w] = row;
    ocol[w] = j;
    ofrac[w] = frac;
  }
}

__global__ __launch_bounds__(CB) void k_runaway(int n, const float* __restrict__ gx,
                                                const float* __restrict__ gy,
                                                const float* __restrict__ gvx,
                                                const float* __restrict__ gvy,
                                                const unsigned char* __restrict__ ast,
                                                const unsigned char* __restrict__ vis,
                                                float bx, float by, float two_gm,
                                                int* __restrict__ flag) {
  const int i = blockIdx.x * CB + threadIdx.x;
  if (i >= n) return;
  const float dx = gx[i] - bx, dy = gy[i] - by;
  float r = hypotf(dx, dy);
  r = r > 1e-6f ? r : 1e-6f;
  const float v2 = gvx[i] * gvx[i] + gvy[i] * gvy[i];
  const float vesc2 = two_gm / r;
  flag[i] = (v2 > 9.0f * vesc2 && ast[i] && vis[i] != 0) ? 1 : 0;
}

__global__ __launch_bounds__(CB) void k_compact_idx(int n, const int* __restrict__ flag,
                                                    const int* __restrict__ pos,
                                                    long long* __restrict__ out) {
  const int i = blockIdx.x * CB + threadIdx.x;
  if (i < n && flag[i]) out[pos[i]] = i;
}

// ---- bounce response (`bounce_deltas`, film_producer.py:916-1019) ----

// Greedy one-contact-per-body filter in list order: a sequential rule, run
// by one thread (the lists hold at most thousands of pairs).
__global__ void k_bounce_greedy(int np, const long long* __restrict__ hi,
                                const long long* __restrict__ hj,
                                unsigned char* __restrict__ used,
                                int* __restrict__ keep) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int t = 0; t < np; ++t) {
    const long long a = hi[t], b = hj[t];
    if (used[a] || used[b]) {
      keep[t] = 0;
      continue;
    }
    used[a] = 1;
    used[b] = 1;
    keep[t] = 1;
  }
}

__global__ __launch_bounds__(CB) void k_zero_u8(long long n, unsigned char* __restrict__ a) {
  const long long i = (long long)blockIdx.x * CB + threadIdx.x;
  if (i < n) a[i] = 0;
}

__global__ __launch_bounds__(CB) void k_zero_f64(long long n, double* __restrict__ a) {
  const long long i = (long long)blockIdx.x * CB + threadIdx.x;
  if (i < n) a[i] = 0.0;
}

// Impulse and overlap push of every kept pair; each body is in at most one
// kept pair, so the dense delta arrays are written without conflicts.
__global__ __launch_bounds__(CB) void k_bounce_pairs(
    int np, const long long* __restrict__ hi, const long long* __restrict__ hj,
    const int* __restrict__ keep, const float* __restrict__ x,
    const float* __restrict__ y, const float* __restrict__ vx,
    const float* __restrict__ vy, const double* __restrict__ rr,
    const double* __restrict__ mass, double dt, double e,
    double* __restrict__ ddx, double* __restrict__ ddy, double* __restrict__ ddvx,
    double* __restrict__ ddvy, int* __restrict__ act, unsigned char* __restrict__ aff) {
  const int t = blockIdx.x * CB + threadIdx.x;
  if (t >= np) return;
  act[t] = 0;
  if (!keep[t]) return;
  const long long a = hi[t], b = hj[t];
  const double dpx = (double)x[b] - (double)x[a];
  const double dpy = (double)y[b] - (double)y[a];
  const double dvx = (double)vx[b] - (double)vx[a];
  const double dvy = (double)vy[b] - (double)vy[a];
  const double dv2 = dvx * dvx + dvy * dvy;
  const double touch = rr[a] + rr[b];
  const double c = dpx * dpx + dpy * dpy - touch * touch;
  const double bq = 2.0 * (dpx * dvx + dpy * dvy);
  const double disc = bq * bq - 4.0 * dv2 * c;
  const double sq = sqrt(disc > 0.0 ? disc : 0.0);
  double tin = (-bq - sq) / (2.0 * (dv2 > 0.0 ? dv2 : 1.0));
  if (tin != tin) tin = 0.0;  // nan_to_num
  else if (isinf(tin)) tin = tin > 0 ? 1.7976931348623157e308 : -1.7976931348623157e308;
  const double tc = c < 0.0 ? 0.0 : (tin < -dt ? -dt : (tin > dt ? dt : tin));
  const bool valid = (c < 0.0) || (disc >= -(touch * touch) * 0.01);
  const double ncx = dpx + dvx * tc, ncy = dpy + dvy * tc;
  double dist = hypot(ncx, ncy);
  dist = dist > 1e-30 ? dist : 1.0;
  const double nx = ncx / dist, ny = ncy / dist;
  const double vrel = dvx * nx + dvy * ny;
  if (!(valid && vrel < 0.0)) return;
  act[t] = 1;
  const double mi = mass[a], mj = mass[b];
  double imp = -(1.0 + e) * vrel / (1.0 / mi + 1.0 / mj);
  if (imp != imp || isinf(imp)) imp = imp != imp ? 0.0 : (imp > 0 ? 1.7976931348623157e308 : -1.7976931348623157e308);
  ddvx[a] -= imp * nx / mi;
  ddvy[a] -= imp * ny / mi;
  ddvx[b] += imp * nx / mj;
  ddvy[b] += imp * ny / mj;
  const double pdist = hypot(dpx, dpy);
  const double overlap = touch - pdist;
  if (overlap > 0.0 && pdist > 1e-30) {
    const double pnx = dpx / pdist, pny = dpy / pdist;
    const double mt = mi + mj;
    const double shift = overlap * 1.1;
    ddx[a] -= pnx * shift * (mj / mt);
    ddy[a] -= pny * shift * (mj / mt);
    ddx[b] += pnx * shift * (mi / mt);
    ddy[b] += pny * shift * (mi / mt);
  }
  aff[a] = 1;
  aff[b] = 1;
}

__global__ __launch_bounds__(CB) void k_u8_to_i32(int n, const unsigned char* __restrict__ a,
                                                  int* __restrict__ b) {
  const int i = blockIdx.x * CB + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// ---- apply_merges on the device ----
// Body arrays of one kind of the nbody state (massive or asteroid): x, y,
// vx, vy, m (FP64), visible flags, radii.
constexpr int AM_THREADS = 256;  // workgroup of k_apply_merges

struct BodyArrays {
  double* f[5];
  unsigned char* vis;
  double* rad;
};

// One 32-byte event record of the reference's ring: "<dIIfIff" = time in
// days, a, b, new mass, kind, x, y.
__device__ void put_event(unsigned char* ev, int k, double t, unsigned int a,
                          unsigned int b, double m, unsigned int kind, double x, double y) {
  unsigned char* r = ev + 32 * (size_t)k;
  *reinterpret_cast<double*>(r) = t;
  *reinterpret_cast<unsigned int*>(r + 8) = a;
  *reinterpret_cast<unsigned int*>(r + 12) = b;
  *reinterpret_cast<float*>(r + 16) = (float)m;
  *reinterpret_cast<unsigned int*>(r + 20) = kind;
  *reinterpret_cast<float*>(r + 24) = (float)x;
  *reinterpret_cast<float*>(r + 28) = (float)y;
}

// The loop of apply_merges (film_producer.py:781-861), pairs in list order
// on one thread: a later pair sees the masses and flags the earlier ones
// left.  The runaway removals after it are independent of one another (each
// touches only its own body), so the workgroup takes them 256 at a time,
// their records placed in list order by a scan of the still visible ones
// (a body listed twice only at its first entry, as the loop would).
// Per pair (i, j, frac): a contact time the nbody kernel saw for j replaces
// frac when it lies within one raster (and is consumed either way); with
// shatter_check, two visible massive non-star bodies meeting at
// |v_j - v_i| >= 1.5 v_esc stop the loop (result[2..3], shatter_t), the
// runaway list included; otherwise the lighter body merges into the heavier
// (mass-weighted position and velocity, volume-conserving radius), which the
// state, the mirrors and a kind-0 event receive.  Then every still visible
// runaway body is removed (kind 2, a = 0xFFFFFFFF).
__global__ void k_apply_merges(int n, const float* __restrict__ smp, int npairs,
                               const int* __restrict__ pi, const int* __restrict__ pj,
                               const double* __restrict__ pfrac, int nrun,
                               const int* __restrict__ run, double* mass,
                               unsigned char* vis, double* real_r,
                               const unsigned char* __restrict__ is_ast,
                               const unsigned char* __restrict__ is_star_bh,
                               double* contact_t, int k_ev, int shatter_check, double G,
                               double t0_days, double raster_days,
                               const int* __restrict__ inv_kind,
                               const int* __restrict__ inv_pos, BodyArrays Ms,
                               BodyArrays As, unsigned char* events, int* result,
                               double* shatter_t) {
  __shared__ int s_state[3];       // records, merges, stopped
  __shared__ int s_wave[AM_THREADS / 64];
  const float* sx = smp;
  const float* sy = smp + n;
  const float* svx = smp + 2 * (size_t)n;
  const float* svy = smp + 3 * (size_t)n;
  auto arrays = [&](int idx) -> BodyArrays& { return inv_kind[idx] == -1 ? Ms : As; };
  auto deactivate = [&](int idx) {
    BodyArrays& B = arrays(idx);
    const int p = inv_pos[idx];
    B.f[4][p] = 0.0;
    B.vis[p] = 0;
  };
  if (threadIdx.x == 0) {
    int nev = 0, merged = 0;
    s_state[2] = 0;
    result[2] = -1;
    result[3] = -1;
    for (int q = 0; q < npairs; ++q) {
      const int mi = pi[q], j = pj[q];
      double frac = pfrac[q];
      const double tk = contact_t[j];
      if (tk == tk) {  // present (not NaN): consumed
        contact_t[j] = __builtin_nan("");
        const double fk = tk - (double)(k_ev + 1);
        if (-1.0 <= fk && fk <= 1.0) frac = fk;
      }
      if (shatter_check && !is_ast[mi] && !is_ast[j] && !is_star_bh[mi] && !is_star_bh[j] &&
          vis[mi] && vis[j] && mass[mi] > 0.0 && mass[j] > 0.0) {
        const float dvx = svx[j] - svx[mi], dvy = svy[j] - svy[mi];
        const double v_imp = (double)hypotf(dvx, dvy);
        const double touch = fmax(1e-12, real_r[mi] + real_r[j]);
        const double v_esc = sqrt(2.0 * G * (mass[mi] + mass[j]) / touch);
        if (v_imp >= 1.5 * v_esc) {
          result[2] = mi;
          result[3] = j;
          *shatter_t = t0_days + ((double)(k_ev + 1) + frac) * raster_days;
          result[0] = nev;
          result[1] = merged;
          s_state[2] = 1;
          break;
        }
      }
      if (!vis[mi] || !vis[j] || mass[mi] <= 0.0) continue;
      const int a = mass[mi] >= mass[j] ? mi : j;
      const int b = a == mi ? j : mi;
      const double ma = mass[a], mb = mass[b];
      const double mg = ma + mb;
      if (mg <= 0.0) continue;
      const double nx = ((double)sx[a] * ma + (double)sx[b] * mb) / mg;
      const double ny = ((double)sy[a] * ma + (double)sy[b] * mb) / mg;
      const double nvx = ((double)svx[a] * ma + (double)svx[b] * mb) / mg;
      const double nvy = ((double)svy[a] * ma + (double)svy[b] * mb) / mg;
      {
        BodyArrays& A = arrays(a);
        const int p = inv_pos[a];
        A.f[0][p] = nx;
        A.f[1][p] = ny;
        A.f[2][p] = nvx;
        A.f[3][p] = nvy;
        A.f[4][p] = mg;
      }
      deactivate(b);
      mass[a] = mg;
      mass[b] = 0.0;
      vis[b] = 0;
      const double ra = real_r[a], rb = real_r[b];
      real_r[a] = pow(ra * ra * ra + rb * rb * rb, 1.0 / 3.0);
      arrays(a).rad[inv_pos[a]] = real_r[a];
      put_event(events, nev++, t0_days + ((double)(k_ev + 1) + frac) * raster_days,
                (unsigned int)a, (unsigned int)b, mg, 0u, nx, ny);
      ++merged;
    }
    s_state[0] = nev;
    s_state[1] = merged;
  }
  __syncthreads();
  if (s_state[2]) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int nev = s_state[0];
  __shared__ int s_j[AM_THREADS];
  for (int base = 0; base < nrun; base += AM_THREADS) {
    const int q = base + (int)threadIdx.x;
    const int j = q < nrun ? run[q] : -1;
    s_j[threadIdx.x] = j;
    __syncthreads();
    // a body listed twice is removed once, at its first entry (earlier
    // chunks already cleared its flag; earlier entries of this chunk here)
    bool act = q < nrun && vis[j] != 0;
    for (int e = 0; act && e < (int)threadIdx.x; ++e) act = s_j[e] != j;
    const unsigned long long bal = __ballot(act);
    if (lane == 0) s_wave[wave] = __popcll(bal);
    __syncthreads();
    int before = __popcll(bal & ((1ull << lane) - 1ull)), total = 0;
    for (int w = 0; w < AM_THREADS / 64; ++w) {
      if (w < wave) before += s_wave[w];
      total += s_wave[w];
    }
    if (act) {
      deactivate(j);
      mass[j] = 0.0;
      vis[j] = 0;
      put_event(events, nev + before, t0_days + (double)(k_ev + 1) * raster_days,
                0xFFFFFFFFu, (unsigned int)j, 0.0, 2u, (double)sx[j], (double)sy[j]);
    }
    nev += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    result[0] = nev;
    result[1] = s_state[1];
  }
}

// Bounce deltas onto the state: body aff[q] (original index) gets
// x, y, vx, vy += d[0..3][aff[q]] in its integrator slot (the host's
// apply_deltas after bounce_deltas, film_producer.py:1021-1026); the
// affected bodies are distinct, so every slot sees one addition.
__global__ void k_add_body_deltas(int naff, const long long* __restrict__ aff,
                                  const double* __restrict__ d0,
                                  const double* __restrict__ d1,
                                  const double* __restrict__ d2,
                                  const double* __restrict__ d3,
                                  const int* __restrict__ inv_kind,
                                  const int* __restrict__ inv_pos, BodyArrays Ms,
                                  BodyArrays As) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= naff) return;
  const long long b = aff[q];
  BodyArrays& B = inv_kind[b] == -1 ? Ms : As;
  const int p = inv_pos[b];
  B.f[0][p] += d0[b];
  B.f[1][p] += d1[b];
  B.f[2][p] += d2[b];
  B.f[3][p] += d3[b];
}

}  // namespace

extern "C" {

size_t ssim_bounce_work_bytes(int n, int npairs) {
  const size_t nn = (size_t)(n > 0 ? n : 1), np = (size_t)(npairs > 0 ? npairs : 1);
  const size_t big = nn > np ? nn : np;
  return ssim::align256(nn) * 2 + ssim::align256(np * 4) * 3 +
         ssim::align256(nn * 4) * 2 + ssim::align256((big / 1024 + 2) * 4) + 256;
}

int ssim_bounce_deltas(int n, int npairs, const long long* hi, const long long* hj,
                       const float* x, const float* y, const float* vx,
                       const float* vy, const double* real_r, const double* mass,
                       double dt_y, double restitution, double* ddx, double* ddy,
                       double* ddvx, double* ddvy, long long* act_pairs,
                       long long* affected, long long* counts, void* work,
                       size_t work_bytes, void* stream) {
  SSIM_REQUIRE(n >= 0 && npairs >= 0, "bad sizes");
  SSIM_REQUIRE(counts && work && work_bytes >= ssim_bounce_work_bytes(n, npairs),
               "workspace too small");
  counts[0] = counts[1] = 0;
  if (npairs == 0 || n == 0) return SSIM_OK;
  SSIM_REQUIRE(hi && hj && x && y && vx && vy && real_r && mass && ddx && ddy &&
                   ddvx && ddvy && act_pairs && affected,
               "null pointer");
  hipStream_t st = ssim::as_stream(stream);
  const size_t nn = (size_t)n, np = (size_t)npairs;
  unsigned char* w = static_cast<unsigned char*>(work);
  size_t o = 0;
  unsigned char* used = w + o; o += ssim::align256(nn);
  unsigned char* aff = w + o; o += ssim::align256(nn);
  int* keep = reinterpret_cast<int*>(w + o); o += ssim::align256(np * 4);
  int* act = reinterpret_cast<int*>(w + o); o += ssim::align256(np * 4);
  int* apos = reinterpret_cast<int*>(w + o); o += ssim::align256(np * 4);
  int* affi = reinterpret_cast<int*>(w + o); o += ssim::align256(nn * 4);
  int* affp = reinterpret_cast<int*>(w + o); o += ssim::align256(nn * 4);
  int* totals = reinterpret_cast<int*>(w + o);
  int* sums = reinterpret_cast<int*>(w + work_bytes - 256);
  const dim3 gn(ssim::ceil_div(n, CB)), gp(ssim::ceil_div(npairs, CB));
  hipLaunchKernelGGL(k_zero_u8, gn, dim3(CB), 0, st, (long long)n, used);
  hipLaunchKernelGGL(k_zero_u8, gn, dim3(CB), 0, st, (long long)n, aff);
  for (double* d : {ddx, ddy, ddvx, ddvy})
    hipLaunchKernelGGL(k_zero_f64, gn, dim3(CB), 0, st, (long long)n, d);
  hipLaunchKernelGGL(k_bounce_greedy, dim3(1), dim3(64), 0, st, npairs, hi, hj,
                     used, keep);
  hipLaunchKernelGGL(k_bounce_pairs, gp, dim3(CB), 0, st, npairs, hi, hj, keep,
                     x, y, vx, vy, real_r, mass, dt_y, restitution, ddx, ddy,
                     ddvx, ddvy, act, aff);
  SSIM_LAUNCHED("k_bounce_pairs");
  // active pairs in list order, affected bodies ascending
  SSIM_TRY(scan(npairs, act, apos, totals, sums, st));
  int na = 0;
  SSIM_HIP(hipMemcpyAsync(&na, sums, sizeof(int), hipMemcpyDeviceToHost, st));
  hipLaunchKernelGGL(k_compact_idx, gp, dim3(CB), 0, st, npairs, act, apos,
                     act_pairs);
  hipLaunchKernelGGL(k_u8_to_i32, gn, dim3(CB), 0, st, n, aff, affi);
  SSIM_TRY(scan(n, affi, affp, totals, sums + 1, st));
  int nf = 0;
  SSIM_HIP(hipMemcpyAsync(&nf, sums + 1, sizeof(int), hipMemcpyDeviceToHost, st));
  hipLaunchKernelGGL(k_compact_idx, gn, dim3(CB), 0, st, n, affi, affp, affected);
  SSIM_LAUNCHED("k_compact_idx");
  SSIM_HIP(hipStreamSynchronize(st));
  counts[0] = na;
  counts[1] = nf;
  return SSIM_OK;
}

size_t ssim_merge_work_bytes(int n, int nm) {
  const size_t slots = (size_t)(nm > 0 ? nm : 1) * ssim::ceil_div(n > 0 ? n : 1, CB);
  const size_t nn = (size_t)(n > 0 ? n : 1);
  return 2 * ssim::align256(slots * 4) + 2 * ssim::align256(nn * 4) +
         ssim::align256(((slots > nn ? slots : nn) / 1024 + 2) * 4) + 256;
}

int ssim_merge_scan(int n, const float* gx, const float* gy, const float* gvx,
                    const float* gvy, const float* px, const float* py,
                    const float* rr, const unsigned char* vis,
                    const unsigned char* is_ast, int nm, const int* m_alive,
                    float dt, int do_runaway, float bx, float by, float two_gm,
                    long long capacity, int* out_row, int* out_col,
                    float* out_frac, long long runaway_capacity,
                    long long* runaway, long long* counts, void* work,
                    size_t work_bytes, void* stream) {
  SSIM_REQUIRE(n >= 0 && nm >= 0 && nm <= 64, "bad sizes (n=%d nm=%d)", n, nm);
  SSIM_REQUIRE(counts && work && work_bytes >= ssim_merge_work_bytes(n, nm),
               "workspace too small");
  SSIM_REQUIRE(n == 0 || (gx && gy && gvx && gvy && px && py && rr && vis && is_ast),
               "null input");
  counts[0] = counts[1] = 0;
  if (n == 0) return SSIM_OK;
  hipStream_t st = ssim::as_stream(stream);
  const int cols = ssim::ceil_div(n, CB);
  const size_t slots = (size_t)(nm > 0 ? nm : 1) * cols;
  const size_t nn = (size_t)n;
  unsigned char* w = static_cast<unsigned char*>(work);
  int* cnt = reinterpret_cast<int*>(w);
  int* off = reinterpret_cast<int*>(w + ssim::align256(slots * 4));
  int* flag = reinterpret_cast<int*>(w + 2 * ssim::align256(slots * 4));
  int* pos = reinterpret_cast<int*>(w + 2 * ssim::align256(slots * 4) + ssim::align256(nn * 4));
  int* totals = reinterpret_cast<int*>(w + 2 * ssim::align256(slots * 4) +
                                       2 * ssim::align256(nn * 4));
  int* sums = reinterpret_cast<int*>(w + work_bytes - 256);
  if (nm > 0) {
    SSIM_REQUIRE(m_alive, "null massive index list");
    MergeIn I{gx, gy, gvx, gvy, px, py, rr, vis, is_ast, m_alive, n, nm, dt};
    hipLaunchKernelGGL(k_merge_rows, dim3(cols, nm), dim3(CB), 0, st, I, 0, cnt,
                       nullptr, nullptr, nullptr, nullptr);
    SSIM_TRY(scan((int)slots, cnt, off, totals, sums, st));
    int nh = 0;
    SSIM_HIP(hipMemcpyAsync(&nh, sums, sizeof(int), hipMemcpyDeviceToHost, st));
    SSIM_HIP(hipStreamSynchronize(st));
    counts[0] = nh;
    if (nh > 0 && nh <= capacity && out_row && out_col && out_frac)
      hipLaunchKernelGGL(k_merge_rows, dim3(cols, nm), dim3(CB), 0, st, I, 1,
                         cnt, off, out_row, out_col, out_frac);
    SSIM_LAUNCHED("k_merge_rows");
  }
  if (do_runaway) {
    hipLaunchKernelGGL(k_runaway, dim3(cols), dim3(CB), 0, st, n, gx, gy, gvx,
                       gvy, is_ast, vis, bx, by, two_gm, flag);
    SSIM_TRY(scan(n, flag, pos, totals, sums, st));
    int nr = 0;
    SSIM_HIP(hipMemcpyAsync(&nr, sums, sizeof(int), hipMemcpyDeviceToHost, st));
    SSIM_HIP(hipStreamSynchronize(st));
    counts[1] = nr;
    if (nr > 0 && nr <= runaway_capacity && runaway)
      hipLaunchKernelGGL(k_compact_idx, dim3(cols), dim3(CB), 0, st, n, flag, pos,
                         runaway);
    SSIM_LAUNCHED("k_runaway");
  }
  SSIM_HIP(hipStreamSynchronize(st));
  return SSIM_OK;
}

size_t ssim_collide_work_bytes(int n, long long f32_hit_capacity) {
  return layout(n, f32_hit_capacity).total;
}

size_t ssim_speed_percentile_work_bytes(int n) {
  const size_t nn = (size_t)(n > 0 ? n : 1);
  return 4 * ssim::align256(nn * 4) + ssim::align256((nn / 1024 + 2) * 4) +
         ssim::align256(64) + ssim::radix_sort_work_bytes((int)nn);
}

int ssim_speed_percentile(int n, const float* gvx, const float* gvy,
                          const unsigned char* is_ast, const unsigned char* vis, float q,
                          float* out, void* work, size_t work_bytes, void* stream) {
  SSIM_REQUIRE(n >= 1, "n must be >= 1");
  SSIM_REQUIRE(q >= 0.0f && q <= 100.0f, "q must be in [0, 100], got %g", q);
  SSIM_REQUIRE(gvx && gvy && is_ast && vis && out, "null pointer");
  SSIM_REQUIRE(work && work_bytes >= ssim_speed_percentile_work_bytes(n),
               "workspace too small");
  hipStream_t st = ssim::as_stream(stream);
  const size_t nn = (size_t)n;
  unsigned char* w = static_cast<unsigned char*>(work);
  int* flag = reinterpret_cast<int*>(w);
  w += ssim::align256(nn * 4);
  int* pos = reinterpret_cast<int*>(w);
  w += ssim::align256(nn * 4);
  int* bits = reinterpret_cast<int*>(w);
  w += ssim::align256(nn * 4);
  int* idx = reinterpret_cast<int*>(w);
  w += ssim::align256(nn * 4);
  int* totals = reinterpret_cast<int*>(w);
  w += ssim::align256((nn / 1024 + 2) * 4);
  int* sums = reinterpret_cast<int*>(w);
  w += ssim::align256(64);
  const dim3 g(ssim::ceil_div(n, CB));
  hipLaunchKernelGGL(k_flag_alive, g, dim3(CB), 0, st, n, is_ast, vis, flag);
  SSIM_TRY(scan(n, flag, pos, totals, sums, st));
  hipLaunchKernelGGL(k_alive_speed, g, dim3(CB), 0, st, n, gvx, gvy, is_ast, vis, pos, bits,
                     idx);
  SSIM_LAUNCHED("k_alive_speed");
  // sort the m alive speeds (m on the device: sort all n slots with the
  // unused ones holding the largest key, so the first m are the alive ones)
  hipLaunchKernelGGL(k_pad_bits, g, dim3(CB), 0, st, n, sums, bits, idx);
  SSIM_TRY(ssim::radix_sort_pairs(n, 31, bits, idx, w, ssim::radix_sort_work_bytes(n), st));
  const float qf = q / 100.0f;
  hipLaunchKernelGGL(k_percentile_lerp, dim3(1), dim3(64), 0, st, bits, sums, qf, out);
  SSIM_LAUNCHED("k_percentile_lerp");
  return SSIM_OK;
}

int ssim_collide_batch(long long candidates) {
  SSIM_REQUIRE(candidates >= CPT && candidates <= CBATCH && candidates % CPT == 0,
               "collision batch must be a multiple of %d in [%d, %lld], got %lld", CPT, CPT,
               CBATCH, candidates);
  g_batch = candidates;
  return SSIM_OK;
}

int ssim_collide_hits(int n, const float* gx, const float* gy, const float* gvx,
                      const float* gvy, const float* rr,
                      const unsigned char* is_ast, const unsigned char* vis,
                      float lo, float hi, float halo, double dt_y, double h,
                      long long f32_hit_capacity, long long pair_capacity,
                      long long* out_i, long long* out_j, long long* counts,
                      void* work, size_t work_bytes, void* stream) {
  SSIM_REQUIRE(n >= 0, "n must be >= 0");
  SSIM_REQUIRE(h > 0.0, "cell size h must be > 0");
  SSIM_REQUIRE(counts, "counts must not be null");
  SSIM_REQUIRE(n == 0 || (gx && gy && gvx && gvy && rr && is_ast && vis),
               "null input");
  SSIM_REQUIRE(work && work_bytes >= ssim_collide_work_bytes(n, f32_hit_capacity),
               "workspace too small");
  counts[0] = counts[1] = counts[2] = 0;
  if (n < 2) return SSIM_OK;
  hipStream_t st = ssim::as_stream(stream);
  const Layout L = layout(n, f32_hit_capacity);
  unsigned char* w = static_cast<unsigned char*>(work);
  int* flag = reinterpret_cast<int*>(w + L.flag);
  int* pos = reinterpret_cast<int*>(w + L.pos);
  int* ai = reinterpret_cast<int*>(w + L.ai);
  long long* ix = reinterpret_cast<long long*>(w + L.ix);
  long long* iy = reinterpret_cast<long long*>(w + L.iy);
  unsigned long long* ukey = reinterpret_cast<unsigned long long*>(w + L.ukey);
  int* rkey = reinterpret_cast<int*>(w + L.rkey);
  int* srow = reinterpret_cast<int*>(w + L.srow);
  unsigned long long* skey = reinterpret_cast<unsigned long long*>(w + L.skey);
  float4* sb = reinterpret_cast<float4*>(w + L.sb);
  float* srr = reinterpret_cast<float*>(w + L.srr);
  int* sidx = reinterpret_cast<int*>(w + L.sidx);
  float4* ra4 = reinterpret_cast<float4*>(w + L.ra4);
  float* rar = reinterpret_cast<float*>(w + L.rar);
  int* sstart = reinterpret_cast<int*>(w + L.sstart);
  int* slen = reinterpret_cast<int*>(w + L.slen);
  int* soff = reinterpret_cast<int*>(w + L.soff);
  unsigned long long* bits = reinterpret_cast<unsigned long long*>(w + L.bits);
  int* wcnt = reinterpret_cast<int*>(w + L.wcnt);
  int* wpos = reinterpret_cast<int*>(w + L.wpos);
  unsigned long long* csum = reinterpret_cast<unsigned long long*>(w + L.sums);
  void* swork = w + L.scan;
  void* radix = w + L.radix;
  const size_t radix_bytes = ssim::radix_sort_work_bytes(n);
  int* ha = reinterpret_cast<int*>(w + L.ha);
  int* hb = reinterpret_cast<int*>(w + L.hb);
  int* hflag = reinterpret_cast<int*>(w + L.hflag);
  int* hpos = reinterpret_cast<int*>(w + L.hpos);
  auto read_int = [&](const int* at, int* v) -> int {
    SSIM_HIP(hipMemcpyAsync(v, at, sizeof(int), hipMemcpyDeviceToHost, st));
    SSIM_HIP(hipStreamSynchronize(st));
    return SSIM_OK;
  };

  // alive rows in ascending body index
  hipLaunchKernelGGL(k_alive_flags, dim3(ssim::ceil_div(n, CB)), dim3(CB), 0, st,
                     n, gx, is_ast, vis, lo, hi + halo, flag);
  SSIM_LAUNCHED("k_alive_flags");
  SSIM_TRY(ssim::scan_exclusive(n, flag, pos, swork, st));
  int na = 0;
  SSIM_TRY(read_int(pos + n, &na));
  if (na < 2) return SSIM_OK;
  const dim3 ga(ssim::ceil_div(na, CB));
  hipLaunchKernelGGL(k_compact_alive, dim3(ssim::ceil_div(n, CB)), dim3(CB), 0,
                     st, n, flag, pos, ai);
  // keys, then a stable LSD radix sort of the rows by the 64-bit key: bits
  // 0..30, 31..61 and 62..63
  hipLaunchKernelGGL(k_keys, ga, dim3(CB), 0, st, na, ai, gx, gy, h, ix, iy, ukey, rkey,
                     srow);
  SSIM_LAUNCHED("k_keys");
  SSIM_TRY(ssim::radix_sort_pairs(na, 31, rkey, srow, radix, radix_bytes, st));
  hipLaunchKernelGGL(k_key_slice, ga, dim3(CB), 0, st, na, 31, 0x7fffffffu, ukey, srow,
                     rkey);
  SSIM_TRY(ssim::radix_sort_pairs(na, 31, rkey, srow, radix, radix_bytes, st));
  hipLaunchKernelGGL(k_key_slice, ga, dim3(CB), 0, st, na, 62, 0x3u, ukey, srow, rkey);
  SSIM_TRY(ssim::radix_sort_pairs(na, 2, rkey, srow, radix, radix_bytes, st));
  hipLaunchKernelGGL(k_sorted_pack, ga, dim3(CB), 0, st, na, ai, srow, ukey, gx, gy, gvx,
                     gvy, rr, skey, sb, srr, sidx, ra4, rar);
  SSIM_LAUNCHED("k_sorted_pack");
  // the candidate range of every (offset, row) and their offsets in the
  // flattened list
  const int nseg = 5 * na;
  hipLaunchKernelGGL(k_segments, dim3(ssim::ceil_div(nseg, CB)), dim3(CB), 0, st, na, ix,
                     iy, skey, sstart, slen);
  SSIM_LAUNCHED("k_segments");
  hipLaunchKernelGGL(k_zero_u64_1, dim3(1), dim3(64), 0, st, csum);
  hipLaunchKernelGGL(k_sum_i32_u64, dim3(ssim::ceil_div(nseg, CB * 16)), dim3(CB), 0, st,
                     nseg, slen, csum);
  SSIM_LAUNCHED("k_sum_i32_u64");
  unsigned long long ncand = 0;
  SSIM_HIP(hipMemcpyAsync(&ncand, csum, sizeof(ncand), hipMemcpyDeviceToHost, st));
  SSIM_HIP(hipStreamSynchronize(st));
  counts[0] = (long long)ncand;
  SSIM_REQUIRE(ncand < (unsigned long long)INT_MAX,
               "collision search: %llu candidate pairs exceed the 2^31 list limit "
               "(cell size too large for the cloud)", ncand);
  SSIM_TRY(ssim::scan_exclusive(nseg, slen, soff, swork, st));  // soff[nseg] = total
  // the prefilter once per candidate, in batches
  Pf P{gx, gy, gvx, gvy, rr, (float)dt_y, 1e-6f};
  long long nh = 0;
  bool room = true;
  const long long batch = g_batch;
  for (long long base = 0; base < (long long)ncand; base += batch) {
    const long long nb = std::min<long long>(batch, (long long)ncand - base);
    long long cpt = (nb + 64 * WAVES_MIN - 1) / (64 * WAVES_MIN);
    cpt = cpt < 1 ? 1 : (cpt > CPT ? CPT : cpt);
    const int nwave = (int)((nb + 64 * cpt - 1) / (64 * cpt));
    const dim3 gb(ssim::ceil_div((long long)nwave * 64, CB));
    hipLaunchKernelGGL(k_cand_waves, gb, dim3(CB), 0, st, na, base, batch, (long long)ncand,
                       soff, sstart, ai, ra4, rar, sb, srr, sidx, P.dt, P.tol, 0, (int)cpt,
                       bits, wcnt, nullptr, 0, nullptr, nullptr);
    SSIM_LAUNCHED("k_cand_waves");
    SSIM_TRY(ssim::scan_exclusive(nwave, wcnt, wpos, swork, st));
    int nbh = 0;
    SSIM_TRY(read_int(wpos + nwave, &nbh));
    if (room && nh + nbh <= f32_hit_capacity) {
      hipLaunchKernelGGL(k_cand_waves, gb, dim3(CB), 0, st, na, base, batch,
                         (long long)ncand, soff, sstart, ai, ra4, rar, sb, srr, sidx, P.dt,
                         P.tol, 1, (int)cpt, bits, wcnt, wpos, (int)nh, ha, hb);
      SSIM_LAUNCHED("k_cand_waves");
    } else {
      room = false;
    }
    nh += nbh;
  }
  counts[1] = nh;
  if (nh == 0) return SSIM_OK;
  if (!room) {  // caller retries with a larger workspace
    counts[2] = -1;
    return SSIM_OK;
  }
  const int nhi = (int)nh;
  hipLaunchKernelGGL(k_recheck, dim3(ssim::ceil_div(nhi, CB)), dim3(CB), 0, st, nhi,
                     ha, hb, P, dt_y, hi, hflag);
  SSIM_TRY(ssim::scan_exclusive(nhi, hflag, hpos, swork, st));
  hipLaunchKernelGGL(k_compact_pairs, dim3(ssim::ceil_div(nhi, CB)), dim3(CB), 0,
                     st, nhi, hflag, hpos, ha, hb, out_i, out_j,
                     (out_i && out_j) ? pair_capacity : 0);
  SSIM_LAUNCHED("k_compact_pairs");
  int np = 0;
  SSIM_TRY(read_int(hpos + nhi, &np));
  counts[2] = np;
  return SSIM_OK;
}

int ssim_add_body_deltas(int naff, const long long* aff, const double* ddx,
                         const double* ddy, const double* ddvx, const double* ddvy,
                         const int* inv_kind, const int* inv_pos, void* const* massive,
                         void* const* asteroids, void* stream) {
  SSIM_REQUIRE(naff >= 0, "naff must be >= 0");
  if (naff == 0) return SSIM_OK;
  SSIM_REQUIRE(aff && ddx && ddy && ddvx && ddvy && inv_kind && inv_pos && massive &&
                   asteroids,
               "null pointer");
  BodyArrays M{}, A{};
  for (int f = 0; f < 5; ++f) {
    M.f[f] = static_cast<double*>(massive[f]);
    A.f[f] = static_cast<double*>(asteroids[f]);
  }
  hipLaunchKernelGGL(k_add_body_deltas, dim3(ssim::ceil_div(naff, 256)), dim3(256), 0,
                     ssim::as_stream(stream), naff, aff, ddx, ddy, ddvx, ddvy, inv_kind,
                     inv_pos, M, A);
  SSIM_LAUNCHED("k_add_body_deltas");
  return SSIM_OK;
}

int ssim_apply_merges(int n, const float* sample, int npairs, const int* pair_i,
                      const int* pair_j, const double* pair_frac, int nrun,
                      const int* runaway, double* mass, unsigned char* vis,
                      double* real_r, const unsigned char* is_ast,
                      const unsigned char* is_star_bh, double* contact_t, int k_ev,
                      int shatter_check, double G, double t0_days, double raster_days,
                      const int* inv_kind, const int* inv_pos, void* const* massive,
                      void* const* asteroids, unsigned char* events, int* result,
                      double* shatter_t, void* stream) {
  SSIM_REQUIRE(n >= 1 && npairs >= 0 && nrun >= 0, "bad sizes");
  SSIM_REQUIRE(sample && mass && vis && real_r && is_ast && is_star_bh && contact_t &&
                   inv_kind && inv_pos && massive && asteroids && events && result &&
                   shatter_t && (npairs == 0 || (pair_i && pair_j && pair_frac)) &&
                   (nrun == 0 || runaway),
               "null pointer");
  BodyArrays M{}, A{};
  for (int f = 0; f < 5; ++f) {
    M.f[f] = static_cast<double*>(massive[f]);
    A.f[f] = static_cast<double*>(asteroids[f]);
  }
  M.vis = static_cast<unsigned char*>(massive[5]);
  M.rad = static_cast<double*>(massive[6]);
  A.vis = static_cast<unsigned char*>(asteroids[5]);
  A.rad = static_cast<double*>(asteroids[6]);
  hipStream_t st = ssim::as_stream(stream);
  hipLaunchKernelGGL(k_apply_merges, dim3(1), dim3(AM_THREADS), 0, st, n, sample, npairs, pair_i,
                     pair_j, pair_frac, nrun, runaway, mass, vis, real_r, is_ast,
                     is_star_bh, contact_t, k_ev, shatter_check, G, t0_days, raster_days,
                     inv_kind, inv_pos, M, A, events, result, shatter_t);
  SSIM_LAUNCHED("k_apply_merges");
  return SSIM_OK;
}

}  // extern "C"
