// Film ring slots written on the device from the raster integrator's
// outputs: positions in original body order, the count and sorted original
// indices of the asteroids with a complete fine track, and their tracks
// body-major (film_producer.py:76-123 `slot_bytes` / `schreibe_slot`, fed
// by nbody_kernel.py:836-859 and 894-934).  One D2H copy of k slots then
// replaces the host scatter, argsort and transpose per sample.
//
// Slot layout (little-endian, every field 4-byte aligned):
//   x[n] f32 | y[n] f32 | u32 nh | u32 idx[sub_max] | f32 track[sub_max][m_sub][2]
// with nh = min(#complete tracks, sub_max); idx entries and tracks past nh
// are left as they were (the reference leaves them too).
//
// The asteroid slots i of the integrator are ascending in original index
// (a_idx[i] strictly increasing), so "sorted by original index, keep the
// lowest sub_max" is a stream compaction in slot order: per block counts,
// one scan per frame, then block-local ranks.
#include "ssim_internal.h"

namespace {

constexpr int RING_BLOCK = 256;
constexpr int RING_PER_THREAD = 4;
constexpr int RING_TILE = RING_BLOCK * RING_PER_THREAD;

// flags of the thread's four consecutive slots (bit j = slot 4t + j complete)
__device__ inline unsigned ring_flags(const unsigned char* __restrict__ subn_f,
                                      int na, int i0, int m_sub) {
  unsigned f = 0;
#pragma unroll
  for (int j = 0; j < RING_PER_THREAD; ++j) {
    const int i = i0 + j;
    if (i < na && subn_f[i] == (unsigned char)m_sub) f |= 1u << j;
  }
  return f;
}

// Inclusive scan of one int per thread over the block (Hillis-Steele in LDS).
__device__ inline int block_scan_incl(int v, int* lds) {
  const int t = threadIdx.x;
  lds[t] = v;
  __syncthreads();
  for (int d = 1; d < RING_BLOCK; d <<= 1) {
    const int add = t >= d ? lds[t - d] : 0;
    __syncthreads();
    lds[t] += add;
    __syncthreads();
  }
  return lds[t];
}

__global__ __launch_bounds__(RING_BLOCK) void k_ring_count(
    int na, int m_sub, const unsigned char* __restrict__ subn, int* __restrict__ bcount) {
  __shared__ int lds[RING_BLOCK];
  const int f = blockIdx.y, nblk = gridDim.x;
  const int i0 = blockIdx.x * RING_TILE + threadIdx.x * RING_PER_THREAD;
  const unsigned fl = ring_flags(subn + (long long)f * na, na, i0, m_sub);
  const int tot = block_scan_incl(__popc(fl), lds);
  if (threadIdx.x == RING_BLOCK - 1) bcount[(long long)f * nblk + blockIdx.x] = tot;
}

// One thread per frame: exclusive offsets of the block counts (in place),
// the frame's total, and the slot's count field.
__global__ void k_ring_scan(int k, int nblk, int n, int sub_max, int sub_on,
                            int* __restrict__ bcount, int* __restrict__ totals,
                            unsigned char* __restrict__ slots, long long slot_bytes) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= k) return;
  int run = 0;
  int* bc = bcount + (long long)f * nblk;
  if (sub_on)
    for (int b = 0; b < nblk; ++b) {
      const int c = bc[b];
      bc[b] = run;
      run += c;
    }
  totals[f] = run;
  unsigned* cnt = reinterpret_cast<unsigned*>(slots + f * slot_bytes + 8LL * n);
  *cnt = (unsigned)(run < sub_max ? run : sub_max);
}

__global__ __launch_bounds__(RING_BLOCK) void k_ring_write(
    int n, int na, int m, int m_sub, int sub_max, int sub_on,
    const float* __restrict__ snap, const float* __restrict__ snapM,
    const float* __restrict__ subpos, const unsigned char* __restrict__ subn,
    const unsigned* __restrict__ a_idx, const unsigned* __restrict__ m_idx,
    const int* __restrict__ boff, unsigned char* __restrict__ slots, long long slot_bytes) {
  __shared__ int lds[RING_BLOCK];
  const int f = blockIdx.y, nblk = gridDim.x;
  unsigned char* slot = slots + f * slot_bytes;
  float* xs = reinterpret_cast<float*>(slot);
  const int i0 = blockIdx.x * RING_TILE + threadIdx.x * RING_PER_THREAD;
  // positions of the asteroids (snap is [k][4][na]: x, y, vx, vy rows)
  const float* sx = snap + (long long)f * 4 * na;
#pragma unroll
  for (int j = 0; j < RING_PER_THREAD; ++j) {
    const int i = i0 + j;
    if (i < na) {
      const unsigned o = a_idx[i];
      xs[o] = sx[i];
      xs[n + o] = sx[na + i];
    }
  }
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < m; i += RING_BLOCK) {
      const float* sm = snapM + (long long)f * 4 * m;
      const unsigned o = m_idx[i];
      xs[o] = sm[i];
      xs[n + o] = sm[m + i];
    }
  if (!sub_on) return;
  const unsigned fl = ring_flags(subn + (long long)f * na, na, i0, m_sub);
  const int incl = block_scan_incl(__popc(fl), lds);
  int rank = boff[(long long)f * nblk + blockIdx.x] + incl - __popc(fl);
  if (rank >= sub_max || fl == 0) return;
  unsigned* idxf = reinterpret_cast<unsigned*>(slot + 8LL * n + 4);
  float* track = reinterpret_cast<float*>(slot + 8LL * n + 4 + 4LL * sub_max);
#pragma unroll
  for (int j = 0; j < RING_PER_THREAD; ++j) {
    if (!((fl >> j) & 1u) || rank >= sub_max) continue;
    const int i = i0 + j;
    idxf[rank] = a_idx[i];
    float* tr = track + (long long)rank * m_sub * 2;
    for (int s = 0; s < m_sub; ++s) {          // subpos is [k][m_sub][2][na]
      const long long b = (((long long)f * m_sub + s) * 2) * na + i;
      tr[2 * s] = subpos[b];
      tr[2 * s + 1] = subpos[b + na];
    }
    ++rank;
  }
}

// rows[f][c][idx[i]] = snap[f][c][i] for the asteroid (a_idx) and massive
// (m_idx) slots: the (k, 4n) [x|y|vx|vy] rows of step_batch in original
// order.  blockIdx.y = frame * 4 + component; lanes walk the slots.
__global__ __launch_bounds__(RING_BLOCK) void k_frame_rows(
    int n, int na, int m, const float* __restrict__ snap,
    const float* __restrict__ snapM, const unsigned* __restrict__ a_idx,
    const unsigned* __restrict__ m_idx, float* __restrict__ rows) {
  const int fc = blockIdx.y;                   // frame * 4 + component
  float* row = rows + (long long)fc * n;
  const int i = blockIdx.x * RING_BLOCK + threadIdx.x;
  if (i < na) row[a_idx[i]] = snap[(long long)fc * na + i];
  else if (i < na + m) row[m_idx[i - na]] = snapM[(long long)fc * m + (i - na)];
}

}  // namespace

extern "C" {

int ssim_frame_rows(int k, int n, int na, int m, const float* snap,
                    const float* snap_m, const unsigned* a_idx,
                    const unsigned* m_idx, float* rows, void* stream) {
  SSIM_REQUIRE(k >= 1 && n >= 0 && na >= 0 && m >= 0 && na + m == n,
               "bad sizes k=%d n=%d na=%d m=%d", k, n, na, m);
  SSIM_REQUIRE(rows, "null rows");
  SSIM_REQUIRE(na == 0 || (snap && a_idx), "null asteroid pointer");
  SSIM_REQUIRE(m == 0 || (snap_m && m_idx), "null massive pointer");
  if (n == 0) return SSIM_OK;
  hipLaunchKernelGGL(k_frame_rows, dim3(ssim::ceil_div(n, RING_BLOCK), 4 * k),
                     dim3(RING_BLOCK), 0, ssim::as_stream(stream), n, na, m, snap,
                     snap_m, a_idx, m_idx, rows);
  SSIM_LAUNCHED("k_frame_rows");
  return SSIM_OK;
}

long long ssim_ring_slot_bytes(int n, int m_sub, int sub_max) {
  return 8LL * n + 4 + (long long)sub_max * (4 + 8LL * m_sub);
}

size_t ssim_ring_work_bytes(int k, int na) {
  const long long nblk = ssim::ceil_div(na > 0 ? na : 1, RING_TILE);
  return (size_t)(4LL * k * nblk);
}

int ssim_ring_slots(int k, int n, int na, int m, int m_sub, int sub_max,
                    const float* snap, const float* snapM, const float* subpos,
                    const unsigned char* subn, const unsigned* a_idx,
                    const unsigned* m_idx, unsigned char* slots,
                    int* totals, void* work, size_t work_bytes, void* stream) {
  SSIM_REQUIRE(k >= 1 && n >= 0 && na >= 0 && m >= 0 && na + m == n,
               "bad sizes k=%d n=%d na=%d m=%d", k, n, na, m);
  SSIM_REQUIRE(m_sub >= 0 && sub_max >= 0, "negative m_sub / sub_max");
  SSIM_REQUIRE(slots && totals && work, "null pointer");
  SSIM_REQUIRE(work_bytes >= ssim_ring_work_bytes(k, na), "workspace too small");
  SSIM_REQUIRE(na == 0 || (snap && a_idx), "null asteroid pointer");
  SSIM_REQUIRE(m == 0 || (snapM && m_idx), "null massive pointer");
  const int sub_on = (m_sub > 0 && sub_max > 0 && na > 0) ? 1 : 0;
  SSIM_REQUIRE(!sub_on || (subpos && subn), "null sub-sample pointer");
  const long long sb = ssim_ring_slot_bytes(n, m_sub, sub_max);
  const int nblk = ssim::ceil_div(na > 0 ? na : 1, RING_TILE);
  hipStream_t st = ssim::as_stream(stream);
  int* bcount = static_cast<int*>(work);
  if (sub_on) {
    hipLaunchKernelGGL(k_ring_count, dim3(nblk, k), dim3(RING_BLOCK), 0, st, na,
                       m_sub, subn, bcount);
    SSIM_LAUNCHED("k_ring_count");
  }
  hipLaunchKernelGGL(k_ring_scan, dim3(ssim::ceil_div(k, 64)), dim3(64), 0, st, k,
                     nblk, n, sub_max, sub_on, bcount, totals, slots, sb);
  SSIM_LAUNCHED("k_ring_scan");
  hipLaunchKernelGGL(k_ring_write, dim3(nblk, k), dim3(RING_BLOCK), 0, st, n, na,
                     m, m_sub, sub_max, sub_on, snap, snapM, subpos, subn, a_idx,
                     m_idx, bcount, slots, sb);
  SSIM_LAUNCHED("k_ring_write");
  return SSIM_OK;
}

}  // extern "C"
