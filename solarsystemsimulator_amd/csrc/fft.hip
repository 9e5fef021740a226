// Hand-written real<->complex FFTs for the particle-mesh solver (gfx950).
//
// Replaces the cuFFT calls of the reference (`cp.fft.rfft2/irfft2`,
// pm_kernel.py:60,144-146) and provides the 3-D rfftn/irfftn the Poisson
// flavour needs.  Layout and normalisation follow NumPy: real array of
// shape dims (last axis contiguous), half spectrum of shape
// (dims[0], ..., dims[-1]/2 + 1) complex interleaved; the inverse divides by
// prod(dims).
//
// One kernel launch per axis ("pass").  A workgroup stages a tile of whole
// lines (B lines x L points, <= 4096 complex values = 64 KB in FP64) in LDS,
// runs all log2(L) butterfly stages there and writes the tile back, so each
// pass streams the array through HBM once.  Lines along a strided axis are
// grouped by consecutive inner index so every row of the tile is a
// contiguous segment in HBM (coalesced).  The butterflies are the Stockham
// autosort formulation (radix 4, one radix-2 stage when log2 L is odd):
//   stage (R, p): i in [0, L/R), k = i mod p,
//   x_r = a[i + r L/R] * w^(r k L/(p R)),  y = DFT_R(x),
//   a'[(i - k) R + k + q p] = y_q.
// All reads of a stage land in registers before a barrier, then all writes,
// so one LDS buffer suffices.  Twiddles w^j = exp(-2 pi i j / L) come from a
// table in HBM (L2-resident) filled by ssim_fft_prepare.
#include "ssim_internal.h"

namespace {

constexpr int FFT_THREADS = 256;
constexpr int FFT_TILE_BYTES = 32768;  // LDS tile of one workgroup
constexpr int FFT_MAXB = 64;           // lines per tile along a strided axis
constexpr int FFT_MAX_L = 8192;        // longest line
constexpr int FFT_LINE_L = 4096;       // longest line of the regular pass; longer
                                       // lines: one per workgroup (k_fft_pass<., ., 8192>)

// MODE_POISSON: a strided C2C pass along the first axis of a 3-D half
// spectrum that runs forward, multiplies by the periodic Poisson factor
// and runs inverse on the same tile (see poisson_lines)
enum { MODE_C2C = 0, MODE_R2C = 1, MODE_C2R = 2, MODE_POISSON = 3 };

template <typename F>
struct cpx {
  F re, im;
};

// complex values per tile: 4096 in FP32, 2048 in FP64 (32 KB of LDS plus
// the row padding); a strided pass of 256-point lines then moves rows of 16
// (FP32) or 8 (FP64) consecutive lines, 128 contiguous bytes
template <typename F>
constexpr int tile_elems() {
  return FFT_TILE_BYTES / (int)sizeof(cpx<F>);
}

// the largest tile a workgroup may hold: a single line can be longer than
// the tile (up to MAXL points: 64 KB in FP64 for the regular pass, 128 KB
// for the 8192-point one)
template <typename F, int MAXL = FFT_LINE_L>
constexpr int max_elems() {
  return tile_elems<F>() > MAXL ? tile_elems<F>() : MAXL;
}

template <typename F>
__device__ __forceinline__ cpx<F> cmul(cpx<F> a, cpx<F> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

template <typename F>
__device__ __forceinline__ cpx<F> cadd(cpx<F> a, cpx<F> b) {
  return {a.re + b.re, a.im + b.im};
}

template <typename F>
__device__ __forceinline__ cpx<F> csub(cpx<F> a, cpx<F> b) {
  return {a.re - b.re, a.im - b.im};
}

// multiply by -i (forward) or +i (inverse)
template <typename F>
__device__ __forceinline__ cpx<F> rot(cpx<F> a, bool inverse) {
  return inverse ? cpx<F>{-a.im, a.re} : cpx<F>{a.im, -a.re};
}

int ilog2(int v) {
  int l = 0;
  while ((1 << l) < v) ++l;
  return l;
}

// Tile geometry of one pass (complex-element units).  Only the first
// in_count points of a line are read (the rest are zeros) and only the first
// out_count are written, which prunes the zero padding and the cropped part
// of an isolated convolution.  Contiguous rows r map to offsets
// (r / R1) * s_hi + (r % R1) * s_outer, so a pass can read or write an
// n x n block of rows inside a larger array.
struct Pass {
  int L, log2L;
  int n_outer, n_inner;       // lines = n_outer * n_inner
  long long s_outer_in, s_outer_out, s_elem;
  int lines_per_tile;         // along the inner index (strided) or outer (contig)
  int contig;                 // 1: lines are contiguous rows (s_elem == 1)
  int in_count, out_count;    // points read / written per line
  long long R1, s_in_hi, s_out_hi;
  int tw_lds;                 // 1: the twiddle table is staged in LDS
  int out_es, out_c;          // C2R: element e goes to out[(off + e) * es + c]
  // MODE_POISSON: real grid n0 x n1 x n2 (the pass runs along n0, the
  // inner index is (ky, kz) over n1 x (n2 / 2 + 1)), cell h, G
  int pn0, pn1, pn2;
  double ph, pG;
};

__device__ __forceinline__ long long row_off(long long r, long long R1,
                                             long long s_hi, long long s_lo) {
  return (r / R1) * s_hi + (r % R1) * s_lo;
}

// ---- 256-point lines: four-step FFT with 16-point DFTs in registers ----
// x[n1 + 16 n2] -> Y[n1][k2] = DFT16 over n2 (thread n1), times
// W256^(n1 k2), then X[k2 + 16 k1] = DFT16 over n1 of Y[.][k2] (thread k2):
// two LDS exchanges instead of four radix-4 stages.  Y[n1][k2] sits at
// 16 n1 + (k2 ^ n1) so both exchanges are free of bank conflicts.

template <typename F>
__device__ __forceinline__ void dft4(cpx<F>& x0, cpx<F>& x1, cpx<F>& x2, cpx<F>& x3,
                                     bool inverse) {
  const cpx<F> s02 = cadd(x0, x2), d02 = csub(x0, x2);
  const cpx<F> s13 = cadd(x1, x3), d13 = rot(csub(x1, x3), inverse);
  x0 = cadd(s02, s13);
  x1 = cadd(d02, d13);
  x2 = csub(s02, s13);
  x3 = csub(d02, d13);
}

// in-place 16-point DFT (forward: exp(-2 pi i nk/16)), natural order out
template <typename F>
__device__ __forceinline__ void dft16(cpx<F> v[16], bool inverse) {
  // W16^j, j = 0..9, as (cos, -sin) of 2 pi j / 16
  constexpr double C1 = 0.92387953251128675613, S1 = 0.38268343236508977173;
  constexpr double C2 = 0.70710678118654752440;
  const F sg = inverse ? F(1) : F(-1);
  const cpx<F> w[10] = {{F(1), F(0)},          {F(C1), sg * F(S1)},
                        {F(C2), sg * F(C2)},   {F(S1), sg * F(C1)},
                        {F(0), sg},            {F(-S1), sg * F(C1)},
                        {F(-C2), sg * F(C2)},  {F(-C1), sg * F(S1)},
                        {F(-1), F(0)},         {F(-C1), -sg * F(S1)}};
#pragma unroll
  for (int b = 0; b < 4; ++b) dft4(v[b], v[b + 4], v[b + 8], v[b + 12], inverse);
  // v[b + 4 k1] = U_b[k1]; twiddle by W16^(b k1)
#pragma unroll
  for (int b = 1; b < 4; ++b)
#pragma unroll
    for (int k1 = 1; k1 < 4; ++k1) v[b + 4 * k1] = cmul(v[b + 4 * k1], w[b * k1]);
  cpx<F> o[16];
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    cpx<F> t0 = v[4 * k1], t1 = v[4 * k1 + 1], t2 = v[4 * k1 + 2], t3 = v[4 * k1 + 3];
    dft4(t0, t1, t2, t3, inverse);
    o[k1] = t0;
    o[k1 + 4] = t1;
    o[k1 + 8] = t2;
    o[k1 + 12] = t3;
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = o[k];
}

template <typename F>
__device__ void four_step_256(cpx<F>* s, int LS, int nl, const cpx<F>* tw,
                              bool inverse) {
  const int jobs = nl * 16;
  constexpr int MAXJ = 4096 / 16 / FFT_THREADS > 0 ? 4096 / 16 / FFT_THREADS : 1;
  cpx<F> v[MAXJ][16];
  // step 1: DFT16 over n2 of x[n1 + 16 n2], thread (l, n1)
#pragma unroll
  for (int q = 0; q < MAXJ; ++q) {
    const int job = threadIdx.x + q * FFT_THREADS;
    if (job < jobs) {
      const cpx<F>* a = s + (job >> 4) * LS;
      const int n1 = job & 15;
#pragma unroll
      for (int n2 = 0; n2 < 16; ++n2) v[q][n2] = a[n1 + 16 * n2];
      dft16(v[q], inverse);
#pragma unroll
      for (int k2 = 1; k2 < 16; ++k2) {
        cpx<F> w = tw[n1 * k2];
        if (inverse) w.im = -w.im;
        v[q][k2] = cmul(v[q][k2], w);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MAXJ; ++q) {
    const int job = threadIdx.x + q * FFT_THREADS;
    if (job < jobs) {
      cpx<F>* a = s + (job >> 4) * LS;
      const int n1 = job & 15;
#pragma unroll
      for (int k2 = 0; k2 < 16; ++k2) a[16 * n1 + (k2 ^ n1)] = v[q][k2];
    }
  }
  __syncthreads();
  // step 3: DFT16 over n1 of Y[n1][k2], thread (l, k2)
#pragma unroll
  for (int q = 0; q < MAXJ; ++q) {
    const int job = threadIdx.x + q * FFT_THREADS;
    if (job < jobs) {
      const cpx<F>* a = s + (job >> 4) * LS;
      const int k2 = job & 15;
#pragma unroll
      for (int n1 = 0; n1 < 16; ++n1) v[q][n1] = a[16 * n1 + (k2 ^ n1)];
      dft16(v[q], inverse);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MAXJ; ++q) {
    const int job = threadIdx.x + q * FFT_THREADS;
    if (job < jobs) {
      cpx<F>* a = s + (job >> 4) * LS;
      const int k2 = job & 15;
#pragma unroll
      for (int k1 = 0; k1 < 16; ++k1) a[k2 + 16 * k1] = v[q][k1];
    }
  }
  __syncthreads();
}

// in-place 8-point DFT, natural order out: 4-point DFTs of the even and the
// odd inputs, the odd ones twiddled by W8^k1, then one radix-2 step
template <typename F>
__device__ __forceinline__ void dft8(cpx<F> v[8], bool inverse) {
  constexpr double C2 = 0.70710678118654752440;
  const F sg = inverse ? F(1) : F(-1);
  dft4(v[0], v[2], v[4], v[6], inverse);
  dft4(v[1], v[3], v[5], v[7], inverse);
  v[3] = cmul(v[3], cpx<F>{F(C2), sg * F(C2)});
  v[5] = rot(v[5], inverse);
  v[7] = cmul(v[7], cpx<F>{F(-C2), sg * F(C2)});
  cpx<F> o[8];
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    o[k1] = cadd(v[2 * k1], v[2 * k1 + 1]);
    o[k1 + 4] = csub(v[2 * k1], v[2 * k1 + 1]);
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) v[k] = o[k];
}

// ---- 128-point lines: the four-step FFT with 128 = 16 x 8 ----
// x[n1 + 16 n2] -> Y[n1][k2] = DFT8 over n2 (thread n1), times
// W128^(n1 k2), then X[k2 + 8 k1] = DFT16 over n1 of Y[.][k2] (thread k2):
// one LDS exchange instead of a radix-2 and three radix-4 stages.
// Y[n1][k2] sits at 8 n1 + (k2 ^ (n1 & 7)).
template <typename F>
__device__ void four_step_128(cpx<F>* s, int LS, int nl, const cpx<F>* tw,
                              bool inverse) {
  constexpr int MAXJ1 = 4096 / 8 / FFT_THREADS, MAXJ3 = 4096 / 16 / FFT_THREADS;
  cpx<F> v[MAXJ1][8];
  const int jobs1 = nl * 16, jobs3 = nl * 8;
#pragma unroll
  for (int q = 0; q < MAXJ1; ++q) {
    const int job = threadIdx.x + q * FFT_THREADS;
    if (job < jobs1) {
      const cpx<F>* a = s + (job >> 4) * LS;
      const int n1 = job & 15;
#pragma unroll
      for (int n2 = 0; n2 < 8; ++n2) v[q][n2] = a[n1 + 16 * n2];
      dft8(v[q], inverse);
#pragma unroll
      for (int k2 = 1; k2 < 8; ++k2) {
        cpx<F> w = tw[n1 * k2];
        if (inverse) w.im = -w.im;
        v[q][k2] = cmul(v[q][k2], w);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MAXJ1; ++q) {
    const int job = threadIdx.x + q * FFT_THREADS;
    if (job < jobs1) {
      cpx<F>* a = s + (job >> 4) * LS;
      const int n1 = job & 15;
#pragma unroll
      for (int k2 = 0; k2 < 8; ++k2) a[8 * n1 + (k2 ^ (n1 & 7))] = v[q][k2];
    }
  }
  __syncthreads();
  cpx<F> u[MAXJ3][16];
#pragma unroll
  for (int q = 0; q < MAXJ3; ++q) {
    const int job = threadIdx.x + q * FFT_THREADS;
    if (job < jobs3) {
      const cpx<F>* a = s + (job >> 3) * LS;
      const int k2 = job & 7;
#pragma unroll
      for (int n1 = 0; n1 < 16; ++n1) u[q][n1] = a[8 * n1 + (k2 ^ (n1 & 7))];
      dft16(u[q], inverse);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MAXJ3; ++q) {
    const int job = threadIdx.x + q * FFT_THREADS;
    if (job < jobs3) {
      cpx<F>* a = s + (job >> 3) * LS;
      const int k2 = job & 7;
#pragma unroll
      for (int k1 = 0; k1 < 16; ++k1) a[k2 + 8 * k1] = u[q][k1];
    }
  }
  __syncthreads();
}

// The periodic Poisson factor on a tile of first-axis lines of the 3-D half
// spectrum, after their forward transform: element e of line l is mode
// (kx = e, ky, kz) with (ky, kz) the line's inner index; times
// -4 pi G / (k^2 h^3), k = 2 pi f / (n h) with f the signed frequency, and
// 0 for k = 0 — the operations of k_poisson (pm.hip), so the fused solve
// rounds exactly like the separate passes.
template <typename F>
__device__ void poisson_lines(cpx<F>* s, int LS, int nl, int L, const Pass& P, long long o0,
                              long long b0, int nvalid) {
  const int W = P.pn2 / 2 + 1;
  for (int t = threadIdx.x; t < nl * L; t += FFT_THREADS) {
    const int l = t % nl, e = t / nl;
    if (l >= nvalid) continue;
    const long long inner = b0 + l;
    const int b = (int)(inner / W), c = (int)(inner % W);
    const int f[3] = {e <= P.pn0 / 2 ? e : e - P.pn0, b <= P.pn1 / 2 ? b : b - P.pn1, c};
    const int ns[3] = {P.pn0, P.pn1, P.pn2};
    double k2 = 0.0;
    for (int q = 0; q < 3; ++q) {
      const double k = 2.0 * M_PI * (double)f[q] / ((double)ns[q] * P.ph);
      k2 += k * k;
    }
    const double vol = P.ph * P.ph * P.ph;
    const double g = (k2 > 0.0) ? (-4.0 * M_PI * P.pG) / (k2 * vol) : 0.0;
    cpx<F>& v = s[l * LS + e];
    v = {(F)(v.re * g), (F)(v.im * g)};
  }
  (void)o0;
}

// All lines of a tile transformed in LDS (forward or inverse): the
// four-step forms for 128 and 256 points, Stockham radix-4 (+ one radix-2
// stage) otherwise.
template <typename F, int MAXL = FFT_LINE_L>
__device__ void tile_fft(cpx<F>* s, int LS, int nl, int L, int log2L, const cpx<F>* twp,
                         int inverse) {
  const int E = nl * L;
  if (L == 256) {
    four_step_256(s, LS, nl, twp, inverse != 0);
  } else if (L == 128) {
    four_step_128(s, LS, nl, twp, inverse != 0);
  } else {
    // ---- Stockham stages over all lines of the tile ----
    int p = 1;
    if (log2L & 1) {  // one radix-2 stage (p == 1: k == 0, no twiddle)
      constexpr int NB = max_elems<F, MAXL>() / 2 / FFT_THREADS;
      const int T = L / 2;
      cpx<F> r0[NB], r1[NB];
  #pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
        const int u = threadIdx.x + nb * FFT_THREADS;
        if (u < E / 2) {
          const cpx<F>* a = s + (u / T) * LS;
          const int i = u % T;
          r0[nb] = a[i];
          r1[nb] = a[i + T];
        }
      }
      __syncthreads();
  #pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
        const int u = threadIdx.x + nb * FFT_THREADS;
        if (u < E / 2) {
          cpx<F>* a = s + (u / T) * LS;
          const int i = u % T;
          a[2 * i] = cadd(r0[nb], r1[nb]);
          a[2 * i + 1] = csub(r0[nb], r1[nb]);
        }
      }
      __syncthreads();
      p = 2;
    }
    for (; p < L; p *= 4) {
      constexpr int NB = max_elems<F, MAXL>() / 4 / FFT_THREADS;
      const int T = L / 4;
      const int twstep = L / (4 * p);
      cpx<F> r[NB][4];
  #pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
        const int u = threadIdx.x + nb * FFT_THREADS;
        if (u < E / 4) {
          const int i = u % T;
          const int k = i & (p - 1);
          const cpx<F>* a = s + (u / T) * LS;
  #pragma unroll
          for (int q = 0; q < 4; ++q) {
            cpx<F> x = a[i + q * T];
            if (q > 0 && k > 0) {
              cpx<F> w = twp[q * k * twstep];
              if (inverse) w.im = -w.im;
              x = cmul(x, w);
            }
            r[nb][q] = x;
          }
        }
      }
      __syncthreads();
  #pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
        const int u = threadIdx.x + nb * FFT_THREADS;
        if (u < E / 4) {
          const int i = u % T;
          const int k = i & (p - 1);
          cpx<F>* a = s + (u / T) * LS;
          const cpx<F> s02 = cadd(r[nb][0], r[nb][2]);
          const cpx<F> d02 = csub(r[nb][0], r[nb][2]);
          const cpx<F> s13 = cadd(r[nb][1], r[nb][3]);
          const cpx<F> d13 = rot(csub(r[nb][1], r[nb][3]), inverse != 0);
          const int j = (i - k) * 4 + k;
          a[j] = cadd(s02, s13);
          a[j + p] = cadd(d02, d13);
          a[j + 2 * p] = csub(s02, s13);
          a[j + 3 * p] = csub(d02, d13);
        }
      }
      __syncthreads();
    }
  }

}

template <typename F, int MODE, int MAXL = FFT_LINE_L>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_pass(
    const void* __restrict__ vin, void* __restrict__ vout, Pass P,
    const cpx<F>* __restrict__ tw, int inverse, F scale,
    const F* __restrict__ mul) {
  extern __shared__ unsigned char smem_raw[];
  cpx<F>* s = reinterpret_cast<cpx<F>*>(smem_raw);
  const int L = P.L;
  const int nl = P.lines_per_tile;
  const int E = nl * L;
  // lines sit L + 1 apart in LDS: the strided loads and stores, where
  // consecutive lanes take consecutive lines, then fall in different banks
  const int LS = nl > 1 ? L + 1 : L;

  // which lines does this workgroup own?
  long long o0, b0;
  int nvalid;
  if (P.contig) {
    o0 = (long long)blockIdx.x * nl;
    b0 = 0;
    nvalid = (int)min((long long)nl, (long long)P.n_outer - o0);
  } else {
    const int groups = (P.n_inner + nl - 1) / nl;
    o0 = blockIdx.x / groups;
    b0 = (long long)(blockIdx.x % groups) * nl;
    nvalid = (int)min((long long)nl, (long long)P.n_inner - b0);
  }

  // twiddles: a copy of the table after the tile when it fits in LDS
  cpx<F>* stw = s + nl * LS;
  if (P.tw_lds)
    for (int t = threadIdx.x; t < L; t += FFT_THREADS) stw[t] = tw[t];
  const cpx<F>* twp = P.tw_lds ? stw : tw;

  // ---- load the tile: s[l * LS + e] ----
  if (MODE == MODE_R2C) {
    const F* in = static_cast<const F*>(vin);
    for (int t = threadIdx.x; t < E; t += FFT_THREADS) {
      const int l = t / L, e = t % L;
      F v = 0;
      if (l < nvalid && e < P.in_count)
        v = in[row_off(o0 + l, P.R1, P.s_in_hi, P.s_outer_in) + e];
      s[l * LS + e] = {v, F(0)};
    }
  } else if (MODE == MODE_C2R) {
    // half spectrum of L/2+1 values -> Hermitian line of L values
    const cpx<F>* in = static_cast<const cpx<F>*>(vin);
    const int H = L / 2 + 1;
    for (int t = threadIdx.x; t < E; t += FFT_THREADS) {
      const int l = t / L, e = t % L;
      cpx<F> v = {F(0), F(0)};
      if (l < nvalid) {
        const cpx<F>* row = in + row_off(o0 + l, P.R1, P.s_in_hi, P.s_outer_in);
        if (e < H) {
          v = row[e];
        } else {
          const cpx<F> c = row[L - e];
          v = {c.re, -c.im};
        }
      }
      s[l * LS + e] = v;
    }
  } else {
    const cpx<F>* in = static_cast<const cpx<F>*>(vin);
    if (P.contig) {
      for (int t = threadIdx.x; t < E; t += FFT_THREADS) {
        const int l = t / L, e = t % L;
        s[l * LS + e] = (l < nvalid && e < P.in_count)
                   ? in[row_off(o0 + l, P.R1, P.s_in_hi, P.s_outer_in) + e]
                   : cpx<F>{0, 0};
      }
    } else {
      // consecutive threads walk the inner index: contiguous in HBM
      for (int t = threadIdx.x; t < E; t += FFT_THREADS) {
        const int l = t % nl, e = t / nl;
        cpx<F> v = {0, 0};
        if (l < nvalid && e < P.in_count) {
          const long long idx = o0 * P.s_outer_in + (b0 + l) + e * P.s_elem;
          v = in[idx];
          if (mul) {  // times i * mul (an imaginary spectrum)
            const F m = mul[idx];
            v = {-v.im * m, v.re * m};
          }
        }
        s[l * LS + e] = v;
      }
    }
  }
  __syncthreads();

  tile_fft<F, MAXL>(s, LS, nl, L, P.log2L, twp, inverse);
  if (MODE == MODE_POISSON) {
    poisson_lines(s, LS, nl, L, P, o0, b0, nvalid);
    __syncthreads();
    tile_fft<F, MAXL>(s, LS, nl, L, P.log2L, twp, 1);
  }

  // ---- store ----
  if (MODE == MODE_R2C) {
    cpx<F>* out = static_cast<cpx<F>*>(vout);
    const int H = L / 2 + 1;
    for (int t = threadIdx.x; t < nl * H; t += FFT_THREADS) {
      const int l = t / H, e = t % H;
      if (l < nvalid) {
        cpx<F> v = s[l * LS + e];
        out[row_off(o0 + l, P.R1, P.s_out_hi, P.s_outer_out) + e] = {v.re * scale,
                                                                     v.im * scale};
      }
    }
  } else if (MODE == MODE_C2R) {
    F* out = static_cast<F*>(vout);
    for (int t = threadIdx.x; t < E; t += FFT_THREADS) {
      const int l = t / L, e = t % L;
      if (l < nvalid && e < P.out_count)
        out[(row_off(o0 + l, P.R1, P.s_out_hi, P.s_outer_out) + e) * P.out_es + P.out_c] =
            s[l * LS + e].re * scale;
    }
  } else {
    cpx<F>* out = static_cast<cpx<F>*>(vout);
    if (P.contig) {
      for (int t = threadIdx.x; t < E; t += FFT_THREADS) {
        const int l = t / L, e = t % L;
        if (l < nvalid && e < P.out_count) {
          const cpx<F> v = s[l * LS + e];
          out[row_off(o0 + l, P.R1, P.s_out_hi, P.s_outer_out) + e] = {v.re * scale,
                                                                       v.im * scale};
        }
      }
    } else {
      for (int t = threadIdx.x; t < E; t += FFT_THREADS) {
        const int l = t % nl, e = t / nl;
        if (l < nvalid && e < P.out_count) {
          const cpx<F> v = s[l * LS + e];
          out[o0 * P.s_outer_out + (b0 + l) + e * P.s_elem] = {v.re * scale,
                                                               v.im * scale};
        }
      }
    }
  }
}

// ---- fused x axis of the isolated convolution (256-point lines) ----
// One workgroup owns NL consecutive x lines (consecutive inner index, so
// every row access is NL contiguous values).  Thread (l = t % NL, j = t / NL)
// reads x[j + 16 n2], n2 < 8, of line l straight from HBM (x >= n = 128 is
// the zero padding), and the forward four-step FFT leaves X[j + 16 k1] in
// its registers.  For each component c it multiplies by i kim_c, runs the
// inverse four-step from those registers (the DFT16 inputs of the inverse
// are exactly the values the thread holds) and stores x = j + 16 k1 < n of
// the result.  LDS only carries the two DFT16 exchanges; rows are padded to
// 258 so the exchanges of NL lines are free of bank conflicts.  In HBM this
// replaces four passes (forward x and three product + inverse x) by one
// read of the x < n half, three reads of the kernel spectra and three
// half-size writes.
constexpr int HUB_LS = 258;

template <typename F, int NL, int LS = HUB_LS>
__device__ __forceinline__ void hub_exchange(cpx<F> v[16], cpx<F>* s, int l, int j,
                                             const cpx<F>* tw, bool inverse) {
  // v = DFT16 outputs of index n1 = j; twiddle, exchange, DFT16 over n1
#pragma unroll
  for (int k2 = 1; k2 < 16; ++k2) {
    cpx<F> w = tw[j * k2];
    if (inverse) w.im = -w.im;
    v[k2] = cmul(v[k2], w);
  }
  __syncthreads();  // the previous exchange's reads are done
  cpx<F>* a = s + l * LS;
#pragma unroll
  for (int k2 = 0; k2 < 16; ++k2) a[16 * j + (k2 ^ j)] = v[k2];
  __syncthreads();
#pragma unroll
  for (int n1 = 0; n1 < 16; ++n1) v[n1] = a[16 * n1 + (j ^ n1)];
  dft16(v, inverse);
}

template <typename F, int NL>
__global__ __launch_bounds__(NL * 16) __attribute__((amdgpu_waves_per_eu(3))) void k_conv_x3(
    const cpx<F>* in, const F* __restrict__ kim0,  // in == out0: updated in place
    const F* __restrict__ kim1, const F* __restrict__ kim2, cpx<F>* out0,
    cpx<F>* out1, cpx<F>* out2, const cpx<F>* __restrict__ tw, int lines) {
  // 256-point lines of the padded 2n = 256 mesh: x < 128 read and written;
  // the host launches lines / NL workgroups (lines = 256 * 129 divides)
  constexpr int L = 256;
  __shared__ cpx<F> s[NL * HUB_LS];
  __shared__ cpx<F> stw[L];
  for (int t = threadIdx.x; t < L; t += NL * 16) stw[t] = tw[t];
  __syncthreads();
  const int l = threadIdx.x % NL, j = threadIdx.x / NL;
  // XCD-contiguous: the two 64-byte halves of a kernel-spectrum line are
  // read by neighbouring workgroups, which then share one L2
  const int line = ssim_xcd_block(blockIdx.x, gridDim.x) * NL + l;
  // element (x, line) at line + x * lines (< 2^31, checked by the host)
  const unsigned int base = (unsigned int)line + (unsigned int)j * (unsigned int)lines;
  const unsigned int step = 16u * (unsigned int)lines;
  // Kernel spectra: component c is odd in k_c and even in the other two
  // axes (exactly so: its real-space kernel is exactly odd / even), so only
  // the kx, ky <= 128 corner of each table is read, mirrored with the sign:
  // 8.6 of the 33.8 MB of a table.
  const int ky = line / 129, kz = line - ky * 129;
  const bool ky_hi = ky > 128;
  const unsigned int kcorner = (unsigned int)((ky_hi ? 256 - ky : ky) * 129 + kz);
  // kx = j + 16 k1 reads kx for k1 < 8 and 256 - kx for k1 >= 8 (both 128
  // at the Nyquist plane, where component 0 vanishes): two bases, fixed
  // strides
  auto kload = [&](const F* __restrict__ kim, int c, F km[16]) {
    const F* lo = kim + kcorner + (size_t)j * lines;
    const F* hi = kim + kcorner + (size_t)(256 - j) * lines;
    const bool neg_y = c == 1 && ky_hi;
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) {
      const F val = k1 < 8 ? lo[(size_t)(16 * k1) * lines] : hi[-(ptrdiff_t)(16 * k1) * lines];
      const bool neg = (c == 0 && k1 >= 8) != neg_y;
      km[k1] = neg ? -val : val;
    }
  };
  cpx<F> v[16];
  F km[16];  // kernel spectrum of the next component, loaded one stage early
#pragma unroll
  for (int n2 = 0; n2 < 8; ++n2) v[n2] = in[base + n2 * step];
  kload(kim0, 0, km);
#pragma unroll
  for (int n2 = 8; n2 < 16; ++n2) v[n2] = {F(0), F(0)};
  dft16(v, false);
  hub_exchange<F, NL>(v, s, l, j, stw, false);
  // X[j + 16 k1] = v[k1]
  cpx<F> X[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) X[k] = v[k];
#pragma unroll 1
  for (int c = 0; c < 3; ++c) {
    cpx<F>* out = c == 0 ? out0 : (c == 1 ? out1 : out2);
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) v[k1] = {-X[k1].im * km[k1], X[k1].re * km[k1]};
    if (c < 2) kload(c == 0 ? kim1 : kim2, c + 1, km);  // overlaps this inverse
    dft16(v, true);
    hub_exchange<F, NL>(v, s, l, j, stw, true);
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) out[base + k1 * step] = v[k1];
  }
}

enum { LY_FWD = 0, LY_INV = 1, LZ_R2C = 2, LZ_C2R = 3, LZ_C2R3 = 4 };

// Up to three arrays of one pass: component c's lines are lines
// [c * per, (c + 1) * per) of the launch.
struct Lines3 {
  const void* in[3];
  void* out[3];
};

// load the 16 inputs x[j + 16 n2] of a line (see k_line256)
template <typename F, int KIND>
__device__ __forceinline__ void line_load(const void* vin, int lc, int j, bool live,
                                          cpx<F> v[16]) {
  constexpr int n = 128, H = 129, PH = 256 * 129;
  if (KIND == LY_FWD || KIND == LY_INV) {
    // line = x * H + kz; element ky at x * PH + ky * H + kz
    const int x = lc / H, kz = lc - x * H;
    const cpx<F>* in = static_cast<const cpx<F>*>(vin) + (size_t)x * PH + kz;
    const int nin = KIND == LY_FWD ? 8 : 16;  // n2 < 8: ky = j + 16 n2 < n
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2)
      v[n2] = (n2 < nin && live) ? in[(j + 16 * n2) * H] : cpx<F>{F(0), F(0)};
  } else if (KIND == LZ_R2C) {
    const F* in = static_cast<const F*>(vin) + (size_t)lc * n;
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2)
      v[n2] = {(n2 < 8 && live) ? in[j + 16 * n2] : F(0), F(0)};
  } else {
    // row (x, y) = lc of the half spectrum at x * PH + y * H; e > n mirrored
    const int x = lc / n, y = lc - x * n;
    const cpx<F>* row = static_cast<const cpx<F>*>(vin) + (size_t)x * PH + y * H;
#pragma unroll
    for (int n2 = 0; n2 < 16; ++n2) {
      const int e = j + 16 * n2;
      cpx<F> c = {F(0), F(0)};
      if (live) {
        if (e <= n) {
          c = row[e];
        } else {
          const cpx<F> m = row[256 - e];
          c = {m.re, -m.im};
        }
      }
      v[n2] = c;
    }
  }
}

// ---- 256-point lines of the isolated convolution, register-direct ----
// The y and z passes of the padded 2n = 256 mesh (n = 128, H = 129) in the
// style of k_conv_x3: thread (line l, index j) loads its 16 (or 8) inputs
// straight from HBM, runs the four-step FFT with one LDS exchange, and stores
// its outputs straight from registers; only the x < n planes exist.
//   LY_FWD:  y lines of plane x (kz = line % H), inputs ky < n, all 256 out;
//   LY_INV:  the same lines, all 256 in, ky < n out (`per` lines per
//            component, up to three components in one launch);
//   LZ_R2C:  rows (x, y) of the real n^3 mesh (z < n, zeros beyond), the
//            H = n + 1 half-spectrum values out;
//   LZ_C2R:  rows (x, y) of the half spectrum (Hermitian completion), z < n
//            real out, times `scale`, at out[(row * n + z) * es + ec];
//   LZ_C2R3: the same rows of all three components, written as float4
//            (c0, c1, c2, 0) [n^3]: whole 16-byte slots, one launch.
// y lines: lanes take consecutive lines (consecutive kz: 128-byte rows);
// z rows: lanes take consecutive indices of one row.
template <typename F, int NL, int KIND>
__global__ __launch_bounds__(NL * 16) void k_line256(
    Lines3 P, const cpx<F>* __restrict__ tw,  // y passes: in place
    int lines, int per, F scale, int es, int ec) {
  constexpr int n = 128, H = 129, PH = 256 * 129;
  constexpr bool YL = KIND == LY_FWD || KIND == LY_INV;
  // exchange rows: 258 apart when lanes take consecutive lines, 272 when
  // they take consecutive indices of a line (both free of bank conflicts)
  constexpr int LS = YL ? HUB_LS : 272;
  __shared__ cpx<F> s[NL * LS];
  __shared__ cpx<F> stw[256];
  for (int t = threadIdx.x; t < 256; t += NL * 16) stw[t] = tw[t];
  __syncthreads();
  const int l = YL ? threadIdx.x % NL : threadIdx.x / 16;
  const int j = YL ? threadIdx.x / NL : threadIdx.x % 16;
  const int line = ssim_xcd_block(blockIdx.x, gridDim.x) * NL + l;  // XCD-contiguous
  const bool live = line < lines;
  const int lg = live ? line : 0;
  const int comp = lg / per, lc = lg - comp * per;  // component, line in it
  const bool inverse = KIND == LY_INV || KIND == LZ_C2R || KIND == LZ_C2R3;
  cpx<F> v[16];
  if (KIND == LZ_C2R3) {
    F res[3][8];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      line_load<F, LZ_C2R>(P.in[c], lc, j, live, v);
      dft16(v, true);
      hub_exchange<F, NL, LS>(v, s, l, j, stw, true);
#pragma unroll
      for (int k1 = 0; k1 < 8; ++k1) res[c][k1] = v[k1].re * scale;
    }
    if (!live) return;
    float4* out = static_cast<float4*>(P.out[0]);
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1)
      out[(size_t)lc * n + j + 16 * k1] =
          make_float4((float)res[0][k1], (float)res[1][k1], (float)res[2][k1], 0.f);
    return;
  }
  line_load<F, KIND>(P.in[comp], lc, j, live, v);
  dft16(v, inverse);
  hub_exchange<F, NL, LS>(v, s, l, j, stw, inverse);
  // v[k1] = X[j + 16 k1]
  if (!live) return;
  if (YL) {
    const int x = lc / H, kz = lc - x * H;
    cpx<F>* out = static_cast<cpx<F>*>(P.out[comp]) + (size_t)x * PH + kz;
    const int nout = KIND == LY_FWD ? 16 : 8;
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1)
      if (k1 < nout) out[(j + 16 * k1) * H] = v[k1];
  } else if (KIND == LZ_R2C) {
    const int x = lc / n, y = lc - x * n;
    cpx<F>* out = static_cast<cpx<F>*>(P.out[0]) + (size_t)x * PH + y * H;
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) out[j + 16 * k1] = v[k1];
    if (j == 0) out[n] = v[8];
  } else {
    F* out = static_cast<F*>(P.out[0]);
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1)
      out[((size_t)lc * n + j + 16 * k1) * es + ec] = v[k1].re * scale;
  }
}

template <typename F>
__global__ void k_fft_twiddles(int L, cpx<F>* tw) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= L) return;
  double sn, cs;
  sincospi(-2.0 * (double)j / (double)L, &sn, &cs);
  tw[j] = {(F)cs, (F)sn};
}

size_t elem_bytes(int prec) { return prec == 0 ? 16 : 8; }  // complex

bool valid_dims(int ndim, const int* dims) {
  if (ndim < 1 || ndim > 3 || !dims) return false;
  for (int a = 0; a < ndim; ++a) {
    const int L = dims[a];
    if (L < 2 || L > FFT_MAX_L || (L & (L - 1))) return false;
  }
  return true;
}

// Twiddle table offsets: one table of L_a complex values per axis.
size_t tw_offset(int ndim, const int* dims, int axis, int prec) {
  size_t off = 0;
  for (int a = 0; a < axis; ++a) off += ssim::align256(dims[a] * elem_bytes(prec));
  (void)ndim;
  return off;
}

template <typename F, int MODE>
int launch_pass(const void* in, void* out, Pass P, const void* tw, int inverse,
                F scale, hipStream_t st, const F* mul = nullptr) {
  int nwg;
  if (P.contig)
    nwg = ssim::ceil_div(P.n_outer, P.lines_per_tile);
  else
    nwg = P.n_outer * ssim::ceil_div(P.n_inner, P.lines_per_tile);
  size_t lds = (size_t)P.lines_per_tile *
               (P.lines_per_tile > 1 ? P.L + 1 : P.L) * sizeof(cpx<F>);
  P.tw_lds = lds + (size_t)P.L * sizeof(cpx<F>) <= 65536 ? 1 : 0;
  if (P.tw_lds) lds += (size_t)P.L * sizeof(cpx<F>);
  if (P.L <= FFT_LINE_L) {
    SSIM_TIMED("k_fft_pass", st,
               hipLaunchKernelGGL((k_fft_pass<F, MODE>), dim3(nwg), dim3(FFT_THREADS),
                                  lds, st, in, out, P, static_cast<const cpx<F>*>(tw),
                                  inverse, scale, mul));
  } else {
    // one line per workgroup in more than 64 KB of LDS (8192 points: 64 KB
    // in FP32, 128 KB in FP64), which the device must offer a workgroup
    SSIM_REQUIRE(P.lines_per_tile == 1, "long FFT lines go one per workgroup");
    int dev = 0, max_lds = 0;
    SSIM_HIP(hipGetDevice(&dev));
    SSIM_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    SSIM_REQUIRE(lds <= (size_t)max_lds,
                 "%d-point FFT lines need %zu bytes of LDS per workgroup, the device "
                 "offers %d", P.L, lds, max_lds);
    SSIM_HIP(hipFuncSetAttribute(
        reinterpret_cast<const void*>(&k_fft_pass<F, MODE, FFT_MAX_L>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    SSIM_TIMED("k_fft_pass", st,
               hipLaunchKernelGGL((k_fft_pass<F, MODE, FFT_MAX_L>), dim3(nwg),
                                  dim3(FFT_THREADS), lds, st, in, out, P,
                                  static_cast<const cpx<F>*>(tw), inverse, scale, mul));
  }
  SSIM_LAUNCHED("k_fft_pass");
  return SSIM_OK;
}

// Pass along `axis` of the complex array with shape cdims (in place).
Pass strided_pass(int ndim, const int* cdims, int axis, int tile) {
  Pass P{};
  P.L = cdims[axis];
  P.log2L = ilog2(P.L);
  long long inner = 1, outer = 1;
  for (int a = axis + 1; a < ndim; ++a) inner *= cdims[a];
  for (int a = 0; a < axis; ++a) outer *= cdims[a];
  P.n_outer = (int)outer;
  P.n_inner = (int)inner;
  P.s_elem = inner;
  P.s_outer_in = P.s_outer_out = inner * P.L;
  P.contig = 0;
  P.in_count = P.out_count = P.L;
  P.R1 = 1;
  P.s_in_hi = P.s_out_hi = 0;
  P.out_es = 1;
  P.out_c = 0;
  int b = tile / P.L;
  if (b > FFT_MAXB) b = FFT_MAXB;
  if (b > P.n_inner) {
    b = 1;
    while (b * 2 <= P.n_inner) b *= 2;
  }
  P.lines_per_tile = b < 1 ? 1 : b;
  return P;
}

Pass row_pass(long long rows, int L, long long s_in, long long s_out, int tile) {
  Pass P{};
  P.L = L;
  P.log2L = ilog2(L);
  P.n_outer = (int)rows;
  P.n_inner = 1;
  P.s_elem = 1;
  P.s_outer_in = s_in;
  P.s_outer_out = s_out;
  P.contig = 1;
  P.in_count = P.out_count = L;
  P.R1 = rows > 0 ? rows : 1;
  P.s_in_hi = P.s_out_hi = 0;
  P.out_es = 1;
  P.out_c = 0;
  int b = tile / L;
  if (b < 1) b = 1;
  P.lines_per_tile = b;
  return P;
}

template <typename F>
int fft_r2c(int ndim, const int* dims, const F* in, cpx<F>* out,
            const unsigned char* work, int prec, hipStream_t st) {
  int cdims[3];
  long long rows = 1;
  for (int a = 0; a < ndim; ++a) cdims[a] = dims[a];
  cdims[ndim - 1] = dims[ndim - 1] / 2 + 1;
  for (int a = 0; a < ndim - 1; ++a) rows *= dims[a];
  const int L = dims[ndim - 1];
  SSIM_TRY((launch_pass<F, MODE_R2C>(
      in, out, row_pass(rows, L, L, cdims[ndim - 1], tile_elems<F>()),
      work + tw_offset(ndim, dims, ndim - 1, prec), 0, F(1), st)));
  for (int a = ndim - 2; a >= 0; --a)
    SSIM_TRY((launch_pass<F, MODE_C2C>(
        out, out, strided_pass(ndim, cdims, a, tile_elems<F>()),
        work + tw_offset(ndim, dims, a, prec), 0, F(1), st)));
  return SSIM_OK;
}

template <typename F>
int fft_c2r(int ndim, const int* dims, cpx<F>* in, F* out,
            const unsigned char* work, int prec, hipStream_t st) {
  int cdims[3];
  long long rows = 1;
  double total = 1.0;
  for (int a = 0; a < ndim; ++a) {
    cdims[a] = dims[a];
    total *= dims[a];
  }
  cdims[ndim - 1] = dims[ndim - 1] / 2 + 1;
  for (int a = 0; a < ndim - 1; ++a) rows *= dims[a];
  for (int a = 0; a < ndim - 1; ++a)
    SSIM_TRY((launch_pass<F, MODE_C2C>(
        in, in, strided_pass(ndim, cdims, a, tile_elems<F>()),
        work + tw_offset(ndim, dims, a, prec), 1, F(1), st)));
  const int L = dims[ndim - 1];
  SSIM_TRY((launch_pass<F, MODE_C2R>(
      in, out, row_pass(rows, L, cdims[ndim - 1], L, tile_elems<F>()),
      work + tw_offset(ndim, dims, ndim - 1, prec), 1, F(1.0 / total), st)));
  return SSIM_OK;
}

// The periodic potential of a 3-D density in five passes: R2C rows, forward
// y lines, the fused x pass (forward, Poisson factor, inverse), inverse y
// lines, C2R rows — the same operations as fft_r2c + k_poisson + fft_c2r
// without the two spectrum round trips of the x axis.
template <typename F>
int poisson3(const int* dims, const F* rho, cpx<F>* spec, F* phi, double h, double G,
             const unsigned char* work, int prec, hipStream_t st) {
  const int cdims[3] = {dims[0], dims[1], dims[2] / 2 + 1};
  const long long rows = (long long)dims[0] * dims[1];
  const int L = dims[2];
  const double total = (double)dims[0] * dims[1] * dims[2];
  SSIM_TRY((launch_pass<F, MODE_R2C>(rho, spec, row_pass(rows, L, L, cdims[2], tile_elems<F>()),
                                     work + tw_offset(3, dims, 2, prec), 0, F(1), st)));
  SSIM_TRY((launch_pass<F, MODE_C2C>(spec, spec, strided_pass(3, cdims, 1, tile_elems<F>()),
                                     work + tw_offset(3, dims, 1, prec), 0, F(1), st)));
  Pass px = strided_pass(3, cdims, 0, tile_elems<F>());
  px.pn0 = dims[0];
  px.pn1 = dims[1];
  px.pn2 = dims[2];
  px.ph = h;
  px.pG = G;
  SSIM_TRY((launch_pass<F, MODE_POISSON>(spec, spec, px, work + tw_offset(3, dims, 0, prec),
                                         0, F(1), st)));
  SSIM_TRY((launch_pass<F, MODE_C2C>(spec, spec, strided_pass(3, cdims, 1, tile_elems<F>()),
                                     work + tw_offset(3, dims, 1, prec), 1, F(1), st)));
  SSIM_TRY((launch_pass<F, MODE_C2R>(spec, phi, row_pass(rows, L, cdims[2], L, tile_elems<F>()),
                                     work + tw_offset(3, dims, 2, prec), 1, F(1.0 / total),
                                     st)));
  return SSIM_OK;
}

// The 256^3 padded convolution (n = 128) from register-direct kernels:
// R2C rows, forward y lines, the fused x axis, and per component inverse y
// lines + C2R rows; spec and tmp as in conv_isolated3.
template <typename F>
int conv256(const F* rho, const F* const* kim, F* const* fields, cpx<F>* spec,
            cpx<F>* tmp, const unsigned char* twx, const unsigned char* twy,
            const unsigned char* twz, bool inter, F scale, hipStream_t st) {
  constexpr int n = 128, H = 129, NL = sizeof(F) == 4 ? 16 : 8;
  const cpx<F>* tx = reinterpret_cast<const cpx<F>*>(twx);
  const cpx<F>* ty = reinterpret_cast<const cpx<F>*>(twy);
  const cpx<F>* tz = reinterpret_cast<const cpx<F>*>(twz);
  const int rows = n * n, ylines = n * H;
  const dim3 blk(NL * 16);
  Lines3 L{{rho, nullptr, nullptr}, {spec, nullptr, nullptr}};
  SSIM_TIMED("k_fft_pass", st,
             hipLaunchKernelGGL((k_line256<F, NL, LZ_R2C>), dim3(ssim::ceil_div(rows, NL)),
                                blk, 0, st, L, tz, rows, rows, F(1), 1, 0));
  SSIM_LAUNCHED("k_line256");
  L = Lines3{{spec, nullptr, nullptr}, {spec, nullptr, nullptr}};
  SSIM_TIMED("k_fft_pass", st,
             hipLaunchKernelGGL((k_line256<F, NL, LY_FWD>), dim3(ssim::ceil_div(ylines, NL)),
                                blk, 0, st, L, ty, ylines, ylines, F(1), 1, 0));
  SSIM_LAUNCHED("k_line256");
  // fused x axis: component 0 in place in spec (x < n), components 1 and 2
  // in the two halves of tmp, each an [n][P][H] block
  const long long lines = 256LL * H;
  static_assert((256LL * 129) % NL == 0, "whole tiles of x lines");
  cpx<F>* blk3[3] = {spec, tmp, tmp + (long long)n * lines};
  SSIM_TIMED("k_conv_x3", st,
             hipLaunchKernelGGL((k_conv_x3<F, NL>), dim3((int)(lines / NL)),
                                dim3(NL * 16), 0, st, spec, kim[0], kim[1], kim[2],
                                blk3[0], blk3[1], blk3[2], tx, (int)lines));
  SSIM_LAUNCHED("k_conv_x3");
  // inverse y lines of the three components in one launch
  L = Lines3{{blk3[0], blk3[1], blk3[2]}, {blk3[0], blk3[1], blk3[2]}};
  SSIM_TIMED("k_fft_pass", st,
             hipLaunchKernelGGL((k_line256<F, NL, LY_INV>),
                                dim3(ssim::ceil_div(3 * ylines, NL)), blk, 0, st, L, ty,
                                3 * ylines, ylines, F(1), 1, 0));
  SSIM_LAUNCHED("k_line256");
  if (inter && sizeof(F) == 4) {
    // C2R rows of the three components, whole float4 slots
    L = Lines3{{blk3[0], blk3[1], blk3[2]}, {fields[0], nullptr, nullptr}};
    SSIM_TIMED("k_fft_pass", st,
               hipLaunchKernelGGL((k_line256<F, NL, LZ_C2R3>), dim3(ssim::ceil_div(rows, NL)),
                                  blk, 0, st, L, tz, rows, rows, scale, 4, 0));
    SSIM_LAUNCHED("k_line256");
    return SSIM_OK;
  }
  for (int c = 0; c < 3; ++c) {
    L = Lines3{{blk3[c], nullptr, nullptr}, {inter ? fields[0] : fields[c], nullptr, nullptr}};
    SSIM_TIMED("k_fft_pass", st,
               hipLaunchKernelGGL((k_line256<F, NL, LZ_C2R>), dim3(ssim::ceil_div(rows, NL)),
                                  blk, 0, st, L, tz, rows, rows, scale, inter ? 4 : 1,
                                  inter ? c : 0));
    SSIM_LAUNCHED("k_line256");
  }
  return SSIM_OK;
}

// Isolated convolution of an n^3 mesh with three imaginary kernel spectra
// on the zero-padded P = 2n mesh, keeping only the n^3 corner:
//   forward: R2C rows (x, y < n) read straight from rho (n points, zeros
//            beyond), y lines of the planes x < n (n points read), x lines
//            (n points read);
//   per component: x lines read spec * i kim_c and write x < n into tmp,
//            y lines of the planes x < n write y < n, C2R rows (x, y < n)
//            write z < n into the field, scaled by 1/P^3.
// The padding never exists in memory and nothing outside the corner is
// computed in the inverse: about 2.6x fewer HBM bytes than pad + rfftn +
// 3 x (product + irfftn + crop).  With 256-point lines (n = 128) the
// forward x lines and the three product + inverse x lines are one kernel
// (k_conv_x3) that keeps each line's spectrum in registers; its outputs
// go to spec (component 0, in place) and the two halves of tmp.
// fields[1] == fields[2] == nullptr: the three results are interleaved in
// fields[0] as [n^3][4] (component c at 4 i + c, slot 3 untouched).
template <typename F>
int conv_isolated3(int n, const F* rho, const F* const* kim, F* const* fields,
                   cpx<F>* spec, cpx<F>* tmp, const unsigned char* tw, int prec,
                   hipStream_t st) {
  const bool inter = fields[1] == nullptr && fields[2] == nullptr;
  const int P = 2 * n, H = n + 1;
  const int dims[3] = {P, P, P};
  const int cd[3] = {P, P, H};
  const unsigned char* twx = tw + tw_offset(3, dims, 0, prec);
  const unsigned char* twy = tw + tw_offset(3, dims, 1, prec);
  const unsigned char* twz = tw + tw_offset(3, dims, 2, prec);
  if (P == 256)
    return conv256<F>(rho, kim, fields, spec, tmp, twx, twy, twz, inter,
                      F(1.0 / ((double)P * P * P)), st);
  const long long rows = (long long)n * n;
  // row passes use half tiles: twice the workgroups, so loads, butterflies
  // and stores of different workgroups overlap (a full-tile row pass of the
  // 128^2 rows is a single round of 1024 workgroups)
  Pass fz = row_pass(rows, P, n, H, tile_elems<F>() / 2);
  fz.R1 = n;
  fz.s_in_hi = rows;
  fz.s_out_hi = (long long)P * H;
  fz.in_count = n;
  SSIM_TRY((launch_pass<F, MODE_R2C>(rho, spec, fz, twz, 0, F(1), st)));
  Pass fy = strided_pass(3, cd, 1, tile_elems<F>());
  fy.n_outer = n;
  fy.in_count = n;
  SSIM_TRY((launch_pass<F, MODE_C2C>(spec, spec, fy, twy, 0, F(1), st)));
  const F scale = F(1.0 / ((double)P * P * P));
  Pass fx = strided_pass(3, cd, 0, tile_elems<F>());
  fx.in_count = n;
  SSIM_TRY((launch_pass<F, MODE_C2C>(spec, spec, fx, twx, 0, F(1), st)));
  for (int c = 0; c < 3; ++c) {
    Pass ix = strided_pass(3, cd, 0, tile_elems<F>());
    ix.out_count = n;
    SSIM_TRY((launch_pass<F, MODE_C2C>(spec, tmp, ix, twx, 1, F(1), st, kim[c])));
    Pass iy = strided_pass(3, cd, 1, tile_elems<F>());
    iy.n_outer = n;
    iy.out_count = n;
    SSIM_TRY((launch_pass<F, MODE_C2C>(tmp, tmp, iy, twy, 1, F(1), st)));
    Pass iz = row_pass(rows, P, H, n, tile_elems<F>() / 2);
    iz.R1 = n;
    iz.s_in_hi = (long long)P * H;
    iz.s_out_hi = rows;
    iz.out_count = n;
    if (inter) {
      iz.out_es = 4;
      iz.out_c = c;
    }
    SSIM_TRY((launch_pass<F, MODE_C2R>(tmp, inter ? fields[0] : fields[c], iz, twz, 1,
                                       scale, st)));
  }
  return SSIM_OK;
}

}  // namespace

namespace ssim {

size_t fft_work_bytes(int ndim, const int* dims, int prec) {
  if (!valid_dims(ndim, dims)) return 0;
  return tw_offset(ndim, dims, ndim, prec);
}

int fft_prepare(int ndim, const int* dims, int prec, void* work,
                size_t work_bytes, hipStream_t st) {
  SSIM_REQUIRE(valid_dims(ndim, dims),
               "FFT dims must be 1..3 powers of two in [2, %d]", FFT_MAX_L);
  SSIM_REQUIRE(prec == SSIM_PREC_F64 || prec == SSIM_PREC_F32C,
               "unknown precision %d", prec);
  SSIM_REQUIRE(work && work_bytes >= fft_work_bytes(ndim, dims, prec),
               "FFT workspace too small");
  unsigned char* w = static_cast<unsigned char*>(work);
  for (int a = 0; a < ndim; ++a) {
    const int L = dims[a];
    void* t = w + tw_offset(ndim, dims, a, prec);
    if (prec == SSIM_PREC_F64)
      hipLaunchKernelGGL(k_fft_twiddles<double>, dim3(ceil_div(L, 256)),
                         dim3(256), 0, st, L, static_cast<cpx<double>*>(t));
    else
      hipLaunchKernelGGL(k_fft_twiddles<float>, dim3(ceil_div(L, 256)),
                         dim3(256), 0, st, L, static_cast<cpx<float>*>(t));
    SSIM_LAUNCHED("k_fft_twiddles");
  }
  return SSIM_OK;
}

int fft_r2c(int ndim, const int* dims, int prec, const void* in, void* out,
            const void* work, size_t work_bytes, hipStream_t st) {
  SSIM_REQUIRE(valid_dims(ndim, dims),
               "FFT dims must be 1..3 powers of two in [2, %d]", FFT_MAX_L);
  SSIM_REQUIRE(in && out && work, "null pointer");
  SSIM_REQUIRE(work_bytes >= fft_work_bytes(ndim, dims, prec),
               "FFT workspace too small");
  const unsigned char* w = static_cast<const unsigned char*>(work);
  if (prec == SSIM_PREC_F64)
    return ::fft_r2c<double>(ndim, dims, static_cast<const double*>(in),
                             static_cast<cpx<double>*>(out), w, prec, st);
  return ::fft_r2c<float>(ndim, dims, static_cast<const float*>(in),
                          static_cast<cpx<float>*>(out), w, prec, st);
}

int fft_c2r(int ndim, const int* dims, int prec, void* in, void* out,
            const void* work, size_t work_bytes, hipStream_t st) {
  SSIM_REQUIRE(valid_dims(ndim, dims),
               "FFT dims must be 1..3 powers of two in [2, %d]", FFT_MAX_L);
  SSIM_REQUIRE(in && out && work, "null pointer");
  SSIM_REQUIRE(work_bytes >= fft_work_bytes(ndim, dims, prec),
               "FFT workspace too small");
  const unsigned char* w = static_cast<const unsigned char*>(work);
  if (prec == SSIM_PREC_F64)
    return ::fft_c2r<double>(ndim, dims, static_cast<cpx<double>*>(in),
                             static_cast<double*>(out), w, prec, st);
  return ::fft_c2r<float>(ndim, dims, static_cast<cpx<float>*>(in),
                          static_cast<float*>(out), w, prec, st);
}

int pm_potential_periodic(const int* dims, int prec, const void* rho, void* spec,
                          void* phi, double h, double G, const void* work,
                          size_t work_bytes, hipStream_t st) {
  SSIM_REQUIRE(valid_dims(3, dims),
               "FFT dims must be powers of two in [2, %d]", FFT_MAX_L);
  SSIM_REQUIRE(rho && spec && phi && work, "null pointer");
  SSIM_REQUIRE(h > 0.0, "h must be > 0");
  SSIM_REQUIRE(work_bytes >= fft_work_bytes(3, dims, prec), "FFT workspace too small");
  const unsigned char* w = static_cast<const unsigned char*>(work);
  if (prec == SSIM_PREC_F64)
    return ::poisson3<double>(dims, static_cast<const double*>(rho),
                              static_cast<cpx<double>*>(spec), static_cast<double*>(phi), h,
                              G, w, prec, st);
  return ::poisson3<float>(dims, static_cast<const float*>(rho),
                           static_cast<cpx<float>*>(spec), static_cast<float*>(phi), h, G, w,
                           prec, st);
}

}  // namespace ssim

extern "C" {

size_t ssim_fft_work_bytes(int ndim, const int* dims, int prec) {
  return ssim::fft_work_bytes(ndim, dims, prec);
}

int ssim_pm_potential_periodic(const int* grid_n, int prec, const void* rho, void* spec,
                               void* phi, double h, double G, const void* fft_work,
                               size_t fft_work_bytes, void* stream) {
  return ssim::pm_potential_periodic(grid_n, prec, rho, spec, phi, h, G, fft_work,
                                     fft_work_bytes, ssim::as_stream(stream));
}

int ssim_pm_conv_isolated3(int n, int prec, const void* rho, const void* kim_x,
                           const void* kim_y, const void* kim_z, void* fx,
                           void* fy, void* fz, void* spec, void* tmp,
                           const void* fft_work, size_t fft_work_bytes,
                           void* stream) {
  SSIM_REQUIRE(n >= 1 && 2 * n <= FFT_MAX_L && (n & (n - 1)) == 0,
               "n must be a power of two with 2n <= %d", FFT_MAX_L);
  SSIM_REQUIRE(prec == SSIM_PREC_F64 || prec == SSIM_PREC_F32C,
               "unknown precision %d", prec);
  SSIM_REQUIRE(rho && kim_x && kim_y && kim_z && fx && spec && tmp && fft_work,
               "null pointer");
  SSIM_REQUIRE((fy == nullptr) == (fz == nullptr),
               "fy and fz are both given or both NULL (interleaved)");
  const int dims[3] = {2 * n, 2 * n, 2 * n};
  SSIM_REQUIRE(fft_work_bytes >= ssim::fft_work_bytes(3, dims, prec),
               "FFT workspace too small (prepare it for the padded 2n^3 mesh)");
  const unsigned char* tw = static_cast<const unsigned char*>(fft_work);
  hipStream_t st = ssim::as_stream(stream);
  if (prec == SSIM_PREC_F64) {
    const double* k[3] = {static_cast<const double*>(kim_x),
                          static_cast<const double*>(kim_y),
                          static_cast<const double*>(kim_z)};
    double* f[3] = {static_cast<double*>(fx), static_cast<double*>(fy),
                    static_cast<double*>(fz)};
    return conv_isolated3<double>(n, static_cast<const double*>(rho), k, f,
                                  static_cast<cpx<double>*>(spec),
                                  static_cast<cpx<double>*>(tmp), tw, prec, st);
  }
  const float* k[3] = {static_cast<const float*>(kim_x),
                       static_cast<const float*>(kim_y),
                       static_cast<const float*>(kim_z)};
  float* f[3] = {static_cast<float*>(fx), static_cast<float*>(fy),
                 static_cast<float*>(fz)};
  return conv_isolated3<float>(n, static_cast<const float*>(rho), k, f,
                               static_cast<cpx<float>*>(spec),
                               static_cast<cpx<float>*>(tmp), tw, prec, st);
}

int ssim_fft_prepare(int ndim, const int* dims, int prec, void* work,
                     size_t work_bytes, void* stream) {
  return ssim::fft_prepare(ndim, dims, prec, work, work_bytes,
                           ssim::as_stream(stream));
}

int ssim_fft_r2c(int ndim, const int* dims, int prec, const void* in,
                 void* out, const void* work, size_t work_bytes, void* stream) {
  return ssim::fft_r2c(ndim, dims, prec, in, out, work, work_bytes,
                       ssim::as_stream(stream));
}

int ssim_fft_c2r(int ndim, const int* dims, int prec, void* in, void* out,
                 const void* work, size_t work_bytes, void* stream) {
  return ssim::fft_c2r(ndim, dims, prec, in, out, work, work_bytes,
                       ssim::as_stream(stream));
}

}  // extern "C"
