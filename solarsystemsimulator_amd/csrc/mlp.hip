// Learned near-field correction: feature extraction and a 32-64-64-64-16
// multilayer perceptron (SiLU) with BF16 MFMA and FP32 accumulation, plus
// its training kernels (backward, weight gradients, Adam).  No reference
// counterpart: the reference has no learned component.
//
// MFMA formulation (v_mfma_f32_16x16x16_bf16, wave64).  Every layer is
// computed transposed, H^T[out x 16 bodies] = W[out x in] X^T[in x 16], so
// the 16x16 accumulator tile of one layer (lane l: rows 4(l/16)+i, column
// l%16) is already the B-operand layout of the next layer (lane l: k rows
// 4(l/16)+i, column l%16): bias + SiLU + BF16 rounding happen in registers
// and no LDS shuffle is needed between layers.  A-operand weight tiles
// (lane l: row l%16, k 4(l/16)+i) are 4 consecutive BF16 of one weight row,
// loaded once per wave; all four layers' weights stay in ~88 VGPRs.
//
// Parameter block (FP32 master copy, same layout in BF16 for the weights):
//   W1[64][32] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[16][64] b4[16]
// with rows 3..15 of W4 / b4 zero (three outputs).
#include "ssim_internal.h"

namespace {

typedef short s4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

constexpr int NIN = 32, NH = 64, NOUT = 16, NOUT_USED = 3;
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + NH * NIN;
constexpr int OFF_W2 = OFF_B1 + NH;
constexpr int OFF_B2 = OFF_W2 + NH * NH;
constexpr int OFF_W3 = OFF_B2 + NH;
constexpr int OFF_B3 = OFF_W3 + NH * NH;
constexpr int OFF_W4 = OFF_B3 + NH;
constexpr int OFF_B4 = OFF_W4 + NOUT * NH;
constexpr int NPARAM = OFF_B4 + NOUT;

constexpr int WAVES = 4;  // waves per workgroup
constexpr int MB = 256;

// FP32 -> BF16 bits, round to nearest even (finite inputs).
__device__ __forceinline__ unsigned short bf16_bits(float f) {
  unsigned int u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_float(unsigned short b) {
  return __uint_as_float(((unsigned int)b) << 16);
}

// sigmoid with the hardware reciprocal (1 ulp) instead of a full-precision
// division: the BF16 rounding of the activations is 2^-9, far coarser
__device__ __forceinline__ float sigmoid(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}
__device__ __forceinline__ float silu(float z) { return z * sigmoid(z); }
__device__ __forceinline__ float silu_grad(float z) {
  const float s = sigmoid(z);
  return s * (1.0f + z * (1.0f - s));
}

__device__ __forceinline__ f4_t mfma(s4_t a, s4_t b, f4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// A tile (row block rt, k block kb) of a row-major weight matrix [rows][K].
__device__ __forceinline__ s4_t load_a(const unsigned short* __restrict__ w,
                                       int K, int rt, int kb, int lane) {
  const unsigned short* p = w + (size_t)(16 * rt + (lane & 15)) * K + 16 * kb +
                            4 * (lane >> 4);
  s4_t v;
  v.x = (short)p[0];
  v.y = (short)p[1];
  v.z = (short)p[2];
  v.w = (short)p[3];
  return v;
}

struct Weights {
  s4_t w1[4][2], w2[4][4], w3[4][4], w4[4];
};

__device__ __forceinline__ void load_weights(const unsigned short* __restrict__ wb,
                                             int lane, Weights& W) {
#pragma unroll
  for (int rt = 0; rt < 4; ++rt) {
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) W.w1[rt][kb] = load_a(wb + OFF_W1, NIN, rt, kb, lane);
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
      W.w2[rt][kb] = load_a(wb + OFF_W2, NH, rt, kb, lane);
      W.w3[rt][kb] = load_a(wb + OFF_W3, NH, rt, kb, lane);
    }
    W.w4[rt] = load_a(wb + OFF_W4, NH, 0, rt, lane);
  }
}

// bias + SiLU on one accumulator tile; returns the BF16 B operand of the
// next layer; optionally records Z (FP32) and H (BF16) rows of the body.
__device__ __forceinline__ s4_t activate(f4_t acc, const float* __restrict__ bias,
                                         int rt, int lane, float* zrow,
                                         unsigned short* hrow) {
  s4_t out;
  short o[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = 16 * rt + 4 * (lane >> 4) + i;
    const float z = acc[i] + bias[r];
    const unsigned short hb = bf16_bits(silu(z));
    o[i] = (short)hb;
    if (zrow) zrow[r] = z;
    if (hrow) hrow[r] = hb;
  }
  out.x = o[0];
  out.y = o[1];
  out.z = o[2];
  out.w = o[3];
  return out;
}

// Forward pass of 16 bodies (columns) starting at b0.  x: BF16 features
// [nb][32].  Writes the 3 outputs (unscaled) and, when zs/hs != nullptr,
// the layer activations for training: zs[l][body][64] FP32, hs[l][body][64]
// BF16, l = 0..2.
// `bias` holds b1 | b2 | b3 (64 each) | b4 (16) contiguously (LDS).
__device__ void forward16(const Weights& W, const float* __restrict__ bias,
                          const unsigned short* __restrict__ x, int nb, int b0,
                          int lane, float out3[3], bool& valid, float* zs,
                          unsigned short* hs) {
  const int body = b0 + (lane & 15);
  valid = body < nb;
  s4_t xb[2];
#pragma unroll
  for (int kb = 0; kb < 2; ++kb) {
    // 4 consecutive BF16 features: one 8-byte load
    xb[kb] = valid ? *reinterpret_cast<const s4_t*>(x + (size_t)body * NIN + 16 * kb +
                                                    4 * (lane >> 4))
                   : s4_t{0, 0, 0, 0};
  }
  const size_t zoff = (size_t)body * NH;
  const size_t lstride = (size_t)nb * NH;
  s4_t h1[4], h2[4], h3[4];
#pragma unroll
  for (int rt = 0; rt < 4; ++rt) {
    f4_t acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma(W.w1[rt][0], xb[0], acc);
    acc = mfma(W.w1[rt][1], xb[1], acc);
    h1[rt] = activate(acc, bias, rt, lane,
                      (zs && valid) ? zs + zoff : nullptr,
                      (hs && valid) ? hs + zoff : nullptr);
  }
#pragma unroll
  for (int rt = 0; rt < 4; ++rt) {
    f4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) acc = mfma(W.w2[rt][kb], h1[kb], acc);
    h2[rt] = activate(acc, bias + NH, rt, lane,
                      (zs && valid) ? zs + lstride + zoff : nullptr,
                      (hs && valid) ? hs + lstride + zoff : nullptr);
  }
#pragma unroll
  for (int rt = 0; rt < 4; ++rt) {
    f4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) acc = mfma(W.w3[rt][kb], h2[kb], acc);
    h3[rt] = activate(acc, bias + 2 * NH, rt, lane,
                      (zs && valid) ? zs + 2 * lstride + zoff : nullptr,
                      (hs && valid) ? hs + 2 * lstride + zoff : nullptr);
  }
  f4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) acc = mfma(W.w4[kb], h3[kb], acc);
  // rows 0..2 live in lanes 0..15 (4(l/16) + i with l/16 == 0)
#pragma unroll
  for (int i = 0; i < 3; ++i) out3[i] = acc[i] + bias[3 * NH + i];
}

__global__ __launch_bounds__(64 * WAVES) void k_mlp_forward(
    int nb, const unsigned short* __restrict__ x,
    const unsigned short* __restrict__ wbf, const float* __restrict__ params,
    float scale, int relative, float* __restrict__ out, double* __restrict__ a0,
    double* __restrict__ a1, double* __restrict__ a2, const double* __restrict__ v0,
    const double* __restrict__ v1, const double* __restrict__ v2, int n_proj,
    float* __restrict__ zs, unsigned short* __restrict__ hs) {
  __shared__ float sbias[3 * NH + NOUT];
  for (int i = threadIdx.x; i < 3 * NH + NOUT; i += 64 * WAVES) {
    const int off = i < NH ? OFF_B1 + i
                  : i < 2 * NH ? OFF_B2 + i - NH
                  : i < 3 * NH ? OFF_B3 + i - 2 * NH
                  : OFF_B4 + i - 3 * NH;
    sbias[i] = params[off];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * WAVES + (threadIdx.x >> 6);
  const int nwaves = gridDim.x * WAVES;
  Weights W;
  load_weights(wbf, lane, W);
  // one body's result: store / add the (projected) correction
  auto emit = [&](int body, const float o[3]) {
    if (out) {
      out[(size_t)body * 3 + 0] = o[0];
      out[(size_t)body * 3 + 1] = o[1];
      out[(size_t)body * 3 + 2] = o[2];
    }
    if (a0) {
      // relative: the output is in units of scale * |a| of the force it
      // corrects (read before the update)
      const double u = relative ? (double)scale * sqrt(a0[body] * a0[body] +
                                                       a1[body] * a1[body] +
                                                       a2[body] * a2[body])
                                : (double)scale;
      double d0 = u * (double)o[0], d1 = u * (double)o[1], d2 = u * (double)o[2];
      if (body < n_proj) {
        // no component along the velocity: the correction does no work
        const double w0 = v0[body], w1 = v1[body], w2 = v2[body];
        const double vv = w0 * w0 + w1 * w1 + w2 * w2;
        if (vv > 0.0) {
          const double c = (d0 * w0 + d1 * w1 + d2 * w2) / vv;
          d0 -= c * w0;
          d1 -= c * w1;
          d2 -= c * w2;
        }
      }
      a0[body] += d0;
      a1[body] += d1;
      a2[body] += d2;
    }
  };
  for (int b0 = wave * 16; b0 < nb; b0 += nwaves * 16) {
    float o[3];
    bool valid;
    forward16(W, sbias, x, nb, b0, lane, o, valid, zs, hs);
    if (lane < 16 && valid) emit(b0 + lane, o);
  }
}

// ---- inference ----
// The hidden layers carry t = -log2(e) z instead of z: sigmoid(z) =
// 1 / (1 + 2^t) is one v_exp_f32 and one v_rcp_f32, and the activation
// kept (BF16, rounded by v_cvt_pk_bf16_f32) is h' = t sigmoid(z) =
// -log2(e) silu(z).  The next layer's W h' + b' with b' = -log2(e) b is then
// exactly its own t (log2(e) ln(2) = 1), so the weights stay as trained:
// layer 1 scales its accumulator once, and layer 4 gives
// out = -ln(2) (W4 h3') + b4.  Biases enter as the MFMA's C operand.
typedef __bf16 bf2_t __attribute__((ext_vector_type(2)));
typedef float f2_t __attribute__((ext_vector_type(2)));

constexpr int IU = 4;  // 16-body column blocks per iteration (independent MFMA chains)
constexpr float NEG_LOG2E = -1.4426950408889634f;
constexpr float NEG_LN2 = -0.6931471805599453f;

__device__ __forceinline__ s4_t act_t(f4_t t) {
  float h[4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
    h[i] = t[i] * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(t[i]));
  const bf2_t lo = __builtin_convertvector((f2_t{h[0], h[1]}), bf2_t);
  const bf2_t hi = __builtin_convertvector((f2_t{h[2], h[3]}), bf2_t);
  const unsigned int ul = __builtin_bit_cast(unsigned int, lo);
  const unsigned int uh = __builtin_bit_cast(unsigned int, hi);
  s4_t out;
  out.x = (short)(ul & 0xFFFFu);
  out.y = (short)(ul >> 16);
  out.z = (short)(uh & 0xFFFFu);
  out.w = (short)(uh >> 16);
  return out;
}

typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));
__device__ __forceinline__ f4_t mfma32(bf8_t a, bf8_t b, f4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ bf8_t cat8(s4_t lo, s4_t hi) {
  typedef short s8_t __attribute__((ext_vector_type(8)));
  const s8_t v = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
  return __builtin_bit_cast(bf8_t, v);
}
// The inference kernel uses v_mfma_f32_16x16x32_bf16 (half the MFMA count of
// the K = 16 form; lane l: A row l%16 / B column l%16, k 8(l/16)+i, i < 8;
// accumulator as above).  A tiles: 8 BF16 of one weight row.
// The B operand of a hidden layer is the pair of accumulator tiles of row
// blocks 2j, 2j+1 (lane l: units 32j + 4(l/16) + i and 32j + 16 + 4(l/16) + i),
// so the weight tile takes its columns in that order.
struct Weights32 {
  bf8_t w1[4], w2[4][2], w3[4][2], w4[2];
};
__device__ __forceinline__ bf8_t load_a_pair(const unsigned short* __restrict__ w, int K,
                                             int rt, int j, int lane) {
  const unsigned short* p = w + (size_t)(16 * rt + (lane & 15)) * K + 32 * j + 4 * (lane >> 4);
  const s4_t lo = *reinterpret_cast<const s4_t*>(p);
  const s4_t hi = *reinterpret_cast<const s4_t*>(p + 16);
  return cat8(lo, hi);
}
__device__ __forceinline__ void load_weights32(const unsigned short* __restrict__ wb,
                                               int lane, Weights32& W) {
#pragma unroll
  for (int rt = 0; rt < 4; ++rt) {
    const unsigned short* p = wb + OFF_W1 + (size_t)(16 * rt + (lane & 15)) * NIN + 8 * (lane >> 4);
    W.w1[rt] = *reinterpret_cast<const bf8_t*>(p);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      W.w2[rt][j] = load_a_pair(wb + OFF_W2, NH, rt, j, lane);
      W.w3[rt][j] = load_a_pair(wb + OFF_W3, NH, rt, j, lane);
    }
  }
#pragma unroll
  for (int j = 0; j < 2; ++j) W.w4[j] = load_a_pair(wb + OFF_W4, NH, 0, j, lane);
}

template <int U>
__device__ __forceinline__ void infer16(const Weights32& W, const float* __restrict__ sb,
                                        const unsigned short* __restrict__ x, int nb,
                                        int b0, int lane, float out3[U][3], bool valid[U]) {
  bf8_t xb[U];
  s4_t h1[U][4], h2[U][4], h3[U][4];
  const int g4 = 4 * (lane >> 4);
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int body = b0 + 16 * u + (lane & 15);
    valid[u] = body < nb;
    const int bl = valid[u] ? body : 0;
    typedef short s8_t __attribute__((ext_vector_type(8)));
    const s8_t v = *reinterpret_cast<const s8_t*>(x + (size_t)bl * NIN + 8 * (lane >> 4));
    xb[u] = __builtin_bit_cast(bf8_t, valid[u] ? v : s8_t{0, 0, 0, 0, 0, 0, 0, 0});
  }
#pragma unroll
  for (int rt = 0; rt < 4; ++rt) {
    const f4_t b = *reinterpret_cast<const f4_t*>(sb + 16 * rt + g4);
#pragma unroll
    for (int u = 0; u < U; ++u) h1[u][rt] = act_t(mfma32(W.w1[rt], xb[u], b) * NEG_LOG2E);
  }
#pragma unroll
  for (int rt = 0; rt < 4; ++rt) {
    const f4_t b = *reinterpret_cast<const f4_t*>(sb + NH + 16 * rt + g4);
    f4_t acc[U];
#pragma unroll
    for (int u = 0; u < U; ++u) acc[u] = b;
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int u = 0; u < U; ++u)
        acc[u] = mfma32(W.w2[rt][j], cat8(h1[u][2 * j], h1[u][2 * j + 1]), acc[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) h2[u][rt] = act_t(acc[u]);
  }
#pragma unroll
  for (int rt = 0; rt < 4; ++rt) {
    const f4_t b = *reinterpret_cast<const f4_t*>(sb + 2 * NH + 16 * rt + g4);
    f4_t acc[U];
#pragma unroll
    for (int u = 0; u < U; ++u) acc[u] = b;
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int u = 0; u < U; ++u)
        acc[u] = mfma32(W.w3[rt][j], cat8(h2[u][2 * j], h2[u][2 * j + 1]), acc[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) h3[u][rt] = act_t(acc[u]);
  }
  f4_t acc[U];
#pragma unroll
  for (int u = 0; u < U; ++u) acc[u] = f4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int u = 0; u < U; ++u) acc[u] = mfma32(W.w4[j], cat8(h3[u][2 * j], h3[u][2 * j + 1]), acc[u]);
#pragma unroll
  for (int u = 0; u < U; ++u)
#pragma unroll
    for (int i = 0; i < 3; ++i) out3[u][i] = fmaf(acc[u][i], NEG_LN2, sb[3 * NH + i]);
}

// ADD: add the (scaled, projected) correction to a as k_mlp_forward does;
// otherwise only store the raw outputs.
template <bool ADD>
__global__ __launch_bounds__(64 * WAVES) void k_mlp_infer(
    int nb, const unsigned short* __restrict__ x, const unsigned short* __restrict__ wbf,
    const float* __restrict__ params, float scale, int relative, float* __restrict__ out,
    double* __restrict__ a0, double* __restrict__ a1, double* __restrict__ a2,
    const double* __restrict__ v0, const double* __restrict__ v1,
    const double* __restrict__ v2, int n_proj) {
  __shared__ __align__(16) float sb[3 * NH + NOUT];
  for (int i = threadIdx.x; i < 3 * NH + NOUT; i += 64 * WAVES) {
    const int off = i < NH ? OFF_B1 + i
                  : i < 2 * NH ? OFF_B2 + i - NH
                  : i < 3 * NH ? OFF_B3 + i - 2 * NH
                  : OFF_B4 + i - 3 * NH;
    const float b = params[off];
    sb[i] = (i >= NH && i < 3 * NH) ? NEG_LOG2E * b : b;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * WAVES + (threadIdx.x >> 6);
  const int nwaves = gridDim.x * WAVES;
  Weights32 W;
  load_weights32(wbf, lane, W);
  for (int b0 = wave * 16 * IU; b0 < nb; b0 += nwaves * 16 * IU) {
    float o[IU][3];
    bool valid[IU];
    infer16<IU>(W, sb, x, nb, b0, lane, o, valid);
    if (lane >= 16) continue;
#pragma unroll
    for (int u = 0; u < IU; ++u) {
      if (!valid[u]) continue;
      const int body = b0 + 16 * u + lane;
      if (out) {
        out[(size_t)body * 3 + 0] = o[u][0];
        out[(size_t)body * 3 + 1] = o[u][1];
        out[(size_t)body * 3 + 2] = o[u][2];
      }
      if (ADD) {
        // relative: the output is in units of scale * |a| of the force it
        // corrects (read before the update)
        const double un = relative ? (double)scale * sqrt(a0[body] * a0[body] +
                                                          a1[body] * a1[body] +
                                                          a2[body] * a2[body])
                                   : (double)scale;
        double d0 = un * (double)o[u][0], d1 = un * (double)o[u][1],
               d2 = un * (double)o[u][2];
        if (body < n_proj) {
          // no component along the velocity: the correction does no work
          const double w0 = v0[body], w1 = v1[body], w2 = v2[body];
          const double vv = w0 * w0 + w1 * w1 + w2 * w2;
          if (vv > 0.0) {
            const double c = (d0 * w0 + d1 * w1 + d2 * w2) / vv;
            d0 -= c * w0;
            d1 -= c * w1;
            d2 -= c * w2;
          }
        }
        a0[body] += d0;
        a1[body] += d1;
        a2[body] += d2;
      }
    }
  }
}

// Backward through the layers for 16 bodies: given the stored activations
// and dL/dout (3 per body), writes delta_l (BF16, [body][64]) for l = 1..3
// (pre-activation gradients) used by the weight-gradient kernel.
__global__ __launch_bounds__(64 * WAVES) void k_mlp_backward(
    int nb, const unsigned short* __restrict__ wbf,
    const float* __restrict__ zs, const float* __restrict__ gout,
    unsigned short* __restrict__ deltas) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * WAVES + (threadIdx.x >> 6);
  const int nwaves = gridDim.x * WAVES;
  // transposed weights as A operands: A[row = in][k = out] = W[out][in]
  s4_t w4t[4], w3t[4][4], w2t[4][4];
  for (int rt = 0; rt < 4; ++rt) {
    const int in = 16 * rt + (lane & 15);
    short v[4];
    for (int i = 0; i < 4; ++i)
      v[i] = (short)wbf[OFF_W4 + (size_t)(4 * (lane >> 4) + i) * NH + in];
    w4t[rt] = s4_t{v[0], v[1], v[2], v[3]};
    for (int kb = 0; kb < 4; ++kb) {
      short u[4], q[4];
      for (int i = 0; i < 4; ++i) {
        const int o = 16 * kb + 4 * (lane >> 4) + i;
        u[i] = (short)wbf[OFF_W3 + (size_t)o * NH + in];
        q[i] = (short)wbf[OFF_W2 + (size_t)o * NH + in];
      }
      w3t[rt][kb] = s4_t{u[0], u[1], u[2], u[3]};
      w2t[rt][kb] = s4_t{q[0], q[1], q[2], q[3]};
    }
  }
  const size_t lstride = (size_t)nb * NH;
  for (int b0 = wave * 16; b0 < nb; b0 += nwaves * 16) {
    const int body = b0 + (lane & 15);
    const bool valid = body < nb;
    // delta4^T: rows = outputs (0..2 used), lane l holds rows 4(l/16)+i
    short g4[4];
    for (int i = 0; i < 4; ++i) {
      const int r = 4 * (lane >> 4) + i;
      const float g = (valid && r < NOUT_USED) ? gout[(size_t)body * 3 + r] : 0.f;
      g4[i] = (short)bf16_bits(g);
    }
    const s4_t d4 = s4_t{g4[0], g4[1], g4[2], g4[3]};
    s4_t d3[4], d2[4];
    const float* z3 = zs + 2 * lstride + (size_t)body * NH;
    const float* z2 = zs + lstride + (size_t)body * NH;
    const float* z1 = zs + (size_t)body * NH;
    for (int rt = 0; rt < 4; ++rt) {
      f4_t acc = {0.f, 0.f, 0.f, 0.f};
      acc = mfma(w4t[rt], d4, acc);
      short o[4];
      for (int i = 0; i < 4; ++i) {
        const int r = 16 * rt + 4 * (lane >> 4) + i;
        const float d = valid ? acc[i] * silu_grad(z3[r]) : 0.f;
        const unsigned short db = bf16_bits(d);
        o[i] = (short)db;
        if (valid) deltas[2 * lstride + (size_t)body * NH + r] = db;
      }
      d3[rt] = s4_t{o[0], o[1], o[2], o[3]};
    }
    for (int rt = 0; rt < 4; ++rt) {
      f4_t acc = {0.f, 0.f, 0.f, 0.f};
      for (int kb = 0; kb < 4; ++kb) acc = mfma(w3t[rt][kb], d3[kb], acc);
      short o[4];
      for (int i = 0; i < 4; ++i) {
        const int r = 16 * rt + 4 * (lane >> 4) + i;
        const float d = valid ? acc[i] * silu_grad(z2[r]) : 0.f;
        const unsigned short db = bf16_bits(d);
        o[i] = (short)db;
        if (valid) deltas[lstride + (size_t)body * NH + r] = db;
      }
      d2[rt] = s4_t{o[0], o[1], o[2], o[3]};
    }
    for (int rt = 0; rt < 4; ++rt) {
      f4_t acc = {0.f, 0.f, 0.f, 0.f};
      for (int kb = 0; kb < 4; ++kb) acc = mfma(w2t[rt][kb], d2[kb], acc);
      for (int i = 0; i < 4; ++i) {
        const int r = 16 * rt + 4 * (lane >> 4) + i;
        const float d = valid ? acc[i] * silu_grad(z1[r]) : 0.f;
        if (valid) deltas[(size_t)body * NH + r] = bf16_bits(d);
      }
    }
  }
}

// Weight gradients of one layer, dW[out][in] = sum_b d[b][out] h[b][in],
// db[out] = sum_b d[b][out].  Workgroup g takes bodies [g*chunk, ...), each
// of its waves every WAVES-th 16-body block; wave partials go to
// part[g*WAVES + w][(out*K + in) | bias] and are summed in a fixed order by
// k_mlp_wgrad_reduce.  d: BF16 [nb][dstride] (FP32 for the output layer),
// h: BF16 [nb][K], the layer's inputs.
template <int NO, int K, bool DFLOAT>
__global__ __launch_bounds__(64 * WAVES) void k_mlp_wgrad(
    int nb, int chunk, const void* __restrict__ dv, int dstride,
    const unsigned short* __restrict__ h, float* __restrict__ part) {
  constexpr int RT = (NO + 15) / 16, CT = K / 16;
  constexpr int TOTAL = RT * 16 * K + RT * 16;
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int c0 = blockIdx.x * chunk;
  const int c1 = min(nb, c0 + chunk);
  f4_t acc[RT][CT];
  float dbias[RT];
#pragma unroll
  for (int rt = 0; rt < RT; ++rt) {
    dbias[rt] = 0.f;
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) acc[rt][ct] = f4_t{0.f, 0.f, 0.f, 0.f};
  }
  for (int b0 = c0 + 16 * w; b0 < c1; b0 += 16 * WAVES) {
    // A[row = out][k = body]: lane holds 4 bodies of one output
    // B[k = body][col = in]:  lane holds 4 bodies of one input
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
      const int o = 16 * rt + (lane & 15);
      short av[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int b = b0 + 4 * (lane >> 4) + i;
        float d = 0.f;
        if (b < c1 && o < NO) {
          if (DFLOAT)
            d = static_cast<const float*>(dv)[(size_t)b * dstride + o];
          else
            d = bf16_float(static_cast<const unsigned short*>(dv)[(size_t)b * dstride + o]);
        }
        av[i] = (short)bf16_bits(d);
        dbias[rt] += bf16_float((unsigned short)av[i]);
      }
      const s4_t a = s4_t{av[0], av[1], av[2], av[3]};
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) {
        const int in = 16 * ct + (lane & 15);
        short bv[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int b = b0 + 4 * (lane >> 4) + i;
          bv[i] = (b < c1) ? (short)h[(size_t)b * K + in] : (short)0;
        }
        acc[rt][ct] = mfma(a, s4_t{bv[0], bv[1], bv[2], bv[3]}, acc[rt][ct]);
      }
    }
  }
  float* r = part + (size_t)(blockIdx.x * WAVES + w) * TOTAL;
#pragma unroll
  for (int rt = 0; rt < RT; ++rt) {
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int o = 16 * rt + 4 * (lane >> 4) + i;
        const int in = 16 * ct + (lane & 15);
        r[o * K + in] = acc[rt][ct][i];
      }
    // bias: lanes l, l^16, l^32, l^48 hold the same output
    float s = dbias[rt];
    s += __shfl_xor(s, 16);
    s += __shfl_xor(s, 32);
    if (lane < 16) r[RT * 16 * K + 16 * rt + lane] = s;
  }
}

// Sum the chunk partials in order and scatter into the gradient block.
__global__ __launch_bounds__(MB) void k_mlp_wgrad_reduce(
    int nchunks, int total, int NOv, int K, int off_w, int off_b,
    const float* __restrict__ part, float* __restrict__ grad) {
  const int e = blockIdx.x * MB + threadIdx.x;
  if (e >= total) return;
  float s = 0.f;
  for (int c = 0; c < nchunks; ++c) s += part[(size_t)c * total + e];
  const int wsz = ((NOv + 15) / 16) * 16 * K;
  if (e < wsz) {
    const int o = e / K, in = e % K;
    if (o < NOv) grad[off_w + o * K + in] = s;
  } else {
    const int o = e - wsz;
    if (o < NOv) grad[off_b + o] = s;
  }
}

// Adam step on the FP32 master parameters, then refresh the BF16 copy.
__global__ __launch_bounds__(MB) void k_adam(int n, float* __restrict__ p,
                                             const float* __restrict__ g,
                                             float* __restrict__ m,
                                             float* __restrict__ v, float lr,
                                             float b1, float b2, float eps,
                                             float c1, float c2,
                                             unsigned short* __restrict__ wbf) {
  const int i = blockIdx.x * MB + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float upd = lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  const float np = p[i] - upd;
  p[i] = np;
  wbf[i] = bf16_bits(np);
}

__global__ __launch_bounds__(MB) void k_to_bf16(int n, const float* __restrict__ p,
                                                unsigned short* __restrict__ wbf) {
  const int i = blockIdx.x * MB + threadIdx.x;
  if (i < n) wbf[i] = bf16_bits(p[i]);
}

// Features of body p: the 27 node masses (/ m_unit) of the 3x3x3 stencil
// around its nearest mesh node, then its offset from that node in cells
// (times 2, in [-1, 1]), 1 for a mass (p < n_mass, its own mass is in the
// stencil) or 0 for a tracer, and a zero; BF16 [nb][32].
template <typename M>
__global__ __launch_bounds__(MB) void k_mlp_features(
    int nb, int n_mass, const int* __restrict__ order, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z, int n, double ox,
    double oy, double oz, double h, const M* __restrict__ rho, double inv_unit,
    unsigned short* __restrict__ feat) {
  const int i_ = blockIdx.x * MB + threadIdx.x;
  if (i_ >= nb) return;
  const int p = order ? order[i_] : i_;
  const double f[3] = {(x[p] - ox) / h, (y[p] - oy) / h, (z[p] - oz) / h};
  int c[3];
  float off[3];
  for (int a = 0; a < 3; ++a) {
    const double r = floor(f[a] + 0.5);
    c[a] = min(max((int)r, 0), n - 1);
    off[a] = (float)(2.0 * (f[a] - r));
  }
  // branch-free stencil: nodes outside the mesh are read clamped and
  // count as 0, so all 27 loads can be in flight at once
  int ix[3][3];
  bool ok[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int i = c[a] + d - 1;
      ok[a][d] = i >= 0 && i < n;
      ix[a][d] = ok[a][d] ? i : 0;
    }
  M m[27];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int d = 0; d < 3; ++d)
        m[9 * a + 3 * b + d] = rho[((size_t)ix[0][a] * n + ix[1][b]) * n + ix[2][d]];
  // 32 BF16 values packed into 16 words, written as four 16-byte stores
  unsigned int wd[16];
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    const bool in = ok[0][q / 9] && ok[1][(q / 3) % 3] && ok[2][q % 3];
    const double v = in ? (double)m[q] : 0.0;
    const unsigned int b = bf16_bits((float)(v * inv_unit));
    if (q & 1) wd[q >> 1] |= b << 16; else wd[q >> 1] = b;
  }
  wd[13] |= (unsigned int)bf16_bits(off[0]) << 16;                 // 27
  wd[14] = bf16_bits(off[1]) | ((unsigned int)bf16_bits(off[2]) << 16);
  wd[15] = p < n_mass ? 0x3F80u : 0u;                              // 30, 31
  uint4* o = reinterpret_cast<uint4*>(feat + (size_t)p * NIN);
  for (int v = 0; v < 4; ++v)
    o[v] = make_uint4(wd[4 * v], wd[4 * v + 1], wd[4 * v + 2], wd[4 * v + 3]);
}

#ifndef SSIM_MLP_WG_CAP
#define SSIM_MLP_WG_CAP 1024
#endif
int grid_waves(int nb) {
  // workgroups: one per 16-body batch per wave, at most SSIM_MLP_WG_CAP:
  // every wave then loads its 88 registers of weights once for several
  // batches
  const int batches = ssim::ceil_div(nb, 16);
  int wg = ssim::ceil_div(batches, WAVES);
  return wg < 1 ? 1 : (wg > SSIM_MLP_WG_CAP ? SSIM_MLP_WG_CAP : wg);
}

}  // namespace

extern "C" {

int ssim_mlp_param_count(void) { return NPARAM; }

int ssim_mlp_to_bf16(const float* params, unsigned short* wbf, void* stream) {
  SSIM_REQUIRE(params && wbf, "null pointer");
  hipLaunchKernelGGL(k_to_bf16, dim3(ssim::ceil_div(NPARAM, MB)), dim3(MB), 0,
                     ssim::as_stream(stream), NPARAM, params, wbf);
  SSIM_LAUNCHED("k_to_bf16");
  return SSIM_OK;
}

int ssim_mlp_features(int nb, int n_mass, const int* order, const double* x,
                      const double* y, const double* z, int mesh_n,
                      const double* origin, double h,
                      int mesh_prec, const void* rho, double mass_unit,
                      unsigned short* feat, void* stream) {
  SSIM_REQUIRE(nb >= 0 && mesh_n >= 1 && h > 0.0 && mass_unit > 0.0,
               "bad feature arguments");
  if (nb == 0) return SSIM_OK;
  SSIM_REQUIRE(x && y && z && origin && rho && feat, "null pointer");
  const bool f64 = mesh_prec == SSIM_PREC_F64;
  hipStream_t st = ssim::as_stream(stream);
  if (f64)
    SSIM_TIMED("k_mlp_features", st,
               hipLaunchKernelGGL(k_mlp_features<double>, dim3(ssim::ceil_div(nb, MB)),
                                  dim3(MB), 0, st, nb, n_mass, order, x, y, z, mesh_n,
                                  origin[0], origin[1], origin[2], h,
                                  static_cast<const double*>(rho), 1.0 / mass_unit, feat));
  else
    SSIM_TIMED("k_mlp_features", st,
               hipLaunchKernelGGL(k_mlp_features<float>, dim3(ssim::ceil_div(nb, MB)),
                                  dim3(MB), 0, st, nb, n_mass, order, x, y, z, mesh_n,
                                  origin[0], origin[1], origin[2], h,
                                  static_cast<const float*>(rho), 1.0 / mass_unit, feat));
  SSIM_LAUNCHED("k_mlp_features");
  return SSIM_OK;
}

int ssim_mlp_forward(int nb, const unsigned short* feat,
                     const unsigned short* wbf, const float* params,
                     float scale, int relative, float* out, double* ax,
                     double* ay, double* az, const double* vx, const double* vy,
                     const double* vz, int n_proj, float* zs, unsigned short* hs,
                     void* stream) {
  SSIM_REQUIRE(nb >= 0, "nb must be >= 0");
  if (nb == 0) return SSIM_OK;
  SSIM_REQUIRE(feat && wbf && params, "null pointer");
  SSIM_REQUIRE(!ax || (ay && az), "need all three acceleration arrays");
  SSIM_REQUIRE((zs == nullptr) == (hs == nullptr), "zs and hs go together");
  SSIM_REQUIRE(n_proj >= 0 && n_proj <= nb, "n_proj must be in [0, nb]");
  SSIM_REQUIRE(n_proj == 0 || (ax && vx && vy && vz),
               "the projection needs the accelerations and velocities");
  hipStream_t st = ssim::as_stream(stream);
  if (zs == nullptr) {
    const int grid = grid_waves(ssim::ceil_div(nb, IU));  // 16 IU bodies per wave
    if (ax)
      SSIM_TIMED("k_mlp_forward", st,
                 hipLaunchKernelGGL(k_mlp_infer<true>, dim3(grid), dim3(64 * WAVES), 0, st,
                                    nb, feat, wbf, params, scale, relative, out, ax, ay,
                                    az, vx, vy, vz, n_proj));
    else
      SSIM_TIMED("k_mlp_forward", st,
                 hipLaunchKernelGGL(k_mlp_infer<false>, dim3(grid), dim3(64 * WAVES), 0,
                                    st, nb, feat, wbf, params, scale, relative, out, ax,
                                    ay, az, vx, vy, vz, n_proj));
    SSIM_LAUNCHED("k_mlp_infer");
    return SSIM_OK;
  }
  SSIM_TIMED("k_mlp_forward", st,
             hipLaunchKernelGGL(k_mlp_forward, dim3(grid_waves(nb)),
                                dim3(64 * WAVES), 0, st, nb, feat, wbf, params,
                                scale, relative, out, ax, ay, az, vx, vy, vz,
                                n_proj, zs, hs));
  SSIM_LAUNCHED("k_mlp_forward");
  return SSIM_OK;
}

size_t ssim_mlp_train_work_bytes(int nb) {
  const size_t nbs = (size_t)(nb > 0 ? nb : 1);
  const size_t nchunks = (size_t)ssim::ceil_div(nbs, 2048);
  const size_t part = nchunks * WAVES * (size_t)(64 * 64 + 64) * sizeof(float);
  return ssim::align256(3 * nbs * NH * 2) + ssim::align256(part);
}

int ssim_mlp_backward(int nb, const unsigned short* feat,
                      const unsigned short* wbf, const float* zs,
                      const unsigned short* hs, const float* gout, float* grad,
                      void* work, size_t work_bytes, void* stream) {
  SSIM_REQUIRE(nb >= 1, "empty batch");
  SSIM_REQUIRE(feat && wbf && zs && hs && gout && grad && work, "null pointer");
  SSIM_REQUIRE(work_bytes >= ssim_mlp_train_work_bytes(nb), "workspace too small");
  hipStream_t st = ssim::as_stream(stream);
  unsigned short* deltas = static_cast<unsigned short*>(work);
  float* part = reinterpret_cast<float*>(static_cast<unsigned char*>(work) +
                                         ssim::align256((size_t)3 * nb * NH * 2));
  hipLaunchKernelGGL(k_mlp_backward, dim3(grid_waves(nb)), dim3(64 * WAVES), 0,
                     st, nb, wbf, zs, gout, deltas);
  SSIM_LAUNCHED("k_mlp_backward");
  const int chunk = 2048;
  const int nch = ssim::ceil_div(nb, chunk);
  const size_t lstride = (size_t)nb * NH;
  // layer 4: d = gout (FP32, 3 used), h = H3
  {
    const int total = 16 * NH + 16;
    hipLaunchKernelGGL((k_mlp_wgrad<NOUT_USED, NH, true>), dim3(nch),
                       dim3(64 * WAVES), 0, st, nb, chunk, gout, 3,
                       hs + 2 * lstride, part);
    hipLaunchKernelGGL(k_mlp_wgrad_reduce, dim3(ssim::ceil_div(total, MB)),
                       dim3(MB), 0, st, nch * WAVES, total, NOUT_USED, NH, OFF_W4,
                       OFF_B4, part, grad);
  }
  // layers 3, 2: d = delta_l, h = H_{l-1}
  const int total_h = NH * NH + NH;
  hipLaunchKernelGGL((k_mlp_wgrad<NH, NH, false>), dim3(nch), dim3(64 * WAVES),
                     0, st, nb, chunk, deltas + 2 * lstride, NH, hs + lstride,
                     part);
  hipLaunchKernelGGL(k_mlp_wgrad_reduce, dim3(ssim::ceil_div(total_h, MB)),
                     dim3(MB), 0, st, nch * WAVES, total_h, NH, NH, OFF_W3, OFF_B3, part,
                     grad);
  hipLaunchKernelGGL((k_mlp_wgrad<NH, NH, false>), dim3(nch), dim3(64 * WAVES),
                     0, st, nb, chunk, deltas + lstride, NH, hs, part);
  hipLaunchKernelGGL(k_mlp_wgrad_reduce, dim3(ssim::ceil_div(total_h, MB)),
                     dim3(MB), 0, st, nch * WAVES, total_h, NH, NH, OFF_W2, OFF_B2, part,
                     grad);
  // layer 1: d = delta_1, h = features
  const int total_1 = NH * NIN + NH;
  hipLaunchKernelGGL((k_mlp_wgrad<NH, NIN, false>), dim3(nch), dim3(64 * WAVES),
                     0, st, nb, chunk, deltas, NH, feat, part);
  hipLaunchKernelGGL(k_mlp_wgrad_reduce, dim3(ssim::ceil_div(total_1, MB)),
                     dim3(MB), 0, st, nch * WAVES, total_1, NH, NIN, OFF_W1, OFF_B1,
                     part, grad);
  SSIM_LAUNCHED("k_mlp_wgrad");
  return SSIM_OK;
}

int ssim_mlp_adam(float* params, const float* grad, float* m, float* v,
                  float lr, float beta1, float beta2, float eps, int step,
                  unsigned short* wbf, void* stream) {
  SSIM_REQUIRE(params && grad && m && v && wbf && step >= 1, "bad Adam args");
  const float c1 = 1.f - powf(beta1, (float)step);
  const float c2 = 1.f - powf(beta2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(ssim::ceil_div(NPARAM, MB)), dim3(MB), 0,
                     ssim::as_stream(stream), NPARAM, params, grad, m, v, lr,
                     beta1, beta2, eps, c1, c2, wbf);
  SSIM_LAUNCHED("k_adam");
  return SSIM_OK;
}

}  // extern "C"

namespace {

// gout = 2 (out - y) / nb  and per-block partial sums of (out - y)^2.
__global__ __launch_bounds__(MB) void k_mse_grad(int nb, const float* __restrict__ out,
                                                 const float* __restrict__ y,
                                                 float* __restrict__ gout,
                                                 float* __restrict__ loss_part) {
  __shared__ float red[MB];
  const int i = blockIdx.x * MB + threadIdx.x;
  float l = 0.f;
  if (i < nb) {
    for (int c = 0; c < 3; ++c) {
      const float d = out[(size_t)i * 3 + c] - y[(size_t)i * 3 + c];
      gout[(size_t)i * 3 + c] = 2.0f * d / (float)nb;
      l += d * d;
    }
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int w = MB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

}  // namespace

namespace {
__global__ __launch_bounds__(MB) void k_residual(int nb, const double* __restrict__ r0,
                                                 const double* __restrict__ r1,
                                                 const double* __restrict__ r2,
                                                 const double* __restrict__ m0,
                                                 const double* __restrict__ m1,
                                                 const double* __restrict__ m2,
                                                 double unit, int relative,
                                                 double clip, float* __restrict__ out) {
  const int i = blockIdx.x * MB + threadIdx.x;
  if (i >= nb) return;
  double u = unit;
  if (relative) u *= sqrt(m0[i] * m0[i] + m1[i] * m1[i] + m2[i] * m2[i]);
  const double d[3] = {r0[i] - m0[i], r1[i] - m1[i], r2[i] - m2[i]};
  for (int c = 0; c < 3; ++c) {
    double v = u > 0.0 ? d[c] / u : 0.0;
    if (clip > 0.0) v = fmin(fmax(v, -clip), clip);
    out[(size_t)i * 3 + c] = (float)v;
  }
}
}  // namespace

extern "C" int ssim_mlp_residual(int nb, const double* ref_x, const double* ref_y,
                                 const double* ref_z, const double* mod_x,
                                 const double* mod_y, const double* mod_z,
                                 double a_unit, int relative, double clip,
                                 float* out, void* stream) {
  SSIM_REQUIRE(nb >= 0 && a_unit > 0.0 && clip >= 0.0, "bad arguments");
  if (nb == 0) return SSIM_OK;
  SSIM_REQUIRE(ref_x && ref_y && ref_z && mod_x && mod_y && mod_z && out,
               "null pointer");
  hipLaunchKernelGGL(k_residual, dim3(ssim::ceil_div(nb, MB)), dim3(MB), 0,
                     ssim::as_stream(stream), nb, ref_x, ref_y, ref_z, mod_x,
                     mod_y, mod_z, a_unit, relative, clip, out);
  SSIM_LAUNCHED("k_residual");
  return SSIM_OK;
}

extern "C" int ssim_mlp_mse_grad(int nb, const float* out, const float* target,
                                 float* gout, float* loss_part, void* stream) {
  SSIM_REQUIRE(nb >= 1 && out && target && gout && loss_part, "bad arguments");
  hipLaunchKernelGGL(k_mse_grad, dim3(ssim::ceil_div(nb, MB)), dim3(MB), 0,
                     ssim::as_stream(stream), nb, out, target, gout, loss_part);
  SSIM_LAUNCHED("k_mse_grad");
  return SSIM_OK;
}
