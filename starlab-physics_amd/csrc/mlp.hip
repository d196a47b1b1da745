// Learned warm start: a per-vertex MLP (24 -> 64 -> 64 -> 64 -> 3, GELU) that
// predicts the converged Jacobi correction y* - y_pred from 1-ring features, on
// BF16 MFMA (v_mfma_f32_32x32x16_bf16) with FP32 accumulation.  No reference
// counterpart: the reference has no learned component; the plain Jacobi solve
// (framework/physics/XPBD.py:140-184, 383-387) stays the judge of the residual.
//
// Features of vertex v (FP32, s = the mesh's mean rest length), in order:
//   0-2   (y_v - x_v)/s                 predicted displacement
//   3-5   v_v dt / s                    previous displacement
//   6-11  stretch strains (|y_v-y_u| - l0)/l0 of the first 6 stretch entries
//   12-17 bending strains (|x_v-x_u| - l0)/l0 of the first 6 bending entries
//   18-20 sum over stretch entries of strain * unit(y_v - y_u)
//   21-23 sum over bending entries of strain * unit(x_v - x_u)
// normalised by per-feature mean / std of the training set and clamped to +-8.
//
// Parameters (one FP32 master array, one BF16 working copy), K and N padded to
// the MFMA tile: W1[64][32] b1[64] W2[64][64] b2[64] W3[64][64] b3[64]
// W4[32][64] b4[32] (weights stored [out][in]); pads are zero and stay zero.
//
// MFMA 32x32x16 operand layout (lane l): A row l%32, k = 8*(l/32)+i;
// B column l%32, k = 8*(l/32)+i; D element j: row (j/4)*8 + (l/32)*4 + j%4,
// column l%32.
#include <math.h>

#include "slx_common.h"

namespace slx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kIn = 24, kInP = 32, kH = 64, kOut = 3, kOutP = 32;
constexpr int oW1 = 0, ob1 = oW1 + kH * kInP, oW2 = ob1 + kH, ob2 = oW2 + kH * kH, oW3 = ob2 + kH,
              ob3 = oW3 + kH * kH, oW4 = ob3 + kH, ob4 = oW4 + kOutP * kH, kParams = ob4 + kOutP;
constexpr int kFeat = 24;

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
  return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

// C[32x32] += A[32 x 16ks] B[16ks x 32];  A(m,k) = A[m*ars + k*aks], B(k,n) = B[k*bks + n*bcs]
__device__ __forceinline__ void mfma_tile(const __bf16* A, int ars, int aks, const __bf16* B, int bks, int bcs,
                                          int ksteps, f32x16& c) {
  const int l = threadIdx.x & 63, r = l & 31, kh = (l >> 5) * 8;
  for (int s = 0; s < ksteps; ++s) {
    bf16x8 a, b;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int k = s * 16 + kh + i;
      a[i] = A[r * ars + k * aks];
      b[i] = B[k * bks + r * bcs];
    }
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
  }
}
__device__ __forceinline__ int d_row(int j) { return (j >> 2) * 8 + ((threadIdx.x & 63) >> 5) * 4 + (j & 3); }
__device__ __forceinline__ int d_col() { return threadIdx.x & 31; }

// forward of one 32-row tile; act0 [32][32] bf16 in LDS; writes a1..a3 (bf16) and
// z1..z3 (fp32, nullable) to LDS; returns the output tile (col < 3 meaningful)
struct FwdLds {
  __bf16* a0;   // [32][32]
  __bf16* a1;   // [32][64]
  __bf16* a2;
  __bf16* a3;
  float* z1;    // [32][64] or null
  float* z2;
  float* z3;
};

// biases in LDS: bias[0..63] b1, [64..127] b2, [128..191] b3, [192..223] b4
constexpr int kBias = 3 * kH + kOutP;

__device__ void mlp_forward_tile(const __bf16* W, const float* bias, FwdLds L, f32x16& out) {
  const __bf16* Wl[3] = {W + oW1, W + oW2, W + oW3};
  const float* bl[3] = {bias, bias + kH, bias + 2 * kH};
  const __bf16* ain[3] = {L.a0, L.a1, L.a2};
  __bf16* aout[3] = {L.a1, L.a2, L.a3};
  float* zout[3] = {L.z1, L.z2, L.z3};
  const int K[3] = {kInP, kH, kH};
  for (int layer = 0; layer < 3; ++layer) {
    for (int nb = 0; nb < 2; ++nb) {
      f32x16 c = {};
      mfma_tile(ain[layer], K[layer], 1, Wl[layer] + nb * 32 * K[layer], 1, K[layer], K[layer] / 16, c);
      const int col = d_col() + 32 * nb;
      const float b = bl[layer][col];
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int row = d_row(j);
        float z = c[j] + b;
        if (zout[layer]) zout[layer][row * kH + col] = z;
        aout[layer][row * kH + col] = (__bf16)gelu(z);
      }
    }
    __syncthreads();
  }
  f32x16 c = {};
  mfma_tile(L.a3, kH, 1, W + oW4, 1, kH, kH / 16, c);
  const float b = bias[3 * kH + d_col()];
#pragma unroll
  for (int j = 0; j < 16; ++j) c[j] += b;
  out = c;
}

// BF16 weights (16-byte loads: kParams is a multiple of 8 and both arrays are
// 16-byte aligned) and FP32 biases (from the master copy) into LDS
static_assert(kParams % 8 == 0, "weights are loaded eight bf16 at a time");
__device__ __forceinline__ void load_weights(__bf16* sW, float* sB, const __bf16* W, const float* Wf) {
  const uint4* src = reinterpret_cast<const uint4*>(W);
  uint4* dst = reinterpret_cast<uint4*>(sW);
  for (int i = threadIdx.x; i < kParams / 8; i += blockDim.x) dst[i] = src[i];
  for (int i = threadIdx.x; i < kH; i += blockDim.x) {
    sB[i] = Wf[ob1 + i];
    sB[kH + i] = Wf[ob2 + i];
    sB[2 * kH + i] = Wf[ob3 + i];
  }
  for (int i = threadIdx.x; i < kOutP; i += blockDim.x) sB[3 * kH + i] = Wf[ob4 + i];
}

// normalised input, clamped to +-kClip: the strain features are heavy-tailed (a
// few vertices near pins sit tens of standard deviations out)
constexpr float kClip = 8.0f;
__device__ __forceinline__ float norm_feature(float f, float mean, float istd) {
  return fminf(fmaxf((f - mean) * istd, -kClip), kClip);
}

// ---- features ----
__device__ void vertex_features(int v, const float4* x4, const float4* v4, const float4* y4, const int32_t* soff,
                                const int32_t* swid, const int2* ent, float dt, float inv_s, float* f) {
  const int s = v >> 6, lane = v & 63, base = soff[s] + lane, width = swid[s];
  const float4 xv = x4[v], yv = y4[v], vv = v4[v];
  f[0] = (yv.x - xv.x) * inv_s; f[1] = (yv.y - xv.y) * inv_s; f[2] = (yv.z - xv.z) * inv_s;
  f[3] = vv.x * dt * inv_s; f[4] = vv.y * dt * inv_s; f[5] = vv.z * dt * inv_s;
  for (int i = 6; i < kFeat; ++i) f[i] = 0.f;
  int ns = 0, nb = 0;
  for (int k = 0; k < width; ++k) {
    int2 e = ent[base + 64 * k];
    if (e.x == -1) break;
    const bool bend = e.x < 0;
    const int u = e.x & 0x7fffffff;
    const float l0 = __int_as_float(e.y);
    float3 d = bend ? sub3(xyz(xv), xyz(x4[u])) : sub3(xyz(yv), xyz(y4[u]));
    float len = sqrtf(dot3(d, d));
    float strain = (len - l0) / l0;
    float3 n = len > 0.f ? mul3(1.0f / len, d) : make_float3(0.f, 0.f, 0.f);
    if (!bend) {
      if (ns < 6) f[6 + ns] = strain;
      ++ns;
      f[18] += strain * n.x; f[19] += strain * n.y; f[20] += strain * n.z;
    } else {
      if (nb < 6) f[12 + nb] = strain;
      ++nb;
      f[21] += strain * n.x; f[22] += strain * n.y; f[23] += strain * n.z;
    }
  }
}

__global__ void k_features(int nv, const float4* __restrict__ x4, const float4* __restrict__ v4,
                           const float4* __restrict__ y4, const int32_t* __restrict__ soff,
                           const int32_t* __restrict__ swid, const int2* __restrict__ ent, float dt, float inv_s,
                           float* __restrict__ feat) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  float f[kFeat];
  vertex_features(v, x4, v4, y4, soff, swid, ent, dt, inv_s, f);
  for (int i = 0; i < kFeat; ++i) feat[(size_t)v * kFeat + i] = f[i];
}

// ---- inference: features -> MLP -> y = y_pred + f * out * out_scale ----
// Four wavefronts per workgroup share the weights in LDS (loaded once per
// workgroup) and each takes its own 32-vertex tile with its own activation
// buffers; the workgroup steps over groups of four tiles so every wavefront
// reaches the forward's workgroup barriers the same number of times.
constexpr int kInfWaves = 4;
__global__ void __launch_bounds__(64 * kInfWaves)
k_warm_start(int nv, const float4* __restrict__ x4, const float4* __restrict__ v4, const float4* __restrict__ y4,
             float4* __restrict__ y_out, const float* __restrict__ fixed, const int32_t* __restrict__ soff,
             const int32_t* __restrict__ swid, const int2* __restrict__ ent, float dt, float inv_s,
             const __bf16* __restrict__ W, const float* __restrict__ Wf, const float* __restrict__ fmean,
             const float* __restrict__ fistd, float out_scale) {
  __shared__ __align__(16) __bf16 sW[kParams];
  __shared__ float sB[kBias];
  __shared__ __bf16 a0[kInfWaves][32 * kInP];
  __shared__ __bf16 a1[kInfWaves][32 * kH], a2[kInfWaves][32 * kH];
  __shared__ float so[kInfWaves][32 * 4];
  load_weights(sW, sB, W, Wf);
  __syncthreads();
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int ntiles = (nv + 31) / 32, ngroups = (ntiles + kInfWaves - 1) / kInfWaves;
  for (int gi = blockIdx.x; gi < ngroups; gi += gridDim.x) {
    const int v = (gi * kInfWaves + w) * 32 + (l & 31);
    if (l < 32) {
      float f[kFeat];
      if (v < nv) vertex_features(v, x4, v4, y4, soff, swid, ent, dt, inv_s, f);
      else for (int i = 0; i < kFeat; ++i) f[i] = 0.f;
      for (int i = 0; i < kInP; ++i)
        a0[w][l * kInP + i] = (__bf16)(i < kFeat && v < nv ? norm_feature(f[i], fmean[i], fistd[i]) : 0.f);
    }
    __syncthreads();
    // the third layer's output reuses a1 (the second layer has read it by then)
    FwdLds L{a0[w], a1[w], a2[w], a1[w], nullptr, nullptr, nullptr};
    f32x16 out;
    mlp_forward_tile(sW, sB, L, out);
    if (d_col() < kOut) {
#pragma unroll
      for (int j = 0; j < 16; ++j) so[w][d_row(j) * 4 + d_col()] = out[j];
    }
    __syncthreads();
    if (l < 32 && v < nv) {
      const float fv = fixed[v];
      const float4 y = y4[v];
      y_out[v] = make_float4(y.x + fv * (so[w][l * 4 + 0] * out_scale), y.y + fv * (so[w][l * 4 + 1] * out_scale),
                             y.z + fv * (so[w][l * 4 + 2] * out_scale), y.w);
    }
    __syncthreads();
  }
}

// ---- training: forward + backward over tiles of 32 samples, per-block gradient partials ----
__global__ void __launch_bounds__(64)
k_train_grad(int n, int tiles_per_block, const float* __restrict__ feat, const float* __restrict__ target,
             const __bf16* __restrict__ W, const float* __restrict__ Wf, const float* __restrict__ fmean,
             const float* __restrict__ fistd, float inv_n, float* __restrict__ part, float* __restrict__ loss_part) {
  __shared__ __align__(16) __bf16 sW[kParams];
  __shared__ float sB[kBias];
  __shared__ float g[kParams];                  // this block's gradient
  __shared__ __bf16 a0[32 * kInP];
  __shared__ __bf16 a1[32 * kH], a2[32 * kH], a3[32 * kH];
  __shared__ float z[3][32 * kH];
  __shared__ __bf16 dz[32 * kH];                // dL/dz of the current layer (bf16 operand)
  __shared__ float dzf[32 * kH];                // same, FP32 (bias gradient)
  load_weights(sW, sB, W, Wf);
  for (int i = threadIdx.x; i < kParams; i += 64) g[i] = 0.f;
  __syncthreads();
  const int l = threadIdx.x;
  float loss = 0.f;
  for (int tt = 0; tt < tiles_per_block; ++tt) {
    const int t = blockIdx.x * tiles_per_block + tt;
    const int m0 = t * 32;
    if (m0 >= n) break;
    if (l < 32) {
      const int m = m0 + l;
      for (int i = 0; i < kInP; ++i)
        a0[l * kInP + i] =
            (__bf16)(i < kFeat && m < n ? norm_feature(feat[(size_t)m * kFeat + i], fmean[i], fistd[i]) : 0.f);
    }
    __syncthreads();
    FwdLds L{a0, a1, a2, a3, z[0], z[1], z[2]};
    f32x16 out;
    mlp_forward_tile(sW, sB, L, out);
    // dL/dout = 2 (out - target) / n on the 3 real outputs
    {
      const int col = d_col();
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int row = d_row(j), m = m0 + row;
        float gval = 0.f;
        if (col < kOut && m < n) {
          float diff = out[j] - target[(size_t)m * kOut + col];
          loss += diff * diff;
          gval = 2.0f * diff * inv_n;
        }
        dzf[row * kH + col] = gval;
        dz[row * kH + col] = (__bf16)gval;
      }
    }
    __syncthreads();
    // layer 4: dW4[n][k] += sum_m dz[m][n] a3[m][k]  (N = 32, K = 64); db4
    for (int kb = 0; kb < 2; ++kb) {
      f32x16 c = {};
      mfma_tile(dz, 1, kH, a3 + kb * 32, kH, 1, 2, c);
#pragma unroll
      for (int j = 0; j < 16; ++j) g[oW4 + d_row(j) * kH + kb * 32 + d_col()] += c[j];
    }
    if (l < kOutP) {
      float s = 0.f;
      for (int m = 0; m < 32; ++m) s += dzf[m * kH + l];
      g[ob4 + l] += s;
    }
    // da3[m][k] = sum_n dz[m][n] W4[n][k]  -> dz3 = da3 * gelu'(z3)
    f32x16 da[2];
    for (int kb = 0; kb < 2; ++kb) {
      da[kb] = f32x16{};
      mfma_tile(dz, kH, 1, sW + oW4 + kb * 32, kH, 1, 2, da[kb]);
    }
    __syncthreads();
    for (int layer = 2; layer >= 0; --layer) {
      const __bf16* a_in = layer == 0 ? a0 : (layer == 1 ? a1 : a2);
      for (int kb = 0; kb < 2; ++kb)
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          const int row = d_row(j), col = d_col() + 32 * kb;
          float v = da[kb][j] * gelu_grad(z[layer][row * kH + col]);
          dzf[row * kH + col] = v;
          dz[row * kH + col] = (__bf16)v;
        }
      __syncthreads();
      const int K = layer == 0 ? kInP : kH;
      const int woff = layer == 0 ? oW1 : (layer == 1 ? oW2 : oW3);
      const int boff = layer == 0 ? ob1 : (layer == 1 ? ob2 : ob3);
      // dW[n][k] += sum_m dz[m][n] a_in[m][k]  (N = 64, K)
      for (int nb = 0; nb < 2; ++nb)
        for (int kb = 0; kb < K / 32; ++kb) {
          f32x16 c = {};
          mfma_tile(dz + nb * 32, 1, kH, a_in + kb * 32, K, 1, 2, c);
#pragma unroll
          for (int j = 0; j < 16; ++j) g[woff + (nb * 32 + d_row(j)) * K + kb * 32 + d_col()] += c[j];
        }
      {
        float s = 0.f;
        for (int m = 0; m < 32; ++m) s += dzf[m * kH + l];
        g[boff + l] += s;
      }
      if (layer > 0) {
        // da_in[m][k] = sum_n dz[m][n] W[n][k]  (K = 64)
        for (int kb = 0; kb < 2; ++kb) {
          da[kb] = f32x16{};
          mfma_tile(dz, kH, 1, sW + woff + kb * 32, K, 1, 4, da[kb]);
        }
      }
      __syncthreads();
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kParams; i += 64) part[(size_t)blockIdx.x * kParams + i] = g[i];
  // deterministic loss partial: per-lane partial sums reduced in lane order
  __shared__ float sl[64];
  sl[l] = loss;
  __syncthreads();
  if (l == 0) {
    float s = 0.f;
    for (int i = 0; i < 64; ++i) s += sl[i];
    loss_part[blockIdx.x] = s * inv_n;
  }
}

// sum of the block partials in block order, then Adam on the FP32 master weights
// and the BF16 working copy
__global__ void k_adam(int nblocks, const float* __restrict__ part, float* __restrict__ Wf, __bf16* __restrict__ W,
                       float* __restrict__ m1, float* __restrict__ m2, float lr, float b1, float b2, float eps,
                       float bc1, float bc2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kParams) return;
  float gsum = 0.f;
  for (int b = 0; b < nblocks; ++b) gsum += part[(size_t)b * kParams + i];
  float a = b1 * m1[i] + (1.f - b1) * gsum;
  float v = b2 * m2[i] + (1.f - b2) * gsum * gsum;
  m1[i] = a;
  m2[i] = v;
  float w = Wf[i] - lr * (a / bc1) / (sqrtf(v / bc2) + eps);
  Wf[i] = w;
  W[i] = (__bf16)w;
}

__global__ void k_sum_grad(int nblocks, const float* __restrict__ part, float* __restrict__ grad) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kParams) return;
  float s = 0.f;
  for (int b = 0; b < nblocks; ++b) s += part[(size_t)b * kParams + i];
  grad[i] = s;
}

__global__ void k_to_bf16(int n, const float* __restrict__ a, __bf16* __restrict__ b) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (__bf16)a[i];
}

// max_v |y_a - y_b|_inf over vertices (exact: max is order-independent)
__global__ void k_max_update(int nv, const float4* __restrict__ a, const float4* __restrict__ b,
                             unsigned int* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  float m = 0.f;
  if (i < nv) {
    float4 p = a[i], q = b[i];
    m = fmaxf(fmaxf(fabsf(p.x - q.x), fabsf(p.y - q.y)), fabsf(p.z - q.z));
  }
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));   // non-negative floats order as uints
}

}  // namespace slx

using namespace slx;

extern "C" int slx_mlp_num_params(void) { return kParams; }

extern "C" int slx_features(int nv, const float* x4, const float* v4, const float* y4, const int32_t* slice_off,
                            const int32_t* slice_width, const int32_t* entries, float dt, float inv_s, float* feat,
                            slx_stream_t stream) {
  SLX_REQUIRE(nv >= 0, "slx_features: nv < 0");
  if (nv == 0) return SLX_OK;
  SLX_REQUIRE(x4 && v4 && y4 && slice_off && slice_width && entries && feat, "slx_features: null pointer");
  SLX_LAUNCH("features", k_features, dim3(blocks_for(nv, 256)), dim3(256), 0, (hipStream_t)stream, nv,
             (const float4*)x4, (const float4*)v4, (const float4*)y4, slice_off, slice_width, (const int2*)entries, dt,
             inv_s, feat);
  return check_launch("slx_features");
}

extern "C" int slx_warm_start(int nv, const float* x4, const float* v4, const float* y4_pred, float* y4_out,
                              const float* fixed, const int32_t* slice_off, const int32_t* slice_width,
                              const int32_t* entries, float dt, float inv_s, const void* w_bf16, const float* w_f32,
                              const float* fmean, const float* fistd, float out_scale, slx_stream_t stream) {
  SLX_REQUIRE(nv >= 0, "slx_warm_start: nv < 0");
  if (nv == 0) return SLX_OK;
  SLX_REQUIRE(x4 && v4 && y4_pred && y4_out && fixed && slice_off && slice_width && entries && w_bf16 && w_f32 &&
                  fmean && fistd,
              "slx_warm_start: null pointer");
  SLX_REQUIRE(y4_pred != y4_out, "slx_warm_start: y4_pred and y4_out must differ");
  const int groups = ((nv + 31) / 32 + kInfWaves - 1) / kInfWaves;
  const int grid = groups < 512 ? groups : 512;
  SLX_LAUNCH("warm_start", k_warm_start, dim3(grid), dim3(64 * kInfWaves), 0, (hipStream_t)stream, nv, (const float4*)x4,
             (const float4*)v4, (const float4*)y4_pred, (float4*)y4_out, fixed, slice_off, slice_width,
             (const int2*)entries, dt, inv_s,
             (const __bf16*)w_bf16, w_f32, fmean, fistd, out_scale);
  return check_launch("slx_warm_start");
}

extern "C" int slx_mlp_grad(int n, const float* feat, const float* target, const void* w_bf16, const float* w_f32,
                            const float* fmean, const float* fistd, int n_blocks, float* part, float* loss_part,
                            float* grad, slx_stream_t stream) {
  SLX_REQUIRE(n > 0 && n_blocks > 0, "slx_mlp_grad: bad sizes");
  SLX_REQUIRE(feat && target && w_bf16 && w_f32 && fmean && fistd && part && loss_part, "slx_mlp_grad: null pointer");
  int tiles = (n + 31) / 32;
  int tpb = (tiles + n_blocks - 1) / n_blocks;
  hipStream_t st = (hipStream_t)stream;
  SLX_LAUNCH("mlp_grad", k_train_grad, dim3(n_blocks), dim3(64), 0, st, n, tpb, feat, target, (const __bf16*)w_bf16,
             w_f32, fmean, fistd, 1.0f / (float)n, part, loss_part);
  if (grad)
    SLX_LAUNCH("mlp_sum", k_sum_grad, dim3(blocks_for(kParams, 256)), dim3(256), 0, st, n_blocks, part, grad);
  return check_launch("slx_mlp_grad");
}

extern "C" int slx_mlp_adam(int n_blocks, const float* part, float* w_f32, void* w_bf16, float* m1, float* m2,
                            float lr, float beta1, float beta2, float eps, int step, slx_stream_t stream) {
  SLX_REQUIRE(n_blocks > 0 && step >= 1, "slx_mlp_adam: bad sizes");
  SLX_REQUIRE(part && w_f32 && w_bf16 && m1 && m2, "slx_mlp_adam: null pointer");
  float bc1 = 1.f - powf(beta1, (float)step), bc2 = 1.f - powf(beta2, (float)step);
  SLX_LAUNCH("mlp_adam", k_adam, dim3(blocks_for(kParams, 256)), dim3(256), 0, (hipStream_t)stream, n_blocks, part,
             w_f32, (__bf16*)w_bf16, m1, m2, lr, beta1, beta2, eps, bc1, bc2);
  return check_launch("slx_mlp_adam");
}

extern "C" int slx_mlp_to_bf16(const float* w_f32, void* w_bf16, slx_stream_t stream) {
  SLX_REQUIRE(w_f32 && w_bf16, "slx_mlp_to_bf16: null pointer");
  SLX_LAUNCH("to_bf16", k_to_bf16, dim3(blocks_for(kParams, 256)), dim3(256), 0, (hipStream_t)stream, kParams, w_f32,
             (__bf16*)w_bf16);
  return check_launch("slx_mlp_to_bf16");
}

extern "C" int slx_max_update(int nv, const float* a4, const float* b4, uint32_t* out, slx_stream_t stream) {
  SLX_REQUIRE(nv >= 0 && (nv == 0 || (a4 && b4 && out)), "slx_max_update: bad arguments");
  if (nv == 0) return SLX_OK;
  SLX_LAUNCH("max_update", k_max_update, dim3(blocks_for(nv, 256)), dim3(256), 0, (hipStream_t)stream, nv,
             (const float4*)a4, (const float4*)b4, out);
  return check_launch("slx_max_update");
}
