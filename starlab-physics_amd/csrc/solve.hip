// Solve-to-residual passes for the time-to-residual workload (plain and
// warm-started).  No reference counterpart: the reference runs one Jacobi pass
// per substep (framework/physics/XPBD.py:443-469), and repeating that pass is not
// a convergent iteration (its bending term acts on x, which does not change within
// a substep, so every repeat adds the same bending displacement again).
//
// The residual solve takes the reference pass's result y1 as the inertial anchor
// and solves the substep's implicit-Euler stretch problem
//     minimise  sum_v |y_v - y1_v|^2 / (2 w_v)  +  sum_e kappa C_e(y)^2 / 2
// (w_v = fixed * m_inv, kappa = stretch stiffness * dt^2, C_e = |y_u - y_v| - l0_e,
// the stretch constraints of solve_spring_constraints_x, XPBD.py:146-167) by
// Jacobi iteration with the diagonal Gauss-Newton step
//     y_v <- y_v - f_v (y_v - y1_v + w_v g_v) / (1 + w_v sum_e kappa),
//     g_v  = sum over the vertex's stretch constraints of kappa C_e n_e,
// n_e the unit vector from the neighbour to v.  The residual of an iterate is
// the largest step, max_v |step_v|_inf, zero exactly at the solution.
//
// Pass k reads y_k and writes y_{k+1}; it also max-reduces the residual of y_k
// into res[k] (float bits; res starts at zero).  A pass first looks at the
// previous one: if res[k-1] <= tol, or conv already names an earlier pass, it
// records conv = min(conv, k-1) and does nothing else, so a chain of passes
// launched without host synchronisation stops by itself.
#include <limits.h>

#include "slx_common.h"

namespace slx {

__global__ void __launch_bounds__(256)
k_implicit_pass(int nv, const float* __restrict__ fixed, const float4* __restrict__ y_anchor,
                const float4* __restrict__ y_in, float4* __restrict__ y_out, const int32_t* __restrict__ slice_off,
                const int32_t* __restrict__ slice_width, const int2* __restrict__ ent, float kappa, int k, float tol,
                uint32_t* __restrict__ res, int32_t* __restrict__ conv) {
  if (k > 0) {
    const int c = conv[0];
    const bool done = c < k || __uint_as_float(res[k - 1]) <= tol;
    if (done) {
      if (blockIdx.x == 0 && threadIdx.x == 0 && c > k - 1) atomicMin(conv, k - 1);
      return;
    }
  }
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  float rmax = 0.f;
  if (v < nv) {
    const int s = v >> 6, lane = v & 63;
    const int base = slice_off[s] + lane, width = slice_width[s];
    const float4 yv4 = y_in[v];
    const float3 yv = xyz(yv4);
    float3 gs = make_float3(0.f, 0.f, 0.f);
    float ks = 0.f;
    for (int j = 0; j < width; ++j) {
      const int2 e = ent[base + 64 * j];
      if (e.x == -1) break;
      if (e.x < 0) continue;                      // bending entry: not part of this solve
      const float3 pu = xyz(y_in[e.x]);
      const float l0 = __int_as_float(e.y);
      const float3 d = sub3(yv, pu);
      const float len = sqrtf(dot3(d, d));
      const float C = len - l0;
      const float inv = 1.0f / len;
      const float3 n = mul3(inv, d);
      gs = add3(gs, mul3(kappa * C, n));
      ks = ks + kappa;
    }
    const float w = yv4.w;
    const float3 ya = xyz(y_anchor[v]);
    const float den = 1.0f + w * ks;
    const float f = fixed[v];
    const float3 step = make_float3(f * (((yv.x - ya.x) + w * gs.x) / den), f * (((yv.y - ya.y) + w * gs.y) / den),
                                    f * (((yv.z - ya.z) + w * gs.z) / den));
    rmax = fmaxf(fmaxf(fabsf(step.x), fabsf(step.y)), fabsf(step.z));
    y_out[v] = make_float4(yv.x - step.x, yv.y - step.y, yv.z - step.z, yv4.w);
  }
  // wavefront max, one atomic per wavefront (non-negative floats order as their bits)
  for (int o = 32; o > 0; o >>= 1) rmax = fmaxf(rmax, __shfl_xor(rmax, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(&res[k], __float_as_uint(rmax));
}

__global__ void k_solve_reset(int n_res, uint32_t* __restrict__ res, int32_t* __restrict__ conv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_res) res[i] = 0u;
  if (i == 0) conv[0] = INT_MAX;
}

}  // namespace slx

using namespace slx;

extern "C" int slx_solve_reset(int n_res, uint32_t* res, int32_t* conv, slx_stream_t stream) {
  SLX_REQUIRE(n_res >= 0 && res && conv, "slx_solve_reset: bad arguments");
  SLX_LAUNCH("solve_reset", k_solve_reset, dim3(blocks_for(n_res > 0 ? n_res : 1, 256)), dim3(256), 0,
             (hipStream_t)stream, n_res, res, conv);
  return check_launch("slx_solve_reset");
}

extern "C" int slx_implicit_passes(int nv, const float* fixed, const float* y_anchor, float* ya, float* yb,
                                   const int32_t* slice_off, const int32_t* slice_width, const int32_t* entries,
                                   float kappa, int k0, int count, float tol, uint32_t* res, int32_t* conv,
                                   slx_stream_t stream) {
  SLX_REQUIRE(nv >= 0 && k0 >= 0 && count >= 0, "slx_implicit_passes: bad sizes");
  SLX_REQUIRE(nv == 0 || (fixed && y_anchor && ya && yb && slice_off && slice_width && entries && res && conv),
              "slx_implicit_passes: null pointer");
  if (nv == 0) return SLX_OK;
  hipStream_t st = (hipStream_t)stream;
  for (int j = 0; j < count; ++j) {
    const int k = k0 + j;
    const float4* in = (const float4*)((k & 1) ? yb : ya);
    float4* out = (float4*)((k & 1) ? ya : yb);
    SLX_LAUNCH("implicit_pass", k_implicit_pass, dim3(blocks_for(nv, 256)), dim3(256), 0, st, nv, fixed,
               (const float4*)y_anchor, in, out, slice_off, slice_width, (const int2*)entries, kappa, k, tol, res,
               conv);
  }
  return check_launch("slx_implicit_passes");
}
