// XPBD Jacobi solve on gfx950: predict, stretch + bending distance constraints
// gathered per vertex through the SELL-64 adjacency (no float atomics), Jacobi
// averaging, velocity and position update.
//
// Reference kernels: framework/physics/XPBD.py compute_y (115-118),
// solve_spring_constraints_x (140-184), update_dx (383-387), compute_velocity
// (340-343), update_x (332-337), set_fixed_vertices (389-395).
//
// Memory layout (HBM): float4 per vertex for x/y/v (one 16-byte load per
// gather), SELL-64 adjacency: slice s of 64 consecutive vertices stores entry k
// of its rows at slice_off[s] + 64*k + lane, so each load instruction of a
// wavefront reads 64 consecutive 8-byte entries (512 contiguous bytes).
#include "slx_common.h"

namespace slx {

// ---- per-vertex formulas, op order identical to the serial restatement ----

// y = x + f*v*dt + g*dt*dt  (XPBD.py:118)
__device__ __forceinline__ float3 predict(float4 x, float4 v, float dt, float3 gdt2) {
  float f = v.w;
  return make_float3((x.x + (f * v.x) * dt) + gdt2.x, (x.y + (f * v.y) * dt) + gdt2.y,
                     (x.z + (f * v.z) * dt) + gdt2.z);
}

// Contribution of one distance constraint to vertex v (XPBD.py:150-162 seen from
// either endpoint): d = p_v - p_u, n = d/|d|, S = (fw_v + fw_u) n.n,
// ld = k C / (k S + 1); the vertex's delta decreases by (fw_v ld) n.
__device__ __forceinline__ float3 constraint_term(float3 pv, float3 pu, float fwv, float fwu,
                                                  float l0, float k) {
  float3 d = sub3(pv, pu);
  float len = sqrtf(dot3(d, d));
  float C = len - l0;
  float inv = 1.0f / len;
  float3 n = mul3(inv, d);
  float schur = (fwv + fwu) * dot3(n, n);
  float ld = (k * C) / ((k * schur) + 1.0f);
  return mul3(fwv * ld, n);
}

// y + f*(dx/nc) where nc > 0 (XPBD.py:386-387)
__device__ __forceinline__ float3 jacobi_apply(float3 y, float3 dx, float nc, float f) {
  if (nc > 0.0f)
    return make_float3(y.x + f * (dx.x / nc), y.y + f * (dx.y / nc), y.z + f * (dx.z / nc));
  return y;
}

// v = (1-damping) f (y-x) / dt ; x += dt v   (XPBD.py:343, 337)
__device__ __forceinline__ void finalize(float3 y, float4 x, float f, float damping, float dt,
                                         float4& xo, float4& vo) {
  float a = 1.0f - damping;
  float af = a * f;
  float vx = (af * (y.x - x.x)) / dt, vy = (af * (y.y - x.y)) / dt, vz = (af * (y.z - x.z)) / dt;
  xo = make_float4(x.x + dt * vx, x.y + dt * vy, x.z + dt * vz, x.w);
  vo = make_float4(vx, vy, vz, f);
}

__global__ void k_compute_y(int nv, const float4* __restrict__ x4, const float4* __restrict__ v4,
                            float4* __restrict__ y4, float dt, float3 gdt2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  float4 x = x4[i];
  float3 y = predict(x, v4[i], dt, gdt2);
  y4[i] = make_float4(y.x, y.y, y.z, x.w);
}

__global__ void k_update_vx(int nv, float4* __restrict__ x4, float4* __restrict__ v4,
                            const float4* __restrict__ y4, const float* __restrict__ fixed,
                            float damping, float dt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  float4 xo, vo;
  finalize(xyz(y4[i]), x4[i], fixed[i], damping, dt, xo, vo);
  x4[i] = xo;
  v4[i] = vo;
}

__global__ void k_set_fixed(int nv, const float* __restrict__ fixed, const float* __restrict__ m_inv,
                            float4* x4, float4* v4, float4* y4) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  float f = fixed[i];
  float fw = f * m_inv[i];
  x4[i].w = fw;
  if (y4) y4[i].w = fw;
  v4[i].w = f;
}

// One Jacobi spring pass for vertex v.
//  PRED:  neighbour/own positions for stretch are predicted on the fly from x4/v4
//         (the first pass of a substep, compute_y fused in);
//         otherwise they are read from y_in.
//  FINAL: the result is turned into the substep's velocity and position
//         (compute_velocity + update_x fused in) and written to x_out/v_out;
//         otherwise it is written to y_out.
template <bool PRED, bool FINAL>
__device__ __forceinline__ void spring_vertex(int v, const float4* __restrict__ x4, const float4* __restrict__ v4,
                                              const float* __restrict__ fixed, const float4* __restrict__ y_in,
                                              float4* __restrict__ y_out, float4* __restrict__ x_out,
                                              float4* __restrict__ v_out, const int32_t* __restrict__ slice_off,
                                              const int32_t* __restrict__ slice_width, const int2* __restrict__ ent,
                                              float k_s, float k_b, float dt, float3 gdt2, float damping) {
  const int s = v >> 6, lane = v & 63;
  const int base = slice_off[s] + lane, width = slice_width[s];
  const float4 xv = x4[v];
  float3 yv;
  if constexpr (PRED) yv = predict(xv, v4[v], dt, gdt2);
  else yv = xyz(y_in[v]);
  const float fwv = xv.w;
  float3 dx = make_float3(0.f, 0.f, 0.f);
  float nc = 0.f;
  // Entries are consumed in groups of 4: all entry loads, then all neighbour
  // gathers, then the arithmetic in entry order, so four independent gathers are
  // in flight per lane.  Rows are left-aligned: an entry of -1 ends the row.
  constexpr int G = 4;
  for (int k0 = 0; k0 < width; k0 += G) {
    int2 e[G];
#pragma unroll
    for (int q = 0; q < G; ++q) e[q] = (k0 + q < width) ? ent[base + 64 * (k0 + q)] : make_int2(-1, 0);
    float4 pa[G], pb[G];
#pragma unroll
    for (int q = 0; q < G; ++q) {
      const int u = e[q].x == -1 ? v : (e[q].x & 0x7fffffff);
      if (e[q].x < 0) {            // bending (or padding): x of the neighbour (XPBD.py:169)
        pa[q] = x4[u];
      } else if constexpr (PRED) { // stretch on the predicted y, from x and v
        pa[q] = x4[u];
        pb[q] = v4[u];
      } else {                     // stretch on y (XPBD.py:150)
        pa[q] = y_in[u];
      }
    }
    bool done = false;
#pragma unroll
    for (int q = 0; q < G; ++q) {
      if (e[q].x == -1) {
        done = true;
        break;
      }
      const float l0 = __int_as_float(e[q].y);
      float3 pv, pu;
      float kk;
      if (e[q].x < 0) {
        pv = xyz(xv);
        pu = xyz(pa[q]);
        kk = k_b;
      } else {
        if constexpr (PRED) pu = predict(pa[q], pb[q], dt, gdt2);
        else pu = xyz(pa[q]);
        pv = yv;
        kk = k_s;
      }
      float3 t = constraint_term(pv, pu, fwv, pa[q].w, l0, kk);
      dx = sub3(dx, t);
      nc = nc + 1.0f;
    }
    if (done) break;
  }
  const float f = fixed[v];
  float3 yn = jacobi_apply(yv, dx, nc, f);
  if constexpr (FINAL) {
    float4 xo, vo;
    finalize(yn, xv, f, damping, dt, xo, vo);
    x_out[v] = xo;
    v_out[v] = vo;
  } else {
    y_out[v] = make_float4(yn.x, yn.y, yn.z, fwv);
  }
}

// VPT vertices per thread: block b covers [b*blockDim*VPT, (b+1)*blockDim*VPT), thread t
// takes t, t + blockDim, ... so every load instruction stays coalesced per wave.
template <bool PRED, bool FINAL, int VPT>
__global__ void __launch_bounds__(256)
k_spring(int nv, const float4* __restrict__ x4, const float4* __restrict__ v4,
         const float* __restrict__ fixed, const float4* __restrict__ y_in,
         float4* __restrict__ y_out, float4* __restrict__ x_out, float4* __restrict__ v_out,
         const int32_t* __restrict__ slice_off, const int32_t* __restrict__ slice_width,
         const int2* __restrict__ ent, float k_s, float k_b, float dt, float3 gdt2,
         float damping) {
  const int base = xcd_block() * (int)blockDim.x * VPT + (int)threadIdx.x;
#pragma unroll
  for (int j = 0; j < VPT; ++j) {
    const int v = base + j * (int)blockDim.x;
    if (v < nv)
      spring_vertex<PRED, FINAL>(v, x4, v4, fixed, y_in, y_out, x_out, v_out, slice_off, slice_width, ent, k_s, k_b,
                                 dt, gdt2, damping);
  }
}


int spring_launch(bool pred, bool final, int nv, const float4* x4, const float4* v4, const float* fixed,
                  const float4* y_in, float4* y_out, float4* x_out, float4* v_out, const int32_t* slice_off,
                  const int32_t* slice_width, const int32_t* entries, float k_s, float k_b, float dt, float gx,
                  float gy, float gz, float damping, hipStream_t st) {
  if (nv <= 0) return SLX_OK;
  const float3 gdt2 = make_float3((gx * dt) * dt, (gy * dt) * dt, (gz * dt) * dt);
  // launch shape (autotuned by the host: block x vertices/thread)
  const int bs = knobs().spring_block, vpt = knobs().spring_vpt;
  const dim3 grid(blocks_for(nv, bs * vpt)), blk(bs);
  const int2* E = (const int2*)entries;
#define SLX_SPRING(NAME, P, F, V)                                                                              \
  SLX_LAUNCH(NAME, (k_spring<P, F, V>), grid, blk, 0, st, nv, x4, v4, fixed, y_in, y_out, x_out, v_out, slice_off, \
             slice_width, E, k_s, k_b, dt, gdt2, damping)
#define SLX_SPRING_V(NAME, P, F)   \
  do {                             \
    if (vpt == 4)                  \
      SLX_SPRING(NAME, P, F, 4);   \
    else if (vpt == 2)             \
      SLX_SPRING(NAME, P, F, 2);   \
    else                           \
      SLX_SPRING(NAME, P, F, 1);   \
  } while (0)
  if (pred && final) SLX_SPRING_V("spring_pred_final", true, true);
  else if (pred) SLX_SPRING_V("spring_pred", true, false);
  else if (final) SLX_SPRING_V("spring_final", false, true);
  else SLX_SPRING_V("spring", false, false);
#undef SLX_SPRING_V
#undef SLX_SPRING
  return check_launch("k_spring");
}

}  // namespace slx

using namespace slx;

static inline hipStream_t S(slx_stream_t s) { return (hipStream_t)s; }

extern "C" int slx_compute_y(int nv, const float* x4, const float* v4, float* y4, float dt,
                             float gx, float gy, float gz, slx_stream_t stream) {
  SLX_REQUIRE(nv >= 0, "slx_compute_y: nv < 0");
  if (nv == 0) return SLX_OK;
  SLX_REQUIRE(x4 && v4 && y4, "slx_compute_y: null pointer");
  float3 gdt2 = make_float3((gx * dt) * dt, (gy * dt) * dt, (gz * dt) * dt);
  SLX_LAUNCH("compute_y", k_compute_y, dim3(blocks_for(nv, 256)), dim3(256), 0, S(stream), nv,
                     (const float4*)x4, (const float4*)v4, (float4*)y4, dt, gdt2);
  return check_launch("k_compute_y");
}

extern "C" int slx_update_vx(int nv, float* x4, float* v4, const float* y4, const float* fixed,
                             float damping, float dt, slx_stream_t stream) {
  SLX_REQUIRE(nv >= 0, "slx_update_vx: nv < 0");
  if (nv == 0) return SLX_OK;
  SLX_REQUIRE(x4 && v4 && y4 && fixed, "slx_update_vx: null pointer");
  SLX_LAUNCH("update_vx", k_update_vx, dim3(blocks_for(nv, 256)), dim3(256), 0, S(stream), nv,
                     (float4*)x4, (float4*)v4, (const float4*)y4, fixed, damping, dt);
  return check_launch("k_update_vx");
}

extern "C" int slx_set_fixed(int nv, const float* fixed, const float* m_inv, float* x4, float* v4,
                             float* y4, slx_stream_t stream) {
  SLX_REQUIRE(nv >= 0, "slx_set_fixed: nv < 0");
  if (nv == 0) return SLX_OK;
  SLX_REQUIRE(fixed && m_inv && x4 && v4, "slx_set_fixed: null pointer");
  SLX_LAUNCH("set_fixed", k_set_fixed, dim3(blocks_for(nv, 256)), dim3(256), 0, S(stream), nv, fixed,
                     m_inv, (float4*)x4, (float4*)v4, (float4*)y4);
  return check_launch("k_set_fixed");
}

extern "C" int slx_spring_pass(int nv, const float* x4, const float* fixed, const float* y4_in,
                               float* y4_out, const int32_t* slice_off,
                               const int32_t* slice_width, const int32_t* entries, float k_s,
                               float k_b, slx_stream_t stream) {
  SLX_REQUIRE(nv >= 0, "slx_spring_pass: nv < 0");
  if (nv == 0) return SLX_OK;
  SLX_REQUIRE(x4 && fixed && y4_in && y4_out && slice_off && slice_width && entries,
              "slx_spring_pass: null pointer");
  SLX_REQUIRE(y4_in != y4_out, "slx_spring_pass: y4_in and y4_out must differ (Jacobi)");
  return spring_launch(false, false, nv, (const float4*)x4, nullptr, fixed, (const float4*)y4_in, (float4*)y4_out,
                       nullptr, nullptr, slice_off, slice_width, entries, k_s, k_b, 0.f, 0.f, 0.f, 0.f, 0.f,
                       S(stream));
}

extern "C" int slx_substeps_plain(int nv, float* x4, float* v4, float* x4_alt, float* v4_alt,
                                  float* y4a, float* y4b, const float* fixed,
                                  const int32_t* slice_off, const int32_t* slice_width,
                                  const int32_t* entries, int n_substeps, int n_iters, float dt,
                                  float gx, float gy, float gz, float k_s, float k_b,
                                  float damping, slx_stream_t stream) {
  SLX_REQUIRE(nv >= 0 && n_substeps >= 0 && n_iters >= 1, "slx_substeps_plain: bad sizes");
  if (nv == 0 || n_substeps == 0) return SLX_OK;
  SLX_REQUIRE(x4 && v4 && x4_alt && v4_alt && fixed && slice_off && slice_width && entries,
              "slx_substeps_plain: null pointer");
  SLX_REQUIRE(n_iters == 1 || (y4a && y4b), "slx_substeps_plain: y buffers required");
  hipStream_t st = S(stream);
  float4 *cx = (float4*)x4, *cv = (float4*)v4, *ax = (float4*)x4_alt, *av = (float4*)v4_alt;
  float4 *ya = (float4*)y4a, *yb = (float4*)y4b;
  int rc = SLX_OK;
  for (int s = 0; s < n_substeps && rc == SLX_OK; ++s) {
    if (n_iters == 1) {
      rc = spring_launch(true, true, nv, cx, cv, fixed, nullptr, nullptr, ax, av, slice_off, slice_width, entries,
                         k_s, k_b, dt, gx, gy, gz, damping, st);
    } else {
      rc = spring_launch(true, false, nv, cx, cv, fixed, nullptr, ya, nullptr, nullptr, slice_off, slice_width,
                         entries, k_s, k_b, dt, gx, gy, gz, damping, st);
      for (int it = 1; it + 1 < n_iters && rc == SLX_OK; ++it) {
        rc = spring_launch(false, false, nv, cx, cv, fixed, ya, yb, nullptr, nullptr, slice_off, slice_width,
                           entries, k_s, k_b, dt, gx, gy, gz, damping, st);
        float4* t = ya; ya = yb; yb = t;
      }
      if (rc == SLX_OK)
        rc = spring_launch(false, true, nv, cx, cv, fixed, ya, nullptr, ax, av, slice_off, slice_width, entries,
                           k_s, k_b, dt, gx, gy, gz, damping, st);
    }
    float4* t = cx; cx = ax; ax = t;
    t = cv; cv = av; av = t;
  }
  if (rc) return rc;
  if (cx != (float4*)x4) {
    if (hipMemcpyAsync(x4, cx, sizeof(float4) * (size_t)nv, hipMemcpyDeviceToDevice, st) != hipSuccess ||
        hipMemcpyAsync(v4, cv, sizeof(float4) * (size_t)nv, hipMemcpyDeviceToDevice, st) != hipSuccess)
      return fail(SLX_EHIP, "slx_substeps_plain: copy-back failed");
  }
  return SLX_OK;
}

extern "C" int slx_set_spring_launch(int block, int vpt) {
  SLX_REQUIRE(block == 64 || block == 128 || block == 256, "slx_set_spring_launch: block must be 64, 128 or 256");
  SLX_REQUIRE(vpt == 1 || vpt == 2 || vpt == 4, "slx_set_spring_launch: vertices per thread must be 1, 2 or 4");
  slx::knob_defaults().spring_block = block;
  slx::knob_defaults().spring_vpt = vpt;
  return SLX_OK;
}
