// Graph-coloured Gauss-Seidel variant of the spring pass.
//
// Colouring (host): stretch edges are coloured greedily in edge order, each
// taking the lowest colour unused by an already-coloured edge sharing a vertex,
// then listed colour by colour with ascending edge id inside a colour
// (framework/utilities/graph_coloring.py:131-149 greedy, 172-201 stable
// counting sort).  The reference colours stretch edges only and has no
// Gauss-Seidel sweep (README "[WIP]"); the sweep here is this package's.
//
// One GS iteration: (1) the bending constraints, which act on x
// (XPBD.py:166-184) and are therefore fixed within a substep, are applied as a
// Jacobi-averaged pass (y += f * sum/n_bend); (2) for each colour in order, every
// stretch edge of that colour is projected on the current y and both endpoints
// are moved at once (y_v += f_v * (-(f_v w_v ld) n_v)).  No two edges of one
// colour share a vertex, so a colour's edges are independent and the sweep
// equals a serial loop in colour order.
#include <string.h>

#include <vector>

#include "slx_common.h"

namespace slx {

// bending-only Jacobi pass over the SELL-64 adjacency, in place on y
__global__ void __launch_bounds__(256)
k_bend_pass(int nv, const float4* __restrict__ x4, const float* __restrict__ fixed, float4* __restrict__ y4,
            const int32_t* __restrict__ slice_off, const int32_t* __restrict__ slice_width,
            const int2* __restrict__ ent, float k_b) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const int s = v >> 6, lane = v & 63;
  const int base = slice_off[s] + lane, width = slice_width[s];
  const float4 xv = x4[v];
  float3 dx = make_float3(0.f, 0.f, 0.f);
  float nc = 0.f;
  for (int k = 0; k < width; ++k) {
    int2 e = ent[base + 64 * k];
    if (e.x == -1) break;
    if (e.x >= 0) continue;                // stretch entry
    const int u = e.x & 0x7fffffff;
    const float4 xu = x4[u];
    float3 d = sub3(xyz(xv), xyz(xu));
    float len = sqrtf(dot3(d, d));
    float C = len - __int_as_float(e.y);
    float inv = 1.0f / len;
    float3 n = mul3(inv, d);
    float schur = (xv.w + xu.w) * dot3(n, n);
    float ld = (k_b * C) / ((k_b * schur) + 1.0f);
    dx = sub3(dx, mul3(xv.w * ld, n));
    nc = nc + 1.0f;
  }
  if (nc > 0.f) {
    const float f = fixed[v];
    float4 y = y4[v];
    y4[v] = make_float4(y.x + f * (dx.x / nc), y.y + f * (dx.y / nc), y.z + f * (dx.z / nc), y.w);
  }
}

// one colour of stretch edges; gs_edges[k] = (v0, v1, l0 bits, 0) in colour order
__global__ void k_gs_color(int begin, int end, const int4* __restrict__ gs_edges, const float* __restrict__ fixed,
                           float4* __restrict__ y4, float k_s) {
  int k = begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= end) return;
  const int4 e = gs_edges[k];
  const float4 p0 = y4[e.x], p1 = y4[e.y];
  float3 d = sub3(xyz(p0), xyz(p1));
  float len = sqrtf(dot3(d, d));
  float C = len - __int_as_float(e.z);
  float inv = 1.0f / len;
  float3 n = mul3(inv, d);
  float schur = (p0.w + p1.w) * dot3(n, n);
  float ld = (k_s * C) / ((k_s * schur) + 1.0f);
  // endpoint 0 moves by -(fw0 ld) n, endpoint 1 by +(fw1 ld) n (XPBD.py:159-160)
  float3 t0 = mul3(p0.w * ld, n), t1 = mul3(p1.w * ld, n);
  const float f0 = fixed[e.x], f1 = fixed[e.y];
  float3 d0 = make_float3(0.f - t0.x, 0.f - t0.y, 0.f - t0.z);
  float3 d1 = make_float3(0.f + t1.x, 0.f + t1.y, 0.f + t1.z);
  y4[e.x] = make_float4(p0.x + f0 * d0.x, p0.y + f0 * d0.y, p0.z + f0 * d0.z, p0.w);
  y4[e.y] = make_float4(p1.x + f1 * d1.x, p1.y + f1 * d1.y, p1.z + f1 * d1.z, p1.w);
}

int gs_pass(int nv, const float4* x4, const float* fixed, float4* y4, const int32_t* slice_off,
            const int32_t* slice_width, const int32_t* entries, const int4* gs_edges, const int32_t* color_off_host,
            int n_colors, float k_s, float k_b, hipStream_t st) {
  if (nv <= 0) return SLX_OK;
  SLX_LAUNCH("gs_bend", k_bend_pass, dim3(blocks_for(nv, 256)), dim3(256), 0, st, nv, x4, fixed, y4, slice_off,
             slice_width, (const int2*)entries, k_b);
  for (int c = 0; c < n_colors; ++c) {
    int b = color_off_host[c], e = color_off_host[c + 1];
    if (e <= b) continue;
    SLX_LAUNCH("gs_color", k_gs_color, dim3(blocks_for(e - b, 256)), dim3(256), 0, st, b, e, gs_edges, fixed, y4,
               k_s);
  }
  return check_launch("gs_pass");
}

}  // namespace slx

using namespace slx;

extern "C" int slx_greedy_edge_coloring(int nv, int ne, const int32_t* edges, int max_colors, int32_t* color,
                                        int32_t* order, int32_t* offsets) {
  SLX_REQUIRE(nv >= 0 && ne >= 0 && max_colors > 0 && max_colors <= 1024, "slx_greedy_edge_coloring: bad sizes");
  SLX_REQUIRE(ne == 0 || (edges && color && order && offsets), "slx_greedy_edge_coloring: null pointer");
  // incident edges per vertex
  std::vector<int32_t> start((size_t)nv + 1, 0);
  for (int e = 0; e < ne; ++e) {
    int a = edges[2 * e], b = edges[2 * e + 1];
    SLX_REQUIRE(a >= 0 && a < nv && b >= 0 && b < nv, "slx_greedy_edge_coloring: vertex id out of range");
    start[a + 1]++;
    start[b + 1]++;
  }
  for (int v = 0; v < nv; ++v) start[v + 1] += start[v];
  std::vector<int32_t> inc((size_t)2 * ne), fillp(start.begin(), start.end() - 1);
  for (int e = 0; e < ne; ++e) {
    inc[fillp[edges[2 * e]]++] = e;
    inc[fillp[edges[2 * e + 1]]++] = e;
  }
  std::vector<int32_t> stamp((size_t)max_colors, -1);
  int ncol = 0;
  for (int e = 0; e < ne; ++e) color[e] = -1;
  for (int e = 0; e < ne; ++e) {
    for (int s = 0; s < 2; ++s) {
      int v = edges[2 * e + s];
      for (int k = start[v]; k < start[v + 1]; ++k) {
        int o = inc[k];
        if (o != e && color[o] >= 0) stamp[color[o]] = e;   // colour used by a neighbour
      }
    }
    int c = 0;
    while (c < max_colors && stamp[c] == e) ++c;
    if (c == max_colors) return fail(SLX_EINVAL, "slx_greedy_edge_coloring: more than %d colours needed", max_colors);
    color[e] = c;
    if (c + 1 > ncol) ncol = c + 1;
  }
  memset(offsets, 0, sizeof(int32_t) * (size_t)(ncol + 1));
  for (int e = 0; e < ne; ++e) offsets[color[e] + 1]++;
  for (int c = 0; c < ncol; ++c) offsets[c + 1] += offsets[c];
  std::vector<int32_t> pos(offsets, offsets + ncol);
  for (int e = 0; e < ne; ++e) order[pos[color[e]]++] = e;   // ascending id inside a colour
  return ncol;
}

extern "C" int slx_gs_pass(int nv, const float* x4, const float* fixed, float* y4, const int32_t* slice_off,
                           const int32_t* slice_width, const int32_t* entries, const int32_t* gs_edges,
                           const int32_t* color_off_host, int n_colors, float k_s, float k_b, slx_stream_t stream) {
  SLX_REQUIRE(nv >= 0 && n_colors >= 0, "slx_gs_pass: bad sizes");
  if (nv == 0) return SLX_OK;
  SLX_REQUIRE(x4 && fixed && y4 && slice_off && slice_width && entries && (n_colors == 0 || (gs_edges && color_off_host)),
              "slx_gs_pass: null pointer");
  return gs_pass(nv, (const float4*)x4, fixed, (float4*)y4, slice_off, slice_width, entries, (const int4*)gs_edges,
                 color_off_host, n_colors, k_s, k_b, (hipStream_t)stream);
}
