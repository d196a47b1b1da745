// Edge-edge constraints on gfx950: position-level (after the vertex-triangle
// contacts, before update_dx) and velocity-level (after the vertex-triangle
// pairs, before update_dv), with deterministic accumulation.
//
// Reference (all of it commented out of the reference's step, kept callable
// there): framework/physics/XPBD.py:75-89 (pair fields), 217-232 (edge broad
// phase), 266-281 and 313-330 (loops); framework/physics/
// collision_constraints_x.py:312-507 (__ee_dy, __ee_st), collision_constraints_v.py
// 573-878 (__ee_st, __ee_dy); framework/collision/distance.py:49-124 (d_type_EE,
// d_EE; the gradient is derived here).
//
// The edge sets come after the vertex-triangle sets in the reference's serial
// order, so a vertex's total is (its vertex-triangle sum) + its edge-edge deltas
// in edge order.  The vertex-triangle records are first folded into each
// vertex's running sum (k_bucket_fold, same summation order), the buckets are
// emptied, and the edge queries put their deltas in the buckets with keys in
// their own serial order:
//   set 0 (static):  ((e * cap + j) * 3 + side) * 2 + slot     (slots x0, x1)
//   set 1 (dynamic): 6 ne cap + ((e * cap + j) * 3 + side) * 4 + slot
// and for the velocity pass, over pair records j:
//   set 0: (e * pair_cap + j) * 2 + slot,  set 1: 2 ne pair_cap + (e * pair_cap + j) * 4 + slot.
#include "slx_common.h"

namespace slx {

// d_type_EE, distance.py:49-99 (op order as the serial restatement)
__device__ int d_type_ee(float3 x0, float3 x1, float3 x2, float3 x3) {
  const float3 u = sub3(x1, x0), v = sub3(x3, x2), w = sub3(x0, x2);
  const float a = dot3(u, u), b = dot3(u, v), c = dot3(v, v), d = dot3(u, w), e = dot3(v, w);
  const float D = a * c - b * b;
  float tD = D, tN;
  int ty = 8;
  const float sN = b * e - c * d;
  if (sN <= 0.f) {
    tN = e;
    tD = c;
    ty = 2;
  } else if (sN >= D) {
    tN = e + b;
    tD = c;
    ty = 5;
  } else {
    tN = a * e - b * d;
    const float3 uv = cross3(u, v);
    if (tN > 0.f && tN < tD && (dot3(uv, w) < 1e-4f || dot3(uv, uv) < (1.0e-20f * a) * c)) {
      if (sN < D / 2.0f) {
        tN = e;
        tD = c;
        ty = 2;
      } else {
        tN = e + b;
        tD = c;
        ty = 5;
      }
    }
  }
  if (tN <= 0.f) {
    if (-d <= 0.f) ty = 0;
    else if (-d >= a) ty = 3;
    else ty = 6;
  } else if (tN >= tD) {
    if ((-d + b) <= 0.f) ty = 1;
    else if ((-d + b) >= a) ty = 4;
    else ty = 7;
  }
  return ty;
}

// d_EE, distance.py:118-124: d = s^2/N, u = x1-x0, v = x3-x2, w = x2-x0, b = u x v,
// s = w.b, N = b.b; gradient g_k = (2s/N) ds_k - (s^2/N^2) dN_k with
// ds = (-b - v x w, v x w, b - w x u, w x u), dN = 2 (-(v x b), v x b, -(b x u), b x u).
__device__ __forceinline__ float dist_ee(float3 x0, float3 x1, float3 x2, float3 x3, float3 (&g)[4]) {
  const float3 u = sub3(x1, x0), v = sub3(x3, x2), w = sub3(x2, x0);
  const float3 b = cross3(u, v);
  const float s = dot3(w, b), N = dot3(b, b);
  const float d = (s * s) / N;
  const float3 vw = cross3(v, w), wu = cross3(w, u), vb = cross3(v, b), bu = cross3(b, u);
  const float p = (2.0f * s) / N, q = (s * s) / (N * N);
  g[0] = make_float3(p * (-b.x - vw.x) - q * (2.0f * -vb.x), p * (-b.y - vw.y) - q * (2.0f * -vb.y),
                     p * (-b.z - vw.z) - q * (2.0f * -vb.z));
  g[1] = make_float3(p * vw.x - q * (2.0f * vb.x), p * vw.y - q * (2.0f * vb.y), p * vw.z - q * (2.0f * vb.z));
  g[2] = make_float3(p * (b.x - wu.x) - q * (2.0f * -bu.x), p * (b.y - wu.y) - q * (2.0f * -bu.y),
                     p * (b.z - wu.z) - q * (2.0f * -bu.z));
  g[3] = make_float3(p * wu.x - q * (2.0f * bu.x), p * wu.y - q * (2.0f * bu.y), p * wu.z - q * (2.0f * bu.z));
  return d;
}

struct EdgeArgs {
  int ne, cap, pcap;
  const uint32_t* gate;          // the edge gate's state (nullable: no gate)
  const int32_t *edges_dy, *edges_st, *fe_dy, *fe_st;
  const int32_t *c_st, *n_st, *c_dy, *n_dy;
  int32_t* pair_num;
  int32_t* pair_cnt;
  float4* pair_rec;
};

// (the edge gate's predicate; the gate itself is described further below)
__device__ __forceinline__ uint32_t max_key(float v) {          // v >= 0 or NaN
  return v != v ? 0x7f800000u : __float_as_uint(v);
}

// edge e of set `set` (0 static, 1 dynamic) with |u| = lu cannot have an acting pair
__device__ __forceinline__ bool ee_gated(const uint32_t* __restrict__ gate, int set, float lu) {
  if (!gate || !(gate[4] & (set ? 2u : 1u))) return false;
  const float pad = __uint_as_float(gate[5]);
  const float C = fmaxf(fmaxf(__uint_as_float(gate[1]), __uint_as_float(gate[3])), __uint_as_float(gate[7]));
  const float slack = ldexpf(C, -20);
  float b;
  if (set) {
    const float LX = __uint_as_float(gate[0]), D = __uint_as_float(gate[2]);
    const float LV = LX + 3.4641016f * D;                           // |v| <= LX + 2 sqrt(3) D
    b = (lu * (LV + slack)) * (lu + 1.7320508f * (((LX + 2.0f * pad) + D) + slack));
  } else {
    const float LS = __uint_as_float(gate[6]);
    b = (lu * (LS + slack)) * (lu + 1.7320508f * ((LS + 2.0f * pad) + slack));
  }
  return b * 1.001f < 1e-4f;                                       // false for NaN
}

// every edge of set `set` is gated: the bound grows with |u|, and |u| <= LX +
// 2 sqrt(3) D for every dynamic edge at y (with an allowance for rounding), so
// when that longest possible edge is gated, each edge is
__device__ __forceinline__ bool ee_set_gated(const uint32_t* __restrict__ gate, int set) {
  if (!gate || !(gate[4] & (set ? 2u : 1u))) return false;
  const float lu_max = (__uint_as_float(gate[0]) + 3.4641016f * __uint_as_float(gate[2])) * 1.0001f +
                       ldexpf(fmaxf(__uint_as_float(gate[1]), __uint_as_float(gate[3])), -20);
  return ee_gated(gate, set, lu_max);
}

// A vertex's bucket records so far folded into its running sum (in key order,
// as the apply kernels would add them) and count; the bucket is emptied.
__global__ void __launch_bounds__(256)
k_bucket_fold(int nv, int bucket, int32_t* __restrict__ bcnt, const float4* __restrict__ brec,
              float4* __restrict__ self_acc) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const float4 sa = self_acc[v];
  int n = bcnt[v];
  if (n > bucket) n = bucket;
  const float3 acc = bucket_sum(make_float3(sa.x, sa.y, sa.z), brec + (size_t)v * bucket, n);
  self_acc[v] = make_float4(acc.x, acc.y, acc.z, sa.w + (float)n);
  bcnt[v] = 0;
}

__device__ __forceinline__ void push_rec(int t, float3 d, uint32_t key, int bucket, int32_t* __restrict__ bcnt,
                                         float4* __restrict__ brec, bool& over) {
  const int at = atomicAdd(&bcnt[t], 1);
  if (at < bucket) brec[(size_t)t * bucket + at] = make_float4(d.x, d.y, d.z, __uint_as_float(key));
  else over = true;
}

// One thread per (set, edge): the sides of its candidate faces in order
// (collision_constraints_x.py:312-507).  (Threads take the edges in index order:
// the candidate lists are read coalesced.)
__global__ void __launch_bounds__(256)
k_ee_contacts(EdgeArgs A, const float4* __restrict__ y4, const float4* __restrict__ xs4,
              const float* __restrict__ fixed, const float* __restrict__ m_inv, float dhat, float k, int bucket,
              int32_t* __restrict__ bcnt, float4* __restrict__ brec, int32_t* status) {
  const int nq = 2 * A.ne;
  const int qg = blockIdx.x * blockDim.x + threadIdx.x;
  if (qg >= nq) return;
  const int set = qg >= A.ne ? 1 : 0;
  const int e = qg - set * A.ne;
  const int32_t* C = set ? A.c_dy : A.c_st;
  int pairs = 0, stored = 0;
  bool over = false, pover = false;
  if (C && !ee_set_gated(A.gate, set)) {   // a gated set: no pair of it can act
    const int32_t* N = set ? A.n_dy : A.n_st;
    const int32_t* fe = set ? A.fe_dy : A.fe_st;
    const int32_t* E2 = set ? A.edges_dy : A.edges_st;
    const float4* P2 = set ? y4 : xs4;
    int n = N[e];
    if (n > A.cap) n = A.cap;
    const int v0 = A.edges_dy[2 * e], v1 = A.edges_dy[2 * e + 1];
    const float3 x0 = xyz(y4[v0]), x1 = xyz(y4[v1]);
    const float3 x01 = sub3(x0, x1);
    if (A.gate) {                          // an edge the gate dropped (it was not walked): no pair can act
      const float3 u = sub3(x1, x0);
      if (ee_gated(A.gate, set, sqrtf(dot3(u, u)))) n = 0;
    }
    const uint32_t base = set ? 6u * (uint32_t)A.ne * (uint32_t)A.cap : 0u;
    const uint32_t nslot = set ? 4u : 2u;
    for (int j = 0; j < n; ++j) {
      const int f = C[(size_t)j * A.ne + e];
      const int3 sides = make_int3(fe[3 * f], fe[3 * f + 1], fe[3 * f + 2]);
#pragma unroll
      for (int kk = 0; kk < 3; ++kk) {
        const int e2 = kk == 0 ? sides.x : (kk == 1 ? sides.y : sides.z);
        const int v2 = E2[2 * e2], v3 = E2[2 * e2 + 1];
        if (set == 1 && (v0 == v2 || v0 == v3 || v1 == v2 || v1 == v3)) continue;
        const float3 x2 = xyz(P2[v2]), x3 = xyz(P2[v3]);
        const int dtype = d_type_ee(x0, x1, x2, x3);
        if (dtype != 8) continue;
        const float3 cr = cross3(x01, sub3(x2, x3));
        if (!(sqrtf(dot3(cr, cr)) > 1e-3f)) continue;
        float3 g[4];
        const float d = dist_ee(x0, x1, x2, x3, g);
        if (!(d < dhat)) continue;
        const int vid[4] = {v0, v1, v2, v3};
        const int nmov = set ? 4 : 2;
        float fw[4];
        float schur = 0.f;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          if (s >= nmov) break;
          fw[s] = fixed[vid[s]] * m_inv[vid[s]];
          const float term = fw[s] * dot3(g[s], g[s]);
          schur = s == 0 ? term : schur + term;
        }
        const float ld = (k * (dhat - d)) / ((k * schur) + 1.0f);
        const uint32_t o0 = base + (((uint32_t)e * (uint32_t)A.cap + (uint32_t)j) * 3u + (uint32_t)kk) * nslot;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          if (s >= nmov) break;
          push_rec(vid[s], mul3(fw[s] * ld, g[s]), o0 + (uint32_t)s, bucket, bcnt, brec, over);
        }
        if (A.pair_rec) {
          if (stored < A.pcap) {
            const size_t col = (size_t)stored * 4u * (size_t)nq + (size_t)qg;
            A.pair_rec[col] = make_float4(g[0].x, g[0].y, g[0].z, schur);
            A.pair_rec[col + (size_t)nq] = make_float4(g[1].x, g[1].y, g[1].z, __int_as_float(e2));
            A.pair_rec[col + 2 * (size_t)nq] = make_float4(g[2].x, g[2].y, g[2].z, __int_as_float(dtype));
            A.pair_rec[col + 3 * (size_t)nq] = make_float4(g[3].x, g[3].y, g[3].z, 0.f);
            ++stored;
          } else {
            pover = true;
          }
        }
        ++pairs;
      }
    }
  }
  if (A.pair_num) A.pair_num[qg] = pairs;
  if (A.pair_rec) A.pair_cnt[qg] = stored;
  if (over && status) atomicOr(status, kStatusRecOverflow);
  if (pover && status) atomicOr(status, kStatusPairOverflow);
}

// One thread per (set, edge): its recorded pairs in order
// (collision_constraints_v.py:573-878, normal part only).
__global__ void __launch_bounds__(256)
k_vel_ee_contacts(EdgeArgs A, const float4* __restrict__ v4, const float4* __restrict__ vs4,
                  const float* __restrict__ fixed, const float* __restrict__ m_inv, int bucket,
                  int32_t* __restrict__ bcnt, float4* __restrict__ brec, int32_t* status) {
  const int nq = 2 * A.ne;
  const int qg = blockIdx.x * blockDim.x + threadIdx.x;
  if (qg >= nq) return;
  const int set = qg >= A.ne ? 1 : 0;
  const int e = qg - set * A.ne;
  int n = A.pair_cnt[qg];
  if (n > A.pcap) n = A.pcap;
  if (n <= 0) return;
  const int v0 = A.edges_dy[2 * e], v1 = A.edges_dy[2 * e + 1];
  const float3 u0 = xyz(v4[v0]), u1 = xyz(v4[v1]);
  const uint32_t base = set ? 2u * (uint32_t)A.ne * (uint32_t)A.pcap : 0u;
  const uint32_t nslot = set ? 4u : 2u;
  bool over = false;
  for (int j = 0; j < n; ++j) {
    const size_t col = (size_t)j * 4u * (size_t)nq + (size_t)qg;
    const float4 r0 = A.pair_rec[col], r1 = A.pair_rec[col + (size_t)nq], r2 = A.pair_rec[col + 2 * (size_t)nq],
                 r3 = A.pair_rec[col + 3 * (size_t)nq];
    const int e2 = __float_as_int(r1.w);
    const float3 g0 = xyz(r0), g1 = xyz(r1), g2 = xyz(r2), g3 = xyz(r3);
    const float schur = r0.w;
    const uint32_t o0 = base + ((uint32_t)e * (uint32_t)A.pcap + (uint32_t)j) * nslot;
    if (set == 0) {
      // __ee_st: Cv = g0.v0 + g1.v1 + g3.vs3 (g2 does not enter, collision_constraints_v.py:782)
      const int v3 = A.edges_st[2 * e2 + 1];
      const float3 u3 = vs4 ? xyz(vs4[v3]) : make_float3(0.f, 0.f, 0.f);
      const float Cv = (dot3(g0, u0) + dot3(g1, u1)) + dot3(g3, u3);
      if (!(Cv < 0.f)) continue;
      const float ld = -Cv / schur;
      push_rec(v0, mul3((fixed[v0] * m_inv[v0]) * ld, g0), o0, bucket, bcnt, brec, over);
      push_rec(v1, mul3((fixed[v1] * m_inv[v1]) * ld, g1), o0 + 1u, bucket, bcnt, brec, over);
    } else {
      // __ee_dy: dvn = sum g.v; dv -= f w g (dvn / schur) on all four points
      const int v2 = A.edges_dy[2 * e2], v3 = A.edges_dy[2 * e2 + 1];
      const float dvn = ((dot3(g0, u0) + dot3(g1, u1)) + dot3(g2, xyz(v4[v2]))) + dot3(g3, xyz(v4[v3]));
      if (!(dvn < 0.f)) continue;
      const float ld = dvn / schur;
      const int vid[4] = {v0, v1, v2, v3};
      const float3 g[4] = {g0, g1, g2, g3};
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const float a = fixed[vid[s]] * m_inv[vid[s]];
        const float3 d = make_float3(-((a * g[s].x) * ld), -((a * g[s].y) * ld), -((a * g[s].z) * ld));
        push_rec(vid[s], d, o0 + (uint32_t)s, bucket, bcnt, brec, over);
      }
    }
  }
  if (over && status) atomicOr(status, kStatusRecOverflow);
}

// ---------------- the edge gate ----------------
//
// Type 8 needs (u x v).w >= 1e-4 in d_type_EE (u, v the two edges, w = x0 - x2),
// and |(u x v).w| <= |u| |v| |w|.  For a candidate face f of edge e's query box
// Q (y -+ pad), f's box (its vertices at the tree's positions X, -+ pad) overlaps
// Q, so per axis |y0 - X2| <= |u| + ext(f) + 2 pad and |y2 - X2| <= D, hence
//   |w| <= |u| + sqrt(3) (LX + 2 pad + D),   |v| <= LX + 2 sqrt(3) D,
// with LX the longest edge at X and D = max |y - X|_inf over the vertices (for
// the static mesh, whose tree and narrow phase both use xs: LS and D = 0).  When
// that bound on the triple product, with allowances for rounding (2^-20 of the
// largest coordinate per axis, a factor 1.001), stays below 1e-4, none of e's
// pairs can act: the edge is not walked, and k_ee_contacts, evaluating the same
// predicate, tests nothing for it (pair count 0).  NaN or infinite inputs never
// pass the test.  The dynamic set is gated only when its tree's boxes come from
// the current x (X = x).
//
// gate[] (float bits; maxima kept with integer atomics: non-negative floats order
// as their bits, NaN counts as +inf): 0 LX, 1 max |x|, 2 D, 3 max |y|, 4 flags (1
// static set gated, 2 dynamic set gated), 5 padding, 6 LS, 7 max |xs|.
__device__ __forceinline__ void block_max2(float a, float b, uint32_t* ga, uint32_t* gb) {
  __shared__ uint32_t sa[4], sb[4];
  uint32_t ka = max_key(a), kb = max_key(b);
  for (int o = 32; o > 0; o >>= 1) {
    ka = max(ka, (uint32_t)__shfl_xor((int)ka, o, 64));
    kb = max(kb, (uint32_t)__shfl_xor((int)kb, o, 64));
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sa[w] = ka;
    sb[w] = kb;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) {
      ka = max(ka, sa[i]);
      kb = max(kb, sb[i]);
    }
    if (ka > *ga) atomicMax(ga, ka);
    if (kb > *gb) atomicMax(gb, kb);
  }
}

// zero gate[z0 .. z1), set the flags and the padding, zero the live counts
__global__ void k_gate_set(uint32_t* __restrict__ gate, int z0, int z1, uint32_t flags, float pad,
                           int32_t* __restrict__ c0, int32_t* __restrict__ c1) {
  const int i = threadIdx.x;
  if (i >= z0 && i < z1) gate[i] = 0u;
  if (i == 4) gate[4] = flags;
  if (i == 5) gate[5] = __float_as_uint(pad);
  if (i == 0 && c0) *c0 = 0;
  if (i == 0 && c1) *c1 = 0;
}

// gate[slot] = longest edge of `edges` at p4, gate[slot + 1] = largest |coordinate|
// of their end points (2048 edges per workgroup of 256 threads)
__global__ void __launch_bounds__(256)
k_gate_edges(int ne, const int32_t* __restrict__ edges, const float4* __restrict__ p4, uint32_t* gate, int slot) {
  float lm = 0.f, cm = 0.f;
  const int base = blockIdx.x * 2048 + threadIdx.x;
  for (int i = 0; i < 8; ++i) {
    const int e = base + i * 256;
    if (e >= ne) break;
    const float3 a = xyz(p4[edges[2 * e]]), b = xyz(p4[edges[2 * e + 1]]);
    const float3 d = sub3(b, a);
    const float l = sqrtf(dot3(d, d));
    const float c = fmaxf(fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fmaxf(fabsf(a.z), fabsf(b.x))),
                          fmaxf(fabsf(b.y), fabsf(b.z)));
    lm = (l != l || lm != lm) ? l + lm : fmaxf(lm, l);            // a NaN sticks
    cm = (c != c || cm != cm) ? c + cm : fmaxf(cm, c);
  }
  block_max2(lm, cm, gate + slot, gate + slot + 1);
}

// gate[2] = max |y - x|_inf, gate[3] = max |y| over the dynamic vertices (2048 per
// workgroup of 256 threads)
__global__ void __launch_bounds__(256)
k_gate_disp(int nv, const float4* __restrict__ x4, const float4* __restrict__ y4, uint32_t* gate) {
  float dm = 0.f, cm = 0.f;
  const int base = blockIdx.x * 2048 + threadIdx.x;
  for (int i = 0; i < 8; ++i) {
    const int v = base + i * 256;
    if (v >= nv) break;
    const float4 x = x4[v], y = y4[v];
    const float d = fmaxf(fmaxf(fabsf(y.x - x.x), fabsf(y.y - x.y)), fabsf(y.z - x.z));
    const float c = fmaxf(fmaxf(fabsf(y.x), fabsf(y.y)), fabsf(y.z));
    dm = (d != d || dm != dm) ? d + dm : fmaxf(dm, d);
    cm = (c != c || cm != cm) ? c + cm : fmaxf(cm, c);
  }
  block_max2(dm, cm, gate + 2, gate + 3);
}

// the edges (taken in `order`) whose pairs could act, appended per gated set to
// its list ([0] = count, then edge ids)
__global__ void __launch_bounds__(256)
k_ee_gate(int ne, const int32_t* __restrict__ order, const int32_t* __restrict__ edges, const float4* __restrict__ y4,
          const uint32_t* __restrict__ gate, int32_t* __restrict__ live_st, int32_t* __restrict__ live_dy) {
  if ((!live_st || ee_set_gated(gate, 0)) && (!live_dy || ee_set_gated(gate, 1))) return;
  const int t = blockIdx.x * 256 + threadIdx.x;
  const bool active = t < ne;
  const int e = active ? (order ? order[t] : t) : 0;
  bool live0 = false, live1 = false;
  if (active) {
    const float3 u = sub3(xyz(y4[edges[2 * e + 1]]), xyz(y4[edges[2 * e]]));
    const float lu = sqrtf(dot3(u, u));
    live0 = !ee_gated(gate, 0, lu);
    live1 = !ee_gated(gate, 1, lu);
  }
  const int lane = threadIdx.x & 63;
  auto append = [&](bool need, int32_t* list) {
    if (!list) return;
    const unsigned long long m = __ballot(need);
    if (!m) return;
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(&list[0], __popcll(m));
    base = __shfl(base, leader, 64);
    if (need) list[1 + base + __popcll(m & ((1ull << lane) - 1ull))] = e;
  };
  append(active && live0, live_st);
  append(active && live1, live_dy);
}

int ee_gate_static(const slx_edge_sets* ee, int nes, const float4* xs4, hipStream_t st) {
  SLX_LAUNCH("ee_gate", k_gate_set, dim3(1), dim3(64), 0, st, ee->gate, 6, 8, 0u, 0.f, (int32_t*)nullptr,
             (int32_t*)nullptr);
  if (nes > 0)
    SLX_LAUNCH("ee_gate", k_gate_edges, dim3(blocks_for(nes, 256 * 8)), dim3(256), 0, st, nes, ee->edges_st, xs4,
               ee->gate, 6);
  return check_launch("ee_gate_static");
}

// live_st / live_dy: the lists of the gated sets (null: that set is not gated);
// new_x: x changed since the last call (the substep's first pass)
int ee_gate(const slx_edge_sets* ee, int nv, const float4* x4, const float4* y4, float pad, bool new_x,
            int32_t* live_st, int32_t* live_dy, hipStream_t st) {
  const uint32_t flags = (live_st ? 1u : 0u) | (live_dy ? 2u : 0u);
  SLX_LAUNCH("ee_gate", k_gate_set, dim3(1), dim3(64), 0, st, ee->gate, new_x ? 0 : 2, 4, flags, pad, live_st,
             live_dy);
  if (!flags) return check_launch("ee_gate");
  if (new_x)
    SLX_LAUNCH("ee_gate", k_gate_edges, dim3(blocks_for(ee->ne, 256 * 8)), dim3(256), 0, st, ee->ne, ee->edges_dy,
               x4, ee->gate, 0);
  SLX_LAUNCH("ee_gate", k_gate_disp, dim3(blocks_for(nv, 256 * 8)), dim3(256), 0, st, nv, x4, y4, ee->gate);
  SLX_LAUNCH("ee_gate", k_ee_gate, dim3(blocks_for(ee->ne, 256)), dim3(256), 0, st, ee->ne, ee->order,
             ee->edges_dy, y4, ee->gate, live_st, live_dy);
  return check_launch("ee_gate");
}

static EdgeArgs edge_args(const slx_edge_sets* ee) {
  EdgeArgs A;
  A.ne = ee->ne;
  A.cap = ee->cap;
  A.pcap = ee->pair_cap;
  A.gate = ee->gate;
  A.edges_dy = ee->edges_dy;
  A.edges_st = ee->edges_st;
  A.fe_dy = ee->fe_dy;
  A.fe_st = ee->fe_st;
  A.c_st = ee->c_st;
  A.n_st = ee->n_st;
  A.c_dy = ee->c_dy;
  A.n_dy = ee->n_dy;
  A.pair_num = ee->pair_num;
  A.pair_cnt = ee->pair_cnt;
  A.pair_rec = (float4*)ee->pair_rec;
  return A;
}

int ee_check(const slx_edge_sets* ee, bool velocity) {
  SLX_REQUIRE(ee->ne >= 0 && ee->cap >= 0 && ee->pair_cap >= 0, "edge sets: bad sizes");
  if (ee->ne == 0) return SLX_OK;
  SLX_REQUIRE(ee->edges_dy, "edge sets: null edges_dy");
  if (velocity) {
    SLX_REQUIRE(ee->pair_cnt && ee->pair_rec, "edge sets: the velocity pass needs pair_cnt and pair_rec");
    SLX_REQUIRE(ee->edges_st || !ee->c_st, "edge sets: null edges_st");
    SLX_REQUIRE((long long)ee->ne * (long long)ee->pair_cap * 6ll < (1ll << 32),
                "edge sets: too many edges x pairs for 32-bit order keys");
    return SLX_OK;
  }
  SLX_REQUIRE(!ee->c_st || (ee->n_st && ee->fe_st && ee->edges_st), "edge sets: static set incomplete");
  SLX_REQUIRE(!ee->c_dy || (ee->n_dy && ee->fe_dy), "edge sets: dynamic set incomplete");
  SLX_REQUIRE(!ee->pair_rec || ee->pair_cnt, "edge sets: pair records need pair_cnt");
  SLX_REQUIRE((long long)ee->ne * (long long)ee->cap * 18ll < (1ll << 32),
              "edge sets: too many edges x candidates for 32-bit order keys");
  return SLX_OK;
}

// Fold the vertex-triangle records, then the edge-edge deltas into the buckets.
int ee_position_launch(int nv, const slx_edge_sets* ee, const float4* y4, const float4* xs4, const float* fixed,
                       const float* m_inv, float dhat, float k, const ContactWs& w, int bucket, int32_t* status,
                       hipStream_t st) {
  SLX_LAUNCH("bucket_fold", k_bucket_fold, dim3(blocks_for(nv, 256)), dim3(256), 0, st, nv, bucket, w.bcnt, w.brec,
             w.self_acc);
  if (ee->ne == 0) return check_launch("ee_contacts");
  SLX_REQUIRE(!ee->c_st || xs4, "slx_collision_pass_ee: the static set needs static positions");
  SLX_LAUNCH("ee_contacts", k_ee_contacts, dim3(blocks_for(2ll * ee->ne, 256)), dim3(256), 0, st, edge_args(ee), y4,
             xs4, fixed, m_inv, dhat, k, bucket, w.bcnt, w.brec, status);
  return check_launch("ee_contacts");
}

int ee_velocity_launch(int nv, const slx_edge_sets* ee, const float4* v4, const float4* vs4, const float* fixed,
                       const float* m_inv, const ContactWs& w, int bucket, int32_t* status, hipStream_t st) {
  SLX_LAUNCH("bucket_fold", k_bucket_fold, dim3(blocks_for(nv, 256)), dim3(256), 0, st, nv, bucket, w.bcnt, w.brec,
             w.self_acc);
  if (ee->ne == 0) return check_launch("vel_ee_contacts");
  SLX_LAUNCH("vel_ee_contacts", k_vel_ee_contacts, dim3(blocks_for(2ll * ee->ne, 256)), dim3(256), 0, st,
             edge_args(ee), v4, vs4, fixed, m_inv, bucket, w.bcnt, w.brec, status);
  return check_launch("vel_ee_contacts");
}

}  // namespace slx
