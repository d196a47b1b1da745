/*
 * CPU oracle for the XPBD cloth step.  TEST INFRASTRUCTURE ONLY (see oracle.h):
 * loaded by tests, the smoke check and the benchmark's CPU-baseline leg as the
 * checker; never part of the product path.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline float sqrt_f(float a) { return sqrtf(a); }
static inline double sqrt_d(double a) { return sqrt(a); }
static inline float fabs_f(float a) { return fabsf(a); }
static inline double fabs_d(double a) { return fabs(a); }

#define R float
#define S f
#include "oracle_real.h"
#undef R
#undef S
#define R double
#define S d
#include "oracle_real.h"
#undef R
#undef S
/* the OpenMP instantiation (suffix p, float): element loops on threads,
   atomic scatter -- a timing baseline only, never a checker */
static inline float sqrt_p(float a) { return sqrtf(a); }
static inline float fabs_p(float a) { return fabsf(a); }
#define ORC_PAR 1
#define R float
#define S p
#include "oracle_real.h"
#undef R
#undef S
#undef ORC_PAR

void orc_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}

int orc_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------ LBVH -------------------------------------- */

#define GRID 32
#define NCELLS (GRID * GRID * GRID)
#define STACK_CAP 64

uint32_t orc_expand_bits(uint32_t v) {
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

static inline float clampf(float a, float lo, float hi) { return fminf(fmaxf(a, lo), hi); }

int32_t orc_morton3d(float x, float y, float z) {
  x = clampf(x * 1024.0f, 0.0f, 1023.0f);
  y = clampf(y * 1024.0f, 0.0f, 1023.0f);
  z = clampf(z * 1024.0f, 0.0f, 1023.0f);
  uint32_t xx = orc_expand_bits((uint32_t)x), yy = orc_expand_bits((uint32_t)y),
           zz = orc_expand_bits((uint32_t)z);
  return (int32_t)(xx | (yy << 1) | (zz << 2));
}

/* Normalised Morton code of a point inside [lo, hi] (lbvh_cell.py:139-148). */
static inline int32_t morton_in_box(const float* pos, const float* lo, const float* hi) {
  float t[3];
  for (int c = 0; c < 3; ++c) t[c] = clampf((pos[c] - lo[c]) / (hi[c] - lo[c]), 0.0f, 1.0f);
  return orc_morton3d(t[0], t[1], t[2]);
}

/* Triangle trees: the box's largest extent on every axis (lo + e, the same float
   operations as the library), so thin sheets keep their codes' bits for their width. */
static inline int32_t morton_in_cube(const float* pos, const float* lo, const float* hi) {
  const float e = fmaxf(fmaxf(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
  const float hc[3] = {lo[0] + e, lo[1] + e, lo[2] + e};
  return morton_in_box(pos, lo, hc);
}

static void cell_grid(const float* lo, const float* hi, float* cs, float* centers) {
  for (int c = 0; c < 3; ++c) cs[c] = (hi[c] - lo[c]) / (float)GRID;
  if (!centers) return;
  for (int i = 0; i < NCELLS; ++i) {
    int ijk[3] = {i / (GRID * GRID), (i % (GRID * GRID)) / GRID, i % GRID};
    for (int c = 0; c < 3; ++c)
      centers[3 * i + c] = ((cs[c] * (float)ijk[c]) + (0.5f * cs[c])) + lo[c];
  }
}

void orc_cell_morton(const float* lo, const float* hi, int32_t* codes, float* centers,
                     float* cell_size) {
  cell_grid(lo, hi, cell_size, centers);
  for (int i = 0; i < NCELLS; ++i) codes[i] = morton_in_box(centers + 3 * i, lo, hi);
}

void orc_radix_sort(int n, int32_t* keys, int32_t* ids, int bits, int passes) {
  int radix = 1 << bits;
  int32_t* cnt = (int32_t*)calloc((size_t)radix, sizeof(int32_t));
  int32_t* k2 = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  int32_t* i2 = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  for (int p = 0; p < passes; ++p) {
    memset(cnt, 0, sizeof(int32_t) * (size_t)radix);
    int sh = p * bits;
    for (int i = 0; i < n; ++i) cnt[((uint32_t)keys[i] >> sh) & (uint32_t)(radix - 1)] += 1;
    for (int d = 1; d < radix; ++d) cnt[d] += cnt[d - 1];          /* inclusive */
    for (int I = n - 1; I >= 0; --I) {
      int d = (int)(((uint32_t)keys[I] >> sh) & (uint32_t)(radix - 1));
      int idx = --cnt[d];
      k2[idx] = keys[I];
      i2[idx] = ids[I];
    }
    memcpy(keys, k2, sizeof(int32_t) * (size_t)n);
    memcpy(ids, i2, sizeof(int32_t) * (size_t)n);
  }
  free(cnt);
  free(k2);
  free(i2);
}

static inline int clz64(uint64_t v) {
  if (!v) return 64;
  return __builtin_clzll(v);
}

/* Length of the common prefix of keys i and j, -1 when j is outside [0, n). */
static inline int prefix_len(const int32_t* codes, int n, int wide, int i, int j) {
  if (j < 0 || j >= n) return -1;
  if (wide) {
    uint64_t a = ((uint64_t)(uint32_t)codes[i] << 32) | (uint32_t)i;
    uint64_t b = ((uint64_t)(uint32_t)codes[j] << 32) | (uint32_t)j;
    return clz64(a ^ b);
  }
  uint32_t x = (uint32_t)codes[i] ^ (uint32_t)codes[j];
  return x ? __builtin_clz(x) : 32;
}

void orc_karras(int n, const int32_t* codes, int wide, int32_t* child_a, int32_t* child_b,
                int32_t* parent, int32_t* range_l, int32_t* range_r) {
  if (n <= 0) return;
  for (int k = 0; k < 2 * n - 1; ++k) parent[k] = -1;
  for (int i = 0; i < n - 1; ++i) {
    int dl = prefix_len(codes, n, wide, i, i - 1);
    int dr = prefix_len(codes, n, wide, i, i + 1);
    int d = (dr < dl) ? -1 : 1;
    int dmin = (dr < dl) ? dr : dl;
    /* exponential then binary search for the far end of the range */
    int hi = 2;
    while (prefix_len(codes, n, wide, i, i + hi * d) > dmin) hi *= 2;
    int len = 0;
    for (int t = hi / 2; t > 0; t /= 2)
      if (prefix_len(codes, n, wide, i, i + (len + t) * d) > dmin) len += t;
    int j = i + len * d;
    int first = d > 0 ? i : j, last = d > 0 ? j : i;
    /* split: last index in [first, last) sharing more than the node's prefix with first */
    int split;
    int node_prefix = prefix_len(codes, n, wide, first, last);
    if (!wide && node_prefix == 32) {
      split = (first + last) / 2;
    } else {
      split = first;
      int step = last - first;
      do {
        step = (step + 1) / 2;
        int cand = split + step;
        if (cand < last && prefix_len(codes, n, wide, first, cand) > node_prefix) split = cand;
      } while (step > 1);
    }
    int a = (split == first) ? split + (n - 1) : split;
    int b = (split + 1 == last) ? split + 1 + (n - 1) : split + 1;
    child_a[i] = a;
    child_b[i] = b;
    parent[a] = i;
    parent[b] = i;
    range_l[i] = first;
    range_r[i] = last;
  }
}

void orc_face_aabbs(int nf, const int* faces, const float* x, float padding, float* fmin,
                    float* fmax) {
  for (int f = 0; f < nf; ++f) {
    const float *a = x + 3 * faces[3 * f], *b = x + 3 * faces[3 * f + 1], *c = x + 3 * faces[3 * f + 2];
    for (int k = 0; k < 3; ++k) {
      float mn = fminf(fminf(a[k], b[k]), c[k]);
      float mx = fmaxf(fmaxf(a[k], b[k]), c[k]);
      fmin[3 * f + k] = mn - padding * 1.0f;
      fmax[3 * f + k] = mx + padding * 1.0f;
    }
  }
}

void orc_mesh_aabb(int nv, const float* x, float padding, float* lo, float* hi) {
  for (int k = 0; k < 3; ++k) { lo[k] = 1e5f; hi[k] = -1e5f; }
  for (int i = 0; i < nv; ++i)
    for (int k = 0; k < 3; ++k) {
      hi[k] = fmaxf(hi[k], x[3 * i + k]);
      lo[k] = fminf(lo[k], x[3 * i + k]);
    }
  for (int k = 0; k < 3; ++k) { lo[k] = lo[k] - padding * 1.0f; hi[k] = hi[k] + padding * 1.0f; }
}

/* float -> int32 truncation with NaN -> INT32_MIN and saturation, identical on
   host and device. */
static inline int32_t trunc_i32(float v) {
  if (!(v > -2147483648.0f)) return INT32_MIN;
  if (v >= 2147483648.0f) return INT32_MAX;
  return (int32_t)v;
}

int32_t orc_flat_cell_id(const float* pos, const float* cs, const float* origin) {
  uint32_t ijk[3];
  for (int c = 0; c < 3; ++c) ijk[c] = (uint32_t)trunc_i32((pos[c] - origin[c]) / cs[c]);
  return (int32_t)(ijk[0] * (uint32_t)(GRID * GRID) + ijk[1] * (uint32_t)GRID + ijk[2]);
}

static inline int overlap(const float* amin, const float* amax, const float* bmin, const float* bmax) {
  return amin[0] <= bmax[0] && amax[0] >= bmin[0] && amin[1] <= bmax[1] && amax[1] >= bmin[1] &&
         amin[2] <= bmax[2] && amax[2] >= bmin[2];
}

/* Children-before-parent order of internal nodes: reverse breadth-first order. */
static int internal_postorder(int nl, const int32_t* ca, const int32_t* cb, int32_t* out) {
  int ni = nl - 1, head = 0, tail = 0;
  if (ni <= 0) return 0;
  out[tail++] = 0;
  while (head < tail) {
    int i = out[head++];
    if (ca[i] < ni) out[tail++] = ca[i];
    if (cb[i] < ni) out[tail++] = cb[i];
  }
  for (int a = 0, b = tail - 1; a < b; ++a, --b) { int t = out[a]; out[a] = out[b]; out[b] = t; }
  return tail;
}

/* lbvh_cell.py:490-522: refit of one internal node from its two children. */
static void refit_node(int i, const int32_t* ca, const int32_t* cb, int32_t* num, float* nmin,
                       float* nmax) {
  int l = ca[i], r = cb[i];
  const float *min0 = nmin + 3 * l, *max0 = nmax + 3 * l, *min1 = nmin + 3 * r, *max1 = nmax + 3 * r;
  num[i] = num[l] + num[r];
  for (int c = 0; c < 3; ++c) {
    float mn, mx;
    if (num[l] > 0 && num[r] > 0) {
      mn = fminf(min0[c], min1[c]);
      mx = fmaxf(max0[c], max1[c]);
    } else if (num[l] > 0) {
      mn = min0[c];
      mx = max0[c];
    } else if (num[r] > 0) {
      mn = min1[c];
      mx = max1[c];
    } else {
      mn = mx = 0.25f * (((min0[c] + max0[c]) + min1[c]) + max1[c]);
    }
    nmin[3 * i + c] = mn;
    nmax[3 * i + c] = mx;
  }
}

int orc_cell_build(int nf, const float* fmin, const float* fmax, const float* lo, const float* hi,
                   const int32_t* ca, const int32_t* cb, const int32_t* parent, float* cs,
                   int32_t* face_cell, int32_t* prefix, int32_t* sorted_faces, int32_t* num,
                   int32_t* leaf_l, int32_t* leaf_r, float* nmin, float* nmax) {
  (void)parent;
  float* centers = (float*)malloc(sizeof(float) * 3 * NCELLS);
  cell_grid(lo, hi, cs, centers);
  int bad = 0;
  memset(prefix, 0, sizeof(int32_t) * NCELLS);
  for (int f = 0; f < nf; ++f) {
    float pos[3];
    for (int c = 0; c < 3; ++c) pos[c] = 0.5f * (fmin[3 * f + c] + fmax[3 * f + c]);
    int32_t id = orc_flat_cell_id(pos, cs, lo);
    if (id < 0 || id >= NCELLS) { bad++; id = id < 0 ? 0 : NCELLS - 1; }
    face_cell[f] = id;
    prefix[id] += 1;
  }
  for (int i = 1; i < NCELLS; ++i) prefix[i] += prefix[i - 1];   /* inclusive */
  int32_t* tmp = (int32_t*)malloc(sizeof(int32_t) * NCELLS);
  memcpy(tmp, prefix, sizeof(int32_t) * NCELLS);
  for (int I = nf - 1; I >= 0; --I) sorted_faces[--tmp[face_cell[I]]] = I;
  free(tmp);
  int nn = 2 * NCELLS - 1;
  for (int k = 0; k < nn; ++k) num[k] = 0;
  for (int i = 0; i < NCELLS; ++i) {
    int node = i + NCELLS - 1;
    int l = i ? prefix[i - 1] : 0, r = prefix[i] - 1;
    leaf_l[i] = l;
    leaf_r[i] = r;
    num[node] = r - l + 1;
    float mn[3], mx[3];
    for (int c = 0; c < 3; ++c) mn[c] = mx[c] = centers[3 * i + c];
    for (int j = l; j <= r; ++j) {
      int f = sorted_faces[j];
      for (int c = 0; c < 3; ++c) {
        mn[c] = fminf(mn[c], fmin[3 * f + c]);
        mx[c] = fmaxf(mx[c], fmax[3 * f + c]);
      }
    }
    for (int c = 0; c < 3; ++c) { nmin[3 * node + c] = mn[c]; nmax[3 * node + c] = mx[c]; }
  }
  free(centers);
  int32_t* ord = (int32_t*)malloc(sizeof(int32_t) * NCELLS);
  int cnt = internal_postorder(NCELLS, ca, cb, ord);
  for (int k = 0; k < cnt; ++k) refit_node(ord[k], ca, cb, num, nmin, nmax);
  free(ord);
  return bad;
}

void orc_tri_build(int nf, const float* fmin, const float* fmax, const float* lo, const float* hi,
                   int32_t* codes, int32_t* sorted_faces, int32_t* ca, int32_t* cb, int32_t* parent,
                   float* nmin, float* nmax) {
  if (nf <= 0) return;
  int32_t* keys = (int32_t*)malloc(sizeof(int32_t) * (size_t)nf);
  for (int f = 0; f < nf; ++f) {
    float pos[3];
    for (int c = 0; c < 3; ++c) pos[c] = 0.5f * (fmin[3 * f + c] + fmax[3 * f + c]);
    codes[f] = morton_in_cube(pos, lo, hi);
    keys[f] = codes[f];
    sorted_faces[f] = f;
  }
  orc_radix_sort(nf, keys, sorted_faces, 8, 4);
  int32_t* rl = (int32_t*)malloc(sizeof(int32_t) * (size_t)nf);
  int32_t* rr = (int32_t*)malloc(sizeof(int32_t) * (size_t)nf);
  orc_karras(nf, keys, 1, ca, cb, parent, rl, rr);
  free(rl);
  free(rr);
  free(keys);
  int32_t* num = (int32_t*)malloc(sizeof(int32_t) * (size_t)(2 * nf - 1));
  for (int k = 0; k < nf; ++k) {
    int node = k + nf - 1, f = sorted_faces[k];
    num[node] = 1;
    for (int c = 0; c < 3; ++c) { nmin[3 * node + c] = fmin[3 * f + c]; nmax[3 * node + c] = fmax[3 * f + c]; }
  }
  int32_t* ord = (int32_t*)malloc(sizeof(int32_t) * (size_t)nf);
  int cnt = internal_postorder(nf, ca, cb, ord);
  for (int k = 0; k < cnt; ++k) refit_node(ord[k], ca, cb, num, nmin, nmax);
  free(ord);
  free(num);
}

void orc_tri_refit(int nf, const float* fmin, const float* fmax, const int32_t* sorted_faces,
                   const int32_t* ca, const int32_t* cb, float* nmin, float* nmax) {
  if (nf <= 0) return;
  int32_t* num = (int32_t*)malloc(sizeof(int32_t) * (size_t)(2 * nf - 1));
  for (int k = 0; k < nf; ++k) {
    int node = k + nf - 1, f = sorted_faces[k];
    num[node] = 1;
    for (int c = 0; c < 3; ++c) { nmin[3 * node + c] = fmin[3 * f + c]; nmax[3 * node + c] = fmax[3 * f + c]; }
  }
  int32_t* ord = (int32_t*)malloc(sizeof(int32_t) * (size_t)nf);
  int cnt = internal_postorder(nf, ca, cb, ord);
  for (int k = 0; k < cnt; ++k) refit_node(ord[k], ca, cb, num, nmin, nmax);
  free(ord);
  free(num);
}

int orc_traverse(int mode, int nq, const float* qmin, const float* qmax, int nl,
                 const int32_t* ca, const int32_t* cb, const int32_t* num, const int32_t* leaf_l,
                 const int32_t* leaf_r, const int32_t* sorted_faces, const float* nmin,
                 const float* nmax, const float* fmin, const float* fmax, const float* cs,
                 const float* origin, int cap, int32_t* cand, int32_t* cand_n) {
  int flagged = 0;
  /* queries are independent (each writes its own list and count): the loop
     runs on OpenMP threads with results identical to the serial loop */
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : flagged)
  for (int q = 0; q < nq; ++q) {
    const float *a = qmin + 3 * q, *b = qmax + 3 * q;
    int cnt = 0, over = 0;
    if (mode == 0) {
      float pos[3];
      for (int c = 0; c < 3; ++c) pos[c] = 0.5f * (a[c] + b[c]);
      if (orc_flat_cell_id(pos, cs, origin) > NCELLS - 1) { cand_n[q] = 0; continue; }
    }
    if (nl <= 0) { cand_n[q] = 0; continue; }
    int32_t stack[STACK_CAP];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
      int idx = stack[--sp];
      if (mode == 0 && num[idx] <= 0) continue;
      if (!overlap(a, b, nmin + 3 * idx, nmax + 3 * idx)) continue;
      if (idx >= nl - 1) {
        int leaf = idx - (nl - 1);
        int l = mode == 0 ? leaf_l[leaf] : leaf, r = mode == 0 ? leaf_r[leaf] : leaf;
        for (int j = l; j <= r; ++j) {
          int f = sorted_faces[j];
          if (overlap(a, b, fmin + 3 * f, fmax + 3 * f)) {
            if (cnt < cap) cand[(size_t)q * cap + cnt] = f; else over = 1;
            cnt++;
          }
        }
      } else {
        int l = ca[idx], r = cb[idx];
        if (mode != 0 || num[l] > 0) { if (sp < STACK_CAP) stack[sp++] = l; else over = 1; }
        if (mode != 0 || num[r] > 0) { if (sp < STACK_CAP) stack[sp++] = r; else over = 1; }
      }
    }
    cand_n[q] = cnt < cap ? cnt : cap;
    flagged += over;
  }
  return flagged;
}

/* ---------------------------- colouring ----------------------------------- */

int orc_greedy_coloring(int nv, int ne, const int* edges, int max_colors, int32_t* color,
                        int32_t* order, int32_t* offsets) {
  /* vertex -> incident edges (CSR) */
  int32_t* deg = (int32_t*)calloc((size_t)nv + 1, sizeof(int32_t));
  for (int e = 0; e < ne; ++e) { deg[edges[2 * e] + 1]++; deg[edges[2 * e + 1] + 1]++; }
  for (int v = 0; v < nv; ++v) deg[v + 1] += deg[v];
  int32_t* fill = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nv + 1));
  memcpy(fill, deg, sizeof(int32_t) * (size_t)(nv + 1));
  int32_t* inc = (int32_t*)malloc(sizeof(int32_t) * (size_t)(2 * ne + 1));
  for (int e = 0; e < ne; ++e) { inc[fill[edges[2 * e]]++] = e; inc[fill[edges[2 * e + 1]]++] = e; }
  free(fill);
  unsigned char* used = (unsigned char*)calloc((size_t)max_colors, 1);
  int ncol = 0, err = 0;
  for (int e = 0; e < ne; ++e) color[e] = -1;
  for (int e = 0; e < ne && !err; ++e) {
    memset(used, 0, (size_t)max_colors);
    for (int s = 0; s < 2; ++s) {
      int v = edges[2 * e + s];
      for (int k = deg[v]; k < deg[v + 1]; ++k) {
        int o = inc[k];
        if (o != e && color[o] >= 0) used[color[o]] = 1;
      }
    }
    int c = 0;
    while (c < max_colors && used[c]) ++c;
    if (c == max_colors) { err = 1; break; }
    color[e] = c;
    if (c + 1 > ncol) ncol = c + 1;
  }
  free(used);
  free(inc);
  free(deg);
  if (err) return -1;
  for (int c = 0; c <= ncol; ++c) offsets[c] = 0;
  for (int e = 0; e < ne; ++e) offsets[color[e] + 1]++;
  for (int c = 0; c < ncol; ++c) offsets[c + 1] += offsets[c];
  int32_t* pos = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ncol + 1));
  for (int c = 0; c < ncol; ++c) pos[c] = offsets[c + 1];
  for (int e = ne - 1; e >= 0; --e) order[--pos[color[e]]] = e;
  free(pos);
  return ncol;
}
