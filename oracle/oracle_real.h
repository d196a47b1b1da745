/*
 * Oracle solver and narrow phase, written once for a real type R with suffix
 * S; included by oracle.c for float (S=f) and double (S=d).  TEST
 * INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Floating-point op order is spelled out with explicit parentheses and the
 * file is compiled with -ffp-contract=off, so the float instantiation is an
 * exact IEEE restatement of each formula.
 */
#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(CAT(orc_, name), CAT(_, S))
#define SQRT CAT(sqrt_, S)

static inline R CAT(dot3_, S)(const R* a, const R* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
static inline void CAT(sub3_, S)(const R* a, const R* b, R* o) { o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2]; }
static inline void CAT(cross3_, S)(const R* a, const R* b, R* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
#define DOT CAT(dot3_, S)
#define SUB CAT(sub3_, S)
#define CROSS CAT(cross3_, S)

/* ACC(lvalue, value): the scatter of a delta into dx / dv / nc.  The serial
   instantiations add in loop order (lvalue = lvalue + value, the reference's
   serial order); the OpenMP instantiation (ORC_PAR, suffix p) runs the
   element loops marked PFOR on threads and adds with an atomic, as a
   parallel CPU backend does -- its sums are then in no fixed order. */
#undef ACC
#undef PFOR
#ifdef ORC_PAR
#define ACC(lv, val)                      \
  do {                                    \
    R acc_v_ = (val);                     \
    _Pragma("omp atomic") (lv) += acc_v_; \
  } while (0)
#define PFOR _Pragma("omp parallel for schedule(dynamic, 512)")
#else
#define ACC(lv, val) ((lv) = (lv) + (val))
#define PFOR
#endif

/* XPBD.py:115-118: y = x + f*v*dt + g*dt*dt, evaluated left to right. */
void FN(compute_y)(int nv, const R* x, const R* v, const R* fixed, R* y, R dt, const R* g) {
  PFOR
  for (int i = 0; i < nv; ++i) {
    for (int c = 0; c < 3; ++c) {
      R t = (fixed[i] * v[3 * i + c]) * dt;
      R gg = (g[c] * dt) * dt;
      y[3 * i + c] = (x[3 * i + c] + t) + gg;
    }
  }
}

/* One distance constraint of solve_spring_constraints_x (XPBD.py:146-184),
   scattered into dx/nc in the serial order of the reference loop. */
static inline void CAT(spring_one_, S)(int v0, int v1, R l0, const R* p, R k, const R* fixed,
                                      const R* m_inv, R* dx, R* nc) {
  R d[3];
  SUB(p + 3 * v0, p + 3 * v1, d);
  R len = SQRT(DOT(d, d));
  R C = len - l0;
  R inv = (R)1 / len;
  R n[3] = {inv * d[0], inv * d[1], inv * d[2]};
  R fw0 = fixed[v0] * m_inv[v0], fw1 = fixed[v1] * m_inv[v1];
  R schur = (fw0 + fw1) * DOT(n, n);
  R ld = (k * C) / ((k * schur) + (R)1);
  for (int c = 0; c < 3; ++c) {
    ACC(dx[3 * v0 + c], -((fw0 * ld) * n[c]));
    ACC(dx[3 * v1 + c], (fw1 * ld) * n[c]);
  }
  ACC(nc[v0], (R)1);
  ACC(nc[v1], (R)1);
}

/* XPBD.py:383-387 update_dx: y += f * (dx / nc) where nc > 0. */
static void CAT(update_dx_, S)(int nv, R* y, const R* fixed, const R* dx, const R* nc) {
  PFOR
  for (int i = 0; i < nv; ++i) {
    if (nc[i] > (R)0) {
      for (int c = 0; c < 3; ++c) y[3 * i + c] = y[3 * i + c] + fixed[i] * (dx[3 * i + c] / nc[i]);
    }
  }
}

/* XPBD.py:350-359: init_variables, solve_spring_constraints_x(k_s, k_b), update_dx.
   Stretch constraints (i < ne) act on y, bending constraints on x (XPBD.py:169). */
void FN(spring_pass)(int nv, int ne, int nb, const int* edges, const R* l0, const int* bend,
                     const R* bl0, const R* x, R* y, const R* fixed, const R* m_inv, R k_s, R k_b,
                     R* dx, R* nc) {
  PFOR
  for (int i = 0; i < 3 * nv; ++i) dx[i] = 0;
  PFOR
  for (int i = 0; i < nv; ++i) nc[i] = 0;
  PFOR
  for (int i = 0; i < ne; ++i)
    CAT(spring_one_, S)(edges[2 * i], edges[2 * i + 1], l0[i], y, k_s, fixed, m_inv, dx, nc);
  PFOR
  for (int i = 0; i < nb; ++i)
    CAT(spring_one_, S)(bend[2 * i], bend[2 * i + 1], bl0[i], x, k_b, fixed, m_inv, dx, nc);
  CAT(update_dx_, S)(nv, y, fixed, dx, nc);
}

/* Residual solve pass (implicit-Euler stretch problem anchored at ya, Jacobi with
   the diagonal Gauss-Newton step): per vertex g = sum over stretch edges of
   (kappa C) n with n from the neighbour to the vertex, ks = sum of kappa,
   step = f ((y - ya) + w g) / (1 + w ks), y -= step; *res = max |step| over all
   components of the input y's steps.  Scattered in edge order, which is each
   vertex's constraint order. */
void FN(implicit_pass)(int nv, int ne, const int* edges, const R* l0, const R* ya, R* y, const R* fixed,
                       const R* m_inv, R kappa, R* gs, R* ks, R* res) {
  for (int i = 0; i < 3 * nv; ++i) gs[i] = 0;
  for (int i = 0; i < nv; ++i) ks[i] = 0;
  for (int i = 0; i < ne; ++i) {
    const int a = edges[2 * i], b = edges[2 * i + 1];
    R d[3];
    SUB(y + 3 * a, y + 3 * b, d);
    R len = SQRT(DOT(d, d));
    R C = len - l0[i];
    R inv = (R)1 / len;
    R kc = kappa * C;
    for (int c = 0; c < 3; ++c) {
      R t = kc * (inv * d[c]);
      gs[3 * a + c] = gs[3 * a + c] + t;
      gs[3 * b + c] = gs[3 * b + c] - t;
    }
    ks[a] = ks[a] + kappa;
    ks[b] = ks[b] + kappa;
  }
  R r = 0;
  for (int v = 0; v < nv; ++v) {
    R w = fixed[v] * m_inv[v];
    R den = (R)1 + w * ks[v];
    for (int c = 0; c < 3; ++c) {
      R st = fixed[v] * (((y[3 * v + c] - ya[3 * v + c]) + w * gs[3 * v + c]) / den);
      R as = st < (R)0 ? -st : st;
      if (as > r) r = as;
      y[3 * v + c] = y[3 * v + c] - st;
    }
  }
  *res = r;
}

/* Graph-coloured Gauss-Seidel iteration (the reference colours stretch edges,
   graph_coloring.py:131-201, but has no sweep): bending constraints on x as one
   Jacobi-averaged pass, then for each colour, each edge of the colour in order,
   the stretch projection applied to both endpoints at once with nc = 1. */
void FN(gs_pass)(int nv, int nb, const int* bend, const R* bl0, const R* x, R* y, const R* fixed,
                 const R* m_inv, R k_b, const int* gs_order, const int* edges, const R* l0,
                 const int* color_off, int n_colors, R k_s, R* dx, R* nc) {
  for (int i = 0; i < 3 * nv; ++i) dx[i] = 0;
  for (int i = 0; i < nv; ++i) nc[i] = 0;
  for (int i = 0; i < nb; ++i)
    CAT(spring_one_, S)(bend[2 * i], bend[2 * i + 1], bl0[i], x, k_b, fixed, m_inv, dx, nc);
  CAT(update_dx_, S)(nv, y, fixed, dx, nc);
  for (int c = 0; c < n_colors; ++c) {
    for (int k = color_off[c]; k < color_off[c + 1]; ++k) {
      int e = gs_order[k], v0 = edges[2 * e], v1 = edges[2 * e + 1];
      R d1[6] = {0, 0, 0, 0, 0, 0}, n1[2] = {0, 0};
      /* reuse the scatter form on a two-vertex scratch: dx/nc of just this edge */
      R* p = y;
      R d[3];
      SUB(p + 3 * v0, p + 3 * v1, d);
      R len = SQRT(DOT(d, d));
      R C = len - l0[e];
      R inv = (R)1 / len;
      R n[3] = {inv * d[0], inv * d[1], inv * d[2]};
      R fw0 = fixed[v0] * m_inv[v0], fw1 = fixed[v1] * m_inv[v1];
      R schur = (fw0 + fw1) * DOT(n, n);
      R ld = (k_s * C) / ((k_s * schur) + (R)1);
      for (int q = 0; q < 3; ++q) {
        d1[q] = d1[q] - (fw0 * ld) * n[q];
        d1[3 + q] = d1[3 + q] + (fw1 * ld) * n[q];
      }
      n1[0] = n1[1] = (R)1;
      for (int q = 0; q < 3; ++q) {
        y[3 * v0 + q] = y[3 * v0 + q] + fixed[v0] * (d1[q] / n1[0]);
        y[3 * v1 + q] = y[3 * v1 + q] + fixed[v1] * (d1[3 + q] / n1[1]);
      }
    }
  }
}

/* XPBD.py:340-343 then 332-337:
   v = (1 - damping) * f * (y - x) / dt;  x += dt * v. */
void FN(update_vx)(int nv, R* x, R* v, const R* y, const R* fixed, R damping, R dt) {
  R a = (R)1 - damping;
  PFOR
  for (int i = 0; i < nv; ++i) {
    for (int c = 0; c < 3; ++c) {
      R vv = ((a * fixed[i]) * (y[3 * i + c] - x[3 * i + c])) / dt;
      v[3 * i + c] = vv;
      x[3 * i + c] = x[3 * i + c] + dt * vv;
    }
  }
}

/* ---------------- narrow phase (distance.py, collision_constraints_x.py) -------- */

/* Edge-region parameters of p against the edge a->b of a triangle with normal n:
   solve [e.e, e.en; en.e, en.en] [s t]^T = [e.r, en.r] with e = b - a,
   en = e x n, r = p - a, via the explicit 2x2 inverse (distance.py:7-17). */
static inline void CAT(edge_param_, S)(const R* p, const R* a, const R* b, const R* n, R* s, R* t) {
  R e[3], en[3], r[3];
  SUB(b, a, e);
  CROSS(e, n, en);
  SUB(p, a, r);
  R m00 = DOT(e, e), m01 = DOT(e, en), m10 = DOT(en, e), m11 = DOT(en, en);
  R det = m00 * m11 - m01 * m10;
  R id = (R)1 / det;
  R i00 = id * m11, i01 = id * (-m01), i10 = id * (-m10), i11 = id * m00;
  R b0 = DOT(e, r), b1 = DOT(en, r);
  *s = i00 * b0 + i01 * b1;
  *t = i10 * b0 + i11 * b1;
}

/* distance.py:4-47 d_type_PT: 0/1/2 vertex t0/t1/t2, 3/4/5 edge t0t1/t1t2/t2t0,
   6 face interior. */
int FN(d_type_pt)(const R* p, const R* t0, const R* t1, const R* t2) {
  R e1[3], e2[3], n[3];
  SUB(t1, t0, e1);
  SUB(t2, t0, e2);
  CROSS(e1, e2, n);
  R s0, u0, s1, u1, s2, u2;
  CAT(edge_param_, S)(p, t0, t1, n, &s0, &u0);
  if (s0 > (R)0 && s0 < (R)1 && u0 >= (R)0) return 3;
  CAT(edge_param_, S)(p, t1, t2, n, &s1, &u1);
  if (s1 > (R)0 && s1 < (R)1 && u1 >= (R)0) return 4;
  CAT(edge_param_, S)(p, t2, t0, n, &s2, &u2);
  if (s2 > (R)0 && s2 < (R)1 && u2 >= (R)0) return 5;
  if (s0 <= (R)0 && s2 >= (R)1) return 0;
  if (s1 <= (R)0 && s0 >= (R)1) return 1;
  if (s2 <= (R)0 && s1 >= (R)1) return 2;
  return 6;
}

/* Squared point-point distance and gradient: d = |p-a|^2, dd/dp = 2(p-a). */
static inline R CAT(pp_, S)(const R* p, const R* a, R* gp, R* ga) {
  R d[3];
  SUB(p, a, d);
  for (int c = 0; c < 3; ++c) { gp[c] = (R)2 * d[c]; ga[c] = -gp[c]; }
  return DOT(d, d);
}

/* Squared point-line distance d = |u x w|^2 / |w - u|^2 with u = a - p, w = b - p
   (distance.py:105-109).  Gradient: with c = u x w and L = |w-u|^2,
   dd/du = 2 (w x c)/L + 2 d (w-u)/L,  dd/dw = 2 (c x u)/L - 2 d (w-u)/L,
   dd/dp = -(dd/du + dd/dw). */
static inline R CAT(pe_, S)(const R* p, const R* a, const R* b, R* gp, R* ga, R* gb) {
  R u[3], w[3], c[3], e[3], wc[3], cu[3];
  SUB(a, p, u);
  SUB(b, p, w);
  CROSS(u, w, c);
  SUB(b, a, e);
  R L = DOT(e, e);
  R d = DOT(c, c) / L;
  CROSS(w, c, wc);
  CROSS(c, u, cu);
  R twoL = (R)2 / L;
  R q = twoL * d;
  for (int k = 0; k < 3; ++k) {
    ga[k] = twoL * wc[k] + q * e[k];
    gb[k] = twoL * cu[k] - q * e[k];
    gp[k] = -(ga[k] + gb[k]);
  }
  return d;
}

/* Squared point-plane distance d = ((p-a).n)^2 / |n|^2, n = (b-a) x (c-a)
   (distance.py:111-116).  With r = p - a, s = r.n, N = n.n, q = s/N:
   dd/dp = 2 q n,  dd/db = 2 q (e2 x r) - 2 q^2 (e2 x n),
   dd/dc = 2 q (r x e1) - 2 q^2 (n x e1),  dd/da = -(sum of the others). */
static inline R CAT(pt_, S)(const R* p, const R* a, const R* b, const R* c3, R* gp, R* ga, R* gb, R* gc) {
  R e1[3], e2[3], n[3], r[3], e2r[3], e2n[3], re1[3], ne1[3];
  SUB(b, a, e1);
  SUB(c3, a, e2);
  CROSS(e1, e2, n);
  SUB(p, a, r);
  R s = DOT(r, n);
  R N = DOT(n, n);
  R d = (s * s) / N;
  R q = s / N;
  R q2 = (R)2 * q, qq2 = ((R)2 * q) * q;
  CROSS(e2, r, e2r);
  CROSS(e2, n, e2n);
  CROSS(r, e1, re1);
  CROSS(n, e1, ne1);
  for (int k = 0; k < 3; ++k) {
    gp[k] = q2 * n[k];
    gb[k] = q2 * e2r[k] - qq2 * e2n[k];
    gc[k] = q2 * re1[k] - qq2 * ne1[k];
    ga[k] = -((gp[k] + gb[k]) + gc[k]);
  }
  return d;
}

/* Squared distance of type dtype and its gradient; g12 = [g_p, g_t0, g_t1, g_t2]
   with zeros for points not in the pair (collision_constraints_x.py:24-52). */
R FN(dist)(int dtype, const R* p, const R* t0, const R* t1, const R* t2, R* g) {
  for (int k = 0; k < 12; ++k) g[k] = 0;
  switch (dtype) {
    case 0: return CAT(pp_, S)(p, t0, g, g + 3);
    case 1: return CAT(pp_, S)(p, t1, g, g + 6);
    case 2: return CAT(pp_, S)(p, t2, g, g + 9);
    case 3: return CAT(pe_, S)(p, t0, t1, g, g + 3, g + 6);
    case 4: return CAT(pe_, S)(p, t1, t2, g, g + 6, g + 9);
    case 5: return CAT(pe_, S)(p, t0, t2, g, g + 3, g + 9);
    default: return CAT(pt_, S)(p, t0, t1, t2, g, g + 3, g + 6, g + 9);
  }
}

/* Slots of the pair's points among (p, t0, t1, t2) for each type. */
static const unsigned char CAT(slots_, S)[7] = {0x3, 0x5, 0x9, 0x7, 0xD, 0xB, 0xF};

/* collision_constraints_x.py:5-69 __vt_st: dynamic vertex vs static triangle. */
/* Pair record of query q (collision_constraints_x.py:60-68, 164-172, 301-309):
   while the query holds fewer than cap records, store (fid, dtype), g0..g3 and
   schur at its next slot.  pi/pg/ps null: count only.  Layout [query][cap] as
   the reference's pair fields: pi int[2], pg R[12], ps R[1] per slot. */
static void CAT(record_, S)(int q, int cap, int* pnum, int* pi, R* pg, R* ps, int fid, int dtype,
                            const R* g, R schur) {
  int j = pnum[q];
  if (j >= cap) return;
  if (pi) {
    size_t at = (size_t)q * (size_t)cap + (size_t)j;
    pi[2 * at] = fid;
    pi[2 * at + 1] = dtype;
    for (int c = 0; c < 12; ++c) pg[12 * at + c] = g[c];
    ps[at] = schur;
  }
  pnum[q] = j + 1;
}

static void CAT(vt_st_, S)(int vd, int fs, const int* faces_st, const R* y, const R* xs,
                           const R* m_inv, R dhat, R k, R* dx, R* nc, int cap, int* pnum,
                           int* pi, R* pg, R* ps) {
  const R* p = y + 3 * vd;
  const R *t0 = xs + 3 * faces_st[3 * fs], *t1 = xs + 3 * faces_st[3 * fs + 1],
          *t2 = xs + 3 * faces_st[3 * fs + 2];
  int dtype = FN(d_type_pt)(p, t0, t1, t2);
  R g[12];
  R d = FN(dist)(dtype, p, t0, t1, t2, g);
  if (d < dhat) {
    R w = m_inv[vd];
    R schur = w * DOT(g, g);
    R ld = (k * (dhat - d)) / ((k * schur) + (R)1);
    for (int c = 0; c < 3; ++c) ACC(dx[3 * vd + c], (w * ld) * g[c]);
    ACC(nc[vd], (R)1);
    CAT(record_, S)(vd, cap, pnum, pi, pg, ps, fs, dtype, g, schur);
  }
}

/* collision_constraints_x.py:71-173 __tv_st: static vertex vs dynamic triangle.
   Only the triangle's vertices move; the pair is recorded whatever d is. */
static void CAT(tv_st_, S)(int fd, int vs, const int* faces_dy, R* y, const R* xs,
                           const R* fixed, const R* m_inv, R dhat, R k, R* dx, R* nc, int cap,
                           int* pnum, int* pi, R* pg, R* ps) {
  const R* p = xs + 3 * vs;
  int vid[4] = {-1, faces_dy[3 * fd], faces_dy[3 * fd + 1], faces_dy[3 * fd + 2]};
  const R *t0 = y + 3 * vid[1], *t1 = y + 3 * vid[2], *t2 = y + 3 * vid[3];
  int dtype = FN(d_type_pt)(p, t0, t1, t2);
  R g[12];
  R d = FN(dist)(dtype, p, t0, t1, t2, g);
  if (d < dhat) {
    unsigned m = CAT(slots_, S)[dtype];
    R schur = 0;
    int first = 1;
    for (int s = 1; s < 4; ++s) {
      if (!(m & (1u << s))) continue;
      int v = vid[s];
      R term = (fixed[v] * m_inv[v]) * DOT(g + 3 * s, g + 3 * s);
      schur = first ? term : schur + term;
      first = 0;
    }
    R ld = (k * (dhat - d)) / ((k * schur) + (R)1);
    for (int s = 1; s < 4; ++s) {
      if (!(m & (1u << s))) continue;
      int v = vid[s];
      R a = (fixed[v] * m_inv[v]) * ld;
      for (int c = 0; c < 3; ++c) ACC(dx[3 * v + c], a * g[3 * s + c]);
    }
    for (int s = 1; s < 4; ++s)
      if (m & (1u << s)) ACC(nc[vid[s]], (R)1);
    CAT(record_, S)(vs, cap, pnum, pi, pg, ps, fd, dtype, g, schur);
    return;
  }
  /* out of range: recorded with zero gradients and schur (collision_constraints_x.py:86-87) */
  R z[12] = {0};
  CAT(record_, S)(vs, cap, pnum, pi, pg, ps, fd, dtype, z, (R)0);
}

/* collision_constraints_x.py:175-310 __vt_dy: dynamic vertex vs dynamic triangle
   (self contact), hard constraint with ld = (dhat - d) / (sum f w |g|^2 + 1e-4).
   Type 0 omits f on both points' deltas (collision_constraints_x.py:199-200). */
static void CAT(vt_dy_, S)(int vd, int fd, const int* faces_dy, R* y, const R* fixed,
                           const R* m_inv, R dhat, R* dx, R* nc, int cap, int* pnum, int* pi, R* pg,
                           R* ps) {
  int vid[4] = {vd, faces_dy[3 * fd], faces_dy[3 * fd + 1], faces_dy[3 * fd + 2]};
  const R *p = y + 3 * vid[0], *t0 = y + 3 * vid[1], *t1 = y + 3 * vid[2], *t2 = y + 3 * vid[3];
  int dtype = FN(d_type_pt)(p, t0, t1, t2);
  R g[12];
  R d = FN(dist)(dtype, p, t0, t1, t2, g);
  if (d < dhat) {
    unsigned m = CAT(slots_, S)[dtype];
    R schur = 0;
    int first = 1;
    for (int s = 0; s < 4; ++s) {
      if (!(m & (1u << s))) continue;
      int v = vid[s];
      R term = (fixed[v] * m_inv[v]) * DOT(g + 3 * s, g + 3 * s);
      schur = first ? term : schur + term;
      first = 0;
    }
    schur = schur + (R)1e-4f;
    R ld = (dhat - d) / schur;
    for (int s = 0; s < 4; ++s) {
      if (!(m & (1u << s))) continue;
      int v = vid[s];
      R a = (dtype == 0) ? m_inv[v] * ld : (fixed[v] * m_inv[v]) * ld;
      for (int c = 0; c < 3; ++c) ACC(dx[3 * v + c], a * g[3 * s + c]);
    }
    for (int s = 0; s < 4; ++s)
      if (m & (1u << s)) ACC(nc[vid[s]], (R)1);
    CAT(record_, S)(vd, cap, pnum, pi, pg, ps, fd, dtype, g, schur);
    return;
  }
  R z[12] = {0};
  CAT(record_, S)(vd, cap, pnum, pi, pg, ps, fd, dtype, z, (R)0);
}

/* The three vertex-triangle loops of solve_collision_constraints_x
   (XPBD.py:236-264), dx/nc/pair counts zeroed first. */
static void CAT(vt_loops_, S)(int nv, int nvs, const int* faces_dy, const int* faces_st, R* y,
                              const R* xs, const R* fixed, const R* m_inv, R dhat, R k_col, int cap,
                              const int* vt_st_c, const int* vt_st_n, const int* vt_dy_c,
                              const int* vt_dy_n, const int* tv_st_c, const int* tv_st_n, R* dx, R* nc,
                              int* pn_vt_st, int* pn_vt_dy, int* pn_tv_st, int* pi0, R* pg0, R* ps0,
                              int* pi1, R* pg1, R* ps1, int* pi2, R* pg2, R* ps2) {
  PFOR
  for (int i = 0; i < 3 * nv; ++i) dx[i] = 0;
  PFOR
  for (int i = 0; i < nv; ++i) { nc[i] = 0; pn_vt_st[i] = 0; pn_vt_dy[i] = 0; }
  for (int i = 0; i < nvs; ++i) pn_tv_st[i] = 0;
  if (vt_st_c) {
    PFOR
    for (int v = 0; v < nv; ++v) {
      int n = vt_st_n[v] < cap ? vt_st_n[v] : cap;
      for (int j = 0; j < n; ++j)
        CAT(vt_st_, S)(v, vt_st_c[v * cap + j], faces_st, y, xs, m_inv, dhat, k_col, dx, nc, cap, pn_vt_st,
                       pi0, pg0, ps0);
    }
  }
  if (vt_dy_c) {
    PFOR
    for (int v = 0; v < nv; ++v) {
      int n = vt_dy_n[v] < cap ? vt_dy_n[v] : cap;
      for (int j = 0; j < n; ++j) {
        int f = vt_dy_c[v * cap + j];
        if (faces_dy[3 * f] == v || faces_dy[3 * f + 1] == v || faces_dy[3 * f + 2] == v) continue;
        CAT(vt_dy_, S)(v, f, faces_dy, y, fixed, m_inv, dhat, dx, nc, cap, pn_vt_dy, pi1, pg1, ps1);
      }
    }
  }
  if (tv_st_c) {
    PFOR
    for (int v = 0; v < nvs; ++v) {
      int n = tv_st_n[v] < cap ? tv_st_n[v] : cap;
      for (int j = 0; j < n; ++j)
        CAT(tv_st_, S)(tv_st_c[v * cap + j], v, faces_dy, y, xs, fixed, m_inv, dhat, k_col, dx, nc, cap, pn_tv_st,
                       pi2, pg2, ps2);
    }
  }
}

/* XPBD.py:236-264 + update_dx.  pi_k/pg_k/ps_k (nullable) receive the pair
   records of query set k (0 vt_st, 1 vt_dy, 2 tv_st) for the velocity pass. */
void FN(collision_pass_rec)(int nv, int nvs, const int* faces_dy, const int* faces_st, R* y,
                            const R* xs, const R* fixed, const R* m_inv, R dhat, R k_col, int cap,
                            const int* vt_st_c, const int* vt_st_n, const int* vt_dy_c,
                            const int* vt_dy_n, const int* tv_st_c, const int* tv_st_n, R* dx, R* nc,
                            int* pn_vt_st, int* pn_vt_dy, int* pn_tv_st, int* pi0, R* pg0, R* ps0,
                            int* pi1, R* pg1, R* ps1, int* pi2, R* pg2, R* ps2) {
  CAT(vt_loops_, S)(nv, nvs, faces_dy, faces_st, y, xs, fixed, m_inv, dhat, k_col, cap, vt_st_c, vt_st_n,
                    vt_dy_c, vt_dy_n, tv_st_c, tv_st_n, dx, nc, pn_vt_st, pn_vt_dy, pn_tv_st, pi0, pg0, ps0, pi1,
                    pg1, ps1, pi2, pg2, ps2);
  CAT(update_dx_, S)(nv, y, fixed, dx, nc);
}

/* ---------------- edge-edge (distance.py:49-124; collision_constraints_x.py:312-507) ---- */

/* distance.py:49-99 d_type_EE: region of the closest points of segments
   (x0, x1) and (x2, x3) (Sunday's segment-segment parametrisation); 8 = both
   interiors, the only type the edge-edge constraints act on. */
int FN(d_type_ee)(const R* x0, const R* x1, const R* x2, const R* x3) {
  R u[3], v[3], w[3], uv[3];
  SUB(x1, x0, u);
  SUB(x3, x2, v);
  SUB(x0, x2, w);
  R a = DOT(u, u), b = DOT(u, v), c = DOT(v, v), d = DOT(u, w), e = DOT(v, w);
  R D = a * c - b * b;
  R tD = D, tN;
  int ty = 8;
  R sN = b * e - c * d;
  if (sN <= (R)0) {
    tN = e;
    tD = c;
    ty = 2;
  } else if (sN >= D) {
    tN = e + b;
    tD = c;
    ty = 5;
  } else {
    tN = a * e - b * d;
    CROSS(u, v, uv);
    if (tN > (R)0 && tN < tD && (DOT(uv, w) < (R)1e-4f || DOT(uv, uv) < ((R)1.0e-20f * a) * c)) {
      if (sN < D / (R)2) {
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
  if (tN <= (R)0) {
    if (-d <= (R)0) ty = 0;
    else if (-d >= a) ty = 3;
    else ty = 6;
  } else if (tN >= tD) {
    if ((-d + b) <= (R)0) ty = 1;
    else if ((-d + b) >= a) ty = 4;
    else ty = 7;
  }
  return ty;
}

/* distance.py:118-124 d_EE: squared distance of the lines through (x0, x1) and
   (x2, x3), d = s^2 / N with u = x1 - x0, v = x3 - x2, w = x2 - x0, b = u x v,
   s = w . b, N = b . b.  g[12] = its gradient (derived here): with
   ds/d(x0..x3) = (-b - v x w, v x w, b - w x u, w x u) and
   dN/d(x0..x3) = 2 (-(v x b), v x b, -(b x u), b x u),
   g_k = (2 s / N) ds_k - (s^2 / N^2) dN_k. */
R FN(dist_ee)(const R* x0, const R* x1, const R* x2, const R* x3, R* g) {
  R u[3], v[3], w[3], b[3], vw[3], wu[3], vb[3], bu[3];
  SUB(x1, x0, u);
  SUB(x3, x2, v);
  SUB(x2, x0, w);
  CROSS(u, v, b);
  R s = DOT(w, b), N = DOT(b, b);
  R d = (s * s) / N;
  CROSS(v, w, vw);
  CROSS(w, u, wu);
  CROSS(v, b, vb);
  CROSS(b, u, bu);
  R p = ((R)2 * s) / N, q = (s * s) / (N * N);
  for (int c = 0; c < 3; ++c) {
    g[c] = p * (-b[c] - vw[c]) - q * ((R)2 * -vb[c]);
    g[3 + c] = p * vw[c] - q * ((R)2 * vb[c]);
    g[6 + c] = p * (b[c] - wu[c]) - q * ((R)2 * -bu[c]);
    g[9 + c] = p * wu[c] - q * ((R)2 * bu[c]);
  }
  return d;
}

/* One edge pair, collision_constraints_x.py:312-424 (__ee_dy, st = 0) and
   :426-507 (__ee_st, st = 1).  Only type 8 acts, and only for edges further
   than 1e-3 from parallel (|x01 x x23| > 1e-3); then d = d_EE and, if d < dhat,
   the compliant projection ld = k (dhat - d) / (k schur + 1) moves the dynamic
   points (all four for __ee_dy, the first edge's two for __ee_st; schur over
   the same points).  Returns 1 (and g, schur) when d < dhat. */
static int CAT(ee_pair_, S)(int st, const int* vid, const R* p0, const R* p1, const R* p2, const R* p3,
                            const R* fixed, const R* m_inv, R dhat, R k, R* dx, R* nc, int* dtype_out, R* g,
                            R* schur_out) {
  int dtype = FN(d_type_ee)(p0, p1, p2, p3);
  *dtype_out = dtype;
  for (int c = 0; c < 12; ++c) g[c] = 0;
  *schur_out = 0;
  if (dtype != 8) return 0;
  R a01[3], a23[3], cr[3];
  SUB(p0, p1, a01);
  SUB(p2, p3, a23);
  CROSS(a01, a23, cr);
  if (!(SQRT(DOT(cr, cr)) > (R)1e-3f)) return 0;
  R d = FN(dist_ee)(p0, p1, p2, p3, g);
  if (!(d < dhat)) {
    for (int c = 0; c < 12; ++c) g[c] = 0;
    return 0;
  }
  int nmov = st ? 2 : 4;
  R schur = 0;
  for (int s_ = 0; s_ < nmov; ++s_) {
    R term = (fixed[vid[s_]] * m_inv[vid[s_]]) * DOT(g + 3 * s_, g + 3 * s_);
    schur = s_ == 0 ? term : schur + term;
  }
  R ld = (k * (dhat - d)) / ((k * schur) + (R)1);
  for (int s_ = 0; s_ < nmov; ++s_) {
    int v = vid[s_];
    R a = (fixed[v] * m_inv[v]) * ld;
    for (int c = 0; c < 3; ++c) ACC(dx[3 * v + c], a * g[3 * s_ + c]);
  }
  for (int s_ = 0; s_ < nmov; ++s_) ACC(nc[vid[s_]], (R)1);
  *schur_out = schur;
  return 1;
}

/* The edge-edge loops of solve_collision_constraints_x (XPBD.py:266-281): each
   dynamic edge e against the three edges (face_edge_indices) of each candidate
   face, static set (__ee_st) then dynamic set (__ee_dy, skipping edges that
   share a vertex with e, XPBD.py:130-137).  A contacting pair is counted in
   pn_*[e] (the reference writes past its 40-entry pair fields) and stored, while
   fewer than pcap are, as (other edge, type), g0..g3, schur at [e][pcap]. */
static void CAT(ee_loops_, S)(const R* y, const R* xs, const R* fixed, const R* m_inv, R dhat, R k, R* dx, R* nc,
                              const orc_ee* E) {
  for (int set = 0; set < 2; ++set) {
    const int* C = set == 0 ? E->st_c : E->dy_c;
    const int* Nn = set == 0 ? E->st_n : E->dy_n;
    int* pn = set == 0 ? E->pn_st : E->pn_dy;
    if (!C) continue;
    int* pi = set == 0 ? E->pi_st : E->pi_dy;
    R* pg = (R*)(set == 0 ? E->pg_st : E->pg_dy);
    R* ps = (R*)(set == 0 ? E->ps_st : E->ps_dy);
    const int* fe = set == 0 ? E->fe_st : E->fe_dy;
    const int* E2 = set == 0 ? E->edges_st : E->edges_dy;
    const R* P2 = set == 0 ? xs : y;
    for (int e = 0; e < E->ne; ++e) {
      int n = Nn[e] < E->cap ? Nn[e] : E->cap;
      int v0 = E->edges_dy[2 * e], v1 = E->edges_dy[2 * e + 1];
      for (int j = 0; j < n; ++j) {
        int f = C[e * E->cap + j];
        for (int kk = 0; kk < 3; ++kk) {
          int e2 = fe[3 * f + kk];
          int v2 = E2[2 * e2], v3 = E2[2 * e2 + 1];
          if (set == 1 && (v0 == v2 || v0 == v3 || v1 == v2 || v1 == v3)) continue;
          int vid[4] = {v0, v1, v2, v3};
          R g[12], schur;
          int dtype;
          if (!CAT(ee_pair_, S)(set == 0, vid, y + 3 * v0, y + 3 * v1, P2 + 3 * v2, P2 + 3 * v3, fixed, m_inv, dhat,
                                k, dx, nc, &dtype, g, &schur))
            continue;
          int at = pn[e];
          if (pi && at < E->pcap) {
            size_t r = (size_t)e * (size_t)E->pcap + (size_t)at;
            pi[2 * r] = e2;
            pi[2 * r + 1] = dtype;
            for (int c = 0; c < 12; ++c) pg[12 * r + c] = g[c];
            ps[r] = schur;
          }
          pn[e] = at + 1;
        }
      }
    }
  }
}

/* solve_collision_constraints_x with the edge-edge loops after the
   vertex-triangle ones (XPBD.py:236-281), then update_dx. */
void FN(collision_pass_ee)(int nv, int nvs, const int* faces_dy, const int* faces_st, R* y, const R* xs,
                           const R* fixed, const R* m_inv, R dhat, R k_col, int cap, const int* vt_st_c,
                           const int* vt_st_n, const int* vt_dy_c, const int* vt_dy_n, const int* tv_st_c,
                           const int* tv_st_n, R* dx, R* nc, int* pn_vt_st, int* pn_vt_dy, int* pn_tv_st, int* pi0,
                           R* pg0, R* ps0, int* pi1, R* pg1, R* ps1, int* pi2, R* pg2, R* ps2, const orc_ee* E) {
  CAT(vt_loops_, S)(nv, nvs, faces_dy, faces_st, y, xs, fixed, m_inv, dhat, k_col, cap, vt_st_c, vt_st_n,
                    vt_dy_c, vt_dy_n, tv_st_c, tv_st_n, dx, nc, pn_vt_st, pn_vt_dy, pn_tv_st, pi0, pg0, ps0, pi1,
                    pg1, ps1, pi2, pg2, ps2);
  for (int e = 0; e < E->ne; ++e) {
    if (E->pn_st) E->pn_st[e] = 0;
    if (E->pn_dy) E->pn_dy[e] = 0;
  }
  CAT(ee_loops_, S)(y, xs, fixed, m_inv, dhat, k_col, dx, nc, E);
  CAT(update_dx_, S)(nv, y, fixed, dx, nc);
}

void FN(collision_pass)(int nv, int nvs, const int* faces_dy, const int* faces_st, R* y,
                        const R* xs, const R* fixed, const R* m_inv, R dhat, R k_col, int cap,
                        const int* vt_st_c, const int* vt_st_n, const int* vt_dy_c,
                        const int* vt_dy_n, const int* tv_st_c, const int* tv_st_n, R* dx, R* nc,
                        int* pn_vt_st, int* pn_vt_dy, int* pn_tv_st) {
  FN(collision_pass_rec)(nv, nvs, faces_dy, faces_st, y, xs, fixed, m_inv, dhat, k_col, cap, vt_st_c,
                         vt_st_n, vt_dy_c, vt_dy_n, tv_st_c, tv_st_n, dx, nc, pn_vt_st, pn_vt_dy,
                         pn_tv_st, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL);
}

/* XPBD.py:418-441 solve_sewing_constraints_x, applied as a Jacobi pass of its
   own (zero dx/nc, the pairs in order, update_dx): each pair (v1, v2) pulls its
   points together, C = |y2 - y1|, n = (y1 - y2)/C, S = f1 w1 + f2 w2,
   ld = k C / (k S + 1), dx1 -= f1 w1 ld n, dx2 += f2 w2 ld n, nc += 1 on both.
   A coincident pair (C = 0, where the reference normalises a zero vector into
   NaN) adds no delta and still counts. */
void FN(sewing_pass)(int nv, int np, const int* pairs, R* y, const R* fixed, const R* m_inv, R k, R* dx,
                     R* nc) {
  for (int i = 0; i < 3 * nv; ++i) dx[i] = 0;
  for (int i = 0; i < nv; ++i) nc[i] = 0;
  for (int p = 0; p < np; ++p) {
    int v1 = pairs[2 * p], v2 = pairs[2 * p + 1];
    R d[3];
    SUB(y + 3 * v2, y + 3 * v1, d);
    R C = SQRT(DOT(d, d));
    R fw1 = fixed[v1] * m_inv[v1], fw2 = fixed[v2] * m_inv[v2];
    if (C > (R)0) {
      R inv = (R)1 / C;
      R n[3] = {-(inv * d[0]), -(inv * d[1]), -(inv * d[2])};     /* (y1 - y2) / C */
      R schur = fw1 + fw2;
      R ld = (k * C) / ((k * schur) + (R)1);
      for (int c = 0; c < 3; ++c) {
        ACC(dx[3 * v1 + c], -((fw1 * ld) * n[c]));
        ACC(dx[3 * v2 + c], (fw2 * ld) * n[c]);
      }
    }
    ACC(nc[v1], (R)1);
    ACC(nc[v2], (R)1);
  }
  CAT(update_dx_, S)(nv, y, fixed, dx, nc);
}

/* ---------------- velocity pass (collision_constraints_v.py, XPBD.py:283-311, 374-404) ---- */

/* XPBD.py:340-343: v = (1 - damping) * f * (y - x) / dt. */
void FN(compute_velocity)(int nv, const R* x, R* v, const R* y, const R* fixed, R damping, R dt) {
  R a = (R)1 - damping;
  PFOR
  for (int i = 0; i < nv; ++i)
    for (int c = 0; c < 3; ++c) v[3 * i + c] = ((a * fixed[i]) * (y[3 * i + c] - x[3 * i + c])) / dt;
}

/* XPBD.py:332-337: x += dt * v. */
void FN(update_x)(int nv, R* x, const R* v, R dt) {
  PFOR
  for (int i = 0; i < 3 * nv; ++i) x[i] = x[i] + dt * v[i];
}

static inline R CAT(norm3_, S)(const R* a) { return SQRT(DOT(a, a)); }

static inline void CAT(vadd_, S)(R* dv, R* nc, int v, const R* d) {
  for (int c = 0; c < 3; ++c) ACC(dv[3 * v + c], d[c]);
  ACC(nc[v], (R)1);
}

/* collision_constraints_v.py:5-134 __vt_st: dynamic vertex v0 against static
   triangle fs.  Normal part: ld = -Cv / schur, dv0 += w ld g0 (type 3 takes w
   from the static mesh's inverse masses at index v0, :64 -- 0 past the static
   vertex count, where the reference reads outside that field).  Friction: g0Tan =
   p - vTan0 with p the static contact point's velocity (weighted by gradient
   norms on edges and the face; type 5 uses vTan0 - p, :109), full correction if
   mu |Cv| > cTan, else mu times it. */
static void CAT(vel_vt_st_, S)(int v0, int fs, int dtype, const R* g, R schur, const int* faces_st,
                               const R* v, const R* vs, const R* m_inv, const R* m_inv_st, int nvs, R mu,
                               R* dv, R* nc) {
  int vid[4] = {v0, faces_st[3 * fs], faces_st[3 * fs + 1], faces_st[3 * fs + 2]};
  const R *u0 = v + 3 * v0, *u1 = vs + 3 * vid[1], *u2 = vs + 3 * vid[2], *u3 = vs + 3 * vid[3];
  R Cv = ((DOT(g, u0) + DOT(g + 3, u1)) + DOT(g + 6, u2)) + DOT(g + 9, u3);
  if (!(Cv < (R)0)) return;
  R ld = -Cv / schur;
  R m0 = m_inv[v0];
  R mn = m0;
  if (dtype == 3) mn = (m_inv_st && v0 < nvs) ? m_inv_st[v0] : (R)0;
  R d[3], vt[3], p[3], gt[3];
  for (int c = 0; c < 3; ++c) d[c] = (mn * ld) * g[c];
  CAT(vadd_, S)(dv, nc, v0, d);
  for (int c = 0; c < 3; ++c) vt[c] = u0[c] + (m0 * ld) * g[c];
  if (dtype <= 2) {
    const R* q = dtype == 0 ? u1 : (dtype == 1 ? u2 : u3);
    for (int c = 0; c < 3; ++c) p[c] = q[c];
  } else if (dtype <= 5) {
    const R *ga = dtype == 4 ? g + 6 : g + 3, *gb = dtype == 3 ? g + 6 : g + 9;
    const R *qa = dtype == 4 ? u2 : u1, *qb = dtype == 3 ? u2 : u3;
    R a = CAT(norm3_, S)(ga), b = CAT(norm3_, S)(gb);
    for (int c = 0; c < 3; ++c) p[c] = (a * qa[c] + b * qb[c]) / (a + b);
  } else {
    R a = CAT(norm3_, S)(g + 3), b = CAT(norm3_, S)(g + 6), cc = CAT(norm3_, S)(g + 9);
    for (int c = 0; c < 3; ++c) p[c] = ((a * u1[c] + b * u2[c]) + cc * u3[c]) / ((a + b) + cc);
  }
  for (int c = 0; c < 3; ++c) gt[c] = dtype == 5 ? vt[c] - p[c] : p[c] - vt[c];
  R dd = DOT(gt, gt);
  R cT = (R)0.5f * dd;
  R sT = m0 * dd + (R)1e-4f;
  R lT = cT / sT;
  R mu_l = (mu * CAT(fabs_, S)(Cv) > cT) ? (R)1 : mu;
  for (int c = 0; c < 3; ++c) d[c] = mu_l * ((m0 * lT) * gt[c]);
  CAT(vadd_, S)(dv, nc, v0, d);
}

/* Moving slots among (p, t0, t1, t2) per type, without the point. */
static const unsigned char CAT(tri_slots_, S)[7] = {0x2, 0x4, 0x8, 0x6, 0xC, 0xA, 0xE};

/* collision_constraints_v.py:136-326 __tv_st: static vertex vs0 against dynamic
   triangle fd.  Cv uses g3 for both t1 and t2 (:144).  Normal: dv_k += f w ld g_k;
   friction on vTan_k = v_k + f w g_k ld, against the static vertex's velocity. */
static void CAT(vel_tv_st_, S)(int vs0, int fd, int dtype, const R* g, R schur, const int* faces_dy,
                               const R* v, const R* vs, const R* fixed, const R* m_inv, R mu, R* dv,
                               R* nc) {
  int vid[4] = {vs0, faces_dy[3 * fd], faces_dy[3 * fd + 1], faces_dy[3 * fd + 2]};
  const R* u[4] = {vs + 3 * vs0, v + 3 * vid[1], v + 3 * vid[2], v + 3 * vid[3]};
  R Cv = ((DOT(g, u[0]) + DOT(g + 3, u[1])) + DOT(g + 9, u[2])) + DOT(g + 9, u[3]);
  if (!(Cv < (R)0)) return;
  R ld = -Cv / schur;
  unsigned m = CAT(tri_slots_, S)[dtype];
  R vt[4][3], d[3];
  for (int s = 1; s < 4; ++s) {
    if (!(m & (1u << s))) continue;
    int w = vid[s];
    for (int c = 0; c < 3; ++c) d[c] = ((fixed[w] * m_inv[w]) * ld) * g[3 * s + c];
    CAT(vadd_, S)(dv, nc, w, d);
  }
  for (int s = 1; s < 4; ++s) {
    if (!(m & (1u << s))) continue;
    int w = vid[s];
    for (int c = 0; c < 3; ++c) vt[s][c] = u[s][c] + ((fixed[w] * m_inv[w]) * g[3 * s + c]) * ld;
  }
  R gt[4][3], q[3];
  R cT, sT;
  if (dtype <= 2) {
    int s = dtype + 1, w = vid[s];
    for (int c = 0; c < 3; ++c) gt[s][c] = u[0][c] - vt[s][c];
    R dd = DOT(gt[s], gt[s]);
    cT = (R)0.5f * dd;
    sT = m_inv[w] * dd + (R)1e-4f;
  } else if (dtype <= 5) {
    int sa = dtype == 4 ? 2 : 1, sb = dtype == 3 ? 2 : 3;
    R a = CAT(norm3_, S)(g + 3 * sa), b = CAT(norm3_, S)(g + 3 * sb);
    R ab = a + b;
    for (int c = 0; c < 3; ++c) q[c] = u[0][c] - (a * vt[sa][c] + b * vt[sb][c]) / ab;
    for (int c = 0; c < 3; ++c) {
      gt[sa][c] = (a / ab) * q[c];
      gt[sb][c] = (b / ab) * q[c];
    }
    cT = (R)0.5f * DOT(q, q);
    sT = (m_inv[vid[sa]] * DOT(gt[sa], gt[sa]) + m_inv[vid[sb]] * DOT(gt[sb], gt[sb])) + (R)1e-4f;
  } else {
    R a = CAT(norm3_, S)(g + 3), b = CAT(norm3_, S)(g + 6), cc = CAT(norm3_, S)(g + 9);
    R abc = (a + b) + cc;
    for (int c = 0; c < 3; ++c) q[c] = u[0][c] - ((a * vt[1][c] + b * vt[2][c]) + cc * vt[3][c]) / abc;
    for (int c = 0; c < 3; ++c) {
      gt[1][c] = (a / abc) * q[c];
      gt[2][c] = (b / abc) * q[c];
      gt[3][c] = (cc / abc) * q[c];
    }
    cT = (R)0.5f * DOT(q, q);
    sT = ((m_inv[vid[1]] * DOT(gt[1], gt[1]) + m_inv[vid[2]] * DOT(gt[2], gt[2])) +
          m_inv[vid[3]] * DOT(gt[3], gt[3])) + (R)1e-4f;
  }
  R lT = cT / sT;
  R mu_l = (mu * CAT(fabs_, S)(Cv) > cT) ? (R)1 : mu;
  for (int s = 1; s < 4; ++s) {
    if (!(m & (1u << s))) continue;
    int w = vid[s];
    for (int c = 0; c < 3; ++c) d[c] = mu_l * ((m_inv[w] * lT) * gt[s][c]);
    CAT(vadd_, S)(dv, nc, w, d);
  }
}

/* collision_constraints_v.py:329-570 __vt_dy: dynamic vertex v0 against
   dynamic triangle fd.  Normal: dv_k += f w g_k ld; friction with the relative
   velocity of the two points (types 0-2) or of v0 and the weighted edge/face
   point (types 3-6).  Type 1 scales the face point's friction delta by the
   inverse mass of t0 instead of t1 (:388). */
static void CAT(vel_vt_dy_, S)(int v0, int fd, int dtype, const R* g, R schur, const int* faces_dy,
                               const R* v, const R* fixed, const R* m_inv, R mu, R* dv, R* nc) {
  int vid[4] = {v0, faces_dy[3 * fd], faces_dy[3 * fd + 1], faces_dy[3 * fd + 2]};
  const R* u[4] = {v + 3 * v0, v + 3 * vid[1], v + 3 * vid[2], v + 3 * vid[3]};
  R Cv = ((DOT(g, u[0]) + DOT(g + 3, u[1])) + DOT(g + 6, u[2])) + DOT(g + 9, u[3]);
  if (!(Cv < (R)0)) return;
  R ld = -Cv / schur;
  unsigned m = CAT(slots_, S)[dtype];
  R vt[4][3], d[3], gt[4][3], mt[4];
  for (int s = 0; s < 4; ++s) {
    if (!(m & (1u << s))) continue;
    int w = vid[s];
    for (int c = 0; c < 3; ++c) d[c] = ((fixed[w] * m_inv[w]) * g[3 * s + c]) * ld;
    CAT(vadd_, S)(dv, nc, w, d);
  }
  for (int s = 0; s < 4; ++s) {
    if (!(m & (1u << s))) continue;
    int w = vid[s];
    for (int c = 0; c < 3; ++c) vt[s][c] = u[s][c] + ((fixed[w] * m_inv[w]) * g[3 * s + c]) * ld;
    mt[s] = m_inv[w];
  }
  R cT, sT;
  if (dtype <= 2) {
    int s = dtype + 1;
    for (int c = 0; c < 3; ++c) {
      gt[0][c] = -(vt[0][c] - vt[s][c]);
      gt[s][c] = vt[0][c] - vt[s][c];
    }
    R d0 = DOT(gt[0], gt[0]), d1 = DOT(gt[s], gt[s]);
    cT = (R)0.5f * (d1 + d0);
    sT = (mt[0] * d0 + mt[s] * d1) + (R)1e-4f;
    if (dtype == 1) mt[s] = m_inv[vid[1]];
  } else if (dtype <= 5) {
    int sa = dtype == 4 ? 2 : 1, sb = dtype == 3 ? 2 : 3;
    R a = CAT(norm3_, S)(g + 3 * sa), b = CAT(norm3_, S)(g + 3 * sb);
    R ab = a + b;
    for (int c = 0; c < 3; ++c) gt[0][c] = (a * vt[sa][c] + b * vt[sb][c]) / ab - vt[0][c];
    for (int c = 0; c < 3; ++c) {
      gt[sa][c] = (-(a / ab)) * gt[0][c];
      gt[sb][c] = (-(b / ab)) * gt[0][c];
    }
    R d0 = DOT(gt[0], gt[0]), da = DOT(gt[sa], gt[sa]), db = DOT(gt[sb], gt[sb]);
    cT = (R)0.5f * ((d0 + da) + db);
    sT = ((mt[0] * d0 + mt[sa] * da) + mt[sb] * db) + (R)1e-4f;
  } else {
    R a = CAT(norm3_, S)(g + 3), b = CAT(norm3_, S)(g + 6), cc = CAT(norm3_, S)(g + 9);
    R abc = (a + b) + cc;
    for (int c = 0; c < 3; ++c)
      gt[0][c] = ((a * vt[1][c] + b * vt[2][c]) + cc * vt[3][c]) / abc - vt[0][c];
    for (int c = 0; c < 3; ++c) {
      gt[1][c] = (-(a / abc)) * gt[0][c];
      gt[2][c] = (-(b / abc)) * gt[0][c];
      gt[3][c] = (-(cc / abc)) * gt[0][c];
    }
    R d0 = DOT(gt[0], gt[0]), d1 = DOT(gt[1], gt[1]), d2 = DOT(gt[2], gt[2]), d3 = DOT(gt[3], gt[3]);
    cT = (R)0.5f * (((d0 + d1) + d2) + d3);
    sT = (((mt[0] * d0 + mt[1] * d1) + mt[2] * d2) + mt[3] * d3) + (R)1e-4f;
  }
  R lT = cT / sT;
  R mu_l = (mu * CAT(fabs_, S)(Cv) > cT) ? (R)1 : mu;
  for (int s = 0; s < 4; ++s) {
    if (!(m & (1u << s))) continue;
    for (int c = 0; c < 3; ++c) d[c] = mu_l * ((mt[s] * lT) * gt[s][c]);
    CAT(vadd_, S)(dv, nc, vid[s], d);
  }
}

/* solve_constraints_v (XPBD.py:374-381): dv = nc = 0; the recorded pairs of the
   three query sets in the reference's loop order (XPBD.py:283-311); update_dv
   (XPBD.py:397-404): dv /= nc where nc > 0, v += f dv where m_inv > 0.
   pn_k null: set k absent.  vs: static vertex velocities (zeros in the
   reference). */
static void CAT(vel_vt_loops_, S)(int nv, int nvs, const int* faces_dy, const int* faces_st, R* v, const R* vs,
                                  const R* fixed, const R* m_inv, const R* m_inv_st, R mu, int cap,
                                  const int* pn0, const int* pi0, const R* pg0, const R* ps0,
                                  const int* pn1, const int* pi1, const R* pg1, const R* ps1,
                                  const int* pn2, const int* pi2, const R* pg2, const R* ps2, R* dv, R* nc) {
  PFOR
  for (int i = 0; i < 3 * nv; ++i) dv[i] = 0;
  PFOR
  for (int i = 0; i < nv; ++i) nc[i] = 0;
  if (pn0) {
    PFOR
    for (int q = 0; q < nv; ++q)
      for (int j = 0; j < pn0[q]; ++j) {
        size_t at = (size_t)q * (size_t)cap + (size_t)j;
        CAT(vel_vt_st_, S)(q, pi0[2 * at], pi0[2 * at + 1], pg0 + 12 * at, ps0[at], faces_st, v, vs, m_inv,
                           m_inv_st, nvs, mu, dv, nc);
      }
  }
  if (pn1) {
    PFOR
    for (int q = 0; q < nv; ++q)
      for (int j = 0; j < pn1[q]; ++j) {
        size_t at = (size_t)q * (size_t)cap + (size_t)j;
        CAT(vel_vt_dy_, S)(q, pi1[2 * at], pi1[2 * at + 1], pg1 + 12 * at, ps1[at], faces_dy, v, fixed, m_inv,
                           mu, dv, nc);
      }
  }
  if (pn2) {
    PFOR
    for (int q = 0; q < nvs; ++q)
      for (int j = 0; j < pn2[q]; ++j) {
        size_t at = (size_t)q * (size_t)cap + (size_t)j;
        CAT(vel_tv_st_, S)(q, pi2[2 * at], pi2[2 * at + 1], pg2 + 12 * at, ps2[at], faces_dy, v, vs, fixed,
                           m_inv, mu, dv, nc);
      }
  }
}

/* update_dv (XPBD.py:397-404): dv /= nc where nc > 0, v += f dv where m_inv > 0. */
static void CAT(update_dv_, S)(int nv, R* v, const R* fixed, const R* m_inv, R* dv, const R* nc) {
  PFOR
  for (int i = 0; i < nv; ++i) {
    if (nc[i] > (R)0)
      for (int c = 0; c < 3; ++c) dv[3 * i + c] = dv[3 * i + c] / nc[i];
    if (m_inv[i] > (R)0)
      for (int c = 0; c < 3; ++c) v[3 * i + c] = v[3 * i + c] + fixed[i] * dv[3 * i + c];
  }
}

void FN(velocity_pass)(int nv, int nvs, const int* faces_dy, const int* faces_st, R* v, const R* vs,
                       const R* fixed, const R* m_inv, const R* m_inv_st, R mu, int cap,
                       const int* pn0, const int* pi0, const R* pg0, const R* ps0,
                       const int* pn1, const int* pi1, const R* pg1, const R* ps1,
                       const int* pn2, const int* pi2, const R* pg2, const R* ps2, R* dv, R* nc) {
  CAT(vel_vt_loops_, S)(nv, nvs, faces_dy, faces_st, v, vs, fixed, m_inv, m_inv_st, mu, cap, pn0, pi0, pg0, ps0,
                        pn1, pi1, pg1, ps1, pn2, pi2, pg2, ps2, dv, nc);
  CAT(update_dv_, S)(nv, v, fixed, m_inv, dv, nc);
}

/* collision_constraints_v.py:573-819 __ee_st (type 8, the only one recorded):
   Cv = g0.v0 + g1.v1 + g3.vs3 -- g2 does not enter (:782) -- and if Cv < 0,
   ld = -Cv / schur, dv_k += f_k w_k ld g_k on the dynamic edge's points.
   :821-878 __ee_dy: dvn = sum g_k.v_k and if dvn < 0, ld = dvn / schur,
   dv_k -= f_k w_k g_k ld on all four points.  Normal only, nc += 1 per point. */
static void CAT(vel_ee_loops_, S)(const R* v, const R* vs, const R* fixed, const R* m_inv, R* dv, R* nc,
                                  const orc_ee* E) {
  for (int set = 0; set < 2; ++set) {
    const int* pn = set == 0 ? E->pn_st : E->pn_dy;
    if (!pn) continue;
    const int* pi = set == 0 ? E->pi_st : E->pi_dy;
    const R* pg = (const R*)(set == 0 ? E->pg_st : E->pg_dy);
    const R* ps = (const R*)(set == 0 ? E->ps_st : E->ps_dy);
    for (int e = 0; e < E->ne; ++e) {
      int n = pn[e] < E->pcap ? pn[e] : E->pcap;
      int v0 = E->edges_dy[2 * e], v1 = E->edges_dy[2 * e + 1];
      for (int j = 0; j < n; ++j) {
        size_t r = (size_t)e * (size_t)E->pcap + (size_t)j;
        int e2 = pi[2 * r];
        const R* g = pg + 12 * r;
        R schur = ps[r];
        if (set == 0) {
          int v3 = E->edges_st[2 * e2 + 1];
          R Cv = (DOT(g, v + 3 * v0) + DOT(g + 3, v + 3 * v1)) + DOT(g + 9, vs + 3 * v3);
          if (!(Cv < (R)0)) continue;
          R ld = -Cv / schur;
          int vid[2] = {v0, v1};
          for (int s_ = 0; s_ < 2; ++s_) {
            R d[3], a = (fixed[vid[s_]] * m_inv[vid[s_]]) * ld;
            for (int c = 0; c < 3; ++c) d[c] = a * g[3 * s_ + c];
            CAT(vadd_, S)(dv, nc, vid[s_], d);
          }
        } else {
          int vid[4] = {v0, v1, E->edges_dy[2 * e2], E->edges_dy[2 * e2 + 1]};
          R dvn = ((DOT(g, v + 3 * vid[0]) + DOT(g + 3, v + 3 * vid[1])) + DOT(g + 6, v + 3 * vid[2])) +
                  DOT(g + 9, v + 3 * vid[3]);
          if (!(dvn < (R)0)) continue;
          R ld = dvn / schur;
          for (int s_ = 0; s_ < 4; ++s_) {
            R d[3], a = fixed[vid[s_]] * m_inv[vid[s_]];
            for (int c = 0; c < 3; ++c) d[c] = -((a * g[3 * s_ + c]) * ld);
            CAT(vadd_, S)(dv, nc, vid[s_], d);
          }
        }
      }
    }
  }
}

/* solve_constraints_v with the edge-edge pairs after the vertex-triangle ones
   (XPBD.py:283-330), then update_dv. */
void FN(velocity_pass_ee)(int nv, int nvs, const int* faces_dy, const int* faces_st, R* v, const R* vs,
                          const R* fixed, const R* m_inv, const R* m_inv_st, R mu, int cap,
                          const int* pn0, const int* pi0, const R* pg0, const R* ps0,
                          const int* pn1, const int* pi1, const R* pg1, const R* ps1,
                          const int* pn2, const int* pi2, const R* pg2, const R* ps2, R* dv, R* nc,
                          const orc_ee* E) {
  CAT(vel_vt_loops_, S)(nv, nvs, faces_dy, faces_st, v, vs, fixed, m_inv, m_inv_st, mu, cap, pn0, pi0, pg0, ps0,
                        pn1, pi1, pg1, ps1, pn2, pi2, pg2, ps2, dv, nc);
  CAT(vel_ee_loops_, S)(v, vs, fixed, m_inv, dv, nc, E);
  CAT(update_dv_, S)(nv, v, fixed, m_inv, dv, nc);
}

#undef FN
#undef SQRT
#undef DOT
#undef SUB
#undef CROSS
