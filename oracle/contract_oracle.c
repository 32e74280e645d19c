/*
 * contract_oracle.c -- CPU FP64 oracle for the propagator-contraction path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library
 * (contract_amd/csrc) is compared against.  Only tests/, the smoke check in
 * __graft_entry__.py and the cpu_baseline leg of bench.py may load it.  The
 * product path never calls into it and has no CPU fallback.
 *
 * It restates, in plain C99, the per-site semantics of the reference's
 * contraction kernels, written from their closed forms rather than from the
 * kernels' lane loops:
 *
 *   layout   P[site][s1*4+s2][c1*3+c2] complex128, s1 = sink spin, s2 = source
 *            spin, c1 = sink colour, c2 = source colour
 *            (include/load_store.cuh:99-109, include/meson.cuh:33-34)
 *   gamma    G(n) = g1^{n&1} g2^{(n>>1)&1} g3^{(n>>2)&1} g4^{(n>>3)&1} in the
 *            DeGrand-Rossi basis (include/gamma.cuh:41-127,176-271)
 *   epsilon  ee(X,Y,Z) = sum eps_abc eps_def X[a,d] Y[b,e] Z[c,f]
 *            (include/kernel.cuh:28-44)
 *
 * Every function states the reference lines it follows.  Sums over sites that
 * the new fused API adds (timeslice / momentum projection) use Neumaier
 * compensated FP64 accumulation.
 *
 * Parity pins (tests/test_oracle_*.py): dense numpy einsum golden vectors,
 * the reference's own gamma tables, the reference library run on a GPU
 * (tests/golden/ref_vectors.npz), and identities that tie entry points
 * together (general == sparse, diquark+meson == baryon, sum P*seq == C).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  double re, im;
} cplx;

static inline cplx cx(double re, double im)
{
  cplx z = {re, im};
  return z;
}
static inline cplx cadd(cplx a, cplx b) { return cx(a.re + b.re, a.im + b.im); }
static inline cplx csub(cplx a, cplx b) { return cx(a.re - b.re, a.im - b.im); }
static inline cplx cmul(cplx a, cplx b) { return cx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline cplx cconj(cplx a) { return cx(a.re, -a.im); }
static inline int cnz(cplx a) { return a.re != 0.0 || a.im != 0.0; }

enum { NS = 4, NC = 3, SITE = 144 };

/* ------------------------------------------------------------------ */
/* Gamma matrices (include/gamma.cuh:41-127 sparse tables, 176-223 dense) */
/* ------------------------------------------------------------------ */

/* The four Euclidean DeGrand-Rossi matrices, rows = sink index. */
static void basis_gamma(int mu, cplx out[4][4])
{
  memset(out, 0, sizeof(cplx) * 16);
  switch (mu) {
  case 0: /* gamma_1 */
    out[0][3] = cx(0, 1); out[1][2] = cx(0, 1); out[2][1] = cx(0, -1); out[3][0] = cx(0, -1);
    break;
  case 1: /* gamma_2 */
    out[0][3] = cx(-1, 0); out[1][2] = cx(1, 0); out[2][1] = cx(1, 0); out[3][0] = cx(-1, 0);
    break;
  case 2: /* gamma_3 */
    out[0][2] = cx(0, 1); out[1][3] = cx(0, -1); out[2][0] = cx(0, -1); out[3][1] = cx(0, 1);
    break;
  default: /* gamma_4 */
    out[0][2] = cx(1, 0); out[1][3] = cx(1, 0); out[2][0] = cx(1, 0); out[3][1] = cx(1, 0);
    break;
  }
}

/* Dense G(n), n = 0..15: product of the basis matrices selected by the bits of
 * n, multiplied left to right in the order g1 g2 g3 g4. */
void oracle_gamma(int n, double *out /* [4][4][2] */)
{
  cplx acc[4][4], tmp[4][4], b[4][4];
  memset(acc, 0, sizeof(acc));
  for (int r = 0; r < 4; ++r) acc[r][r] = cx(1, 0);
  for (int mu = 0; mu < 4; ++mu) {
    if (!((n >> mu) & 1)) continue;
    basis_gamma(mu, b);
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) {
        cplx s = cx(0, 0);
        for (int k = 0; k < 4; ++k) s = cadd(s, cmul(acc[r][k], b[k][c]));
        tmp[r][c] = s;
      }
    memcpy(acc, tmp, sizeof(acc));
  }
  memcpy(out, acc, sizeof(acc));
}

static void gamma_dense(int n, cplx g[4][4]) { oracle_gamma(n, (double *)g); }

/* gamma_5 = G(15) = diag(1, 1, -1, -1). */
static double gamma5_diag(int s) { return s < 2 ? 1.0 : -1.0; }

/* ------------------------------------------------------------------ */
/* Colour epsilon-epsilon (include/kernel.cuh:28-44)                   */
/* ------------------------------------------------------------------ */

static const int PERM[6][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}, {0, 2, 1}, {2, 1, 0}, {1, 0, 2}};
static const double PSIGN[6] = {1, 1, 1, -1, -1, -1};

/* ee(X,Y,Z) = sum_{abc,def} eps_abc eps_def X[a,d] Y[b,e] Z[c,f]; 3x3 colour
 * blocks stored row-major [c_sink*3 + c_source]. */
static cplx epseps(const cplx *X, const cplx *Y, const cplx *Z)
{
  cplx s = cx(0, 0);
  for (int p = 0; p < 6; ++p) {
    const int a = PERM[p][0], b = PERM[p][1], c = PERM[p][2];
    for (int q = 0; q < 6; ++q) {
      const int d = PERM[q][0], e = PERM[q][1], f = PERM[q][2];
      cplx t = cmul(cmul(X[a * 3 + d], Y[b * 3 + e]), Z[c * 3 + f]);
      const double sg = PSIGN[p] * PSIGN[q];
      s.re += sg * t.re;
      s.im += sg * t.im;
    }
  }
  return s;
}

/* Colour cofactor with one slot open: out[a,d] = sum eps_abc eps_def Y[b,e] Z[c,f],
 * so that ee(X,Y,Z) = sum_ad X[a,d] out[a,d]. */
static void epseps_open(cplx out[9], const cplx *Y, const cplx *Z)
{
  for (int i = 0; i < 9; ++i) out[i] = cx(0, 0);
  for (int p = 0; p < 6; ++p) {
    const int a = PERM[p][0], b = PERM[p][1], c = PERM[p][2];
    for (int q = 0; q < 6; ++q) {
      const int d = PERM[q][0], e = PERM[q][1], f = PERM[q][2];
      cplx t = cmul(Y[b * 3 + e], Z[c * 3 + f]);
      const double sg = PSIGN[p] * PSIGN[q];
      out[a * 3 + d].re += sg * t.re;
      out[a * 3 + d].im += sg * t.im;
    }
  }
}

static inline const cplx *blk(const cplx *site, int s1, int s2) { return site + (s1 * 4 + s2) * 9; }

/* ------------------------------------------------------------------ */
/* Meson two-point (include/meson.cuh:32-50, include/meson_all.cuh:52-103) */
/* ------------------------------------------------------------------ */

/* c(x) = Tr_s Tr_c[(g5 Gij) P_i (g5 Gkl) P_j^dagger]
 *      = sum_{ijkl} (g5 Gij)[j,i] (g5 Gkl)[k,l] sum_ad P_i[i,k](a,d) conj(P_j[j,l](a,d)).
 * The reference reads these factors from gamma_gamma5_data<true>, i.e.
 * (g5 G)^T (include/gamma.cuh:249-269), at lane (i,l) (meson.cuh:38-49). */
static cplx meson_site(const cplx *pi, const cplx *pj, const cplx A[4][4], const cplx B[4][4])
{
  cplx s = cx(0, 0);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      if (!cnz(A[j][i])) continue;
      for (int l = 0; l < 4; ++l)
        for (int k = 0; k < 4; ++k) {
          if (!cnz(B[k][l])) continue;
          const cplx *x = blk(pi, i, k), *y = blk(pj, j, l);
          cplx t = cx(0, 0);
          for (int ad = 0; ad < 9; ++ad) t = cadd(t, cmul(x[ad], cconj(y[ad])));
          s = cadd(s, cmul(cmul(A[j][i], B[k][l]), t));
        }
    }
  return s;
}

static void g5_times(int g, cplx out[4][4])
{
  cplx G[4][4];
  gamma_dense(g, G);
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) out[r][c] = cx(gamma5_diag(r) * G[r][c].re, gamma5_diag(r) * G[r][c].im);
}

/* meson_two_point (include/contract.h:16, lib/meson/meson.cu:6-18) */
void oracle_meson_two_point(double *correl, const double *propag_i, const double *propag_j, size_t volume, int gamma_ij,
                            int gamma_kl)
{
  cplx A[4][4], B[4][4];
  g5_times(gamma_ij, A);
  g5_times(gamma_kl, B);
  const cplx *pi = (const cplx *)propag_i, *pj = (const cplx *)propag_j;
  cplx *out = (cplx *)correl;
#pragma omp parallel for schedule(static)
  for (long x = 0; x < (long)volume; ++x) out[x] = meson_site(pi + x * SITE, pj + x * SITE, A, B);
}

/* meson_all_source_two_point: out[g](x) = meson(Gij = gamma_ij, Gkl = g)
 * (include/meson_all.cuh:66-72); correl is [16][volume]. */
void oracle_meson_all_source_two_point(double *correl, const double *propag_i, const double *propag_j, size_t volume,
                                       int gamma_ij)
{
  for (int g = 0; g < 16; ++g)
    oracle_meson_two_point(correl + (size_t)g * volume * 2, propag_i, propag_j, volume, gamma_ij, g);
}

/* meson_all_sink_two_point: out[g](x) = meson(Gij = g, Gkl = gamma_kl)
 * (include/meson_all.cuh:96-102). */
void oracle_meson_all_sink_two_point(double *correl, const double *propag_i, const double *propag_j, size_t volume,
                                     int gamma_kl)
{
  for (int g = 0; g < 16; ++g)
    oracle_meson_two_point(correl + (size_t)g * volume * 2, propag_i, propag_j, volume, g, gamma_kl);
}

/* ------------------------------------------------------------------ */
/* Diquark (include/diquark.cuh:32-58)                                 */
/* ------------------------------------------------------------------ */

/* D[m,l](a,d) = g5[m] g5[l] conj( sum_{ijk} Gij[i,j] Gkl[k,l]
 *               sum_{bc,ef} eps_abc eps_def P_i[i,k](b,e) P_j[j,m](c,f) ). */
void oracle_baryon_diquark(double *diquark, const double *propag_i, const double *propag_j, size_t volume, int gamma_ij,
                           int gamma_kl)
{
  cplx Gij[4][4], Gkl[4][4];
  gamma_dense(gamma_ij, Gij);
  gamma_dense(gamma_kl, Gkl);
  const cplx *pi = (const cplx *)propag_i, *pj = (const cplx *)propag_j;
  cplx *out = (cplx *)diquark;
#pragma omp parallel for schedule(static)
  for (long x = 0; x < (long)volume; ++x) {
    const cplx *si = pi + x * SITE, *sj = pj + x * SITE;
    cplx *so = out + x * SITE;
    for (int m = 0; m < 4; ++m)
      for (int l = 0; l < 4; ++l) {
        cplx acc[9], cof[9];
        for (int ad = 0; ad < 9; ++ad) acc[ad] = cx(0, 0);
        for (int i = 0; i < 4; ++i)
          for (int j = 0; j < 4; ++j) {
            if (!cnz(Gij[i][j])) continue;
            for (int k = 0; k < 4; ++k) {
              if (!cnz(Gkl[k][l])) continue;
              epseps_open(cof, blk(si, i, k), blk(sj, j, m));
              cplx w = cmul(Gij[i][j], Gkl[k][l]);
              for (int ad = 0; ad < 9; ++ad) acc[ad] = cadd(acc[ad], cmul(w, cof[ad]));
            }
          }
        const double s5 = gamma5_diag(m) * gamma5_diag(l);
        for (int ad = 0; ad < 9; ++ad) {
          cplx v = cconj(acc[ad]);
          so[(m * 4 + l) * 9 + ad] = cx(s5 * v.re, s5 * v.im);
        }
      }
  }
}

/* ------------------------------------------------------------------ */
/* Baryon two-point (include/baryon.cuh:34-105)                        */
/* ------------------------------------------------------------------ */

/* The six contraction types name which source-spin index each propagator
 * carries: type "I<X>_J<Y>_N<Z>" contracts P_i[i,X] P_j[j,Y] P_n[n,Z] with
 * X,Y,Z a permutation of k,l,m (include/contract_enum.h:1-7).  Together with
 * the SWAP_IJ / SWAP_KL template switches of baryon.cuh:41-52,95-105 this is
 *   C(x) = sum Gij[i,j] Gkl[k,l] Gmn[m,n] ee(P_i[i,X], P_j[j,Y], P_n[n,Z]). */
enum { SK = 0, SL = 1, SM = 2 };
static const int SRC_SLOT[6][3] = {
  {SK, SL, SM}, /* IK_JL_NM */
  {SK, SM, SL}, /* IK_JM_NL */
  {SL, SK, SM}, /* IL_JK_NM */
  {SL, SM, SK}, /* IL_JM_NK */
  {SM, SK, SL}, /* IM_JK_NL */
  {SM, SL, SK}, /* IM_JL_NK */
};

/* One site with a dense 4x4 matrix W in the Gmn slot (W = G(gamma_mn) for the
 * sparse entry point, W = project_mn for the general one,
 * include/baryon_general.cuh:62-67,77-84). */
static cplx baryon_site(const cplx *pi, const cplx *pj, const cplx *pn, int type, const cplx Gij[4][4],
                        const cplx Gkl[4][4], const cplx W[4][4])
{
  const int *slot = SRC_SLOT[type];
  cplx s = cx(0, 0);
  int idx[3];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      if (!cnz(Gij[i][j])) continue;
      for (int k = 0; k < 4; ++k)
        for (int l = 0; l < 4; ++l) {
          if (!cnz(Gkl[k][l])) continue;
          const cplx gg = cmul(Gij[i][j], Gkl[k][l]);
          for (int m = 0; m < 4; ++m)
            for (int n = 0; n < 4; ++n) {
              if (!cnz(W[m][n])) continue;
              idx[SK] = k;
              idx[SL] = l;
              idx[SM] = m;
              cplx e = epseps(blk(pi, i, idx[slot[0]]), blk(pj, j, idx[slot[1]]), blk(pn, n, idx[slot[2]]));
              s = cadd(s, cmul(cmul(gg, W[m][n]), e));
            }
        }
    }
  return s;
}

static int valid_type(int t) { return t >= 0 && t < 6; }

/* baryon_two_point (include/contract.h:20-21, lib/baryon/baryon.cu:7-67) */
void oracle_baryon_two_point(double *correl, const double *propag_i, const double *propag_j, const double *propag_n,
                             int contract_type, size_t volume, int gamma_ij, int gamma_kl, int gamma_mn)
{
  if (!valid_type(contract_type)) return;
  cplx Gij[4][4], Gkl[4][4], Gmn[4][4];
  gamma_dense(gamma_ij, Gij);
  gamma_dense(gamma_kl, Gkl);
  gamma_dense(gamma_mn, Gmn);
  const cplx *pi = (const cplx *)propag_i, *pj = (const cplx *)propag_j, *pn = (const cplx *)propag_n;
  cplx *out = (cplx *)correl;
#pragma omp parallel for schedule(static)
  for (long x = 0; x < (long)volume; ++x)
    out[x] = baryon_site(pi + x * SITE, pj + x * SITE, pn + x * SITE, contract_type, Gij, Gkl, Gmn);
}

/* baryon_general_two_point (include/contract.h:22-24, lib/baryon/baryon_general.cu:7-51);
 * project_mn is row-major [m][n] (baryon_general.cuh:62-67). */
void oracle_baryon_general_two_point(double *correl, const double *propag_i, const double *propag_j,
                                     const double *propag_n, int contract_type, size_t volume, int gamma_ij,
                                     int gamma_kl, const double *project_mn)
{
  if (!valid_type(contract_type)) return;
  cplx Gij[4][4], Gkl[4][4], W[4][4];
  gamma_dense(gamma_ij, Gij);
  gamma_dense(gamma_kl, Gkl);
  memcpy(W, project_mn, sizeof(W));
  const cplx *pi = (const cplx *)propag_i, *pj = (const cplx *)propag_j, *pn = (const cplx *)propag_n;
  cplx *out = (cplx *)correl;
#pragma omp parallel for schedule(static)
  for (long x = 0; x < (long)volume; ++x)
    out[x] = baryon_site(pi + x * SITE, pj + x * SITE, pn + x * SITE, contract_type, Gij, Gkl, W);
}

/* ------------------------------------------------------------------ */
/* Baryon sequential source (include/baryon_sequential.cuh:36-267,     */
/* lib/contract.cpp:46-63)                                             */
/* ------------------------------------------------------------------ */

/* seq_X(x) = dC(x)/dP_X(x) element-wise over [16][9], X = i, j or n.  The
 * result overwrites P_X's buffer, which is never read (baryon_sequential.cuh:
 * 177-199, 214-236, 251-266).  Rows with no contribution are written as 0
 * (the zeroing at :110,136 and the dense-zero gamma factor at :92). */
void oracle_baryon_sequential_two_point(double *propag_i, double *propag_j, double *propag_n, int contract_type,
                                        int sequential_type, size_t volume, int gamma_ij, int gamma_kl, int gamma_mn)
{
  if (!valid_type(contract_type)) return;
  if (sequential_type < 0 || sequential_type > 2) return; /* lib/contract.cpp:60 */
  cplx Gij[4][4], Gkl[4][4], Gmn[4][4];
  gamma_dense(gamma_ij, Gij);
  gamma_dense(gamma_kl, Gkl);
  gamma_dense(gamma_mn, Gmn);
  cplx *P[3] = {(cplx *)propag_i, (cplx *)propag_j, (cplx *)propag_n};
  const int *slot = SRC_SLOT[contract_type];
  const int X = sequential_type;
#pragma omp parallel for schedule(static)
  for (long x = 0; x < (long)volume; ++x) {
    cplx out[SITE];
    for (int e = 0; e < SITE; ++e) out[e] = cx(0, 0);
    const cplx *site[3] = {P[0] + x * SITE, P[1] + x * SITE, P[2] + x * SITE};
    int idx[3], sink[3];
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        if (!cnz(Gij[i][j])) continue;
        for (int k = 0; k < 4; ++k)
          for (int l = 0; l < 4; ++l) {
            if (!cnz(Gkl[k][l])) continue;
            for (int m = 0; m < 4; ++m)
              for (int n = 0; n < 4; ++n) {
                if (!cnz(Gmn[m][n])) continue;
                const cplx w = cmul(cmul(Gij[i][j], Gkl[k][l]), Gmn[m][n]);
                idx[SK] = k;
                idx[SL] = l;
                idx[SM] = m;
                sink[0] = i;
                sink[1] = j;
                sink[2] = n;
                /* the two propagators that stay, in their (i, j, n) order */
                const int o1 = (X == 0) ? 1 : 0, o2 = (X == 2) ? 1 : 2;
                cplx cof[9];
                epseps_open(cof, blk(site[o1], sink[o1], idx[slot[o1]]), blk(site[o2], sink[o2], idx[slot[o2]]));
                cplx *dst = out + (sink[X] * 4 + idx[slot[X]]) * 9;
                for (int ad = 0; ad < 9; ++ad) dst[ad] = cadd(dst[ad], cmul(w, cof[ad]));
              }
          }
      }
    memcpy(P[X] + x * SITE, out, sizeof(out));
  }
}

/* ------------------------------------------------------------------ */
/* Timeslice and momentum projection (new API; no reference counterpart) */
/* ------------------------------------------------------------------ */

/* Site index -> (x, y, z, t).  order 0: lexicographic [t][z][y][x].
 * order 1: even-odd [parity][t][z][y][x/2] with parity = (x+y+z+t) % 2, the
 * layout of PyQUDA propagators (baryon_2pt.py:31 "wtzyx"). */
void oracle_site_coords(size_t s, const int dims[4], int order, int c[4])
{
  const size_t Lx = dims[0], Ly = dims[1], Lz = dims[2];
  if (order == 0) {
    c[0] = (int)(s % Lx);
    s /= Lx;
    c[1] = (int)(s % Ly);
    s /= Ly;
    c[2] = (int)(s % Lz);
    c[3] = (int)(s / Lz);
  } else {
    const size_t half = (size_t)dims[0] * dims[1] * dims[2] * dims[3] / 2;
    const int parity = (int)(s / half);
    size_t r = s % half;
    const size_t Lxh = Lx / 2;
    const int xh = (int)(r % Lxh);
    r /= Lxh;
    c[1] = (int)(r % Ly);
    r /= Ly;
    c[2] = (int)(r % Lz);
    c[3] = (int)(r / Lz);
    c[0] = 2 * xh + ((parity + c[1] + c[2] + c[3]) & 1);
  }
}

static void neumaier_add(double *sum, double *comp, double v)
{
  const double t = *sum + v;
  if (fabs(*sum) >= fabs(v))
    *comp += (*sum - t) + v;
  else
    *comp += (v - t) + *sum;
  *sum = t;
}

/* out[q][p][t] = sum_{x in timeslice t} exp(-i 2pi sum_mu p_mu X_mu / G_mu) c_q(x)
 * for q < nq site correlators c given as site_corr[V][nq], momenta mom[np][3]
 * (integer components, any sign), X = x + origin the global coordinate and G
 * the global extent (global = NULL or 0: G = L; origin = NULL: 0).  The phase
 * of each direction is taken from a table w_mu[k] = exp(-i 2pi k / G_mu) with
 * k = p_mu X_mu mod G_mu, and the three factors are multiplied x, then y, then
 * z. */
void oracle_project(double *out, const double *site_corr, int nq, const int dims[4], int order, const int *mom, int np,
                    const int *global, const int *origin)
{
  const int Lt = dims[3];
  const size_t V = (size_t)dims[0] * dims[1] * dims[2] * dims[3];
  int G[3], O[3];
  for (int mu = 0; mu < 3; ++mu) {
    G[mu] = (global && global[mu] > 0) ? global[mu] : dims[mu];
    O[mu] = origin ? origin[mu] : 0;
  }
  const cplx *c = (const cplx *)site_corr;
  double *sum = (double *)calloc((size_t)nq * np * Lt * 4, sizeof(double));
  cplx *tab[3];
  for (int mu = 0; mu < 3; ++mu) {
    tab[mu] = (cplx *)malloc(sizeof(cplx) * G[mu]);
    for (int k = 0; k < G[mu]; ++k) {
      const double th = 2.0 * M_PI * (double)k / (double)G[mu];
      tab[mu][k] = cx(cos(th), -sin(th));
    }
  }
  for (size_t s = 0; s < V; ++s) {
    int co[4];
    oracle_site_coords(s, dims, order, co);
    for (int p = 0; p < np; ++p) {
      cplx ph = cx(1, 0);
      for (int mu = 0; mu < 3; ++mu) {
        long k = ((long)mom[p * 3 + mu] * (co[mu] + O[mu])) % G[mu];
        if (k < 0) k += G[mu];
        ph = (mu == 0) ? tab[0][k] : cmul(ph, tab[mu][k]);
      }
      for (int q = 0; q < nq; ++q) {
        const cplx v = cmul(ph, c[s * nq + q]);
        double *acc = sum + (((size_t)q * np + p) * Lt + co[3]) * 4;
        neumaier_add(&acc[0], &acc[1], v.re);
        neumaier_add(&acc[2], &acc[3], v.im);
      }
    }
  }
  for (size_t e = 0; e < (size_t)nq * np * Lt; ++e) {
    out[2 * e] = sum[4 * e] + sum[4 * e + 1];
    out[2 * e + 1] = sum[4 * e + 2] + sum[4 * e + 3];
  }
  for (int mu = 0; mu < 3; ++mu) free(tab[mu]);
  free(sum);
}

/* Fused baryon two-point workload (new API): for each projector W_w (dense
 * 4x4, row-major [m][n], W = G(g) for an insertion gamma), the site correlator
 * is sum_t weight[t] * baryon_general_two_point(type_t, W_w), then projected as
 * in oracle_project.  out is [nproj][np][Lt]. */
void oracle_baryon_twopoint_project(double *out, const double *propag_i, const double *propag_j,
                                    const double *propag_n, const int *types, const double *weights, int nterm,
                                    int gamma_ij, int gamma_kl, const double *projectors, int nproj, const int dims[4],
                                    int order, const int *mom, int np, const int *global, const int *origin)
{
  const size_t V = (size_t)dims[0] * dims[1] * dims[2] * dims[3];
  double *site = (double *)calloc(V * nproj * 2, sizeof(double));
  double *tmp = (double *)malloc(V * 2 * sizeof(double));
  for (int w = 0; w < nproj; ++w)
    for (int t = 0; t < nterm; ++t) {
      oracle_baryon_general_two_point(tmp, propag_i, propag_j, propag_n, types[t], V, gamma_ij, gamma_kl,
                                      projectors + (size_t)w * 32);
      const double wr = weights[2 * t], wi = weights[2 * t + 1];
      for (size_t x = 0; x < V; ++x) {
        const double a = tmp[2 * x], b = tmp[2 * x + 1];
        site[(x * nproj + w) * 2] += wr * a - wi * b;
        site[(x * nproj + w) * 2 + 1] += wr * b + wi * a;
      }
    }
  oracle_project(out, site, nproj, dims, order, mom, np, global, origin);
  free(tmp);
  free(site);
}

/* ------------------------------------------------------------------ */
/* Three-point step after the sequential source (new API).  The        */
/* reference builds the sequential source (include/baryon_sequential.cuh:165-267, */
/* lib/contract.cpp:46-63); the caller solves on it for the sequential */
/* propagator Sigma and contracts Sigma with the forward propagator F  */
/* of the quark line the current Gamma is inserted on.                 */
/* ------------------------------------------------------------------ */

/* c_w(x) = Tr_s Tr_c[ g5 Gamma_w F(x) Sigma(x)^dagger ]
 *        = sum_{a,b,s,c,c'} (g5 Gamma_w)[a][b] F[b][s](c,c') conj(Sigma[a][s](c,c'))
 * for each dense insertion Gamma_w ([4][4] row-major), projected as in
 * oracle_project.  out is [nins][np][Lt]. */
void oracle_threepoint_project(double *out, const double *fwd, const double *seq, const double *insertions, int nins,
                               const int dims[4], int order, const int *mom, int np, const int *global,
                               const int *origin)
{
  const size_t V = (size_t)dims[0] * dims[1] * dims[2] * dims[3];
  cplx *site = (cplx *)calloc(V * (size_t)nins, sizeof(cplx));
  const cplx *F = (const cplx *)fwd, *S = (const cplx *)seq, *Gm = (const cplx *)insertions;
#pragma omp parallel for schedule(static)
  for (long x = 0; x < (long)V; ++x) {
    const cplx *f = F + x * SITE, *sg = S + x * SITE;
    for (int w = 0; w < nins; ++w) {
      cplx acc = cx(0, 0);
      for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) {
          const cplx g = Gm[(size_t)w * 16 + a * 4 + b];
          const cplx A = cx(gamma5_diag(a) * g.re, gamma5_diag(a) * g.im);
          if (!cnz(A)) continue;
          cplx t = cx(0, 0);
          for (int s = 0; s < 4; ++s) {
            const cplx *fb = blk(f, b, s), *sa = blk(sg, a, s);
            for (int cc = 0; cc < 9; ++cc) t = cadd(t, cmul(fb[cc], cconj(sa[cc])));
          }
          acc = cadd(acc, cmul(A, t));
        }
      site[(size_t)x * nins + w] = acc;
    }
  }
  oracle_project(out, (const double *)site, nins, dims, order, mom, np, global, origin);
  free(site);
}

int oracle_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Threads of the following calls (n <= 0: the OpenMP default). */
void oracle_set_num_threads(int n)
{
#ifdef _OPENMP
  static int dflt = 0;
  if (!dflt) dflt = omp_get_max_threads();
  omp_set_num_threads(n > 0 ? n : dflt);
#else
  (void)n;
#endif
}
