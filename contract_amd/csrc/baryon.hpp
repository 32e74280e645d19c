// baryon.hpp -- kernel argument blocks and launchers of every kernel family.
// Arguments travel by value in the kernel-argument segment; nothing is copied
// to __constant__ memory, so launches are asynchronous and thread-safe.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include "device.hpp"

namespace contract_amd
{

constexpr int MAX_TERMS = 6;
constexpr int MAX_MOM = 64;
constexpr int MAX_PROJ_PER_LAUNCH = 8;
constexpr int MAX_PHASE_TAB = 3 * 128; // Lx + Ly + Lz of a projected lattice

int device_cu_count();
// Grid of a persistent staged-tile kernel: min(ntiles, CUs x waves_per_cu).
size_t persistent_grid(size_t ntiles, int waves_per_cu);

// ---- baryon, drop-in per-site form -------------------------------------
struct BaryonSiteArgs {
  const void *propag_i, *propag_j, *propag_n;
  int gamma_ij; // selects a compile-time Gij form where one is instantiated
  void *correl;
  size_t volume;
  SparseSpin gij;  // Gij: row i -> (j, Gij[i,j])
  SparseSpin gkl;  // Gkl: row k -> (l, Gkl[k,l])
  SparseSpin gklT; // Gkl^T: row l -> (k, Gkl[k,l])
  double w_re[16], w_im[16]; // projector W[m][n] at q = m*4 + n
};
void launch_baryon_site(int type, const BaryonSiteArgs &a, hipStream_t stream);

// ---- timeslice / momentum projection shared by the fused kernels --------
constexpr int MAX_BATCH = 16; // requests (propagator triples) per fused launch

// One request of a fused launch: its propagators, lattice geometry and work
// decomposition.  Work item = a run of `chunk` sites inside one slice (a
// timeslice, or one parity of it), so every item feeds exactly one timeslice.
struct ProjectReq {
  const void *propag_i, *propag_j, *propag_n;
  const signed char *exp_i, *exp_j, *exp_n; // site exponents (FP32 site-scaled storage only)
  double2 *partial;   // [items][np slots][2][16]: C and S of each momentum slot
  size_t slice_sites; // sites per (parity, t) slice
  int dims[4];        // Lx, Ly, Lz, Lt
  int gdims[3];       // global spatial extents (phases are exp(-2 pi i p (x + origin) / gdims))
  int origin[3];      // global coordinates of local site 0
  int order;          // 0 lexicographic [t][z][y][x]; 1 even-odd [parity][t][z][y][x/2]
  int chunk;          // sites per work item (multiple of 4)
  int chunks_per_slice;
  unsigned item0;     // first work item of this request in the launch
};

// A launch runs one wave per resident slot; wave w takes work items w,
// w + gridDim.x, ... over the concatenated items of all requests, prefetching
// the next item's first tile while it finishes the current one.  Which wave
// computes an item does not change the item's partial.
struct ProjectArgs {
  int np; // momentum slots (one per +-k pair)
  int mom[MAX_MOM][3];
  int nreq;
  unsigned total_items; // over all requests
  int share_n;          // every request has propag_n == propag_i (and exp_n == exp_i)
  ProjectReq req[MAX_BATCH];
};

struct BaryonProjectArgs {
  int gamma_ij; // selects a compile-time Gij form where one is instantiated
  SparseSpin gij, gkl, gklT;
  int nterm;
  int types[MAX_TERMS];
  double w_re[MAX_TERMS], w_im[MAX_TERMS];
};
void launch_baryon_project(const BaryonProjectArgs &a, const ProjectArgs &pa, hipStream_t stream);
// Same with complex64 propagators: FP32 per-site contraction on blocks rescaled
// by exact powers of two, FP64 projection and sums.
void launch_baryon_project_fp32(const BaryonProjectArgs &a, const ProjectArgs &pa, hipStream_t stream);
// Same on FP32 site-scaled propagators: no per-site rescaling in the kernel,
// 2^(e_i + e_j + e_n) applied to each site's result in FP64.
void launch_baryon_project_fp32_scaled(const BaryonProjectArgs &a, const ProjectArgs &pa, hipStream_t stream);

// ---- FP32 site-scaled storage -------------------------------------------
constexpr int SCALED_EXP_MAX = 127;      // site exponents are int8 in [-127, 127]
constexpr size_t SCALED_SITE_BYTES = 1152; // complex64 [4][4][3][3]; the exponents follow all V sites
struct ScaleArgs {
  const void *src; // complex128 [V][4][4][3][3]
  void *dst;       // complex64 [V][4][4][3][3]
  void *exps;      // int8 [V]
  size_t volume;
};
void launch_scale_fp32(const ScaleArgs &a, hipStream_t stream);

// partial[items][np][nq] -> open[t][np][nq]: per-timeslice sum over the
// timeslice's work items (both parities), fixed order, Neumaier compensated.
struct ReduceReq {
  const double2 *partial;
  double2 *open;
  int lt, order, chunks_per_slice;
  int t0; // first timeslice of this request in the launch's grid
};
struct ReduceArgs {
  int np, nq, nreq;
  ReduceReq req[MAX_BATCH];
};
void launch_project_reduce(const ReduceArgs &r, int total_lt, hipStream_t stream);

// open[t][slot][2][16] holds, per momentum slot k, C = sum cos(k.x) R and
// S = -sum sin(k.x) R; requested momentum p is slot[p] with sign sgn[p]:
// out[w][p][t] = sum_q W_w[q] (C[q] + sgn i S[q]), for up to
// MAX_PROJ_PER_LAUNCH projectors and MAX_BATCH requests per launch.
struct ApplyReq {
  const double2 *open;
  double2 *out;  // [nproj_total][nmom][lt]
  int lt;
  unsigned e0;   // first output element of this request in the launch
};
struct ApplyArgs {
  int nslot;
  int nmom;      // requested momenta (outputs)
  int nproj;     // projectors in this launch
  int proj_base; // index of the first one in the output
  int slot[MAX_MOM], sgn[MAX_MOM];
  double w_re[MAX_PROJ_PER_LAUNCH][16], w_im[MAX_PROJ_PER_LAUNCH][16];
  int nreq;
  unsigned total;
  ApplyReq req[MAX_BATCH];
};
void launch_project_apply(const ApplyArgs &f, hipStream_t stream);

// ---- meson --------------------------------------------------------------
struct MesonArgs {
  const void *propag_i, *propag_j;
  size_t volume;
  SparseSpin a; // (g5 Gij)^T : row i -> (j, (g5 Gij)[j,i])
  SparseSpin b; // (g5 Gkl)^T : row l -> (k, (g5 Gkl)[k,l])
  void *correl;
};
void launch_meson(const MesonArgs &a, hipStream_t stream);

// 16 correlators of one fixed and one running insertion (meson_all_*).
struct MesonAllArgs {
  const void *propag_i, *propag_j;
  size_t volume;
  int sink;              // 0: source form (Gij fixed), 1: sink form (Gkl fixed)
  SparseSpin fixed;      // the fixed (g5 G)^T
  SparseSpin run[16];    // (g5 G(g))^T for g = 0..15
  void *correl[16];
};
void launch_meson_all(const MesonAllArgs &a, hipStream_t stream);

// ---- diquark ------------------------------------------------------------
struct DiquarkArgs {
  const void *propag_i, *propag_j;
  void *diquark;
  size_t volume;
  SparseSpin gij, gklT;
  // FP32 site-scaled storage only: the site exponents of P_i, P_j and the output
  const signed char *exp_i, *exp_j;
  signed char *exp_out;
};
void launch_diquark(const DiquarkArgs &a, hipStream_t stream);
// Same on FP32 site-scaled propagators, the output in that form too.
void launch_diquark_fp32_scaled(const DiquarkArgs &a, hipStream_t stream);

// ---- sequential ---------------------------------------------------------
struct SequentialArgs {
  const void *in0, *in1; // the two propagators that stay, in (i, j, n) order
  void *out;             // buffer of the removed propagator (never read)
  size_t volume;
  // FP32 site-scaled storage only: the site exponents of in0, in1 and out
  const signed char *exp0, *exp1;
  signed char *exp_out;
  // each spin matrix in both orientations: G (row r -> column, G[r][c]) and
  // G^T (row c -> r, G[r][c])
  SparseSpin gij, gijT, gkl, gklT, gmn, gmnT;
};
void launch_sequential(int type, int target, const SequentialArgs &a, hipStream_t stream);
// Same on FP32 site-scaled propagators (complex64 mantissas + int8 site
// exponents): FP32 arithmetic on the O(1) mantissas, the output written in
// the same form with each site renormalised.
void launch_sequential_fp32_scaled(int type, int target, const SequentialArgs &a, hipStream_t stream);

} // namespace contract_amd
