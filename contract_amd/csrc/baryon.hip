// baryon.hip -- baryon two-point contractions on gfx950.
//
// Every contraction type contracts three propagators with two colour epsilons
// and three spin matrices:
//
//   C(x) = sum Gij[i,j] Gkl[k,l] Gmn[m,n] ee(P_i[i,X], P_j[j,Y], P_n[n,Z])
//
// where type I<X>_J<Y>_N<Z> names the source-spin index each propagator
// carries (include/contract_enum.h:1-7; the SWAP_IJ / SWAP_KL forms of
// include/baryon.cuh:41-52,95-105 reduce to this table).  All six types
// factor through one intermediate, the spin-open diquark of P_i and P_j
//
//   D[x][y](a,d) = sum_i Gij[i,j(i)] cof(P_i[i,x], P_j[j(i),y])[a,d]
//
// (cof = the colour cofactor with one epsilon slot open), after which each
// type is a cheap contraction of D with P_n and Gkl into the open-spin
// correlator R[n][m] = dC/dGmn[m,n]; any insertion or projector W then gives
// C = sum_mn W[m,n] R[n,m].  D costs 64 cofactors per site and is shared by
// every contraction term and every insertion; the reference's lane loop
// (include/baryon.cuh:55-82) redoes its 36-144 cofactors per site for each
// type and each insertion.
//
// Lane q of a 16-lane site group holds D[x][y] for q = 4x + y, and R[n][m]
// for q = 4m + n.

#include <atomic>

#include "baryon.hpp"
#include "device.hpp"
#include "project.hpp"
#include "runtime.hpp"
#include "spin.hpp"

namespace contract_amd
{

enum : int { LK = 0, LL = 1, LM = 2 };
__host__ __device__ constexpr int type_slot(int type, int prop)
{
  // (P_i, P_j, P_n) source letters for the six types, in enum order
  constexpr int tab[6][3] = {{LK, LL, LM}, {LK, LM, LL}, {LL, LK, LM}, {LL, LM, LK}, {LM, LK, LL}, {LM, LL, LK}};
  return tab[type][prop];
}

// D[x][y] for lane q = 4x + y.  The i loop is uniform across the wave, so the
// gamma row lookups are scalar.
template <typename C>
__device__ __forceinline__ void diquark_tensor(C (&D)[9], const C *__restrict__ si, const C *__restrict__ sj, int q,
                                               const SparseSpin &gij)
{
  const int x = q >> 2, y = q & 3;
#pragma unroll
  for (int e = 0; e < 9; ++e) D[e] = czero<C>();
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const C w = cmake<C>((decltype(C::x))gij.re[i], (decltype(C::x))gij.im[i]);
    C b[9];
    load_block(b, sj, gij.col[i], y);
#pragma unroll
    for (int e = 0; e < 9; ++e) b[e] = cmul(w, b[e]);
    cof_acc_stream(D, si + (i * 4 + x) * 9, b);
  }
}

// Rows of a gamma matrix known at compile time: row i has its one nonzero in
// column col[i], equal to i^rot[i].  G(5) = Cg5 in the DeGrand-Rossi basis
// (gamma_1 gamma_3), the diquark projector of the nucleon, is instantiated;
// any other Gij takes the runtime path (the reference instantiates its kernels
// per gamma index in the same way, lib/baryon/baryon.cu:22-67).
template <int G> struct GammaRows;
template <> struct GammaRows<5> {
  static constexpr int col[4] = {1, 0, 3, 2};
  static constexpr int rot[4] = {2, 0, 2, 0}; // -1, +1, -1, +1
};

// diquark_tensor for a compile-time Gij: the weights become operand signs.
template <int GIJ, int I, typename C>
__device__ __forceinline__ void diquark_row(C (&D)[9], const C *__restrict__ si, const C *__restrict__ sj, int x, int y)
{
  C b[9];
  load_block(b, sj, GammaRows<GIJ>::col[I], y);
  cof_acc_stream_rot<GammaRows<GIJ>::rot[I]>(D, si + (I * 4 + x) * 9, b);
}

template <int GIJ, typename C>
__device__ __forceinline__ void diquark_tensor_static(C (&D)[9], const C *__restrict__ si, const C *__restrict__ sj,
                                                      int q)
{
  const int x = q >> 2, y = q & 3;
#pragma unroll
  for (int e = 0; e < 9; ++e) D[e] = czero<C>();
  diquark_row<GIJ, 0>(D, si, sj, x, y);
  diquark_row<GIJ, 1>(D, si, sj, x, y);
  diquark_row<GIJ, 2>(D, si, sj, x, y);
  diquark_row<GIJ, 3>(D, si, sj, x, y);
}

// Per-lane constants of the Gkl tails, computed once per kernel.  For lane
// (x, y): kx = col(x), ky = col(y) of Gkl (its pattern is symmetric), and the
// values Gkl[x][kx] (u), Gkl[kx][x] (v), likewise for y.
template <typename C> struct KlLane {
  int kx, ky;
  C ux, vx, uy, vy;
  bool on0, on2; // lanes whose D enters the Z = m tails (IK_JL_NM, IL_JK_NM)
  __device__ __forceinline__ KlLane(const SparseSpin &gkl, const SparseSpin &gklT, int q)
  {
    const int x = q >> 2, y = q & 3;
    kx = sparse_col(gkl, x);
    ky = sparse_col(gkl, y);
    ux = sparse_val<C>(gkl, x);
    vx = sparse_val<C>(gklT, x);
    uy = sparse_val<C>(gkl, y);
    vy = sparse_val<C>(gklT, y);
    on0 = x == ky; // D[k(l)][l] with l = y
    on2 = y == kx; // D[l][k(l)] with l = x
  }
  // weight of this lane's D in the Z = m tail of type T (0 off-pattern)
  template <int T> __device__ __forceinline__ C wz() const
  {
    return (T == 0) ? (on0 ? vy : czero<C>()) : (on2 ? vx : czero<C>());
  }
};

// Open-spin correlator of type T from D, returned in the R layout q = 4m + n.
// `scr` is 45 complex of LDS owned by this 16-lane group (the group's
// consumed P_i / P_j staging area), used by the Z = m types to sum D over
// four lanes without a 9-entry lane all-reduce.
// LDS_SUM = false sums with a 16-lane all-reduce instead (fewer registers
// in the single-type drop-in kernels, where the LDS round trips cost more).
template <int T, typename C, bool LDS_SUM = true>
__device__ __forceinline__ C baryon_tail(const C (&D)[9], const C *__restrict__ sn, C *__restrict__ scr, int lane,
                                         const KlLane<C> &kl)
{
  const int q = lane & 15;
  constexpr int Z = type_slot(T, 2);
  C a[9];
  if constexpr (Z == LM && !LDS_SUM) {
    const C w = kl.template wz<T>();
    C x1[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) x1[e] = group16_allreduce(cmul(w, D[e]));
    load_block(a, sn, q & 3, q >> 2);
    return dot9(a, x1);
  } else if constexpr (Z == LM) {
    // R[n][m] = sum_ad P_n[n][m] X1, X1 = sum_l Gkl[k,l] D[k][l] (T = 0) or D[l][k] (T = 2):
    // the four lanes holding a D[k(l)][l] (or D[l][k(l)]) post it weighted to
    // LDS slot l, nine lanes sum the slots entry-wise, every lane reads X1.
    const C w = kl.template wz<T>();
    const int slot = (T == 0) ? (q & 3) : (q >> 2);
    __syncthreads();
    if (w.x != 0 || w.y != 0)
#pragma unroll
      for (int e = 0; e < 9; ++e) scr[slot * 9 + e] = cmul(w, D[e]);
    __syncthreads();
    if (q < 9) scr[36 + q] = cadd(cadd(scr[q], scr[9 + q]), cadd(scr[18 + q], scr[27 + q]));
    __syncthreads();
    C x1[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) x1[e] = scr[36 + e];
    load_block(a, sn, q & 3, q >> 2);
    return dot9(a, x1);
  } else {
    // lane (x, y) owns one (l, m) term of  R[n][m] = sum_l w_l sum_ad P_n[n][c] D[.][.]
    //   T=1 IK_JM_NL: D[k][m], k = x, l = kx, c = l,  w = u_x   -> reduce over x
    //   T=3 IL_JM_NK: D[l][m], l = x, c = kx,         w = v_x   -> reduce over x
    //   T=4 IM_JK_NL: D[m][k], k = y, l = ky, c = l,  w = u_y   -> reduce over y
    //   T=5 IM_JL_NK: D[m][l], l = y, c = ky,         w = v_y   -> reduce over y
    constexpr bool over_x = (T == 1 || T == 3);
    const int c = over_x ? kl.kx : kl.ky;
    const C w = (T == 1) ? kl.ux : (T == 3) ? kl.vx : (T == 4) ? kl.uy : kl.vy;
    C r[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      load_block(a, sn, n, c);
      r[n] = cmul(w, dot9(a, D));
    }
    if constexpr (over_x) {
      // lane (x, y) ends with R[n = x][m = y]; move it to lane 4m + n
      const C t = reduce_scatter4_hi(r, q);
      const int src = (lane & ~15) | ((q & 3) << 2) | (q >> 2);
      return cmake<C>(__shfl(t.x, src, 64), __shfl(t.y, src, 64));
    } else {
      return reduce_scatter4_lo(r, q); // lane (x, y) = (m, n)
    }
  }
}

// IK_JL_NM tail whose LDS scratch is P_n's own staging area (for kernels that
// have no other area free): every lane reads its P_n block into registers
// first, then the group's site area holds the X1 sum, formed in the same
// order as baryon_tail<0> so the two give bit-identical results.  Any other
// tail that reads P_n runs before this one.
template <typename C>
__device__ __forceinline__ C baryon_tail0_in_place(const C (&D)[9], C *__restrict__ sn, int lane, const KlLane<C> &kl)
{
  const int q = lane & 15;
  C a[9];
  load_block(a, sn, q & 3, q >> 2);
  const C w = kl.template wz<0>();
  const int slot = q & 3;
  wave_sync(); // the group's P_n reads are done
  if (w.x != 0 || w.y != 0)
#pragma unroll
    for (int e = 0; e < 9; ++e) sn[slot * 9 + e] = cmul(w, D[e]);
  wave_sync();
  if (q < 9) sn[36 + q] = cadd(cadd(sn[q], sn[9 + q]), cadd(sn[18 + q], sn[27 + q]));
  wave_sync();
  C x1[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) x1[e] = sn[36 + e];
  return dot9(a, x1);
}

template <typename C>
__device__ __forceinline__ C baryon_tail_rt(int type, const C (&D)[9], const C *sn, C *scr, int lane,
                                            const KlLane<C> &kl)
{
  switch (type) {
  case 0: return baryon_tail<0, C>(D, sn, scr, lane, kl);
  case 1: return baryon_tail<1, C>(D, sn, scr, lane, kl);
  case 2: return baryon_tail<2, C>(D, sn, scr, lane, kl);
  case 3: return baryon_tail<3, C>(D, sn, scr, lane, kl);
  case 4: return baryon_tail<4, C>(D, sn, scr, lane, kl);
  default: return baryon_tail<5, C>(D, sn, scr, lane, kl);
  }
}

// ---------------------------------------------------------------------------
// Drop-in per-site kernel: correl[x] = sum_q W[q] R_q(x)   (q = m*4 + n)
// baryon_two_point passes W = G(gamma_mn), baryon_general_two_point passes
// W = project_mn (include/baryon_general.cuh:62-67).
// ---------------------------------------------------------------------------
// Staged like the fused nucleon kernel, with its load pipeline: two
// propagators in LDS at a time (18.5 KB per wave instead of 27.8 KB), so
// eight waves fit per CU, two per SIMD, and every next-tile load in flight
// across at least half a step.  Each wave walks one contiguous run of tiles.
// P_j is register-prefetched a tile ahead and staged at the top of the step;
// P_i moves HBM -> LDS directly (dma_tile) into an unpadded area of its own,
// issued once the diquark has consumed the current tile; P_n is stored into
// P_j's consumed area after the diquark and its next tile is loaded into the
// freed registers at once.  Steps order LDS with wave_sync: a
// __syncthreads() would also wait for P_i's load.  Past the run the loads
// take an empty range (no memory is touched; the registers and P_i's area
// receive zeros), so no load sits on one side of a branch.
template <int T, int GIJ>
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(2))) void baryon_site_kernel(
  BaryonSiteArgs a)
{
  using C = double2;
  constexpr int SL = site_lds<C>();
  __shared__ C lds[SITES_PER_WAVE * SL];  // P_j, then P_n (register-staged, padded sites)
  __shared__ C lds_i[Tile<C>::DMA_ELEMS]; // P_i (dma_tile's unpadded layout)
  const int lane = threadIdx.x;
  const int g = lane >> 4, q = lane & 15;
  const C w = cmake<C>(a.w_re[q], a.w_im[q]);
  const KlLane<C> kl(a.gkl, a.gklT, q);
  const C *Pi = (const C *)a.propag_i, *Pj = (const C *)a.propag_j, *Pn = (const C *)a.propag_n;
  const size_t ntiles = (a.volume + SITES_PER_WAVE - 1) / SITES_PER_WAVE;
  // this wave's run of tiles [t0, t1): the launch has at most ntiles waves,
  // so every run holds at least one tile
  const size_t t0 = ntiles * blockIdx.x / gridDim.x, t1 = ntiles * (blockIdx.x + 1) / gridDim.x;
  if (t0 >= t1) return;
  auto nsites_of = [&](size_t t) {
    const size_t left = a.volume - t * SITES_PER_WAVE;
    return left < (size_t)SITES_PER_WAVE ? (int)left : SITES_PER_WAVE;
  };
  const C *si = lds_i + g * SITE_ELEMS;
  C *sj = lds + g * SL;
  Tile<C> bj, bn;
  bj.load(Pj, t0 * SITES_PER_WAVE, nsites_of(t0), lane);
  dma_tile(lds_i, Pi, t0 * SITES_PER_WAVE, nsites_of(t0), lane);
  bn.load(Pn, t0 * SITES_PER_WAVE, nsites_of(t0), lane);
  for (size_t tile = t0; tile < t1; ++tile) {
    const size_t first = tile * SITES_PER_WAVE, next = first + SITES_PER_WAVE;
    const int nsites = nsites_of(tile);
    const int nnext = tile + 1 < t1 ? nsites_of(tile + 1) : 0; // an empty range past the run
    bj.store(lds, lane);
    wave_sync();
    bj.load(Pj, next, nnext, lane);
    // vmcnt(9): everything issued before P_j's loads just issued, P_i's tile
    // included, has completed
    dma_wait_keep<Tile<C>::PER_LANE>();
    C D[9];
    if constexpr (GIJ >= 0)
      diquark_tensor_static<GIJ>(D, si, sj, q);
    else
      diquark_tensor(D, si, sj, q, a.gij);
    wave_sync(); // every lane's P_i and P_j reads are done
    dma_tile(lds_i, Pi, next, nnext, lane);
    bn.store(lds, lane);
    bn.load(Pn, next, nnext, lane);
    wave_sync();
    const C R = baryon_tail<T, C, false>(D, sj, sj, lane, kl); // no LDS scratch with the all-reduce sum
    const C c = group16_allreduce(cmul(w, R));
    if (q == 0 && g < nsites) static_cast<C *>(a.correl)[first + g] = c;
    wave_sync(); // the tail's P_n reads are done before the next P_j store
  }
  dma_wait(); // the last (empty) P_i load has landed before the wave's LDS is released
}

template <int GIJ>
static void launch_site_g(int type, const BaryonSiteArgs &a, dim3 grid, hipStream_t stream)
{
  const dim3 block(WAVE);
  switch (type) {
  case 0: hipLaunchKernelGGL((baryon_site_kernel<0, GIJ>), grid, block, 0, stream, a); break;
  case 1: hipLaunchKernelGGL((baryon_site_kernel<1, GIJ>), grid, block, 0, stream, a); break;
  case 2: hipLaunchKernelGGL((baryon_site_kernel<2, GIJ>), grid, block, 0, stream, a); break;
  case 3: hipLaunchKernelGGL((baryon_site_kernel<3, GIJ>), grid, block, 0, stream, a); break;
  case 4: hipLaunchKernelGGL((baryon_site_kernel<4, GIJ>), grid, block, 0, stream, a); break;
  default: hipLaunchKernelGGL((baryon_site_kernel<5, GIJ>), grid, block, 0, stream, a); break;
  }
}

void launch_baryon_site(int type, const BaryonSiteArgs &a, hipStream_t stream)
{
  const size_t ntiles = (a.volume + SITES_PER_WAVE - 1) / SITES_PER_WAVE;
  if (ntiles == 0) return;
  const dim3 grid((unsigned)persistent_grid(ntiles, 8));
  if (a.gamma_ij == 5)
    launch_site_g<5>(type, a, grid, stream);
  else
    launch_site_g<-1>(type, a, grid, stream);
}

// The one-wave FP64 forms with up to two momentum tiles (the runtime term
// dispatch; the nucleon at two tiles) run persistent waves that prefetch
// across work items.  The staged FP64 nucleon (one tile, two waves per SIMD,
// below), FP32 (two waves per SIMD) and the larger FP64 tiles have no
// register room for the cross-item state: they run one wave per item.
template <typename Real, int NT> constexpr bool project_persistent() { return sizeof(Real) == 8 && NT <= 2; }

// Term sets known at compile time.  The runtime form dispatches each term to
// one of the six tails; with all six inlined into the step the register
// allocator must hold the union of their live ranges.  The nucleon's direct +
// exchange pair (IK_JL_NM, IM_JL_NK, the proton/neutron two-point of
// baryon_2pt.py) is instantiated on its own: for the FP32 forms 183 instead
// of 241 VGPRs, 3-4% faster at 32^3x64; for FP64 it is what lets the staged
// form fit two waves per SIMD.
enum { TERMS_RUNTIME = 0, TERMS_NUCLEON = 1 };

// SHARE: every request of the launch has propag_n == propag_i (a nucleon's
// two same-flavour quarks from one propagator, (u, d, u)).  That propagator
// is read from HBM and staged once per site and serves as both P_i and P_n:
// 4608 instead of 6912 bytes per site in FP64.  The tails' scratch is then
// P_j's consumed staging area, since P_i's is still being read as P_n.
//
// The staged form (the FP64 nucleon at one momentum tile) holds two
// propagators in LDS at a time and skips the register prefetch of the whole
// next tile, which buys a second wave per SIMD (18.5 KB of LDS per wave,
// eight waves per CU, half the register file).  With three distinct
// propagators every load of the next tile is in flight across at least half
// a step: P_j is register-prefetched before the diquark; P_i goes HBM -> LDS
// directly (dma_tile) into an area of its own once the diquark has consumed
// it; P_n is stored into P_j's consumed area mid-step and its next tile is
// loaded into the freed registers right after.  Steps order LDS with
// wave_sync, since a __syncthreads() would wait for the P_i load.  With
// P_n = P_i (shared) the one propagator stays in LDS through the tails and
// is loaded at the top of the step.  The one-wave FP64 forms, bound by their
// own step latency, hide the loads behind the step instead.  (The FP32 forms
// already run two waves per SIMD with the full prefetch; staged at three
// waves they measured 8-17% slower.)
template <typename Real, int NT, int TERMS> constexpr bool project_staged()
{
  return sizeof(Real) == 8 && TERMS == TERMS_NUCLEON && NT == 1;
}

// Waves per SIMD the register allocation targets: two for the staged FP64
// nucleon and the one-tile FP32 forms, one for the other FP64 forms and the
// FP32 forms with more momentum tiles.
template <typename Real, int NT, int TERMS> constexpr int project_waves_per_eu()
{
  return project_staged<Real, NT, TERMS>() ? 2 : (sizeof(Real) == 4 && NT == 1) ? 2 : 1;
}

template <int NT, typename Real, int GIJ = -1, bool SCALED = false, int TERMS = TERMS_RUNTIME, bool SHARE = false>
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(project_waves_per_eu<Real, NT, TERMS>()))) void
baryon_project_kernel(
  BaryonProjectArgs a, ProjectArgs args)
{
  using C = cplx<Real>;
  constexpr int SL = site_lds<C>();
  constexpr bool STAGED = project_staged<Real, NT, TERMS>();
  constexpr bool PREFETCH = !STAGED;
  constexpr bool PERSIST = PREFETCH && project_persistent<Real, NT>();
  constexpr bool PF_J = STAGED; // staged form: P_j a step ahead, P_i (and P_n) at the top of the step
  constexpr int NP = SHARE ? 2 : 3;   // distinct propagators read per site
  constexpr bool MID_N = STAGED && !SHARE; // P_n staged after the diquark
  // Staged three-propagator form: P_i goes HBM -> LDS directly into an area
  // of its own, issued once the diquark has consumed the previous tile (half
  // a step ahead); P_j and then P_n share the register-staged area.
  constexpr bool DMA_I = MID_N;
  constexpr int NL = DMA_I ? 1 : (STAGED ? 2 : NP); // register-staged areas
  __shared__ C lds[NL][SITES_PER_WAVE * SL];
  __shared__ C lds_i[DMA_I ? Tile<C>::DMA_ELEMS : 1]; // unpadded: dma_tile's layout
  extern __shared__ double2 tab[]; // Gx + Gy + Gz phase entries, sized at launch
  const int lane = threadIdx.x;
  const int g = lane >> 4, q = lane & 15;
  unsigned item = blockIdx.x; // global work item (uniform)
  if (item >= args.total_items) return;
  int r = 0; // its request
  while (r + 1 < args.nreq && item >= args.req[r + 1].item0) ++r;
  size_t first, last;
  item_range(args.req[r], item - args.req[r].item0, first, last);
  load_phase_tab(tab, args.req[r], lane);
  const KlLane<C> kl(a.gkl, a.gklT, q);

  // one-wave workgroups: with P_i's HBM -> LDS load in flight across the
  // step, a __syncthreads() would wait for it (see wave_sync)
  auto step_sync = [] {
    if constexpr (DMA_I)
      wave_sync();
    else
      __syncthreads();
  };
  auto tile_sites = [](size_t s, size_t end) {
    const size_t left = end - s;
    return left < (size_t)SITES_PER_WAVE ? (int)left : SITES_PER_WAVE;
  };
  // FP32 site-scaled storage: the site exponents e_i, e_j, e_n of this
  // lane's site in the tile at st (padded lanes repeat the last site; their
  // R is zero), kept apart and summed where the step applies them: a sum
  // formed as they load would wait for them right after the next tile's
  // loads are issued, and so (loads count down in issue order) for the
  // whole prefetch
  struct SiteExp {
    int i = 0, j = 0, n = 0;
  };
  auto site_exp = [&](const ProjectReq &rq, size_t st, int n) {
    const size_t k = st + (size_t)(g < n ? g : n - 1);
    SiteExp e;
    e.i = rq.exp_i[k];
    e.j = rq.exp_j[k];
    if constexpr (!SHARE) e.n = rq.exp_n[k];
    return e;
  };
  auto exp_sum = [](const SiteExp &e) { return SHARE ? 2 * e.i + e.j : e.i + e.j + e.n; };
  Tile<C> buf[NP];
  SiteExp ex_buf; // site exponents of the tile in buf
  if constexpr (PREFETCH) {
    const ProjectReq &pa = args.req[r];
    const int n = tile_sites(first, last);
    buf[0].load((const C *)pa.propag_i, first, n, lane);
    buf[1].load((const C *)pa.propag_j, first, n, lane);
    if constexpr (!SHARE) buf[NP - 1].load((const C *)pa.propag_n, first, n, lane);
    if constexpr (SCALED) ex_buf = site_exp(pa, first, n);
  } else if constexpr (PF_J) {
    buf[1].load((const C *)args.req[r].propag_j, first, tile_sites(first, last), lane);
    if constexpr (DMA_I) {
      dma_tile(lds_i, (const C *)args.req[r].propag_i, first, tile_sites(first, last), lane);
      buf[2].load((const C *)args.req[r].propag_n, first, tile_sites(first, last), lane);
    }
  }

  for (;;) {
    const ProjectReq &pa = args.req[r];
    // the wave's next item: its first tile is prefetched during this item's last step
    const unsigned item_next = item + gridDim.x;
    const bool more = PERSIST && item_next < args.total_items;
    int r_next = r;
    size_t first_next = 0, last_next = 0;
    if (more) {
      while (r_next + 1 < args.nreq && item_next >= args.req[r_next + 1].item0) ++r_next;
      item_range(args.req[r_next], item_next - args.req[r_next].item0, first_next, last_next);
    }
    const C *Pi = (const C *)pa.propag_i, *Pj = (const C *)pa.propag_j, *Pn = (const C *)pa.propag_n;
    const C *Qi = (const C *)args.req[r_next].propag_i, *Qj = (const C *)args.req[r_next].propag_j,
            *Qn = (const C *)args.req[r_next].propag_n;

    PhaseWalk<NT> walk;
    walk.init(args, pa, first, g, q);
    dbl4 acc[NT][4];
#pragma unroll
    for (int tt = 0; tt < NT; ++tt)
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[tt][u] = dbl4{0.0, 0.0, 0.0, 0.0};

    for (size_t s = first; s < last; s += SITES_PER_WAVE) {
      const int nhere = tile_sites(s, last);
      if constexpr (STAGED) {
        if constexpr (!DMA_I) buf[0].load(Pi, s, nhere, lane);
        if constexpr (!PF_J) buf[1].load(Pj, s, nhere, lane);
        if constexpr (!SHARE && !DMA_I) buf[2].load(Pn, s, nhere, lane);
        if constexpr (SCALED) ex_buf = site_exp(pa, s, nhere);
      }
      if constexpr (DMA_I) {
        buf[1].store(lds[0], lane);
      } else {
#pragma unroll
        for (int p = 0; p < NL; ++p) buf[p].store(lds[p], lane);
      }
      step_sync();
      int ex = 0;
      const SiteExp ex_tile = ex_buf;
      const size_t next = s + SITES_PER_WAVE;
      const bool here = next < last; // else the next item's first tile (uniform)
      if constexpr (PF_J)
        if (here) buf[1].load(Pj, next, tile_sites(next, last), lane);
      // vmcnt(9): everything issued before P_j's loads just issued, P_i's
      // tile included, has completed
      if constexpr (DMA_I) dma_wait_keep<Tile<C>::PER_LANE>();
      if (PREFETCH && (here || more)) {
        const size_t st = here ? next : first_next;
        const int n = here ? tile_sites(next, last) : tile_sites(first_next, last_next);
        buf[0].load(here ? Pi : Qi, st, n, lane);
        buf[1].load(here ? Pj : Qj, st, n, lane);
        if constexpr (!SHARE) buf[NP - 1].load(here ? Pn : Qn, st, n, lane);
        if constexpr (SCALED) ex_buf = site_exp(here ? pa : args.req[r_next], st, n);
      }

      C *si = DMA_I ? lds_i + g * SITE_ELEMS : lds[0] + g * SL;
      C *sj = lds[DMA_I ? 0 : 1] + g * SL, *sn = lds[NL == 3 ? 2 : 0] + g * SL;
      C *scr = NL == 2 ? sj : si; // consumed once D is formed (unused with DMA_I)
      if constexpr (sizeof(Real) == 4 && !SCALED) {
        // P_i and P_j only: D of the rescaled blocks is O(1), so the tails'
        // products with P_n are as large as P_n's own entries and cannot
        // underflow unless a stored entry already is subnormal.  A shared
        // P_n = P_i is read rescaled, which its exponent counts twice.
        const int ei = normalise_site(si, q);
        ex = (SHARE ? 2 * ei : ei) + normalise_site(sj, q);
        __syncthreads();
      }
      C D[9];
      if constexpr (GIJ >= 0)
        diquark_tensor_static<GIJ>(D, si, sj, q);
      else
        diquark_tensor(D, si, sj, q, a.gij);
      // DMA_I: the phase table (dynamic LDS, which the compiler cannot tell
      // apart from P_i's area) is read before P_i's next load is issued, or
      // its reads would wait for that load
      double ph_re[NT], ph_im[NT];
      if constexpr (DMA_I) project_phases<NT>(ph_re, ph_im, args.np, pa, tab, walk);
      if constexpr (MID_N) { // P_n into P_j's consumed area; the next P_i into its own
        step_sync();
        dma_tile(lds_i, Pi, next, here ? tile_sites(next, last) : 0, lane); // an empty range past the item
        buf[2].store(lds[0], lane);
        if constexpr (DMA_I) // the next P_n, a whole step ahead
          buf[2].load(Pn, next, here ? tile_sites(next, last) : 0, lane);
        step_sync();
      }
      double2 R = make_double2(0.0, 0.0);
      if constexpr (TERMS == TERMS_NUCLEON) { // types {IK_JL_NM, IM_JL_NK}, in that order
        // With DMA_I no LDS area is free for the IK_JL_NM sum: it runs last,
        // in P_n's area (the terms still enter R in type order).
        C r0, r5;
        if constexpr (DMA_I) {
          r5 = baryon_tail<5, C>(D, sn, scr, lane, kl);
          r0 = baryon_tail0_in_place<C>(D, sn, lane, kl);
        } else {
          r0 = baryon_tail<0, C>(D, sn, scr, lane, kl);
          r5 = baryon_tail<5, C>(D, sn, scr, lane, kl);
        }
        R = cfma(R, make_double2(a.w_re[0], a.w_im[0]), make_double2((double)r0.x, (double)r0.y));
        R = cfma(R, make_double2(a.w_re[1], a.w_im[1]), make_double2((double)r5.x, (double)r5.y));
      } else {
        for (int t = 0; t < a.nterm; ++t) {
          const C rt = baryon_tail_rt<C>(a.types[t], D, sn, scr, lane, kl);
          R = cfma(R, make_double2(a.w_re[t], a.w_im[t]), make_double2((double)rt.x, (double)rt.y));
        }
      }
      if constexpr (SCALED) ex = exp_sum(ex_tile);
      if constexpr (sizeof(Real) == 4) R = make_double2(ldexp(R.x, ex), ldexp(R.y, ex));
      if constexpr (DMA_I) { // a ragged tile's padded sites read stale P_i slots
        R = g < nhere ? R : make_double2(0.0, 0.0);
        project_mfma<NT>(acc, ph_re, ph_im, R);
      } else {
        project_accumulate<NT>(acc, args.np, pa, tab, R, walk); // padded lanes carry R = 0
      }
      walk.advance(pa);
      step_sync();
    }
    project_store<NT>(acc, args.np, pa, item - pa.item0, lane);
    if (!more) break;
    if (r_next != r) {
      load_phase_tab(tab, args.req[r_next], lane); // first read after the next step's barrier
    }
    item = item_next;
    r = r_next;
    first = first_next;
    last = last_next;
  }
}

template <int NT, typename Real, int GIJ, bool SCALED, int TERMS, bool SHARE = false>
static void launch_project_nt(const BaryonProjectArgs &a, const ProjectArgs &pa, size_t tab_bytes,
                              hipStream_t stream)
{
  auto kernel = baryon_project_kernel<NT, Real, GIJ, SCALED, TERMS, SHARE>;
  unsigned grid = pa.total_items;
  if (project_persistent<Real, NT>() && !project_staged<Real, NT, TERMS>()) {
    // as many waves as the chip holds at once: resident blocks per CU (queried
    // once per device, with the largest phase table) x CUs
    static std::atomic<int> cache[64];
    int dev = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    if ((per_cu = cache[dev].load(std::memory_order_relaxed)) <= 0) {
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, WAVE, sizeof(double2) * MAX_PHASE_TAB) !=
            hipSuccess ||
          per_cu < 1)
        per_cu = 1;
      cache[dev].store(per_cu, std::memory_order_relaxed);
    }
    const unsigned w = (unsigned)per_cu * (unsigned)device_cu_count();
    grid = grid < w ? grid : w;
  }
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(WAVE), tab_bytes, stream, a, pa);
}

template <typename Real, int GIJ, bool SCALED, int TERMS>
static void launch_project_g(const BaryonProjectArgs &a, const ProjectArgs &pa, hipStream_t stream)
{
  const int nt = (pa.np + 15) / 16;
  int tab = 0;
  for (int r = 0; r < pa.nreq; ++r) {
    const int n = pa.req[r].gdims[0] + pa.req[r].gdims[1] + pa.req[r].gdims[2];
    tab = tab > n ? tab : n;
  }
  const size_t tab_bytes = sizeof(double2) * tab;
  // the shared-P_n form is instantiated for the nucleon's Cg5 diquark at up
  // to 32 momentum slots; other launches read P_n as a third propagator
  if constexpr (GIJ == 5)
    if (pa.share_n && nt <= 2) {
      if (nt == 1)
        launch_project_nt<1, Real, GIJ, SCALED, TERMS, true>(a, pa, tab_bytes, stream);
      else
        launch_project_nt<2, Real, GIJ, SCALED, TERMS, true>(a, pa, tab_bytes, stream);
      return;
    }
  switch (nt) {
  case 1: launch_project_nt<1, Real, GIJ, SCALED, TERMS>(a, pa, tab_bytes, stream); break;
  case 2: launch_project_nt<2, Real, GIJ, SCALED, TERMS>(a, pa, tab_bytes, stream); break;
  case 3: launch_project_nt<3, Real, GIJ, SCALED, TERMS>(a, pa, tab_bytes, stream); break;
  default: launch_project_nt<4, Real, GIJ, SCALED, TERMS>(a, pa, tab_bytes, stream); break;
  }
}

template <typename Real, bool SCALED = false>
static void launch_project_t(const BaryonProjectArgs &a, const ProjectArgs &pa, hipStream_t stream)
{
  // FP64 takes the nucleon kernel where it runs two waves per SIMD (one
  // momentum tile) or shares P_n; at more tiles the one-wave FP64 kernels
  // measured 1% faster in the runtime form
  const bool nucleon = a.nterm == 2 && a.types[0] == 0 && a.types[1] == 5 &&
                       (sizeof(Real) == 4 || pa.share_n || pa.np <= 16);
  if (a.gamma_ij == 5 && nucleon) return launch_project_g<Real, 5, SCALED, TERMS_NUCLEON>(a, pa, stream);
  if (a.gamma_ij == 5)
    launch_project_g<Real, 5, SCALED, TERMS_RUNTIME>(a, pa, stream);
  else
    launch_project_g<Real, -1, SCALED, TERMS_RUNTIME>(a, pa, stream);
}

void launch_baryon_project(const BaryonProjectArgs &a, const ProjectArgs &pa, hipStream_t stream)
{
  launch_project_t<double>(a, pa, stream);
}

void launch_baryon_project_fp32(const BaryonProjectArgs &a, const ProjectArgs &pa, hipStream_t stream)
{
  launch_project_t<float>(a, pa, stream);
}

void launch_baryon_project_fp32_scaled(const BaryonProjectArgs &a, const ProjectArgs &pa, hipStream_t stream)
{
  launch_project_t<float, true>(a, pa, stream);
}

} // namespace contract_amd
