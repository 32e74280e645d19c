// diquark.hip -- diquark and baryon sequential-source contractions on gfx950.
// Both produce a propagator-shaped [V][16][9] result, one 3x3 colour row per
// lane of a 16-lane site group, written back through LDS so the HBM stores
// are coalesced like the loads.

#include <type_traits>

#include "baryon.hpp"
#include "device.hpp"

namespace contract_amd
{

// FP32 site-scaled output (both kernels below): the staged mantissas are
// O(1) (each site's largest component in [1, 2)), so the FP32 cofactor sums
// cannot underflow; each output site is renormalised the same way, its
// exponent byte e_out = e_0 + e_1 + (the sum's own binary exponent), clamped
// to [-127, 127] like the converter's, the mantissas scaled by 2^(e_0 + e_1 -
// e_out) (exact unless the clamp pushes them below FP32's normal range).
// acc: this lane's 9 entries of its site's [16][9]; lane q = 0 of the site
// group writes e_out.
__device__ __forceinline__ void renormalise_output_site(float2 (&acc)[9], const signed char *e0,
                                                        const signed char *e1, signed char *eo, size_t first,
                                                        int nsites, int g, int q)
{
  float mx = 0.0f;
#pragma unroll
  for (int e = 0; e < 9; ++e) mx = fmaxf(mx, fmaxf(fabsf(acc[e].x), fabsf(acc[e].y)));
#pragma unroll
  for (int m = 8; m >= 1; m >>= 1) mx = fmaxf(mx, __shfl_xor(mx, m, 64));
  const size_t site = first + (size_t)(g < nsites ? g : nsites - 1); // padded lanes repeat the last site
  const int ein = (int)e0[site] + (int)e1[site];
  int eout = 0;
  if (mx > 0.0f) {
    eout = ein + ilogbf(mx);
    eout = eout < -SCALED_EXP_MAX ? -SCALED_EXP_MAX : (eout > SCALED_EXP_MAX ? SCALED_EXP_MAX : eout);
  }
  const int sh = mx > 0.0f ? ein - eout : 0;
#pragma unroll
  for (int e = 0; e < 9; ++e) acc[e] = make_float2(ldexpf(acc[e].x, sh), ldexpf(acc[e].y, sh));
  if (q == 0 && g < nsites) eo[first + g] = (signed char)eout;
}

// D[m,l](a,d) = g5[m] g5[l] conj( sum_i Gij[i,j] Gkl[k,l] cof(P_i[i,k], P_j[j,m])[a,d] )
// with j = col_ij(i), k = col_kl^T(l)  (include/diquark.cuh:32-58); lane q = m*4 + l.
// SCALED: FP32 site-scaled storage in and out.
template <bool SCALED>
__global__ __launch_bounds__(WAVE) void diquark_kernel(DiquarkArgs a)
{
  using C = typename std::conditional<SCALED, float2, double2>::type;
  using R = decltype(C::x);
  constexpr int SL = site_lds<C>();
  __shared__ C lds[2 * SITES_PER_WAVE * SL];
  const int lane = threadIdx.x;
  const int g = lane >> 4, q = lane & 15;
  const int m = q >> 2, l = q & 3;
  const int k = sparse_col(a.gklT, l);
  const C gl = sparse_val<C>(a.gklT, l);
  const R s5 = ((m >> 1) ^ (l >> 1)) ? (R)-1 : (R)1; // g5[m] g5[l]
  const C *P[2] = {(const C *)a.propag_i, (const C *)a.propag_j};
  persistent_tiles<2>(lds, P, a.volume, lane, [&](size_t first, int nsites) {
    const C *si = lds + g * SL, *sj = lds + (SITES_PER_WAVE + g) * SL;
    C acc[9], x[9], y[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) acc[e] = czero<C>();
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int j = sparse_col(a.gij, i);
      load_block(x, si, i, k);
      load_block(y, sj, j, m);
      cof_acc(acc, x, y, cmul(sparse_val<C>(a.gij, i), gl));
    }
    if constexpr (SCALED) renormalise_output_site(acc, a.exp_i, a.exp_j, a.exp_out, first, nsites, g, q);
    __syncthreads();
    C *row = lds + g * SL + q * 9;
#pragma unroll
    for (int e = 0; e < 9; ++e) row[e] = cmake<C>(s5 * acc[e].x, -s5 * acc[e].y);
    __syncthreads();
    unstage_sites(static_cast<C *>(a.diquark), lds, first, nsites, lane);
  });
}

template <bool SCALED>
static void launch_diquark_form(const DiquarkArgs &a, hipStream_t stream)
{
  const size_t ntiles = (a.volume + SITES_PER_WAVE - 1) / SITES_PER_WAVE;
  if (ntiles == 0) return;
  hipLaunchKernelGGL(diquark_kernel<SCALED>, dim3((unsigned)persistent_grid(ntiles, 8)), dim3(WAVE), 0, stream, a);
}

void launch_diquark(const DiquarkArgs &a, hipStream_t stream) { launch_diquark_form<false>(a, stream); }

void launch_diquark_fp32_scaled(const DiquarkArgs &a, hipStream_t stream) { launch_diquark_form<true>(a, stream); }

// ---------------------------------------------------------------------------
// Sequential source: seq_X = dC/dP_X for X in {i, j, n}
// (include/baryon_sequential.cuh:36-267, lib/contract.cpp:46-63).
//
// Lane q = r*4 + c owns output row [r][c] of P_X (sink r, source c).  Fixing
// the row pins three of the six spin indices; one more index stays free and
// the row is a sum of four colour cofactors of the two remaining propagators,
// gathered by the lane that owns it -- no atomics.  The one exception is the
// target P_n with source letter m (types IK_JL_NM, IL_JK_NM): there both i and
// l are free, so the 16 cofactors are spread over the 16 lanes and summed with
// a lane all-reduce; only the four rows with Gmn[c,r] != 0 are nonzero.
// ---------------------------------------------------------------------------
enum : int { SLK = 0, SLL = 1, SLM = 2 };
__host__ __device__ constexpr int seq_slot(int type, int prop)
{
  constexpr int tab[6][3] = {{SLK, SLL, SLM}, {SLK, SLM, SLL}, {SLL, SLK, SLM},
                             {SLL, SLM, SLK}, {SLM, SLK, SLL}, {SLM, SLL, SLK}};
  return tab[type][prop];
}

// Given the spin indices (i, j, k, l, m, n) of one term, accumulate
// w * cof(other1 block, other2 block) into acc.
template <int T, int X, typename C>
__device__ __forceinline__ void seq_term(C (&acc)[9], const C *s0, const C *s1, int i, int j, int k, int l, int m,
                                         int n, C w)
{
  const int sink[3] = {i, j, n};
  const int idx[3] = {k, l, m};
  constexpr int o0 = (X == 0) ? 1 : 0;
  constexpr int o1 = (X == 2) ? 1 : 2;
  C y[9], z[9];
  load_block(y, s0, sink[o0], idx[seq_slot(T, o0)]);
  load_block(z, s1, sink[o1], idx[seq_slot(T, o1)]);
  cof_acc(acc, y, z, w);
}

// SCALED: FP32 site-scaled storage in and out (renormalise_output_site).
template <int T, int X, bool SCALED = false>
__global__ __launch_bounds__(WAVE) void sequential_kernel(SequentialArgs a)
{
  using C = typename std::conditional<SCALED, float2, double2>::type;
  constexpr int SL = site_lds<C>();
  __shared__ C lds[2 * SITES_PER_WAVE * SL];
  const int lane = threadIdx.x;
  const int g = lane >> 4, q = lane & 15;
  const C *P[2] = {(const C *)a.in0, (const C *)a.in1};
  persistent_tiles<2>(lds, P, a.volume, lane, [&](size_t first, int nsites) {
    const C *s0 = lds + g * SL, *s1 = lds + (SITES_PER_WAVE + g) * SL;
    const int r = q >> 2, c = q & 3;
    constexpr int LX = seq_slot(T, X);
    C acc[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) acc[e] = czero<C>();

    if constexpr (X == 2 && LX == SLM) {
      // dC/dP_n[n,m] = Gmn[m,n] * sum_{il} Gij[i,j] Gkl[k,l] cof(P_i[i,X], P_j[j,Y])
      const int i = q >> 2, l = q & 3;
      const int j = sparse_col(a.gij, i), k = sparse_col(a.gklT, l);
      const C w = cmul(sparse_val<C>(a.gij, i), sparse_val<C>(a.gklT, l));
      seq_term<T, X>(acc, s0, s1, i, j, k, l, 0, 0, w);
#pragma unroll
      for (int e = 0; e < 9; ++e) acc[e] = group16_allreduce(acc[e]);
      const bool live = (sparse_col(a.gmnT, r) == c); // row [n = r][m = c]
      const C gm = live ? sparse_val<C>(a.gmnT, r) : czero<C>();
#pragma unroll
      for (int e = 0; e < 9; ++e) acc[e] = cmul(gm, acc[e]);
    } else {
#pragma unroll
      for (int f = 0; f < 4; ++f) {
        int i = 0, j = 0, k = 0, l = 0, m = 0, n = 0;
        C w;
        // indices fixed by the output row
        if constexpr (X == 0) {
          i = r;
          j = sparse_col(a.gij, i);
          w = sparse_val<C>(a.gij, i);
        } else if constexpr (X == 1) {
          j = r;
          i = sparse_col(a.gijT, j);
          w = sparse_val<C>(a.gijT, j);
        } else {
          n = r;
          m = sparse_col(a.gmnT, n);
          w = sparse_val<C>(a.gmnT, n);
        }
        if constexpr (LX == SLK) {
          k = c;
          l = sparse_col(a.gkl, k);
          w = cmul(w, sparse_val<C>(a.gkl, k));
        } else if constexpr (LX == SLL) {
          l = c;
          k = sparse_col(a.gklT, l);
          w = cmul(w, sparse_val<C>(a.gklT, l));
        } else {
          m = c;
          n = sparse_col(a.gmn, m);
          w = cmul(w, sparse_val<C>(a.gmn, m));
        }
        // the free index
        if constexpr (X == 2) {
          i = f;
          j = sparse_col(a.gij, i);
          w = cmul(w, sparse_val<C>(a.gij, i));
        } else if constexpr (LX == SLM) {
          l = f;
          k = sparse_col(a.gklT, l);
          w = cmul(w, sparse_val<C>(a.gklT, l));
        } else {
          n = f;
          m = sparse_col(a.gmnT, n);
          w = cmul(w, sparse_val<C>(a.gmnT, n));
        }
        seq_term<T, X>(acc, s0, s1, i, j, k, l, m, n, w);
      }
    }
    if constexpr (SCALED) renormalise_output_site(acc, a.exp0, a.exp1, a.exp_out, first, nsites, g, q);
    __syncthreads();
    C *row = lds + g * SL + q * 9;
#pragma unroll
    for (int e = 0; e < 9; ++e) row[e] = acc[e];
    __syncthreads();
    unstage_sites(static_cast<C *>(a.out), lds, first, nsites, lane);
  });
}

template <int T, bool SCALED>
static void launch_seq_t(int target, const SequentialArgs &a, dim3 grid, hipStream_t stream)
{
  switch (target) {
  case 0: hipLaunchKernelGGL((sequential_kernel<T, 0, SCALED>), grid, dim3(WAVE), 0, stream, a); break;
  case 1: hipLaunchKernelGGL((sequential_kernel<T, 1, SCALED>), grid, dim3(WAVE), 0, stream, a); break;
  default: hipLaunchKernelGGL((sequential_kernel<T, 2, SCALED>), grid, dim3(WAVE), 0, stream, a); break;
  }
}

template <bool SCALED>
static void launch_sequential_form(int type, int target, const SequentialArgs &a, hipStream_t stream)
{
  const size_t ntiles = (a.volume + SITES_PER_WAVE - 1) / SITES_PER_WAVE;
  if (ntiles == 0) return;
  const dim3 grid((unsigned)persistent_grid(ntiles, 8));
  switch (type) {
  case 0: launch_seq_t<0, SCALED>(target, a, grid, stream); break;
  case 1: launch_seq_t<1, SCALED>(target, a, grid, stream); break;
  case 2: launch_seq_t<2, SCALED>(target, a, grid, stream); break;
  case 3: launch_seq_t<3, SCALED>(target, a, grid, stream); break;
  case 4: launch_seq_t<4, SCALED>(target, a, grid, stream); break;
  default: launch_seq_t<5, SCALED>(target, a, grid, stream); break;
  }
}

void launch_sequential(int type, int target, const SequentialArgs &a, hipStream_t stream)
{
  launch_sequential_form<false>(type, target, a, stream);
}

void launch_sequential_fp32_scaled(int type, int target, const SequentialArgs &a, hipStream_t stream)
{
  launch_sequential_form<true>(type, target, a, stream);
}

} // namespace contract_amd
