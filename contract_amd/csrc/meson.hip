// meson.hip -- meson two-point contractions on gfx950.
//
//   c(x) = Tr_s Tr_c[(g5 Gij) P_i (g5 Gkl) P_j^dagger]
//        = sum_{il} (g5 Gij)[j,i] (g5 Gkl)[k,l] sum_ad P_i[i,k](a,d) conj(P_j[j,l](a,d))
//
// (include/meson.cuh:32-50; the gamma factors are the reference's
// gamma_gamma5_data<true> = (g5 G)^T, include/gamma.cuh:249-269).  Lane
// q = i*4 + l of a 16-lane site group owns one (i, l) spin pair.

#include "baryon.hpp"
#include "device.hpp"

namespace contract_amd
{

__global__ __launch_bounds__(WAVE) void meson_kernel(MesonArgs a)
{
  using C = double2;
  constexpr int SL = site_lds<C>();
  __shared__ C lds[2 * SITES_PER_WAVE * SL];
  const int lane = threadIdx.x;
  const int g = lane >> 4, q = lane & 15;
  const int i = q >> 2, l = q & 3;
  const int j = sparse_col(a.a, i), k = sparse_col(a.b, l);
  const C w = cmul(sparse_val<C>(a.a, i), sparse_val<C>(a.b, l));
  const C *P[2] = {(const C *)a.propag_i, (const C *)a.propag_j};
  persistent_tiles<2>(lds, P, a.volume, lane, [&](size_t first, int nsites) {
    C x[9], y[9];
    load_block(x, lds + g * SL, i, k);
    load_block(y, lds + (SITES_PER_WAVE + g) * SL, j, l);
    C t = czero<C>();
#pragma unroll
    for (int e = 0; e < 9; ++e) t = cfma_conj(t, x[e], y[e]);
    const C c = group16_allreduce(cmul(w, t));
    if (q == 0 && g < nsites) static_cast<C *>(a.correl)[first + g] = c;
  });
}

void launch_meson(const MesonArgs &a, hipStream_t stream)
{
  const size_t ntiles = (a.volume + SITES_PER_WAVE - 1) / SITES_PER_WAVE;
  if (ntiles == 0) return;
  hipLaunchKernelGGL(meson_kernel, dim3((unsigned)persistent_grid(ntiles, 8)), dim3(WAVE), 0, stream, a);
}

// meson_all_source (Gij fixed, Gkl = G(g)) and meson_all_sink (Gkl fixed,
// Gij = G(g)) for all g at once (include/meson_all.cuh:66-72,96-102).  Lane
// (i, l) forms the four colour traces any of the sixteen running gammas can
// select,
//   source: K_il[k] = tr(P_i[i][k], P_j[j(i)][l])   (j(i) the fixed column)
//   sink:   K_il[j] = tr(P_i[i][k(l)], P_j[j][l])   (k(l) the fixed column),
// posts them to LDS (the site's consumed P_i area), and lane g then gathers
// the sixteen it needs for correlator g:
//   source: c_g = sum_l G_g[l] sum_i Gij[i] K_il[col_g(l)]
//   sink:   c_g = sum_i G_g[i] sum_l Gkl[l] K_il[col_g(i)]
// (G_g[r] and col_g(r) the value and column of row r of (g5 G(g))^T).  One
// LDS hand-over instead of a 16-lane reduce-scatter of sixteen weighted,
// selected terms per lane.  SINK is compile-time: one form's registers per
// kernel.
template <bool SINK> __global__ __launch_bounds__(WAVE) void meson_all_kernel(MesonAllArgs a)
{
  using C = double2;
  constexpr int SL = site_lds<C>();
  __shared__ C lds[2 * SITES_PER_WAVE * SL];
  const int lane = threadIdx.x;
  const int g = lane >> 4, q = lane & 15;
  const int i = q >> 2, l = q & 3;
  // this lane's gather (correlator q): the running gamma's value and column
  // per summed spin row, and the fixed gamma's values (uniform)
  C run_val[4], fix_val[4];
  int run_col[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    run_val[r] = sparse_val<C>(a.run[q], r);
    run_col[r] = sparse_col(a.run[q], r);
    fix_val[r] = make_double2(a.fixed.re[r], a.fixed.im[r]);
  }
  // read once: a lane-indexed kernel argument is a memory load, and one in
  // the step would make the output store wait for the next tile's loads
  C *const out = static_cast<C *>(a.correl[q]);
  const C *P[2] = {(const C *)a.propag_i, (const C *)a.propag_j};
  persistent_tiles<2>(lds, P, a.volume, lane, [&](size_t first, int nsites) {
    C *si = lds + g * SL;
    const C *sj = lds + (SITES_PER_WAVE + g) * SL;
    C x[9], y[9], K[4];
    if constexpr (!SINK) {
      load_block(y, sj, sparse_col(a.fixed, i), l);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        load_block(x, si, i, k);
        C t = czero<C>();
#pragma unroll
        for (int e = 0; e < 9; ++e) t = cfma_conj(t, x[e], y[e]);
        K[k] = t;
      }
    } else {
      load_block(x, si, i, sparse_col(a.fixed, l));
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        load_block(y, sj, j, l);
        C t = czero<C>();
#pragma unroll
        for (int e = 0; e < 9; ++e) t = cfma_conj(t, x[e], y[e]);
        K[j] = t;
      }
    }
    wave_sync(); // every lane's tile reads are done: the site's P_i area takes K
#pragma unroll
    for (int k = 0; k < 4; ++k) si[q * 4 + k] = K[k]; // K_il[k] at slot 4 (4i + l) + k
    wave_sync();
    C c = czero<C>();
#pragma unroll
    for (int r = 0; r < 4; ++r) { // r = l (source) or i (sink): the row the running gamma weights
      C t = czero<C>();
#pragma unroll
      for (int m = 0; m < 4; ++m) { // m = i (source) or l (sink): the row the fixed gamma weights
        const int il = SINK ? 4 * r + m : 4 * m + r;
        t = cfma(t, fix_val[m], si[il * 4 + run_col[r]]);
      }
      c = cfma(c, run_val[r], t);
    }
    if (g < nsites) out[first + g] = c;
  });
}

void launch_meson_all(const MesonAllArgs &a, hipStream_t stream)
{
  const size_t blocks = (a.volume + SITES_PER_WAVE - 1) / SITES_PER_WAVE;
  if (blocks == 0) return;
  const dim3 grid((unsigned)persistent_grid(blocks, 8));
  if (a.sink)
    hipLaunchKernelGGL(meson_all_kernel<true>, grid, dim3(WAVE), 0, stream, a);
  else
    hipLaunchKernelGGL(meson_all_kernel<false>, grid, dim3(WAVE), 0, stream, a);
}

} // namespace contract_amd
