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
// (i, l) forms the four colour traces that any of the sixteen running gammas
// can select, weights them, and a 16-lane reduce-scatter leaves correlator g
// in lane g.  SINK is compile-time: one form's registers per kernel.
template <bool SINK> __global__ __launch_bounds__(WAVE) void meson_all_kernel(MesonAllArgs a)
{
  using C = double2;
  constexpr int SL = site_lds<C>();
  __shared__ C lds[2 * SITES_PER_WAVE * SL];
  const int lane = threadIdx.x;
  const int g = lane >> 4, q = lane & 15;
  const C *P[2] = {(const C *)a.propag_i, (const C *)a.propag_j};
  persistent_tiles<2>(lds, P, a.volume, lane, [&](size_t first, int nsites) {
    const C *si = lds + g * SL, *sj = lds + (SITES_PER_WAVE + g) * SL;
    const int i = q >> 2, l = q & 3;
    C x[9], y[9], K[4], v[16];
    if constexpr (!SINK) {
      // fixed row pair (i, j = a(i)); running k
      const int j = sparse_col(a.fixed, i);
      load_block(y, sj, j, l);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        load_block(x, si, i, k);
        C t = czero<C>();
#pragma unroll
        for (int e = 0; e < 9; ++e) t = cfma_conj(t, x[e], y[e]);
        K[k] = t;
      }
      const C fa = sparse_val<C>(a.fixed, i);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int k = sparse_col(a.run[r], l);
        const C sel = mux4(K[0], K[1], K[2], K[3], k);
        v[r] = cmul(cmul(fa, sparse_val<C>(a.run[r], l)), sel);
      }
    } else {
      const int k = sparse_col(a.fixed, l);
      load_block(x, si, i, k);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        load_block(y, sj, j, l);
        C t = czero<C>();
#pragma unroll
        for (int e = 0; e < 9; ++e) t = cfma_conj(t, x[e], y[e]);
        K[j] = t;
      }
      const C fb = sparse_val<C>(a.fixed, l);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int j = sparse_col(a.run[r], i);
        const C sel = mux4(K[0], K[1], K[2], K[3], j);
        v[r] = cmul(cmul(sparse_val<C>(a.run[r], i), fb), sel);
      }
    }
    const C c = group16_reduce_scatter(v, q);
    if (g < nsites) static_cast<C *>(a.correl[q])[first + g] = c;
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
