// threept.hip -- the three-point step after the sequential source:
// contract_threepoint_project (include/contract_ext.h).
//
// The reference builds the sequential source seq_X = dC/dP_X of a two-point
// function (include/baryon_sequential.cuh:165-267, lib/contract.cpp:46-63);
// the caller solves on it for the sequential propagator Sigma.  The
// three-point function of a current Gamma inserted at x on quark line X is
//
//   C3(x; Gamma) = Tr_s Tr_c[ g5 Gamma F(x) Sigma(x)^dagger ]
//                = sum_{a,b} (g5 Gamma)[a][b] M[b][a](x),
//   M[b][a](x)   = sum_{s,c,c'} F[b][s](c,c') conj(Sigma[a][s](c,c')),
//
// F being the forward propagator of line X.  M is a 16-component open-spin
// matrix per site, so the kernel is the fused projection of the baryon
// kernels with a cheaper per-site step: lane q = 4a + b of a site's 16 lanes
// forms M[b][a] (36 complex multiply-adds out of LDS) and the FP64 MFMA
// projection accumulates it per momentum; the finish applies W = g5 Gamma
// for every insertion after the timeslice sums.  Two propagators are read
// per site (4608 B FP64, 2304 B FP32); nothing is written per site.

#include <contract_ext.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "baryon.hpp"
#include "device.hpp"
#include "plan.hpp"
#include "project.hpp"
#include "runtime.hpp"
#include "spin.hpp"

namespace contract_amd
{

template <int NT, typename Real, bool SCALED>
__global__ __launch_bounds__(WAVE) void threepoint_project_kernel(ProjectArgs args)
{
  using C = cplx<Real>;
  constexpr int SL = site_lds<C>();
  __shared__ C lds[2][SITES_PER_WAVE * SL];
  extern __shared__ double2 tab[]; // Gx + Gy + Gz phase entries, sized at launch
  const int lane = threadIdx.x;
  const int g = lane >> 4, q = lane & 15;
  const int a = q >> 2, b = q & 3; // lane q holds M[b][a], the R layout of the finish
  const unsigned item = blockIdx.x;
  if (item >= args.total_items) return;
  int r = 0;
  while (r + 1 < args.nreq && item >= args.req[r + 1].item0) ++r;
  const ProjectReq &pa = args.req[r];
  size_t first, last;
  item_range(pa, item - pa.item0, first, last);
  load_phase_tab(tab, pa, lane);
  const C *F = (const C *)pa.propag_i, *S = (const C *)pa.propag_j;

  auto tile_sites = [](size_t s, size_t end) {
    const size_t left = end - s;
    return left < (size_t)SITES_PER_WAVE ? (int)left : SITES_PER_WAVE;
  };
  // site-scaled storage: e_F and e_Sigma of this lane's site (padded lanes
  // repeat the last site; their M is zero), summed where the step applies
  // them: a sum formed as they load would wait for them right after the next
  // tile's loads are issued, and so (loads count down in issue order) for
  // the whole prefetch
  auto site_exp = [&](size_t st, int n) {
    const size_t k = st + (size_t)(g < n ? g : n - 1);
    return make_int2(pa.exp_i[k], pa.exp_j[k]);
  };

  PhaseWalk<NT> walk;
  walk.init(args, pa, first, g, q);
  dbl4 acc[NT][4];
#pragma unroll
  for (int tt = 0; tt < NT; ++tt)
#pragma unroll
    for (int u = 0; u < 4; ++u) acc[tt][u] = dbl4{0.0, 0.0, 0.0, 0.0};

  Tile<C> bf, bs;
  int2 ex_buf = make_int2(0, 0);
  {
    const int n = tile_sites(first, last);
    bf.load(F, first, n, lane);
    bs.load(S, first, n, lane);
    if constexpr (SCALED) ex_buf = site_exp(first, n);
  }
  for (size_t s = first; s < last; s += SITES_PER_WAVE) {
    bf.store(lds[0], lane);
    bs.store(lds[1], lane);
    __syncthreads();
    int ex = 0;
    const int2 ex_tile = ex_buf;
    const size_t next = s + SITES_PER_WAVE;
    if (next < last) {
      const int n = tile_sites(next, last);
      bf.load(F, next, n, lane);
      bs.load(S, next, n, lane);
      if constexpr (SCALED) ex_buf = site_exp(next, n);
    }
    C *sf = lds[0] + g * SL, *ss = lds[1] + g * SL;
    if constexpr (sizeof(Real) == 4 && !SCALED) {
      // plain complex64: rescale both sites by exact powers of two so the
      // FP32 products cannot underflow, and count the exponents in FP64
      ex = normalise_site(sf, q) + normalise_site(ss, q);
      __syncthreads();
    }
    C m = czero<C>();
#pragma unroll
    for (int sp = 0; sp < 4; ++sp) {
      const C *fb = sf + (b * 4 + sp) * 9, *sa = ss + (a * 4 + sp) * 9;
#pragma unroll
      for (int e = 0; e < 9; ++e) m = cfma_conj(m, fb[e], sa[e]);
    }
    double2 R = make_double2((double)m.x, (double)m.y);
    if constexpr (SCALED) ex = ex_tile.x + ex_tile.y;
    if constexpr (sizeof(Real) == 4) R = make_double2(ldexp(R.x, ex), ldexp(R.y, ex));
    project_accumulate<NT>(acc, args.np, pa, tab, R, walk); // padded lanes carry M = 0
    walk.advance(pa);
    __syncthreads();
  }
  project_store<NT>(acc, args.np, pa, item - pa.item0, lane);
}

template <typename Real, bool SCALED>
static void launch_threepoint_t(const ProjectArgs &pa, hipStream_t stream)
{
  int tab = 0;
  for (int r = 0; r < pa.nreq; ++r)
    tab = std::max(tab, pa.req[r].gdims[0] + pa.req[r].gdims[1] + pa.req[r].gdims[2]);
  const size_t tab_bytes = sizeof(double2) * tab;
  const dim3 grid(pa.total_items), block(WAVE);
  switch ((pa.np + 15) / 16) {
  case 1: hipLaunchKernelGGL((threepoint_project_kernel<1, Real, SCALED>), grid, block, tab_bytes, stream, pa); break;
  case 2: hipLaunchKernelGGL((threepoint_project_kernel<2, Real, SCALED>), grid, block, tab_bytes, stream, pa); break;
  case 3: hipLaunchKernelGGL((threepoint_project_kernel<3, Real, SCALED>), grid, block, tab_bytes, stream, pa); break;
  default:
    hipLaunchKernelGGL((threepoint_project_kernel<4, Real, SCALED>), grid, block, tab_bytes, stream, pa);
    break;
  }
}

} // namespace contract_amd

using namespace contract_amd;

extern "C" {

size_t contract_threepoint_workspace_bytes(const contract_threepoint_request *reqs, int nreq, int nmom)
{
  if (!reqs || nreq < 1 || nmom < 1 || nmom > MAX_MOM) return 0;
  size_t total = 0;
  for (int r = 0; r < nreq; ++r) {
    Geometry geo;
    const contract_sublattice &L = reqs[r].lattice;
    if (resolve_geometry(geo, L.dims, L.site_order, L.global, L.origin) != CONTRACT_OK) return 0;
    total += plan_project(geo.dims, geo.order, nmom, 0).total;
  }
  return total;
}

int contract_threepoint_project(const contract_threepoint_request *reqs, int nreq, int precision,
                                const double _Complex *insertions, int nins, const int *momenta, int nmom,
                                void *workspace, size_t workspace_bytes, void *stream_)
{
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  int rc;
  if (!reqs || nreq < 1) {
    set_last_error("nreq = %d must be >= 1 with a non-NULL request array", nreq);
    return CONTRACT_ERR_ARG;
  }
  std::vector<Geometry> geo(nreq);
  for (int r = 0; r < nreq; ++r) {
    const contract_sublattice &L = reqs[r].lattice;
    if ((rc = resolve_geometry(geo[r], L.dims, L.site_order, L.global, L.origin)) != CONTRACT_OK) return rc;
    if (!reqs[r].out || !reqs[r].propag_fwd || !reqs[r].propag_seq) {
      set_last_error("request %d: NULL output or propagator pointer", r);
      return CONTRACT_ERR_ARG;
    }
  }
  if ((rc = check_momenta(momenta, nmom)) != CONTRACT_OK) return rc;
  if (precision != CONTRACT_FP64 && precision != CONTRACT_FP32 && precision != CONTRACT_FP32_SCALED) {
    set_last_error("precision %d is not CONTRACT_FP64, CONTRACT_FP32 or CONTRACT_FP32_SCALED", precision);
    return CONTRACT_ERR_ARG;
  }
  if (nins < 1 || !insertions) {
    set_last_error("nins = %d must be >= 1 with a non-NULL insertions array", nins);
    return CONTRACT_ERR_ARG;
  }

  const std::vector<MomSlot> slots = pair_momenta(momenta, nmom);
  const int nslot = (int)slots.size();
  std::vector<ProjectPlan> plan(nreq);
  std::vector<size_t> offset(nreq);
  size_t total = 0;
  for (int r = 0; r < nreq; ++r) {
    plan[r] = plan_project(geo[r].dims, geo[r].order, nslot, 0);
    offset[r] = total;
    total += plan[r].total;
  }
  void *ws = workspace;
  bool own = false;
  if (ws) {
    if (workspace_bytes < total) {
      set_last_error("workspace of %zu bytes is smaller than the %zu needed", workspace_bytes, total);
      return CONTRACT_ERR_WORKSPACE;
    }
  } else {
    CONTRACT_EXT_HIP(hipMallocAsync(&ws, total, stream));
    own = true;
  }
  char *wsb = static_cast<char *>(ws);

  std::vector<ProjectReq> preq(nreq);
  for (int r = 0; r < nreq; ++r) {
    ProjectReq &R = preq[r];
    std::memset(&R, 0, sizeof(R));
    R.propag_i = reqs[r].propag_fwd;
    R.propag_j = reqs[r].propag_seq;
    R.propag_n = reqs[r].propag_fwd; // unused
    if (precision == CONTRACT_FP32_SCALED) {
      const size_t V = (size_t)geo[r].dims[0] * geo[r].dims[1] * geo[r].dims[2] * geo[r].dims[3];
      R.exp_i = static_cast<const signed char *>(R.propag_i) + V * SCALED_SITE_BYTES;
      R.exp_j = static_cast<const signed char *>(R.propag_j) + V * SCALED_SITE_BYTES;
      R.exp_n = R.exp_i;
    }
    R.partial = reinterpret_cast<double2 *>(wsb + offset[r]);
    R.slice_sites = plan[r].slice_sites;
    for (int mu = 0; mu < 4; ++mu) R.dims[mu] = geo[r].dims[mu];
    for (int mu = 0; mu < 3; ++mu) {
      R.gdims[mu] = geo[r].global[mu];
      R.origin[mu] = geo[r].origin[mu];
    }
    R.order = geo[r].order;
    R.chunk = plan[r].chunk;
    R.chunks_per_slice = plan[r].chunks_per_slice;
  }

  CONTRACT_EXT_HIP(timer_begin(stream));
  for (int r0 = 0; r0 < nreq; r0 += MAX_BATCH) {
    ProjectArgs pa;
    std::memset(&pa, 0, sizeof(pa));
    pa.np = nslot;
    for (int k = 0; k < nslot; ++k)
      for (int mu = 0; mu < 3; ++mu) pa.mom[k][mu] = momenta[3 * slots[k].rep + mu];
    pa.nreq = std::min(nreq - r0, MAX_BATCH);
    size_t items = 0;
    for (int r = 0; r < pa.nreq; ++r) {
      pa.req[r] = preq[r0 + r];
      pa.req[r].item0 = (unsigned)items;
      items += plan[r0 + r].items;
    }
    pa.total_items = (unsigned)items;
    if (precision == CONTRACT_FP64)
      launch_threepoint_t<double, false>(pa, stream);
    else if (precision == CONTRACT_FP32)
      launch_threepoint_t<float, false>(pa, stream);
    else
      launch_threepoint_t<float, true>(pa, stream);
    CONTRACT_EXT_HIP(hipGetLastError());
  }
  CONTRACT_EXT_HIP(timer_end(stream));

  // W_w = g5 Gamma_w: the finish forms sum_{m,n} W[m][n] R[n][m] with R = M
  std::vector<zhost> W((size_t)nins * 16);
  const double *gm = reinterpret_cast<const double *>(insertions);
  for (int w = 0; w < nins; ++w)
    for (int m = 0; m < 4; ++m)
      for (int n = 0; n < 4; ++n) {
        const size_t e = (size_t)w * 16 + m * 4 + n;
        const double g5 = m < 2 ? 1.0 : -1.0;
        W[e] = zhost(g5 * gm[2 * e], g5 * gm[2 * e + 1]);
      }
  for (int r0 = 0; r0 < nreq; r0 += MAX_BATCH) {
    FinishReq fr[MAX_BATCH];
    const int n = std::min(nreq - r0, MAX_BATCH);
    for (int k = 0; k < n; ++k) {
      const int r = r0 + k;
      fr[k].partial = preq[r].partial;
      fr[k].open = reinterpret_cast<double2 *>(wsb + offset[r] + plan[r].partial_bytes);
      fr[k].out = reqs[r].out;
      fr[k].lt = geo[r].dims[3];
      fr[k].order = geo[r].order;
      fr[k].chunks_per_slice = plan[r].chunks_per_slice;
    }
    launch_finish_all(fr, n, nslot, slots, nmom, W.data(), nins, stream);
  }
  CONTRACT_EXT_HIP(hipGetLastError());
  if (own) CONTRACT_EXT_HIP(hipFreeAsync(ws, stream));
  return CONTRACT_OK;
}

} // extern "C"
