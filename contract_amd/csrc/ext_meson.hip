// ext_meson.hip -- host side of the batched meson Gamma x Gamma' grid
// (include/contract_ext.h: contract_meson_grid_project): validation, the work
// decomposition of a batch of uneven requests, and the launch sequence.

#include <contract_ext.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "baryon.hpp"
#include "meson_grid.hpp"
#include "plan.hpp"
#include "runtime.hpp"
#include "spin.hpp"

namespace contract_amd
{

namespace
{

struct GridTuning {
  int chunk, tile, bpc, unit; // unit: 1 vector-ALU traces, 2 matrix-core traces
};

// Library defaults, each the fastest plan for its storage on a batch of 24
// lattices from 16^3x32 to 32^3x64 at 7 momenta: 1024 sites per item; FP64
// the shared-tile matrix-core form, tile 4; both FP32 forms the wave-local
// form (plain FP32 scaled per site in the kernel).  An explicit vector-ALU
// plan for plain FP32 defaults to tile 8.  A zero blocks_per_cu takes the
// default of the form each pass runs (pass_blocks_per_cu).
GridTuning resolve_tuning(const contract_meson_grid_tuning *t, int precision)
{
  GridTuning g{1024, 0, 0, 0};
  if (t) {
    g.tile = t->tile_sites;
    if (t->sites_per_item) g.chunk = t->sites_per_item;
    g.bpc = t->blocks_per_cu;
    g.unit = t->trace_unit;
  }
  if (g.unit == 0) g.unit = precision == CONTRACT_FP64 ? 2 : 3;
  if (g.tile == 0) g.tile = precision == CONTRACT_FP32 && g.unit == 1 ? 8 : 4;
  return g;
}

// Persistent workgroups per CU of a pass: the one-wave workgroups of the
// wave-local form fill one SIMD each (two for the FP32 forms, which fit two
// waves per SIMD); three four-wave workgroups for the
// shared-tile matrix-core form, two for the vector-ALU form.
int pass_blocks_per_cu(const GridTuning &g, bool wave_pass, int precision)
{
  if (g.bpc) return g.bpc;
  if (wave_pass) return precision == CONTRACT_FP64 ? 4 : 8;
  return g.unit >= 2 ? 3 : 2;
}

int check_tuning(const GridTuning &g, int precision)
{
  if (g.unit < 1 || g.unit > 3) {
    set_last_error("trace_unit = %d is not 0, 1, 2 or 3", g.unit);
    return CONTRACT_ERR_ARG;
  }
  if (g.unit == 2 && precision == CONTRACT_FP32) {
    set_last_error("trace_unit = 2 (shared-tile matrix-core traces) needs CONTRACT_FP64 or CONTRACT_FP32_SCALED "
                   "storage; plain CONTRACT_FP32 takes 1 or 3");
    return CONTRACT_ERR_ARG;
  }
  if (g.tile != 4 && g.tile != 8) {
    set_last_error("tile_sites = %d is not 4 or 8", g.tile);
    return CONTRACT_ERR_ARG;
  }
  if (g.chunk < 1 || g.chunk > (1 << 24)) {
    set_last_error("sites_per_item = %d out of range 1..%d", g.chunk, 1 << 24);
    return CONTRACT_ERR_ARG;
  }
  if (g.bpc < 0 || g.bpc > 8) {
    set_last_error("blocks_per_cu = %d out of range 1..8 (0: the default)", g.bpc);
    return CONTRACT_ERR_ARG;
  }
  return CONTRACT_OK;
}

int chunk_of(const GridTuning &g) { return (g.chunk + g.tile - 1) / g.tile * g.tile; }

// Work items of request r: whole slices (a timeslice, or one parity of it)
// cut into runs of `chunk` sites.
struct ReqPlan {
  int slices;
  unsigned slice_sites;
  int chunks_per_slice;
  size_t items;
};

ReqPlan plan_request(const contract_meson_request &r, int chunk)
{
  ReqPlan p{};
  const size_t V = (size_t)r.dims[0] * r.dims[1] * r.dims[2] * r.dims[3];
  p.slices = r.dims[3] * (r.site_order ? 2 : 1);
  p.slice_sites = (unsigned)(V / p.slices);
  p.chunks_per_slice = (int)((p.slice_sites + chunk - 1) / chunk);
  p.items = (size_t)p.slices * p.chunks_per_slice;
  return p;
}

int check_requests(const contract_meson_request *reqs, int nreq)
{
  if (nreq < 1 || !reqs) {
    set_last_error("nreq = %d must be >= 1 with a non-NULL request array", nreq);
    return CONTRACT_ERR_ARG;
  }
  for (int r = 0; r < nreq; ++r) {
    Geometry geo;
    int rc = resolve_geometry(geo, reqs[r].dims, reqs[r].site_order, reqs[r].global, reqs[r].origin);
    if (rc != CONTRACT_OK) return rc;
    const int *d = reqs[r].dims;
    if (d[0] + d[1] + d[2] > MAX_PHASE_TAB) {
      set_last_error("request %d: Lx + Ly + Lz = %d exceeds %d", r, d[0] + d[1] + d[2], MAX_PHASE_TAB);
      return CONTRACT_ERR_ARG;
    }
    const size_t V = (size_t)d[0] * d[1] * d[2] * d[3];
    if (V / d[3] > (size_t)1 << 31) {
      set_last_error("request %d: timeslice of %zu sites is too large", r, V / d[3]);
      return CONTRACT_ERR_ARG;
    }
    if (!reqs[r].propag_i || !reqs[r].propag_j || !reqs[r].out) {
      set_last_error("request %d: NULL propagator or output pointer", r);
      return CONTRACT_ERR_ARG;
    }
  }
  return CONTRACT_OK;
}

// Largest item count of any launch group (GRID_MAX_REQ consecutive requests).
size_t max_group_items(const contract_meson_request *reqs, int nreq, int chunk)
{
  size_t best = 0;
  for (int r0 = 0; r0 < nreq; r0 += GRID_MAX_REQ) {
    size_t n = 0;
    for (int r = r0; r < std::min(nreq, r0 + GRID_MAX_REQ); ++r) n += plan_request(reqs[r], chunk).items;
    best = std::max(best, n);
  }
  return best;
}

} // namespace

} // namespace contract_amd

using namespace contract_amd;

// Row tiles of each work item's partial: two when a call may have more than
// eight momentum slots (its passes then take up to sixteen), from the
// momentum count alone so the workspace size needs no momenta.
static int partial_tiles(int nmom) { return nmom > GRID_MOM_PER_TILE ? GRID_MAX_TILES : 1; }

extern "C" {

size_t contract_meson_grid_workspace_bytes(const contract_meson_request *reqs, int nreq, int nmom,
                                           const contract_meson_grid_tuning *tuning)
{
  const GridTuning g = resolve_tuning(tuning, CONTRACT_FP64); // the unit does not change the workspace
  if (check_tuning(g, CONTRACT_FP64) != CONTRACT_OK || check_requests(reqs, nreq) != CONTRACT_OK || nmom < 1 ||
      nmom > MAX_MOM)
    return 0;
  return max_group_items(reqs, nreq, chunk_of(g)) * partial_tiles(nmom) * GRID_PARTIAL_TILE * sizeof(double);
}

int contract_meson_grid_project(const contract_meson_request *reqs, int nreq, int precision, const int *momenta,
                                int nmom, const contract_meson_grid_tuning *tuning, void *workspace,
                                size_t workspace_bytes, void *stream_)
{
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  int rc;
  if (precision != CONTRACT_FP64 && precision != CONTRACT_FP32 && precision != CONTRACT_FP32_SCALED) {
    set_last_error("precision %d is not CONTRACT_FP64, CONTRACT_FP32 or CONTRACT_FP32_SCALED", precision);
    return CONTRACT_ERR_ARG;
  }
  const GridTuning g = resolve_tuning(tuning, precision);
  if ((rc = check_tuning(g, precision)) != CONTRACT_OK) return rc;
  if ((rc = check_requests(reqs, nreq)) != CONTRACT_OK) return rc;
  if ((rc = check_momenta(momenta, nmom)) != CONTRACT_OK) return rc;
  const int chunk = chunk_of(g);
  const size_t group_items = max_group_items(reqs, nreq, chunk);
  if (group_items >= ((size_t)1 << 32)) {
    set_last_error("batch of %zu work items is too large; raise sites_per_item", group_items);
    return CONTRACT_ERR_ARG;
  }
  const int tiles = partial_tiles(nmom);
  const size_t need = group_items * tiles * GRID_PARTIAL_TILE * sizeof(double);
  void *ws = workspace;
  bool own = false;
  if (ws) {
    if (workspace_bytes < need) {
      set_last_error("workspace of %zu bytes is smaller than the %zu needed", workspace_bytes, need);
      return CONTRACT_ERR_WORKSPACE;
    }
  } else {
    CONTRACT_EXT_HIP(hipMallocAsync(&ws, need, stream));
    own = true;
  }

  const std::vector<MomSlot> slots = pair_momenta(momenta, nmom);
  MesonGridFinish f;
  std::memset(&f, 0, sizeof(f));
  for (int gm = 0; gm < 16; ++gm) f.g5t[gm] = sparse_of(mat_transpose(gamma5_times(gm)));
  f.np_total = nmom;
  f.partial = static_cast<const double *>(ws);

  CONTRACT_EXT_HIP(timer_begin(stream));
  for (int r0 = 0; r0 < nreq; r0 += GRID_MAX_REQ) {
    const int nr = std::min(nreq - r0, GRID_MAX_REQ);
    MesonGridArgs a;
    std::memset(&a, 0, sizeof(a));
    a.nreq = nr;
    a.chunk = chunk;
    a.partial = static_cast<double *>(ws);
    a.partial_tiles = tiles;
    f.partial_tiles = tiles;
    f.nreq = nr;
    size_t items = 0;
    int blocks = 0, tab = 0;
    for (int r = 0; r < nr; ++r) {
      const contract_meson_request &q = reqs[r0 + r];
      const ReqPlan p = plan_request(q, chunk);
      MesonGridReq &R = a.req[r];
      R.propag_i = q.propag_i;
      R.propag_j = q.propag_j;
      if (precision == CONTRACT_FP32_SCALED) { // the site exponents follow the complex64 entries
        const size_t V = (size_t)q.dims[0] * q.dims[1] * q.dims[2] * q.dims[3];
        R.exp_i = static_cast<const signed char *>(q.propag_i) + V * SCALED_SITE_BYTES;
        R.exp_j = static_cast<const signed char *>(q.propag_j) + V * SCALED_SITE_BYTES;
      }
      Geometry geo;
      resolve_geometry(geo, q.dims, q.site_order, q.global, q.origin); // validated above
      for (int mu = 0; mu < 4; ++mu) R.dims[mu] = q.dims[mu];
      for (int mu = 0; mu < 3; ++mu) {
        R.gdims[mu] = geo.global[mu];
        R.origin[mu] = geo.origin[mu];
      }
      R.order = q.site_order;
      R.chunks_per_slice = p.chunks_per_slice;
      R.slice_sites = p.slice_sites;
      R.item0 = (unsigned)items;
      MesonGridOut &O = f.req[r];
      O.out = q.out;
      O.lt = q.dims[3];
      O.order = q.site_order;
      O.chunks_per_slice = p.chunks_per_slice;
      O.item0 = (unsigned)items;
      O.block0 = blocks;
      items += p.items;
      blocks += q.dims[3];
      tab = std::max(tab, q.dims[0] + q.dims[1] + q.dims[2] + 3);
    }
    a.total_items = (unsigned)items;
    const int storage =
        precision == CONTRACT_FP32_SCALED ? GRID_FP32_SCALED : (precision == CONTRACT_FP32 ? GRID_FP32 : GRID_FP64);
    for (size_t p0 = 0; p0 < slots.size(); p0 += GRID_MOM_PER_PASS) {
      a.np = (int)std::min(slots.size() - p0, (size_t)GRID_MOM_PER_PASS);
      int bpc = pass_blocks_per_cu(g, grid_wave_pass(g.unit, a.np), precision);
      if (!g.bpc) {
        // a default plan runs no more persistent workgroups than fit on a
        // CU at once: the rest would start only as resident ones finish
        // their whole item lists
        const int fit = meson_grid_resident_blocks(a, storage, g.unit, g.tile, (size_t)tab * a.np);
        if (fit > 0 && fit < bpc) bpc = fit;
      }
      const size_t grid = std::min(items, (size_t)device_cu_count() * bpc);
      for (int p = 0; p < a.np; ++p) {
        const MomSlot &m = slots[p0 + p];
        for (int mu = 0; mu < 3; ++mu) a.mom[p][mu] = momenta[3 * m.rep + mu];
        f.out_plus[p] = m.plus;
        f.out_minus[p] = m.minus;
      }
      launch_meson_grid(a, storage, g.unit, g.tile, grid, (size_t)tab * a.np, stream);
      CONTRACT_EXT_HIP(hipGetLastError());
      f.np = a.np;
      launch_meson_grid_finish(f, blocks, stream);
      CONTRACT_EXT_HIP(hipGetLastError());
    }
  }
  CONTRACT_EXT_HIP(timer_end(stream));
  if (own) CONTRACT_EXT_HIP(hipFreeAsync(ws, stream));
  return CONTRACT_OK;
}

} // extern "C"
