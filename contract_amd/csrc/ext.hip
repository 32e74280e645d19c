// ext.hip -- extension API (include/contract_ext.h): fused contraction +
// timeslice/momentum projection, error reporting, workspace planning.

#include <contract_ext.h>

#include <atomic>
#include <cstdarg>
#include <cstring>
#include <vector>

#include "baryon.hpp"
#include "plan.hpp"
#include "runtime.hpp"
#include "spin.hpp"

namespace contract_amd
{

static thread_local char g_last_error[512] = "";
static thread_local hipEvent_t g_timer_start = nullptr, g_timer_stop = nullptr;
static thread_local hipStream_t g_abi_stream = nullptr;

hipStream_t abi_stream() { return g_abi_stream; }

hipError_t timer_begin(hipStream_t stream)
{
  return g_timer_start ? hipEventRecord(g_timer_start, stream) : hipSuccess;
}
hipError_t timer_end(hipStream_t stream) { return g_timer_stop ? hipEventRecord(g_timer_stop, stream) : hipSuccess; }

void set_last_error(const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

int device_cu_count()
{
  static std::atomic<int> cache[64]; // per device, 0 = not yet queried
  int dev = 0, n = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (dev >= 0 && dev < 64 && (n = cache[dev].load(std::memory_order_relaxed)) > 0) return n;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) return 256;
  if (dev >= 0 && dev < 64) cache[dev].store(n, std::memory_order_relaxed);
  return n;
}

size_t persistent_grid(size_t ntiles, int waves_per_cu)
{
  const size_t cap = (size_t)device_cu_count() * waves_per_cu;
  return ntiles < cap ? ntiles : cap;
}

static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

std::vector<MomSlot> pair_momenta(const int *mom, int nmom)
{
  std::vector<MomSlot> slots;
  std::vector<char> used(nmom, 0);
  for (int p = 0; p < nmom; ++p) {
    if (used[p]) continue;
    used[p] = 1;
    MomSlot s{p, p, -1};
    const bool zero = mom[3 * p] == 0 && mom[3 * p + 1] == 0 && mom[3 * p + 2] == 0;
    if (!zero)
      for (int q = p + 1; q < nmom; ++q)
        if (!used[q] && mom[3 * q] == -mom[3 * p] && mom[3 * q + 1] == -mom[3 * p + 1] &&
            mom[3 * q + 2] == -mom[3 * p + 2]) {
          used[q] = 1;
          s.minus = q;
          break;
        }
    slots.push_back(s);
  }
  return slots;
}

// Work decomposition of the fused baryon kernels (contract_set_baryon_plan).
// Defaults: 8 work items per CU, one item per resident wave of the two-wave
// staged kernel (one round, no tail of half-empty rounds; 16 measured 3%
// slower for the FP64 proton at 24^3x64 and 0-1.5% slower for the other
// forms), and at least 128 sites (32 four-site steps) per item: a shorter
// item spends a large part of its time on the prologue (phase tables, first
// tile) and the partial store (16^3x32: 4.48 -> 4.76 TB/s single, 4.82 ->
// 5.17 TB/s in a batch of 16).
constexpr int DEFAULT_ITEMS_PER_CU = 8;
constexpr int DEFAULT_MIN_ITEM_SITES = 128;
static thread_local contract_baryon_plan g_baryon_plan = {DEFAULT_ITEMS_PER_CU, DEFAULT_MIN_ITEM_SITES};

ProjectPlan plan_project(const int dims[4], int order, int nmom, int sites_per_item)
{
  ProjectPlan p{};
  const size_t V = (size_t)dims[0] * dims[1] * dims[2] * dims[3];
  p.slices = dims[3] * (order ? 2 : 1);
  p.slice_sites = V / p.slices;
  size_t chunk;
  if (sites_per_item > 0) {
    chunk = (size_t)sites_per_item;
  } else {
    // enough work items to fill the chip
    const size_t target_items = (size_t)device_cu_count() * (size_t)g_baryon_plan.items_per_cu;
    size_t cps = (target_items + p.slices - 1) / p.slices;
    if (cps < 1) cps = 1;
    chunk = (p.slice_sites + cps - 1) / cps;
  }
  if (sites_per_item <= 0 && chunk < (size_t)g_baryon_plan.min_item_sites) chunk = g_baryon_plan.min_item_sites;
  chunk = (chunk + SITES_PER_WAVE - 1) / SITES_PER_WAVE * SITES_PER_WAVE;
  if (chunk < SITES_PER_WAVE) chunk = SITES_PER_WAVE;
  p.chunk = (int)chunk;
  p.chunks_per_slice = (int)((p.slice_sites + chunk - 1) / chunk);
  p.items = (size_t)p.slices * p.chunks_per_slice;
  p.partial_bytes = align256(p.items * (size_t)nmom * 32 * sizeof(double2));
  p.open_bytes = align256((size_t)dims[3] * nmom * 32 * sizeof(double2));
  p.total = p.partial_bytes + p.open_bytes;
  return p;
}

int check_lattice(const int dims[4], int order)
{
  if (!dims) {
    set_last_error("dims is NULL");
    return CONTRACT_ERR_ARG;
  }
  for (int mu = 0; mu < 4; ++mu)
    if (dims[mu] <= 0 || dims[mu] > 4096) {
      set_last_error("dims[%d] = %d out of range", mu, dims[mu]);
      return CONTRACT_ERR_ARG;
    }
  if (order != 0 && order != 1) {
    set_last_error("site_order %d is not 0 (lexicographic) or 1 (even-odd)", order);
    return CONTRACT_ERR_ARG;
  }
  if (order == 1 && (dims[0] % 2)) {
    set_last_error("even-odd order needs an even Lx, got %d", dims[0]);
    return CONTRACT_ERR_ARG;
  }
  return CONTRACT_OK;
}

int check_momenta(const int *momenta, int nmom)
{
  if (nmom < 1 || nmom > MAX_MOM || !momenta) {
    set_last_error("nmom = %d must be in 1..%d with a non-NULL momenta array", nmom, MAX_MOM);
    return CONTRACT_ERR_ARG;
  }
  return CONTRACT_OK;
}

int resolve_geometry(Geometry &g, const int dims[4], int order, const int *global, const int *origin)
{
  int rc = check_lattice(dims, order);
  if (rc != CONTRACT_OK) return rc;
  for (int mu = 0; mu < 4; ++mu) g.dims[mu] = dims[mu];
  g.order = order;
  int tab = 0;
  for (int mu = 0; mu < 3; ++mu) {
    g.global[mu] = (global && global[mu] > 0) ? global[mu] : dims[mu];
    g.origin[mu] = origin ? origin[mu] : 0;
    if (g.global[mu] < dims[mu] || g.origin[mu] < 0 || g.origin[mu] >= g.global[mu]) {
      set_last_error("direction %d: global extent %d and origin %d do not hold a local extent of %d", mu,
                     g.global[mu], g.origin[mu], dims[mu]);
      return CONTRACT_ERR_ARG;
    }
    if (order == 1 && (g.origin[mu] & 1)) {
      set_last_error("even-odd order needs even origins (direction %d has %d)", mu, g.origin[mu]);
      return CONTRACT_ERR_ARG;
    }
    tab += g.global[mu];
  }
  if (tab > MAX_PHASE_TAB) {
    set_last_error("global Lx + Ly + Lz = %d exceeds %d", tab, MAX_PHASE_TAB);
    return CONTRACT_ERR_ARG;
  }
  return CONTRACT_OK;
}

// Per-timeslice reduction of the work-item partials, then the projectors
// (MAX_PROJ_PER_LAUNCH per launch), expanding each momentum slot to its +k
// and -k outputs; up to MAX_BATCH requests share each launch.
void launch_finish_all(const FinishReq *reqs, int nreq, int nslot, const std::vector<MomSlot> &slots, int nmom,
                       const zhost *W, int nproj, hipStream_t stream)
{
  ReduceArgs r;
  std::memset(&r, 0, sizeof(r));
  r.np = nslot;
  r.nq = 32;
  r.nreq = nreq;
  int total_lt = 0;
  for (int k = 0; k < nreq; ++k) {
    ReduceReq &q = r.req[k];
    q.partial = reqs[k].partial;
    q.open = reqs[k].open;
    q.lt = reqs[k].lt;
    q.order = reqs[k].order;
    q.chunks_per_slice = reqs[k].chunks_per_slice;
    q.t0 = total_lt;
    total_lt += q.lt;
  }
  launch_project_reduce(r, total_lt, stream);
  ApplyArgs f;
  std::memset(&f, 0, sizeof(f));
  f.nslot = nslot;
  f.nmom = nmom;
  for (int k = 0; k < (int)slots.size(); ++k) {
    f.slot[slots[k].plus] = k;
    f.sgn[slots[k].plus] = 1;
    if (slots[k].minus >= 0) {
      f.slot[slots[k].minus] = k;
      f.sgn[slots[k].minus] = -1;
    }
  }
  f.nreq = nreq;
  for (int base = 0; base < nproj; base += MAX_PROJ_PER_LAUNCH) {
    f.nproj = nproj - base < MAX_PROJ_PER_LAUNCH ? nproj - base : MAX_PROJ_PER_LAUNCH;
    f.proj_base = base;
    for (int w = 0; w < f.nproj; ++w)
      for (int q = 0; q < 16; ++q) {
        f.w_re[w][q] = W[(base + w) * 16 + q].real();
        f.w_im[w][q] = W[(base + w) * 16 + q].imag();
      }
    unsigned e = 0;
    for (int k = 0; k < nreq; ++k) {
      f.req[k].open = reqs[k].open;
      f.req[k].out = static_cast<double2 *>(reqs[k].out);
      f.req[k].lt = reqs[k].lt;
      f.req[k].e0 = e;
      e += (unsigned)(f.nproj * nmom * reqs[k].lt);
    }
    f.total = e;
    launch_project_apply(f, stream);
  }
}

} // namespace contract_amd

using namespace contract_amd;

extern "C" {

const char *contract_last_error(void) { return g_last_error; }

int contract_set_stream(void *stream)
{
  g_abi_stream = static_cast<hipStream_t>(stream);
  return CONTRACT_OK;
}

int contract_set_baryon_plan(const contract_baryon_plan *plan)
{
  if (!plan) {
    g_baryon_plan = {DEFAULT_ITEMS_PER_CU, DEFAULT_MIN_ITEM_SITES};
    return CONTRACT_OK;
  }
  contract_baryon_plan p = *plan;
  if (p.items_per_cu == 0) p.items_per_cu = DEFAULT_ITEMS_PER_CU;
  if (p.min_item_sites == 0) p.min_item_sites = DEFAULT_MIN_ITEM_SITES;
  if (p.items_per_cu < 1 || p.items_per_cu > 256) {
    set_last_error("items_per_cu = %d out of range 1..256", p.items_per_cu);
    return CONTRACT_ERR_ARG;
  }
  if (p.min_item_sites < 4 || p.min_item_sites > (1 << 20)) {
    set_last_error("min_item_sites = %d out of range 4..%d", p.min_item_sites, 1 << 20);
    return CONTRACT_ERR_ARG;
  }
  g_baryon_plan = p;
  return CONTRACT_OK;
}

int contract_get_baryon_plan(contract_baryon_plan *plan)
{
  if (!plan) {
    set_last_error("plan is NULL");
    return CONTRACT_ERR_ARG;
  }
  *plan = g_baryon_plan;
  return CONTRACT_OK;
}

int contract_set_kernel_timer(void *start, void *stop)
{
  g_timer_start = static_cast<hipEvent_t>(start);
  g_timer_stop = static_cast<hipEvent_t>(stop);
  return CONTRACT_OK;
}

int contract_gamma_matrix(int g, double _Complex *out)
{
  if (g < 0 || g > 15 || !out) {
    set_last_error("gamma index %d out of range 0..15", g);
    return CONTRACT_ERR_ARG;
  }
  const Mat4 G = gamma_matrix(g);
  double *o = reinterpret_cast<double *>(out);
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      o[2 * (r * 4 + c)] = G.v[r][c].real();
      o[2 * (r * 4 + c) + 1] = G.v[r][c].imag();
    }
  return CONTRACT_OK;
}

size_t contract_project_workspace_bytes(const int dims[4], int site_order, int nmom)
{
  if (check_lattice(dims, site_order) != CONTRACT_OK || nmom < 1 || nmom > MAX_MOM) return 0;
  return plan_project(dims, site_order, nmom, 0).total;
}

int contract_baryon_twopoint_project(void *out, const void *propag_i, const void *propag_j, const void *propag_n,
                                     int precision, const int *contract_types, const double _Complex *weights,
                                     int nterm, int gamma_ij, int gamma_kl, const double _Complex *projectors,
                                     int nproj, const int *momenta, int nmom, const int dims[4], int site_order,
                                     void *workspace, size_t workspace_bytes, void *stream)
{
  if (!dims) {
    set_last_error("dims is NULL");
    return CONTRACT_ERR_ARG;
  }
  contract_sublattice lat;
  std::memset(&lat, 0, sizeof(lat));
  for (int mu = 0; mu < 4; ++mu) lat.dims[mu] = dims[mu];
  lat.site_order = site_order;
  return contract_baryon_twopoint_project_sub(out, propag_i, propag_j, propag_n, precision, contract_types, weights,
                                              nterm, gamma_ij, gamma_kl, projectors, nproj, momenta, nmom, &lat,
                                              workspace, workspace_bytes, stream);
}

int contract_baryon_twopoint_project_sub(void *out, const void *propag_i, const void *propag_j,
                                         const void *propag_n, int precision, const int *contract_types,
                                         const double _Complex *weights, int nterm, int gamma_ij, int gamma_kl,
                                         const double _Complex *projectors, int nproj, const int *momenta, int nmom,
                                         const contract_sublattice *lattice, void *workspace, size_t workspace_bytes,
                                         void *stream)
{
  if (!lattice) {
    set_last_error("lattice is NULL");
    return CONTRACT_ERR_ARG;
  }
  contract_baryon_request req;
  req.propag_i = propag_i;
  req.propag_j = propag_j;
  req.propag_n = propag_n;
  req.out = out;
  req.lattice = *lattice;
  return contract_baryon_twopoint_project_batch(&req, 1, precision, contract_types, weights, nterm, gamma_ij,
                                                gamma_kl, projectors, nproj, momenta, nmom, workspace,
                                                workspace_bytes, stream);
}

size_t contract_baryon_batch_workspace_bytes(const contract_baryon_request *reqs, int nreq, int nmom)
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

int contract_baryon_twopoint_project_batch(const contract_baryon_request *reqs, int nreq, int precision,
                                           const int *contract_types, const double _Complex *weights, int nterm,
                                           int gamma_ij, int gamma_kl, const double _Complex *projectors, int nproj,
                                           const int *momenta, int nmom, void *workspace, size_t workspace_bytes,
                                           void *stream_)
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
    if (!reqs[r].out || !reqs[r].propag_i || !reqs[r].propag_j || !reqs[r].propag_n) {
      set_last_error("request %d: NULL output or propagator pointer", r);
      return CONTRACT_ERR_ARG;
    }
  }
  if ((rc = check_momenta(momenta, nmom)) != CONTRACT_OK) return rc;
  if (precision != CONTRACT_FP64 && precision != CONTRACT_FP32 && precision != CONTRACT_FP32_SCALED) {
    set_last_error("precision %d is not CONTRACT_FP64, CONTRACT_FP32 or CONTRACT_FP32_SCALED", precision);
    return CONTRACT_ERR_ARG;
  }
  if (nterm < 1 || nterm > MAX_TERMS || !contract_types || !weights) {
    set_last_error("nterm = %d must be in 1..%d", nterm, MAX_TERMS);
    return CONTRACT_ERR_ARG;
  }
  for (int t = 0; t < nterm; ++t)
    if (contract_types[t] < 0 || contract_types[t] > 5) {
      set_last_error("contract_types[%d] = %d is not a BaryonContractType", t, contract_types[t]);
      return CONTRACT_ERR_ARG;
    }
  if (gamma_ij < 0 || gamma_ij > 15 || gamma_kl < 0 || gamma_kl > 15) {
    set_last_error("gamma_ij = %d / gamma_kl = %d out of range 0..15", gamma_ij, gamma_kl);
    return CONTRACT_ERR_ARG;
  }
  if (nproj < 1 || !projectors) {
    set_last_error("nproj = %d must be >= 1 with a non-NULL projectors array", nproj);
    return CONTRACT_ERR_ARG;
  }

  // per request: its work decomposition and workspace slice (partials, then
  // per-timeslice sums); +k/-k momenta share a slot
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

  BaryonProjectArgs a;
  std::memset(&a, 0, sizeof(a));
  a.gamma_ij = gamma_ij;
  a.gij = sparse_of(gamma_matrix(gamma_ij));
  a.gkl = sparse_of(gamma_matrix(gamma_kl));
  a.gklT = sparse_of(mat_transpose(gamma_matrix(gamma_kl)));
  a.nterm = nterm;
  const double *wt = reinterpret_cast<const double *>(weights);
  for (int t = 0; t < nterm; ++t) {
    a.types[t] = contract_types[t];
    a.w_re[t] = wt[2 * t];
    a.w_im[t] = wt[2 * t + 1];
  }

  std::vector<ProjectReq> preq(nreq);
  for (int r = 0; r < nreq; ++r) {
    ProjectReq &R = preq[r];
    std::memset(&R, 0, sizeof(R));
    R.propag_i = reqs[r].propag_i;
    R.propag_j = reqs[r].propag_j;
    R.propag_n = reqs[r].propag_n;
    if (precision == CONTRACT_FP32_SCALED) { // the site exponents follow the complex64 entries
      const size_t V = (size_t)geo[r].dims[0] * geo[r].dims[1] * geo[r].dims[2] * geo[r].dims[3];
      R.exp_i = static_cast<const signed char *>(R.propag_i) + V * SCALED_SITE_BYTES;
      R.exp_j = static_cast<const signed char *>(R.propag_j) + V * SCALED_SITE_BYTES;
      R.exp_n = static_cast<const signed char *>(R.propag_n) + V * SCALED_SITE_BYTES;
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
    pa.share_n = 1;
    for (int r = 0; r < pa.nreq; ++r) {
      pa.req[r] = preq[r0 + r];
      pa.req[r].item0 = (unsigned)items;
      items += plan[r0 + r].items;
      pa.share_n &= pa.req[r].propag_n == pa.req[r].propag_i;
    }
    pa.total_items = (unsigned)items;
    if (precision == CONTRACT_FP64)
      launch_baryon_project(a, pa, stream);
    else if (precision == CONTRACT_FP32)
      launch_baryon_project_fp32(a, pa, stream);
    else
      launch_baryon_project_fp32_scaled(a, pa, stream);
    CONTRACT_EXT_HIP(hipGetLastError());
  }
  CONTRACT_EXT_HIP(timer_end(stream));

  std::vector<zhost> W((size_t)nproj * 16);
  const double *pm = reinterpret_cast<const double *>(projectors);
  for (size_t e = 0; e < W.size(); ++e) W[e] = zhost(pm[2 * e], pm[2 * e + 1]);
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
    launch_finish_all(fr, n, nslot, slots, nmom, W.data(), nproj, stream);
  }
  CONTRACT_EXT_HIP(hipGetLastError());
  if (own) CONTRACT_EXT_HIP(hipFreeAsync(ws, stream));
  return CONTRACT_OK;
}

size_t contract_scaled_fp32_bytes(size_t volume) { return volume * (SCALED_SITE_BYTES + 1); }

int contract_propagator_to_scaled_fp32(void *dst, const void *src, size_t volume, void *stream)
{
  if (!dst || !src || volume == 0) {
    set_last_error("NULL propagator or zero volume");
    return CONTRACT_ERR_ARG;
  }
  ScaleArgs a;
  a.src = src;
  a.dst = dst;
  a.exps = static_cast<char *>(dst) + volume * SCALED_SITE_BYTES;
  a.volume = volume;
  launch_scale_fp32(a, static_cast<hipStream_t>(stream));
  CONTRACT_EXT_HIP(hipGetLastError());
  return CONTRACT_OK;
}

} // extern "C"
