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
    // enough work items to fill the chip several times over
    const size_t target_items = (size_t)device_cu_count() * 16;
    size_t cps = (target_items + p.slices - 1) / p.slices;
    if (cps < 1) cps = 1;
    chunk = (p.slice_sites + cps - 1) / cps;
  }
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

void fill_project_args(ProjectArgs &pa, const ProjectPlan &plan, const Geometry &geo, const int *momenta,
                       const std::vector<MomSlot> &slots, void *ws)
{
  const int nmom = (int)slots.size();
  std::memset(&pa, 0, sizeof(pa));
  for (int mu = 0; mu < 4; ++mu) pa.dims[mu] = geo.dims[mu];
  for (int mu = 0; mu < 3; ++mu) {
    pa.gdims[mu] = geo.global[mu];
    pa.origin[mu] = geo.origin[mu];
  }
  const int order = geo.order;
  pa.order = order;
  pa.np = nmom;
  for (int p = 0; p < nmom; ++p)
    for (int mu = 0; mu < 3; ++mu) pa.mom[p][mu] = momenta[3 * slots[p].rep + mu];
  pa.chunk = plan.chunk;
  pa.chunks_per_slice = plan.chunks_per_slice;
  pa.slice_sites = plan.slice_sites;
  pa.partial = reinterpret_cast<double2 *>(ws);
}

// Per-timeslice reduction of the work-item partials, then the projectors
// (MAX_PROJ_PER_LAUNCH per launch), expanding each momentum slot to its +k
// and -k outputs.
void launch_finish_all(const ProjectArgs &pa, const ProjectPlan &plan, const std::vector<MomSlot> &slots, int nmom,
                       const zhost *W, int nproj, void *out, void *ws, hipStream_t stream)
{
  ReduceArgs r;
  r.partial = pa.partial;
  r.open = reinterpret_cast<double2 *>(static_cast<char *>(ws) + plan.partial_bytes);
  r.np = pa.np;
  r.lt = pa.dims[3];
  r.nq = 32;
  r.order = pa.order;
  r.chunks_per_slice = plan.chunks_per_slice;
  launch_project_reduce(r, stream);
  for (int base = 0; base < nproj; base += MAX_PROJ_PER_LAUNCH) {
    ApplyArgs f;
    std::memset(&f, 0, sizeof(f));
    f.open = r.open;
    f.nslot = pa.np;
    f.lt = pa.dims[3];
    f.nmom = nmom;
    for (int k = 0; k < (int)slots.size(); ++k) {
      f.slot[slots[k].plus] = k;
      f.sgn[slots[k].plus] = 1;
      if (slots[k].minus >= 0) {
        f.slot[slots[k].minus] = k;
        f.sgn[slots[k].minus] = -1;
      }
    }
    f.nproj = nproj - base < MAX_PROJ_PER_LAUNCH ? nproj - base : MAX_PROJ_PER_LAUNCH;
    f.proj_base = base;
    for (int w = 0; w < f.nproj; ++w)
      for (int q = 0; q < 16; ++q) {
        f.w_re[w][q] = W[(base + w) * 16 + q].real();
        f.w_im[w][q] = W[(base + w) * 16 + q].imag();
      }
    f.out = static_cast<double2 *>(out);
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
                                         void *stream_)
{
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  int rc;
  if (!lattice) {
    set_last_error("lattice is NULL");
    return CONTRACT_ERR_ARG;
  }
  Geometry geo;
  if ((rc = resolve_geometry(geo, lattice->dims, lattice->site_order, lattice->global, lattice->origin)) !=
      CONTRACT_OK)
    return rc;
  const int *dims = geo.dims;
  const int site_order = geo.order;
  if ((rc = check_momenta(momenta, nmom)) != CONTRACT_OK) return rc;
  if (!out || !propag_i || !propag_j || !propag_n) {
    set_last_error("NULL output or propagator pointer");
    return CONTRACT_ERR_ARG;
  }
  if (precision != CONTRACT_FP64 && precision != CONTRACT_FP32) {
    set_last_error("precision %d is not CONTRACT_FP64 or CONTRACT_FP32", precision);
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

  const std::vector<MomSlot> slots = pair_momenta(momenta, nmom);
  const ProjectPlan plan = plan_project(dims, site_order, (int)slots.size(), 0);
  void *ws = workspace;
  bool own = false;
  if (ws) {
    if (workspace_bytes < plan.total) {
      set_last_error("workspace of %zu bytes is smaller than the %zu needed", workspace_bytes, plan.total);
      return CONTRACT_ERR_WORKSPACE;
    }
  } else {
    CONTRACT_EXT_HIP(hipMallocAsync(&ws, plan.total, stream));
    own = true;
  }

  ProjectArgs pa;
  fill_project_args(pa, plan, geo, momenta, slots, ws);

  BaryonProjectArgs a;
  std::memset(&a, 0, sizeof(a));
  a.propag_i = propag_i;
  a.propag_j = propag_j;
  a.propag_n = propag_n;
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
  CONTRACT_EXT_HIP(timer_begin(stream));
  if (precision == CONTRACT_FP64)
    launch_baryon_project(a, pa, plan.items, stream);
  else
    launch_baryon_project_fp32(a, pa, plan.items, stream);
  CONTRACT_EXT_HIP(hipGetLastError());
  CONTRACT_EXT_HIP(timer_end(stream));

  std::vector<zhost> W((size_t)nproj * 16);
  const double *pm = reinterpret_cast<const double *>(projectors);
  for (size_t e = 0; e < W.size(); ++e) W[e] = zhost(pm[2 * e], pm[2 * e + 1]);
  launch_finish_all(pa, plan, slots, nmom, W.data(), nproj, out, ws, stream);
  CONTRACT_EXT_HIP(hipGetLastError());
  if (own) CONTRACT_EXT_HIP(hipFreeAsync(ws, stream));
  return CONTRACT_OK;
}

} // extern "C"
