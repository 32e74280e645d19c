// noalloc.hip -- the fused baryon two-point for callers that cannot allocate
// (contract_baryon_twopoint_project_noalloc, include/contract_ext.h).
//
// The fused entry points need a workspace for the work items' partial sums
// and the per-timeslice sums.  Here it is a fixed device array of this file,
// `scratch`, resolved once per device (at init(), else on the device's first
// call) to its device address with hipGetSymbolAddress; only that address is
// ever passed to kernels.  Every call on one GPU shares that GPU's array:
// calls that overlap on one GPU (different streams, or host threads) race.
// Nothing in it has to start at zero: every work item writes its whole
// partial and the reduction writes every per-timeslice sum before they are
// read.

#include <contract_ext.h>

#include <mutex>

#include "runtime.hpp"

namespace contract_amd
{

constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20; // per GPU, held until the process ends
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[NOALLOC_MAX_DEVICES];

// Device address of this device's `scratch` (resolved with the device
// current, once), or an error.
static int scratch_of_current_device(void **out)
{
  int dev = 0;
  CONTRACT_EXT_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
    set_last_error("device %d is beyond the %d devices the scratch table holds", dev, NOALLOC_MAX_DEVICES);
    return CONTRACT_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    CONTRACT_EXT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return CONTRACT_OK;
}

// init()'s set-up step: resolve the current device's scratch address.
int noalloc_setup() { void *p; return scratch_of_current_device(&p); }

} // namespace contract_amd

using namespace contract_amd;

extern "C" {

size_t contract_noalloc_scratch_bytes(void) { return NOALLOC_SCRATCH_BYTES; }

int contract_baryon_twopoint_project_noalloc(void *out, const void *propag_i, const void *propag_j,
                                             const void *propag_n, int precision, const int *contract_types,
                                             const double _Complex *weights, int nterm, int gamma_ij, int gamma_kl,
                                             const double _Complex *projectors, int nproj, const int *momenta,
                                             int nmom, const contract_sublattice *lattice, void *stream)
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
  const size_t need = contract_baryon_batch_workspace_bytes(&req, 1, nmom);
  if (need > NOALLOC_SCRATCH_BYTES) {
    set_last_error("this call needs %zu bytes of scratch, more than the %zu the no-allocation entry point holds "
                   "(use contract_baryon_twopoint_project with a workspace)",
                   need, NOALLOC_SCRATCH_BYTES);
    return CONTRACT_ERR_WORKSPACE;
  }
  void *ws = nullptr;
  const int rc = scratch_of_current_device(&ws);
  if (rc != CONTRACT_OK) return rc;
  // need == 0: the arguments are invalid, and the batch entry point names the
  // error; with a workspace given it never allocates (a valid call would fail
  // the size check instead)
  return contract_baryon_twopoint_project_batch(&req, 1, precision, contract_types, weights, nterm, gamma_ij,
                                                gamma_kl, projectors, nproj, momenta, nmom, ws,
                                                need == 0 ? 0 : NOALLOC_SCRATCH_BYTES, stream);
}

} // extern "C"
