// plan.hpp -- work decomposition of the fused (timeslice x momentum) kernels.
#pragma once

#include <stddef.h>

#include <vector>

#include "baryon.hpp"
#include "spin.hpp"

namespace contract_amd
{

// A work item is a contiguous run of `chunk` sites inside one slice (a
// timeslice, or a timeslice of one parity for even-odd order), so every item
// contributes to exactly one timeslice.  Workspace = one [nmom][16] complex
// partial per item + the [Lt][nmom][16] per-timeslice sums.
struct ProjectPlan {
  int slices;
  size_t slice_sites;
  int chunk;
  int chunks_per_slice;
  size_t items;
  size_t partial_bytes; // [items][nslot][2][16] complex
  size_t open_bytes;    // [Lt][nslot][2][16] complex, the per-timeslice sums
  size_t total;
};

// A momentum k and its negative share one slot: the projection keeps
// C = sum cos(k.x) c(x) and S = -sum sin(k.x) c(x), and P(+-k) = C +- iS.
// One slot per distinct pair (exact integer negation; the zero momentum and
// duplicates get their own), with the output index of +k and -k (-1 if not
// requested).
struct MomSlot {
  int rep, plus, minus;
};
std::vector<MomSlot> pair_momenta(const int *mom, int nmom);

// sites_per_item <= 0 picks ~8 items per compute unit; nslot momentum slots.
ProjectPlan plan_project(const int dims[4], int order, int nmom, int sites_per_item);
int check_lattice(const int dims[4], int order);
int check_momenta(const int *momenta, int nmom);
// Resolved sub-lattice geometry (global = dims, origin = 0 for a whole lattice).
struct Geometry {
  int dims[4], global[3], origin[3], order;
};
int resolve_geometry(Geometry &g, const int dims[4], int order, const int *global, const int *origin);
// Per-timeslice sums of each request's work-item partials into its `open`,
// then the projectors, each momentum slot expanded to its +k and -k outputs
// (nreq <= MAX_BATCH requests per call).
struct FinishReq {
  const double2 *partial;
  double2 *open;
  void *out;
  int lt, order, chunks_per_slice;
};
void launch_finish_all(const FinishReq *reqs, int nreq, int nslot, const std::vector<MomSlot> &slots, int nmom,
                       const zhost *W, int nproj, hipStream_t stream);

} // namespace contract_amd
