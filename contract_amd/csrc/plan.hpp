// plan.hpp -- work decomposition of the fused (timeslice x momentum) kernels.
#pragma once

#include <stddef.h>

#include "baryon.hpp"

namespace contract_amd
{

// A work item is a contiguous run of `chunk` sites inside one slice (a
// timeslice, or a timeslice of one parity for even-odd order), so every item
// contributes to exactly one timeslice.  Workspace = phase table + one
// [nmom][16] complex partial per item.
struct ProjectPlan {
  int slices;
  size_t slice_sites;
  int chunk;
  int chunks_per_slice;
  size_t items;
  size_t tab_bytes;
  size_t partial_bytes;
  size_t total;
};

// sites_per_item <= 0 picks ~16 items per compute unit.
ProjectPlan plan_project(const int dims[4], int order, int nmom, int sites_per_item);
int check_lattice(const int dims[4], int order);
int check_momenta(const int *momenta, int nmom);
void fill_project_args(ProjectArgs &pa, const ProjectPlan &plan, const int dims[4], int order, const int *momenta,
                       int nmom, void *ws);

} // namespace contract_amd
