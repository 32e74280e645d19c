// meson_grid.hpp -- batched meson Gamma x Gamma' grids projected to
// timeslices and momenta (many requests, uneven lattices, one launch).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include "device.hpp"

namespace contract_amd
{

constexpr int GRID_THREADS = 256;       // four waves share one staged tile
constexpr int GRID_MAX_REQ = 32;        // requests per launch
constexpr int GRID_MOM_PER_TILE = 8;    // 16 MFMA rows = 8 momentum slots x (cos, -sin)
constexpr int GRID_MAX_TILES = 2;       // row tiles per pass
constexpr int GRID_MOM_PER_PASS = GRID_MOM_PER_TILE * GRID_MAX_TILES;
constexpr int GRID_COLS = 256;          // K[i,k,j,l] columns, (k*4 + l)*16 + i*4 + j
// doubles per work item and row tile: [row][K re/im][col]; a work item's
// partial holds `partial_tiles` of them, rows 2p (cos) and 2p + 1 (-sin) of
// momentum slot p running on across the tiles
constexpr int GRID_PARTIAL_TILE = 16 * 2 * GRID_COLS;

struct MesonGridReq {
  const void *propag_i, *propag_j;
  const signed char *exp_i, *exp_j; // site exponents (FP32 site-scaled storage only)
  int dims[4];
  int gdims[3], origin[3]; // phases exp(-2 pi i p (x + origin) / gdims)
  int order;            // 0 lexicographic, 1 even-odd
  int chunks_per_slice; // work items per (parity, t) slice
  unsigned slice_sites;
  unsigned item0;       // first work item of this request in the batch
};

struct MesonGridArgs {
  MesonGridReq req[GRID_MAX_REQ];
  int nreq;
  int np;    // momentum slots in this pass (<= GRID_MOM_PER_PASS)
  int chunk; // sites per work item, a multiple of the tile
  int mom[GRID_MOM_PER_PASS][3];
  unsigned total_items;
  int partial_tiles; // row tiles per work item in `partial`, >= the pass's row tiles
  double *partial;   // [total_items][partial_tiles * GRID_PARTIAL_TILE]
};

struct MesonGridOut {
  void *out; // complex128 [16][16][np_total][Lt]
  int lt, order, chunks_per_slice;
  unsigned item0;
  int block0; // first finish block (one per timeslice)
};

struct MesonGridFinish {
  MesonGridOut req[GRID_MAX_REQ];
  int nreq, np, np_total;
  int partial_tiles; // as in MesonGridArgs
  // slot p of the pass holds momentum k: output index of +k and of -k (the
  // rows give both, P(+-k) = A +- iB), -1 where that sign was not requested
  int out_plus[GRID_MOM_PER_PASS], out_minus[GRID_MOM_PER_PASS];
  const double *partial;
  SparseSpin g5t[16]; // (g5 G(g))^T: row i -> (j, (g5 G(g))[j][i])
};

// unit: 1 vector-ALU traces, 2 matrix-core traces (four waves share a staged
// tile), 3 matrix-core traces with a wave-local K (one-wave workgroups;
// passes of at most 4 momentum slots, others run unit 2); tile_sites: sites
// staged per step (4 or 8, units 1 and 2); blocks: persistent grid size
// (workgroups); tab_entries: np x max(Lx + Ly + Lz + 3), the per-momentum
// phase tables in LDS.
enum GridStorage { GRID_FP64 = 0, GRID_FP32 = 1, GRID_FP32_SCALED = 2 };
// whether launch_meson_grid runs a pass of np momentum slots in the
// wave-local form (one-wave workgroups)
inline bool grid_wave_pass(int unit, int np) { return unit == 3 && np <= GRID_MOM_PER_TILE / 2; }
void launch_meson_grid(const MesonGridArgs &a, int storage, int unit, int tile_sites, size_t blocks,
                       size_t tab_entries, hipStream_t stream);
// Workgroups of the kernel launch_meson_grid would run for this pass that fit
// on one CU at once (registers, LDS with the phase tables), 0 if the runtime
// cannot tell.
int meson_grid_resident_blocks(const MesonGridArgs &a, int storage, int unit, int tile_sites, size_t tab_entries);
void launch_meson_grid_finish(const MesonGridFinish &f, int nblocks, hipStream_t stream);

} // namespace contract_amd
