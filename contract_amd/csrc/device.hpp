// device.hpp -- complex arithmetic, lane collectives and HBM->LDS staging for
// the gfx950 contraction kernels.
//
// Work mapping shared by every per-site kernel: a site is handled by 16 lanes
// of a wave64 (lane q = lane & 15 indexes a 4x4 spin pair), so one wave covers
// SITES_PER_WAVE = 4 sites.  The propagators of those sites are staged into
// LDS with fully coalesced 16-byte loads (a wave reads 1 KiB contiguous per
// instruction), then each lane reads the 3x3 colour blocks it needs into
// registers and works from there.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace contract_amd
{

constexpr int NS = 4;
constexpr int NC = 3;
constexpr int SITE_ELEMS = 144; // [4][4][3][3] complex entries per site
constexpr int LANES_PER_SITE = 16;
constexpr int WAVE = 64;
constexpr int SITES_PER_WAVE = WAVE / LANES_PER_SITE;

// ---------------------------------------------------------------------------
// complex helpers (double2 / float2 hold (re, im))
// ---------------------------------------------------------------------------
template <typename R> struct cplx_t;
template <> struct cplx_t<double> {
  using type = double2;
};
template <> struct cplx_t<float> {
  using type = float2;
};
template <typename R> using cplx = typename cplx_t<R>::type;

template <typename C> __device__ __forceinline__ C cmake(decltype(C::x) re, decltype(C::x) im)
{
  C z;
  z.x = re;
  z.y = im;
  return z;
}
template <typename C> __device__ __forceinline__ C czero() { return cmake<C>(0, 0); }
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return cmake<C>(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return cmake<C>(a.x - b.x, a.y - b.y); }
// The fused form is spelled out: left to the compiler, `a.x * b.x - a.y * b.y`
// may contract into either FMA, and the choice can differ between kernels
// that compute the same value (the bit-identity of kernel variants relies on
// one rounding).
template <typename C> __device__ __forceinline__ C cmul(C a, C b)
{
  return cmake<C>(fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x));
}
// acc + a * b
template <typename C> __device__ __forceinline__ C cfma(C acc, C a, C b)
{
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
  return acc;
}
// acc - a * b
template <typename C> __device__ __forceinline__ C cfms(C acc, C a, C b)
{
  acc.x = fma(-a.x, b.x, acc.x);
  acc.x = fma(a.y, b.y, acc.x);
  acc.y = fma(-a.x, b.y, acc.y);
  acc.y = fma(-a.y, b.x, acc.y);
  return acc;
}
// acc + a * conj(b)
template <typename C> __device__ __forceinline__ C cfma_conj(C acc, C a, C b)
{
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(a.y, b.y, acc.x);
  acc.y = fma(a.y, b.x, acc.y);
  acc.y = fma(-a.x, b.y, acc.y);
  return acc;
}
// acc + i^R * a * b for a compile-time unit i^R (R = 0..3): the unit only
// moves and negates FMA operands
template <int R, typename C> __device__ __forceinline__ C cfma_rot(C acc, C a, C b)
{
  constexpr int r = R & 3;
  if constexpr (r == 0) return cfma(acc, a, b);
  if constexpr (r == 2) return cfms(acc, a, b);
  if constexpr (r == 1) { // acc.x -= Im(ab), acc.y += Re(ab)
    acc.x = fma(-a.x, b.y, acc.x);
    acc.x = fma(-a.y, b.x, acc.x);
    acc.y = fma(a.x, b.x, acc.y);
    acc.y = fma(-a.y, b.y, acc.y);
  } else { // r == 3: acc.x += Im(ab), acc.y -= Re(ab)
    acc.x = fma(a.x, b.y, acc.x);
    acc.x = fma(a.y, b.x, acc.x);
    acc.y = fma(-a.x, b.x, acc.y);
    acc.y = fma(a.y, b.y, acc.y);
  }
  return acc;
}
template <typename C> __device__ __forceinline__ C cconj(C a) { return cmake<C>(a.x, -a.y); }
template <typename C> __device__ __forceinline__ C cscale(C a, decltype(C::x) s) { return cmake<C>(a.x * s, a.y * s); }

// ---------------------------------------------------------------------------
// lane collectives inside a 16-lane site group
// ---------------------------------------------------------------------------
__device__ __forceinline__ double shfl_xor_d(double v, int mask) { return __shfl_xor(v, mask, 64); }
__device__ __forceinline__ float shfl_xor_d(float v, int mask) { return __shfl_xor(v, mask, 64); }

// Lane exchanges inside a 16-lane row on the vector ALUs (DPP moves) rather
// than through the LDS crossbar (ds_bpermute, what __shfl_xor compiles to):
// the reductions below are chains of dependent exchanges, and a DPP move
// costs a VALU issue where a permute costs an LDS round trip.  gfx9 DPP
// controls: quad_perm [1,0,3,2] = 0xb1 is xor 1, [2,3,0,1] = 0x4e is xor 2;
// row_ror:n = 0x120 + n, where lane i reads lane (i - n) mod 16 of its row,
// so row_ror:8 is xor 8, and xor 4 is row_ror:12 on the lanes with bit 2
// clear and row_ror:4 on the others.
template <int CTRL> __device__ __forceinline__ int dpp_i32(int v)
{
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false);
}
template <int CTRL> __device__ __forceinline__ float dpp_r(float v)
{
  return __builtin_bit_cast(float, dpp_i32<CTRL>(__builtin_bit_cast(int, v)));
}
template <int CTRL> __device__ __forceinline__ double dpp_r(double v)
{
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)dpp_i32<CTRL>((int)(unsigned)b);
  const unsigned hi = (unsigned)dpp_i32<CTRL>((int)(unsigned)(b >> 32));
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
// the value of lane q ^ M of this lane's 16-lane row (q = lane & 15)
template <int M, typename R> __device__ __forceinline__ R lane_xor(R v, int q)
{
  static_assert(M == 1 || M == 2 || M == 4 || M == 8, "an exchange inside a 16-lane row");
  if constexpr (M == 1) return dpp_r<0xb1>(v);
  else if constexpr (M == 2) return dpp_r<0x4e>(v);
  else if constexpr (M == 8) return dpp_r<0x128>(v);
  else {
    const R up = dpp_r<0x12c>(v), down = dpp_r<0x124>(v);
    return (q & 4) ? down : up;
  }
}

template <int M, typename C> __device__ __forceinline__ C shfl_xor_c(C v)
{
  static_assert(M != 4, "xor 4 depends on the lane: shfl_xor4_c");
  return cmake<C>(lane_xor<M>(v.x, 0), lane_xor<M>(v.y, 0));
}
template <typename C> __device__ __forceinline__ C shfl_xor4_c(C v, int q)
{
  return cmake<C>(lane_xor<4>(v.x, q), lane_xor<4>(v.y, q));
}

// Sum over the 16 lanes of a site group; every lane of the group gets the
// same bits.  The xor-4 step is a rotation by 4: after the xor-8 step lanes
// i and i + 8 hold equal values, so every lane adds the same two values as
// the xor butterfly would.
template <typename C> __device__ __forceinline__ C group16_allreduce(C v)
{
  v = cadd(v, shfl_xor_c<8>(v));
  v = cadd(v, cmake<C>(dpp_r<0x124>(v.x), dpp_r<0x124>(v.y)));
  v = cadd(v, shfl_xor_c<2>(v));
  return cadd(v, shfl_xor_c<1>(v));
}

// Reduce-scatter over a 16-lane group: every lane holds v[0..15]; on return
// lane q holds sum over the group of v[q].  Fifteen exchanges instead of the
// 4 x 16 of sixteen separate all-reduces.
template <typename C> __device__ __forceinline__ C group16_reduce_scatter(C (&v)[16], int q)
{
  C a8[8];
  {
    const bool hi = q & 8;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      C keep = hi ? v[k + 8] : v[k];
      C send = hi ? v[k] : v[k + 8];
      a8[k] = cadd(keep, shfl_xor_c<8>(send));
    }
  }
  C a4[4];
  {
    const bool hi = q & 4;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      C keep = hi ? a8[k + 4] : a8[k];
      C send = hi ? a8[k] : a8[k + 4];
      a4[k] = cadd(keep, shfl_xor4_c(send, q));
    }
  }
  C a2[2];
  {
    const bool hi = q & 2;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      C keep = hi ? a4[k + 2] : a4[k];
      C send = hi ? a4[k] : a4[k + 2];
      a2[k] = cadd(keep, shfl_xor_c<2>(send));
    }
  }
  const bool hi = q & 1;
  C keep = hi ? a2[1] : a2[0];
  C send = hi ? a2[0] : a2[1];
  return cadd(keep, shfl_xor_c<1>(send));
}

// ---------------------------------------------------------------------------
// HBM -> registers -> LDS staging of one tile (SITES_PER_WAVE consecutive
// sites) of one propagator.  The ABI layout is site-major [V][144] complex,
// so a tile is one contiguous run of 4 x 144 complex; it moves as 16-byte
// chunks, chunk e = it * 64 + lane, so every load instruction reads 1 KiB of
// consecutive HBM.  Loading into registers first lets a wave keep the next
// tile in flight while it computes on the current one from LDS.  In LDS each
// site is padded by one 16-byte chunk so the four sites of a tile start on
// different banks.  Sites past `nsites` (ragged tail) are zero-filled.
// ---------------------------------------------------------------------------
// Tile loads go through a buffer resource whose range ends at the tile's last
// valid byte: the hardware returns zero past it, so a ragged tile needs no lane
// mask, bounds select or zero fill (a masked or selected load on one side of a
// branch makes the compiler wait for it where the paths join).  The streams
// are read once: the loads are non-temporal, which keeps them out of the
// caches' reuse set (+1-5% over plain loads).
constexpr int BUFFER_RSRC_WORD3 = 0x00020000; // gfx9 raw buffer: 32-bit data format
constexpr int BUFFER_AUX_NT = 2;              // cache policy: non-temporal

__device__ __forceinline__ __amdgpu_buffer_rsrc_t tile_rsrc(const void *base, unsigned bytes)
{
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, (int)bytes, BUFFER_RSRC_WORD3);
}

__device__ __forceinline__ uint4 stream_load(__amdgpu_buffer_rsrc_t rsrc, int byte_offset)
{
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  const u32x4 w = __builtin_amdgcn_raw_buffer_load_b128(rsrc, byte_offset, 0, BUFFER_AUX_NT);
  return make_uint4(w.x, w.y, w.z, w.w);
}

// Plain non-temporal load of one 16-byte chunk.
__device__ __forceinline__ uint4 stream_load(const uint4 *p)
{
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  const u32x4 w = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(p));
  return make_uint4(w.x, w.y, w.z, w.w);
}

// Propagator-shaped outputs are written once: non-temporal stores (+5% for
// the diquark and sequential-source kernels).
__device__ __forceinline__ void stream_store(uint4 *p, uint4 v)
{
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  __builtin_nontemporal_store(u32x4{v.x, v.y, v.z, v.w}, reinterpret_cast<u32x4 *>(p));
}

template <typename C, int NSITE = SITES_PER_WAVE> struct Tile {
  static constexpr int SITE16 = SITE_ELEMS * (int)sizeof(C) / 16;     // 144 (FP64) / 72 (FP32)
  static constexpr int LDS16 = SITE16 + 1;                             // padded site stride
  static constexpr int LDS_ELEMS = LDS16 * 16 / (int)sizeof(C);       // per site, in complex
  static constexpr int CHUNKS = NSITE * SITE16;                       // per tile
  static constexpr int PER_LANE = (CHUNKS + WAVE - 1) / WAVE;         // 4 sites: 9 / 5; 2 sites: 5 / 3
  static constexpr int DMA_ELEMS = PER_LANE * WAVE * 16 / (int)sizeof(C); // a dma_tile destination, in complex
  uint4 v[PER_LANE];

  // Chunks past nsites (ragged tile) and lanes past the tile read as zero;
  // nsites = 0 reads an all-zero tile without touching memory.
  __device__ __forceinline__ void load(const C *__restrict__ src, size_t first_site, int nsites, int lane)
  {
    const __amdgpu_buffer_rsrc_t rsrc = tile_rsrc(src + first_site * SITE_ELEMS, (unsigned)(nsites * SITE16 * 16));
#pragma unroll
    for (int it = 0; it < PER_LANE; ++it) v[it] = stream_load(rsrc, (it * WAVE + lane) * 16);
  }

  __device__ __forceinline__ void store(C *__restrict__ lds, int lane) const
  {
    uint4 *l = reinterpret_cast<uint4 *>(lds);
#pragma unroll
    for (int it = 0; it < PER_LANE; ++it) {
      const int e = it * WAVE + lane;
      if (CHUNKS % WAVE == 0 || it + 1 < PER_LANE || e < CHUNKS) { // lanes past the tile hold no chunk
        const int s = e / SITE16;
        l[s * LDS16 + (e - s * SITE16)] = v[it];
      }
    }
  }
};

// HBM -> LDS without registers (gfx950 `buffer_load_dwordx4 ... lds`): lane l
// of instruction it writes chunk e = it * 64 + l to the 16-byte LDS slot e, so
// the tile lands unpadded (site stride SITE16 chunks).  The destination holds
// Tile::DMA_ELEMS (PER_LANE * 64 slots: whole instructions, an FP32 tile's
// last one is half past the tile).  Range-limited like Tile::load (the
// instruction offset goes in the scalar offset, which the range check
// includes): chunks past `nsites` are not read and land as zeros (both
// measured with an in-bounds probe).  Completion is tracked by vmcnt; see
// dma_wait.
template <typename C>
__device__ __forceinline__ void dma_tile(C *__restrict__ lds, const C *__restrict__ src, size_t first_site,
                                         int nsites, int lane)
{
  using T = Tile<C>;
  typedef __attribute__((address_space(3))) void lds_void;
  const __amdgpu_buffer_rsrc_t rsrc = tile_rsrc(src + first_site * SITE_ELEMS, (unsigned)(nsites * T::SITE16 * 16));
  uint4 *l16 = reinterpret_cast<uint4 *>(lds);
#pragma unroll
  for (int it = 0; it < T::PER_LANE; ++it)
    __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, (lds_void *)(l16 + it * WAVE), 16, lane * 16, it * WAVE * 16, 0,
                                             BUFFER_AUX_NT); // lane offset in a VGPR, instruction offset in an SGPR
}

// dma_tile of the same sites of two propagators, its instructions spread
// over the NW waves of a workgroup (wave w issues instructions w, w + NW, ...
// of each propagator).  Each wave waits for its own share (dma_wait); a
// workgroup barrier after that makes the whole tile visible.  Every
// instruction names its destination array directly, so the compiler can
// tell reads of another array from these loads.
template <int NW, int NSITE, typename C>
__device__ __forceinline__ void dma_tile_pair_split(C *__restrict__ lds_i, C *__restrict__ lds_j,
                                                    const C *__restrict__ src_i, const C *__restrict__ src_j,
                                                    size_t first_site, int nsites, int lane, int wave)
{
  using T = Tile<C, NSITE>;
  typedef __attribute__((address_space(3))) void lds_void;
  const unsigned bytes = (unsigned)(nsites * T::SITE16 * 16);
  const __amdgpu_buffer_rsrc_t ri = tile_rsrc(src_i + first_site * SITE_ELEMS, bytes);
  const __amdgpu_buffer_rsrc_t rj = tile_rsrc(src_j + first_site * SITE_ELEMS, bytes);
  uint4 *li = reinterpret_cast<uint4 *>(lds_i), *lj = reinterpret_cast<uint4 *>(lds_j);
#pragma unroll
  for (int m = 0; m < (T::PER_LANE + NW - 1) / NW; ++m) {
    const int it = wave + m * NW;
    if (it < T::PER_LANE) {
      __builtin_amdgcn_raw_ptr_buffer_load_lds(ri, (lds_void *)(li + it * WAVE), 16, lane * 16, it * WAVE * 16, 0,
                                               BUFFER_AUX_NT);
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rj, (lds_void *)(lj + it * WAVE), 16, lane * 16, it * WAVE * 16, 0,
                                               BUFFER_AUX_NT);
    }
  }
}

// Orders the LDS traffic between the lanes of a one-wave workgroup (a wave's
// LDS instructions execute in issue order, so finishing the earlier ones is
// enough).  __syncthreads() is a workgroup release fence, which also waits for
// every HBM -> LDS load in flight (vmcnt(0)); a kernel that keeps dma_tile
// loads in flight across such a point uses this instead, and waits for a
// dma_tile itself (dma_wait, dma_wait_keep) before the first read of its
// destination.
__device__ __forceinline__ void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_s_waitcnt(0xc07f); // lgkmcnt(0): the wave's earlier LDS reads and writes are done
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Waits for every HBM -> LDS load (dma_tile) of the wave, and orders the
// LDS reads that follow after it.  The compiler sometimes inserts this wait
// before the first read of a dma_tile destination, but not reliably (the
// meson grid's wave-local kernel read its tiles with no wait, and so did the
// staged baryon kernels once the module changed), so every reader states
// it.  Out-of-range lanes of a dma_tile write zeros (measured).
__device__ __forceinline__ void dma_wait()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_s_waitcnt(0x0f70); // vmcnt(0)
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// dma_wait that leaves the KEEP most recently issued vector-memory
// operations in flight (vmcnt(KEEP); loads count down in issue order): for a
// dma_tile followed by at most KEEP loads or stores the wave does not need
// yet.  Every kernel that reads a dma_tile destination states its wait: the
// compiler's own placement of one is not reliable (an unrelated kernel added
// to the same module once turned a generic-to-LDS cast into a null-checked
// one, and the wait before the first read of the destination was dropped).
template <int KEEP> __device__ __forceinline__ void dma_wait_keep()
{
  static_assert(KEEP >= 0 && KEEP < 64, "vmcnt has six bits");
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_s_waitcnt(0x0f70 | (KEEP & 15) | ((KEEP >> 4) << 14)); // vmcnt(KEEP)
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename C> constexpr int site_lds() { return Tile<C>::LDS_ELEMS; }
constexpr int SITE_LDS = Tile<double2>::LDS_ELEMS; // FP64 padded site stride

template <typename C>
__device__ __forceinline__ void stage_sites(C *__restrict__ lds, const C *__restrict__ src, size_t first_site,
                                            int nsites, int lane)
{
  Tile<C> t;
  t.load(src, first_site, nsites, lane);
  t.store(lds, lane);
}

// LDS -> HBM: the inverse of staging, for propagator-shaped outputs.
template <typename C>
__device__ __forceinline__ void unstage_sites(C *__restrict__ dst, const C *__restrict__ lds, size_t first_site,
                                              int nsites, int lane)
{
  using T = Tile<C>;
  uint4 *base = reinterpret_cast<uint4 *>(dst + first_site * SITE_ELEMS);
  const uint4 *l = reinterpret_cast<const uint4 *>(lds);
  const int total = nsites * T::SITE16;
#pragma unroll
  for (int it = 0; it < T::PER_LANE; ++it) {
    const int e = it * WAVE + lane;
    if (e < total) {
      const int s = e / T::SITE16;
      stream_store(base + e, l[s * T::LDS16 + (e - s * T::SITE16)]);
    }
  }
}

// Persistent grid-stride loop over the 4-site tiles of `volume` sites for a
// kernel that stages NP propagators.  The loads of tile k+1 are issued into
// registers before body(k) runs, so they overlap the contraction of tile k;
// body(first_site, nsites) reads the staged sites from lds[p * 4 * SL ...].
template <int NP, typename C, typename Body>
__device__ __forceinline__ void persistent_tiles(C *__restrict__ lds, const C *const (&P)[NP], size_t volume,
                                                 int lane, Body &&body)
{
  constexpr int SL = Tile<C>::LDS_ELEMS;
  const size_t ntiles = (volume + SITES_PER_WAVE - 1) / SITES_PER_WAVE;
  auto nsites_of = [&](size_t t) {
    const size_t left = volume - t * SITES_PER_WAVE;
    return left < (size_t)SITES_PER_WAVE ? (int)left : SITES_PER_WAVE;
  };
  Tile<C> buf[NP];
  size_t tile = blockIdx.x;
  if (tile < ntiles)
#pragma unroll
    for (int p = 0; p < NP; ++p) buf[p].load(P[p], tile * SITES_PER_WAVE, nsites_of(tile), lane);
  for (; tile < ntiles; tile += gridDim.x) {
#pragma unroll
    for (int p = 0; p < NP; ++p) buf[p].store(lds + p * SITES_PER_WAVE * SL, lane);
    __syncthreads();
    const size_t next = tile + gridDim.x;
    if (next < ntiles)
#pragma unroll
      for (int p = 0; p < NP; ++p) buf[p].load(P[p], next * SITES_PER_WAVE, nsites_of(next), lane);
    body(tile * SITES_PER_WAVE, nsites_of(tile));
    __syncthreads();
  }
}

// Read one 3x3 colour block [s1][s2] of a staged site into registers.
template <typename C> __device__ __forceinline__ void load_block(C (&b)[9], const C *__restrict__ site_lds, int s1, int s2)
{
  const C *p = site_lds + (s1 * 4 + s2) * 9;
#pragma unroll
  for (int e = 0; e < 9; ++e) b[e] = p[e];
}

// ---------------------------------------------------------------------------
// colour algebra
// ---------------------------------------------------------------------------
// cof(Y,Z)[a,d] = sum_{bc,ef} eps_abc eps_def Y[b,e] Z[c,f], accumulated with
// weight w:  acc[a,d] += w * cof(Y,Z)[a,d].  Symmetric in Y and Z; the triple
// product ee(X,Y,Z) = sum_ad X[a,d] cof(Y,Z)[a,d].
template <typename C> __device__ __forceinline__ void cof_acc(C (&acc)[9], const C (&Y)[9], const C (&Z)[9], C w)
{
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int b = (a + 1) % 3, c = (a + 2) % 3;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int e = (d + 1) % 3, f = (d + 2) % 3;
      C t = cmul(Y[b * 3 + e], Z[c * 3 + f]);
      t = cfma(t, Y[c * 3 + f], Z[b * 3 + e]);
      t = cfms(t, Y[b * 3 + f], Z[c * 3 + e]);
      t = cfms(t, Y[c * 3 + e], Z[b * 3 + f]);
      acc[a * 3 + d] = cfma(acc[a * 3 + d], w, t);
    }
  }
}

// acc += cof(Y, Z) with Y read one entry at a time from LDS and Z held in
// registers: every Y[b][e] feeds the four entries (a, d) with a != b, d != e,
//   acc[a1][d1] += Y Z[a2][d2]   acc[a1][d2] -= Y Z[a2][d1]
//   acc[a2][d1] -= Y Z[a1][d2]   acc[a2][d2] += Y Z[a1][d1]
// with a1 = b+1, a2 = b+2, d1 = e+1, d2 = e+2 (mod 3).  Same 36 complex
// products as cof_acc with one block fewer live in registers.
template <typename C>
__device__ __forceinline__ void cof_acc_stream(C (&acc)[9], const C *__restrict__ Yp, const C (&Z)[9])
{
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    const int a1 = (b + 1) % 3, a2 = (b + 2) % 3;
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      const int d1 = (e + 1) % 3, d2 = (e + 2) % 3;
      const C y = Yp[b * 3 + e];
      acc[a1 * 3 + d1] = cfma(acc[a1 * 3 + d1], y, Z[a2 * 3 + d2]);
      acc[a1 * 3 + d2] = cfms(acc[a1 * 3 + d2], y, Z[a2 * 3 + d1]);
      acc[a2 * 3 + d1] = cfms(acc[a2 * 3 + d1], y, Z[a1 * 3 + d2]);
      acc[a2 * 3 + d2] = cfma(acc[a2 * 3 + d2], y, Z[a1 * 3 + d1]);
    }
  }
}

// acc += i^R cof(Y, Z), the unit folded into the FMAs (see cof_acc_stream)
template <int R, typename C>
__device__ __forceinline__ void cof_acc_stream_rot(C (&acc)[9], const C *__restrict__ Yp, const C (&Z)[9])
{
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    const int a1 = (b + 1) % 3, a2 = (b + 2) % 3;
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      const int d1 = (e + 1) % 3, d2 = (e + 2) % 3;
      const C y = Yp[b * 3 + e];
      acc[a1 * 3 + d1] = cfma_rot<R>(acc[a1 * 3 + d1], y, Z[a2 * 3 + d2]);
      acc[a1 * 3 + d2] = cfma_rot<R + 2>(acc[a1 * 3 + d2], y, Z[a2 * 3 + d1]);
      acc[a2 * 3 + d1] = cfma_rot<R + 2>(acc[a2 * 3 + d1], y, Z[a1 * 3 + d2]);
      acc[a2 * 3 + d2] = cfma_rot<R>(acc[a2 * 3 + d2], y, Z[a1 * 3 + d1]);
    }
  }
}

template <typename C> __device__ __forceinline__ C dot9(const C (&X)[9], const C (&Y)[9])
{
  C s = cmul(X[0], Y[0]);
#pragma unroll
  for (int e = 1; e < 9; ++e) s = cfma(s, X[e], Y[e]);
  return s;
}

// ---------------------------------------------------------------------------
// sparse spin matrices: every Gamma in the DeGrand-Rossi basis (and gamma_5
// times one) has exactly one nonzero per row, so a 4x4 matrix is stored as
// the column of that entry per row and its value.
// ---------------------------------------------------------------------------
struct SparseSpin {
  int col[4];
  double re[4];
  double im[4];
};

// Row lookups with a lane-dependent row go through selects, so the argument
// block stays in scalar registers instead of being copied to scratch for
// indexing.
template <typename T> __device__ __forceinline__ T mux4(T a0, T a1, T a2, T a3, int i)
{
  // two levels of bit-selects; an equality chain gets turned back into a
  // dynamically indexed stack array (scratch) by the optimiser
  const T lo = (i & 1) ? a1 : a0;
  const T hi = (i & 1) ? a3 : a2;
  return (i & 2) ? hi : lo;
}
template <typename T> __device__ __forceinline__ T sel4(const T (&a)[4], int i) { return mux4(a[0], a[1], a[2], a[3], i); }

// Reduce-scatter over the four lanes of a site group that differ in bits 2
// and 3 of q (i = q >> 2): lane i returns sum over the four lanes of r[i].
template <typename C> __device__ __forceinline__ C reduce_scatter4_hi(const C (&r)[4], int q)
{
  const bool h1 = q & 8, h0 = q & 4;
  C a0 = cadd(h1 ? r[2] : r[0], shfl_xor_c<8>(h1 ? r[0] : r[2]));
  C a1 = cadd(h1 ? r[3] : r[1], shfl_xor_c<8>(h1 ? r[1] : r[3]));
  return cadd(h0 ? a1 : a0, shfl_xor4_c(h0 ? a0 : a1, q));
}

// Same over the lanes that differ in bits 0 and 1 (y = q & 3): lane y returns
// sum over the four lanes of r[y].
template <typename C> __device__ __forceinline__ C reduce_scatter4_lo(const C (&r)[4], int q)
{
  const bool h1 = q & 2, h0 = q & 1;
  C a0 = cadd(h1 ? r[2] : r[0], shfl_xor_c<2>(h1 ? r[0] : r[2]));
  C a1 = cadd(h1 ? r[3] : r[1], shfl_xor_c<2>(h1 ? r[1] : r[3]));
  return cadd(h0 ? a1 : a0, shfl_xor_c<1>(h0 ? a0 : a1));
}

__device__ __forceinline__ int sparse_col(const SparseSpin &g, int row) { return sel4(g.col, row); }

template <typename C> __device__ __forceinline__ C sparse_val(const SparseSpin &g, int row)
{
  return cmake<C>((decltype(C::x))sel4(g.re, row), (decltype(C::x))sel4(g.im, row));
}

} // namespace contract_amd
