// project.hpp -- device side of the fused timeslice x momentum projection
// shared by the fused contraction kernels (baryon two-point, three-point):
// phase tables and their incremental walk over a work item, the FP64 MFMA
// accumulation of a per-site 16-component open-spin vector, the partial
// store, work-item ranges and the FP32 site rescaling.
#pragma once

#include <hip/hip_runtime.h>

#include "baryon.hpp"
#include "device.hpp"

namespace contract_amd
{

// ---------------------------------------------------------------------------
// Fused kernel: open-spin correlator summed over contraction terms, projected
// onto momenta per timeslice with FP64 MFMA.
//
// Work item = a contiguous run of sites inside one timeslice (and parity for
// even-odd order).  Per 4-site step the wave forms, for each 16-momentum tile,
//   D[p][q] += sum_{s<4} phase[p][s] * R[s][q]
// with v_mfma_f64_16x16x4f64: A = phases (lane: row p = lane&15, k = site
// lane>>4), B = R (lane: k = site lane>>4, column q = lane&15) -- exactly the
// registers the contraction leaves in each lane.  The result D[p][q] lands in
// lane 16*(p % 4) + q, accumulator register p / 4.  Complex products take four
// real MFMAs.  Partial sums go to partial[item][p][16]; project_finish sums
// them per timeslice with Neumaier compensation in a fixed order.
// ---------------------------------------------------------------------------
typedef double dbl4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int nonneg_mod(int a, int L)
{
  const int r = a % L;
  return r < 0 ? r + L : r;
}

// Phases exp(-2 pi i p.x / L) of this lane's MFMA rows (momenta p = tt*16 + q)
// at this lane's site, walked along the work item without integer division:
// the lane's site advances by SITES_PER_WAVE per step, and the table indices
// k_mu = (p_mu x_mu) mod L_mu move by fixed increments (a wrap of x_mu by L_mu
// leaves k_mu unchanged).  A slice is [z][y][xr] with xr = x (lexicographic)
// or x / 2 (even-odd: x = 2 xr + ((parity + y + z + t) & 1)).
// (a + b) mod G for 0 <= a, b < G
__device__ __forceinline__ int add_mod(int a, int b, int G)
{
  const int r = a + b;
  return r >= G ? r - G : r;
}

template <int NT> struct PhaseWalk {
  static_assert(MAX_PHASE_TAB < 512, "table indices are packed in 9-bit fields");
  int xr, y, z;  // lane coordinates (local)
  int Lr, pbase; // row length of a slice, parity + t (even-odd)
  // per momentum tile, three 9-bit fields per register (global extents G_mu < 512):
  //   k = kx | ky << 9 | kz << 18, kx = p_x (xs xr + o_x) mod Gx with xs = 2 for
  //   even-odd (the parity bit is added in phase()) and 1 otherwise,
  //   ky = p_y (y + o_y) mod Gy, kz = p_z (z + o_z) mod Gz;
  //   m = p_x | p_y << 9 | p_z << 18 (p_mu mod G_mu; 0 for rows past np);
  //   d = the increments, each reduced mod G_mu: xs * 4 * p_x per 4-site
  //   step | -p_x Lx when x wraps (y steps on) << 9 | -p_y Ly when y wraps
  //   (z steps on) << 18, so a step is adds and conditional subtracts
  int k[NT], m[NT], d[NT];

  __device__ __forceinline__ void init(const ProjectArgs &args, const ProjectReq &pa, size_t first, int g, int q)
  {
    const size_t slice = first / pa.slice_sites;
    const unsigned r0 = (unsigned)(first - slice * pa.slice_sites) + (unsigned)g;
    const int lt = pa.dims[3];
    pbase = (int)(slice / lt) + (int)(slice % lt);
    Lr = pa.order ? pa.dims[0] >> 1 : pa.dims[0];
    xr = (int)(r0 % (unsigned)Lr);
    const unsigned rest = r0 / (unsigned)Lr;
    y = (int)(rest % (unsigned)pa.dims[1]);
    z = (int)(rest / (unsigned)pa.dims[1]);
    const int xs = pa.order ? 2 : 1;
#pragma unroll
    for (int tt = 0; tt < NT; ++tt) {
      const int p = tt * 16 + q;
      const bool live = p < args.np;
      const int mx = live ? nonneg_mod(args.mom[p][0], pa.gdims[0]) : 0;
      const int my = live ? nonneg_mod(args.mom[p][1], pa.gdims[1]) : 0;
      const int mz = live ? nonneg_mod(args.mom[p][2], pa.gdims[2]) : 0;
      const int kx = (mx * ((xs * xr + pa.origin[0]) % pa.gdims[0])) % pa.gdims[0];
      const int ky = (my * ((y + pa.origin[1]) % pa.gdims[1])) % pa.gdims[1];
      const int kz = (mz * ((z + pa.origin[2]) % pa.gdims[2])) % pa.gdims[2];
      k[tt] = kx | ky << 9 | kz << 18;
      m[tt] = mx | my << 9 | mz << 18;
      const int gx = pa.gdims[0], gy = pa.gdims[1];
      const int dkx = (xs * SITES_PER_WAVE * mx) % gx;
      const int wx = (gx - (mx * pa.dims[0]) % gx) % gx;
      const int wy = (gy - (my * pa.dims[1]) % gy) % gy;
      d[tt] = dkx | wx << 9 | wy << 18;
    }
  }

  __device__ __forceinline__ double2 phase(const ProjectReq &pa, const double2 *tab, int tt) const
  {
    int kx = k[tt] & 511;
    if (pa.order && ((pbase + y + z) & 1)) {
      kx += m[tt] & 511;
      kx -= kx >= pa.gdims[0] ? pa.gdims[0] : 0;
    }
    return cmul(cmul(tab[kx], tab[pa.gdims[0] + ((k[tt] >> 9) & 511)]),
                tab[pa.gdims[0] + pa.gdims[1] + (k[tt] >> 18)]);
  }

  __device__ __forceinline__ void advance(const ProjectReq &pa)
  {
    const int Gx = pa.gdims[0], Gy = pa.gdims[1], Gz = pa.gdims[2];
#pragma unroll
    for (int tt = 0; tt < NT; ++tt) k[tt] = (k[tt] & ~511) | add_mod(k[tt] & 511, d[tt] & 511, Gx);
    xr += SITES_PER_WAVE;
    while (xr >= Lr) {
      // x steps back by Lx: kx -= p_x Lx (0 mod Gx for a whole lattice), y steps on
      xr -= Lr;
      ++y;
#pragma unroll
      for (int tt = 0; tt < NT; ++tt) {
        const int kx = add_mod(k[tt] & 511, (d[tt] >> 9) & 511, Gx);
        const int ky = add_mod((k[tt] >> 9) & 511, (m[tt] >> 9) & 511, Gy);
        k[tt] = (k[tt] & ~((1 << 18) - 1)) | ky << 9 | kx;
      }
      if (y == pa.dims[1]) {
        // y steps back by Ly: ky -= p_y Ly, z steps on
        y = 0;
        ++z;
#pragma unroll
        for (int tt = 0; tt < NT; ++tt) {
          const int ky = add_mod((k[tt] >> 9) & 511, d[tt] >> 18, Gy);
          const int kz = add_mod(k[tt] >> 18, m[tt] >> 18, Gz);
          k[tt] = (k[tt] & 511) | ky << 9 | kz << 18;
        }
      }
    }
  }
};

// Momentum slot k (one per +-k pair) keeps C = sum cos(k.x) R and
// S = -sum sin(k.x) R in four accumulators (C re, C im, S re, S im); the finish
// forms P(+-k) = C +- iS.
// This lane's phase rows (cos, -sin) at its site of the step; rows past np
// are zero.
template <int NT>
__device__ __forceinline__ void project_phases(double (&pr)[NT], double (&pi)[NT], int np, const ProjectReq &pa,
                                               const double2 *tab, const PhaseWalk<NT> &walk)
{
#pragma unroll
  for (int tt = 0; tt < NT; ++tt) {
    const double2 ph = walk.phase(pa, tab, tt); // (cos, -sin)
    const bool live = tt * 16 + (int)(threadIdx.x & 15) < np; // rows past np stay zero
    pr[tt] = live ? ph.x : 0.0;
    pi[tt] = live ? ph.y : 0.0;
  }
}

template <int NT>
__device__ __forceinline__ void project_mfma(dbl4 (&acc)[NT][4], const double (&pr)[NT], const double (&pi)[NT],
                                             double2 R)
{
#pragma unroll
  for (int tt = 0; tt < NT; ++tt) {
    acc[tt][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(pr[tt], R.x, acc[tt][0], 0, 0, 0);
    acc[tt][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(pr[tt], R.y, acc[tt][1], 0, 0, 0);
    acc[tt][2] = __builtin_amdgcn_mfma_f64_16x16x4f64(pi[tt], R.x, acc[tt][2], 0, 0, 0);
    acc[tt][3] = __builtin_amdgcn_mfma_f64_16x16x4f64(pi[tt], R.y, acc[tt][3], 0, 0, 0);
  }
}

template <int NT>
__device__ __forceinline__ void project_accumulate(dbl4 (&acc)[NT][4], int np, const ProjectReq &pa,
                                                   const double2 *tab, double2 R, const PhaseWalk<NT> &walk)
{
  double pr[NT], pi[NT];
  project_phases<NT>(pr, pi, np, pa, tab, walk);
#pragma unroll
  for (int tt = 0; tt < NT; ++tt) {
    acc[tt][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(pr[tt], R.x, acc[tt][0], 0, 0, 0);
    acc[tt][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(pr[tt], R.y, acc[tt][1], 0, 0, 0);
    acc[tt][2] = __builtin_amdgcn_mfma_f64_16x16x4f64(pi[tt], R.x, acc[tt][2], 0, 0, 0);
    acc[tt][3] = __builtin_amdgcn_mfma_f64_16x16x4f64(pi[tt], R.y, acc[tt][3], 0, 0, 0);
  }
}

// Build the phase tables exp(-2 pi i k / G_mu) of the three spatial directions
// (global extents) in LDS ([Gx | Gy | Gz] entries); a few sincospi per lane per
// work item.
__device__ __forceinline__ void load_phase_tab(double2 *tab, const ProjectReq &pa, int lane)
{
  const int G0 = pa.gdims[0], G1 = pa.gdims[1], G2 = pa.gdims[2];
  const int n = G0 + G1 + G2;
  for (int e = lane; e < n; e += WAVE) {
    const int L = e < G0 ? G0 : (e < G0 + G1 ? G1 : G2);
    const int k = e < G0 ? e : (e < G0 + G1 ? e - G0 : e - G0 - G1);
    double sn, cs;
    sincospi(2.0 * (double)k / (double)L, &sn, &cs);
    tab[e] = make_double2(cs, -sn);
  }
}

template <int NT>
__device__ __forceinline__ void project_store(const dbl4 (&acc)[NT][4], int np, const ProjectReq &pa, size_t item,
                                              int lane)
{
  const int g = lane >> 4, q = lane & 15;
  double2 *dst = pa.partial + item * (size_t)np * 32;
#pragma unroll
  for (int tt = 0; tt < NT; ++tt)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int p = tt * 16 + g + 4 * r; // v_mfma_f64_16x16x4f64: lane (g, q) register r holds D[g + 4r][q]
      if (p < np) {
        dst[p * 32 + q] = make_double2(acc[tt][0][r], acc[tt][1][r]);      // C
        dst[p * 32 + 16 + q] = make_double2(acc[tt][2][r], acc[tt][3][r]); // S
      }
    }
}

// A wave-uniform value computed on the vector ALU (the 64-bit divisions of
// item_range) moved to scalar registers.
__device__ __forceinline__ size_t uniform_size(size_t v)
{
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)v);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v >> 32));
  return (size_t)hi << 32 | lo;
}

__device__ __forceinline__ void item_range(const ProjectReq &pa, size_t item, size_t &first, size_t &last)
{
  const size_t slice = item / pa.chunks_per_slice;
  const size_t c = item - slice * pa.chunks_per_slice;
  const size_t s0 = slice * pa.slice_sites;
  first = s0 + c * (size_t)pa.chunk;
  last = first + pa.chunk;
  const size_t s1 = s0 + pa.slice_sites;
  last = uniform_size(last < s1 ? last : s1);
  first = uniform_size(first < s1 ? first : s1);
}

// FP32 storage: rescale a staged site of a propagator by an exact power of
// two so its largest entry lies in [1, 2).  The contraction is trilinear, so
// rescaling P_i and P_j makes the FP32 result times 2^(e_i + e_j) the true
// value; the product of three steeply decaying propagators (~1e-66 at the far
// end of a 64-timeslice lattice) therefore never underflows FP32.
__device__ __forceinline__ int normalise_site(float2 *site_lds, int q)
{
  float2 *p = site_lds + q * 9;
  float mx = 0.0f;
#pragma unroll
  for (int e = 0; e < 9; ++e) mx = fmaxf(mx, fmaxf(fabsf(p[e].x), fabsf(p[e].y)));
#pragma unroll
  for (int m = 8; m >= 1; m >>= 1) mx = fmaxf(mx, __shfl_xor(mx, m, 64));
  int ex = (mx > 0.0f) ? ilogbf(mx) : 0;
  ex = ex < -126 ? -126 : (ex > 126 ? 126 : ex);
  const float sc = ldexpf(1.0f, -ex);
#pragma unroll
  for (int e = 0; e < 9; ++e) p[e] = make_float2(p[e].x * sc, p[e].y * sc);
  return ex;
}

} // namespace contract_amd
