// meson_grid.hip -- batched meson two-point Gamma x Gamma' grids projected to
// timeslices and momenta, many requests (uneven lattices) in one persistent
// launch.
//
// For one site the 16 x 16 grid of meson_two_point values
// (include/meson.cuh:32-50, meson(x; Gij = g1, Gkl = g2)) is a linear image of
// the 256 spin-open colour traces
//
//   K[i,k,j,l] = sum_ad P_i[i,k](a,d) conj(P_j[j,l](a,d)),
//   c(g1, g2)  = sum_{i,l} (g5 G(g1))[j(i), i] (g5 G(g2))[k(l), l] K[i,k(l),j(i),l],
//
// where j(i) / k(l) are the nonzero rows of column i / l of g5 G.  The kernel
// therefore projects K (not the 256 correlators) to (timeslice, momentum) and
// a small epilogue applies the gammas.  The projection is an FP64 MFMA GEMM
// per 4-site step, out[row][col] += sum_s A[row][s] K[s][col] with rows =
// (momentum, re/im of the phase) and 16 columns per MFMA.  All 256 columns x
// (K re, K im) x 16 rows are 8192 accumulators, so a workgroup of four waves
// shares one staged tile and wave w owns the columns with k = w: lane (i, j)
// of a site group computes K[i,w,j,l] for l = 0..3.
//
// Work item = a run of sites inside one timeslice of one request; the
// persistent grid walks the flat item space of the whole batch, so uneven
// requests balance across the chip.

#include <type_traits>

#include "meson_grid.hpp"

namespace contract_amd
{

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int nonneg_mod(int a, int L)
{
  const int r = a % L;
  return r < 0 ? r + L : r;
}

// FP32 storage: the X block a lane holds is rescaled in registers by an
// exact power of two so its largest entry lies in [1, 2); the trace with the
// stored Y block then stays far above FP32's underflow threshold (it is as
// large as Y's entries), and the FP64 result is scaled back exactly.
__device__ __forceinline__ int scale_block(float2 (&b)[9])
{
  float mx = 0.0f;
#pragma unroll
  for (int e = 0; e < 9; ++e) mx = fmaxf(mx, fmaxf(fabsf(b[e].x), fabsf(b[e].y)));
  int ex = (mx > 0.0f) ? ilogbf(mx) : 0;
  ex = ex < -126 ? -126 : (ex > 126 ? 126 : ex);
  const float sc = ldexpf(1.0f, -ex);
#pragma unroll
  for (int e = 0; e < 9; ++e) b[e] = make_float2(b[e].x * sc, b[e].y * sc);
  return ex;
}

// Per-site traces K[ik][jl] = sum_c P_i[ik][c] conj(P_j[jl][c]) (c = the
// nine colour pairs) are a complex 16 x 9 x 16 product per site.  Two forms:
//  KMFMA = false: on the vector ALUs, lane (i, j) of wave k forms K[i k][j l]
//    for l = 0..3 from one P_i block and four P_j blocks read from LDS;
//  KMFMA = true: on the matrix cores, wave w forms all of K for site w of each
//    four-site step with 6 (FP64) or 8 (FP32) v_mfma 16x16x4 (2 colour
//    chunks x 3 or 4 real products; A = P_i rows, B = P_j rows) and colour
//    pair 8 on the vector ALUs, each staged entry read from LDS once; K goes
//    through LDS once more to become the projection's B operand (4 sites x 16
//    columns).  Needs FP64 or site-scaled FP32 storage: plain FP32 products of
//    two steeply decaying propagators could underflow without a per-row scale.
// RT row tiles of 8 momentum slots each (one pass holds up to 8 RT slots):
// every staged site is read and traced once for all of them.  Two tiles
// double the projection accumulators, which fit only at one wave per SIMD.
// Q4 (matrix-core traces, passes of at most 4 slots, e.g. the origin and
// the six unit momenta): 8 of the 16 rows of a 16x16x4 projection MFMA would
// be zero, so the projection runs as two v_mfma_f64_4x4x4f64 (four 4x4
// blocks, a quarter of the cycles each) over rows 0-3 and 4-7.  Their lane
// layout (measured with an in-bounds lane probe): A[b][i][k] from lane
// 16k + 4b + i, B[b][k][j] from lane 16k + 4b + j, D[b][i][j] in lane
// 16i + 4b + j.  With k = site g and 4b + j = column q, B is the 16x16x4
// form's operand unchanged; A is the phase of row 4 half + (q & 3) at site
// g; lane (g, q) ends with rows 4 half + g of column q.
// FP64 matrix-core traces with two row tiles (passes of 9-16 momentum
// slots): the tile moves HBM -> LDS directly (dma_tile_pair_split, unpadded
// sites), issued once every wave has traced the current one, which drops the
// registers of the prefetch and fits two waves per SIMD.
template <typename Real, bool KMFMA, int RT> constexpr bool grid_dma()
{
  return KMFMA && RT == 2 && sizeof(Real) == 8;
}

template <typename Real, int NS, bool SCALED = false, bool KMFMA = false, int RT = 1, bool Q4 = false>
__global__ __launch_bounds__(GRID_THREADS)
__attribute__((amdgpu_waves_per_eu(grid_dma<Real, KMFMA, RT>() ? 2
                                   : ((RT > 1 || (sizeof(Real) == 8 && NS == 8 && !KMFMA)) ? 1 : 2)))) void
meson_grid_kernel(MesonGridArgs a)
{
  static_assert(NS % SITES_PER_WAVE == 0, "tile is a whole number of MFMA site steps");
  static_assert(!Q4 || (KMFMA && RT == 1), "the quarter projection is a form of the one-tile matrix-core kernel");
  using C = cplx<Real>;
  using T = Tile<C>;
  constexpr bool DMA = grid_dma<Real, KMFMA, RT>();
  static_assert(!DMA || Tile<C, NS>::DMA_ELEMS == NS * SITE_ELEMS, "a DMA tile is whole instructions");
  constexpr int SL = DMA ? SITE_ELEMS : T::LDS_ELEMS; // site stride (complex entries): padded unless DMA
  constexpr int CH = NS * T::SITE16;             // 16-byte chunks per propagator tile
  constexpr int PER_PROP = (CH + GRID_THREADS - 1) / GRID_THREADS; // chunks per thread and propagator
  constexpr int TILE16 = NS * T::LDS16;          // one propagator tile in LDS, 16-byte units
  __shared__ C lds[2][NS * SL];                  // [propagator]
  // K of every site of the staged tile, [four-site group][site][ik][jl]:
  // each wave forms K of its sites w, w + 4, ... of the tile, one barrier
  // hands them over, then every wave projects its columns of all of them;
  // the next tile's writes follow the barrier at the top of its step.  Rows
  // of 16 entries and sites of 256 (unpadded): a ds_read/write_b128 of a wave
  // is conflict-free when each 16-lane group touches 16 distinct banks and
  // lane q of every group the same bank (group offsets a multiple of 256 B;
  // measured with an in-bounds LDS probe), which both the MFMA-layout writes
  // (groups on rows g + 4v) and the B-operand reads (groups on sites) meet
  constexpr int KROW = 16, KSITE = 16 * KROW;
  constexpr int NH = NS / SITES_PER_WAVE; // four-site groups per tile
  __shared__ double2 kbuf[KMFMA ? NH : 1][KMFMA ? SITES_PER_WAVE : 1][KSITE];
  extern __shared__ double2 tab[];
  const int tid = threadIdx.x;
  const int w = tid >> 6, lane = tid & 63;
  const int g = lane >> 4, q = lane & 15;
  const int i = q >> 2, j = q & 3;

  // this thread's chunks of a propagator tile, e = it * GRID_THREADS + tid
  // (a wave's loads all come from one propagator, so each takes one buffer
  // resource), and their LDS slots (the same every step)
  int dst_off[PER_PROP];
#pragma unroll
  for (int it = 0; it < PER_PROP; ++it) {
    const int e = it * GRID_THREADS + tid;
    const int s = e / T::SITE16;
    dst_off[it] = s * T::LDS16 + (e - s * T::SITE16);
  }
  auto active = [&](int it) { return CH % GRID_THREADS == 0 || it + 1 < PER_PROP || tid < CH % GRID_THREADS; };

  for (size_t item = blockIdx.x; item < a.total_items; item += gridDim.x) {
    int r = 0; // request of this item (uniform over the workgroup)
    while (r + 1 < a.nreq && item >= a.req[r + 1].item0) ++r;
    const MesonGridReq &R = a.req[r];
    const size_t local = item - R.item0;
    const size_t slice = local / R.chunks_per_slice;
    const size_t c = local - slice * R.chunks_per_slice;
    const size_t s0 = slice * R.slice_sites;
    const size_t first = s0 + c * (size_t)a.chunk;
    size_t last = first + a.chunk;
    last = last < s0 + R.slice_sites ? last : s0 + R.slice_sites;

    __syncthreads(); // the previous item's readers of tab are done
    // per-momentum phase tables exp(-2 pi i p_mu (k + o_mu) / G_mu) for local
    // k < L_mu (G, o: global extent and origin): [np][Lx + 1] | [np][Ly + 1] |
    // [np][Lz + 1], so a lane's phase needs no division; the padding entry
    // keeps rows of power-of-two length from starting on the same LDS bank
    // (the eight rows a wave reads at once)
    const int Lx = R.dims[0], Ly = R.dims[1], Lz = R.dims[2];
    const int offy = a.np * (Lx + 1), offz = a.np * (Lx + Ly + 2);
    const int ntab = a.np * (Lx + Ly + Lz + 3);
    for (int e = tid; e < ntab; e += GRID_THREADS) {
      const int mu = e < offy ? 0 : (e < offz ? 1 : 2);
      const int L = R.dims[mu], G = R.gdims[mu];
      const int rel = e - (mu == 0 ? 0 : (mu == 1 ? offy : offz));
      const int p = rel / (L + 1), k = rel - p * (L + 1);
      const int m = nonneg_mod(a.mom[p][mu], G);
      double sn, cs;
      sincospi(2.0 * (double)((m * ((k + R.origin[mu]) % G)) % G) / (double)G, &sn, &cs);
      tab[e] = make_double2(cs, -sn);
    }
    // this lane's MFMA row of tile rt: momentum slot 8 rt + (q >> 1), phase
    // part q & 1 (0 re, 1 im)
    bool live[RT];
    int prow[RT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
      const int pr = rt * GRID_MOM_PER_TILE + (q >> 1);
      live[rt] = pr < a.np;
      prow[rt] = live[rt] ? pr : 0;
    }
    // Q4: this lane's A rows 4 half + (q & 3): slot 2 half + ((q >> 1) & 1), part q & 1
    bool live4[2];
    int prow4[2];
#pragma unroll
    for (int hf = 0; hf < 2; ++hf) {
      const int pr = 2 * hf + ((q >> 1) & 1);
      live4[hf] = pr < a.np;
      prow4[hf] = live4[hf] ? pr : 0;
    }

    // coordinates of the step's first site, advanced incrementally (uniform):
    // a slice is [z][y][xr] with xr = x (lexicographic) or x / 2 (even-odd,
    // where x = 2 xr + ((parity + y + z + t) & 1))
    const int lt = R.dims[3];
    const int par = (int)(slice / lt), tsl = (int)(slice - (size_t)par * lt);
    const int Lr = R.order ? Lx >> 1 : Lx;
    int x0, y0, z0;
    {
      const unsigned r0 = (unsigned)(c * (size_t)a.chunk);
      x0 = (int)(r0 % (unsigned)Lr);
      const unsigned rest = r0 / (unsigned)Lr;
      y0 = (int)(rest % (unsigned)Ly);
      z0 = (int)(rest / (unsigned)Ly);
    }
    const int pbase = par + tsl;

    // column tiles of k = w: l = 0..3 (vector traces, lane q = (i, j)) or
    // i = 0..3 (matrix-core traces, lane q = (j, l)); B = K re / K im
    dbl4 acc_re[RT][4], acc_im[RT][4];
    double acc4_re[4][2], acc4_im[4][2]; // Q4: [i][half]
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int l = 0; l < 4; ++l) acc_re[rt][l] = acc_im[rt][l] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int l = 0; l < 4; ++l) acc4_re[l][0] = acc4_re[l][1] = acc4_im[l][0] = acc4_im[l][1] = 0.0;

    // the workgroup stages NS sites of both propagators per step; the next
    // step's chunks are in flight in registers while this one is computed
    const C *P0 = static_cast<const C *>(R.propag_i), *P1 = static_cast<const C *>(R.propag_j);
    uint4 buf[2][PER_PROP];
    auto load = [&](size_t st) { // chunks past the item's end read as zero
      const size_t left = last - st;
      const unsigned bytes = (unsigned)((left < (size_t)NS ? left : (size_t)NS) * T::SITE16 * 16);
      const __amdgpu_buffer_rsrc_t r0 = tile_rsrc(P0 + st * SITE_ELEMS, bytes);
      const __amdgpu_buffer_rsrc_t r1 = tile_rsrc(P1 + st * SITE_ELEMS, bytes);
#pragma unroll
      for (int it = 0; it < PER_PROP; ++it) {
        buf[0][it] = stream_load(r0, (it * GRID_THREADS + tid) * 16);
        buf[1][it] = stream_load(r1, (it * GRID_THREADS + tid) * 16);
      }
    };
    auto store = [&]() {
      uint4 *l16 = reinterpret_cast<uint4 *>(&lds[0][0]);
#pragma unroll
      for (int it = 0; it < PER_PROP; ++it)
        if (active(it)) {
          l16[dst_off[it]] = buf[0][it];
          l16[TILE16 + dst_off[it]] = buf[1][it];
        }
    };
    auto tile_sites = [&](size_t st) {
      const size_t left = last - st;
      return left < (size_t)NS ? (int)left : NS;
    };
    if constexpr (DMA)
      dma_tile_pair_split<GRID_THREADS / WAVE, NS>(lds[0], lds[1], P0, P1, first, tile_sites(first), lane, w);
    else
      load(first);
    // phases of this lane's site in four-site group h of the step (sites past
    // the item's end carry K = 0; their coordinates only need to stay inside
    // the tables)
    auto site_phases = [&](int h, double (&av)[RT], double (&av4)[2]) {
      const int ls = h * SITES_PER_WAVE + g;
      int xr = x0 + ls, y = y0, z = z0;
      while (xr >= Lr) {
        xr -= Lr;
        if (++y == Ly) {
          y = 0;
          ++z;
        }
      }
      z = z < Lz ? z : 0;
      const int x = R.order ? 2 * xr + ((pbase + y + z) & 1) : xr;
      auto phase = [&](int pr) {
        const double2 ph =
          cmul(cmul(tab[pr * (Lx + 1) + x], tab[offy + pr * (Ly + 1) + y]), tab[offz + pr * (Lz + 1) + z]);
        return (q & 1) ? ph.y : ph.x;
      };
      if constexpr (Q4) {
#pragma unroll
        for (int hf = 0; hf < 2; ++hf) av4[hf] = live4[hf] ? phase(prow4[hf]) : 0.0;
      } else {
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) av[rt] = live[rt] ? phase(prow[rt]) : 0.0;
      }
    };
    for (size_t st = first; st < last; st += NS) {
      if constexpr (DMA) {
        dma_wait();      // this wave's share of the tile
        __syncthreads(); // every share has landed; the readers of the previous K are done
      } else {
        __syncthreads(); // the readers of the previous tile are done
        store();
        if (st + NS < last) load(st + NS);
        __syncthreads();
      }
      if constexpr (KMFMA) {
        // traces of this wave's sites of the tile, into kbuf
#pragma unroll
        for (int h = 0; h < NH; ++h) {
          const int sk = h * SITES_PER_WAVE + w; // this wave's site
          using KAcc = typename std::conditional<sizeof(Real) == 8, dbl4, flt4>::type;
          const C *xs = lds[0] + sk * SL, *ys = lds[1] + sk * SL;
          // FP64: three real products per colour chunk (Gauss), T1 = Xr Yr^T,
          // T2 = Xi Yi^T, T3 = (Xr + Xi)(Yr - Yi)^T, so K = X conj(Y)^T has
          // Re = T1 + T2, Im = T3 - T1 + T2 (6 FP64 MFMAs per site: FP64
          // matrix issue is the scarcer resource).  FP32: four products (FP32
          // MFMA issue is twice as fast, and the Gauss form's cancellation in
          // Im costs FP32 accuracy)
          KAcc kr = KAcc{0, 0, 0, 0}, ki = KAcc{0, 0, 0, 0}, t3 = KAcc{0, 0, 0, 0};
#pragma unroll
          for (int ch = 0; ch < 2; ++ch) {
            // A[ik = q][kk = g] = P_i[q][c], B[kk = g][jl = q] = P_j[q][c], c = 4 ch + g
            const int cc = 4 * ch + g;
            const C xv = xs[q * 9 + cc], yv = ys[q * 9 + cc];
            if constexpr (sizeof(Real) == 8) {
              kr = __builtin_amdgcn_mfma_f64_16x16x4f64(xv.x, yv.x, kr, 0, 0, 0); // T1
              ki = __builtin_amdgcn_mfma_f64_16x16x4f64(xv.y, yv.y, ki, 0, 0, 0); // T2
              t3 = __builtin_amdgcn_mfma_f64_16x16x4f64(xv.x + xv.y, yv.x - yv.y, t3, 0, 0, 0);
            } else {
              kr = __builtin_amdgcn_mfma_f32_16x16x4f32(xv.x, yv.x, kr, 0, 0, 0);
              kr = __builtin_amdgcn_mfma_f32_16x16x4f32(xv.y, yv.y, kr, 0, 0, 0);
              ki = __builtin_amdgcn_mfma_f32_16x16x4f32(xv.y, yv.x, ki, 0, 0, 0);
              ki = __builtin_amdgcn_mfma_f32_16x16x4f32(-xv.x, yv.y, ki, 0, 0, 0);
            }
          }
          if constexpr (sizeof(Real) == 8) {
            const KAcc t1 = kr, t2 = ki;
            kr = t1 + t2;
            ki = t3 - t1 + t2;
          }
          {
            // colour pair 8 on the vector ALUs, in the MFMA output layout: a
            // third colour chunk would multiply three zero pads per entry
            const C yv = ys[q * 9 + 8];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const int row = sizeof(Real) == 8 ? g + 4 * v : 4 * g + v;
              const C t = cfma_conj(cmake<C>(kr[v], ki[v]), xs[row * 9 + 8], yv);
              kr[v] = t.x;
              ki[v] = t.y;
            }
          }
          int ex = 0;
          if constexpr (SCALED) {
            const size_t site = st + (size_t)sk < last ? st + (size_t)sk : last - 1;
            ex = (int)R.exp_i[site] + (int)R.exp_j[site];
          }
          // 16x16x4 output: lane (g, q) register v holds D[g + 4v][q] (f64) or
          // D[4g + v][q] (f32)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int row = sizeof(Real) == 8 ? g + 4 * v : 4 * g + v;
            double re = (double)kr[v], im = (double)ki[v];
            if constexpr (SCALED) {
              re = ldexp(re, ex);
              im = ldexp(im, ex);
            }
            kbuf[h][w][row * KROW + q] = make_double2(re, im);
          }
        }
        __syncthreads();
      }
      // DMA: the phases of every site group (table reads before the next
      // tile's loads are issued: a dynamic-LDS read after them would wait for
      // them), then the next tile into the traced one's place
      double avd[DMA ? NS / SITES_PER_WAVE : 1][RT];
      if constexpr (DMA) {
#pragma unroll
        for (int h = 0; h < NS / SITES_PER_WAVE; ++h) {
          double unused4[2];
          site_phases(h, avd[h], unused4);
        }
        if (st + NS < last)
          dma_tile_pair_split<GRID_THREADS / WAVE, NS>(lds[0], lds[1], P0, P1, st + NS, tile_sites(st + NS), lane, w);
      }
#pragma unroll
      for (int h = 0; h < NS / SITES_PER_WAVE; ++h) {
        const int ls = h * SITES_PER_WAVE + g; // this lane's site in the tile
        double av[RT], av4[2];
        if constexpr (DMA) {
#pragma unroll
          for (int rt = 0; rt < RT; ++rt) av[rt] = avd[h][rt];
        } else {
          site_phases(h, av, av4);
        }
        if constexpr (KMFMA) {
          // rows ik = 4 i + k of this wave's k = w; B[site g][jl = q]
#pragma unroll
          for (int i2 = 0; i2 < 4; ++i2) {
            const double2 kv = kbuf[h][g][(4 * i2 + w) * KROW + q];
            if constexpr (Q4) {
#pragma unroll
              for (int hf = 0; hf < 2; ++hf) {
                acc4_re[i2][hf] = __builtin_amdgcn_mfma_f64_4x4x4f64(av4[hf], kv.x, acc4_re[i2][hf], 0, 0, 0);
                acc4_im[i2][hf] = __builtin_amdgcn_mfma_f64_4x4x4f64(av4[hf], kv.y, acc4_im[i2][hf], 0, 0, 0);
              }
            } else {
#pragma unroll
              for (int rt = 0; rt < RT; ++rt) {
                acc_re[rt][i2] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[rt], kv.x, acc_re[rt][i2], 0, 0, 0);
                acc_im[rt][i2] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[rt], kv.y, acc_im[rt][i2], 0, 0, 0);
              }
            }
          }
          continue;
        }
        const C *si = lds[0] + ls * SL, *sj = lds[1] + ls * SL;
        C xb[9], yb[9];
        load_block(xb, si, i, w);
        int ex = 0;
        if constexpr (SCALED) { // site-scaled storage: the site's exponents (a site past the item reads the last one)
          const size_t site = st + (size_t)ls < last ? st + (size_t)ls : last - 1;
          ex = (int)R.exp_i[site] + (int)R.exp_j[site];
        } else if constexpr (sizeof(Real) == 4) {
          ex = scale_block(xb);
        }
#pragma unroll
        for (int l = 0; l < 4; ++l) {
          load_block(yb, sj, j, l);
          C t = czero<C>();
#pragma unroll
          for (int e = 0; e < 9; ++e) t = cfma_conj(t, xb[e], yb[e]);
          double kr = (double)t.x, ki = (double)t.y;
          if constexpr (sizeof(Real) == 4) {
            kr = ldexp(kr, ex);
            ki = ldexp(ki, ex);
          }
#pragma unroll
          for (int rt = 0; rt < RT; ++rt) {
            acc_re[rt][l] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[rt], kr, acc_re[rt][l], 0, 0, 0);
            acc_im[rt][l] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[rt], ki, acc_im[rt][l], 0, 0, 0);
          }
        }
      }
      x0 += NS;
      while (x0 >= Lr) {
        x0 -= Lr;
        if (++y0 == Ly) {
          y0 = 0;
          ++z0;
        }
      }
    }
    // v_mfma_f64_16x16x4f64: lane (g, q) register v holds D[g + 4v][q];
    // partial column (k l) * 16 + (i j) with k = w
    double *dst = a.partial + item * (size_t)a.partial_tiles * GRID_PARTIAL_TILE;
    if constexpr (Q4) {
      // lane (g, q) holds rows 4 half + g of column q; rows 8-15 (slots 4-7)
      // are not read for a pass of at most 4 slots
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int hf = 0; hf < 2; ++hf) {
          const int row = 4 * hf + g;
          const int col = (w * 4 + (q & 3)) * 16 + t * 4 + (q >> 2);
          dst[(row * 2 + 0) * GRID_COLS + col] = acc4_re[t][hf];
          dst[(row * 2 + 1) * GRID_COLS + col] = acc4_im[t][hf];
        }
    } else {
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int row = rt * 16 + g + 4 * v;
            const int col = KMFMA ? (w * 4 + (q & 3)) * 16 + t * 4 + (q >> 2) : (w * 4 + t) * 16 + q;
            dst[(row * 2 + 0) * GRID_COLS + col] = acc_re[rt][t][v];
            dst[(row * 2 + 1) * GRID_COLS + col] = acc_im[rt][t][v];
          }
    }
  }
}

// The two 32-bit halves of a double, for the lane-row swaps below.
__device__ __forceinline__ void swap_rows32(double &a, double &b)
{
  const unsigned long long ua = __builtin_bit_cast(unsigned long long, a);
  const unsigned long long ub = __builtin_bit_cast(unsigned long long, b);
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)ua, (unsigned)ub, false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32), false, false);
  a = __builtin_bit_cast(double, (unsigned long long)hi[0] << 32 | lo[0]);
  b = __builtin_bit_cast(double, (unsigned long long)hi[1] << 32 | lo[1]);
}
__device__ __forceinline__ void swap_rows16(double &a, double &b)
{
  const unsigned long long ua = __builtin_bit_cast(unsigned long long, a);
  const unsigned long long ub = __builtin_bit_cast(unsigned long long, b);
  const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)ua, (unsigned)ub, false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32), false, false);
  a = __builtin_bit_cast(double, (unsigned long long)hi[0] << 32 | lo[0]);
  b = __builtin_bit_cast(double, (unsigned long long)hi[1] << 32 | lo[1]);
}
__device__ __forceinline__ void swap_rows32(float &a, float &b)
{
  const auto r = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(unsigned, a), __builtin_bit_cast(unsigned, b),
                                                  false, false);
  a = __builtin_bit_cast(float, (unsigned)r[0]);
  b = __builtin_bit_cast(float, (unsigned)r[1]);
}
__device__ __forceinline__ void swap_rows16(float &a, float &b)
{
  const auto r = __builtin_amdgcn_permlane16_swap(__builtin_bit_cast(unsigned, a), __builtin_bit_cast(unsigned, b),
                                                  false, false);
  a = __builtin_bit_cast(float, (unsigned)r[0]);
  b = __builtin_bit_cast(float, (unsigned)r[1]);
}

// Transpose of 4 registers x 4 lane rows (16 lanes each): on return row s of
// r[g] holds what row g of r[s] held.  v_permlane32_swap(a, b) exchanges the
// upper half of a with the lower half of b, v_permlane16_swap(a, b) the odd
// rows of a with the even rows of b (lane layouts measured with an in-bounds
// probe), so swaps of (0,2), (1,3) and then (0,1), (2,3) transpose.
template <typename R> __device__ __forceinline__ void transpose_rows4(R &r0, R &r1, R &r2, R &r3)
{
  swap_rows32(r0, r2);
  swap_rows32(r1, r3);
  swap_rows16(r0, r1);
  swap_rows16(r2, r3);
}

// Matrix-core traces with the K hand-over in registers (trace_unit 3), for
// passes of at most 4 momentum slots: a workgroup is one wave, which stages
// a 4-site tile of both propagators (register prefetch of the next), forms K
// of all four sites (the 16x16x4 trace MFMAs above, lane (g, q) register v
// holding K_s[row(g, v)][q]), turns the four sites' registers into the
// projection's B operand (lane row k = site k) with lane-row transposes
// instead of an LDS round trip and a workgroup barrier, and accumulates all
// 256 columns x (K re, K im) x 8 rows itself with v_mfma_f64_4x4x4f64 (the
// quarter-tile projection of meson_grid_kernel; A = the phase of row
// 4 half + (q & 3) at the site of lane row g).  The 64 accumulators per lane
// live in the matrix-core registers.  Waves are independent: no LDS is shared
// and no barrier is taken.  Partials land in the same layout as
// meson_grid_kernel's, so the finish kernel is shared.
//
// DMA (site-scaled FP32): the tiles move HBM -> LDS directly (dma_tile), the
// next one issued once this step's traces have read the current one; without
// the 40 registers of the prefetch the kernel fits two waves per SIMD.  The
// step's phases and site exponents are read before that load is issued (a
// later read of the dynamic-LDS table or a later global load would wait for
// it), and a ragged tile's padded sites get K = 0.
// The site terms of a wave-kernel step (inside its loop): the site exponent
// sum EX of lane row g's site (site-scaled storage) and the phases AV[half]
// of this lane's A rows at that site.  A macro so that each kernel form can
// place it where its register pressure allows.
#define GRID_SITE_TERMS(EX, AV)                                                                             \
  do {                                                                                                      \
    if constexpr (SCALED) { /* a site past the item reads the last one; its K is 0 */                      \
      const size_t site_ = st + (size_t)g < last ? st + (size_t)g : last - 1;                               \
      EX = (int)R.exp_i[site_] + (int)R.exp_j[site_];                                                       \
    }                                                                                                       \
    int xr_ = x0 + g, y_ = y0, z_ = z0;                                                                     \
    while (xr_ >= Lr) {                                                                                     \
      xr_ -= Lr;                                                                                            \
      if (++y_ == Ly) {                                                                                     \
        y_ = 0;                                                                                             \
        ++z_;                                                                                               \
      }                                                                                                     \
    }                                                                                                       \
    z_ = z_ < Lz ? z_ : 0;                                                                                  \
    const int x_ = R.order ? 2 * xr_ + ((pbase + y_ + z_) & 1) : xr_;                                       \
    for (int hf_ = 0; hf_ < 2; ++hf_) {                                                                     \
      const int pr_ = prow4[hf_];                                                                           \
      const double2 ph_ = cmul(cmul(tab[pr_ * (Lx + 1) + x_], tab[offy + pr_ * (Ly + 1) + y_]),           \
                               tab[offz + pr_ * (Lz + 1) + z_]);                                            \
      AV[hf_] = live4[hf_] ? ((q & 1) ? ph_.y : ph_.x) : 0.0;                                               \
    }                                                                                                       \
  } while (0)

template <int CTRL> __device__ __forceinline__ float dpp_move(float x)
{
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), CTRL, 0xf, 0xf, false));
}

// Maximum over each 16-lane row, returned to every lane of the row: DPP
// quad_perm [1,0,3,2] and [2,3,0,1], row_half_mirror, row_mirror.
__device__ __forceinline__ float row_max16(float v)
{
  v = fmaxf(v, dpp_move<0xb1>(v));
  v = fmaxf(v, dpp_move<0x4e>(v));
  v = fmaxf(v, dpp_move<0x141>(v));
  return fmaxf(v, dpp_move<0x140>(v));
}

// Plain FP32 storage in the wave-local form: for lane row g, e = ilogb of
// the largest entry of site g of a staged 4-site tile (the per-site exponent
// that FP32 site-scaled storage holds: 0 for an all-zero site, within FP32's
// normal range).
__device__ __forceinline__ int site_exponent(const float2 *tile, int lane)
{
  constexpr int SLOTS = SITE_ELEMS / 2; // 16-byte slots per site
  constexpr int M = (SLOTS + 15) / 16;
  const int g = lane >> 4, q = lane & 15;
  const float4 *site = reinterpret_cast<const float4 *>(tile + g * SITE_ELEMS);
  float mx = 0.0f;
#pragma unroll
  for (int m = 0; m < M; ++m)
    if (m + 1 < M || q + 16 * m < SLOTS) {
      const float4 v = site[q + 16 * m];
      mx = fmaxf(mx, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))));
    }
  mx = row_max16(mx);
  const int e = mx > 0.0f ? ilogbf(mx) : 0;
  return e < -126 ? -126 : (e > 126 ? 126 : e);
}

// 2^-e of site sk (held by lane row sk) for the whole wave
__device__ __forceinline__ float site_scale(float sc, int sk)
{
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, sc), 16 * sk));
}

// AUTO: plain FP32 storage, scaled in the kernel.  Once a tile has landed,
// lane row g finds the exponents of site g (site_exponent); the traces
// scale every operand of a site by 2^-e of its propagator as they read it,
// so their FP32 products stay far above the underflow threshold however far
// the propagators have decayed, and the operands are the values FP32
// site-scaled storage holds (exact powers of two: the results equal that
// form's bit for bit).  The results are scaled back in FP64 as for
// site-scaled storage.
template <typename Real, bool SCALED, bool DMA, bool AUTO = false>
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(DMA ? 2 : 1, DMA ? 2 : 1))) void
meson_grid_wave_kernel(MesonGridArgs a)
{
  static_assert(!AUTO || (DMA && !SCALED && sizeof(Real) == 4), "in-kernel scaling is a form of the FP32 DMA kernel");
  using C = cplx<Real>;
  using T = Tile<C>;
  using KAcc = typename std::conditional<sizeof(Real) == 8, dbl4, flt4>::type;
  constexpr int SL = DMA ? SITE_ELEMS : T::LDS_ELEMS; // site stride of the staged tiles
  __shared__ C lds[DMA ? 1 : 2][DMA ? 1 : SITES_PER_WAVE * SL]; // [propagator], register-staged
  __shared__ C dma_i[DMA ? T::DMA_ELEMS : 1], dma_j[DMA ? T::DMA_ELEMS : 1];
  C *const tile_i = DMA ? dma_i : lds[0];
  C *const tile_j = DMA ? dma_j : lds[DMA ? 0 : 1];
  extern __shared__ double2 tab[];
  const int lane = threadIdx.x;
  const int g = lane >> 4, q = lane & 15;
  // Q4: this lane's A rows 4 half + (q & 3): slot 2 half + ((q >> 1) & 1), part q & 1
  bool live4[2];
  int prow4[2];
#pragma unroll
  for (int hf = 0; hf < 2; ++hf) {
    const int pr = 2 * hf + ((q >> 1) & 1);
    live4[hf] = pr < a.np;
    prow4[hf] = live4[hf] ? pr : 0;
  }
  int rtab = -1; // request whose phase tables are in LDS

  for (size_t item = blockIdx.x; item < a.total_items; item += gridDim.x) {
    int r = 0;
    while (r + 1 < a.nreq && item >= a.req[r + 1].item0) ++r;
    const MesonGridReq &R = a.req[r];
    const size_t local = item - R.item0;
    const size_t slice = local / R.chunks_per_slice;
    const size_t c = local - slice * R.chunks_per_slice;
    const size_t s0 = slice * R.slice_sites;
    const size_t first = s0 + c * (size_t)a.chunk;
    size_t last = first + a.chunk;
    last = last < s0 + R.slice_sites ? last : s0 + R.slice_sites;

    const int Lx = R.dims[0], Ly = R.dims[1], Lz = R.dims[2];
    const int offy = a.np * (Lx + 1), offz = a.np * (Lx + Ly + 2);
    if (r != rtab) { // the tables depend on the request only
      wave_sync();   // the previous item's readers of tab are done
      const int ntab = a.np * (Lx + Ly + Lz + 3);
      for (int e = lane; e < ntab; e += WAVE) {
        const int mu = e < offy ? 0 : (e < offz ? 1 : 2);
        const int L = R.dims[mu], G = R.gdims[mu];
        const int rel = e - (mu == 0 ? 0 : (mu == 1 ? offy : offz));
        const int p = rel / (L + 1), k = rel - p * (L + 1);
        const int m = nonneg_mod(a.mom[p][mu], G);
        double sn, cs;
        sincospi(2.0 * (double)((m * ((k + R.origin[mu]) % G)) % G) / (double)G, &sn, &cs);
        tab[e] = make_double2(cs, -sn);
      }
      rtab = r;
    }

    // coordinates of the step's first site, advanced incrementally (as in
    // meson_grid_kernel)
    const int lt = R.dims[3];
    const int par = (int)(slice / lt), tsl = (int)(slice - (size_t)par * lt);
    const int Lr = R.order ? Lx >> 1 : Lx;
    int x0, y0, z0;
    {
      const unsigned r0 = (unsigned)(c * (size_t)a.chunk);
      x0 = (int)(r0 % (unsigned)Lr);
      const unsigned rest = r0 / (unsigned)Lr;
      y0 = (int)(rest % (unsigned)Ly);
      z0 = (int)(rest / (unsigned)Ly);
    }
    const int pbase = par + tsl;

    double acc[16][2][2]; // [row ik of K][K re, K im][half]
#pragma unroll
    for (int ik = 0; ik < 16; ++ik)
#pragma unroll
      for (int u = 0; u < 2; ++u) acc[ik][u][0] = acc[ik][u][1] = 0.0;

    const C *P0 = static_cast<const C *>(R.propag_i), *P1 = static_cast<const C *>(R.propag_j);
    auto tile_n = [&](size_t st) { // sites of the tile at st (past the item's end read as zero)
      const size_t left = last - st;
      return left < (size_t)SITES_PER_WAVE ? (int)left : SITES_PER_WAVE;
    };
    T bi, bj;
    if constexpr (DMA) {
      dma_wait(); // the previous item's last (empty) tile load has landed
      dma_tile(dma_i, P0, first, tile_n(first), lane);
      dma_tile(dma_j, P1, first, tile_n(first), lane);
    } else {
      bi.load(P0, first, tile_n(first), lane);
      bj.load(P1, first, tile_n(first), lane);
    }
    for (size_t st = first; st < last; st += SITES_PER_WAVE) {
      const size_t next = st + SITES_PER_WAVE;
      const int nnext = next < last ? tile_n(next) : 0; // an empty range past the item
      int ex_auto = 0;               // AUTO: lane row g's site exponents, P_i's plus P_j's
      float sci = 1.0f, scj = 1.0f; // AUTO: 2^-exponent of lane row g's site
      if constexpr (DMA) {
        dma_wait(); // this tile's loads, the only ones in flight here
        if constexpr (AUTO) {
          const int ei = site_exponent(reinterpret_cast<const float2 *>(dma_i), lane);
          const int ej = site_exponent(reinterpret_cast<const float2 *>(dma_j), lane);
          ex_auto = ei + ej;
          sci = ldexpf(1.0f, -ei);
          scj = ldexpf(1.0f, -ej);
        }
      } else {
        wave_sync(); // the previous step's readers of the tile are done
        bi.store(lds[0], lane);
        bj.store(lds[1], lane);
        bi.load(P0, next < last ? next : st, nnext, lane);
        bj.load(P1, next < last ? next : st, nnext, lane);
        wave_sync();
      }

      // traces of the four sites
      KAcc Kr[SITES_PER_WAVE], Ki[SITES_PER_WAVE];
#pragma unroll
      for (int sk = 0; sk < SITES_PER_WAVE; ++sk) {
        const C *xs = tile_i + sk * SL, *ys = tile_j + sk * SL;
        KAcc kr = KAcc{0, 0, 0, 0}, ki = KAcc{0, 0, 0, 0}, t3 = KAcc{0, 0, 0, 0};
        Real si = 1, sj = 1; // AUTO: this site's scales
        if constexpr (AUTO) {
          si = site_scale(sci, sk);
          sj = site_scale(scj, sk);
        }
#pragma unroll
        for (int ch = 0; ch < 2; ++ch) {
          const int cc = 4 * ch + g;
          const C xv = AUTO ? cscale(xs[q * 9 + cc], si) : xs[q * 9 + cc];
          const C yv = AUTO ? cscale(ys[q * 9 + cc], sj) : ys[q * 9 + cc];
          if constexpr (sizeof(Real) == 8) {
            kr = __builtin_amdgcn_mfma_f64_16x16x4f64(xv.x, yv.x, kr, 0, 0, 0);
            ki = __builtin_amdgcn_mfma_f64_16x16x4f64(xv.y, yv.y, ki, 0, 0, 0);
            t3 = __builtin_amdgcn_mfma_f64_16x16x4f64(xv.x + xv.y, yv.x - yv.y, t3, 0, 0, 0);
          } else {
            kr = __builtin_amdgcn_mfma_f32_16x16x4f32(xv.x, yv.x, kr, 0, 0, 0);
            kr = __builtin_amdgcn_mfma_f32_16x16x4f32(xv.y, yv.y, kr, 0, 0, 0);
            ki = __builtin_amdgcn_mfma_f32_16x16x4f32(xv.y, yv.x, ki, 0, 0, 0);
            ki = __builtin_amdgcn_mfma_f32_16x16x4f32(-xv.x, yv.y, ki, 0, 0, 0);
          }
        }
        if constexpr (sizeof(Real) == 8) {
          const KAcc t1 = kr, t2 = ki;
          kr = t1 + t2;
          ki = t3 - t1 + t2;
        }
        const C yv = AUTO ? cscale(ys[q * 9 + 8], sj) : ys[q * 9 + 8]; // colour pair 8 on the vector ALUs
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = sizeof(Real) == 8 ? g + 4 * v : 4 * g + v;
          const C t = cfma_conj(cmake<C>(kr[v], ki[v]), AUTO ? cscale(xs[row * 9 + 8], si) : xs[row * 9 + 8], yv);
          kr[v] = t.x;
          ki[v] = t.y;
        }
        if constexpr (DMA) { // padded sites of a ragged tile (zero-filled by dma_tile; kept exact here)
          if (st + (size_t)sk >= last) kr = ki = KAcc{0, 0, 0, 0};
        }
        Kr[sk] = kr;
        Ki[sk] = ki;
      }
      // this lane row's site exponents, coordinates and phases: with DMA read
      // before the next tile's load is issued (a dynamic-LDS or global read
      // after it would wait for it), otherwise after the transposes (fewer
      // registers live across them)
      int ex = 0;
      double av4[2];
      if constexpr (DMA) {
        GRID_SITE_TERMS(ex, av4);
      }
      if constexpr (AUTO) ex = ex_auto; // lane row g's site
      if constexpr (DMA) { // the tile has been read: the next one into its place
        wave_sync();
        // past the item an empty range (its zeros land before the next
        // item's first tile is loaded: that load waits for them)
        dma_tile(dma_i, P0, next < last ? next : st, nnext, lane);
        dma_tile(dma_j, P1, next < last ? next : st, nnext, lane);
      }
      // lane row g of site register sk holds rows row(g, v) of K_sk; after the
      // transpose, lane row k of register g holds row row(g, v) of K_k
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        Real r0 = Kr[0][v], r1 = Kr[1][v], r2 = Kr[2][v], r3 = Kr[3][v];
        transpose_rows4(r0, r1, r2, r3);
        Kr[0][v] = r0, Kr[1][v] = r1, Kr[2][v] = r2, Kr[3][v] = r3;
        Real i0 = Ki[0][v], i1 = Ki[1][v], i2 = Ki[2][v], i3 = Ki[3][v];
        transpose_rows4(i0, i1, i2, i3);
        Ki[0][v] = i0, Ki[1][v] = i1, Ki[2][v] = i2, Ki[3][v] = i3;
      }
      if constexpr (!DMA) {
        GRID_SITE_TERMS(ex, av4);
      }
#pragma unroll
      for (int gg = 0; gg < 4; ++gg)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int ik = sizeof(Real) == 8 ? gg + 4 * v : 4 * gg + v;
          double re = (double)Kr[gg][v], im = (double)Ki[gg][v];
          if constexpr (SCALED || AUTO) {
            re = ldexp(re, ex);
            im = ldexp(im, ex);
          }
#pragma unroll
          for (int hf = 0; hf < 2; ++hf) {
            acc[ik][0][hf] = __builtin_amdgcn_mfma_f64_4x4x4f64(av4[hf], re, acc[ik][0][hf], 0, 0, 0);
            acc[ik][1][hf] = __builtin_amdgcn_mfma_f64_4x4x4f64(av4[hf], im, acc[ik][1][hf], 0, 0, 0);
          }
        }
      x0 += SITES_PER_WAVE;
      while (x0 >= Lr) {
        x0 -= Lr;
        if (++y0 == Ly) {
          y0 = 0;
          ++z0;
        }
      }
    }
    // D[b][i][j] of v_mfma_f64_4x4x4f64 in lane 16 i + 4 b + j: lane (g, q)
    // holds row 4 half + g of column q = jl of row ik; partial column
    // (k l) * 16 + (i j) for ik = 4 i + k, jl = 4 j + l
    double *dst = a.partial + item * (size_t)a.partial_tiles * GRID_PARTIAL_TILE;
#pragma unroll
    for (int ik = 0; ik < 16; ++ik)
#pragma unroll
      for (int hf = 0; hf < 2; ++hf) {
        const int row = 4 * hf + g;
        const int col = ((ik & 3) * 4 + (q & 3)) * 16 + (ik >> 2) * 4 + (q >> 2);
        dst[(row * 2 + 0) * GRID_COLS + col] = acc[ik][0][hf];
        dst[(row * 2 + 1) * GRID_COLS + col] = acc[ik][1][hf];
      }
  }
}

#undef GRID_SITE_TERMS

// The kernel a pass runs, handed to f(kernel, threads per workgroup): the
// launch and the occupancy query below select it the same way.
template <typename Real, bool SCALED, bool KMFMA, int RT, bool Q4 = false, typename F>
static void with_grid_rt(int tile_sites, F &&f)
{
  if (tile_sites == 8)
    f(meson_grid_kernel<Real, 8, SCALED, KMFMA, RT, Q4>, GRID_THREADS);
  else
    f(meson_grid_kernel<Real, 4, SCALED, KMFMA, RT, Q4>, GRID_THREADS);
}

template <typename Real, bool SCALED, bool KMFMA, typename F>
static void with_grid_t(const MesonGridArgs &a, int tile_sites, F &&f)
{
  if (a.np > GRID_MOM_PER_TILE)
    with_grid_rt<Real, SCALED, KMFMA, 2>(tile_sites, f);
  else if (KMFMA && a.np <= GRID_MOM_PER_TILE / 2)
    with_grid_rt<Real, SCALED, KMFMA, 1, KMFMA>(tile_sites, f);
  else
    with_grid_rt<Real, SCALED, KMFMA, 1>(tile_sites, f);
}

template <typename F> static void with_grid_kernel(const MesonGridArgs &a, int storage, int unit, int tile_sites, F &&f)
{
  const bool mfma_traces = unit >= 2;
  if (grid_wave_pass(unit, a.np)) { // wave-local K; other passes take the shared-tile form
    if (storage == GRID_FP32_SCALED)
      f(meson_grid_wave_kernel<float, true, true>, WAVE);
    else if (storage == GRID_FP32)
      f(meson_grid_wave_kernel<float, false, true, true>, WAVE);
    else
      f(meson_grid_wave_kernel<double, false, false>, WAVE);
    return;
  }
  if (storage == GRID_FP32_SCALED) {
    if (mfma_traces)
      with_grid_t<float, true, true>(a, tile_sites, f);
    else
      with_grid_t<float, true, false>(a, tile_sites, f);
  } else if (storage == GRID_FP32) {
    with_grid_t<float, false, false>(a, tile_sites, f);
  } else {
    if (mfma_traces)
      with_grid_t<double, false, true>(a, tile_sites, f);
    else
      with_grid_t<double, false, false>(a, tile_sites, f);
  }
}

void launch_meson_grid(const MesonGridArgs &a, int storage, int unit, int tile_sites, size_t blocks,
                       size_t tab_entries, hipStream_t stream)
{
  const size_t shmem = tab_entries * sizeof(double2);
  with_grid_kernel(a, storage, unit, tile_sites, [&](auto kernel, int threads) {
    hipLaunchKernelGGL(kernel, dim3((unsigned)blocks), dim3(threads), shmem, stream, a);
  });
}

int meson_grid_resident_blocks(const MesonGridArgs &a, int storage, int unit, int tile_sites, size_t tab_entries)
{
  const size_t shmem = tab_entries * sizeof(double2);
  int n = 0;
  with_grid_kernel(a, storage, unit, tile_sites, [&](auto kernel, int threads) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, threads, shmem) != hipSuccess) n = 0;
  });
  return n;
}

__device__ __forceinline__ void nsum(double &s, double &c, double v)
{
  const double t = s + v;
  c += (fabs(s) >= fabs(v)) ? ((s - t) + v) : ((v - t) + s);
  s = t;
}

// One block per (request, timeslice, momentum slot), 512 threads = two chunk
// groups x one thread per K column: each thread sums every other work-item
// partial of the timeslice in item order (Neumaier compensated), the two
// groups are combined in a fixed order (so the result is bit-reproducible),
// and the rows A = sum cos(k.x) K, B = -sum sin(k.x) K give the projected K of
// momentum +k (A + iB) and of -k (A - iB); thread (g1, g2) applies the gamma
// pair to each requested sign.  Output out[g1][g2][momentum][t].
constexpr int FINISH_GROUPS = 2;
constexpr int FINISH_UNROLL = 4;

__global__ __launch_bounds__(FINISH_GROUPS * GRID_COLS) void meson_grid_finish_kernel(MesonGridFinish f)
{
  __shared__ double part[FINISH_GROUPS][4][2][GRID_COLS];
  __shared__ double2 kp[2][GRID_COLS];
  const int col = threadIdx.x % GRID_COLS, cg = threadIdx.x / GRID_COLS;
  const int p = blockIdx.y;
  int r = 0;
  while (r + 1 < f.nreq && (int)blockIdx.x >= f.req[r + 1].block0) ++r;
  const MesonGridOut &R = f.req[r];
  const int t = (int)blockIdx.x - R.block0;
  const int nslices = R.order ? 2 : 1;
  double s[4] = {0, 0, 0, 0}, cc[4] = {0, 0, 0, 0};
  for (int par = 0; par < nslices; ++par) {
    const size_t slice = (size_t)par * R.lt + t;
    const size_t stride = (size_t)f.partial_tiles * GRID_PARTIAL_TILE;
    const double *base = f.partial + (R.item0 + slice * R.chunks_per_slice) * stride + (size_t)(4 * p) * GRID_COLS + col;
    // FINISH_UNROLL chunks' loads in flight at once, summed in chunk order
    for (int ch0 = cg; ch0 < R.chunks_per_slice; ch0 += FINISH_GROUPS * FINISH_UNROLL) {
      double v[FINISH_UNROLL][4];
#pragma unroll
      for (int k = 0; k < FINISH_UNROLL; ++k) {
        const int ch = ch0 + k * FINISH_GROUPS;
        const double *src = base + (size_t)(ch < R.chunks_per_slice ? ch : ch0) * stride;
        // row 2p: cos weight, row 2p + 1: -sin weight; each with K re / K im
#pragma unroll
        for (int u = 0; u < 4; ++u) v[k][u] = src[u * GRID_COLS];
      }
#pragma unroll
      for (int k = 0; k < FINISH_UNROLL; ++k)
        if (ch0 + k * FINISH_GROUPS < R.chunks_per_slice)
#pragma unroll
          for (int u = 0; u < 4; ++u) nsum(s[u], cc[u], v[k][u]);
    }
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    part[cg][u][0][col] = s[u];
    part[cg][u][1][col] = cc[u];
  }
  __syncthreads();
  if (cg == 0) {
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      double S = 0, Cc = 0;
      for (int gq = 0; gq < FINISH_GROUPS; ++gq) {
        nsum(S, Cc, part[gq][u][0][col]);
        Cc += part[gq][u][1][col];
      }
      v[u] = S + Cc;
    }
    // A = (v0, v1), B = (v2, v3): A + iB and A - iB
    kp[0][col] = make_double2(v[0] - v[3], v[1] + v[2]);
    kp[1][col] = make_double2(v[0] + v[3], v[1] - v[2]);
  }
  __syncthreads();
  if (cg == 0) {
    const int g1 = col >> 4, g2 = col & 15;
    const SparseSpin a1 = f.g5t[g1], a2 = f.g5t[g2];
    double2 *out = static_cast<double2 *>(R.out);
#pragma unroll
    for (int sgn = 0; sgn < 2; ++sgn) {
      const int po = sgn ? f.out_minus[p] : f.out_plus[p];
      if (po < 0) continue;
      double2 acc = make_double2(0.0, 0.0);
      for (int ii = 0; ii < 4; ++ii)
        for (int l = 0; l < 4; ++l) {
          const int jj = a1.col[ii], kk = a2.col[l];
          const double2 wgt = cmul(make_double2(a1.re[ii], a1.im[ii]), make_double2(a2.re[l], a2.im[l]));
          acc = cfma(acc, wgt, kp[sgn][(kk * 4 + l) * 16 + ii * 4 + jj]);
        }
      out[((size_t)(g1 * 16 + g2) * f.np_total + po) * R.lt + t] = acc;
    }
  }
}

void launch_meson_grid_finish(const MesonGridFinish &f, int nblocks, hipStream_t stream)
{
  hipLaunchKernelGGL(meson_grid_finish_kernel, dim3(nblocks, f.np), dim3(FINISH_GROUPS * GRID_COLS), 0, stream, f);
}

} // namespace contract_amd
