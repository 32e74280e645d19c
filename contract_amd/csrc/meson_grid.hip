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

#include "meson_grid.hpp"

namespace contract_amd
{

typedef double dbl4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void grid_coords(size_t s, const int (&dims)[4], int order, int &x, int &y, int &z)
{
  if (order == 0) {
    x = (int)(s % dims[0]);
    s /= dims[0];
    y = (int)(s % dims[1]);
    s /= dims[1];
    z = (int)(s % dims[2]);
  } else {
    const size_t half = (size_t)dims[0] * dims[1] * dims[2] * dims[3] / 2;
    const int parity = (int)(s / half);
    size_t r = s - (size_t)parity * half;
    const int lxh = dims[0] >> 1;
    const int xh = (int)(r % lxh);
    r /= lxh;
    y = (int)(r % dims[1]);
    r /= dims[1];
    z = (int)(r % dims[2]);
    const int t = (int)(r / dims[2]);
    x = 2 * xh + ((parity + y + z + t) & 1);
  }
}

__device__ __forceinline__ int nonneg_mod(int a, int L)
{
  const int r = a % L;
  return r < 0 ? r + L : r;
}

// Scale one staged FP32 site (16 lanes, 9 entries each) by an exact power of
// two so its largest entry lies in [1, 2); returns the exponent.  The meson
// traces are bilinear, so K of the scaled blocks times 2^(e_i + e_j) is exact
// up to FP32 rounding and never underflows.
__device__ __forceinline__ int grid_normalise(float2 *site, int q)
{
  float2 *p = site + q * 9;
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

template <typename Real, int NS>
__global__ __launch_bounds__(GRID_THREADS)
__attribute__((amdgpu_waves_per_eu((sizeof(Real) == 8 && NS == 8) ? 1 : 2))) void meson_grid_kernel(
    MesonGridArgs a)
{
  static_assert(NS % SITES_PER_WAVE == 0, "tile is a whole number of MFMA site steps");
  using C = cplx<Real>;
  using T = Tile<C>;
  constexpr int SL = T::LDS_ELEMS;           // padded site stride (complex entries)
  constexpr int CH = NS * T::SITE16;         // 16-byte chunks per propagator tile
  constexpr int PER_THREAD = (2 * CH + GRID_THREADS - 1) / GRID_THREADS;
  __shared__ C lds[2][NS * SL];
  __shared__ int sexp[2][NS];
  extern __shared__ double2 tab[];
  const int tid = threadIdx.x;
  const int w = tid >> 6, lane = tid & 63;
  const int g = lane >> 4, q = lane & 15;
  const int i = q >> 2, j = q & 3;

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
    const int ntab = R.dims[0] + R.dims[1] + R.dims[2];
    for (int e = tid; e < ntab; e += GRID_THREADS) {
      const int L = e < R.dims[0] ? R.dims[0] : (e < R.dims[0] + R.dims[1] ? R.dims[1] : R.dims[2]);
      const int k = e < R.dims[0] ? e : (e < R.dims[0] + R.dims[1] ? e - R.dims[0] : e - R.dims[0] - R.dims[1]);
      double sn, cs;
      sincospi(2.0 * (double)k / (double)L, &sn, &cs);
      tab[e] = make_double2(cs, -sn); // exp(-2 pi i k / L)
    }
    // this lane's MFMA row: momentum q >> 1, phase part q & 1 (0 re, 1 im)
    const int pr = q >> 1;
    const bool live = pr < a.np;
    const int mx = live ? nonneg_mod(a.mom[pr][0], R.dims[0]) : 0;
    const int my = live ? nonneg_mod(a.mom[pr][1], R.dims[1]) : 0;
    const int mz = live ? nonneg_mod(a.mom[pr][2], R.dims[2]) : 0;

    dbl4 acc_re[4], acc_im[4]; // column tiles l = 0..3 of k = w; B = K re / K im
#pragma unroll
    for (int l = 0; l < 4; ++l) acc_re[l] = acc_im[l] = dbl4{0.0, 0.0, 0.0, 0.0};

    // the workgroup stages NS sites of both propagators per step; the next
    // step's chunks are in flight in registers while this one is computed
    const C *P0 = static_cast<const C *>(R.propag_i), *P1 = static_cast<const C *>(R.propag_j);
    uint4 buf[PER_THREAD];
    auto load = [&](size_t st) {
      const size_t left = last - st;
      const int total = (left < (size_t)NS ? (int)left : NS) * T::SITE16;
#pragma unroll
      for (int it = 0; it < PER_THREAD; ++it) {
        const int e = it * GRID_THREADS + tid;
        if ((2 * CH) % GRID_THREADS == 0 || e < 2 * CH) {
          const int p = e >= CH, ec = e - p * CH;
          const uint4 *base = reinterpret_cast<const uint4 *>((p ? P1 : P0) + st * SITE_ELEMS);
          buf[it] = (ec < total) ? base[ec] : make_uint4(0, 0, 0, 0);
        }
      }
    };
    auto store = [&]() {
#pragma unroll
      for (int it = 0; it < PER_THREAD; ++it) {
        const int e = it * GRID_THREADS + tid;
        if ((2 * CH) % GRID_THREADS == 0 || e < 2 * CH) {
          const int p = e >= CH, ec = e - p * CH;
          const int s = ec / T::SITE16;
          reinterpret_cast<uint4 *>(lds[p])[s * T::LDS16 + (ec - s * T::SITE16)] = buf[it];
        }
      }
    };
    if (first < last) load(first);
    for (size_t st = first; st < last; st += NS) {
      __syncthreads(); // readers of the previous tile are done
      store();
      if (st + NS < last) load(st + NS);
      __syncthreads();
      if constexpr (sizeof(Real) == 4) {
        for (int u = w; u < NS / 2; u += 4) { // (propagator, 4-site group) per wave
          const int e1 = grid_normalise(reinterpret_cast<float2 *>(lds[u & 1]) + ((u >> 1) * 4 + g) * SL, q);
          if (q == 0) sexp[u & 1][(u >> 1) * 4 + g] = e1;
        }
        __syncthreads();
      }
#pragma unroll
      for (int h = 0; h < NS / SITES_PER_WAVE; ++h) {
        const int ls = h * SITES_PER_WAVE + g; // this lane's site in the tile
        // phase of this lane's MFMA row at its site (padded sites carry K = 0)
        const size_t site = st + ls < last ? st + ls : st;
        int x, y, z;
        grid_coords(site, R.dims, R.order, x, y, z);
        double av = 0.0;
        if (live) {
          const double2 ph = cmul(cmul(tab[(mx * x) % R.dims[0]], tab[R.dims[0] + (my * y) % R.dims[1]]),
                                  tab[R.dims[0] + R.dims[1] + (mz * z) % R.dims[2]]);
          av = (q & 1) ? ph.y : ph.x;
        }
        const C *si = lds[0] + ls * SL, *sj = lds[1] + ls * SL;
        int ex = 0;
        if constexpr (sizeof(Real) == 4) ex = sexp[0][ls] + sexp[1][ls];
        C xb[9], yb[9];
        load_block(xb, si, i, w);
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
          acc_re[l] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, kr, acc_re[l], 0, 0, 0);
          acc_im[l] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, ki, acc_im[l], 0, 0, 0);
        }
      }
    }
    // v_mfma_f64_16x16x4f64: lane (g, q) register v holds D[g + 4v][q]
    double *dst = a.partial + item * (size_t)GRID_PARTIAL;
#pragma unroll
    for (int l = 0; l < 4; ++l)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = g + 4 * v;
        const int col = (w * 4 + l) * 16 + q;
        dst[(row * 2 + 0) * GRID_COLS + col] = acc_re[l][v];
        dst[(row * 2 + 1) * GRID_COLS + col] = acc_im[l][v];
      }
  }
}

template <typename Real>
static void launch_grid_t(const MesonGridArgs &a, int tile_sites, size_t blocks, size_t shmem, hipStream_t stream)
{
  if (tile_sites == 8)
    hipLaunchKernelGGL((meson_grid_kernel<Real, 8>), dim3((unsigned)blocks), dim3(GRID_THREADS), shmem, stream, a);
  else
    hipLaunchKernelGGL((meson_grid_kernel<Real, 4>), dim3((unsigned)blocks), dim3(GRID_THREADS), shmem, stream, a);
}

void launch_meson_grid(const MesonGridArgs &a, bool fp32, int tile_sites, size_t blocks, size_t tab_entries,
                       hipStream_t stream)
{
  const size_t shmem = tab_entries * sizeof(double2);
  if (fp32)
    launch_grid_t<float>(a, tile_sites, blocks, shmem, stream);
  else
    launch_grid_t<double>(a, tile_sites, blocks, shmem, stream);
}

__device__ __forceinline__ void nsum(double &s, double &c, double v)
{
  const double t = s + v;
  c += (fabs(s) >= fabs(v)) ? ((s - t) + v) : ((v - t) + s);
  s = t;
}

// One block of 256 threads per (request, timeslice), one thread per K column:
// sum the timeslice's work-item partials in item order (Neumaier compensated,
// so the result is bit-reproducible), combine the (phase part, K part) rows
// into the complex momentum-projected K, then apply the gamma pairs, thread
// (g1, g2).  Output out[g1][g2][p_base + p][t].
__global__ __launch_bounds__(256) void meson_grid_finish_kernel(MesonGridFinish f)
{
  __shared__ double2 kp[GRID_MOM_PER_PASS][GRID_COLS];
  const int col = threadIdx.x;
  int r = 0;
  while (r + 1 < f.nreq && (int)blockIdx.x >= f.req[r + 1].block0) ++r;
  const MesonGridOut &R = f.req[r];
  const int t = (int)blockIdx.x - R.block0;
  const int nslices = R.order ? 2 : 1;
  for (int p = 0; p < f.np; ++p) {
    double s[4] = {0, 0, 0, 0}, cc[4] = {0, 0, 0, 0};
    for (int par = 0; par < nslices; ++par) {
      const size_t slice = (size_t)par * R.lt + t;
      for (int ch = 0; ch < R.chunks_per_slice; ++ch) {
        const double *src = f.partial + (R.item0 + slice * R.chunks_per_slice + ch) * (size_t)GRID_PARTIAL;
        // row 2p: phase re, row 2p + 1: phase im; each with K re / K im
#pragma unroll
        for (int u = 0; u < 4; ++u) nsum(s[u], cc[u], src[(4 * p + u) * GRID_COLS + col]);
      }
    }
    const double rr = s[0] + cc[0], ri = s[1] + cc[1], ir = s[2] + cc[2], ii = s[3] + cc[3];
    kp[p][col] = make_double2(rr - ii, ri + ir); // (ph_re + i ph_im)(K_re + i K_im)
  }
  __syncthreads();
  const int g1 = col >> 4, g2 = col & 15;
  const SparseSpin a1 = f.g5t[g1], a2 = f.g5t[g2];
  double2 *out = static_cast<double2 *>(R.out);
  for (int p = 0; p < f.np; ++p) {
    double2 acc = make_double2(0.0, 0.0);
    for (int ii = 0; ii < 4; ++ii)
      for (int l = 0; l < 4; ++l) {
        const int jj = a1.col[ii], kk = a2.col[l];
        const double2 wgt = cmul(make_double2(a1.re[ii], a1.im[ii]), make_double2(a2.re[l], a2.im[l]));
        acc = cfma(acc, wgt, kp[p][(kk * 4 + l) * 16 + ii * 4 + jj]);
      }
    out[((size_t)(g1 * 16 + g2) * f.np_total + f.p_base + p) * R.lt + t] = acc;
  }
}

void launch_meson_grid_finish(const MesonGridFinish &f, int nblocks, hipStream_t stream)
{
  hipLaunchKernelGGL(meson_grid_finish_kernel, dim3(nblocks), dim3(256), 0, stream, f);
}

} // namespace contract_amd
