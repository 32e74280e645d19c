// project.hip -- phase tables and the deterministic per-timeslice reduction
// that closes every fused (timeslice x momentum) contraction.

#include "baryon.hpp"
#include "device.hpp"

namespace contract_amd
{

// tab = [exp(-2 pi i k/Lx), k < Lx | same for Ly | same for Lz]
__global__ void phase_table_kernel(double2 *tab, int lx, int ly, int lz)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  int L, k;
  if (e < lx) {
    L = lx;
    k = e;
  } else if (e < lx + ly) {
    L = ly;
    k = e - lx;
  } else if (e < lx + ly + lz) {
    L = lz;
    k = e - lx - ly;
  } else {
    return;
  }
  double s, c;
  sincospi(2.0 * (double)k / (double)L, &s, &c);
  tab[e] = make_double2(c, -s);
}

void launch_phase_table(double2 *tab, const int dims[3], hipStream_t stream)
{
  const int n = dims[0] + dims[1] + dims[2];
  hipLaunchKernelGGL(phase_table_kernel, dim3((n + 63) / 64), dim3(64), 0, stream, tab, dims[0], dims[1], dims[2]);
}

__device__ __forceinline__ void neumaier(double &sum, double &comp, double v)
{
  const double t = sum + v;
  comp += (fabs(sum) >= fabs(v)) ? ((sum - t) + v) : ((v - t) + sum);
  sum = t;
}

// One block per (momentum p, timeslice t).  Lanes 0..15 sum the work-item
// partials of component q over the timeslice's items (both parities for
// even-odd order) in item order, with Neumaier compensation; lanes then apply
// the projectors W_w[q].
__global__ __launch_bounds__(64) void project_finish_kernel(FinishArgs f)
{
  __shared__ double2 rp[16];
  const int p = blockIdx.x, t = blockIdx.y, lane = threadIdx.x;
  if (lane < 16) {
    double sr = 0, cr = 0, si = 0, ci = 0;
    const int nslices = f.order ? 2 : 1;
    for (int par = 0; par < nslices; ++par) {
      const size_t slice = (size_t)par * f.lt + t;
      for (int c = 0; c < f.chunks_per_slice; ++c) {
        const size_t item = slice * f.chunks_per_slice + c;
        const double2 v = f.partial[(item * f.np + p) * 16 + lane];
        neumaier(sr, cr, v.x);
        neumaier(si, ci, v.y);
      }
    }
    rp[lane] = make_double2(sr + cr, si + ci);
  }
  __syncthreads();
  for (int w = lane; w < f.nproj; w += 64) {
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int q = 0; q < 16; ++q) acc = cfma(acc, make_double2(f.w_re[w][q], f.w_im[w][q]), rp[q]);
    f.out[((size_t)(f.proj_base + w) * f.np + p) * f.lt + t] = acc;
  }
}

void launch_project_finish(const FinishArgs &f, hipStream_t stream)
{
  hipLaunchKernelGGL(project_finish_kernel, dim3(f.np, f.lt), dim3(64), 0, stream, f);
}

} // namespace contract_amd
