// project.hip -- the deterministic per-timeslice reduction and projector
// application that close every fused (timeslice x momentum) contraction.

#include "baryon.hpp"
#include "device.hpp"

namespace contract_amd
{

__device__ __forceinline__ void neumaier(double &sum, double &comp, double v)
{
  const double t = sum + v;
  comp += (fabs(sum) >= fabs(v)) ? ((sum - t) + v) : ((v - t) + sum);
  sum = t;
}

// One block of 256 threads per (momentum p, timeslice t): thread (c0, q)
// sums component q over the work items c0, c0 + 16, ... of the timeslice
// (both parities for even-odd order); the 16 column partials are then
// combined in index order.  Every sum is Neumaier compensated and the order
// is fixed, so results are bit-reproducible.
__global__ __launch_bounds__(256) void project_reduce_kernel(ReduceArgs r)
{
  __shared__ double part[16][16][4];
  const int p = blockIdx.x, t = blockIdx.y;
  const int q = threadIdx.x & 15, c0 = threadIdx.x >> 4;
  double sr = 0, cr = 0, si = 0, ci = 0;
  if (q < r.nq) {
    const int nslices = r.order ? 2 : 1;
    for (int par = 0; par < nslices; ++par) {
      const size_t slice = (size_t)par * r.lt + t;
      for (int c = c0; c < r.chunks_per_slice; c += 16) {
        const size_t item = slice * r.chunks_per_slice + c;
        const double2 v = r.partial[(item * r.np + p) * r.nq + q];
        neumaier(sr, cr, v.x);
        neumaier(si, ci, v.y);
      }
    }
  }
  part[c0][q][0] = sr;
  part[c0][q][1] = cr;
  part[c0][q][2] = si;
  part[c0][q][3] = ci;
  __syncthreads();
  if (c0 == 0 && q < r.nq) {
    double Sr = 0, Cr = 0, Si = 0, Ci = 0;
    for (int c = 0; c < 16; ++c) {
      neumaier(Sr, Cr, part[c][q][0]);
      Cr += part[c][q][1];
      neumaier(Si, Ci, part[c][q][2]);
      Ci += part[c][q][3];
    }
    r.open[((size_t)t * r.np + p) * r.nq + q] = make_double2(Sr + Cr, Si + Ci);
  }
}

void launch_project_reduce(const ReduceArgs &r, hipStream_t stream)
{
  hipLaunchKernelGGL(project_reduce_kernel, dim3(r.np, r.lt), dim3(256), 0, stream, r);
}

__global__ __launch_bounds__(256) void project_apply_kernel(ApplyArgs f)
{
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; // (w, p, t)
  const size_t total = (size_t)f.nproj * f.np * f.lt;
  if (e >= total) return;
  const int t = (int)(e % f.lt);
  const int p = (int)((e / f.lt) % f.np);
  const int w = (int)(e / ((size_t)f.lt * f.np));
  const double2 *o = f.open + ((size_t)t * f.np + p) * f.nq;
  double2 acc = make_double2(0.0, 0.0);
  for (int q = 0; q < f.nq; ++q) acc = cfma(acc, make_double2(f.w_re[w][q], f.w_im[w][q]), o[q]);
  f.out[((size_t)(f.proj_base + w) * f.np + p) * f.lt + t] = acc;
}

void launch_project_apply(const ApplyArgs &f, hipStream_t stream)
{
  const size_t total = (size_t)f.nproj * f.np * f.lt;
  hipLaunchKernelGGL(project_apply_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream, f);
}

} // namespace contract_amd
