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

// One block of 512 threads per (momentum slot p, timeslice t): thread (c0, q)
// sums component q (< nq <= 32) over the work items c0, c0 + 16, ... of the
// timeslice (both parities for even-odd order); the 16 column partials are
// then combined in index order.  Every sum is Neumaier compensated and the
// order is fixed, so results are bit-reproducible.
constexpr int REDUCE_Q = 32;

__global__ __launch_bounds__(16 * REDUCE_Q) void project_reduce_kernel(ReduceArgs r)
{
  __shared__ double part[16][REDUCE_Q][4];
  const int p = blockIdx.x, t = blockIdx.y;
  const int q = threadIdx.x % REDUCE_Q, c0 = threadIdx.x / REDUCE_Q;
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
  hipLaunchKernelGGL(project_reduce_kernel, dim3(r.np, r.lt), dim3(16 * REDUCE_Q), 0, stream, r);
}

__global__ __launch_bounds__(256) void project_apply_kernel(ApplyArgs f)
{
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; // (w, p, t)
  const size_t total = (size_t)f.nproj * f.nmom * f.lt;
  if (e >= total) return;
  const int t = (int)(e % f.lt);
  const int p = (int)((e / f.lt) % f.nmom);
  const int w = (int)(e / ((size_t)f.lt * f.nmom));
  const double2 *o = f.open + ((size_t)t * f.nslot + f.slot[p]) * 32; // C[16], then S[16]
  const double sg = (double)f.sgn[p];
  double2 acc = make_double2(0.0, 0.0);
  for (int q = 0; q < 16; ++q) {
    const double2 v = make_double2(o[q].x - sg * o[16 + q].y, o[q].y + sg * o[16 + q].x); // C + sg i S
    acc = cfma(acc, make_double2(f.w_re[w][q], f.w_im[w][q]), v);
  }
  f.out[((size_t)(f.proj_base + w) * f.nmom + p) * f.lt + t] = acc;
}

void launch_project_apply(const ApplyArgs &f, hipStream_t stream)
{
  const size_t total = (size_t)f.nproj * f.nmom * f.lt;
  hipLaunchKernelGGL(project_apply_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream, f);
}

} // namespace contract_amd
