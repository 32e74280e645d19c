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

// One block of 512 threads per (momentum slot p, timeslice t of one
// request): thread (c0, q) sums component q (< nq <= 32) over the work items
// c0, c0 + 16, ... of the timeslice (both parities for even-odd order); the 16
// column partials are then combined in index order.  Every sum is Neumaier
// compensated and the order is fixed, so results are bit-reproducible and
// independent of the other requests of the launch.
constexpr int REDUCE_Q = 32;

__global__ __launch_bounds__(16 * REDUCE_Q) void project_reduce_kernel(ReduceArgs r)
{
  __shared__ double part[16][REDUCE_Q][4];
  const int p = blockIdx.y;
  int k = 0; // request of this timeslice (uniform)
  while (k + 1 < r.nreq && (int)blockIdx.x >= r.req[k + 1].t0) ++k;
  const ReduceReq &rq = r.req[k];
  const int t = (int)blockIdx.x - rq.t0;
  const int q = threadIdx.x % REDUCE_Q, c0 = threadIdx.x / REDUCE_Q;
  double sr = 0, cr = 0, si = 0, ci = 0;
  if (q < r.nq) {
    const int nslices = rq.order ? 2 : 1;
    for (int par = 0; par < nslices; ++par) {
      const size_t slice = (size_t)par * rq.lt + t;
      for (int c = c0; c < rq.chunks_per_slice; c += 16) {
        const size_t item = slice * rq.chunks_per_slice + c;
        const double2 v = rq.partial[(item * r.np + p) * r.nq + q];
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
    rq.open[((size_t)t * r.np + p) * r.nq + q] = make_double2(Sr + Cr, Si + Ci);
  }
}

void launch_project_reduce(const ReduceArgs &r, int total_lt, hipStream_t stream)
{
  hipLaunchKernelGGL(project_reduce_kernel, dim3(total_lt, r.np), dim3(16 * REDUCE_Q), 0, stream, r);
}

__global__ __launch_bounds__(256) void project_apply_kernel(ApplyArgs f)
{
  const unsigned g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= f.total) return;
  int k = 0; // request of this element
  while (k + 1 < f.nreq && g >= f.req[k + 1].e0) ++k;
  const ApplyReq &rq = f.req[k];
  const unsigned e = g - rq.e0; // (w, p, t) of this request
  const int t = (int)(e % rq.lt);
  const int p = (int)((e / rq.lt) % f.nmom);
  const int w = (int)(e / ((unsigned)rq.lt * f.nmom));
  const double2 *o = rq.open + ((size_t)t * f.nslot + f.slot[p]) * 32; // C[16], then S[16]
  const double sg = (double)f.sgn[p];
  double2 acc = make_double2(0.0, 0.0);
  for (int q = 0; q < 16; ++q) {
    const double2 v = make_double2(o[q].x - sg * o[16 + q].y, o[q].y + sg * o[16 + q].x); // C + sg i S
    acc = cfma(acc, make_double2(f.w_re[w][q], f.w_im[w][q]), v);
  }
  rq.out[((size_t)(f.proj_base + w) * f.nmom + p) * rq.lt + t] = acc;
}

void launch_project_apply(const ApplyArgs &f, hipStream_t stream)
{
  hipLaunchKernelGGL(project_apply_kernel, dim3((f.total + 255) / 256), dim3(256), 0, stream, f);
}

} // namespace contract_amd
