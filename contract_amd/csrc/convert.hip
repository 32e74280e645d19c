// convert.hip -- FP64 propagators to the FP32 site-scaled storage form
// (include/contract_ext.h, CONTRACT_FP32_SCALED): complex64 P(x) 2^-e(x)
// followed by the int8 site exponents e(x).

#include "baryon.hpp"
#include "device.hpp"

namespace contract_amd
{

// One 16-lane group per site: lane q takes spin block q (9 colour entries),
// the group max gives the site exponent, the scaled block goes to LDS in the
// FP32 tile layout and back to HBM with coalesced stores.
__global__ __launch_bounds__(WAVE) void scale_fp32_kernel(ScaleArgs a)
{
  constexpr int SL = site_lds<double2>();
  constexpr int SL32 = site_lds<float2>();
  __shared__ double2 lds[SITES_PER_WAVE * SL];
  __shared__ float2 out[SITES_PER_WAVE * SL32];
  const int lane = threadIdx.x;
  const int g = lane >> 4, q = lane & 15;
  const double2 *P[1] = {static_cast<const double2 *>(a.src)};
  persistent_tiles<1>(lds, P, a.volume, lane, [&](size_t first, int nsites) {
    const double2 *b = lds + g * SL + q * 9;
    double mx = 0.0;
#pragma unroll
    for (int e = 0; e < 9; ++e) mx = fmax(mx, fmax(fabs(b[e].x), fabs(b[e].y)));
#pragma unroll
    for (int m = 8; m >= 1; m >>= 1) mx = fmax(mx, shfl_xor_d(mx, m));
    int ex = mx > 0.0 ? ilogb(mx) : 0;
    ex = ex < -SCALED_EXP_MAX ? -SCALED_EXP_MAX : (ex > SCALED_EXP_MAX ? SCALED_EXP_MAX : ex);
    const double sc = ldexp(1.0, -ex);
    float2 *o = out + g * SL32 + q * 9;
#pragma unroll
    for (int e = 0; e < 9; ++e) o[e] = make_float2((float)(b[e].x * sc), (float)(b[e].y * sc));
    if (q == 0 && g < nsites) static_cast<signed char *>(a.exps)[first + g] = (signed char)ex;
    __syncthreads();
    unstage_sites(static_cast<float2 *>(a.dst), out, first, nsites, lane);
  });
}

void launch_scale_fp32(const ScaleArgs &a, hipStream_t stream)
{
  const size_t ntiles = (a.volume + SITES_PER_WAVE - 1) / SITES_PER_WAVE;
  if (ntiles == 0) return;
  hipLaunchKernelGGL(scale_fp32_kernel, dim3((unsigned)persistent_grid(ntiles, 8)), dim3(WAVE), 0, stream, a);
}

} // namespace contract_amd
