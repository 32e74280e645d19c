// abi.hip -- the drop-in C ABI (include/contract.h).  Each entry point checks
// its arguments the way the reference's launcher does, builds a by-value
// argument block and launches one kernel on the null stream (or on the
// calling thread's stream set by contract_set_stream).  No entry point
// allocates, copies to __constant__ memory, or synchronises.

#include <contract.h>
#include <contract_ext.h>

#include <cstring>

#include "baryon.hpp"
#include "runtime.hpp"
#include "spin.hpp"

using namespace contract_amd;

namespace
{

void check_launch() { CONTRACT_HIP_CHECK(hipGetLastError()); }

void check_contract_type(int t)
{
  // lib/baryon/baryon.cu:61-64: exit on an unknown type
  if (t < 0 || t > 5) {
    std::fprintf(stderr, "Error: Invalid contract_type value %d\n", t);
    std::exit(-1);
  }
}

void check_gamma_mn(int g)
{
  // lib/baryon/baryon.cu:43-46
  if (g < 0 || g > 15) {
    std::fprintf(stderr, "Error: Invalid gamma_mn value %d\n", g);
    std::exit(-1);
  }
}

SparseSpin sp(int g) { return sparse_of(gamma_matrix(g)); }
SparseSpin spT(int g) { return sparse_of(mat_transpose(gamma_matrix(g))); }
SparseSpin g5spT(int g) { return sparse_of(mat_transpose(gamma5_times(g))); }

} // namespace

extern "C" {

void init(int device)
{
  CONTRACT_HIP_CHECK(hipSetDevice(device));
  // resolve this device's scratch address for the no-allocation entry point
  if (contract_amd::noalloc_setup() != contract_amd::CONTRACT_OK) {
    std::fprintf(stderr, "HIP error in init: %s\n", contract_last_error());
    std::exit(-1);
  }
}

void meson_two_point(void *correl, void *propag_i, void *propag_j, unsigned long volume, int gamma_ij, int gamma_kl)
{
  abi_check_volume(volume);
  abi_check_gamma(gamma_ij, "gamma_ij");
  abi_check_gamma(gamma_kl, "gamma_kl");
  MesonArgs a;
  a.propag_i = propag_i;
  a.propag_j = propag_j;
  a.volume = volume;
  a.a = g5spT(gamma_ij);
  a.b = g5spT(gamma_kl);
  a.correl = correl;
  launch_meson(a, abi_stream());
  check_launch();
}

static void meson_all(void **correl, void *propag_i, void *propag_j, unsigned long volume, int gamma, int sink)
{
  abi_check_volume(volume);
  abi_check_gamma(gamma, sink ? "gamma_kl" : "gamma_ij");
  MesonAllArgs a;
  a.propag_i = propag_i;
  a.propag_j = propag_j;
  a.volume = volume;
  a.sink = sink;
  a.fixed = g5spT(gamma);
  for (int g = 0; g < 16; ++g) {
    a.run[g] = g5spT(g);
    a.correl[g] = correl[g];
  }
  launch_meson_all(a, abi_stream());
  check_launch();
}

void meson_all_source_two_point(void **correl, void *propag_i, void *propag_j, unsigned long volume, int gamma_ij)
{
  meson_all(correl, propag_i, propag_j, volume, gamma_ij, 0);
}

void meson_all_sink_two_point(void **correl, void *propag_i, void *propag_j, unsigned long volume, int gamma_kl)
{
  meson_all(correl, propag_i, propag_j, volume, gamma_kl, 1);
}

void baryon_diquark(void *diquark, void *propag_i, void *propag_j, size_t volume, int gamma_ij, int gamma_kl)
{
  if (gamma_ij < 0 || gamma_ij > 15) return; // lib/baryon/diquark.cu:40: silent no-op
  abi_check_volume(volume);
  abi_check_gamma(gamma_kl, "gamma_kl");
  DiquarkArgs a = {};
  a.propag_i = propag_i;
  a.propag_j = propag_j;
  a.diquark = diquark;
  a.volume = volume;
  a.gij = sp(gamma_ij);
  a.gklT = spT(gamma_kl);
  launch_diquark(a, abi_stream());
  check_launch();
}

static void baryon_site(void *correl, void *propag_i, void *propag_j, void *propag_n, int contract_type,
                        size_t volume, int gamma_ij, int gamma_kl, const zhost W[16])
{
  BaryonSiteArgs a;
  a.propag_i = propag_i;
  a.propag_j = propag_j;
  a.propag_n = propag_n;
  a.correl = correl;
  a.volume = volume;
  a.gamma_ij = gamma_ij;
  a.gij = sp(gamma_ij);
  a.gkl = sp(gamma_kl);
  a.gklT = spT(gamma_kl);
  for (int q = 0; q < 16; ++q) {
    a.w_re[q] = W[q].real();
    a.w_im[q] = W[q].imag();
  }
  launch_baryon_site(contract_type, a, abi_stream());
  check_launch();
}

void baryon_two_point(void *correl, void *propag_i, void *propag_j, void *propag_n, BaryonContractType contract_type,
                      unsigned long volume, int gamma_ij, int gamma_kl, int gamma_mn)
{
  check_contract_type((int)contract_type);
  check_gamma_mn(gamma_mn);
  abi_check_volume(volume);
  abi_check_gamma(gamma_ij, "gamma_ij");
  abi_check_gamma(gamma_kl, "gamma_kl");
  const Mat4 G = gamma_matrix(gamma_mn);
  zhost W[16];
  for (int m = 0; m < 4; ++m)
    for (int n = 0; n < 4; ++n) W[m * 4 + n] = G.v[m][n];
  baryon_site(correl, propag_i, propag_j, propag_n, (int)contract_type, volume, gamma_ij, gamma_kl, W);
}

void baryon_general_two_point(void *correl, void *propag_i, void *propag_j, void *propag_n,
                              BaryonContractType contract_type, size_t volume, int gamma_ij, int gamma_kl,
                              double _Complex *project_mn)
{
  check_contract_type((int)contract_type);
  abi_check_volume(volume);
  abi_check_gamma(gamma_ij, "gamma_ij");
  abi_check_gamma(gamma_kl, "gamma_kl");
  const double *pm = reinterpret_cast<const double *>(project_mn);
  zhost W[16];
  for (int q = 0; q < 16; ++q) W[q] = zhost(pm[2 * q], pm[2 * q + 1]);
  baryon_site(correl, propag_i, propag_j, propag_n, (int)contract_type, volume, gamma_ij, gamma_kl, W);
}

void baryon_sequential_two_point(void *propag_i, void *propag_j, void *propag_n, BaryonContractType contract_type,
                                 BaryonSequentialType sequential_type, size_t volume, int gamma_ij, int gamma_kl,
                                 int gamma_mn)
{
  const int target = (int)sequential_type;
  if (target < 0 || target > 2) return; // lib/contract.cpp:60: silent no-op
  check_contract_type((int)contract_type);
  check_gamma_mn(gamma_mn);
  abi_check_volume(volume);
  abi_check_gamma(gamma_ij, "gamma_ij");
  abi_check_gamma(gamma_kl, "gamma_kl");
  void *P[3] = {propag_i, propag_j, propag_n};
  SequentialArgs a;
  a.in0 = P[target == 0 ? 1 : 0];
  a.in1 = P[target == 2 ? 1 : 2];
  a.out = P[target];
  a.volume = volume;
  a.gij = sp(gamma_ij);
  a.gijT = spT(gamma_ij);
  a.gkl = sp(gamma_kl);
  a.gklT = spT(gamma_kl);
  a.gmn = sp(gamma_mn);
  a.gmnT = spT(gamma_mn);
  launch_sequential((int)contract_type, target, a, abi_stream());
  check_launch();
}

int contract_baryon_diquark(void *diquark, const void *propag_i, const void *propag_j, int precision, size_t volume,
                            int gamma_ij, int gamma_kl, void *stream)
{
  if (precision != CONTRACT_FP64 && precision != CONTRACT_FP32_SCALED) {
    set_last_error("precision %d is not CONTRACT_FP64 or CONTRACT_FP32_SCALED", precision);
    return CONTRACT_ERR_ARG;
  }
  if (gamma_ij < 0 || gamma_ij > 15 || gamma_kl < 0 || gamma_kl > 15) {
    set_last_error("gamma_ij = %d / gamma_kl = %d out of range 0..15", gamma_ij, gamma_kl);
    return CONTRACT_ERR_ARG;
  }
  if (!diquark || !propag_i || !propag_j) {
    set_last_error("NULL propagator pointer");
    return CONTRACT_ERR_ARG;
  }
  if (volume == 0) return CONTRACT_OK;
  DiquarkArgs a = {};
  a.propag_i = propag_i;
  a.propag_j = propag_j;
  a.diquark = diquark;
  a.volume = volume;
  a.gij = sp(gamma_ij);
  a.gklT = spT(gamma_kl);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (precision == CONTRACT_FP64) {
    launch_diquark(a, s);
  } else { // the site exponents follow the complex64 entries of every buffer
    a.exp_i = static_cast<const signed char *>(propag_i) + volume * SCALED_SITE_BYTES;
    a.exp_j = static_cast<const signed char *>(propag_j) + volume * SCALED_SITE_BYTES;
    a.exp_out = static_cast<signed char *>(diquark) + volume * SCALED_SITE_BYTES;
    launch_diquark_fp32_scaled(a, s);
  }
  CONTRACT_EXT_HIP(hipGetLastError());
  return CONTRACT_OK;
}

int contract_baryon_sequential(void *propag_i, void *propag_j, void *propag_n, int precision, int contract_type,
                               int sequential_type, size_t volume, int gamma_ij, int gamma_kl, int gamma_mn,
                               void *stream)
{
  if (precision != CONTRACT_FP64 && precision != CONTRACT_FP32_SCALED) {
    set_last_error("precision %d is not CONTRACT_FP64 or CONTRACT_FP32_SCALED", precision);
    return CONTRACT_ERR_ARG;
  }
  if (sequential_type < 0 || sequential_type > 2) {
    set_last_error("sequential_type %d is not SEQUENTIAL_I, SEQUENTIAL_J or SEQUENTIAL_N", sequential_type);
    return CONTRACT_ERR_ARG;
  }
  if (contract_type < 0 || contract_type > 5) {
    set_last_error("contract_type %d is not a BaryonContractType", contract_type);
    return CONTRACT_ERR_ARG;
  }
  if (gamma_ij < 0 || gamma_ij > 15 || gamma_kl < 0 || gamma_kl > 15 || gamma_mn < 0 || gamma_mn > 15) {
    set_last_error("gamma_ij = %d / gamma_kl = %d / gamma_mn = %d out of range 0..15", gamma_ij, gamma_kl, gamma_mn);
    return CONTRACT_ERR_ARG;
  }
  if (!propag_i || !propag_j || !propag_n) {
    set_last_error("NULL propagator pointer");
    return CONTRACT_ERR_ARG;
  }
  if (volume == 0) return CONTRACT_OK;
  void *P[3] = {propag_i, propag_j, propag_n};
  const int target = sequential_type;
  SequentialArgs a;
  std::memset(&a, 0, sizeof(a));
  a.in0 = P[target == 0 ? 1 : 0];
  a.in1 = P[target == 2 ? 1 : 2];
  a.out = P[target];
  a.volume = volume;
  a.gij = sp(gamma_ij);
  a.gijT = spT(gamma_ij);
  a.gkl = sp(gamma_kl);
  a.gklT = spT(gamma_kl);
  a.gmn = sp(gamma_mn);
  a.gmnT = spT(gamma_mn);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (precision == CONTRACT_FP64) {
    launch_sequential(contract_type, target, a, s);
  } else { // the site exponents follow the complex64 entries of every buffer
    a.exp0 = static_cast<const signed char *>(a.in0) + volume * SCALED_SITE_BYTES;
    a.exp1 = static_cast<const signed char *>(a.in1) + volume * SCALED_SITE_BYTES;
    a.exp_out = static_cast<signed char *>(a.out) + volume * SCALED_SITE_BYTES;
    launch_sequential_fp32_scaled(contract_type, target, a, s);
  }
  CONTRACT_EXT_HIP(hipGetLastError());
  return CONTRACT_OK;
}

} // extern "C"
