/*
 * abi_check.c -- a C99 consumer of the installed libcontract.so: includes
 * contract.h and contract_ext.h, links -lcontract, calls the drop-in entry
 * points the reference's own C tests call (test/test_baryon.cpp:74-121,
 * test/test_meson.cpp:25-44) and two extension entry points on device
 * buffers, and compares every result with the committed fixture
 * tests/golden/c_abi_fixture.bin (the FP64 oracle on the same inputs, made
 * by tests/golden/make_c_abi_fixture.py).
 *
 *   abi_check FIXTURE    exit 0 when every result is within 1e-12
 *
 * Inputs: lattice 4x2x2x4 (V = 64), three propagators whose entries are
 * 2u - 1 for u = splitmix64(seed << 32 | k) / 2^64 (k = index of the real
 * or imaginary part), seeds 1, 2, 3.
 */
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <contract.h>
#include <contract_ext.h>

#define LX 4
#define LY 2
#define LZ 2
#define LT 4
#define V (LX * LY * LZ * LT)
#define PROP_DOUBLES ((size_t)V * 144 * 2)

static uint64_t splitmix64(uint64_t x)
{
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static void fill(double *p, uint64_t seed)
{
  for (size_t k = 0; k < PROP_DOUBLES; ++k) p[k] = 2.0 * (double)(splitmix64(seed << 32 | k) >> 11) * 0x1p-53 - 1.0;
}

static void *dev_alloc(size_t bytes)
{
  void *p = NULL;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    fprintf(stderr, "hipMalloc of %zu bytes failed\n", bytes);
    exit(2);
  }
  return p;
}

static void to_dev(void *d, const void *h, size_t bytes)
{
  if (hipMemcpy(d, h, bytes, hipMemcpyHostToDevice) != hipSuccess) exit(2);
}

static void to_host(void *h, const void *d, size_t bytes)
{
  if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost) != hipSuccess) exit(2);
}

/* The fixture holds the expected results one after another, each as
 * complex128, in the order of the checks below. */
static double *fixture;
static size_t fixture_len, fixture_pos;
static int failures;

static void check(const char *name, const double *got, size_t n_complex)
{
  if (fixture_pos + 2 * n_complex > fixture_len) {
    fprintf(stderr, "%s: fixture too short\n", name);
    exit(2);
  }
  const double *want = fixture + fixture_pos;
  fixture_pos += 2 * n_complex;
  double num = 0.0, den = 0.0;
  for (size_t k = 0; k < 2 * n_complex; ++k) {
    num = fmax(num, fabs(got[k] - want[k]));
    den = fmax(den, fabs(want[k]));
  }
  const double rel = den > 0 ? num / den : num;
  const int ok = rel <= 1e-12 && den > 0;
  printf("%-48s rel err %.3e  %s\n", name, rel, ok ? "PASS" : "FAIL");
  failures += !ok;
}

int main(int argc, char **argv)
{
  if (argc != 2) {
    fprintf(stderr, "usage: %s FIXTURE\n", argv[0]);
    return 2;
  }
  FILE *f = fopen(argv[1], "rb");
  if (!f) {
    perror(argv[1]);
    return 2;
  }
  fseek(f, 0, SEEK_END);
  fixture_len = (size_t)ftell(f) / sizeof(double);
  fseek(f, 0, SEEK_SET);
  fixture = malloc(fixture_len * sizeof(double));
  if (fread(fixture, sizeof(double), fixture_len, f) != fixture_len) return 2;
  fclose(f);

  init(0);
  const size_t pbytes = PROP_DOUBLES * sizeof(double);
  double *h[3], *host_out = malloc(pbytes);
  void *d[3];
  for (int k = 0; k < 3; ++k) {
    h[k] = malloc(pbytes);
    fill(h[k], (uint64_t)(k + 1));
    d[k] = dev_alloc(pbytes);
    to_dev(d[k], h[k], pbytes);
  }
  void *correl = dev_alloc(V * 16);

  /* drop-in entry points, as the reference's tests call them */
  baryon_two_point(correl, d[0], d[1], d[2], IK_JL_NM, V, 5, 5, 8);
  to_host(host_out, correl, V * 16);
  check("baryon_two_point IK_JL_NM Cg5 Cg5 G(8)", host_out, V);
  baryon_two_point(correl, d[0], d[1], d[2], IN_JL_MK, V, 5, 5, 8);
  to_host(host_out, correl, V * 16);
  check("baryon_two_point IN_JL_MK (= IM_JL_NK) Cg5 Cg5 G(8)", host_out, V);

  double _Complex project_mn[16];
  memset(project_mn, 0, sizeof(project_mn));
  for (int i = 0; i < 4; ++i) project_mn[i * 4 + (i + 2) % 4] = 1; /* g4, test/test_baryon.cpp:79-80 */
  baryon_general_two_point(correl, d[0], d[1], d[2], IK_JN_ML, V, 5, 5, project_mn);
  to_host(host_out, correl, V * 16);
  check("baryon_general_two_point IK_JN_ML projector g4", host_out, V);

  void *all[16];
  void *all_buf = dev_alloc(16 * V * 16);
  for (int g = 0; g < 16; ++g) all[g] = (char *)all_buf + (size_t)g * V * 16;
  meson_all_source_two_point(all, d[0], d[1], V, 5);
  to_host(host_out, all_buf, 16 * V * 16);
  check("meson_all_source_two_point Gij = G(5), 16 Gkl", host_out, 16 * V);
  meson_two_point(correl, d[0], d[1], V, 15, 15);
  to_host(host_out, correl, V * 16);
  check("meson_two_point G(15) G(15)", host_out, V);

  void *out_prop = dev_alloc(pbytes);
  baryon_diquark(out_prop, d[0], d[1], V, 5, 5);
  to_host(host_out, out_prop, pbytes);
  check("baryon_diquark Cg5 Cg5", host_out, (size_t)V * 144);
  to_dev(out_prop, h[2], pbytes); /* the sequential target is overwritten, never read */
  baryon_sequential_two_point(d[0], d[1], out_prop, IK_JL_NM, SEQUENTIAL_N, V, 5, 5, 8);
  to_host(host_out, out_prop, pbytes);
  check("baryon_sequential_two_point IK_JL_NM SEQUENTIAL_N", host_out, (size_t)V * 144);

  /* extension: fused proton two-point, 16 insertions, 3 momenta */
  const int dims[4] = {LX, LY, LZ, LT};
  const int types[2] = {IK_JL_NM, IM_JL_NK};
  const double _Complex weights[2] = {1, 1};
  double _Complex proj[16][16];
  for (int g = 0; g < 16; ++g) contract_gamma_matrix(g, proj[g]);
  const int mom[3][3] = {{0, 0, 0}, {1, 0, 0}, {0, -1, 1}};
  void *proj_out = dev_alloc(16 * 3 * LT * 16);
  if (contract_baryon_twopoint_project(proj_out, d[0], d[1], d[2], CONTRACT_FP64, types, weights, 2, 5, 5, &proj[0][0],
                                       16, &mom[0][0], 3, dims, CONTRACT_LEXICOGRAPHIC, NULL, 0, NULL) != 0) {
    fprintf(stderr, "contract_baryon_twopoint_project: %s\n", contract_last_error());
    return 1;
  }
  to_host(host_out, proj_out, 16 * 3 * LT * 16);
  check("contract_baryon_twopoint_project proton 16x3", host_out, 16 * 3 * LT);

  /* extension: three-point step on (F = P_0, Sigma = P_1), 16 insertions */
  contract_threepoint_request req;
  memset(&req, 0, sizeof(req));
  req.propag_fwd = d[0];
  req.propag_seq = d[1];
  req.out = proj_out;
  for (int mu = 0; mu < 4; ++mu) req.lattice.dims[mu] = dims[mu];
  if (contract_threepoint_project(&req, 1, CONTRACT_FP64, &proj[0][0], 16, &mom[0][0], 3, NULL, 0, NULL) != 0) {
    fprintf(stderr, "contract_threepoint_project: %s\n", contract_last_error());
    return 1;
  }
  to_host(host_out, proj_out, 16 * 3 * LT * 16);
  check("contract_threepoint_project 16x3", host_out, 16 * 3 * LT);

  if (fixture_pos != fixture_len) {
    fprintf(stderr, "fixture has %zu unread values\n", fixture_len - fixture_pos);
    return 1;
  }
  printf("%s\n", failures ? "FAILED" : "ALL PASS");
  return failures ? 1 : 0;
}
