/*
 * contract_ext.h -- extension API of libcontract.so (no counterpart in the
 * reference's include/contract.h, which stays unchanged).
 *
 * The reference stops at per-site correlators and leaves the timeslice sum
 * and momentum projection to the caller (cupy in PyQUDA).  These entry points
 * fuse them into the contraction: propagators are read from HBM once and the
 * result is the projected correlator
 *
 *   out[w][p][t] = sum_{x in timeslice t} exp(-i 2pi sum_mu p_mu x_mu / L_mu)
 *                  * sum_terms weight * C_{type}(x; Gij, Gkl, W_w)
 *
 * where C_{type}(x; Gij, Gkl, W) is exactly what baryon_general_two_point
 * returns for site x with project_mn = W.
 *
 * Conventions
 *   dims        {Lx, Ly, Lz, Lt} of the local lattice; volume = product.
 *   site_order  0: lexicographic, site = ((t*Lz + z)*Ly + y)*Lx + x.
 *               1: even-odd, site = parity*V/2 + ((t*Lz + z)*Ly + y)*Lx/2 + x/2
 *                  with parity = (x+y+z+t) % 2 (PyQUDA propagators).
 *   momenta     host int[nmom][3], integer components (any sign) in units
 *               of 2pi/L_mu.
 *   projectors  host double _Complex[nproj][4][4], row-major [m][n]; use
 *               contract_gamma_matrix() for G(g).
 *   out         device complex128 [nproj][nmom][Lt].
 *   precision   CONTRACT_FP64: propagators are complex128, all arithmetic FP64.
 *               CONTRACT_FP32: propagators are complex64; the per-site
 *               contraction runs in FP32 on power-of-two rescaled blocks and
 *               the timeslice/momentum sums stay FP64 (compensated).
 *               CONTRACT_FP32_SCALED: propagators in the FP32 site-scaled form
 *               of contract_propagator_to_scaled_fp32(); FP32 contraction with
 *               no per-site rescaling in the kernel, FP64 sums.
 *   workspace   device scratch of contract_project_workspace_bytes() bytes,
 *               or NULL to let the library allocate it stream-ordered.
 *   stream      a hipStream_t (NULL = null stream).
 *
 * Return value: 0 on success, nonzero on error; contract_last_error()
 * describes the most recent error of the calling thread.  These entry points
 * never exit the process.
 */
#pragma once

#include <stddef.h>
#ifndef __cplusplus
#include <complex.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

enum { CONTRACT_FP64 = 0, CONTRACT_FP32 = 1, CONTRACT_FP32_SCALED = 2 };
enum { CONTRACT_LEXICOGRAPHIC = 0, CONTRACT_EVEN_ODD = 1 };

/* Message of the calling thread's most recent error ("" if none). */
const char *contract_last_error(void);

/* FP32 site-scaled propagator storage (CONTRACT_FP32_SCALED), V sites in the
 * caller's site order: complex64 [V][4][4][3][3] holding P(x) * 2^-e(x),
 * followed by int8 e(x)[V], where e(x) = floor(log2 of the largest |Re| or
 * |Im| of site x) clamped to [-127, 127] (0 for an all-zero site), so every
 * stored site's largest component lies in [1, 2).  A site whose components
 * all lie within 2^-127 .. 2^128 is stored to FP32 rounding; the contraction
 * applies 2^(e_i + e_j + e_n) to each site's result in FP64.  The form holds
 * propagators decaying over many decades without FP32 underflow in the
 * products of three of them. */
size_t contract_scaled_fp32_bytes(size_t volume); /* V * 1153 */

/* complex128 [V][4][4][3][3] (device) -> FP32 site-scaled form (device, of
 * contract_scaled_fp32_bytes(V) bytes), stream-ordered. */
int contract_propagator_to_scaled_fp32(void *dst, const void *src, size_t volume, void *stream);

/* Dense DeGrand-Rossi G(g), g = 0..15, written as [4][4] row-major. */
int contract_gamma_matrix(int g, double _Complex *out);

/* Device scratch needed by the projected entry points for this lattice. */
size_t contract_project_workspace_bytes(const int dims[4], int site_order, int nmom);

/* Baryon two-point, summed over `nterm` contraction terms (e.g. the direct and
 * exchange terms of a proton), for `nproj` insertions, projected to
 * timeslices and momenta.  contract_types[t] in 0..5 (BaryonContractType).
 * propag_n may be the same buffer as propag_i (a nucleon's (u, d, u)): it is
 * then read once per site, with results identical to passing a copy. */
int contract_baryon_twopoint_project(void *out, const void *propag_i, const void *propag_j, const void *propag_n,
                                     int precision, const int *contract_types, const double _Complex *weights,
                                     int nterm, int gamma_ij, int gamma_kl, const double _Complex *projectors,
                                     int nproj, const int *momenta, int nmom, const int dims[4], int site_order,
                                     void *workspace, size_t workspace_bytes, void *stream);

/* The local lattice as one rank's part of a larger one (PyQUDA's MPI grid):
 * momenta are in units of 2pi / global[mu] and the phase of local site x is
 * exp(-i 2pi sum_mu p_mu (x_mu + origin[mu]) / global[mu]); out keeps the local
 * timeslices (global timeslice = t + origin[3]).  A zero global[mu] means
 * global[mu] = dims[mu].  Even-odd order needs even origins (local parity =
 * global parity).  The plain entry point is the case global = dims, origin = 0. */
typedef struct {
  int dims[4];   /* local Lx, Ly, Lz, Lt */
  int global[4]; /* global extents, >= dims; 0 = dims */
  int origin[4]; /* global coordinates of local site (0, 0, 0, 0), 0 <= origin < global */
  int site_order;
} contract_sublattice;

int contract_baryon_twopoint_project_sub(void *out, const void *propag_i, const void *propag_j,
                                         const void *propag_n, int precision, const int *contract_types,
                                         const double _Complex *weights, int nterm, int gamma_ij, int gamma_kl,
                                         const double _Complex *projectors, int nproj, const int *momenta, int nmom,
                                         const contract_sublattice *lattice, void *workspace, size_t workspace_bytes,
                                         void *stream);

/* Several baryon two-points (e.g. source positions, flavour combinations or
 * lattices of different size) in one fused launch: every request has its
 * own propagators, output out[nproj][nmom][Lt] and lattice; terms, gammas,
 * projectors, momenta and precision are shared.  Up to 16 requests share a
 * launch.  A request's result does not depend on the rest of the batch. */
typedef struct {
  const void *propag_i, *propag_j, *propag_n;
  void *out;
  contract_sublattice lattice;
} contract_baryon_request;

size_t contract_baryon_batch_workspace_bytes(const contract_baryon_request *reqs, int nreq, int nmom);

int contract_baryon_twopoint_project_batch(const contract_baryon_request *reqs, int nreq, int precision,
                                           const int *contract_types, const double _Complex *weights, int nterm,
                                           int gamma_ij, int gamma_kl, const double _Complex *projectors, int nproj,
                                           const int *momenta, int nmom, void *workspace, size_t workspace_bytes,
                                           void *stream);

/* ---- diquark in FP64 or FP32 site-scaled storage ----------------------
 * baryon_diquark (include/contract.h; the reference's lib/contract.cpp:24-28
 * -> lib/baryon/diquark.cu:21-42) with a precision flag and an error code:
 * CONTRACT_FP64 is the drop-in call on complex128 buffers (the same kernel,
 * bit-identical); CONTRACT_FP32_SCALED takes both propagators in the FP32
 * site-scaled form and writes the diquark in that form (FP32 arithmetic on
 * the stored mantissas, each output site renormalised, its exponent byte
 * after the V complex64 sites).  Unlike baryon_diquark, an out-of-range
 * gamma_ij is an error, not a silent no-op. */
int contract_baryon_diquark(void *diquark, const void *propag_i, const void *propag_j, int precision, size_t volume,
                            int gamma_ij, int gamma_kl, void *stream);

/* ---- sequential source in FP64 or FP32 site-scaled storage -------------
 * baryon_sequential_two_point (include/contract.h; the reference's
 * lib/contract.cpp:46-63) with a precision flag and an error code instead of
 * exit(): CONTRACT_FP64 is the drop-in call on complex128 buffers;
 * CONTRACT_FP32_SCALED takes all three buffers in the FP32 site-scaled form
 * (contract_scaled_fp32_bytes(volume) bytes each) and writes the sequential
 * source seq_X = dC/dP_X into the selected buffer in that form, mantissas
 * and site exponents both (FP32 arithmetic on the stored mantissas; an
 * output site below 2^-253 of scale loses its low bits).  The selected
 * buffer is overwritten and never read. */
int contract_baryon_sequential(void *propag_i, void *propag_j, void *propag_n, int precision, int contract_type,
                               int sequential_type, size_t volume, int gamma_ij, int gamma_kl, int gamma_mn,
                               void *stream);

/* ---- for callers that cannot allocate ----------------------------------
 * No counterpart in the reference.  For code under HIP graph capture, tight
 * serving loops and harnesses that forbid device allocation: the fused
 * baryon two-point of contract_baryon_twopoint_project_sub with its
 * workspace in a fixed device array of the library (64 MiB per GPU, held
 * until the process ends; contract_noalloc_scratch_bytes()) instead of a
 * caller workspace or a stream-ordered allocation.  The call allocates no
 * device memory.  A call needing more scratch than the array holds
 * (contract_baryon_batch_workspace_bytes of its one request) fails with
 * CONTRACT_ERR_WORKSPACE.  Every call on one GPU uses that GPU's one array:
 * calls that overlap on one GPU (other streams or host threads) race; order
 * them on one stream. */
size_t contract_noalloc_scratch_bytes(void);

int contract_baryon_twopoint_project_noalloc(void *out, const void *propag_i, const void *propag_j,
                                             const void *propag_n, int precision, const int *contract_types,
                                             const double _Complex *weights, int nterm, int gamma_ij, int gamma_kl,
                                             const double _Complex *projectors, int nproj, const int *momenta,
                                             int nmom, const contract_sublattice *lattice, void *stream);

/* ---- three-point step after the sequential source ----------------------
 * baryon_sequential_two_point (include/contract.h) leaves the sequential
 * source seq_X = dC/dP_X of a two-point function in the buffer of P_X.  A
 * solver turns it into the sequential propagator Sigma = D^-1 sigma, with
 * source sigma(y)[s1][s2](c1,c2) = g5[s1] conj(seq_X(y)[s1][s2](c1,c2)).
 * The three-point function of a current Gamma inserted at x on quark line
 * X, with F the forward propagator of that line (F = P_X), is
 *
 *   out[w][p][t] = sum_{x in t} exp(-i 2pi p.x / L) Tr_s Tr_c[ g5 Gamma_w F(x) Sigma(x)^dagger ]
 *
 * (= meson_two_point(F, Sigma, volume, Gamma_w, 15) per site for a gamma
 * index; any dense Gamma_w here), out being device complex128
 * [nins][nmom][Lt] per request.  insertions: host double _Complex
 * [nins][4][4] row-major (contract_gamma_matrix() for G(g)).  F and Sigma
 * share a storage form (precision as above).  Up to 16 requests share a
 * launch; momenta and insertions are shared; the workspace and the work
 * plan (contract_set_baryon_plan) are those of the fused baryon entry
 * points.  With D = 1 (Sigma = sigma) out equals the projected two-point
 * function with P_X replaced by Gamma P_X. */
typedef struct {
  const void *propag_fwd; /* F, [V][4][4][3][3] */
  const void *propag_seq; /* Sigma, [V][4][4][3][3] */
  void *out;
  contract_sublattice lattice;
} contract_threepoint_request;

size_t contract_threepoint_workspace_bytes(const contract_threepoint_request *reqs, int nreq, int nmom);

int contract_threepoint_project(const contract_threepoint_request *reqs, int nreq, int precision,
                                const double _Complex *insertions, int nins, const int *momenta, int nmom,
                                void *workspace, size_t workspace_bytes, void *stream);

/* ---- batched meson Gamma x Gamma' grids -------------------------------
 * One request = two propagators on one lattice.  For every request the full
 * 16 x 16 grid of meson_two_point(propag_i, propag_j, volume, g1, g2) (the
 * reference's include/contract.h:16 per-site correlator, one call per gamma
 * pair there) is summed per timeslice and projected to momenta:
 *
 *   out[g1][g2][p][t] = sum_{x in t} exp(-i 2pi p.x / L) meson(x; g1, g2),
 *
 * out being device complex128 [16][16][nmom][Lt] of that request.  All
 * requests of one call share precision, momenta and the workspace, and go
 * through one persistent launch per group of 32 requests and 16 +-k momentum
 * slots (a momentum and its negative share a slot), so
 * small and large lattices balance across the chip. */
typedef struct {
  const void *propag_i; /* device [V][4][4][3][3], complex128 or complex64 */
  const void *propag_j;
  void *out;            /* device complex128 [16][16][nmom][Lt] */
  int dims[4];          /* Lx, Ly, Lz, Lt */
  int site_order;       /* CONTRACT_LEXICOGRAPHIC or CONTRACT_EVEN_ODD */
  int global[3];        /* spatial global extents, 0 = dims (see contract_sublattice) */
  int origin[3];        /* spatial global coordinates of local site 0 */
} contract_meson_request;

/* Work decomposition; a zero field takes the library default. */
typedef struct {
  int sites_per_item; /* sites per work item (rounded up to the tile) */
  int tile_sites;     /* sites staged per step: 4 or 8; 0: 8 for plain FP32
                         with vector-ALU traces, else 4 */
  int blocks_per_cu;  /* persistent workgroups per compute unit: 1..8; 0:
                         3 four-wave workgroups for matrix-core traces, 2 for
                         vector-ALU traces, or for a wave-local pass 4
                         one-wave workgroups (8 for the FP32 forms) */
  int trace_unit;     /* per-site traces on 1: the vector ALUs, 2: the matrix
                         cores (FP64 and FP32 site-scaled storage only),
                         3: the matrix cores with one wave per 4-site tile
                         holding all of K (any storage; plain FP32 tiles are
                         scaled per site as they land; passes of at most 4
                         momentum slots, other passes run as 2, or as 1 for
                         plain FP32; tile_sites ignored, blocks_per_cu
                         counts one-wave workgroups);
                         0: the library's choice (2 for FP64, 3 for both
                         FP32 forms) */
} contract_meson_grid_tuning;

size_t contract_meson_grid_workspace_bytes(const contract_meson_request *reqs, int nreq, int nmom,
                                           const contract_meson_grid_tuning *tuning);

/* The kernel timer (contract_set_kernel_timer) brackets the whole batch. */
int contract_meson_grid_project(const contract_meson_request *reqs, int nreq, int precision, const int *momenta,
                                int nmom, const contract_meson_grid_tuning *tuning, void *workspace,
                                size_t workspace_bytes, void *stream);

/* Stream of the drop-in entry points of include/contract.h called from this
 * thread (the reference launches them on the null stream, which stays the
 * default).  Lets a caller overlap reference-ABI work on its own streams. */
int contract_set_stream(void *stream);

/* Work decomposition of the fused baryon entry points above, for calls from
 * this thread (including their workspace sizes): a lattice is cut into about
 * items_per_cu work items per compute unit, each at least min_item_sites
 * sites inside one timeslice.  A zero field takes the default (8 and 128);
 * NULL restores both.  Results agree to rounding for any plan; for one plan
 * they are bit-reproducible. */
typedef struct {
  int items_per_cu;   /* 1..256 */
  int min_item_sites; /* 4..2^20 */
} contract_baryon_plan;

int contract_set_baryon_plan(const contract_baryon_plan *plan);
int contract_get_baryon_plan(contract_baryon_plan *plan);

/* Profiling hook: while set, every projected entry point called from this
 * thread records `start` (a hipEvent_t) on its stream right before its main
 * contraction kernel and `stop` right after it, so a caller can time that
 * kernel alone.  Pass NULL, NULL to clear. */
int contract_set_kernel_timer(void *start, void *stop);

#ifdef __cplusplus
}
#endif
