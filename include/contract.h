/*
 * contract.h -- drop-in C ABI of libcontract.so (MI355X / gfx950 build).
 *
 * The declarations are those of the reference's include/contract.h:15-27, so
 * `python -m pyquda_plugins -i contract.h -l contract` generates the same
 * binding.  Semantics per entry point (reference file:line replaced):
 *
 *   init                         lib/contract.cpp:65-69        select the HIP device
 *   meson_two_point              lib/meson/meson.cu:6-18       correl[V] = Tr[(g5 Gij) P_i (g5 Gkl) P_j^+]
 *   meson_all_source_two_point   lib/meson/meson_all.cu:6-18   correl[g][V], Gkl = G(g), g = 0..15
 *   meson_all_sink_two_point     lib/meson/meson_all.cu:22-34  correl[g][V], Gij = G(g), g = 0..15
 *   baryon_diquark               lib/baryon/diquark.cu:7-42    diquark[V][16][9]
 *   baryon_two_point             lib/baryon/baryon.cu:7-67     correl[V]
 *   baryon_general_two_point     lib/baryon/baryon_general.cu:7-51  correl[V], dense projector
 *   baryon_sequential_two_point  lib/contract.cpp:46-63        overwrites the selected propagator
 *
 * All pointers except `correl` of meson_all_* (a host array of 16 device
 * pointers) and `project_mn` (a host array of 16 complex, row-major [m][n])
 * are device pointers to complex128 data; propagators are
 * [volume][4 sink spin][4 source spin][3 sink colour][3 source colour].
 * Launches are asynchronous on the null stream; the caller synchronises.
 * Invalid volume (% 4 != 0), contract type, gamma_mn or gamma index
 * prints an error and exits the process, like the reference; an invalid
 * sequential_type or diquark gamma_ij is a silent no-op, like the reference.
 */
#pragma once

#include <stdio.h>
#include <stdlib.h>
#ifndef __cplusplus
#include <complex.h>
#endif

#include <contract_enum.h>

#ifdef __cplusplus
extern "C" {
#endif

void init(int device);
void meson_two_point(void *correl, void *propag_i, void *propag_j, unsigned long volume, int gamma_ij, int gamma_kl);
void meson_all_source_two_point(void **correl, void *propag_i, void *propag_j, unsigned long volume, int gamma_ij);
void meson_all_sink_two_point(void **correl, void *propag_i, void *propag_j, unsigned long volume, int gamma_kl);
void baryon_diquark(void *diquark, void *propag_i, void *propag_j, size_t volume, int gamma_ij, int gamma_kl);
void baryon_two_point(void *correl, void *propag_i, void *propag_j, void *propag_n, BaryonContractType contract_type,
                      unsigned long volume, int gamma_ij, int gamma_kl, int gamma_mn);
void baryon_general_two_point(void *correl, void *propag_i, void *propag_j, void *propag_n,
                              BaryonContractType contract_type, size_t volume, int gamma_ij, int gamma_kl,
                              double _Complex *project_mn);
void baryon_sequential_two_point(void *propag_i, void *propag_j, void *propag_n, BaryonContractType contract_type,
                                 BaryonSequentialType sequential_type, size_t volume, int gamma_ij, int gamma_kl,
                                 int gamma_mn);

#ifdef __cplusplus
}
#endif
