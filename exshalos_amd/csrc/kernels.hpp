// Launchers shared between the translation units of the library.
#pragma once

#include <array>
#include <functional>
#include <limits>

#include "common.hpp"
#include "host_tables.hpp"

namespace exh {

// field_lpt.hip
template <typename T>
void launch_field(exh_ctx* c, const exh_box* b, const exh_field_src* s, double k_smooth,
                  void* delta, void* deltak_raw,
                  const std::function<void()>& after_modes = {});
template <typename T>
void launch_lpt(exh_ctx* c, const exh_box* b, const exh_field_src* s, double k_smooth,
                int do_2lpt, int out_pos, T* S, T* V, T* P = nullptr);
template <typename T>
void launch_field_adjoint(exh_ctx* c, const exh_box* b, const exh_field_src* s, double k_smooth,
                          const void* gk, T* gnoise);
// adjoint.hip: dL/dnoise of the monopole P(k) of the LPT particles of the
// modes (gP: host [Nk] dL/dP_b)
template <typename T>
void run_pk_adjoint(exh_ctx* c, const exh_box* b, const exh_field_src* s, double k_smooth,
                    int do_2lpt, int window, int Nk, double k_min, double k_max,
                    int kbin_precision, const double* gP, T* gnoise, double* P_out,
                    long long* count_out);

// halos.hip
struct HaloResult {
  long long nh;
};
template <typename T>
long long run_find_halos(exh_ctx* c, const exh_box* b, const T* delta, const exh_halo_opts* o,
                         const std::function<void(int)>& phase = {});
template <typename T>
void run_halo_posh(exh_ctx* c, const exh_box* b, const T* S, const T* V, int lagrangian);
void run_halo_flag(exh_ctx* c, const exh_box* b, long long* flag_dev);

// sort.hip: Lomuto quicksort emulation (ascending, "<=" left, pivot last).
// With m keys below `floor`, positions m.. hold exactly what the full sort
// puts there; positions below m are unspecified (segments of keys below
// `floor` are not sorted further).
template <typename K>
void lomuto_sort(exh_ctx* c, K* keys, int* vals, long long n,
                 double floor = -std::numeric_limits<double>::infinity());

// spectrum.hip
// scratch_dev: device memory of NOALLOC_GRID_WORDS words for the deposit's
// counters instead of the context's buffer (the allocation-free entry point)
template <typename T>
void run_grid(exh_ctx* c, const T* pos, const T* vel, const T* mass, const int* type,
              long long np, int ntype, int nd, double L, double Om0, double z, int direction,
              int window, double R, double R_times, int interlacing, int folds, T* grid,
              void* scratch_dev = nullptr);
// The deposit scratch of the allocation-free entry point on `device` (which
// must be current): the address of a static device array, looked up once
// per device.
constexpr int NOALLOC_GRID_WORDS = 128;
constexpr int NOALLOC_MAX_DEVICES = 64;
void* noalloc_grid_scratch(int device);
template <typename T>
void run_power(exh_ctx* c, const T* grid, int ntype, int nd, double L, int window, double R,
               int interlacing, int Nk, double k_min, double k_max, int l_max, int direction,
               int kbin_precision, double* k_out, double* P_out, long long* count_out,
               bool deferred = false);
// differential halo abundance of masses Mh (device, nh); outputs on the host
template <typename T>
void run_abundance(exh_ctx* c, const T* Mh, long long nh, double Mmin, double Mmax, int Nm,
                   double Lc, int ndx, int ndy, int ndz, double* Mmean_out, double* dn_out,
                   double* dn_err_out);
// bispectra and shell power spectra (bi_compute); returns the number of
// configurations written (cap: capacity of kB, Bk, Ntri_out)
template <typename T>
int run_bispectrum(exh_ctx* c, const T* grid, int ntype, int nd, double L, int window, double R,
                   int interlacing, int Nk, double k_min, double k_max, long long cap,
                   double* kP, double* Pk, double* Nk_out, double* kB, double* Bk,
                   double* Ntri_out);
void bispectrum_configs(int Nk, std::vector<std::array<int, 3>>& tri);
// adjoints: dL/dgrid of the monopole P(k) of one grid (gP: host [Nk]), and
// dL/dS of the LPT particles (nL^3 lattice, spacing Lc, displacements S)
// of a normalised unweighted grid from dL/dgrid
template <typename T>
void run_power_adjoint(exh_ctx* c, const T* grid, int nd, double L, int window, double R,
                       int Nk, double k_min, double k_max, int kbin_precision, const double* gP,
                       T* g, double* P_out, long long* count_out);
template <typename T>
void run_deposit_adjoint(exh_ctx* c, const T* S, int nL, double Lc, int nd, double L, int window,
                         const T* g, T* gS);
// host outputs of a deferred run_power (waits for the context's stream)
void power_copy_out(exh_ctx* c, double* k_out, double* P_out, long long* count_out);

}  // namespace exh
