// Host-side parameter derivation and tables; see host_tables.cpp.
#pragma once

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace exh {

int box_init(exh_box* b, int ndx, int ndy, int ndz, double Lc_or_Mc, double Om0, double z,
             double dc, int precision);

void spline_second_derivs(const double* x, const double* y, int n, double* m);
double spline_eval_host(const double* x, const double* y, const double* m, int n, double xv,
                        int* err);

// Returns Ncells (>= 0), or -1 when sigma(M) was evaluated outside its knots.
// Knots of the sigma(M) and Tinker tables: sigma_radii, or Nk for the
// reference's table.
inline int halo_table_radii(const exh_halo_opts* o) {
  return o->sigma_radii > 0 ? o->sigma_radii : o->Nk;
}
int halo_tables(const exh_box* b, const exh_halo_opts* o, void* sig_grid, void* Bt, int nB,
                double* Mknots, double* Int);

template <typename R>
R barrier_value(R S, int do_eb, R dc, R a, R beta, R alpha);

std::vector<int> cells_in_spheres(int smax);
void shell_offsets(int smax, int tmax, std::vector<int>& start, std::vector<unsigned>& packed);

// Inter_Power's truncation: xi(r) of the log-spaced P(k) by FFTLog, zero
// beyond R_max, back to P(k) on the transform's grid (k_out, P_out [n]).
void pk_truncate(const double* k, const double* P, int n, double R_max, double* k_out,
                 double* P_out);

void taus_jump_tables(uint32_t* mats);
void taus_seed_state(unsigned long s, uint32_t st[3]);

}  // namespace exh
