// Drop-in replacement of the reference's field-level Fermions.h / Lattice.h
// functions, implemented on the MI355X library through the C ABI (sfh.h).
//
// Same names, signatures, argument meaning, ownership and exceptions as the
// reference (Fermions.h:31-70, :105, :130; Lattice.h:75-77; Smearing.h:29-31):
//   - inputs are const& host vectors in the reference layouts, results are
//     returned by value with exact sizes 6V or 4V*18;
//   - apply_DDdagger writes into the caller's pre-sized out_field;
//   - the leapfrog stages mutate P_field and the links in place;
//   - CG failures and the hermiticity guard throw std::runtime_error, the
//     exponential's guards std::invalid_argument, with the reference's texts.
// The reference's compile-time constants (Parameters.h:9-24, :51-52) are
// runtime values here: set them once with sfh_compat::configure().
// Each host thread gets its own library context (the reference calls these
// from inside an OpenMP loop over coupling values, Main.cpp:174).
#pragma once

#include <array>
#include <complex>
#include <cstddef>
#include <string>
#include <utility>
#include <vector>

#include "sfh_compat.h"

using Link_array = std::vector<double>;
using complex = std::complex<double>;

std::vector<double> generate_phi_field(const Link_array& arr);
std::vector<double> generate_conjugate_field_configuration();
double field_inner_product_real(const std::vector<double>& A, const std::vector<double>& B);
void apply_DDdagger(const std::vector<double>& in_field, const Link_array& arr, std::vector<double>& out_field);
std::vector<double> solve_CG_DDdagger(const std::vector<double>& phi_field, const Link_array& arr);
std::vector<double> solve_X_field(const std::vector<double>& phi_field, const Link_array& arr);
std::vector<double> solve_Y_field(const std::vector<double>& X_field, const Link_array& arr);
std::vector<double> compute_full_force_field(const std::vector<double>& X_field, const std::vector<double>& Y_field,
                                             const Link_array& arr, double beta);
std::vector<double> compute_fermion_only_force_field(const std::vector<double>& X_field,
                                                     const std::vector<double>& Y_field, const Link_array& arr);
std::vector<double> compute_gauge_only_force_field(const Link_array& arr, double beta);
double compute_Tr_P2(const std::vector<double>& P_field);
void apply_initial_step_to_conjugate_field(std::vector<double>& P_field, double beta,
                                           const std::vector<double>& phi_field, const Link_array& arr);
void apply_intermediate_update_to_gauge_field_and_conjugate_field(std::vector<double>& P_field, double beta,
                                                                  const std::vector<double>& phi_field,
                                                                  Link_array& arr);
std::vector<double> apply_final_step(std::vector<double>& P_field, double beta, const std::vector<double>& phi_field,
                                     Link_array& arr);
void Monte_carlo_step(double initial_Tr_P2, double initial_phi_field_action, const std::vector<double>& P_field_final,
                      const std::vector<double>& phi_field, double gauge_field_initial_action,
                      Link_array& gauge_field_initial, Link_array& gauge_field_final,
                      const std::vector<double>& X_field, double beta);
void single_sweep_monte_carlo_update_WITH_FERMIONS(Link_array& gauge_field, double beta);

// eps / n_steps copies of the stages (Fermions.h:131-141)
void local_initial_half_step(std::vector<double>& P_field, double beta, const std::vector<double>& phi_field,
                             const Link_array& arr, double eps);
void local_leapfrog_steps(std::vector<double>& P_field, double beta, const std::vector<double>& phi_field,
                          Link_array& arr, double eps, int n_steps);
std::vector<double> local_final_step(std::vector<double>& P_field, double beta, const std::vector<double>& phi_field,
                                     Link_array& arr, double eps);

double compute_action(const Link_array& arr, double beta);
double compute_avg_plaq(const Link_array& arr);
// One Cabibbo-Marinari heatbath sweep of the Symanzik action (c_11, c_12 of
// the configuration) in place; returns (average plaquette, action density)
// as the reference does.  Sweeps of one thread take consecutive sweep
// indices of the configured seed's Philox streams; the links are visited in
// the device's colour order, so a sweep differs from the reference's
// lexicographic one draw for draw but samples the same distribution.
std::pair<double, double> heatbath_update(Link_array& arr, double beta);

// HEX smearing in place with alpha1..3 of the configuration (Smearing.h:7-9),
// each level V = exp(proj(alpha/2 Gamma U^dagger)) U as the reference writes
// it; the exponential is exp of that anti-hermitian matrix (the reference's
// exp(iQ) routine rejects it) and the level-3 sum starts at zero.
void Apply_HEX_smearing(Link_array& U_array);
void Apply_4HEX_smearing(Link_array& U_array);
std::vector<Link_array> Apply_4HEX_smearing_with_history(const Link_array& U_thin);

// Measurement step of Main.cpp:219-228 (Lattice.h:84-85): the Polyakov grid
// Tr prod_t U_3 at every spatial site [(x*L + y)*L + z] (Lattice.cpp:868-884)
// and the correlator over the displacements with d.d = r_squared
// (Lattice.cpp:886-948), both computed on the device.
std::vector<complex> precompute_polyakov_grid(const Link_array& arr);
complex correlator_over_fixed_distance_fast(const std::vector<complex>& P_grid, int r_squared);

// Checkpoints in the reference's byte format (Lattice.h:43-44,
// Lattice.cpp:810-847): the CONFIG arrays back to back as raw doubles; false
// (and the reference's message on std::cerr) when the file cannot be opened
// or written / read completely.  CONFIG is deduced from the array.
namespace sfh_compat {
bool save_links(const Link_array* arrays, size_t n, const std::string& filename);
bool load_links(Link_array* arrays, size_t n, const std::string& filename);
}  // namespace sfh_compat

template <size_t N>
bool save_all_lattice_configs(const std::array<Link_array, N>& links_at_coupling, const std::string& filename) {
  return sfh_compat::save_links(links_at_coupling.data(), N, filename);
}
template <size_t N>
bool load_all_lattice_configs(std::array<Link_array, N>& links_at_coupling, const std::string& filename) {
  return sfh_compat::load_links(links_at_coupling.data(), N, filename);
}
