// Runtime configuration of the drop-in layer: the reference's compile-time
// constants (Parameters.h:9-24, :51-52) as values set once at start-up.
// This header is safe to include next to the reference's own headers.
#pragma once

namespace sfh_compat {

struct Config {
  int Spatial_Size = 4;              // Parameters.h:9
  int temporal_size = 4;             // Parameters.h:10
  double fermion_mass = 1.0;         // Parameters.h:18
  double epsilon = 1e-2;             // Parameters.h:21
  int number_of_steps_fermion = 100; // Parameters.h:22
  double tol = 1e-12;                // Parameters.h:23
  int max_iter = 50000;              // Parameters.h:24
  double c_11 = 5.0 / 3.0;           // Parameters.h:51
  double c_12 = -1.0 / 12.0;         // Parameters.h:52
  double alpha1 = 0.95;              // Smearing.h:7
  double alpha2 = 0.76;              // Smearing.h:8
  double alpha3 = 0.38;              // Smearing.h:9
  int precision = 0;                 // 0 FP64 (reference), 1 FP32 inner, 2 FP16 inner
  int device = 0;
  unsigned long long seed = 0x5f3759dfULL;  // Philox key of the field generators
  bool verbose = true;               // print the H components like Fermions.cpp:798-804
};

void configure(const Config& c);
const Config& config();
int volume();
// Random streams: every field generator, heatbath sweep and Metropolis
// uniform of a host thread draws from streams keyed by (Config::seed, the
// thread's stream index, a per-thread counter), so a run with the same seed
// gives the same fields in every process.  The stream index is the OpenMP
// thread number when called from an OpenMP region (one stream per OpenMP
// thread, as the reference's per-thread generators, rng.cpp:8-16), else the
// order in which threads first draw; set_thread_stream(k) fixes it to k for
// the calling thread and restarts its counters.
void set_thread_stream(int k);

}  // namespace sfh_compat
