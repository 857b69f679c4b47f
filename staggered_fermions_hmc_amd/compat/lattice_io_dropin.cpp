// save_all_lattice_configs / load_all_lattice_configs with the reference's
// exact non-template signatures (Lattice.h:43-44), for the reference's own
// build: CONFIG comes from its Parameters.h (Parameters.h:30) when that is
// on the include path (compile this file inside the reference's build with
// -I<reference dir>), else from -DSFH_CONFIG=<n>.  The byte format and the
// messages are the reference's (Lattice.cpp:810-847); the I/O runs through
// the library (sfh_links_save / sfh_links_load via sfh_compat).
#include <array>
#include <string>
#include <vector>

#if __has_include("Parameters.h")
#include "Parameters.h"
#define SFH_IO_CONFIG CONFIG
#elif defined(SFH_CONFIG)
#define SFH_IO_CONFIG SFH_CONFIG
#else
#error "CONFIG: put the reference's Parameters.h on the include path or define SFH_CONFIG"
#endif

#include "Fermions_gpu.h"

bool save_all_lattice_configs(const std::array<Link_array, SFH_IO_CONFIG>& links_at_coupling,
                              const std::string& filename) {
  return sfh_compat::save_links(links_at_coupling.data(), SFH_IO_CONFIG, filename);
}

bool load_all_lattice_configs(std::array<Link_array, SFH_IO_CONFIG>& links_at_coupling, const std::string& filename) {
  return sfh_compat::load_links(links_at_coupling.data(), SFH_IO_CONFIG, filename);
}
