// Input-deck driver (namelist + lattice -> GPU run).  Filled in below.
#include <string>

#include "g4mi_host.h"

namespace g4mi {
void set_error(const std::string &msg);
}

extern "C" int g4mi_run_deck(const char *input_file, int device, int quiet) {
    (void)input_file;
    (void)device;
    (void)quiet;
    g4mi::set_error("g4mi_run_deck: not available in this build");
    return G4MI_E_STATE;
}
