// main.cpp -- the `fluidsim` executable: main() of the reference
// (src/main.cu:17-30) on a case directory.
//
//   fluidsim [case_dir]      (default: the current directory)
//
// The case directory holds config/input.ini and data/*.dat; .vtk files are
// written under <case_dir>/out/{fluid,solid,virtual}.
#include <cstdio>

#include "fluidsim_host.h"

int main(int argc, char **argv) {
    const char *root = argc > 1 ? argv[1] : ".";
    const int rc = fsh_run(root, 1);
    if (rc != FSA_OK) {
        std::fprintf(stderr, "fluidsim: %s\n", fsh_last_error());
        return 1;
    }
    return 0;
}
