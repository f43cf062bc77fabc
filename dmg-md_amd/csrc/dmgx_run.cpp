// dmgx_run.cpp -- command-line runner: `dmgx-run [directory [run_file [model_file]]]`
// runs a run.in / model.xyz input directory (the reference's dmg-md
// executable, src/dmgmd_main.cpp, on one GPU) through the C ABI.
#include <cstdio>

#include "dmgx.h"

int main(int argc, char** argv)
{
  if (argc > 4 || (argc > 1 && (argv[1][0] == '-'))) {
    std::fprintf(stderr, "usage: dmgx-run [directory [run_file [model_file]]]\n");
    return 2;
  }
  const char* dir = argc > 1 ? argv[1] : ".";
  const int st = dmgx_run_directory(dir, argc > 2 ? argv[2] : nullptr, argc > 3 ? argv[3] : nullptr);
  if (st != DMGX_OK) {
    std::fprintf(stderr, "dmgx-run: %s\n", dmgx_last_error());
    return 1;
  }
  return 0;
}
