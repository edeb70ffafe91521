// multiphaselbm <inputfile>: the reference command line (src/main.cc:4-17)
// on the MI355X library.
#include <cstdio>

#include "mplbm_host.h"

int main(int argc, char const* argv[]) {
  if (argc != 2) {
    std::printf("Usage: %s commandFile\n", argv[0]);
    return 0;
  }
  return mplbm_host_run(argv[1]) == 0 ? 0 : 1;
}
