// main.cpp -- the `gpue` executable: GPUE's main() (src/init.cu:499-571) over
// the MI355X step.  Same flags, data directory layout and output files.
#include <cstdio>
#include <ctime>
#include <exception>
#include <iostream>

#include "gpue_host.h"

int main(int argc, char** argv) {
  using namespace gpue_host;
  try {
    Grid par = parse_args(argc, argv);
    time_t start, fin;
    time(&start);
    std::printf("Start: %s\n", ctime(&start));
    gpue_ctx* ctx = init(par);
    const int gsteps = par.ival("gsteps"), esteps = par.ival("esteps");
    if (par.bval("write_file")) FileIO::writeOutParam(par, par.sval("data_dir") + "Params.dat");
    std::string buffer;
    if (gsteps > 0) {
      std::cout << "Imaginary-time evolution started..." << '\n';
      set_variables(par, ctx, 0);
      evolve(par, ctx, gsteps, 0, buffer);
    }
    if (esteps > 0) {
      std::cout << "real-time evolution started..." << '\n';
      set_variables(par, ctx, 1);
      evolve(par, ctx, esteps, 1, buffer);
    }
    std::cout << "done evolving" << '\n';
    gpue_destroy(ctx);
    time(&fin);
    std::printf("Finish: %s\n", ctime(&fin));
    std::printf("Total time: %ld seconds\n ", (long)(fin - start));
    std::cout << '\n';
  } catch (const std::exception& e) {
    std::fprintf(stderr, "gpue: %s\n", e.what());
    return 1;
  }
  return 0;
}
