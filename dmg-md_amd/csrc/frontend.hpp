// frontend.hpp -- the run.in / model.xyz front end and the output writers.
//
// Runs an input directory the way the reference's executable does
// (src/dmgmd_main.cpp, run_m1_replicated in src/runtime.cu): run.in is a list
// of commands (run_parser.cpp:158-315); model.xyz an extended-XYZ model
// (model_parser.cpp:279-416); every `run N` is one MD segment with the
// measurements declared since the previous run; thermo.out, dump_xyz files,
// restart.xyz and neighbor.out are written in the reference's formats
// (runtime.cu:1148-1313, nep.cu:1156-1176).
#pragma once

#include <stdexcept>
#include <string>

namespace dmgx {

// An input error with its file and line.
struct InputError : std::invalid_argument {
  using std::invalid_argument::invalid_argument;
};

// Runs <dir>/<run_file> with <dir>/<model_file>; the potential path in run.in
// and every output file are relative to dir.
void run_directory(const std::string& dir, const std::string& run_file = "run.in",
                   const std::string& model_file = "model.xyz");

}  // namespace dmgx
