// Training-set preparation: the reference's offline train.in tools, with
// their compile-time constants turned into options.
//   DFT data -> train.in        tools/construct_train.in/construct_train_in.cpp:45-222
//   train.in cells x 2 x 2 x 2  tools/construct_train.in/expand222.cpp:41-151
// Host only.  The float arithmetic and the "%g" text are the tools' own, so
// the output is the same file byte for byte.
#pragma once

#include <string>

namespace gpuga {

struct ConvertOptions {
  int cells[3] = {4, 4, 3};   // energy/virial cell replicated NA x NB x NC (:47-49)
  float delta_e = 1.31684f;   // energy shift per atom, eV (:52)
  bool stress_in_kbar = true; // VASP stress in kbar -> virial in eV (:51)
};

// list: the tool's standard input (Nc_force Nc_energy, then Nc DFT file
// names; a relative name is taken relative to dir).  Returns train.in.
// Force files: Na, box (9), Na x (type x y z fx fy fz).  Energy files: Na,
// energy, virial (6), box (9), Na x (type x y z).
std::string construct_train_in(const std::string& list, const std::string& dir,
                               const ConvertOptions& opt);

// Every configuration of a train.in replicated cells[0] x cells[1] x cells[2]
// (expand222 is {2, 2, 2}).  Returns the new train.in.
std::string expand_train_in(const std::string& text, const int cells[3]);

} // namespace gpuga
