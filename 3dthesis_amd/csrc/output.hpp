// The Name/Data output layout: <data_dir>/<Name>.<tag>.csv with the column set
// and order of the reference's Grid constructor (Grid.h:120-280), rows in
// ascending point index for points whose output flag is set, values in the
// default 6-significant-digit stream format (Grid.cpp:72-99).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "model.hpp"

namespace thesis {

// Host copy of the per-point state a run reports. Arrays that the output
// selection does not need may be empty.
struct HostState {
  std::vector<double> T, tSol, G, V, Gx, Gy, Gz, dTdt, eqFrac, depth;
  std::vector<uint16_t> numMelt;
  std::vector<uint8_t> out_flag;  // empty: every point is written
};

struct Column {
  const char* name;
  int id;
};
std::vector<Column> output_columns(const Problem& pb);

// Writes the CSV; with full_dir non-empty also "<full_dir>/<Name>.<tag>.full.txt"
// (point index, then every column at 17 significant digits).
void write_output(const Problem& pb, const HostState& st, const std::string& tag, const std::string& full_dir);

}  // namespace thesis
