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
  std::vector<uint16_t> numMelt;  // with RDF: the number of recorded melt times
  std::vector<double> mp_w, mp_l, mp_d;  // MP_Stats (empty: zeros)
  std::vector<uint8_t> out_flag;  // empty: every point is written
};

// RDF events of a run, per point in event order (Grid.h:103-105).
struct RdfEvents {
  std::vector<std::vector<double>> tm, tl, cr;
};

struct Column {
  const char* name;
  int id;
};
std::vector<Column> output_columns(const Problem& pb);

// Writes the CSV; with full_dir non-empty also "<full_dir>/<Name>.<tag>.full.txt"
// (point index, then every column at 17 significant digits).
void write_output(const Problem& pb, const HostState& st, const std::string& tag, const std::string& full_dir);

// Grid::Output_T_hist (Grid.cpp:101-174) for histories recorded at the times
// t[e] for every point (T[e * pnum + p]). With the tracking keyword None: one
// table "<Name>T.hist.csv" (no dot after the name), a column per time, times
// and temperatures through std::to_string ("%f"), coordinates in the 6-digit
// stream format. Otherwise "<Name>.Points.csv" (p,x,y,z of the points whose
// output flag is set; empty flags: every point) and "<Name>.T_hist.<p>.csv"
// (t,T) per such point. With full_dir non-empty also
// "<full_dir>/<Name>.T_hist.full.txt" (p t T, 17 digits).
// Grid::Output_RDF (Grid.cpp:176-201): "<Name>.RDF.Final.csv", one row
// x,y,z,tm,tl,cr per recorded melt time of each output-flagged point. A melt
// still open at the end has no liquidus time in the reference (it reads past
// its list); 0 is written there. With full_dir non-empty also
// "<full_dir>/<Name>.RDF.full.txt" (p tm tl cr, 17 digits).
void write_rdf(const Problem& pb, const RdfEvents& ev, const std::vector<uint8_t>& out_flag,
               const std::string& full_dir);

// Grid::Output_RRDF_csv (Grid.cpp:203-239): "<Name>.RRDF.csv", one row per
// Stork record; with full_dir non-empty also "<full_dir>/<Name>.RRDF.full.txt"
// (ci cj ck t_prev t_cur and the 16 temperatures, 17 digits).
void write_rrdf(const Problem& pb, const std::vector<uint32_t>& idx, const std::vector<double>& ts,
                const std::vector<double>& Ts, const std::string& full_dir);

void write_thist(const Problem& pb, const std::vector<double>& t, const std::vector<double>& T,
                 const std::vector<uint8_t>& out_flag, const std::string& full_dir);

// Histories recorded per point by the tracked drivers (Grid.cpp:300-304):
// records sorted by point, each point's in evaluation order.
struct HistRecords {
  std::vector<int> p;
  std::vector<double> t, T;
};
// Grid::Output_T_hist's tracked branch (Grid.cpp:138-172) for such records:
// "<Name>.Points.csv" and "<Name>.T_hist.<p>.csv" per output-flagged point
// (t,T in the 6-digit stream format, no row for a point never evaluated with
// set); with full_dir non-empty also "<full_dir>/<Name>.T_hist.full.txt"
// (p t T, 17 digits, every record).
void write_thist_tracked(const Problem& pb, const HistRecords& h, const std::vector<uint8_t>& out_flag,
                         const std::string& full_dir);

}  // namespace thesis
