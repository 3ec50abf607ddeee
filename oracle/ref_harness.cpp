// Golden-vector harness for the reference 3DThesis library.
//
// This file is test infrastructure. It is compiled together with the
// reference's own library sources (src/*.cpp, read in place under the
// reference checkout, never copied) by oracle/Makefile into oracle/_ref/.
// It only calls the reference's public interfaces:
//   Init::GetFileNames / Init::ReadSimParams   (Init.cpp:233-303)
//   Grid::Grid / Grid::Calc_T / Output_Table   (Grid.h:120-290, Grid.cpp:54-70, :275-314)
//   Calc::Integrate_Serial                     (Calc.cpp:48-65)
//   Run::Simulate                              (Run.cpp:27-57)
//   Grid::Output_T_hist, get_t_hist/get_T_hist (Grid.cpp:101-174, Grid.h:357-358)
//   Grid::Output_RDF, Output_RRDF_csv and their getters (Grid.cpp:176-239, Grid.h:360-370)
// and mirrors the CLI flow of bin/Main.cpp:44-124 (which itself cannot be
// compiled here without CMake-generated ThesisConfig.h).
//
// Sub-commands:
//   run    <ParamInput>                 full run like the CLI; prints
//                                       "Execution time (s): X" and writes the
//                                       reference CSVs, plus <Data>/<Name>.final.full.txt
//                                       (every output row, %.17g, with its point index)
//                                       and, with T_hist on, <Data>/<Name>.T_hist.full.txt
//                                       (p t T per recorded evaluation, %.17g); with RDF
//                                       <Name>.RDF.full.txt (p tm tl cr), with Stork
//                                       <Name>.RRDF.full.txt (ci cj ck t_prev t_cur T x16)
//   nodes  <ParamInput> <t> <isSol> <out>   node list from Integrate_Serial
//   tgrid  <ParamInput> <t> <out>           nodes at t, then Calc_T at every grid point
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "Calc.h"
#include "DataStructs.h"
#include "Grid.h"
#include "Init.h"
#include "Run.h"
#include "Util.h"

namespace {

void load(Simdat& sim, const std::string& input) {
  sim.print = false;
  Init::GetFileNames(sim.files, input, sim.print);
  Init::ReadSimParams(sim);
}

int cmd_run(const std::string& input) {
  Simdat sim;
  load(sim, input);
  Grid grid(sim);
  sim.util.approxEndTime = sim.util.allScansEndTime;
  auto t0 = std::chrono::high_resolution_clock::now();
  Run::Simulate(grid, sim);
  auto t1 = std::chrono::high_resolution_clock::now();
  std::printf("Execution time (s): %.6f\n",
              std::chrono::duration<double>(t1 - t0).count());
  if (sim.param.mode == "Solidification") {  // Main.cpp:104-112
    if (sim.param.tracking != "Stork") {
      grid.Output(sim, "Solidification.Final");
    } else {
      grid.Output_RRDF_csv(sim, "RRDF");
      const std::string rr = sim.files.dataDir + "/" + sim.files.name + ".RRDF.full.txt";
      FILE* r = std::fopen(rr.c_str(), "w");
      if (!r) { std::perror("open RRDF dump"); return 2; }
      const std::vector<uint32_t>& idx = grid.get_RRDF_idxs();
      const std::vector<double>& ts = grid.get_RRDF_ts();
      const std::vector<double>& Ts = grid.get_RRDF_Ts();
      for (size_t e = 0; e < ts.size() / 2; ++e) {
        std::fprintf(r, "%u %u %u %.17g %.17g", idx[3 * e], idx[3 * e + 1], idx[3 * e + 2], ts[2 * e], ts[2 * e + 1]);
        for (int n = 0; n < 16; ++n) std::fprintf(r, " %.17g", Ts[16 * e + n]);
        std::fprintf(r, "\n");
      }
      std::fclose(r);
    }
  }
  const std::string full = sim.files.dataDir + "/" + sim.files.name + ".final.full.txt";
  FILE* f = std::fopen(full.c_str(), "w");
  if (!f) { std::perror("open full dump"); return 2; }
  std::vector<std::vector<double>> tab = grid.Output_Table(sim);
  for (size_t p = 0; p < tab.size(); ++p) {
    if (tab[p].empty()) continue;
    std::fprintf(f, "%zu", p);
    for (double v : tab[p]) std::fprintf(f, " %.17g", v);
    std::fprintf(f, "\n");
  }
  std::fclose(f);
  if (sim.output.T_hist) {  // Main.cpp:113-115, then every recorded (t, T) in full precision
    grid.Output_T_hist(sim, "T.hist");
    const std::string hist = sim.files.dataDir + "/" + sim.files.name + ".T_hist.full.txt";
    FILE* h = std::fopen(hist.c_str(), "w");
    if (!h) { std::perror("open T_hist dump"); return 2; }
    for (int p = 0; p < sim.domain.pnum; ++p) {
      const std::vector<double> ts = grid.get_t_hist(p), Ts = grid.get_T_hist(p);
      for (size_t e = 0; e < ts.size(); ++e) std::fprintf(h, "%d %.17g %.17g\n", p, ts[e], Ts[e]);
    }
    std::fclose(h);
  }
  if (sim.output.RDF && sim.param.mode == "Solidification") {  // Main.cpp:116-118, then every melt event
    grid.Output_RDF(sim, "RDF.Final");
    const std::string rdf = sim.files.dataDir + "/" + sim.files.name + ".RDF.full.txt";
    FILE* h = std::fopen(rdf.c_str(), "w");
    if (!h) { std::perror("open RDF dump"); return 2; }
    for (int p = 0; p < sim.domain.pnum; ++p) {
      if (!grid.get_output_flag(p)) continue;
      const std::vector<double> tm = grid.get_RDF_tm(p), tl = grid.get_RDF_tl(p), cr = grid.get_RDF_cr(p);
      for (size_t e = 0; e < tm.size(); ++e)
        std::fprintf(h, "%d %.17g %.17g %.17g\n", p, tm[e], e < tl.size() ? tl[e] : 0.0, e < cr.size() ? cr[e] : 0.0);
    }
    std::fclose(h);
  }
  return 0;
}

int cmd_nodes(const std::string& input, double t, int is_sol, const std::string& out) {
  Simdat sim;
  load(sim, input);
  Nodes nodes;
  Calc::Integrate_Serial(nodes, sim, t, is_sol != 0);
  FILE* f = std::fopen(out.c_str(), "w");
  if (!f) { std::perror("open node dump"); return 2; }
  std::fprintf(f, "%zu\n", nodes.size);
  for (size_t n = 0; n < nodes.size; ++n) {
    std::fprintf(f, "%.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n",
                 nodes.xb[n], nodes.yb[n], nodes.zb[n], nodes.phix[n],
                 nodes.phiy[n], nodes.phiz[n], nodes.dtau[n], nodes.expmod[n]);
  }
  std::fclose(f);
  return 0;
}

int cmd_tgrid(const std::string& input, double t, const std::string& out) {
  Simdat sim;
  load(sim, input);
  Grid grid(sim);
  Nodes nodes;
  Calc::Integrate_Serial(nodes, sim, t, false);
  FILE* f = std::fopen(out.c_str(), "w");
  if (!f) { std::perror("open T dump"); return 2; }
  std::fprintf(f, "%d %zu\n", sim.domain.pnum, nodes.size);
  for (int p = 0; p < sim.domain.pnum; ++p) {
    std::fprintf(f, "%.17g\n", grid.Calc_T(t, nodes, sim, false, p));
  }
  std::fclose(f);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s run|nodes|tgrid <ParamInput> ...\n", argv[0]);
    return 1;
  }
  const std::string cmd = argv[1];
  if (cmd == "run") return cmd_run(argv[2]);
  if (cmd == "nodes" && argc == 6) return cmd_nodes(argv[2], std::atof(argv[3]), std::atoi(argv[4]), argv[5]);
  if (cmd == "tgrid" && argc == 5) return cmd_tgrid(argv[2], std::atof(argv[3]), argv[4]);
  std::fprintf(stderr, "bad arguments\n");
  return 1;
}
