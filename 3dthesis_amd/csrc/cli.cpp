// `thesis` command line: the drop-in for the reference's `3DThesis <ParamInput.txt>`
// (bin/Main.cpp:44-124). Same input files, same Data/<Name>.<tag>.csv outputs.
//
//   thesis ParamInput.txt [--fp32c] [--fabs] [--full DIR] [--progress]
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/thesis_amd.h"

int main(int argc, char** argv) {
  if (argc <= 1) {
    std::fprintf(stderr, "Input file argument required: thesis ParamInput.txt\n");
    return 1;
  }
  const char* input = nullptr;
  int precision = THESIS_FP64, abs_mode = 0, progress = 0;
  const char* full = nullptr;
  const char* usage = "usage: thesis ParamInput.txt [--fp32c] [--fabs] [--full DIR] [--progress]\n";
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--fp32c")) precision = THESIS_FP32C;
    else if (!std::strcmp(argv[i], "--fabs")) abs_mode = 1;
    else if (!std::strcmp(argv[i], "--progress")) progress = 1;
    else if (!std::strcmp(argv[i], "--full")) {
      if (i + 1 >= argc || argv[i + 1][0] == '-') {
        std::fprintf(stderr, "--full needs a directory\n%s", usage);
        return 1;
      }
      full = argv[++i];
    } else if (argv[i][0] == '-' && argv[i][1] == '-') {
      std::fprintf(stderr, "unknown option %s\n%s", argv[i], usage);
      return 1;
    } else if (input) {
      std::fprintf(stderr, "more than one input file (%s, %s)\n%s", input, argv[i], usage);
      return 1;
    } else {
      input = argv[i];
    }
  }
  if (!input) {
    std::fprintf(stderr, "Input file argument required\n%s", usage);
    return 1;
  }
  const auto t0 = std::chrono::steady_clock::now();
  thesis_ctx* ctx = nullptr;
  int rc = thesis_ctx_create(input, &ctx);
  if (rc == THESIS_OK) rc = thesis_set_option(ctx, "precision", precision);
  if (rc == THESIS_OK) rc = thesis_set_option(ctx, "abs_mode", abs_mode);
  if (rc == THESIS_OK) rc = thesis_set_option(ctx, "progress", progress);
  if (rc == THESIS_OK && full) rc = thesis_set_full_dump(ctx, full);
  const auto t1 = std::chrono::steady_clock::now();
  if (rc != THESIS_OK) {
    std::fprintf(stderr, "Fatal Error: %s\n", thesis_last_error());
    thesis_ctx_destroy(ctx);
    return 1;
  }
  std::printf("Initialization time (s): %g\n\n", std::chrono::duration<double>(t1 - t0).count());
  rc = thesis_run(ctx);
  if (rc != THESIS_OK) {
    std::fprintf(stderr, "Fatal Error: %s\n", thesis_last_error());
    thesis_ctx_destroy(ctx);
    return 1;
  }
  thesis_stats st{};
  thesis_get_stats(ctx, &st);
  std::printf("Execution time (s): %g\n\n", st.wall_s);
  std::printf("Point-node evaluations: %.6g (%.4g /s)\n", st.evals, st.wall_s > 0 ? st.evals / st.wall_s : 0.0);
  thesis_ctx_destroy(ctx);
  return 0;
}
