// gpuga command-line driver: `gpuga < input.txt`, as src/main.cu:25-80.
// stdin holds the number of job directories, then the directories; each is
// fitted in turn (ga.in, potential.in, train.in in; ga.out, force.out,
// energy.out, virial.out out).
//
// The training-set tools run as the reference's do, in the current directory:
//   gpuga construct_train_in [--cells NA NB NC] [--delta-e E] [--stress-ev] < input.txt
//       writes train.in (tools/construct_train.in/construct_train_in.cpp)
//   gpuga expand222 [--cells NA NB NC]
//       reads train.in, writes train.out (tools/construct_train.in/expand222.cpp)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>

#include "../../include/gpuga.h"

static void line1() { std::printf("\n---------------------------------------------------------------\n"); }
static void line2() { std::printf("---------------------------------------------------------------\n\n"); }

static void print_error(const char* s)
{
  std::printf("ERROR: %s", s);
  std::exit(EXIT_FAILURE);
}

static int tool(int argc, char** argv)
{
  gpuga_convert_options o;
  gpuga_default_convert_options(&o);
  int cells[3] = {2, 2, 2};
  const bool construct = std::strcmp(argv[1], "construct_train_in") == 0;
  if (construct) {
    for (int k = 0; k < 3; ++k) cells[k] = o.cells[k];
  } else if (std::strcmp(argv[1], "expand222") != 0) {
    std::fprintf(stderr, "usage: gpuga < input.txt | gpuga construct_train_in [--cells NA NB NC] "
                         "[--delta-e E] [--stress-ev] < input.txt | gpuga expand222 [--cells NA NB NC]\n");
    return EXIT_FAILURE;
  }
  for (int i = 2; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--cells") && i + 3 < argc) {
      for (int k = 0; k < 3; ++k) cells[k] = std::atoi(argv[++i]);
    } else if (construct && !std::strcmp(argv[i], "--delta-e") && i + 1 < argc) {
      o.delta_e = std::strtof(argv[++i], nullptr);
    } else if (construct && !std::strcmp(argv[i], "--stress-ev")) {
      o.stress_in_kbar = 0;
    } else {
      std::fprintf(stderr, "unknown option %s\n", argv[i]);
      return EXIT_FAILURE;
    }
  }
  int rc;
  if (construct) {
    for (int k = 0; k < 3; ++k) o.cells[k] = cells[k];
    std::string list;
    char buf[4096];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, stdin)) > 0) list.append(buf, n);
    rc = gpuga_construct_train_in(list.c_str(), ".", "train.in", &o);
  } else {
    rc = gpuga_expand_train_in("train.in", "train.out", cells);
  }
  if (rc != 0) {
    std::fprintf(stderr, "%s\n", gpuga_last_error());
    return EXIT_FAILURE;
  }
  return EXIT_SUCCESS;
}

int main(int argc, char** argv)
{
  if (argc > 1) return tool(argc, argv);
  std::printf("\n***************************************************************\n");
  std::printf("*   GPUGA fitness evaluation on AMD Instinct (HIP, gfx950)    *\n");
  std::printf("***************************************************************\n\n");
  int n = 0;
  if (std::scanf("%d", &n) != 1) print_error("reading error for number of inputs.\n");
  gpuga_options opt;
  gpuga_default_options(&opt);
  if (const char* b = std::getenv("GPUGA_BACKEND")) opt.backend = std::atoi(b);
  for (int k = 0; k < n; ++k) {
    char dir[4096];
    if (std::scanf("%4095s", dir) != 1) print_error("reading error for input directory.\n");
    line1();
    std::printf("Run simulation for '%s'.\n", dir);
    line2();
    const std::clock_t t0 = std::clock();
    double spg = 0.0;
    if (gpuga_run(dir, &opt, &spg) != 0) {
      std::printf("ERROR: %s\n", gpuga_last_error());
      return EXIT_FAILURE;
    }
    const float used = (std::clock() - t0) / float(CLOCKS_PER_SEC);
    line1();
    std::printf("Time used for '%s' = %f s (%.6f s per generation).\n", dir, used, spg);
    line2();
  }
  line1();
  std::printf("Finished running gpuga.\n");
  line2();
  return EXIT_SUCCESS;
}
