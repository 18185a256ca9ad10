// gpuga command-line driver: `gpuga < input.txt`, as src/main.cu:25-80.
// stdin holds the number of job directories, then the directories; each is
// fitted in turn (ga.in, potential.in, train.in in; ga.out, force.out,
// energy.out, virial.out out).
#include <cstdio>
#include <cstdlib>
#include <ctime>

#include "../../include/gpuga.h"

static void line1() { std::printf("\n---------------------------------------------------------------\n"); }
static void line2() { std::printf("---------------------------------------------------------------\n\n"); }

static void print_error(const char* s)
{
  std::printf("ERROR: %s", s);
  std::exit(EXIT_FAILURE);
}

int main()
{
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
