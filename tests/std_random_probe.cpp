// Test helper: prints the random streams the reference's GA draws from, made
// with the C++ standard library itself (std::mt19937,
// std::uniform_real_distribution<float>, std::uniform_int_distribution<int>),
// so the explicit algorithms of the oracle and of the library can be checked
// against libstdc++.  Built by tests/test_ga_rng.py with g++.
//
// Output lines: "u <float bits as hex>" for U[0,1) floats and "i <int>" for
// integers, in the order of the schedule given on the command line:
//   probe SEED SCHEDULE   where SCHEDULE is a string of tokens
//   f            one uniform_real_distribution<float>(0,1) draw
//   iA:B         one uniform_int_distribution<int>(A,B) draw
//   sN,          N uniform_real_distribution<float>(0,1) draws, not printed
//   mA:B,        one mutation written as the reference writes it,
//                `genes[r1(rng)] = r2(rng);` with r1 uniform_int(A, B) and
//                r2 uniform_real<float>(0, 1) (src/ga.cu:193-197): the order
//                in which the two draws happen is the compiler's (C++14
//                leaves it unspecified, C++17 draws the value first); prints
//                "m <index> <value bits as hex>"
//   cM:G:g:R,    the mutation count of generation g, M genes, G generations,
//                rate R, with the reference's expression types
//                (src/ga.cu:190-192): "c <count>"
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

int main(int argc, char** argv)
{
  if (argc < 3) return 2;
  std::mt19937 rng((unsigned)std::strtoul(argv[1], nullptr, 10));
  std::string sched = argv[2];
  size_t p = 0;
  while (p < sched.size()) {
    if (sched[p] == 'f') {
      std::uniform_real_distribution<float> d(0, 1);
      float x = d(rng);
      unsigned u;
      std::memcpy(&u, &x, 4);
      std::printf("u %08x\n", u);
      ++p;
    } else if (sched[p] == 'i') {
      size_t q = sched.find(',', p);
      if (q == std::string::npos) q = sched.size();
      int a = 0, b = 0;
      std::sscanf(sched.c_str() + p + 1, "%d:%d", &a, &b);
      std::uniform_int_distribution<int> d(a, b);
      std::printf("i %d\n", d(rng));
      p = q;
    } else if (sched[p] == 's') {
      size_t q = sched.find(',', p);
      if (q == std::string::npos) q = sched.size();
      long n = std::strtol(sched.c_str() + p + 1, nullptr, 10);
      std::uniform_real_distribution<float> d(0, 1);
      for (long k = 0; k < n; ++k) (void)d(rng);
      p = q;
    } else if (sched[p] == 'm') {
      size_t q = sched.find(',', p);
      if (q == std::string::npos) q = sched.size();
      int a = 0, b = 0;
      std::sscanf(sched.c_str() + p + 1, "%d:%d", &a, &b);
      static std::vector<float> genes;
      if ((int)genes.size() <= b) genes.assign(b + 1, -1.0f);
      std::uniform_int_distribution<int> r1(a, b);
      std::uniform_real_distribution<float> r2(0, 1);
      genes[r1(rng)] = r2(rng);
      for (int k = a; k <= b; ++k)
        if (genes[k] != -1.0f) {
          unsigned u;
          std::memcpy(&u, &genes[k], 4);
          std::printf("m %d %08x\n", k, u);
          genes[k] = -1.0f;
        }
      p = q;
    } else if (sched[p] == 'c') {
      size_t q = sched.find(',', p);
      if (q == std::string::npos) q = sched.size();
      int m = 0, G = 1, g = 0;
      float mutation_rate = 0.0f;
      std::sscanf(sched.c_str() + p + 1, "%d:%d:%d:%f", &m, &G, &g, &mutation_rate);
      float rate_g = mutation_rate * (1.0 - float(g) / G);
      int count = round(m * rate_g);
      std::printf("c %d\n", count);
      p = q;
    } else {
      ++p;
    }
  }
  return 0;
}
