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
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>

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
    } else {
      ++p;
    }
  }
  return 0;
}
