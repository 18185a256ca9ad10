// Training-set preparation tools (see convert.h).  Each number is read as
// fscanf's %d / %f read it (strtol / strtof, correctly rounded) and written
// as "%g" of the float, with the tools' own float expressions evaluated in
// the same order (this file is compiled without FMA contraction).
#include "convert.h"

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <vector>

#include "common.h"

namespace gpuga {

namespace {

std::string slurp(const std::string& path)
{
  std::ifstream f(path, std::ios::binary);
  // the tools' my_fopen message (construct_train_in.cpp:31-41)
  if (!f) fail(strfmt("Failed to open %s!\n%s", path.c_str(), std::strerror(errno)));
  std::ostringstream s;
  s << f.rdbuf();
  return s.str();
}

// Whitespace-separated fields of one file, read in order.
class Fields {
public:
  Fields(const std::string& text, std::string name) : s_(text), p_(s_.c_str()), name_(std::move(name)) {}
  int i(const char* what)
  {
    skip();
    char* e = nullptr;
    errno = 0;
    const long v = std::strtol(p_, &e, 10);
    if (e == p_ || errno != 0 || v < -2147483647L - 1 || v > 2147483647L) bad(what);
    p_ = e;
    return (int)v;
  }
  float f(const char* what)
  {
    skip();
    char* e = nullptr;
    const float v = std::strtof(p_, &e);
    if (e == p_) bad(what);
    p_ = e;
    return v;
  }
  std::string word(const char* what)
  {
    skip();
    const char* b = p_;
    while (*p_ && !is_space(*p_)) ++p_;
    if (p_ == b) bad(what);
    return std::string(b, p_);
  }

private:
  static bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }
  void skip()
  {
    while (*p_ && is_space(*p_)) ++p_;
  }
  [[noreturn]] void bad(const char* what) const
  {
    fail(strfmt("Reading error: %s: cannot read %s", name_.c_str(), what));
  }
  std::string s_;
  const char* p_;
  std::string name_;
};

class Text {
public:
  void g(float x, char sep)
  {
    char b[48];
    const int n = std::snprintf(b, sizeof b, "%g%c", (double)x, sep);
    s.append(b, (size_t)n);
  }
  void d(int x, char sep)
  {
    char b[24];
    const int n = std::snprintf(b, sizeof b, "%d%c", x, sep);
    s.append(b, (size_t)n);
  }
  std::string s;
};

std::string join(const std::string& dir, const std::string& name)
{
  if (name.empty() || name[0] == '/' || dir.empty()) return name;
  return dir.back() == '/' ? dir + name : dir + "/" + name;
}

void check_cells(const int cells[3])
{
  for (int k = 0; k < 3; ++k)
    if (cells[k] < 1) fail(strfmt("cell counts must be positive (got %d %d %d)", cells[0], cells[1], cells[2]));
  if ((long)cells[0] * cells[1] * cells[2] > 1000000) fail("too many cells");
}

// image offset n_a a + n_b b + n_c c of a box whose rows are a, b, c
// (construct_train_in.cpp:203-205, expand222.cpp:86-88)
inline void offset(const float* box, int na, int nb, int nc, float* d)
{
  d[0] = box[0] * (float)na + box[3] * (float)nb + box[6] * (float)nc;
  d[1] = box[1] * (float)na + box[4] * (float)nb + box[7] * (float)nc;
  d[2] = box[2] * (float)na + box[5] * (float)nb + box[8] * (float)nc;
}

} // namespace

std::string construct_train_in(const std::string& list, const std::string& dir,
                               const ConvertOptions& opt)
{
  check_cells(opt.cells);
  const int nabc = opt.cells[0] * opt.cells[1] * opt.cells[2];
  Fields in(list, "standard input");
  const int nc_force = in.i("the number of force configurations");
  const int nc_energy = in.i("the number of energy configurations");
  if (nc_force < 0 || nc_energy < 0) fail("configuration counts must not be negative");
  const int nc = nc_force + nc_energy;
  std::vector<std::string> files(nc), texts(nc);
  for (int n = 0; n < nc; ++n) files[n] = join(dir, in.word("a DFT file name"));
  for (int n = 0; n < nc; ++n) texts[n] = slurp(files[n]);

  Text out;
  out.d(nc, ' ');
  out.d(nc_force, '\n');
  // atom counts (construct_train_in.cpp:103-115)
  std::vector<int> na(nc);
  for (int n = 0; n < nc; ++n) {
    Fields f(texts[n], files[n]);
    na[n] = f.i("the number of atoms");
    if (na[n] < 1) fail(strfmt("%s: the number of atoms must be positive", files[n].c_str()));
    if (n >= nc_force && (long)na[n] * nabc > 2147483647L) fail("too many atoms");
    out.d(n >= nc_force ? na[n] * nabc : na[n], '\n');
  }
  // force configurations, copied (:118-150)
  for (int n = 0; n < nc_force; ++n) {
    Fields f(texts[n], files[n]);
    f.i("the number of atoms");
    for (int k = 0; k < 9; ++k) out.g(f.f("the box"), ' ');
    out.s += '\n';
    for (int a = 0; a < na[n]; ++a) {
      out.d(f.i("a type"), ' ');
      for (int k = 0; k < 6; ++k) out.g(f.f("a position or force"), k < 5 ? ' ' : '\n');
    }
  }
  // energy/virial configurations, replicated (:153-222)
  for (int n = nc_force; n < nc; ++n) {
    Fields f(texts[n], files[n]);
    const int m = f.i("the number of atoms");
    float energy = f.f("the energy");
    float virial[6], box[9];
    for (int d = 0; d < 6; ++d) virial[d] = f.f("the virial");
    for (int k = 0; k < 9; ++k) box[k] = f.f("the box");
    energy = ((energy / (float)m) + opt.delta_e) * (float)m * (float)nabc;
    for (int d = 0; d < 6; ++d) virial[d] *= (float)nabc;
    if (opt.stress_in_kbar) {
      const float vol = box[0] * (box[4] * box[8] - box[5] * box[7]) +
                        box[1] * (box[5] * box[6] - box[3] * box[8]) +
                        box[2] * (box[3] * box[7] - box[4] * box[6]);
      // kbar x A^3 -> eV: the ratio is formed in double, the product rounded
      // to float (:186-187)
      const double s = (double)vol / 1602.18;
      for (int d = 0; d < 6; ++d) virial[d] = (float)((double)virial[d] * s);
    }
    out.g(energy, ' ');
    for (int d = 0; d < 6; ++d) out.g(virial[d], d < 5 ? ' ' : '\n');
    for (int k = 0; k < 9; ++k) out.g(box[k] * (float)opt.cells[k / 3], ' ');
    out.s += '\n';
    std::vector<int> type(m);
    std::vector<float> r(3 * (size_t)m);
    for (int a = 0; a < m; ++a) {
      type[a] = f.i("a type");
      for (int k = 0; k < 3; ++k) r[3 * a + k] = f.f("a position");
    }
    for (int ia = 0; ia < opt.cells[0]; ++ia)
      for (int ib = 0; ib < opt.cells[1]; ++ib)
        for (int ic = 0; ic < opt.cells[2]; ++ic) {
          float d[3];
          offset(box, ia, ib, ic, d);
          for (int a = 0; a < m; ++a) {
            out.d(type[a], ' ');
            for (int k = 0; k < 3; ++k) out.g(r[3 * a + k] + d[k], k < 2 ? ' ' : '\n');
          }
        }
  }
  return std::move(out.s);
}

std::string expand_train_in(const std::string& text, const int cells[3])
{
  check_cells(cells);
  const int prod = cells[0] * cells[1] * cells[2];
  Fields in(text, "train.in");
  const int nc = in.i("Nc"), nc_force = in.i("Nc_force");
  if (nc < 0 || nc_force < 0 || nc_force > nc) fail("train.in: bad configuration counts");
  Text out;
  out.d(nc, ' ');
  out.d(nc_force, '\n');
  std::vector<int> na(nc);
  for (int n = 0; n < nc; ++n) {
    na[n] = in.i("Na");
    if (na[n] < 0 || (long)na[n] * prod > 2147483647L) fail("train.in: bad atom count");
    out.d(na[n] * prod, '\n');
  }
  // (expand222.cpp:61-96: images of each atom next to each other)
  for (int n = 0; n < nc; ++n) {
    const bool force = n < nc_force;
    if (!force) {
      for (int k = 0; k < 7; ++k) out.g(in.f("energy and virial") * (float)prod, ' ');
      out.s += '\n';
    }
    float box[9];
    for (int k = 0; k < 9; ++k) {
      box[k] = in.f("the box");
      out.g(box[k] * (float)cells[k / 3], ' ');
    }
    out.s += '\n';
    for (int a = 0; a < na[n]; ++a) {
      const int type = in.i("a type");
      float v[6];
      for (int k = 0; k < (force ? 6 : 3); ++k) v[k] = in.f(force ? "a position or force" : "a position");
      for (int ia = 0; ia < cells[0]; ++ia)
        for (int ib = 0; ib < cells[1]; ++ib)
          for (int ic = 0; ic < cells[2]; ++ic) {
            float d[3];
            offset(box, ia, ib, ic, d);
            out.d(type, ' ');
            for (int k = 0; k < 3; ++k) out.g(v[k] + d[k], k < 2 || force ? ' ' : '\n');
            if (force)
              for (int k = 3; k < 6; ++k) out.g(v[k], k < 5 ? ' ' : '\n');
          }
    }
  }
  return std::move(out.s);
}

} // namespace gpuga
