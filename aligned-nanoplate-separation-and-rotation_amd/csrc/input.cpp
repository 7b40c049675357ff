// bcp.input reader of libscft_amd.so, restating read_input
// (code/read_input.cpp:7-242): a positional file where each value line is
// followed by free text and each section by a blank line and a header.
// Lines holding per-dimension values (grid, box, plate centre and
// orientation) carry `dim` numbers; further numbers on the line are ignored,
// as the reference's fgets() does.  Negative chi values stay negative here;
// the sign is the reference's negative-chi flag.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/scft.h"

namespace {

struct Lines {
  std::vector<std::string> v;
  size_t at = 0;
  bool ok = true;
  std::istringstream cur;
  void skip(int n = 1) { at += n; }
  // the next value line, as a token stream
  std::istringstream& next() {
    if (at >= v.size()) ok = false;
    cur.clear();
    cur.str(at < v.size() ? v[at] : std::string());
    at++;
    return cur;
  }
  template <typename T>
  void get(std::istringstream& s, T& x) {
    if (!(s >> x)) ok = false;
  }
};

}  // namespace

extern "C" int scft_read_input(const char* path, int dim, scft_params* o) {
  if (!path || !o || (dim != 2 && dim != 3)) return SCFT_EINVAL;
  std::ifstream f(path);
  if (!f) return SCFT_EIO;
  Lines L;
  for (std::string line; std::getline(f, line);) L.v.push_back(line);
  std::memset(o, 0, sizeof *o);
  o->dim = dim;
  L.skip();  // "# Main parameters"
  { auto& s = L.next(); L.get(s, o->Nda); L.get(s, o->Ndb); }
  { auto& s = L.next(); L.get(s, o->phiH); L.get(s, o->Nah); }
  { auto& s = L.next(); L.get(s, o->a_smear); }
  { auto& s = L.next(); L.get(s, o->rho0); }
  L.skip(2);
  { auto& s = L.next(); L.get(s, o->chi_ab); }
  { auto& s = L.next(); L.get(s, o->chi_bc); }
  { auto& s = L.next(); L.get(s, o->chi_ac); }
  { auto& s = L.next(); L.get(s, o->kappa); }
  L.skip(2);
  for (int j = 0; j < 2; j++) {
    auto& s = L.next();
    L.get(s, o->ic_flag[j]);
    L.get(s, o->ic_pre[j]);
    L.get(s, o->ic_dir[j]);
    L.get(s, o->ic_period[j]);
  }
  L.skip(2);
  { auto& s = L.next(); for (int i = 0; i < dim; i++) L.get(s, o->Nx[i]); }
  { auto& s = L.next(); for (int i = 0; i < dim; i++) L.get(s, o->L[i]); }
  { auto& s = L.next(); L.get(s, o->lam_imag); }
  { auto& s = L.next(); L.get(s, o->lam_real); }
  { auto& s = L.next(); L.get(s, o->itermax); }
  { auto& s = L.next(); L.get(s, o->print_freq); }
  { auto& s = L.next(); L.get(s, o->sample_freq); L.get(s, o->sample_wait); }
  { auto& s = L.next(); L.get(s, o->error_tol); }
  { auto& s = L.next(); L.get(s, o->update_scheme); }
  { auto& s = L.next(); L.get(s, o->do_CL); }
  L.skip(2);
  { auto& s = L.next(); L.get(s, o->do_film); }
  { auto& s = L.next(); L.get(s, o->wallT); L.get(s, o->wallXi); }
  { auto& s = L.next(); L.get(s, o->top_wall_lamA); L.get(s, o->top_wall_lamB); }
  { auto& s = L.next(); L.get(s, o->bot_wall_lamA); L.get(s, o->bot_wall_lamB); }
  L.skip(2);
  { auto& s = L.next(); L.get(s, o->do_channel); L.get(s, o->hor_dir); L.get(s, o->vert_dir); }
  { auto& s = L.next(); L.get(s, o->wall_width); L.get(s, o->channel_width); L.get(s, o->channel_xi); }
  L.skip(2);
  { auto& s = L.next(); L.get(s, o->n_exp_nr); }
  { auto& s = L.next(); L.get(s, o->do_fld_np); }
  { auto& s = L.next(); L.get(s, o->np_type); }
  { auto& s = L.next(); L.get(s, o->np_chem); }
  { auto& s = L.next(); L.get(s, o->sigma); }
  { auto& s = L.next(); L.get(s, o->Ng); }
  if (o->Ng == 0) o->sigma = 0.0;
  { auto& s = L.next(); L.get(s, o->Nu); }
  { auto& s = L.next(); L.get(s, o->np_frac); }
  if (!o->do_fld_np) o->np_frac = 0.0;
  { auto& s = L.next(); L.get(s, o->L_nr); L.get(s, o->R_nr); L.get(s, o->xi_nr); }
  for (int j = 0; j < 2; j++) {
    { auto& s = L.next(); for (int i = 0; i < dim; i++) L.get(s, o->exp_nr_c[j][i]); }
    auto& s = L.next();
    double m2 = 0.0;
    for (int i = 0; i < dim; i++) {
      L.get(s, o->exp_nr_u[j][i]);
      m2 += o->exp_nr_u[j][i] * o->exp_nr_u[j][i];
    }
    const double m = std::sqrt(m2);
    if (m > 0.0)
      for (int i = 0; i < dim; i++) o->exp_nr_u[j][i] /= m;
    if (o->n_exp_nr > 0) {
      for (int i = 0; i < dim; i++)
        if (o->exp_nr_c[j][i] < 0.0 || o->exp_nr_c[j][i] > 1.0) return SCFT_EINVAL;
      if (m == 0.0) return SCFT_EINVAL;
    }
  }
  return L.ok ? SCFT_OK : SCFT_EINVAL;
}
