// Readers of the files a SIROCCO cycle hands PLUTO
// (Src/LineDriven/line_connect.c:43-497): directional fluxes, force
// multiplier fits, heating / cooling state and prefactors.  Host code only.
//
// The record formats, the matching of a record to grid cells (relative
// tolerance on r and theta, every matching cell takes the record, one
// record's values are consumed per matching cell) and the stdio calls
// that parse them follow the reference, so that a file parses to the same
// values and the same cells here as there.  Where the reference exits on
// a malformed record these return an error; where it prints a note and
// goes on (missing file, prefactor below zero) they report it through the
// out-parameters.
#include <cmath>
#include <cstdio>
#include <string>

#include "plutomx.h"
#include "pmx_common.h"

namespace {

constexpr int LINE = 800;          // LINELENGTH, line_connect.c:26
constexpr double C_MP = 1.67262171e-24;
constexpr int NG = 2;              // ghost cells in the descriptor's grid arrays

std::string path_of(const char* dir, const char* name) {
  if (!dir || !*dir) return name;
  std::string p(dir);
  if (p.back() != '/') p += '/';
  return p + name;
}

struct File {
  FILE* f;
  explicit File(const std::string& p) : f(std::fopen(p.c_str(), "r")) {}
  ~File() { if (f) std::fclose(f); }
};

bool grid_ok(const pmx_wind_desc* g) {
  if (!g || g->nx1 < 1 || g->nx2 < 1 || !g->x1 || !g->x2 || !(g->unit_length > 0.0)) {
    pmx::set_error("grid descriptor without nx1, nx2, x1, x2 or unit_length");
    return false;
  }
  return true;
}

// Records matched by line_connect.c:126-128 and :222-224: r within a
// relative 1e-6 of x1 (in cm), theta within a relative 1e-6 of x2.
inline bool match_flux(const pmx_wind_desc* g, int j, int i, double x1in, double x2in) {
  const double tol = 1e-6;
  return std::fabs(1.0 - (x1in / g->unit_length / g->x1[i + NG])) < tol &&
         std::fabs(1.0 - (x2in / g->x2[j + NG])) < tol;
}

// Records matched by line_connect.c:347-348 and :443-444: relative to the
// record's own r and theta, tolerance 1e-5.
inline bool match_hc(const pmx_wind_desc* g, int j, int i, double rcen, double thetacen) {
  const double tol = 1e-5;
  return std::fabs(((rcen / g->unit_length) - g->x1[i + NG]) / (rcen / g->unit_length)) < tol &&
         std::fabs((thetacen - g->x2[j + NG]) / thetacen) < tol;
}

// The record loop shared by the flux and fit files (line_connect.c:122-161,
// :219-250): "i j [k] r theta v_0 .. v_{n-1}"; each matching cell reads
// the next n values into out[q * nint + cell].
int read_records(FILE* f, const pmx_wind_desc* g, int n, bool has_k, double* out, long* count,
                 const char* what) {
  const long long nint = (long long)g->nx1 * g->nx2;
  long ii, jj;
  double x1in, x2in, temp;
  long icount = 0;
  while (std::fscanf(f, "%ld ", &ii) != EOF) {
    const int got = has_k ? std::fscanf(f, "%ld %*ld %le %le", &jj, &x1in, &x2in)
                          : std::fscanf(f, "%ld %le %le", &jj, &x1in, &x2in);
    if (got != 3) { pmx::set_error(std::string("error in reading ") + what); return -1; }
    bool match = false;
    for (int j = 0; j < g->nx2; j++)
      for (int i = 0; i < g->nx1; i++) {
        if (!match_flux(g, j, i, x1in, x2in)) continue;
        for (int q = 0; q < n; q++) {
          if (std::fscanf(f, "%le", &temp) != 1) {
            pmx::set_error(std::string("error in reading ") + what);
            return -1;
          }
          out[(long long)q * nint + (long long)j * g->nx1 + i] = temp;
        }
        match = true;
        icount++;
      }
    if (!match)
      for (int q = 0; q < n; q++)
        if (std::fscanf(f, "%le", &temp) != 1) {
          pmx::set_error(std::string("error in reading ") + what);
          return -1;
        }
  }
  if (count) *count = icount;
  return 0;
}

const char* const FLUX_FILES[3] = {"directional_flux_r.dat", "directional_flux_theta.dat",
                                   "directional_flux_phi.dat"};

}  // namespace

extern "C" {

int pmx_sirocco_flux_angles(const char* dir, int* nangles) {
  if (!nangles) { pmx::set_error("null argument"); return -1; }
  *nangles = 0;
  File fl(path_of(dir, FLUX_FILES[0]));
  if (!fl.f) return 0;
  char aline[LINE];
  long ii;
  if (!std::fgets(aline, LINE, fl.f) || !std::fgets(aline, LINE, fl.f)) {
    pmx::set_error("directional_flux_r.dat: header lines missing");
    return -1;
  }
  if (std::sscanf(aline, "%*s %*s %ld", &ii) != 1 || ii < 1) {
    pmx::set_error("directional_flux_r.dat: flux header improperly formatted");
    return -1;
  }
  *nangles = (int)ii;
  return 0;
}

int pmx_sirocco_read_fluxes(const char* dir, const pmx_wind_desc* grid, int nangles,
                            double* flux_r, double* flux_t, double* flux_p, int* found,
                            long* matched) {
  if (!grid_ok(grid)) return -1;
  if (!flux_r || !flux_t || !flux_p || nangles < 1) {
    pmx::set_error("flux arrays missing or nangles < 1");
    return -1;
  }
  double* out[3] = {flux_r, flux_t, flux_p};
  int fnd = 0;
  for (int ax = 0; ax < 3; ax++) {
    if (matched) matched[ax] = 0;
    File fl(path_of(dir, FLUX_FILES[ax]));
    if (!fl.f) continue;                        // "No flux file" (line_connect.c:77-80)
    char aline[LINE];
    long ii;
    if (!std::fgets(aline, LINE, fl.f) || !std::fgets(aline, LINE, fl.f)) {
      pmx::set_error(std::string(FLUX_FILES[ax]) + ": header lines missing");
      return -1;
    }
    if (std::sscanf(aline, "%*s %*s %ld", &ii) != 1) {
      pmx::set_error(std::string(FLUX_FILES[ax]) + ": flux header improperly formatted");
      return -1;
    }
    if (ii != nangles) {                        // line_connect.c:100-114
      pmx::set_error(std::string(FLUX_FILES[ax]) + " doesn't agree in NFLUX_ANGLES");
      return -1;
    }
    if (read_records(fl.f, grid, nangles, true, out[ax], matched ? &matched[ax] : nullptr,
                     FLUX_FILES[ax]))
      return -1;
    fnd |= 1 << ax;
  }
  if (found) *found = fnd;
  return 0;
}

int pmx_sirocco_fm_points(const char* dir, int* points) {
  if (!points) { pmx::set_error("null argument"); return -1; }
  *points = 0;
  File fl(path_of(dir, "M_UV_data.dat"));
  if (!fl.f) return 0;
  char aline[LINE];
  long ii;
  if (!std::fgets(aline, LINE, fl.f)) { pmx::set_error("M_UV_data.dat: no header"); return -1; }
  if (std::sscanf(aline, "%*s %ld", &ii) != 1 || ii < 1) {
    pmx::set_error("M_UV_data.dat: header without the number of points");
    return -1;
  }
  *points = (int)ii;
  return 0;
}

int pmx_sirocco_read_force_multiplier(const char* dir, const pmx_wind_desc* grid, int points,
                                      double* t, double* m, int* found, long* matched) {
  if (!grid_ok(grid)) return -1;
  if (!t || !m || points < 1) { pmx::set_error("fit arrays missing or points < 1"); return -1; }
  if (found) *found = 0;
  if (matched) *matched = 0;
  File fl(path_of(dir, "M_UV_data.dat"));
  if (!fl.f) return 0;                          // "No force multiplier file"
  char aline[LINE];
  long ii;
  if (!std::fgets(aline, LINE, fl.f)) { pmx::set_error("M_UV_data.dat: no header"); return -1; }
  if (std::sscanf(aline, "%*s %ld", &ii) != 1 || ii != points) {
    pmx::set_error("M_UV_data.dat: number of points differs from the header");
    return -1;
  }
  if (std::fscanf(fl.f, "%*s ") != 0) {         // line_connect.c:202-207
    pmx::set_error("M_UV_data.dat: error in reading the fit geometry");
    return -1;
  }
  double temp;
  for (int q = 0; q < points; q++) {
    if (std::fscanf(fl.f, "%le", &temp) != 1) {
      pmx::set_error("M_UV_data.dat: error in reading the fit points");
      return -1;
    }
    t[q] = temp;
  }
  if (read_records(fl.f, grid, points, false, m, matched, "M_UV_data.dat")) return -1;
  if (found) *found = 1;
  return 0;
}

int pmx_sirocco_read_heatcool(const char* dir, const pmx_wind_desc* grid, const double* rho,
                              double* xi, double* t_r, int* found, long* matched) {
  if (!grid_ok(grid)) return -1;
  if (!xi || !t_r) { pmx::set_error("xi / t_r arrays missing"); return -1; }
  if (found) *found = 0;
  if (matched) *matched = 0;
  const int nr = grid->nx1, nt = grid->nx2;
  File fl(path_of(dir, "py_heatcool.dat"));
  if (!fl.f) {                                  // line_connect.c:313-326
    if (!rho) { pmx::set_error("py_heatcool.dat missing and no density given"); return -1; }
    for (int j = 0; j < nt; j++)
      for (int i = 0; i < nr; i++) {
        const long long c = (long long)j * nr + i;
        const double r = grid->x1[i + NG] * grid->unit_length;
        const double nH = rho[c] * grid->unit_density / (1.43 * C_MP);
        const double lx = grid->l_star * grid->f_xray;
        xi[c] = lx / nH / (r * r);
        t_r[c] = grid->t_x;
      }
    return 0;
  }
  char aline[LINE];
  if (!std::fgets(aline, LINE, fl.f)) {
    pmx::set_error("error reading py_heatcool.dat header");
    return -1;
  }
  long icount = 0;
  while (std::fgets(aline, LINE, fl.f)) {
    int ii, jj;
    double rcen, thetacen, vol, t_e, t_rin, xiin, ne, h1, h2, h3, h4, c1, c2, c3, dens, n_h;
    const int nwords = std::sscanf(
        aline, "%d %d %le %le %le %le %le %le %le %le %le %le %le %le %le %le %le %le", &ii, &jj,
        &rcen, &thetacen, &vol, &t_e, &t_rin, &xiin, &ne, &h1, &h2, &h3, &h4, &c1, &c2, &c3,
        &dens, &n_h);
    if (nwords != 18) {
      pmx::set_error("py_heatcool file incorrectly formatted: " + std::string(aline));
      return -1;
    }
    for (int j = 0; j < nt; j++)
      for (int i = 0; i < nr; i++) {
        if (!match_hc(grid, j, i, rcen, thetacen)) continue;
        const long long c = (long long)j * nr + i;
        icount++;
        xi[c] = xiin;                           // line_connect.c:350-356
        t_r[c] = t_rin;
        if (xi[c] < 1.0) xi[c] = 1.0;
        if (t_r[c] < 1.e3) t_r[c] = 1.e3;
      }
  }
  if (found) *found = 1;
  if (matched) *matched = icount;
  return 0;
}

int pmx_sirocco_read_prefactors(const char* dir, const pmx_wind_desc* grid, double* pre,
                                int* found, long* matched, long* negative) {
  if (!grid_ok(grid)) return -1;
  if (!pre) { pmx::set_error("prefactor array missing"); return -1; }
  if (found) *found = 0;
  if (matched) *matched = 0;
  if (negative) *negative = 0;
  const int nr = grid->nx1, nt = grid->nx2;
  const long long nint = (long long)nr * nt;
  for (long long q = 0; q < 6 * nint; q++) pre[q] = 1.0;   // line_connect.c:385-414
  File fl(path_of(dir, "prefactors.dat"));
  if (!fl.f) return 0;
  char aline[LINE];
  if (!std::fgets(aline, LINE, fl.f)) {
    pmx::set_error("error reading prefactors.dat header");
    return -1;
  }
  long icount = 0;
  while (std::fgets(aline, LINE, fl.f)) {
    int ii, jj;
    double rcen, thetacen, dens, v[6];
    const int nwords = std::sscanf(aline, "%d %le %d %le %le %le %le %le %le %le %le", &ii,
                                   &rcen, &jj, &thetacen, &dens, &v[0], &v[1], &v[2], &v[3],
                                   &v[4], &v[5]);
    if (nwords != 11) {
      pmx::set_error("prefactor file incorrectly formatted: " + std::string(aline));
      return -1;
    }
    for (int j = 0; j < nt; j++)
      for (int i = 0; i < nr; i++) {
        if (!match_hc(grid, j, i, rcen, thetacen)) continue;
        const long long c = (long long)j * nr + i;
        icount++;
        for (int q = 0; q < 6; q++) pre[q * nint + c] = v[q];
      }
  }
  long neg = 0;                                 // line_connect.c:470-491
  for (long long q = 0; q < 6 * nint; q++) neg += pre[q] < 0.0;
  if (found) *found = 1;
  if (matched) *matched = icount;
  if (negative) *negative = neg;
  return 0;
}

}  // extern "C"
