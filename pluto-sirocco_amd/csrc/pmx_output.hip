// Output writers: the .dbl data files with their dbl.out index and the
// restart.out records, in the reference's layouts (Src/write_data.c,
// Src/bin_io.c, Src/restart.c), so that the reference's restart path and
// the SIROCCO coupling scripts read them.  Host code only.
#include <cstdio>
#include <cstring>
#include <string>

#include "plutomx.h"
#include "pmx_common.h"

namespace {

std::string path_in(const char* dir, const std::string& name) {
  if (!dir || !*dir) return name;
  std::string p(dir);
  if (p.back() != '/') p += '/';
  return p + name;
}

bool little_endian() {
  const int one = 1;
  return *(const char*)&one == 1;
}

// Src/structs.h:256-262
struct RestartRecord {
  int nstep;
  int nfile[16];
  double t;
  double dt;
  char fill[40];
};
static_assert(sizeof(RestartRecord) == 128, "Restart record is 128 bytes");

}  // namespace

extern "C" {

int pmx_write_dbl(const char* dir, int nfile, const double* data, int nvar,
                  const char* const* names, int nx1, int nx2, int nx3, double t, double dt,
                  long step) {
  if (!data || !names || nvar < 1 || nx1 < 1 || nx2 < 1 || nx3 < 1 || nfile < 0) {
    pmx::set_error("pmx_write_dbl: bad argument");
    return -1;
  }
  char fname[64];
  std::snprintf(fname, sizeof(fname), "data.%04d.dbl", nfile);
  const std::string dfile = path_in(dir, fname);
  std::remove(dfile.c_str());                 // FileDelete, write_data.c:131
  FILE* f = std::fopen(dfile.c_str(), "wb");
  if (!f) { pmx::set_error("cannot open " + dfile); return -1; }
  const size_t n = (size_t)nx1 * nx2 * nx3 * nvar;
  const size_t w = std::fwrite(data, sizeof(double), n, f);
  std::fclose(f);
  if (w != n) { pmx::set_error("short write to " + dfile); return -1; }

  // dbl.out, write_data.c:420-450
  const std::string lfile = path_in(dir, "dbl.out");
  FILE* fo;
  if (nfile == 0) {
    fo = std::fopen(lfile.c_str(), "w");
  } else {
    fo = std::fopen(lfile.c_str(), "r+");
    if (fo) {
      char sline[512];
      for (int q = 0; q < nfile; q++)
        if (!std::fgets(sline, 512, fo)) break;
      std::fseek(fo, std::ftell(fo), SEEK_SET);
    }
  }
  if (!fo) { pmx::set_error("cannot open " + lfile); return -1; }
  std::fprintf(fo, "%d %12.6e %12.6e %ld ", nfile, t, dt, step);
  std::fprintf(fo, "single_file ");
  std::fprintf(fo, little_endian() ? "little " : "big ");
  for (int q = 0; q < nvar; q++) std::fprintf(fo, "%s ", names[q]);
  std::fprintf(fo, "\n");
  std::fclose(fo);
  return 0;
}

int pmx_write_restart(const char* dir, int counter, long nstep, const int* nfile, double t,
                      double dt) {
  if (!nfile || counter < 0) { pmx::set_error("pmx_write_restart: bad argument"); return -1; }
  RestartRecord r;
  std::memset(&r, 0, sizeof(r));
  r.nstep = (int)nstep;
  for (int q = 0; q < 16; q++) r.nfile[q] = nfile[q];
  r.t = t;
  r.dt = dt;
  const std::string fname = path_in(dir, "restart.out");
  FILE* fr;
  if (counter == 0) {
    fr = std::fopen(fname.c_str(), "wb");
  } else {
    fr = std::fopen(fname.c_str(), "r+b");
    if (fr) std::fseek(fr, (long)counter * (long)sizeof(RestartRecord), SEEK_SET);
  }
  if (!fr) { pmx::set_error("cannot open " + fname); return -1; }
  const size_t w = std::fwrite(&r, sizeof(r), 1, fr);
  std::fclose(fr);
  if (w != 1) { pmx::set_error("short write to " + fname); return -1; }
  return 0;
}

int pmx_read_restart(const char* dir, int nrestart, long* nstep, int* nfile, double* t,
                     double* dt, int* swap) {
  if (nrestart < 0) {                           // restart.c:206-209
    pmx::set_error("negative restart numbers are not supported (as in the reference)");
    return -1;
  }
  // byte order from the dbl.out line of output nrestart, restart.c:60-83
  const std::string lfile = path_in(dir, "dbl.out");
  FILE* fb = std::fopen(lfile.c_str(), "r");
  if (!fb) { pmx::set_error("cannot find " + lfile); return -1; }
  char str[512];
  int nlines = 0;
  while (std::fgets(str, 512, fb) != nullptr) nlines++;
  std::rewind(fb);
  if (nrestart > nlines - 1) {
    std::fclose(fb);
    pmx::set_error("output #" + std::to_string(nrestart) + " does not exist in " + lfile);
    return -1;
  }
  for (int q = nrestart; q--;) {
    int c;
    while ((c = std::fgetc(fb)) != '\n' && c != EOF) {}
  }
  int inv, ist;
  double dd;
  char mode[512], order[512];
  const int got = std::fscanf(fb, "%d  %lf  %lf  %d  %511s  %511s\n", &inv, &dd, &dd, &ist,
                              mode, order);
  std::fclose(fb);
  if (got != 6) { pmx::set_error("malformed line in " + lfile); return -1; }
  const int sw = (!std::strcmp(order, "big") && little_endian()) ||
                 (!std::strcmp(order, "little") && !little_endian());
  // the first record whose .dbl counter is nrestart, restart.c:220-244
  const std::string rfile = path_in(dir, "restart.out");
  FILE* fr = std::fopen(rfile.c_str(), "rb");
  if (!fr) { pmx::set_error("cannot find " + rfile); return -1; }
  RestartRecord r;
  bool found = false;
  while (std::fread(&r, sizeof(r), 1, fr) == 1) {
    int nf0 = r.nfile[0];
    if (sw) nf0 = (int)__builtin_bswap32((unsigned)nf0);
    if (nf0 == nrestart) { found = true; break; }
  }
  std::fclose(fr);
  if (!found) { pmx::set_error("no record of output #" + std::to_string(nrestart)); return -1; }
  if (sw) {
    r.nstep = (int)__builtin_bswap32((unsigned)r.nstep);
    for (int q = 0; q < 16; q++) r.nfile[q] = (int)__builtin_bswap32((unsigned)r.nfile[q]);
    unsigned long long u;
    std::memcpy(&u, &r.t, 8); u = __builtin_bswap64(u); std::memcpy(&r.t, &u, 8);
    std::memcpy(&u, &r.dt, 8); u = __builtin_bswap64(u); std::memcpy(&r.dt, &u, 8);
  }
  if (nstep) *nstep = r.nstep;
  if (nfile) for (int q = 0; q < 16; q++) nfile[q] = r.nfile[q];
  if (t) *t = r.t;
  if (dt) *dt = r.dt;
  if (swap) *swap = sw;
  return 0;
}

int pmx_read_dbl(const char* dir, int nfile, int nvar, int nx1, int nx2, int nx3, int swap,
                 double* data) {
  if (!data || nvar < 1 || nx1 < 1 || nx2 < 1 || nx3 < 1 || nfile < 0) {
    pmx::set_error("pmx_read_dbl: bad argument");
    return -1;
  }
  char fname[64];
  std::snprintf(fname, sizeof(fname), "data.%04d.dbl", nfile);
  const std::string dfile = path_in(dir, fname);
  FILE* f = std::fopen(dfile.c_str(), "rb");
  if (!f) { pmx::set_error("cannot open " + dfile); return -1; }
  const size_t n = (size_t)nx1 * nx2 * nx3 * nvar;
  const size_t got = std::fread(data, sizeof(double), n, f);
  std::fclose(f);
  if (got != n) { pmx::set_error("short read from " + dfile); return -1; }
  if (swap)
    for (size_t q = 0; q < n; q++) {
      unsigned long long u;
      std::memcpy(&u, &data[q], 8);
      u = __builtin_bswap64(u);
      std::memcpy(&data[q], &u, 8);
    }
  return 0;
}

}  // extern "C"
