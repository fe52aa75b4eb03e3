/* xtc_ref_harness.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Thin C entry points over the reference's vendored xdrfile library
 * (coords_file/xdrfile/xdrfile.c, xdrfile_xtc.c; compiled from the reference
 * tree into oracle/_ref by the Makefile's `ref` target).  Tests use it as the
 * writer of XTC trajectories and as the reader whose output the product's own
 * XTC decoder must reproduce bit for bit.  Nothing in the product links it.
 */
#include <stdlib.h>
#include "xdrfile.h"
#include "xdrfile_xtc.h"

/* Writes n_frames frames of natoms atoms (xyz[f][a][3]) at the given
 * precision; box zero, step f, time f.  Returns 0 on success. */
int refx_write(const char* fn, int natoms, int n_frames, const float* xyz, float prec) {
  XDRFILE* xd = xdrfile_open(fn, "w");
  if (!xd) return -1;
  matrix box = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  rvec* x = (rvec*)malloc(sizeof(rvec) * (size_t)natoms);
  int rc = 0;
  for (int f = 0; f < n_frames && rc == 0; ++f) {
    for (int a = 0; a < natoms; ++a)
      for (int e = 0; e < 3; ++e) x[a][e] = xyz[((size_t)f * natoms + a) * 3 + e];
    if (write_xtc(xd, natoms, f, (float)f, box, x, prec) != exdrOK) rc = -2;
  }
  free(x);
  xdrfile_close(xd);
  return rc;
}

/* Reads up to max_frames frames into out[f][a][3]; *n_frames receives the
 * count.  Returns natoms, or a negative value on error. */
int refx_read(const char* fn, int max_frames, float* out, int* n_frames) {
  int natoms = 0;
  if (read_xtc_natoms((char*)fn, &natoms) != exdrOK) return -1;
  XDRFILE* xd = xdrfile_open(fn, "r");
  if (!xd) return -1;
  rvec* x = (rvec*)malloc(sizeof(rvec) * (size_t)natoms);
  matrix box;
  int step, f = 0;
  float time, prec;
  while (f < max_frames && read_xtc(xd, natoms, &step, &time, box, x, &prec) == exdrOK) {
    for (int a = 0; a < natoms; ++a)
      for (int e = 0; e < 3; ++e) out[((size_t)f * natoms + a) * 3 + e] = x[a][e];
    ++f;
  }
  free(x);
  xdrfile_close(xd);
  *n_frames = f;
  return natoms;
}
