/*
 * Solver configuration keywords:
 *   setSolver        reference src/Solv.c:507-528  (cg|bicgstab|bicgstabl1|
 *                                                   gmres|minres|pardiso)
 *   setDataStruct    reference src/Datastruct.c:1707-1752 (csr|card|ellpack|
 *                    csrdcoo|csrc|csrccoo; "card" is the reference's CSRD word)
 *   setSolverConfig  reference src/Solv.c:547-608 (name|tol|maxit|storage
 *                    ... end, words lower-cased, "//" comments a line out)
 *   smachn           reference src/Solv.c:614-622
 */
#include <ctype.h>
#include <stdio.h>
#include <string.h>
#include "sfr3d_setup.h"

double sfr_smachn(void)
{
  volatile double mq = 1.0;
  while (mq + 1.0 != 1.0) mq /= 2.0;
  return mq;
}

short sfr_solver_id(const char *w)
{
  if (!w) return -1;
  if (!strcmp(w, "cg")) return 1;
  if (!strcmp(w, "bicgstab")) return 2;
  if (!strcmp(w, "bicgstabl1")) return 3;
  if (!strcmp(w, "gmres")) return 4;
  if (!strcmp(w, "minres")) return 5;
  if (!strcmp(w, "pardiso")) return 6;
  return -1;
}

short sfr_storage_id(const char *w)
{
  if (!w) return -1;
  if (!strcmp(w, "csr")) return SFR_CSR;
  if (!strcmp(w, "card")) return SFR_CSRD;
  if (!strcmp(w, "ellpack")) return SFR_ELLPACK;
  if (!strcmp(w, "csrdcoo")) return SFR_CSRDCOO;
  if (!strcmp(w, "csrc")) return SFR_CSRC;
  if (!strcmp(w, "csrccoo")) return SFR_CSRCCOO;
  return -1;          /* the reference leaves storage unset on other words */
}

/* next whitespace-delimited word, lower-cased; a word starting with "//"
 * discards the rest of its line and comes back empty */
static int next_word(FILE *f, char *w, size_t cap)
{
  if (fscanf(f, "%255s", w) != 1) return 0;
  (void)cap;
  if (w[0] == '/' && w[1] == '/') {
    int c;
    while ((c = getc(f)) != '\n' && c != EOF) {}
    w[0] = '\0';
    return 1;
  }
  for (char *p = w; *p; p++) *p = (char)tolower((unsigned char)*p);
  return 1;
}

int sfr_read_solver_config(const char *path, short *solver, double *tol,
                           unsigned *maxit, short *storage)
{
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  char w[256];
  int rc = 0;
  if (!next_word(f, w, sizeof w)) { fclose(f); return -2; }
  while (strcmp(w, "end")) {
    if (!strcmp(w, "name")) {
      if (!next_word(f, w, sizeof w)) { rc = -2; break; }
      const short s = sfr_solver_id(w);
      if (s < 0) { rc = -3; break; }
      if (solver) *solver = s;
    } else if (!strcmp(w, "tol")) {
      double t;
      if (fscanf(f, "%lf", &t) != 1) { rc = -2; break; }
      if (t == 0.0) t = sfr_smachn();
      if (tol) *tol = t;
    } else if (!strcmp(w, "maxit")) {
      unsigned m;
      if (fscanf(f, "%u", &m) != 1) { rc = -2; break; }
      if (maxit) *maxit = m;
    } else if (!strcmp(w, "storage")) {
      if (!next_word(f, w, sizeof w)) { rc = -2; break; }
      const short s = sfr_storage_id(w);
      if (s > 0 && storage) *storage = s;
    }
    if (!next_word(f, w, sizeof w)) { rc = -2; break; }
  }
  fclose(f);
  return rc;
}
