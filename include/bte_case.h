/* bte_case.h -- host convenience entry points of the C++ facade: load a case from the reference's
 * input files (same arguments as BTEcmd, src/BTEMain.cpp:56-151) and run it through
 * StaticBTESolver.  Used by language bindings (ctypes) that cannot call C++.  No counterpart in the
 * reference, which only offers the BTEcmd executable and the C++ classes. */
#ifndef STATICBTE_AMD_BTE_CASE_H
#define STATICBTE_AMD_BTE_CASE_H

#include "bte.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bte_case bte_case;

/* geometry: an .mphtxt path, or a decimal integer N for the uniform 1D mesh (BTEMain.cpp:118-122).
 * Returns BTE_OK or a negative status with the message in err. */
int bte_case_open(const char* geometry, double Lx, double Ly, double Lz, const char* band_file,
                  const char* bc_file, int DM, int ntheta, int nphi, double T_ref, int device,
                  bte_case** out, char* err, int errlen);
void bte_case_close(bte_case* c);
/* sizes: N_cell, N_dir, N_band, N_face, n_bc, dim */
void bte_case_sizes(const bte_case* c, int* out6);
/* T_ref (after the 1D override) and solid angle */
void bte_case_params(const bte_case* c, double* out2);
int bte_case_solve(bte_case* c, const bte_opts* opts, int verbose, bte_result* res);
/* Preprocessed geometry and quadrature (the C ABI inputs). Any pointer may be NULL. */
int bte_case_geometry(bte_case* c, double* volume, double* centre, double* area, double* normal, int* nbr);
int bte_case_quadrature(bte_case* c, double* omega, double* dir, double* S);
/* The C ABI context of the case (created and loaded on first use). */
bte_ctx* bte_case_ctx(bte_case* c);
const char* bte_case_error(const bte_case* c);

#ifdef __cplusplus
}
#endif
#endif
