// ref_shim.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Thin extern "C" entry points over the reference's own pure-C++ helper
// sources, compiled where they lie under the reference tree by
// oracle/Makefile into oracle/_ref/libref.so.  Used only by tests to pin the
// oracle and the host library (coefficient design, model padding, SEG-Y bytes,
// depth/time conversion and phase rotation).  Nothing here
// restates or replaces reference code; each wrapper forwards to it.
#include <malloc.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "segy.h"  // from the reference tree (-I)

// Prototypes of the reference functions (definitions live in the reference
// sources linked into this library).
int funMandC(int nthita, int nfdmax, int nfdmin, int nvel, double tao, double h, double df,
             double eps, double fmax, double vmin, double vmax, double dv, int* fdcoeneed,
             int* M, int* Index, int iLSTE, float** cc, double hzx);  // LSMOrCon_rec_2D.cpp:22
void callenfd2d_ls(double* hk, double* disperr, int nfre, int nfdmin, int nfdmax, double vel,
                   double tao, double h, double df, int nthita, double eps, int* lenfd,
                   int iLSTE, double hzx);                          // LSMOrCon_rec_2D.cpp:293
void CAL2DFDCOE_LSM(double* c, double r, double bb, int M, double hzx);  // :236
void order(int N, float* c);                                        // :526
int D2T(const char* File_Name, float* V, float* D, int Nx, int Nz, int Nx_d_st, int Nx_d_end,
        float dx, float dz, float dt);                              // DisToTimeAndTimeToDis1D.cpp:115
int T2D(const char* File_Name, float* V, float* D, int Nx, int Nz, int Nz_V, int Nx_d_st,
        int Nx_d_end, float dx, float dz, float dt);                // DisToTimeAndTimeToDis1D.cpp:31
void phase_correction(float* din, float* dout, int ntr, int nt, float angle);  // phase_correction_ricker_decon.cpp:153
void velocity(char* OutNameVp, float* v, float* v_2, float* r, float* r_1, float* r_2, int NZ,
              int NX, int NXZ, int N2, float tao, float h, int choice);  // GPU_velocity_real.cpp:6

// SGYWrite.cpp relies on <stdio.h> and segy.h being included before it, as
// kernel.cu:1-13 does; it is compiled here in that same context.
#include "SGYWrite.cpp"

extern "C" {

int ref_funMandC(int nthita, int nfdmax, int nfdmin, int nvel, double tao, double h, double df,
                 double eps, double fmax, double vmin, double vmax, double dv, int* need, int* M,
                 int* Index, double hzx, float* c_out, int c_cap) {
  float* cc = NULL;
  int nc = funMandC(nthita, nfdmax, nfdmin, nvel, tao, h, df, eps, fmax, vmin, vmax, dv, need, M,
                    Index, 0, &cc, hzx);
  if (nc <= c_cap) memcpy(c_out, cc, sizeof(float) * nc);
  free(cc);
  return nc <= c_cap ? nc : -1;
}

int ref_lenfd(int nfre, int nfdmin, int nfdmax, double vel, double tao, double h, double df,
              int nthita, double eps, double hzx) {
  double* hk = (double*)malloc(sizeof(double) * nfre);
  double* de = (double*)malloc(sizeof(double) * nfre);
  int len = 0;
  callenfd2d_ls(hk, de, nfre, nfdmin, nfdmax, vel, tao, h, df, nthita, eps, &len, 0, hzx);
  free(hk);
  free(de);
  return len;
}

void ref_lscoef(double r, double bb, int M, double hzx, double* c) { CAL2DFDCOE_LSM(c, r, bb, M, hzx); }

void ref_taylor(int nfdmax, float* c) { order(2 * nfdmax, c); }

void ref_velocity(const char* path, int NZ, int NX, int N2, float tao, float h, int choice,
                  float* v, float* r1) {
  size_t nxz = (size_t)NZ * NX;
  float* v2 = (float*)calloc(nxz, sizeof(float));
  float* r = (float*)calloc(nxz, sizeof(float));
  float* r2 = (float*)calloc(nxz, sizeof(float));
  char name[4096];
  strncpy(name, path, sizeof(name) - 1);
  name[sizeof(name) - 1] = 0;
  velocity(name, v, v2, r, r1, r2, NZ, NX, (int)nxz, N2, tao, h, choice);
  free(v2);
  free(r);
  free(r2);
}

// WriteSGY reads "SGY_Model.sgy" from the working directory (SGYWrite.cpp:132).
void ref_write_sgy(float* data, int ntr, int ns, int dz, float* sx, float* sy, float rx, float ry,
                   float* dsr, const char* path) {
  char name[4096];
  strncpy(name, path, sizeof(name) - 1);
  name[sizeof(name) - 1] = 0;
  WriteSGY(data, ntr, ns, dz, sx, sy, rx, ry, dsr, name);
}

void ref_trace2segy(char* buf, const float* trace, int ns, int format) {
  trace2segy(buf, trace, ns, format);
}
void ref_segy2trace(const char* buf, float* trace, int ns, int format) {
  segy2trace(buf, trace, ns, format);
}

// Depth/time conversion writes its result to a file and returns its length.
// D2T and T2D read t0[i*Nz] without writing it (DisToTimeAndTimeToDis1D.cpp:
// 43-49, 129-135).  At the reference's image sizes that buffer is a fresh
// mmap-backed allocation, i.e. zero; glibc's M_PERTURB (allocations filled
// with the complement of 0xff) gives the small test inputs the same zeros.
int ref_d2t(const char* path, float* V, float* D, int Nx, int Nz, int st, int end, float dx,
            float dz, float dt) {
  mallopt(M_PERTURB, 0xff);
  int n = D2T(path, V, D, Nx, Nz, st, end, dx, dz, dt);
  mallopt(M_PERTURB, 0);
  return n;
}
int ref_t2d(const char* path, float* V, float* D, int Nx, int Nz, int Nz_V, int st, int end,
            float dx, float dz, float dt) {
  mallopt(M_PERTURB, 0xff);
  int n = T2D(path, V, D, Nx, Nz, Nz_V, st, end, dx, dz, dt);
  mallopt(M_PERTURB, 0);
  return n;
}
void ref_phase_correction(float* din, float* dout, int ntr, int nt, float angle) {
  phase_correction(din, dout, ntr, nt, angle);
}
}  // extern "C"
