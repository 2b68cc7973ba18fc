/*
 * mex.h -- a minimal stand-in of the MATLAB MEX / matrix API (R2018a interleaved complex), enough to
 * compile and drive fd2d_noise_amd/mex/fd2d_mex.c on a machine without MATLAB. TEST INFRASTRUCTURE:
 * it is not MathWorks' header; the names and signatures are the documented public ones, the
 * behaviour is what the gateway relies on (column-major data, interleaved complex, mexErrMsg*
 * leaves the gateway, mexCallMATLAB returns the configuration the harness sets).
 */
#ifndef FD2D_TEST_MEX_H
#define FD2D_TEST_MEX_H

#include <stdbool.h>
#include <stddef.h>

typedef size_t mwSize;
typedef enum { mxUNKNOWN_CLASS = 0, mxDOUBLE_CLASS = 6, mxSINGLE_CLASS = 7 } mxClassID;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;
typedef struct mxArray_tag mxArray;

mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID cls, mxComplexity cx);
mxArray* mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID cls, mxComplexity cx);
mxArray* mxCreateDoubleScalar(double v);
void mxDestroyArray(mxArray* a);
void* mxGetData(const mxArray* a);
double* mxGetPr(const mxArray* a);
const mwSize* mxGetDimensions(const mxArray* a);
mwSize mxGetNumberOfDimensions(const mxArray* a);
size_t mxGetM(const mxArray* a);
size_t mxGetN(const mxArray* a);
size_t mxGetNumberOfElements(const mxArray* a);
bool mxIsEmpty(const mxArray* a);
bool mxIsDouble(const mxArray* a);
bool mxIsSingle(const mxArray* a);
bool mxIsComplex(const mxArray* a);
double mxGetScalar(const mxArray* a);
int mexCallMATLAB(int nlhs, mxArray* plhs[], int nrhs, mxArray* prhs[], const char* name);
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...) __attribute__((noreturn));
void mexWarnMsgIdAndTxt(const char* id, const char* fmt, ...);

/* ---- harness side (not part of the MEX API) ---- */
typedef void (*mex_gateway)(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);
/* values returned by input_parameters / input_interferometry / absorb_specs */
void stub_set_config(double Lx, double Lz, int nx, int nz, double dt, int nt, int order, int n_basis_fct,
                     int fw_nth, const double* f_sample, int n_sample, int freq_samp, double absorb_width,
                     const int absorb[4]);
/* calls the gateway as MATLAB would; 0 = returned, 1 = raised (stub_error_id / stub_error_msg) */
int stub_call(mex_gateway fn, int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);
const char* stub_error_id(void);
const char* stub_error_msg(void);
const char* stub_warning_id(void);
long stub_live_arrays(void); /* arrays created and not destroyed */
mxClassID stub_class(const mxArray* a);

#endif
