/*
 * qsm_mex_util.h -- shared helpers of the MEX drop-in wrappers: they keep
 * the reference MEX argument layouts and forward to libqsm_hip.so's C ABI.
 */
#ifndef QSM_MEX_UTIL_H
#define QSM_MEX_UTIL_H

#include <stddef.h>
#include <stdint.h>

#include "mex.h"
#include "qsm_hip.h"

/* The reference reads nx, ny, nz from mxGetDimensions(prhs[k]) and assumes
 * a 3-D array; a 2-D array is treated as nz = 1 here. */
static inline void qsm_mex_dims(const mxArray *a, size_t sz[3])
{
    const mwSize *d = mxGetDimensions(a);
    const mwSize nd = mxGetNumberOfDimensions(a);
    sz[0] = d[0];
    sz[1] = nd > 1 ? d[1] : 1;
    sz[2] = nd > 2 ? d[2] : 1;
}

/* The reference reads vsz through a double pointer whatever its class; the
 * wrappers require a double 3-vector instead of reading garbage. */
static inline const double *qsm_mex_vsz(const mxArray *a)
{
    if (!mxIsDouble(a) || mxGetNumberOfElements(a) != 3)
        mexErrMsgTxt("voxel size must be a double vector with 3 elements");
    return (const double *)mxGetData(a);
}

/* Masks are read as bytes (MATLAB logical or uint8), as in the reference. */
static inline const uint8_t *qsm_mex_mask(const mxArray *a, const size_t sz[3])
{
    if (!(mxIsLogical(a) || mxIsUint8(a)))
        mexErrMsgTxt("mask must be logical or uint8");
    if (mxGetNumberOfElements(a) != sz[0] * sz[1] * sz[2])
        mexErrMsgTxt("mask size does not match the volume");
    return (const uint8_t *)mxGetData(a);
}

/* Single or double, decided by the first argument as in the reference. */
static inline int qsm_mex_dtype(const mxArray *a)
{
    if (mxIsSingle(a)) return QSM_F32;
    if (mxIsDouble(a)) return QSM_F64;
    mexErrMsgTxt("arrays must be single or double");
    return -1;
}

/* Every floating-point array argument must match the first's class and size. */
static inline void qsm_mex_same(const mxArray *a, const mxArray *ref, const size_t sz[3])
{
    if (mxGetClassID(a) != mxGetClassID(ref))
        mexErrMsgTxt("all arrays must have the same class");
    if (mxGetNumberOfElements(a) != sz[0] * sz[1] * sz[2])
        mexErrMsgTxt("all arrays must have the same size");
}

static inline void qsm_mex_print(const char *line, void *ctx)
{
    (void)ctx;
    mexPrintf("%s", line);
}

static inline void qsm_mex_check(int rc)
{
    if (rc != QSM_OK)
        mexErrMsgTxt(qsm_last_error());
}

#endif
