/*
 * mex.h -- minimal stand-in for MATLAB's MEX API, used only to build and
 * test this project's own MEX wrappers (the *_mex.c files one level up) where MATLAB
 * is absent.  With MATLAB present the wrappers compile against MATLAB's own
 * mex.h instead (see ../make.m).  Class IDs follow MATLAB's numbering.
 */
#ifndef QSM_MEX_SHIM_H
#define QSM_MEX_SHIM_H

#include <stdbool.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef size_t mwSize;
typedef struct mxArray_tag mxArray;

typedef enum {
    mxUNKNOWN_CLASS = 0, mxCELL_CLASS, mxSTRUCT_CLASS, mxLOGICAL_CLASS,
    mxCHAR_CLASS, mxVOID_CLASS, mxDOUBLE_CLASS, mxSINGLE_CLASS, mxINT8_CLASS,
    mxUINT8_CLASS, mxINT16_CLASS, mxUINT16_CLASS, mxINT32_CLASS,
    mxUINT32_CLASS, mxINT64_CLASS, mxUINT64_CLASS
} mxClassID;

typedef enum { mxREAL = 0, mxCOMPLEX } mxComplexity;

void *mxGetData(const mxArray *a);
const mwSize *mxGetDimensions(const mxArray *a);
mwSize mxGetNumberOfDimensions(const mxArray *a);
size_t mxGetNumberOfElements(const mxArray *a);
double mxGetScalar(const mxArray *a);
mxClassID mxGetClassID(const mxArray *a);
bool mxIsSingle(const mxArray *a);
bool mxIsDouble(const mxArray *a);
bool mxIsLogical(const mxArray *a);
bool mxIsUint8(const mxArray *a);
mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID cls,
                              mxComplexity flag);
mxArray *mxDuplicateArray(const mxArray *a);
void mxDestroyArray(mxArray *a);
mxArray *mxCreateDoubleScalar(double v);

/* Raises a MATLAB error: does not return to the caller. */
void mexErrMsgTxt(const char *msg);
int mexPrintf(const char *fmt, ...);

#ifdef __cplusplus
}
#endif

#endif
