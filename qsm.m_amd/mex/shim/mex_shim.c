/*
 * mex_shim.c -- implementation of the minimal MEX API stand-in (mex.h in this
 * directory) plus a small driver interface (shim_*) through which tests call
 * a wrapper's mexFunction the way MATLAB would: arrays are zero-initialised
 * on creation, mexErrMsgTxt unwinds back to shim_invoke, and mexPrintf
 * output is collected for inspection.
 */
#include <setjmp.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"

struct mxArray_tag {
    mxClassID cls;
    mwSize ndim;
    mwSize dims[8];
    size_t n, esz;
    void *data;
};

static size_t class_size(mxClassID c)
{
    switch (c) {
    case mxDOUBLE_CLASS: case mxINT64_CLASS: case mxUINT64_CLASS: return 8;
    case mxSINGLE_CLASS: case mxINT32_CLASS: case mxUINT32_CLASS: return 4;
    case mxINT16_CLASS: case mxUINT16_CLASS: case mxCHAR_CLASS: return 2;
    default: return 1;
    }
}

void *mxGetData(const mxArray *a) { return a ? a->data : NULL; }
const mwSize *mxGetDimensions(const mxArray *a) { return a->dims; }
mwSize mxGetNumberOfDimensions(const mxArray *a) { return a->ndim; }
size_t mxGetNumberOfElements(const mxArray *a) { return a->n; }
mxClassID mxGetClassID(const mxArray *a) { return a->cls; }
bool mxIsSingle(const mxArray *a) { return a->cls == mxSINGLE_CLASS; }
bool mxIsDouble(const mxArray *a) { return a->cls == mxDOUBLE_CLASS; }
bool mxIsLogical(const mxArray *a) { return a->cls == mxLOGICAL_CLASS; }
bool mxIsUint8(const mxArray *a) { return a->cls == mxUINT8_CLASS; }

double mxGetScalar(const mxArray *a)
{
    if (!a || a->n == 0) return 0.0;
    switch (a->cls) {
    case mxDOUBLE_CLASS: return ((const double *)a->data)[0];
    case mxSINGLE_CLASS: return ((const float *)a->data)[0];
    case mxUINT32_CLASS: return ((const uint32_t *)a->data)[0];
    case mxINT32_CLASS: return ((const int32_t *)a->data)[0];
    case mxLOGICAL_CLASS: case mxUINT8_CLASS: return ((const uint8_t *)a->data)[0];
    default: return 0.0;
    }
}

mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID cls,
                              mxComplexity flag)
{
    (void)flag;
    mxArray *a = calloc(1, sizeof *a);
    if (!a) return NULL;
    a->cls = cls;
    a->ndim = ndim < 2 ? 2 : (ndim > 8 ? 8 : ndim);
    a->n = 1;
    for (mwSize d = 0; d < a->ndim; ++d) {
        a->dims[d] = d < ndim ? dims[d] : 1;
        a->n *= a->dims[d];
    }
    a->esz = class_size(cls);
    a->data = calloc(a->n ? a->n : 1, a->esz);
    return a;
}

mxArray *mxDuplicateArray(const mxArray *a)
{
    mxArray *b = mxCreateNumericArray(a->ndim, a->dims, a->cls, mxREAL);
    if (b) memcpy(b->data, a->data, a->n * a->esz);
    return b;
}

void mxDestroyArray(mxArray *a)
{
    if (!a) return;
    free(a->data);
    free(a);
}

mxArray *mxCreateDoubleScalar(double v)
{
    const mwSize one[2] = {1, 1};
    mxArray *a = mxCreateNumericArray(2, one, mxDOUBLE_CLASS, mxREAL);
    if (a) ((double *)a->data)[0] = v;
    return a;
}

/* ---- driver interface ---------------------------------------------------- */

static jmp_buf g_jmp;
static int g_armed = 0;
static char g_err[1024];
static char *g_out = NULL;
static size_t g_out_len = 0, g_out_cap = 0;

void mexErrMsgTxt(const char *msg)
{
    snprintf(g_err, sizeof g_err, "%s", msg ? msg : "");
    if (g_armed)
        longjmp(g_jmp, 1);
    fprintf(stderr, "mexErrMsgTxt outside shim_invoke: %s\n", g_err);
    abort();
}

int mexPrintf(const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    int n = vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (n < 0) return n;
    if ((size_t)n >= sizeof buf) n = sizeof buf - 1;
    if (g_out_len + (size_t)n + 1 > g_out_cap) {
        size_t cap = g_out_cap ? 2 * g_out_cap : 4096;
        while (cap < g_out_len + (size_t)n + 1) cap *= 2;
        char *p = realloc(g_out, cap);
        if (!p) return -1;
        g_out = p;
        g_out_cap = cap;
    }
    memcpy(g_out + g_out_len, buf, (size_t)n);
    g_out_len += (size_t)n;
    g_out[g_out_len] = 0;
    return n;
}

typedef void (*mex_fn)(int, mxArray **, int, const mxArray **);

/* Calls fn like MATLAB does; returns 0, or 1 if the function raised
 * mexErrMsgTxt (message in shim_last_error()). */
int shim_invoke(mex_fn fn, int nlhs, mxArray **plhs, int nrhs, const mxArray **prhs)
{
    g_err[0] = 0;
    g_armed = 1;
    if (setjmp(g_jmp)) {
        g_armed = 0;
        return 1;
    }
    fn(nlhs, plhs, nrhs, prhs);
    g_armed = 0;
    return 0;
}

const char *shim_last_error(void) { return g_err; }
const char *shim_output(void) { return g_out ? g_out : ""; }
void shim_clear_output(void) { g_out_len = 0; if (g_out) g_out[0] = 0; }
