/*
 * lapw_mex.c -- drop-in for src/unwrap/mex/lapw_mex.c:21-51:
 *   lapw_mex(d2u, u, h);
 * The interior of d2u is overwritten in place (faces keep the caller's
 * values, zeros from lapw.m); the work runs on the GPU.
 */
#include "qsm_mex_util.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    if ((nrhs != 3) || (nlhs > 1)) {
        mexErrMsgTxt("Usage: lapw_mex(d2u, u, h);");
        return;
    }
    size_t sz[3];
    qsm_mex_dims(prhs[0], sz);
    const int dtype = qsm_mex_dtype(prhs[0]);
    qsm_mex_same(prhs[1], prhs[0], sz);
    const double *h = qsm_mex_vsz(prhs[2]);
    qsm_mex_check(qsm_lapw(dtype, mxGetData(prhs[0]), mxGetData(prhs[1]), sz, h));
    if (nlhs == 1)
        plhs[0] = mxCreateDoubleScalar(1.0);
}
