/*
 * lap1_mex.c -- drop-in for the reference's lap1_mex (src/utils/fd/mex/lap1_mex.c:29-60),
 * called by src/utils/fd/lap.m:
 *   lap1_mex(d2u, u, G, h);
 * d2u is written in place on G (prhs[0]), as the reference does; voxels
 * outside G keep their values.  The work runs on the GPU.
 */
#include "qsm_mex_util.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    if ((nrhs != 4) || (nlhs > 1)) {
        mexErrMsgTxt("Usage: lap_mex(d2u, u, G, h);");
        return;
    }
    size_t sz[3];
    qsm_mex_dims(prhs[0], sz);
    const int dtype = qsm_mex_dtype(prhs[0]);
    qsm_mex_same(prhs[1], prhs[0], sz);
    const uint8_t *G = qsm_mex_mask(prhs[2], sz);
    const double *h = qsm_mex_vsz(prhs[3]);
    qsm_mex_check(qsm_lap1(dtype, mxGetData(prhs[0]), mxGetData(prhs[1]), G, sz, h));
    if (nlhs == 1)
        plhs[0] = mxCreateDoubleScalar(1.0);
}
