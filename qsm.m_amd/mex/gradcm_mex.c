/*
 * gradcm_mex.c -- drop-in for the reference's gradcm_mex (src/utils/fd/mex/gradcm_mex.c:15-50):
 *   gradcm_mex(dx, dy, dz, u, G, h);
 * The three outputs are written in place into the caller's arrays
 * (prhs[0..2]), as the reference does; the work runs on the GPU.
 */
#include "qsm_mex_util.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    if ((nrhs != 6) || (nlhs > 1)) {
        mexErrMsgTxt("Usage: gradcm_mex(dx, dy, dz, u, G, h);");
        return;
    }
    size_t sz[3];
    qsm_mex_dims(prhs[0], sz);
    const int dtype = qsm_mex_dtype(prhs[0]);
    for (int a = 1; a < 4; ++a)
        qsm_mex_same(prhs[a], prhs[0], sz);
    const uint8_t *G = qsm_mex_mask(prhs[4], sz);
    const double *h = qsm_mex_vsz(prhs[5]);
    qsm_mex_check(qsm_gradcm(dtype, mxGetData(prhs[0]), mxGetData(prhs[1]),
                             mxGetData(prhs[2]), mxGetData(prhs[3]), G, sz, h));
    if (nlhs == 1)
        plhs[0] = mxCreateDoubleScalar(1.0);
}
