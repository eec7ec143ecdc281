/*
 * gradc_mex.c -- drop-in for the reference's gradc_mex (src/utils/fd/mex/gradc_mex.c:13-47):
 *   gradc_mex(dx, dy, dz, u, h);
 * The three outputs are written in place into the caller's arrays
 * (prhs[0..2]), as the reference does; the work runs on the GPU.
 * The voxel (nx-1, ny-1, nz-1) keeps its value: the reference never
 * writes it.
 */
#include "qsm_mex_util.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    if ((nrhs != 5) || (nlhs > 1)) {
        mexErrMsgTxt("Usage: gradc_mex(dx, dy, dz, u, h);");
        return;
    }
    size_t sz[3];
    qsm_mex_dims(prhs[0], sz);
    const int dtype = qsm_mex_dtype(prhs[0]);
    for (int a = 1; a < 4; ++a)
        qsm_mex_same(prhs[a], prhs[0], sz);
    const double *h = qsm_mex_vsz(prhs[4]);
    qsm_mex_check(qsm_gradc(dtype, mxGetData(prhs[0]), mxGetData(prhs[1]),
                            mxGetData(prhs[2]), mxGetData(prhs[3]), sz, h));
    if (nlhs == 1)
        plhs[0] = mxCreateDoubleScalar(1.0);
}
