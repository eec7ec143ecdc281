/*
 * gradb_adj_mex.c -- drop-in for the reference's gradb_adj_mex (src/utils/fd/mex/gradb_adj_mex.c:15-49):
 *   gradb_adj_mex(du, x, y, z, h);
 * The output du (prhs[0]) is overwritten in place, as the reference does;
 * the work runs on the GPU.  The voxel (nx-1, ny-1, nz-1) keeps its value:
 * the reference never writes it.
 */
#include "qsm_mex_util.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    if ((nrhs != 5) || (nlhs > 1)) {
        mexErrMsgTxt("Usage: gradb_adj_mex(du, x, y, z, h);");
        return;
    }
    size_t sz[3];
    qsm_mex_dims(prhs[0], sz);
    const int dtype = qsm_mex_dtype(prhs[0]);
    for (int a = 1; a < 4; ++a)
        qsm_mex_same(prhs[a], prhs[0], sz);
    const double *h = qsm_mex_vsz(prhs[4]);
    qsm_mex_check(qsm_gradb_adj(dtype, mxGetData(prhs[0]), mxGetData(prhs[1]),
                                mxGetData(prhs[2]), mxGetData(prhs[3]), sz, h));
    if (nlhs == 1)
        plhs[0] = mxCreateDoubleScalar(1.0);
}
