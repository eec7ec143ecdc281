/*
 * mgpcg_mex.c -- drop-in for src/utils/poisson_solver/mex/mgpcg_mex.c:75-122:
 *   v = mgpcg_mex(v, f, mask, vsz, tolcg, maxitcg,
 *                 tolmg, maxitmg, mu, npre, npost, nboundary, nlevels, verbose);
 * -Lap v = f on the mask by multigrid-preconditioned CG from the initial
 * guess v, on the GPU; the result is a new array of v's class.  With
 * verbose, the residual table is printed when the solve returns.
 */
#include <string.h>

#include "qsm_mex_util.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    if ((nrhs != 14) || (nlhs > 1)) {
        mexErrMsgTxt("Usage: v = mgpcg_mex(v, f, mask, vsz, tolcg, maxitcg, "
                     "tolmg, maxitmg, mu, npre, npost, nboundary, nlevels, verbose)");
        return;
    }
    size_t sz[3];
    qsm_mex_dims(prhs[0], sz);
    const int dtype = qsm_mex_dtype(prhs[0]);
    qsm_mex_same(prhs[1], prhs[0], sz);
    const uint8_t *mask = qsm_mex_mask(prhs[2], sz);
    const double *vsz = qsm_mex_vsz(prhs[3]);
    const double tol = mxGetScalar(prhs[4]);
    const int maxit = (int)mxGetScalar(prhs[5]);
    qsm_mg_opts o;
    qsm_mg_opts_default(&o);
    o.tol = mxGetScalar(prhs[6]);
    o.maxit = (int)mxGetScalar(prhs[7]);
    o.mu = (int)mxGetScalar(prhs[8]);
    o.npre = (int)mxGetScalar(prhs[9]);
    o.npost = (int)mxGetScalar(prhs[10]);
    o.nboundary = (int)mxGetScalar(prhs[11]);
    o.nlevels = (int)mxGetScalar(prhs[12]);
    const int verbose = (int)mxGetScalar(prhs[13]);
    if (o.nlevels > 16)
        mexErrMsgTxt("nlevels too large. Increase MAX_DEPTH.");
    if (maxit < 0)
        mexErrMsgTxt("maxitcg must be >= 0");

    const size_t N = sz[0] * sz[1] * sz[2];
    const size_t es = dtype == QSM_F32 ? 4 : 8;
    mwSize dims[3] = {(mwSize)sz[0], (mwSize)sz[1], (mwSize)sz[2]};
    mxArray *v = mxCreateNumericArray(3, dims, mxGetClassID(prhs[0]), mxREAL);
    memcpy(mxGetData(v), mxGetData(prhs[0]), N * es);
    const mwSize rdims[2] = {maxit > 0 ? (mwSize)maxit : 1, 1};
    mxArray *res = mxCreateNumericArray(2, rdims, mxDOUBLE_CLASS, mxREAL);
    int iters = 0;
    qsm_mex_check(qsm_mgpcg_x0(dtype, mxGetData(prhs[1]), mask, sz, vsz, tol, maxit, &o,
                               mxGetData(v), &iters, (double *)mxGetData(res)));
    if (verbose) {
        const double *r = (const double *)mxGetData(res);
        mexPrintf("\niter%6s\tresidual\n", "");
        for (int i = 0; i < iters; ++i)
            mexPrintf("%5d/%d\t%1.3e\n", i, maxit, r[i]);
    }
    mxDestroyArray(res);
    plhs[0] = v;
}
