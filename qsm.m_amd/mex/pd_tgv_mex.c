/*
 * pd_tgv_mex.c -- drop-in for src/inversion/mex/pd_tgv_mex.c:104-135:
 *   chi = pd_tgv_mex(d2f, mask, vsz, lam, alpha, tol, maxiter, verbose)
 * Same arguments, same usage message, a fresh chi of d2f's class returned;
 * the solve runs on the GPU through qsm_pd_tgv.
 */
#include "qsm_mex_util.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    if ((nrhs != 8) || (nlhs > 1)) {
        mexErrMsgTxt(
            "Usage: chi = pd_tgv_mex(d2f, mask, vsz, lam, alpha, tol, maxiter, verbose);");
        return;
    }
    size_t sz[3];
    qsm_mex_dims(prhs[0], sz);
    const int dtype = qsm_mex_dtype(prhs[0]);
    const uint8_t *m = qsm_mex_mask(prhs[1], sz);
    const double *h = qsm_mex_vsz(prhs[2]);
    const double lam = mxGetScalar(prhs[3]);
    const double alp = mxGetScalar(prhs[4]);
    const double tol = mxGetScalar(prhs[5]);
    const uint32_t maxiter = (uint32_t)mxGetScalar(prhs[6]);
    const uint32_t verbose = (uint32_t)mxGetScalar(prhs[7]);

    mxArray *chi = mxCreateNumericArray(3, sz, mxGetClassID(prhs[0]), mxREAL);
    const int rc = qsm_pd_tgv(dtype, mxGetData(prhs[0]), m, sz, h, lam, alp, tol,
                              maxiter, verbose, mxGetData(chi), qsm_mex_print, NULL);
    if (rc != QSM_OK) {
        mxDestroyArray(chi);
        mexErrMsgTxt(qsm_last_error());
        return;
    }
    plhs[0] = chi;
}
