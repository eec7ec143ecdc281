"""qsm_amd -- MI355X (gfx950) implementation of QSM.m's primal-dual TV/TGV
dipole-inversion path.

Compute runs in libqsm_hip.so (hand-written HIP kernels behind the C ABI in
include/qsm_hip.h); this package is the host-side mirror of the reference's
MEX/MATLAB interface for that path:

  pd.pd_tgv_mex / pd.pd_tv_mex      src/inversion/mex/pd_t{g,}v_mex.c
  fd.gradfp_mex ... fd.lapw_mex      the finite-difference MEX files
  sstgv.sstgv                        src/inversion/sstgv.m
  dipole.dipole_conv                 dipoleKernel.m + fft3/ifft3 (+ padfastfft)
  fastfft.nextfastfft/padfastfft     third_party/nextfastfft.m, padfastfft.m
  tv.tv / rts.rts                    src/inversion/tv.m, rts.m (ADMM)
  bgremove.sharp / bgremove.vsharp   src/bgremove/sharp.m, vsharp.m
  unwrap.unwrap_laplacian            src/unwrap/unwrapLaplacian.m
  unwrap.mgpcg                       src/utils/poisson_solver/mgpcg.m
"""
from ._lib import F32, F64, TGV, TV, QsmAdmissionError, QsmError, lib  # noqa: F401

__all__ = ["F32", "F64", "TGV", "TV", "QsmError", "QsmAdmissionError", "lib"]
