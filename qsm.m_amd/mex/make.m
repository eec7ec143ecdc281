function make(libdir)
%MAKE Build the MI355X drop-in MEX files against MATLAB's mex.h.
%
%   make            % libqsm_hip.so in ../lib
%   make(libdir)    % libqsm_hip.so in libdir
%
%   Produces pd_tgv_mex, pd_tv_mex, gradfp_mex, gradfp_adj_mex, gradf_mex,
%   gradfm_mex, gradf_adj_mex, gradfm_adj_mex, gradb_mex, gradbm_mex,
%   gradb_adj_mex, gradbm_adj_mex, gradc_mex, gradcm_mex, lap1_mex, lapw_mex
%   and mgpcg_mex with the argument layouts of QSM.m's own MEX files; put this directory ahead of QSM.m's
%   mex directories on the MATLAB path so src/inversion runs unchanged.

    here = fileparts(mfilename('fullpath'));
    if nargin < 1
        libdir = fullfile(here, '..', 'lib');
    end
    inc = fullfile(here, '..', '..', 'include');
    names = {'pd_tgv_mex', 'pd_tv_mex', 'gradfp_mex', 'gradfp_adj_mex', ...
             'gradf_mex', 'gradfm_mex', 'gradf_adj_mex', 'gradfm_adj_mex', ...
             'gradb_mex', 'gradbm_mex', 'gradb_adj_mex', 'gradbm_adj_mex', ...
             'gradc_mex', 'gradcm_mex', 'lap1_mex', 'lapw_mex', 'mgpcg_mex'};
    for k = 1:numel(names)
        src = fullfile(here, [names{k}, '.c']);
        mex('-R2017b', 'CFLAGS="$CFLAGS -std=c11 -O2"', ['-I', inc], ...
            ['-L', libdir], '-lqsm_hip', ['LDFLAGS=$LDFLAGS -Wl,-rpath,', libdir], ...
            '-outdir', here, src);
    end
end
