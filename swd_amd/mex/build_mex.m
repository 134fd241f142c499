% Build the SWD MEX shims against libswd (run from this folder in MATLAB).
% Needs MATLAB's C++ MEX support and the in-tree libswd.so (python -m swd_amd.build).
here = fileparts(mfilename('fullpath'));
root = fileparts(fileparts(here));
inc = ['-I' fullfile(root,'include')];
lib = ['-L' fullfile(root,'swd_amd')];
srcs = {'eigen_staggerfd_single.cpp','eigen_staggerfd_double.cpp','eigen_e2drtm_single.cpp', ...
        'eigen_e2drtm_double.cpp','swd_rt_mex.cpp','swd_lhdisppick_mex.cpp','swd_adwdgrad1_mex.cpp'};
for k = 1:numel(srcs)
    mex('-R2018a', inc, lib, '-lswd', fullfile(here, srcs{k}), '-outdir', here);
end
