function [res,cr_pre] = LHDispPick(ml,vmin,cr_0,method,ini)
% Drop-in for core/LHDispPick.m: pickFDC1 tracking pick (method 1) or column
% argmax (method 2) and the mean squared pick misfit, on the MI355X (libswd).
% ml keeps its class: double spectra (SWD.m) are picked in FP64, single ones
% (SWD_single.m) in FP32.  As in core/LHDispPick.m:5,8, method 1 returns a
% column (idx') and method 2 a row (max over the columns of ml).
if ~isa(ml,'single'), ml = double(ml); end
[res,cr_pre] = swd_lhdisppick_mex(ml,double(vmin),double(cr_0(:)),double(method),double(ini));
if method ~= 1, cr_pre = cr_pre.'; end
end
