function [res,cr_pre] = LHDispPick(ml,vmin,cr_0,method,ini)
% Drop-in for core/LHDispPick.m: pickFDC1 tracking pick (method 1) or column
% argmax (method 2) and the mean squared pick misfit, on the MI355X (libswd).
[res,cr_pre] = swd_lhdisppick_mex(double(ml),double(vmin),double(cr_0(:)),double(method),double(ini));
end
