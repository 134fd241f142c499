function [a_data_res] = ADWDgrad_1(nt,ng,ns,npair,is,w,m,M,SoftArgNorm,grad_outputr,grad_outputl,space_M,saveForBackwardr,saveForBackwardl)
% Drop-in for core/ADWDgrad_1.m: SoftMax adjoint source on the MI355X (libswd).
% The device path builds space_M as 1..np in every column, which is what
% SWD.m:133-136 passes; any other space_M is refused rather than ignored.
np = size(space_M,1);
if ~isequal(space_M, repmat((1:np)', 1, size(space_M,2)))
    error('ADWDgrad_1 drop-in: space_M must hold 1..np in every column (SWD.m:133-136)');
end
a_data_res = swd_adwdgrad1_mex(nt,ng,ns,npair,is,w,m,M,SoftArgNorm,double(grad_outputr(:)),double(grad_outputl(:)),saveForBackwardr,saveForBackwardl);
end
