function [a_data_res] = ADWDgrad_1(nt,ng,ns,npair,is,w,m,M,SoftArgNorm,grad_outputr,grad_outputl,space_M,saveForBackwardr,saveForBackwardl)
% Drop-in for core/ADWDgrad_1.m: SoftMax adjoint source on the MI355X (libswd).
% space_M is 1..np in every column (SWD.m:133-136) and is rebuilt on the device.
a_data_res = swd_adwdgrad1_mex(nt,ng,ns,npair,is,w,m,M,SoftArgNorm,double(grad_outputr(:)),double(grad_outputl(:)),saveForBackwardr,saveForBackwardl);
end
