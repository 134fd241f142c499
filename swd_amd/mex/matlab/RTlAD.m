function [ml,dataLen,saveForBackward]=RTlAD(seismo_v1,is,df,dt,np,vmin,vmax,fmin,fmax,a,b,ds,offset,m,M)
% Drop-in for core/RTlAD.m: linear Radon transform + saveForBackward on the MI355X (libswd).
[ml,dataLen,saveForBackward]=swd_rt_mex(1,seismo_v1,is,df,dt,np,vmin,vmax,fmin,fmax,a,b,ds,offset,m,M);
end
