function [ml,dataLen]=RTl(seismo_v1,is,df,dt,np,vmin,vmax,fmin,fmax,a,b,ds,offset,m,M)
% Drop-in for core/RTl.m: linear Radon transform on the MI355X (libswd).
[ml,dataLen]=swd_rt_mex(1,seismo_v1,is,df,dt,np,vmin,vmax,fmin,fmax,a,b,ds,offset,m,M);
end
