// [cl_img, cm_img, illum_div] = eigen_e2drtm_single(iv, temp, ca, cl, cm, cm1, b, b1, adj, fux, fuz, bwx, bwz)
// Drop-in for the reference MEX (core/eigen_e2drtm_single.cpp:67), called by core/e2drtm_eigen.m:122.
#include "swd_e2drtm_mex.hpp"

class MexFunction : public matlab::mex::Function {
public:
    void operator()(matlab::mex::ArgumentList outputs, matlab::mex::ArgumentList inputs)
    {
        swd_e2drtm_mex<float>(*this, outputs, inputs, swd_e2drtm_f32, "eigen_e2drtm_single");
    }
};
