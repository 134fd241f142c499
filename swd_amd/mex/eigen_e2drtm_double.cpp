// [cl_img, cm_img, illum_div] = eigen_e2drtm_double(...)
// Drop-in for the binary-only reference MEX, called by core/e2drtm_eigend.m:122.
#include "swd_e2drtm_mex.hpp"

class MexFunction : public matlab::mex::Function {
public:
    void operator()(matlab::mex::ArgumentList outputs, matlab::mex::ArgumentList inputs)
    {
        swd_e2drtm_mex<double>(*this, outputs, inputs, swd_e2drtm_f64, "eigen_e2drtm_double");
    }
};
