// [seismo_u, seismo_w, S] = eigen_staggerfd_single(iv, temp, ca, cl, cm, cm1, b, b1, s)
// Drop-in for the reference MEX of the same name (core/eigen_staggerfd_single.cpp:68),
// called by core/staggerfd_eigen.m:139.
#include "swd_staggerfd_mex.hpp"

class MexFunction : public matlab::mex::Function {
public:
    void operator()(matlab::mex::ArgumentList outputs, matlab::mex::ArgumentList inputs)
    {
        swd_staggerfd_mex<float>(*this, outputs, inputs, swd_staggerfd_f32, "eigen_staggerfd_single");
    }
};
