// [seismo_u, seismo_w, S] = eigen_staggerfd_double(iv, temp, ca, cl, cm, cm1, b, b1, s)
// Drop-in for the binary-only reference MEX of the same name, called by
// core/staggerfd_eigend.m:139.  Computes in FP64 on the GPU.
#include "swd_staggerfd_mex.hpp"

class MexFunction : public matlab::mex::Function {
public:
    void operator()(matlab::mex::ArgumentList outputs, matlab::mex::ArgumentList inputs)
    {
        swd_staggerfd_mex<double>(*this, outputs, inputs, swd_staggerfd_f64, "eigen_staggerfd_double");
    }
};
