// [res, cr_pre] = swd_lhdisppick_mex(ml, vmin, cr_0, method, ini)
// Gateway behind the drop-in LHDispPick.m (core/LHDispPick.m:2): double
// spectra (SWD.m) are picked in FP64, single ones (SWD_single.m) in FP32.
#include "swd_mex_common.hpp"

class MexFunction : public matlab::mex::Function {
public:
    void operator()(matlab::mex::ArgumentList outputs, matlab::mex::ArgumentList inputs)
    {
        using namespace swdmex;
        matlab::data::ArrayFactory f;
        const bool single = inputs[0].getType() == matlab::data::ArrayType::SINGLE;
        const size_t np = inputs[0].getDimensions()[0], npair = inputs[0].getDimensions()[1];
        const double vmin = matlab::data::TypedArray<double>(inputs[1])[0];
        auto cr0 = matlab::data::TypedArray<double>(std::move(inputs[2]));
        const int method = (int)matlab::data::TypedArray<double>(inputs[3])[0];
        const int ini = (int)matlab::data::TypedArray<double>(inputs[4])[0];
        double res = 0;
        auto cr = f.createBuffer<double>(npair);
        int rc;
        if (single) {
            auto ml = matlab::data::TypedArray<float>(std::move(inputs[0]));
            rc = swd_lhdisppick(data(ml), (int)np, (int)npair, vmin, data(cr0), method, ini, &res, cr.get());
        } else {
            auto ml = matlab::data::TypedArray<double>(std::move(inputs[0]));
            rc = swd_lhdisppick_f64(data(ml), (int)np, (int)npair, vmin, data(cr0), method, ini, &res, cr.get());
        }
        if (rc != SWD_OK) fail(getEngine(), "LHDispPick", rc);
        outputs[0] = f.createScalar<double>(res);
        if (outputs.size() > 1) outputs[1] = wrap<double>(f, {npair, 1}, std::move(cr));
    }
};
