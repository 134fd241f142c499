// [res, cr_pre] = swd_lhdisppick_mex(ml, vmin, cr_0, method, ini)
// Gateway behind the drop-in LHDispPick.m (core/LHDispPick.m:2).
#include "swd_mex_common.hpp"

class MexFunction : public matlab::mex::Function {
public:
    void operator()(matlab::mex::ArgumentList outputs, matlab::mex::ArgumentList inputs)
    {
        using namespace swdmex;
        matlab::data::ArrayFactory f;
        auto ml = matlab::data::TypedArray<double>(std::move(inputs[0]));
        const size_t np = ml.getDimensions()[0], npair = ml.getDimensions()[1];
        std::vector<float> m32(np * npair);
        size_t k = 0;
        for (auto v : ml) m32[k++] = (float)v;
        const double vmin = matlab::data::TypedArray<double>(inputs[1])[0];
        auto cr0 = matlab::data::TypedArray<double>(std::move(inputs[2]));
        const int method = (int)matlab::data::TypedArray<double>(inputs[3])[0];
        const int ini = (int)matlab::data::TypedArray<double>(inputs[4])[0];
        double res = 0;
        auto cr = f.createBuffer<double>(npair);
        const int rc = swd_lhdisppick(m32.data(), (int)np, (int)npair, vmin, data(cr0), method, ini, &res, cr.get());
        if (rc != SWD_OK) fail(getEngine(), "LHDispPick", rc);
        outputs[0] = f.createScalar<double>(res);
        if (outputs.size() > 1) outputs[1] = wrap<double>(f, {npair, 1}, std::move(cr));
    }
};
