// a_data_res = swd_adwdgrad1_mex(nt, ng, ns, npair, is, w, m, M, SoftArgNorm, gor, gol, sfbR, sfbL)
// Gateway behind the drop-in ADWDgrad_1.m (core/ADWDgrad_1.m:1); sfbR/sfbL are
// the structs returned by the drop-in RTrAD / RTlAD (or 0 when a branch is off).
#include <complex>

#include "swd_mex_common.hpp"

namespace {
struct Side {
    std::vector<float> mm, ml;
    int pos = 0, L = 0, np = 0, lf = 1, ccn = 1;
    double vmin = 0, df = 1, ds = 1;
    bool ok = false;
};

Side side_of(const matlab::data::Array& a, bool right)
{
    Side s;
    if (a.getType() != matlab::data::ArrayType::STRUCT) return s;
    matlab::data::StructArray S(a);
    auto num = [&](const char* k) { return (double)matlab::data::TypedArray<double>(S[0][k])[0]; };
    s.lf = (int)num("lf");
    s.ccn = (int)num("ccn");
    s.np = (int)num("np");
    s.vmin = num("vmin");
    s.df = num("df");
    s.ds = num("ds");
    matlab::data::TypedArray<double> ml = S[0][right ? "mlr" : "mll"];
    matlab::data::TypedArray<std::complex<double>> mm = S[0]["mm"];
    matlab::data::TypedArray<double> pos = S[0][right ? "uxtposr" : "uxtposl"];
    const size_t npair = ml.getDimensions()[1];
    s.L = (int)pos.getNumberOfElements();
    s.pos = right ? (int)pos[0][0] : (int)pos[0][s.L - 1];
    s.ml.resize((size_t)s.np * npair);
    s.mm.resize(2 * (size_t)s.np * npair);
    for (size_t k = 0; k < npair; ++k)
        for (int j = 0; j < s.np; ++j) {
            s.ml[k * s.np + j] = (float)ml[j][k];
            const std::complex<double> z = mm[j][s.lf - 1 + k];
            s.mm[2 * (k * s.np + j)] = (float)z.real();
            s.mm[2 * (k * s.np + j) + 1] = (float)z.imag();
        }
    s.ok = true;
    return s;
}
}  // namespace

class MexFunction : public matlab::mex::Function {
public:
    void operator()(matlab::mex::ArgumentList outputs, matlab::mex::ArgumentList inputs)
    {
        using namespace swdmex;
        matlab::data::ArrayFactory f;
        auto num = [&](size_t k) { return (double)matlab::data::TypedArray<double>(inputs[k])[0]; };
        const int nt = (int)num(0), ng = (int)num(1), ns = (int)num(2), npair = (int)num(3), is = (int)num(4);
        const double w = num(5), beta = num(8);
        const int m = (int)num(6), M = (int)num(7);
        auto gor = matlab::data::TypedArray<double>(std::move(inputs[9]));
        auto gol = matlab::data::TypedArray<double>(std::move(inputs[10]));
        const Side R = side_of(inputs[11], true), Lf = side_of(inputs[12], false);
        const Side& ref = R.ok ? R : Lf;
        std::vector<float> a((size_t)nt * ng, 0.f);
        if (ref.ok) {
            const int rc = swd_adwdgrad1(nt, ng, ns, npair, is, w, m, M, beta, data(gor), data(gol),
                                         R.ok ? R.mm.data() : nullptr, R.ok ? R.ml.data() : nullptr, R.pos, R.L,
                                         Lf.ok ? Lf.mm.data() : nullptr, Lf.ok ? Lf.ml.data() : nullptr, Lf.pos, Lf.L,
                                         ref.np, ref.vmin, ref.df, ref.ds, ref.lf, ref.ccn, a.data());
            if (rc != SWD_OK) fail(getEngine(), "ADWDgrad_1", rc);
        }
        auto out = f.createBuffer<double>((size_t)nt * ng);
        for (size_t k = 0; k < a.size(); ++k) out.get()[k] = a[k];
        outputs[0] = wrap<double>(f, {(size_t)nt, (size_t)ng}, std::move(out));
    }
};
