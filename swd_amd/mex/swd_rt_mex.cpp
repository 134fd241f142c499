// [ml, dataLen, sfb] = swd_rt_mex(side, seismo, is, df, dt, np, vmin, vmax, fmin, fmax, a, b, ds, offset, m, M)
// Gateway behind the drop-in RTr.m / RTl.m / RTrAD.m / RTlAD.m (matlab/): the
// arguments of core/RTr.m:2 plus a leading side flag (0 right, 1 left).  With
// three outputs it also returns the saveForBackward struct of core/RTrAD.m:34 /
// core/RTlAD.m:35 (mm, ll0, lf, nf, ccn, mlr|mll, uxtposr|uxtposl) plus the
// scalars swd_adwdgrad1 rebuilds ll0 from (np, vmin, df, ds, m).
#include <complex>

#include "swd_mex_common.hpp"

class MexFunction : public matlab::mex::Function {
public:
    void operator()(matlab::mex::ArgumentList outputs, matlab::mex::ArgumentList inputs)
    {
        using namespace swdmex;
        matlab::data::ArrayFactory f;
        auto num = [&](size_t k) { return (double)matlab::data::TypedArray<double>(inputs[k])[0]; };
        const int side = (int)num(0);
        const matlab::data::Array sg = inputs[1];
        const size_t nt = sg.getDimensions()[0], ng = sg.getDimensions()[1];
        const int is = (int)num(2), np = (int)num(5), offset = (int)num(13), m = (int)num(14), M = (int)num(15);
        const double df = num(3), dt = num(4), vmin = num(6), vmax = num(7), fmin = num(8), fmax = num(9),
                     ds = num(12);
        const int lf = (int)std::floor(fmin / df + 0.5) + 1, nf = (int)std::floor(fmax / df + 0.5) + 1;
        const size_t npair = nf - lf + 1;
        auto ml = f.createBuffer<double>(np * npair);
        std::vector<float> mlf(np * npair), mlr(np * npair), mm(2 * np * npair);
        int L = 0, info[4] = {0, 0, 0, 0};
        int rc;
        if (sg.getType() == matlab::data::ArrayType::SINGLE) {
            auto s = matlab::data::TypedArray<float>(std::move(inputs[1]));
            rc = swd_radon_f32(side, data(s), (int)nt, (int)ng, is, df, dt, np, vmin, vmax, fmin, fmax, ds, offset,
                               m, M, mlf.data(), &L, mlr.data(), mm.data(), info);
        } else {
            auto s = matlab::data::TypedArray<double>(std::move(inputs[1]));
            rc = swd_radon(side, data(s), (int)nt, (int)ng, is, df, dt, np, vmin, vmax, fmin, fmax, ds, offset, m, M,
                           mlf.data(), &L, mlr.data(), mm.data(), info);
        }
        if (rc != SWD_OK) fail(getEngine(), "swd_rt_mex", rc);
        for (size_t k = 0; k < mlf.size(); ++k) ml.get()[k] = mlf[k];
        outputs[0] = wrap<double>(f, {(size_t)np, npair}, std::move(ml));
        if (outputs.size() > 1) outputs[1] = f.createScalar<double>(L);
        if (outputs.size() > 2) {
            const size_t nfz = info[1];
            auto mmc = f.createArray<std::complex<double>>({(size_t)np, nfz});
            auto ll0 = f.createArray<std::complex<double>>({(size_t)np, (size_t)L});
            auto mlrA = f.createArray<double>({(size_t)np, npair});
            auto pos = f.createArray<double>({1, (size_t)L});
            for (size_t k = 0; k < npair; ++k)
                for (int j = 0; j < np; ++j) {
                    mmc[j][info[0] - 1 + k] = std::complex<double>(mm[2 * (k * np + j)], mm[2 * (k * np + j) + 1]);
                    mlrA[j][k] = mlr[k * np + j];
                }
            for (int j = 0; j < np; ++j)
                for (int l = 0; l < L; ++l)
                    ll0[j][l] = std::complex<double>(0.0, 2 * M_PI * df * ((1.0 / (vmin + j)) * ((m + l) * ds)));
            for (int l = 0; l < L; ++l) pos[0][l] = side == 0 ? info[3] + l : info[3] - L + 1 + l;
            matlab::data::StructArray S = f.createStructArray(
                {1, 1}, {"mm", "ll0", "lf", "nf", "ccn", side == 0 ? "mlr" : "mll", side == 0 ? "uxtposr" : "uxtposl",
                         "np", "vmin", "df", "ds", "m"});
            S[0]["mm"] = mmc;
            S[0]["ll0"] = ll0;
            S[0]["lf"] = f.createScalar<double>(info[0]);
            S[0]["nf"] = f.createScalar<double>(info[1]);
            S[0]["ccn"] = f.createScalar<double>(info[2]);
            S[0][side == 0 ? "mlr" : "mll"] = mlrA;
            S[0][side == 0 ? "uxtposr" : "uxtposl"] = pos;
            S[0]["np"] = f.createScalar<double>(np);
            S[0]["vmin"] = f.createScalar<double>(vmin);
            S[0]["df"] = f.createScalar<double>(df);
            S[0]["ds"] = f.createScalar<double>(ds);
            S[0]["m"] = f.createScalar<double>(m);
            outputs[2] = S;
        }
    }
};
