// MexFunction body shared by eigen_staggerfd_single.cpp and
// eigen_staggerfd_double.cpp: the argument list of
// core/eigen_staggerfd_single.cpp:68 forwarded to swd_staggerfd_f32/f64.
#pragma once

#include "swd_mex_common.hpp"

template <typename T, typename Fn>
void swd_staggerfd_mex(matlab::mex::Function& self, matlab::mex::ArgumentList outputs,
                       matlab::mex::ArgumentList inputs, Fn call, const char* who)
{
    using namespace swdmex;
    matlab::data::ArrayFactory f;
    auto iv = take<T>(inputs, 0);
    const T* v = data(iv);
    const size_t nt = (size_t)v[0], ng = (size_t)v[4], in_wf = (size_t)v[12], nz = (size_t)v[13],
                 nx = (size_t)v[14];
    const int format = (int)v[15];
    const size_t nrec = in_wf ? nt / in_wf : 0, nsn = nz * nx * nrec;
    std::vector<matlab::data::TypedArray<T>> a;
    for (size_t k = 1; k <= 8; ++k) a.push_back(take<T>(inputs, k));
    auto su = f.createBuffer<T>(nt * ng), sw = f.createBuffer<T>(nt * ng);
    const bool want_s = outputs.size() > 2;
    // buffer_ptr_t has a function-pointer deleter: no default construction
    const size_t nb = want_s ? nsn : 1;
    matlab::data::buffer_ptr_t<T> sn[4] = {f.createBuffer<T>(nb), f.createBuffer<T>(nb), f.createBuffer<T>(nb),
                                           f.createBuffer<T>(nb)};
    const int rc = call(v, data(a[0]), data(a[1]), data(a[2]), data(a[3]), data(a[4]), data(a[5]), data(a[6]),
                        data(a[7]), su.get(), sw.get(), want_s ? sn[0].get() : nullptr,
                        want_s ? sn[1].get() : nullptr, want_s ? sn[2].get() : nullptr,
                        want_s ? sn[3].get() : nullptr);
    if (rc != SWD_OK) fail(self.getEngine(), who, rc);
    outputs[0] = wrap<T>(f, {nt, ng}, std::move(su));
    if (outputs.size() > 1) outputs[1] = wrap<T>(f, {nt, ng}, std::move(sw));
    if (want_s) {
        std::vector<size_t> d = format == 2 ? std::vector<size_t>{nz, nx * nrec} : std::vector<size_t>{nz, nx, nrec};
        matlab::data::StructArray S = f.createStructArray({1, 1}, {"fux", "fuz", "bwx", "bwz"});
        const char* names[4] = {"fux", "fuz", "bwx", "bwz"};
        for (int k = 0; k < 4; ++k) S[0][names[k]] = wrap<T>(f, d, std::move(sn[k]));
        outputs[2] = S;
    }
}
