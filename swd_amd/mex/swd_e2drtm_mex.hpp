// MexFunction body shared by eigen_e2drtm_single.cpp / eigen_e2drtm_double.cpp:
// the argument list of core/eigen_e2drtm_single.cpp:67 forwarded to swd_e2drtm_f32/f64.
#pragma once

#include "swd_mex_common.hpp"

template <typename T, typename Fn>
void swd_e2drtm_mex(matlab::mex::Function& self, matlab::mex::ArgumentList outputs,
                    matlab::mex::ArgumentList inputs, Fn call, const char* who)
{
    using namespace swdmex;
    matlab::data::ArrayFactory f;
    auto iv = take<T>(inputs, 0);
    const T* v = data(iv);
    const size_t nz = (size_t)v[11], nx = (size_t)v[12];
    std::vector<matlab::data::TypedArray<T>> a;
    for (size_t k = 1; k <= 12; ++k) a.push_back(take<T>(inputs, k));
    matlab::data::buffer_ptr_t<T> o[3] = {f.createBuffer<T>(nz * nx), f.createBuffer<T>(nz * nx),
                                          f.createBuffer<T>(nz * nx)};
    const int rc = call(v, data(a[0]), data(a[1]), data(a[2]), data(a[3]), data(a[4]), data(a[5]), data(a[6]),
                        data(a[7]), data(a[8]), data(a[9]), data(a[10]), data(a[11]), o[0].get(), o[1].get(),
                        o[2].get());
    if (rc != SWD_OK) fail(self.getEngine(), who, rc);
    for (size_t k = 0; k < outputs.size() && k < 3; ++k) outputs[k] = wrap<T>(f, {nz, nx}, std::move(o[k]));
}
