// Shared helpers of the SWD MEX shims (MATLAB C++ Data API).  Built only
// where MATLAB's mex.hpp / MatlabDataArray.hpp exist (see swd_amd/mex/build_mex.m);
// every shim forwards to libswd through the C ABI in include/swd_abi.h.
#pragma once

#include <string>
#include <vector>

#include "MatlabDataArray.hpp"
#include "mex.hpp"
#include "mexAdapter.hpp"
#include "swd_abi.h"

namespace swdmex {

template <typename T>
const T* data(const matlab::data::TypedArray<T>& a)
{
    return &*a.cbegin();
}

template <typename T>
matlab::data::TypedArray<T> take(matlab::mex::ArgumentList& in, size_t k)
{
    return matlab::data::TypedArray<T>(std::move(in[k]));
}

// Raise a MATLAB error carrying the library's message for a failed call.
inline void fail(const std::shared_ptr<matlab::engine::MATLABEngine>& eng, const char* who, int rc)
{
    matlab::data::ArrayFactory f;
    std::string msg = std::string(who) + ": " + swd_last_error() + " (code " + std::to_string(rc) + ")";
    eng->feval(u"error", 0, std::vector<matlab::data::Array>({f.createScalar(msg)}));
}

template <typename T>
matlab::data::Array wrap(matlab::data::ArrayFactory& f, std::vector<size_t> dims, matlab::data::buffer_ptr_t<T> b)
{
    return f.createArrayFromBuffer<T>(matlab::data::ArrayDimensions(dims), std::move(b));
}

}  // namespace swdmex
