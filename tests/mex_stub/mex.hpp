// Test stand-in for the subset of MATLAB's C++ MEX API (mex.hpp) the SWD MEX
// shims use; see MatlabDataArray.hpp beside it.  Syntax checks only.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "MatlabDataArray.hpp"

namespace matlab {
namespace engine {
class MATLABEngine {
public:
    void feval(const std::u16string&, int, const std::vector<matlab::data::Array>&);
};
}  // namespace engine
namespace mex {
class ArgumentList {
public:
    matlab::data::Array& operator[](size_t);
    size_t size() const;
};
class Function {
public:
    virtual ~Function();
    virtual void operator()(ArgumentList outputs, ArgumentList inputs);
    std::shared_ptr<matlab::engine::MATLABEngine> getEngine();
};
}  // namespace mex
}  // namespace matlab
