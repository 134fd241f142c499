// Test stand-in for the subset of the MATLAB C++ Data API (MatlabDataArray.hpp)
// that the SWD MEX shims use: the class, function and member names and
// signatures the shims call, with no implementation behind them.  Used only to
// compile the shims with -fsyntax-only on a machine without MATLAB; nothing
// built against it is run.
#pragma once

#include <complex>
#include <cstddef>
#include <initializer_list>
#include <memory>
#include <string>
#include <vector>

namespace matlab {
namespace data {

enum class ArrayType { LOGICAL, CHAR, DOUBLE, SINGLE, INT8, UINT8, INT16, UINT16, INT32, UINT32, INT64, UINT64,
                       COMPLEX_DOUBLE, COMPLEX_SINGLE, CELL, STRUCT, MATLAB_STRING, UNKNOWN };

using ArrayDimensions = std::vector<size_t>;

template <typename T>
using buffer_ptr_t = std::unique_ptr<T[], void (*)(void*)>;

class Array {
public:
    Array();
    Array(const Array&);
    Array(Array&&);
    Array& operator=(const Array&);
    Array& operator=(Array&&);
    ArrayType getType() const;
    ArrayDimensions getDimensions() const;
    size_t getNumberOfElements() const;
};

template <typename T>
class ArrayElementRef {
public:
    operator T() const;
    ArrayElementRef& operator=(const T&);
    ArrayElementRef operator[](size_t);
};

template <typename T>
class TypedIterator {
public:
    T& operator*() const;
    TypedIterator& operator++();
    bool operator!=(const TypedIterator&) const;
};

template <typename T>
class TypedArray : public Array {
public:
    TypedArray(const Array&);
    TypedArray(Array&&);
    ArrayElementRef<T> operator[](size_t);
    T operator[](size_t) const;
    TypedIterator<T> begin();
    TypedIterator<T> end();
    TypedIterator<const T> cbegin() const;
    TypedIterator<const T> cend() const;
};

class StructArray;
class Reference {
public:
    template <typename T>
    operator TypedArray<T>() const;
    operator Array() const;
    Reference& operator=(const Array&);
};
class StructRef {
public:
    Reference operator[](const std::string&);
};
class StructArray : public Array {
public:
    StructArray(const Array&);
    StructRef operator[](size_t);
};

class ArrayFactory {
public:
    ArrayFactory();
    template <typename T>
    buffer_ptr_t<T> createBuffer(size_t);
    template <typename T>
    TypedArray<T> createArrayFromBuffer(ArrayDimensions, buffer_ptr_t<T>);
    template <typename T>
    TypedArray<T> createArray(ArrayDimensions);
    template <typename T>
    TypedArray<T> createScalar(const T);
    Array createScalar(const std::string&);
    StructArray createStructArray(ArrayDimensions, std::vector<std::string>);
};

}  // namespace data
}  // namespace matlab
