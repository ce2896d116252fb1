// Scalar/index types and small fixed-size vector helpers used across the host API.
// Mirrors the names of the reference's include/base/Type.h:28-44 and
// include/base/vector_op.h:15-140 (Scalar = double, uInt = uint64, vectorNf =
// std::array) so that code written against the reference compiles unchanged.
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

using Scalar = double;
using Scalar64 = double;
using Scalar128 = long double;
typedef int64_t Int;
typedef uint64_t uInt;

typedef std::array<Scalar, 2> vector2f;
typedef std::array<Scalar, 3> vector3f;
typedef std::array<Scalar, 4> vector4f;
typedef std::array<uInt, 2> vector2u;
typedef std::array<uInt, 3> vector3u;
typedef std::array<uInt, 4> vector4u;
typedef std::array<uInt, 5> vector5u;
typedef std::array<uInt, 6> vector6u;
typedef std::array<uInt, 8> vector8u;

template <typename T>
inline void debug(const T& s) { std::cerr << s << std::endl; }
template <typename T>
inline void print(const T& s) { std::cout << s << std::endl; }

// ---- element-wise std::array arithmetic (reference: base/vector_op.h) ----
template <typename T, size_t N>
inline std::array<T, N> operator+(const std::array<T, N>& a, const std::array<T, N>& b) {
    std::array<T, N> r;
    for (size_t i = 0; i < N; ++i) r[i] = a[i] + b[i];
    return r;
}
template <typename T, size_t N>
inline std::array<T, N> operator-(const std::array<T, N>& a, const std::array<T, N>& b) {
    std::array<T, N> r;
    for (size_t i = 0; i < N; ++i) r[i] = a[i] - b[i];
    return r;
}
template <typename T, size_t N>
inline std::array<T, N> operator*(const std::array<T, N>& a, T s) {
    std::array<T, N> r;
    for (size_t i = 0; i < N; ++i) r[i] = a[i] * s;
    return r;
}
template <typename T, size_t N>
inline std::array<T, N> operator*(T s, const std::array<T, N>& a) { return a * s; }
template <typename T, size_t N>
inline std::array<T, N> operator/(const std::array<T, N>& a, T s) {
    std::array<T, N> r;
    for (size_t i = 0; i < N; ++i) r[i] = a[i] / s;
    return r;
}
template <typename T, size_t N>
inline std::array<T, N>& operator+=(std::array<T, N>& a, const std::array<T, N>& b) {
    for (size_t i = 0; i < N; ++i) a[i] += b[i];
    return a;
}
template <typename T, size_t N>
inline std::array<T, N>& operator*=(std::array<T, N>& a, T s) {
    for (size_t i = 0; i < N; ++i) a[i] *= s;
    return a;
}

template <typename T, size_t N>
inline T vec_dot(const std::array<T, N>& a, const std::array<T, N>& b) {
    T s = 0;
    for (size_t i = 0; i < N; ++i) s += a[i] * b[i];
    return s;
}
inline vector3f vec_cross(const vector3f& a, const vector3f& b) {
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
template <typename T, size_t N>
inline T vec_length(const std::array<T, N>& a) { return std::sqrt(vec_dot(a, a)); }
template <typename T, size_t N>
inline std::array<T, N> vec_unit(const std::array<T, N>& a) { return a / vec_length(a); }
