// Complex and vector types of the two precisions of the engine.
#pragma once

#include <hip/hip_runtime.h>

namespace mtg {

// real type of a complex vector type
template <class C> struct RealOf;
template <> struct RealOf<float2> { using type = float; };
template <> struct RealOf<double2> { using type = double; };
template <class C> using real_t = typename RealOf<C>::type;

// complex (float2 / double2) and 4-vector (float4 / double4) of a real type
template <class R> struct VecOf;
template <> struct VecOf<float> { using c2 = float2; using v4 = float4; };
template <> struct VecOf<double> { using c2 = double2; using v4 = double4; };
template <class R> using cplx_t = typename VecOf<R>::c2;
template <class R> using vec4_t = typename VecOf<R>::v4;

} // namespace mtg
