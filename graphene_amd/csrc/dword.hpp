// Double-word floating point ("two<float>"): a value is the unevaluated sum
// hi + lo of two float32 numbers with |lo| <= ulp(hi) / 2, about 48
// significant bits. It is the element type "twofloat32" of the solver
// configurations (lg/common/Type.cpp:69-88), which the reference uses as the
// extended precision of mixed-precision refinement on hardware without
// float64 (iterative-refinement/Solver.cpp:64-110; arithmetic from the
// libtwofloat 0.2.0 package, conanfile.py:64, which is not part of the
// reference checkout).
//
// The operations follow M. Joldes, J.-M. Muller, V. Popescu, "Tight and
// rigorous error bounds for basic building blocks of double-word
// arithmetic", ACM Trans. Math. Softw. 44(2), 2017 (u = 2^-24):
//   twoSum    Knuth's error-free sum, s + t == a + b exactly
//   fastTwoSum  Dekker's error-free sum for |a| >= |b|
//   twoProd   error-free product with one FMA, p + e == a * b exactly
//   add       the accurate double-word sum (two twoSums, two fastTwoSums),
//             relative error <= 3u^2 + 13u^3
//   mul       the double-word product with FMA-corrected cross terms,
//             relative error <= 5u^2
//   sub       add of the negated operand
// Every step is an IEEE float32 operation in program order (the library is
// built without FP contraction; the FMAs are explicit), so host and device
// results are identical bit for bit.
#pragma once

#include <hip/hip_runtime.h>

namespace gph {

struct DWord {
  float hi, lo;
};

__host__ __device__ __forceinline__ float dwFma(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}

__host__ __device__ __forceinline__ DWord dwTwoSum(float a, float b) {
  const float s = a + b;
  const float ap = s - b;
  const float bp = s - ap;
  const float da = a - ap;
  const float db = b - bp;
  return DWord{s, da + db};
}

__host__ __device__ __forceinline__ DWord dwFastTwoSum(float a, float b) {
  const float s = a + b;
  const float z = s - a;
  return DWord{s, b - z};
}

__host__ __device__ __forceinline__ DWord dwTwoProd(float a, float b) {
  const float p = a * b;
  return DWord{p, dwFma(a, b, -p)};
}

__host__ __device__ __forceinline__ DWord dwFrom(float a) { return DWord{a, 0.0f}; }

// nearest double word of a float64 value (hi = fl32(a), lo = fl32(a - hi))
__host__ __device__ __forceinline__ DWord dwFrom(double a) {
  const float h = (float)a;
  return DWord{h, (float)(a - (double)h)};
}

__host__ __device__ __forceinline__ DWord dwAdd(DWord x, DWord y) {
  const DWord s = dwTwoSum(x.hi, y.hi);
  const DWord t = dwTwoSum(x.lo, y.lo);
  const float c = s.lo + t.hi;
  const DWord v = dwFastTwoSum(s.hi, c);
  const float w = t.lo + v.lo;
  return dwFastTwoSum(v.hi, w);
}

__host__ __device__ __forceinline__ DWord dwNeg(DWord x) { return DWord{-x.hi, -x.lo}; }

__host__ __device__ __forceinline__ DWord dwSub(DWord x, DWord y) { return dwAdd(x, dwNeg(y)); }

__host__ __device__ __forceinline__ DWord dwMul(DWord x, DWord y) {
  const DWord c = dwTwoProd(x.hi, y.hi);
  const float tl0 = x.lo * y.lo;
  const float tl1 = dwFma(x.hi, y.lo, tl0);
  const float cl2 = dwFma(x.lo, y.hi, tl1);
  const float cl3 = c.lo + cl2;
  return dwFastTwoSum(c.hi, cl3);
}

// value rounded to float32 (two<float>::eval)
__host__ __device__ __forceinline__ float dwEval(DWord x) { return x.hi + x.lo; }

// value rounded to float64
__host__ __device__ __forceinline__ double dwToDouble(DWord x) {
  return (double)x.hi + (double)x.lo;
}

}  // namespace gph
