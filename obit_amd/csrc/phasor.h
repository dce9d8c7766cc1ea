// phasor.h -- FP32 sin/cos of angles given in turns, shared by the DFT
// kernels (dft.hip, beam.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace obit {

// Hardware transcendental: v_sin_f32 / v_cos_f32 take revolutions.
__device__ __forceinline__ void sincos_turns_hw(float t, float *s, float *c) {
  *s = __builtin_amdgcn_sinf(t);
  *c = __builtin_amdgcn_cosf(t);
}

// Quadrant reduction t = q/4 + r (exact for |t| <= 1/2), then Taylor
// series of sin and cos in x = 2 pi r, |x| <= pi/4, truncated after x^9
// and x^10 (truncation error < 2e-9), and a quarter-turn rotation.
__device__ __forceinline__ void sincos_turns_poly(float t, float *s, float *c) {
  const float q = rintf(4.0f * t);
  const float r = fmaf(-0.25f, q, t);
  const int iq = ((int)q) & 3;
  const float x = r * 6.28318530717958647692f;
  const float x2 = x * x;
  float ps = fmaf(x2, 2.75573192e-6f, -1.98412698e-4f);
  ps = fmaf(x2, ps, 8.33333333e-3f);
  ps = fmaf(x2, ps, -1.66666667e-1f);
  const float sn = fmaf(x * x2, ps, x);
  float pc = fmaf(x2, -2.75573192e-7f, 2.48015873e-5f);
  pc = fmaf(x2, pc, -1.38888889e-3f);
  pc = fmaf(x2, pc, 4.16666667e-2f);
  pc = fmaf(x2, pc, -0.5f);
  const float cs = fmaf(x2, pc, 1.0f);
  // rotate (cs, sn) by iq quarter turns
  const bool odd = iq & 1;
  const float a = odd ? cs : sn;   // sin of rotated angle before sign
  const float b = odd ? sn : cs;   // cos of rotated angle before sign
  *s = (iq == 2 || iq == 3) ? -a : a;
  *c = (iq == 1 || iq == 2) ? -b : b;
}

// sin and cos of pi t for |t| <= 1/4 (an angle within pi/4): the Taylor
// series of sincos_turns_poly without its quadrant reduction.
__device__ __forceinline__ void sincos_halfturns_small(float t, float *s, float *c) {
  const float x = t * 3.14159265358979323846f;
  const float x2 = x * x;
  float ps = fmaf(x2, 2.75573192e-6f, -1.98412698e-4f);
  ps = fmaf(x2, ps, 8.33333333e-3f);
  ps = fmaf(x2, ps, -1.66666667e-1f);
  *s = fmaf(x * x2, ps, x);
  float pc = fmaf(x2, -2.75573192e-7f, 2.48015873e-5f);
  pc = fmaf(x2, pc, -1.38888889e-3f);
  pc = fmaf(x2, pc, 4.16666667e-2f);
  pc = fmaf(x2, pc, -0.5f);
  *c = fmaf(x2, pc, 1.0f);
}

}  // namespace obit
