// Register-resident 256-point FFT (16 x 16) shared by the plan passes and
// the general solver's grid operator (gfx950).
//
// A 256-point line is owned by 16 lanes: lane t holds elements t + 16 j in
// registers; a radix-16 DFT in registers (4 x 4 split), one twiddle multiply,
// one transpose through LDS and a second radix-16 DFT leave lane t holding
// outputs t + 16 n.  Unnormalised transforms: forward exp(-2 pi i jk / 256),
// inverse exp(+2 pi i jk / 256).
#ifndef FBME_FFT16_H
#define FBME_FFT16_H

#include "fbme_device.h"

namespace fbme {

// v * (c + i s)
__device__ __forceinline__ float2 cmulc(float2 v, float c, float s) {
  return make_float2(v.x * c - v.y * s, v.x * s + v.y * c);
}

// W16^e = exp(-+2 pi i e / 16) for the exponents e = n1 k1 of the 4 x 4 split.
template <bool INV, int E>
__device__ __forceinline__ float2 tw16(float2 v) {
  constexpr float c8 = 0.92387953251128675613f, s8 = 0.38268343236508977173f;
  constexpr float h = 0.70710678118654752440f;
  constexpr float sg = INV ? 1.0f : -1.0f;   // sign of the imaginary part
  if constexpr (E == 0) return v;
  else if constexpr (E == 1) return cmulc(v, c8, sg * s8);
  else if constexpr (E == 2) return cmulc(v, h, sg * h);
  else if constexpr (E == 3) return cmulc(v, s8, sg * c8);
  else if constexpr (E == 4) return rot90<INV>(v);
  else if constexpr (E == 6) return cmulc(v, -h, sg * h);
  else return cmulc(v, -c8, -sg * s8);       // E == 9
}

// In-place natural-order 16-point DFT of v (4 x 4 Cooley-Tukey).
template <bool INV>
__device__ __forceinline__ void dft16(float2 (&v)[16]) {
#pragma unroll
  for (int n1 = 0; n1 < 4; ++n1) {
    float2 q[4] = {v[n1], v[n1 + 4], v[n1 + 8], v[n1 + 12]};
    bfly4<INV>(q);
    v[n1] = q[0];
    v[n1 + 4] = q[1];
    v[n1 + 8] = q[2];
    v[n1 + 12] = q[3];
  }
  // v[n1 + 4 k1] *= W16^(n1 k1)
  v[5] = tw16<INV, 1>(v[5]);
  v[9] = tw16<INV, 2>(v[9]);
  v[13] = tw16<INV, 3>(v[13]);
  v[6] = tw16<INV, 2>(v[6]);
  v[10] = tw16<INV, 4>(v[10]);
  v[14] = tw16<INV, 6>(v[14]);
  v[7] = tw16<INV, 3>(v[7]);
  v[11] = tw16<INV, 6>(v[11]);
  v[15] = tw16<INV, 9>(v[15]);
  float2 out[16];
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    float2 q[4] = {v[4 * k1], v[4 * k1 + 1], v[4 * k1 + 2], v[4 * k1 + 3]};
    bfly4<INV>(q);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) out[k1 + 4 * k2] = q[k2];
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) v[i] = out[i];
}

// 256-point transform of the line owned by the 16 lanes of a group.  On entry
// v[j] = a[t + 16 j]; on exit v[n] = b[t + 16 n].  tw: LDS table
// exp(-2 pi i k / 256); x: this group's LDS transpose buffer, element i at
// x[i * XSTRIDE] (groups interleaved with stride XSTRIDE keep the 16 lanes
// of one LDS pass on consecutive slots); every lane of the workgroup must
// call (the barriers are workgroup-wide).
template <bool INV, int XSTRIDE = 1>
__device__ __forceinline__ void fft256(float2 (&v)[16], const float2* tw, float2* x, int t) {
  dft16<INV>(v);
#pragma unroll
  for (int n = 1; n < 16; ++n) {
    float2 ww = tw[t * n];
    if (INV) ww.y = -ww.y;
    v[n] = cmul(v[n], ww);
  }
#pragma unroll
  for (int n = 0; n < 16; ++n) x[(n * 17 + t) * XSTRIDE] = v[n];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 16; ++j) v[j] = x[(t * 17 + j) * XSTRIDE];
  __syncthreads();
  dft16<INV>(v);
}

// Copy the 256-entry twiddle table to LDS; the caller synchronises before use.
__device__ __forceinline__ void load_tw256(float2* tw, const float2* __restrict__ table) {
  for (int i = threadIdx.x; i < 256; i += blockDim.x) tw[i] = table[i];
}

}  // namespace fbme

#endif
