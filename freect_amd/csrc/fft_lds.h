// fft_lds.h -- one butterfly of a Stockham autosort FFT (forward sign).
//
// Stockham's formulation as given by Govindaraju et al., "High Performance
// Discrete Fourier Transforms on Graphics Processors" (SC'08): pass with
// radix R and current span Ns reads v[r] = in[j + r*N/R], multiplies by
// exp(-2 pi i r (j mod Ns) / (Ns R)), applies the R-point DFT and writes
// out[(j / Ns) * Ns * R + (j mod Ns) + r * Ns].  Passes run radix 4 while at
// least two factors of two remain, then one radix-2 pass.  tw[m] holds
// exp(-2 pi i m / N) (built in FP64 on the host).
//
// The same function is used by the kernel (one call per butterfly index j,
// spread over the workgroup) and by host-side checks of the index algebra.
#pragma once

#include "wfbp_math.h"

struct FctCplx {
  float x, y;
};

FCT_HD FctCplx fct_cmul(FctCplx a, FctCplx b) {
  FctCplx r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}

// Radix-4 butterfly j of the pass with span Ns (N = 4 * nb).
FCT_HD void fct_fft_r4(const FctCplx* in, FctCplx* out, const FctCplx* tw, int N, int Ns,
                       int j) {
  const int nb = N >> 2;
  const int k = j & (Ns - 1);
  const int ts = k * (N / (4 * Ns));
  FctCplx v0 = in[j], v1 = in[j + nb], v2 = in[j + 2 * nb], v3 = in[j + 3 * nb];
  if (Ns > 1) {
    v1 = fct_cmul(v1, tw[ts]);
    v2 = fct_cmul(v2, tw[2 * ts]);
    v3 = fct_cmul(v3, tw[3 * ts]);
  }
  const FctCplx s02 = {v0.x + v2.x, v0.y + v2.y};
  const FctCplx d02 = {v0.x - v2.x, v0.y - v2.y};
  const FctCplx s13 = {v1.x + v3.x, v1.y + v3.y};
  const FctCplx d13 = {v1.x - v3.x, v1.y - v3.y};
  const int o = (j / Ns) * Ns * 4 + k;
  out[o] = FctCplx{s02.x + s13.x, s02.y + s13.y};
  out[o + Ns] = FctCplx{d02.x + d13.y, d02.y - d13.x};       // d02 - i d13
  out[o + 2 * Ns] = FctCplx{s02.x - s13.x, s02.y - s13.y};
  out[o + 3 * Ns] = FctCplx{d02.x - d13.y, d02.y + d13.x};   // d02 + i d13
}

// Radix-2 butterfly j of the pass with span Ns (N = 2 * nb).
FCT_HD void fct_fft_r2(const FctCplx* in, FctCplx* out, const FctCplx* tw, int N, int Ns,
                       int j) {
  const int nb = N >> 1;
  const int k = j & (Ns - 1);
  const int ts = k * (N / (2 * Ns));
  const FctCplx v0 = in[j];
  FctCplx v1 = in[j + nb];
  if (Ns > 1) v1 = fct_cmul(v1, tw[ts]);
  const int o = (j / Ns) * Ns * 2 + k;
  out[o] = FctCplx{v0.x + v1.x, v0.y + v1.y};
  out[o + Ns] = FctCplx{v0.x - v1.x, v0.y - v1.y};
}
