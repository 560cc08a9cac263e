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

// Each part as one product and one fused multiply-add (the kernels build
// without FP contraction, which would otherwise give 4 products and 2 sums).
FCT_HD FctCplx fct_cmul(FctCplx a, FctCplx b) {
  FctCplx r;
  r.x = fmaf(a.x, b.x, -(a.y * b.y));
  r.y = fmaf(a.x, b.y, a.y * b.x);
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

// ---------------------------------------------------------------------------
// Register-resident Stockham FFT for N in {512, 1024, 2048, 4096}.
//
// T = N / 8 threads each hold 8 values.  Passes 0..2 have radix 8, a fourth
// pass has radix N / 512 (2, 4 or 8; none for N = 512).  In the pass with
// radix R and span Ns, thread t runs the butterflies j = t + u * T
// (u < 8 / R); register v[u * R + r] holds in[j + r * N / R] on entry and the
// output destined for out[(j / Ns) * Ns * R + (j mod Ns) + r * Ns] on exit --
// the same Stockham algebra as fct_fft_r4 above, with R values per butterfly
// kept in registers and only the exchanges between passes going through LDS.
// In the last pass Ns = N / R, so thread t ends holding the outputs
// t + m * T (m = u + r * 8 / R), which are exactly the inputs thread t needs
// for the first pass of a following transform: a pointwise product between
// two transforms needs no exchange.
// ---------------------------------------------------------------------------
template <int N>
struct FctFftShape {
  static constexpr int T = N / 8;
  static constexpr int kLastRadix = N / 512;
  static constexpr int kPasses = kLastRadix > 1 ? 4 : 3;
  static constexpr int radix(int p) { return p < 3 ? 8 : kLastRadix; }
  static constexpr int span(int p) { return p == 0 ? 1 : p == 1 ? 8 : p == 2 ? 64 : 512; }
};

// LDS slot of element i: one pad slot per 8 elements keeps the strided
// stores of the early passes free of bank conflicts.
FCT_HD int fct_fft_pad(int i) { return i + (i >> 3); }

FCT_HD FctCplx fct_cadd(FctCplx a, FctCplx b) { return FctCplx{a.x + b.x, a.y + b.y}; }
FCT_HD FctCplx fct_csub(FctCplx a, FctCplx b) { return FctCplx{a.x - b.x, a.y - b.y}; }

// In-place forward DFT of 4 values (exp(-2 pi i n k / 4)).
FCT_HD void fct_dft4(FctCplx& a0, FctCplx& a1, FctCplx& a2, FctCplx& a3) {
  const FctCplx s02 = fct_cadd(a0, a2), d02 = fct_csub(a0, a2);
  const FctCplx s13 = fct_cadd(a1, a3), d13 = fct_csub(a1, a3);
  a0 = fct_cadd(s02, s13);
  a2 = fct_csub(s02, s13);
  a1 = FctCplx{d02.x + d13.y, d02.y - d13.x};   // d02 - i d13
  a3 = FctCplx{d02.x - d13.y, d02.y + d13.x};   // d02 + i d13
}

// In-place forward DFT of R values v[0..R-1]: X[k] = A[k mod 4] + W8^k B[k mod 4]
// with A, B the 4-point DFTs of the even and odd inputs (R = 8).
template <int R>
FCT_HD void fct_dft(FctCplx* v) {
  if (R == 2) {
    const FctCplx a = v[0], b = v[1];
    v[0] = fct_cadd(a, b);
    v[1] = fct_csub(a, b);
  } else if (R == 4) {
    fct_dft4(v[0], v[1], v[2], v[3]);
  } else {
    FctCplx a0 = v[0], a1 = v[2], a2 = v[4], a3 = v[6];
    FctCplx b0 = v[1], b1 = v[3], b2 = v[5], b3 = v[7];
    fct_dft4(a0, a1, a2, a3);
    fct_dft4(b0, b1, b2, b3);
    const float h = 0.70710678118654752f;
    b1 = FctCplx{h * (b1.x + b1.y), h * (b1.y - b1.x)};    // * exp(-i pi / 4)
    b2 = FctCplx{b2.y, -b2.x};                             // * -i
    b3 = FctCplx{h * (b3.y - b3.x), -h * (b3.x + b3.y)};   // * exp(-3 i pi / 4)
    v[0] = fct_cadd(a0, b0);
    v[4] = fct_csub(a0, b0);
    v[1] = fct_cadd(a1, b1);
    v[5] = fct_csub(a1, b1);
    v[2] = fct_cadd(a2, b2);
    v[6] = fct_csub(a2, b2);
    v[3] = fct_cadd(a3, b3);
    v[7] = fct_csub(a3, b3);
  }
}

// Index of the tw[] entry (exp(-2 pi i m / N)) applied to value r of
// butterfly j in the pass with radix R and span Ns.
template <int N>
FCT_HD int fct_fft_tw_index(int j, int r, int R, int Ns) {
  return r * (j & (Ns - 1)) * (N / (Ns * R));
}
// Input / output element indices of register u * R + r of thread t.
template <int N>
FCT_HD int fct_fft_in_index(int t, int u, int r, int R) {
  return t + u * FctFftShape<N>::T + r * (N / R);
}
template <int N>
FCT_HD int fct_fft_out_index(int t, int u, int r, int R, int Ns) {
  const int j = t + u * FctFftShape<N>::T;
  return (j / Ns) * Ns * R + (j & (Ns - 1)) + r * Ns;
}

// Twiddles, DFT and (unless last) the LDS store of pass P for thread t.
// twp[u * R + r] holds tw[fct_fft_tw_index(j, r, R, Ns)] (unused for r = 0).
template <int N, int P>
FCT_HD void fct_fft_pass_compute(FctCplx* v, const FctCplx* twp) {
  constexpr int R = FctFftShape<N>::radix(P), Ns = FctFftShape<N>::span(P);
#pragma unroll
  for (int u = 0; u < 8 / R; ++u) {
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) v[u * R + r] = fct_cmul(v[u * R + r], twp[u * R + r]);
    }
    fct_dft<R>(v + u * R);
  }
}
template <int N, int P>
FCT_HD void fct_fft_pass_store(const FctCplx* v, FctCplx* lds, int t) {
  constexpr int R = FctFftShape<N>::radix(P), Ns = FctFftShape<N>::span(P);
#pragma unroll
  for (int u = 0; u < 8 / R; ++u)
#pragma unroll
    for (int r = 0; r < R; ++r)
      lds[fct_fft_pad(fct_fft_out_index<N>(t, u, r, R, Ns))] = v[u * R + r];
}
template <int N, int P>
FCT_HD void fct_fft_pass_load(FctCplx* v, const FctCplx* lds, int t) {
  constexpr int R = FctFftShape<N>::radix(P);
#pragma unroll
  for (int u = 0; u < 8 / R; ++u)
#pragma unroll
    for (int r = 0; r < R; ++r) v[u * R + r] = lds[fct_fft_pad(fct_fft_in_index<N>(t, u, r, R))];
}
// After the last pass, register u * R + r holds output t + m * T with
// m = u + r * 8 / R; this returns that register for m.
template <int N>
FCT_HD constexpr int fct_fft_final_reg(int m) {
  constexpr int R = FctFftShape<N>::kLastRadix > 1 ? FctFftShape<N>::kLastRadix : 8;
  return (m % (8 / R)) * R + m / (8 / R);
}
