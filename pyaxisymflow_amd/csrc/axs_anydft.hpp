// DFT of any length N on one row in LDS, the building block of the row
// transforms for grids whose nz the half-length register and Stockham
// transforms do not cover (odd nz, or nz / 2 with a prime factor above 5).
//
// N that factors into 2, 3 and 5: one Stockham FFT of length N (fft_lds).
// Otherwise Bluestein's chirp-z form (L. Bluestein, "A linear filtering
// approach to the computation of discrete Fourier transform", IEEE Trans.
// Audio Electroacoust. 18(4), 1970): with w_n = e^{-i pi n^2 / N},
//   X_k = sum_n v_n e^{-2 pi i n k / N} = w_k sum_n (v_n w_n) conj(w_{k-n}),
// a circular convolution of length L >= 2N - 1 (the smallest 2,3,5-smooth
// such L) computed by two length-L FFTs and a pointwise product with
// bhat = FFT_L(b) / L, b_m = conj(w_m) for |m| < N (m taken mod L), which
// the plan computes once on the device.  The chirp angle is reduced as
// pi (n^2 mod 2N) / N on the host so that it stays exact for large n.  The
// inverse direction (exponent +1) uses conj(w) and conj(bhat_{-k}).
#pragma once

#include "axs_fft.hpp"

#include <vector>

namespace axs {

struct AnyDftArgs {
  int N = 0;          // DFT length
  int L = 0;          // FFT length: N (direct) or the convolution length
  int bluestein = 0;  // 1: chirp-z form
  FftPlanArgs fp{};   // radix plan of the length-L FFT
  const Cplx<double>* tw = nullptr;     // e^{-2 pi i t / L}, t < L
  const Cplx<double>* chirp = nullptr;  // w_n, n < N (Bluestein)
  const Cplx<double>* bhat = nullptr;   // FFT_L(b) / L (Bluestein)
  // 1: the FFT reads tw from a copy in LDS after the L work entries
  // (any_dft_stage), for rows short enough that two workgroups still share
  // a CU
  int stage = 0;
};

// Host owner of one length's device tables.
struct AnyDft {
  AnyDftArgs args{};
  // the length-L FFT's launch shape: nth threads, PER butterflies each
  int nth = 64, per = 1;
  std::vector<void*> owned;  // device tables
  hipError_t init(int N);
  template <typename T>
  size_t lds_bytes() const { return sizeof(Cplx<T>) * (size_t)args.L * (args.stage ? 2 : 1); }
  ~AnyDft();
};

// Butterflies of the plan's widest Stockham stage: L / (its smallest radix).
// A workgroup of nth threads with PER butterflies each covers the length-L
// FFT only if PER * nth reaches this.
inline int widest_stage(const AnyDftArgs& a) {
  int rmin = a.L > 1 ? a.L : 2;
  for (int s = 0; s < a.fp.nstages; ++s) rmin = a.fp.radix[s] < rmin ? a.fp.radix[s] : rmin;
  return a.L / rmin;
}

// Largest L one workgroup can hold in LDS (FP64: 8192 x 16 B = 128 KiB of
// the CU's 160 KiB); every length N with L <= kAnyMaxL is supported.
constexpr int kAnyMaxL = 8192;
// FFT length the chirp-z form needs for N (N itself when N is 2,3,5-smooth).
int any_dft_length(int N, bool* bluestein);

// Copies the FFT twiddles into LDS after the work entries when the plan
// stages them; call before the barrier that precedes any_dft.
template <typename T>
__device__ __forceinline__ void any_dft_stage(Cplx<T>* work, const AnyDftArgs& a, int tid,
                                              int nth) {
  if (a.stage) stage_twiddles(work + a.L, a.tw, a.L, tid, nth);
}

template <typename T, int PER>
__device__ __forceinline__ void any_fft(Cplx<T>* work, const AnyDftArgs& a, int sgn, int tid,
                                        int nth) {
  if (a.stage)
    fft_lds<T, PER>(work, a.fp, static_cast<const Cplx<T>*>(work + a.L), sgn, tid, nth);
  else
    fft_lds<T, PER>(work, a.fp, a.tw, sgn, tid, nth);
}

// In-place DFT of work[0..N) (exponent sign sgn), result in work[0..N),
// by threads tid of nth (all of them call it: there are block barriers).
// work holds L entries (then the staged twiddles); the caller has written
// work[0..N), called any_dft_stage and issued a barrier.  Ends with a
// barrier.
template <typename T, int PER>
__device__ void any_dft(Cplx<T>* work, const AnyDftArgs& a, int sgn, int tid, int nth) {
  if (!a.bluestein) {
    any_fft<T, PER>(work, a, sgn, tid, nth);
    return;
  }
  const int N = a.N, L = a.L;
  for (int n = tid; n < L; n += nth) {
    if (n < N) {
      const Cplx<double> w = a.chirp[n];
      work[n] = cmul(work[n], Cplx<T>{T(w.x), T(sgn < 0 ? w.y : -w.y)});
    } else {
      work[n] = Cplx<T>{T(0), T(0)};
    }
  }
  __syncthreads();
  any_fft<T, PER>(work, a, -1, tid, nth);
  for (int k = tid; k < L; k += nth) {
    Cplx<double> b;
    if (sgn < 0) {
      b = a.bhat[k];
    } else {
      const Cplx<double> m = a.bhat[k == 0 ? 0 : L - k];
      b = Cplx<double>{m.x, -m.y};
    }
    work[k] = cmul(work[k], Cplx<T>{T(b.x), T(b.y)});
  }
  __syncthreads();
  any_fft<T, PER>(work, a, +1, tid, nth);
  for (int k = tid; k < N; k += nth) {
    const Cplx<double> w = a.chirp[k];
    work[k] = cmul(work[k], Cplx<T>{T(w.x), T(sgn < 0 ? w.y : -w.y)});
  }
  __syncthreads();
}

}  // namespace axs
