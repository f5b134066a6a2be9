// DCT-II / DCT-III of rows with nz = 2048 (half-length complex FFT of
// M = 1024 points): one wavefront per row, the FFT in registers; nz = 4096
// (M = 2048): two wavefronts per row, one 1024-point FFT each and a
// radix-2 join.
//
// Four-step split M = 16 x 64 (n = 64 n1 + n2, k = k1 + 16 k2):
//   pass 1  lane n2 holds x[64 n1 + n2], n1 < 16: DFT-16 over n1 in
//           registers, times W_1024^(n2 k1)             -> Y[k1][n2]
//   pass 2  lane (k1, m) holds Y[k1][m + 4c], c < 16: DFT-16 over c, times
//           W_64^(m k2a)                                 -> E[k2a][m][k1]
//   pass 3  lane (g, k1) holds E[4g + t][m][k1]: four DFT-4 over m, giving
//           X[k1 + 16 (4g + t) + 256 k2b]
// with k2 = k2a + 16 k2b.  Between passes the wave exchanges values through
// LDS; row strides are padded so each group of 16 lanes (one LDS cycle of
// 16-byte accesses) touches 16 distinct bank quads.  All index arithmetic
// is compile-time except the lane number.  The DCT pre/post steps around
// the FFT are the ones of axs_fft.hpp (Makhoul's reorder and twiddles).
#pragma once

#include "axs_fft.hpp"

namespace axs {

constexpr int kF1024M = 1024;
constexpr int kYStride = 68;                          // Y row (k1) stride, complex
constexpr int kEStrideM = 20, kEStrideK = 80;         // E[k2a][m][k1] strides
constexpr int kF1024Lds = 16 * kEStrideK;             // complex entries (>= M, >= 16*68)
static_assert(kF1024Lds >= kF1024M && kF1024Lds >= 16 * kYStride, "LDS buffer");
// M = R1 x 64 = R1 x 8 x 8, R1 = 4, 6, 8 (fft_r64_wave): Y[k1][n2] rows of
// 64 + 4, E[k2a][m][k1]
constexpr int kF512M = 512;
constexpr int kY512 = 68, kE5M = 8, kE5K = 72;
constexpr int kF512Lds = 576;
static_assert(kF512Lds >= 8 * kY512 && kF512Lds >= 7 * kE5K + 7 * kE5M + 8, "LDS buffer");
// LDS entries of the one-wave transform of M points
template <int M>
constexpr int kWaveLds = M == kF1024M ? kF1024Lds : kF512Lds;

template <typename T, int SGN>
__device__ __forceinline__ void dft4n(Cplx<T>& a0, Cplx<T>& a1, Cplx<T>& a2, Cplx<T>& a3) {
  const Cplx<T> y0 = cadd(a0, a2), y1 = csub(a0, a2), y2 = cadd(a1, a3), d = csub(a1, a3);
  const Cplx<T> y3 = {SGN < 0 ? d.y : -d.y, SGN < 0 ? -d.x : d.x};  // (SGN i) d
  a0 = cadd(y0, y2);
  a2 = csub(y0, y2);
  a1 = cadd(y1, y3);
  a3 = csub(y1, y3);
}

// a * W_16^E, W_16 = e^{SGN 2 pi i / 16}
template <typename T, int SGN, int E>
__device__ __forceinline__ Cplx<T> w16(Cplx<T> a) {
  constexpr int e = E & 15;
  if constexpr (e == 0) {
    return a;
  } else if constexpr (e == 4) {
    return SGN < 0 ? Cplx<T>{a.y, -a.x} : Cplx<T>{-a.y, a.x};
  } else if constexpr (e == 8) {
    return Cplx<T>{-a.x, -a.y};
  } else if constexpr (e == 12) {
    return SGN < 0 ? Cplx<T>{-a.y, a.x} : Cplx<T>{a.y, -a.x};
  } else {
    constexpr double C[16] = {1.0,
                              0.92387953251128675613,
                              0.70710678118654752440,
                              0.38268343236508977173,
                              0.0,
                              -0.38268343236508977173,
                              -0.70710678118654752440,
                              -0.92387953251128675613,
                              -1.0,
                              -0.92387953251128675613,
                              -0.70710678118654752440,
                              -0.38268343236508977173,
                              0.0,
                              0.38268343236508977173,
                              0.70710678118654752440,
                              0.92387953251128675613};
    constexpr double S[16] = {0.0,
                              0.38268343236508977173,
                              0.70710678118654752440,
                              0.92387953251128675613,
                              1.0,
                              0.92387953251128675613,
                              0.70710678118654752440,
                              0.38268343236508977173,
                              0.0,
                              -0.38268343236508977173,
                              -0.70710678118654752440,
                              -0.92387953251128675613,
                              -1.0,
                              -0.92387953251128675613,
                              -0.70710678118654752440,
                              -0.38268343236508977173};
    return cmul(a, Cplx<T>{T(C[e]), T(SGN * S[e])});
  }
}

// DFT-16 in place; on return x[4 ka + kb] holds X[ka + 4 kb].
template <typename T, int SGN>
__device__ __forceinline__ void dft16(Cplx<T> (&x)[16]) {
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) dft4n<T, SGN>(x[nb], x[4 + nb], x[8 + nb], x[12 + nb]);
  // x[4 ka + nb] = t[nb][ka]; twiddle by W_16^(nb ka)
  x[5] = w16<T, SGN, 1>(x[5]);
  x[6] = w16<T, SGN, 2>(x[6]);
  x[7] = w16<T, SGN, 3>(x[7]);
  x[9] = w16<T, SGN, 2>(x[9]);
  x[10] = w16<T, SGN, 4>(x[10]);
  x[11] = w16<T, SGN, 6>(x[11]);
  x[13] = w16<T, SGN, 3>(x[13]);
  x[14] = w16<T, SGN, 6>(x[14]);
  x[15] = w16<T, SGN, 9>(x[15]);
#pragma unroll
  for (int ka = 0; ka < 4; ++ka) dft4n<T, SGN>(x[4 * ka], x[4 * ka + 1], x[4 * ka + 2], x[4 * ka + 3]);
}

// DFT-8 in place, natural order: the DFT-4s of the even and odd samples,
// joined by W_8^k = W_16^(2k).
template <typename T, int SGN>
__device__ __forceinline__ void dft8(Cplx<T> (&x)[8]) {
  dft4n<T, SGN>(x[0], x[2], x[4], x[6]);
  dft4n<T, SGN>(x[1], x[3], x[5], x[7]);
  const Cplx<T> o0 = x[1], o1 = w16<T, SGN, 2>(x[3]), o2 = w16<T, SGN, 4>(x[5]),
                o3 = w16<T, SGN, 6>(x[7]);
  const Cplx<T> e0 = x[0], e1 = x[2], e2 = x[4], e3 = x[6];
  x[0] = cadd(e0, o0);
  x[4] = csub(e0, o0);
  x[1] = cadd(e1, o1);
  x[5] = csub(e1, o1);
  x[2] = cadd(e2, o2);
  x[6] = csub(e2, o2);
  x[3] = cadd(e3, o3);
  x[7] = csub(e3, o3);
}

// DFT-6 in place, natural order: the DFT-3s of the even and odd samples,
// joined by W_6^k.
template <typename T, int SGN>
__device__ __forceinline__ void dft6(Cplx<T> (&x)[6]) {
  Cplx<T> e[3] = {x[0], x[2], x[4]}, o[3] = {x[1], x[3], x[5]};
  dft_small<T, 3>(e, SGN);
  dft_small<T, 3>(o, SGN);
  constexpr double h = 0.86602540378443864676;  // sin(pi / 3)
  const Cplx<T> o1 = cmul(o[1], Cplx<T>{T(0.5), T(SGN * h)});
  const Cplx<T> o2 = cmul(o[2], Cplx<T>{T(-0.5), T(SGN * h)});
  x[0] = cadd(e[0], o[0]);
  x[3] = csub(e[0], o[0]);
  x[1] = cadd(e[1], o1);
  x[4] = csub(e[1], o1);
  x[2] = cadd(e[2], o2);
  x[5] = csub(e[2], o2);
}

// DFT-R in place, natural order (R = 2, 3, 4, 6, 8).
template <typename T, int SGN, int R>
__device__ __forceinline__ void dft_r(Cplx<T> (&x)[R]) {
  if constexpr (R == 8)
    dft8<T, SGN>(x);
  else if constexpr (R == 6)
    dft6<T, SGN>(x);
  else if constexpr (R == 4)
    dft4n<T, SGN>(x[0], x[1], x[2], x[3]);
  else
    dft_small<T, R>(x, SGN);
}

template <int SGN>
__device__ __forceinline__ Cplx<double> twid(const Cplx<double>* __restrict__ tw, int t) {
  const Cplx<double> w = tw[t];
  return SGN < 0 ? w : Cplx<double>{w.x, -w.y};
}

// FFT of the M = 1024 complex values in buf (natural order, lane-shared
// LDS); result in natural order in buf.  One wavefront (lane = threadIdx.x
// mod 64) per buffer; every wave of the block calls it (block barriers).
// tw[TS t] = e^{-2 pi i t / 1024} (TS = 2: the table of a 2048-point plan).
template <typename T, int SGN, int TS = 1>
__device__ __forceinline__ void fft1024_wave(Cplx<T>* buf, const Cplx<double>* __restrict__ tw) {
  const int l = threadIdx.x & 63;
  Cplx<T> v[16];
  // ---- pass 1: lane n2 = l
  Cplx<double> t1[15];
#pragma unroll
  for (int k1 = 1; k1 < 16; ++k1) t1[k1 - 1] = twid<SGN>(tw, TS * l * k1);
#pragma unroll
  for (int n1 = 0; n1 < 16; ++n1) v[n1] = buf[64 * n1 + l];
  dft16<T, SGN>(v);
  __syncthreads();
#pragma unroll
  for (int k1 = 0; k1 < 16; ++k1) {
    Cplx<T> b = v[4 * (k1 & 3) + (k1 >> 2)];
    if (k1 > 0) b = cmul(b, Cplx<T>{T(t1[k1 - 1].x), T(t1[k1 - 1].y)});
    buf[k1 * kYStride + l] = b;
  }
  __syncthreads();
  // ---- pass 2: lane (k1, m)
  {
    const int k1 = l >> 2, m = l & 3;
    Cplx<double> t2[15];
#pragma unroll
    for (int k2a = 1; k2a < 16; ++k2a) t2[k2a - 1] = twid<SGN>(tw, TS * 16 * m * k2a);
#pragma unroll
    for (int c = 0; c < 16; ++c) v[c] = buf[k1 * kYStride + m + 4 * c];
    dft16<T, SGN>(v);
    __syncthreads();
#pragma unroll
    for (int k2a = 0; k2a < 16; ++k2a) {
      Cplx<T> e = v[4 * (k2a & 3) + (k2a >> 2)];
      if (k2a > 0) e = cmul(e, Cplx<T>{T(t2[k2a - 1].x), T(t2[k2a - 1].y)});
      buf[k2a * kEStrideK + m * kEStrideM + k1] = e;
    }
  }
  __syncthreads();
  // ---- pass 3: lane (g, k1)
  {
    const int g = l >> 4, k1 = l & 15;
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int m = 0; m < 4; ++m) v[4 * t + m] = buf[(4 * g + t) * kEStrideK + m * kEStrideM + k1];
#pragma unroll
    for (int t = 0; t < 4; ++t) dft4n<T, SGN>(v[4 * t], v[4 * t + 1], v[4 * t + 2], v[4 * t + 3]);
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) buf[k1 + 16 * (4 * g + t) + 256 * kb] = v[4 * t + kb];
  }
  __syncthreads();
}

// FFT of the M = R1 x 64 complex values in buf (R1 = 2, 3, 4, 6, 8), one
// wavefront, three passes (n = n2 + 64 n1, k = k1 + R1 k', n2 = m + 8 c,
// k' = k2a + 8 k2b):
//   pass 1  lane n2: DFT-R1 over n1, times W_M^(n2 k1)          -> Y[k1][n2]
//   pass 2  lane (k1, m), k1 < R1: DFT-8 over c, times W_64^(m k2a)
//                                                               -> E[k2a][m][k1]
//   pass 3  lane (k2a, k1), k1 < R1: DFT-8 over m     -> X[k1 + R1 (k2a + 8 k2b)]
// (lanes with k1 >= R1 idle in passes 2 and 3).  tw[TS t] = e^{-2 pi i t / M}
// (TS = 2: the table of a plan of twice the length).  Natural order in and
// out of buf.
template <typename T, int SGN, int R1, int TS = 1>
__device__ __forceinline__ void fft_r64_wave(Cplx<T>* buf, const Cplx<double>* __restrict__ tw) {
  static_assert(R1 == 2 || R1 == 3 || R1 == 4 || R1 == 6 || R1 == 8, "first radix");
  const int l = threadIdx.x & 63;
  {
    Cplx<T> v[R1];
    Cplx<double> t1[R1 - 1];
#pragma unroll
    for (int k1 = 1; k1 < R1; ++k1) t1[k1 - 1] = twid<SGN>(tw, TS * l * k1);
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) v[n1] = buf[64 * n1 + l];
    dft_r<T, SGN, R1>(v);
    __syncthreads();
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      Cplx<T> b = v[k1];
      if (k1 > 0) b = cmul(b, Cplx<T>{T(t1[k1 - 1].x), T(t1[k1 - 1].y)});
      buf[k1 * kY512 + l] = b;
    }
  }
  __syncthreads();
  {
    const int k1 = l >> 3, m = l & 7;
    Cplx<T> v[8];
    if (k1 < R1) {
      Cplx<double> t2[7];
#pragma unroll
      for (int k2a = 1; k2a < 8; ++k2a) t2[k2a - 1] = twid<SGN>(tw, TS * R1 * m * k2a);
#pragma unroll
      for (int c = 0; c < 8; ++c) v[c] = buf[k1 * kY512 + m + 8 * c];
      dft8<T, SGN>(v);
#pragma unroll
      for (int k2a = 1; k2a < 8; ++k2a)
        v[k2a] = cmul(v[k2a], Cplx<T>{T(t2[k2a - 1].x), T(t2[k2a - 1].y)});
    }
    __syncthreads();
    if (k1 < R1) {
#pragma unroll
      for (int k2a = 0; k2a < 8; ++k2a) buf[k2a * kE5K + m * kE5M + k1] = v[k2a];
    }
  }
  __syncthreads();
  {
    const int k2a = l >> 3, k1 = l & 7;
    Cplx<T> v[8];
    if (k1 < R1) {
#pragma unroll
      for (int m = 0; m < 8; ++m) v[m] = buf[k2a * kE5K + m * kE5M + k1];
      dft8<T, SGN>(v);
    }
    __syncthreads();
    if (k1 < R1) {
#pragma unroll
      for (int k2b = 0; k2b < 8; ++k2b) buf[k1 + R1 * (k2a + 8 * k2b)] = v[k2b];
    }
  }
  __syncthreads();
}

// The one-wave transform of M = 128, 192, 256, 384, 512 or 1024 points
// (twiddle table stride TS as above).
template <typename T, int SGN, int M, int TS = 1>
__device__ __forceinline__ void fft_wave(Cplx<T>* buf, const Cplx<double>* __restrict__ tw) {
  if constexpr (M == kF1024M)
    fft1024_wave<T, SGN, TS>(buf, tw);
  else
    fft_r64_wave<T, SGN, M / 64, TS>(buf, tw);
}

}  // namespace axs

namespace axs {

// DCT-II post step (as dct2_post): lane l takes k = l + 64 q (q < M / 128) with
// its partner M - k, writing outputs k, N - k, M - k and N - (M - k); lane 0
// also takes k = M / 2, its own partner.  The twiddles are loaded before
// any is used.
template <typename T, int M>
__device__ __forceinline__ void dct2_post_pairs(const Cplx<T>* Z, const Cplx<double>* __restrict__ e1,
                                                const Cplx<double>* __restrict__ e2,
                                                T* __restrict__ out) {
  // pairs k < M / 2 over 64 lanes (the last round partial when 64 does
  // not divide M / 2)
  constexpr int N = 2 * M, KP = (M / 2 + 63) / 64;
  constexpr bool kFull = (M / 2) % 64 == 0;
  const int l = threadIdx.x;
  Cplx<double> a1[KP], b1[KP], a2[KP], b2[KP];
#pragma unroll
  for (int q = 0; q < KP; ++q) {
    const int k = l + 64 * q;
    if (kFull || k < M / 2) {
      a1[q] = e1[k];
      b1[q] = e1[M - k];
      a2[q] = e2[k];
      b2[q] = e2[M - k];
    }
  }
  const Cplx<double> ah = e1[M / 2], bh = e2[M / 2];
#pragma unroll
  for (int q = 0; q < KP; ++q) {
    const int k = l + 64 * q, kp = M - k;
    if (!kFull && k >= M / 2) break;
    const Cplx<T> zk = Z[k], zp = Z[kp == M ? 0 : kp];  // Z is periodic in M
    const Cplx<T> sk = cadd(zk, cconj(zp)), dk = csub(zk, cconj(zp));
    const Cplx<T> wdk = cmul(Cplx<T>{T(a1[q].x), T(a1[q].y)}, dk);
    const Cplx<T> Vk = {T(0.5) * (sk.x + wdk.y), T(0.5) * (sk.y - wdk.x)};
    const Cplx<T> wk = cmul(Cplx<T>{T(a2[q].x), T(a2[q].y)}, Vk);
    const Cplx<T> sp = cadd(zp, cconj(zk)), dp = csub(zp, cconj(zk));
    const Cplx<T> wdp = cmul(Cplx<T>{T(b1[q].x), T(b1[q].y)}, dp);
    const Cplx<T> Vp = {T(0.5) * (sp.x + wdp.y), T(0.5) * (sp.y - wdp.x)};
    const Cplx<T> wp = cmul(Cplx<T>{T(b2[q].x), T(b2[q].y)}, Vp);
    out[k] = wk.x;
    out[kp] = wp.x;
    if (k > 0) {
      out[N - k] = -wk.y;
      out[N - kp] = -wp.y;
    }
  }
  if (l == 0) {
    constexpr int k = M / 2;
    const Cplx<T> zk = Z[k];
    const Cplx<T> sk = cadd(zk, cconj(zk)), dk = csub(zk, cconj(zk));
    const Cplx<T> wdk = cmul(Cplx<T>{T(ah.x), T(ah.y)}, dk);
    const Cplx<T> Vk = {T(0.5) * (sk.x + wdk.y), T(0.5) * (sk.y - wdk.x)};
    const Cplx<T> wk = cmul(Cplx<T>{T(bh.x), T(bh.y)}, Vk);
    out[k] = wk.x;
    out[N - k] = -wk.y;
  }
}

}  // namespace axs

namespace axs {

// Two wavefronts per row, M = 2 HM (HM = 1024: nz = 4096; HM = 512:
// nz = 2048): the packed sequence z splits into its even and odd samples,
// each transformed by one wavefront with the one-wave HM-point FFT (the
// first buffer holds z_{2p}, the second z_{2p+1}), joined by one radix-2 step
//   Z_k = E_k + W_M^k O_k,   Z_{k+HM} = E_k - W_M^k O_k   (k < HM)
// written over E_k and O_k: Z_k for k < HM is at k of the first buffer, for
// k >= HM at k - HM of the second.  The inverse joins with the conjugate
// twiddle.
constexpr int kF2048M = 2048;

// Buffer entry of packed complex sample m (before the FFTs).
template <int HM>
__device__ __forceinline__ int split_in_slot(int m) {
  return ((m & 1) ? kWaveLds<HM> : 0) + (m >> 1);
}
// Buffer entry of transform value k (after the join).
template <int HM>
__device__ __forceinline__ int split_out_slot(int k) {
  return k < HM ? k : kWaveLds<HM> + (k - HM);
}

// The join over bufs (two buffers of kWaveLds<HM> entries), 128 threads,
// block barrier after.  tw[k] = e^{-2 pi i k / (2 HM)}.
template <typename T, int SGN, int HM>
__device__ __forceinline__ void split_join(Cplx<T>* bufs, const Cplx<double>* __restrict__ tw) {
  const int t = threadIdx.x;
  Cplx<double> w[HM / 128];
#pragma unroll
  for (int q = 0; q < HM / 128; ++q) w[q] = twid<SGN>(tw, t + 128 * q);
#pragma unroll
  for (int q = 0; q < HM / 128; ++q) {
    const int k = t + 128 * q;
    const Cplx<T> e = bufs[k];
    const Cplx<T> o = cmul(bufs[kWaveLds<HM> + k], Cplx<T>{T(w[q].x), T(w[q].y)});
    bufs[k] = cadd(e, o);
    bufs[kWaveLds<HM> + k] = csub(e, o);
  }
  __syncthreads();
}

// DCT-II post step for M = 2 HM (as dct2_post_pairs), 128 threads: thread t
// takes k = t + 128 q (q < HM / 128) with its partner M - k; thread 0 also
// takes k = M / 2.
template <typename T, int HM>
__device__ __forceinline__ void dct2_post_pairs_split(const Cplx<T>* bufs,
                                                      const Cplx<double>* __restrict__ e1,
                                                      const Cplx<double>* __restrict__ e2,
                                                      T* __restrict__ out) {
  constexpr int M = 2 * HM, N = 2 * M, KP = M / 2 / 128;
  const int t = threadIdx.x;
  Cplx<double> a1[KP], b1[KP], a2[KP], b2[KP];
#pragma unroll
  for (int q = 0; q < KP; ++q) {
    const int k = t + 128 * q;
    a1[q] = e1[k];
    b1[q] = e1[M - k];
    a2[q] = e2[k];
    b2[q] = e2[M - k];
  }
  const Cplx<double> ah = e1[M / 2], bh = e2[M / 2];
#pragma unroll
  for (int q = 0; q < KP; ++q) {
    const int k = t + 128 * q, kp = M - k;
    const Cplx<T> zk = bufs[split_out_slot<HM>(k)], zp = bufs[split_out_slot<HM>(kp & (M - 1))];
    const Cplx<T> sk = cadd(zk, cconj(zp)), dk = csub(zk, cconj(zp));
    const Cplx<T> wdk = cmul(Cplx<T>{T(a1[q].x), T(a1[q].y)}, dk);
    const Cplx<T> Vk = {T(0.5) * (sk.x + wdk.y), T(0.5) * (sk.y - wdk.x)};
    const Cplx<T> wk = cmul(Cplx<T>{T(a2[q].x), T(a2[q].y)}, Vk);
    const Cplx<T> sp = cadd(zp, cconj(zk)), dp = csub(zp, cconj(zk));
    const Cplx<T> wdp = cmul(Cplx<T>{T(b1[q].x), T(b1[q].y)}, dp);
    const Cplx<T> Vp = {T(0.5) * (sp.x + wdp.y), T(0.5) * (sp.y - wdp.x)};
    const Cplx<T> wp = cmul(Cplx<T>{T(b2[q].x), T(b2[q].y)}, Vp);
    out[k] = wk.x;
    out[kp] = wp.x;
    if (k > 0) {
      out[N - k] = -wk.y;
      out[N - kp] = -wp.y;
    }
  }
  if (t == 0) {
    constexpr int k = M / 2;
    const Cplx<T> zk = bufs[split_out_slot<HM>(k)];
    const Cplx<T> sk = cadd(zk, cconj(zk)), dk = csub(zk, cconj(zk));
    const Cplx<T> wdk = cmul(Cplx<T>{T(ah.x), T(ah.y)}, dk);
    const Cplx<T> Vk = {T(0.5) * (sk.x + wdk.y), T(0.5) * (sk.y - wdk.x)};
    const Cplx<T> wk = cmul(Cplx<T>{T(bh.x), T(bh.y)}, Vk);
    out[k] = wk.x;
    out[N - k] = -wk.y;
  }
}

}  // namespace axs
