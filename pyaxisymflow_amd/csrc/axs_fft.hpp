// In-LDS complex FFT (Stockham auto-sort, radices 4/2/3/5) and the real
// DCT-II / DCT-III row transforms built on it.
//
// DCT-II of a real row x[0..N) (N even), X_k = sum_n x_n cos(pi k (2n+1) / (2N)):
//   v_n = x_{2n}, v_{N-1-n} = x_{2n+1}                       (reorder)
//   z_m = v_{2m} + i v_{2m+1}, Z = FFT_M(z), M = N/2           (half-length FFT)
//   V_k = (Z_k + Z*_{M-k})/2 - i e^{-2 pi i k/N} (Z_k - Z*_{M-k})/2
//   X_k = Re(e^{-i pi k/(2N)} V_k),  X_{N-k} = -Im(e^{-i pi k/(2N)} V_k)
// The inverse runs the same steps backwards and divides by N, so
// inverse(forward(x)) == x.
#pragma once

#include "axs_common.hpp"

namespace axs {

constexpr int kMaxFftStages = 24;

struct FftPlanArgs {
  int M;                       // complex FFT length (N/2)
  int nstages;
  int radix[kMaxFftStages];
};

template <typename T>
struct Cplx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ Cplx<T> cmul(Cplx<T> a, Cplx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ Cplx<T> cadd(Cplx<T> a, Cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ Cplx<T> csub(Cplx<T> a, Cplx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ Cplx<T> cconj(Cplx<T> a) { return {a.x, -a.y}; }

// In-place DFT of size R on a[0..R), exponent sign `sgn` (-1 forward).
template <typename T, int R>
__device__ __forceinline__ void dft_small(Cplx<T>* a, int sgn) {
  if constexpr (R == 2) {
    Cplx<T> t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
  } else if constexpr (R == 4) {
    Cplx<T> t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
    Cplx<T> t2 = cadd(a[1], a[3]), t3 = csub(a[1], a[3]);
    // w * t3 with w = sgn * i
    Cplx<T> wt3 = {-T(sgn) * t3.y, T(sgn) * t3.x};
    a[0] = cadd(t0, t2);
    a[2] = csub(t0, t2);
    a[1] = cadd(t1, wt3);
    a[3] = csub(t1, wt3);
  } else if constexpr (R == 3) {
    const T s3 = T(0.86602540378443864676) * T(sgn);
    Cplx<T> t = cadd(a[1], a[2]), d = csub(a[1], a[2]);
    Cplx<T> m = {a[0].x - T(0.5) * t.x, a[0].y - T(0.5) * t.y};
    Cplx<T> jd = {-s3 * d.y, s3 * d.x};
    a[0] = cadd(a[0], t);
    a[1] = cadd(m, jd);
    a[2] = csub(m, jd);
  } else {  // R == 5: direct
    const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
    const T s1 = T(0.95105651629515357212) * T(sgn), s2 = T(0.58778525229247312917) * T(sgn);
    Cplx<T> b1 = cadd(a[1], a[4]), d1 = csub(a[1], a[4]);
    Cplx<T> b2 = cadd(a[2], a[3]), d2 = csub(a[2], a[3]);
    Cplx<T> r1 = {a[0].x + c1 * b1.x + c2 * b2.x, a[0].y + c1 * b1.y + c2 * b2.y};
    Cplx<T> r2 = {a[0].x + c2 * b1.x + c1 * b2.x, a[0].y + c2 * b1.y + c1 * b2.y};
    // i * (s1 d1 + s2 d2) and i * (s2 d1 - s1 d2)
    Cplx<T> i1 = {-(s1 * d1.y + s2 * d2.y), s1 * d1.x + s2 * d2.x};
    Cplx<T> i2 = {-(s2 * d1.y - s1 * d2.y), s2 * d1.x - s1 * d2.x};
    a[0] = {a[0].x + b1.x + b2.x, a[0].y + b1.y + b2.y};
    a[1] = cadd(r1, i1);
    a[4] = csub(r1, i1);
    a[2] = cadd(r2, i2);
    a[3] = csub(r2, i2);
  }
}

// Copies the plan's twiddle table tw[0..M) into LDS at tl, converted to T,
// by threads tid of nth; a block barrier must follow before the stages read
// it.  (The stages' twiddle reads then wait on LDS, not on a global load
// that cannot be issued before the stage's barrier.  This pays for short
// FP32 rows; for long FP64 rows the strided 16-byte reads compete with the
// row's own LDS traffic and the global table is faster.)  Loads go out two
// at a time before their stores: one load latency when M <= 2 nth, and few
// registers held.
template <typename T>
__device__ __forceinline__ void stage_twiddles(Cplx<T>* tl, const Cplx<double>* __restrict__ tw,
                                               int M, int tid, int nth) {
  constexpr int U = 2;
  for (int base = tid; base < M; base += U * nth) {
    Cplx<double> w[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int t = base + u * nth;
      if (t < M) w[u] = tw[t];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int t = base + u * nth;
      if (t < M) tl[t] = Cplx<T>{T(w[u].x), T(w[u].y)};
    }
  }
}

// One Stockham stage of radix R over buf[0..M) in LDS (in place: all reads
// go to registers before a barrier, then all writes).  Each thread handles
// PER butterflies (compile-time, so the staging array stays in registers).
// tw[t] = e^{-2 pi i t/M}: the plan's FP64 table in global memory
// (W = Cplx<double>) or its copy in LDS (W = Cplx<T>, stage_twiddles); both
// give the same bits.  The inverse (sgn = +1) uses its conjugate.
template <typename T, int R, int PER, typename W>
__device__ __forceinline__ void stockham_stage(Cplx<T>* buf, int M, int Ns, const W* tw, int sgn,
                                               int tid, int nth) {
  const int nb = M / R;
  const int tstride = M / (Ns * R);
  // butterfly j: jm = j mod Ns, jq = j div Ns (a shift and a mask when Ns is
  // a power of two, as it is for every stage when the radix-3 and -5
  // stages come last)
  int jm[PER], jq[PER];
  if ((Ns & (Ns - 1)) == 0) {
    const int sh = __builtin_ctz(Ns);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int j = tid + q * nth;
      jm[q] = j & (Ns - 1);
      jq[q] = j >> sh;
    }
  } else {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const unsigned j = tid + q * nth;
      jq[q] = (int)(j / (unsigned)Ns);
      jm[q] = (int)j - jq[q] * Ns;
    }
  }
  Cplx<T> v[PER][R];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int j = tid + q * nth;
    if (j < nb) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        Cplx<T> a = buf[j + r * nb];
        if (r > 0 && jm[q] > 0) {
          const W w = tw[jm[q] * r * tstride];
          const Cplx<T> wt = {T(w.x), T(sgn < 0 ? w.y : -w.y)};
          a = cmul(a, wt);
        }
        v[q][r] = a;
      }
      dft_small<T, R>(v[q], sgn);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int j = tid + q * nth;
    if (j < nb) {
      const int idxD = jq[q] * Ns * R + jm[q];
#pragma unroll
      for (int r = 0; r < R; ++r) buf[idxD + r * Ns] = v[q][r];
    }
  }
  __syncthreads();
}

// Full FFT of length p.M in LDS by nth threads (this one: tid), twiddles tw
// as for stockham_stage.  Requires p.M / min(radix) <= PER * nth.  Every
// thread of the block must call it (stages end in block barriers).
template <typename T, int PER, typename W>
__device__ void fft_lds(Cplx<T>* buf, const FftPlanArgs& p, const W* tw, int sgn, int tid,
                        int nth) {
  int Ns = 1;
  for (int s = 0; s < p.nstages; ++s) {
    const int R = p.radix[s];
    if (R == 4)
      stockham_stage<T, 4, PER, W>(buf, p.M, Ns, tw, sgn, tid, nth);
    else if (R == 2)
      stockham_stage<T, 2, PER, W>(buf, p.M, Ns, tw, sgn, tid, nth);
    else if (R == 3)
      stockham_stage<T, 3, PER, W>(buf, p.M, Ns, tw, sgn, tid, nth);
    else
      stockham_stage<T, 5, PER, W>(buf, p.M, Ns, tw, sgn, tid, nth);
    Ns *= R;
  }
}

// Column of x that feeds v_n (the DCT reorder): v_n = x_{2n} for n < N/2,
// else x_{2(N-1-n)+1}.
__device__ __forceinline__ int dct_src_index(int n, int N) {
  return (n < N / 2) ? 2 * n : 2 * (N - 1 - n) + 1;
}
// Inverse map: position n in v of column j of x.
__device__ __forceinline__ int dct_dst_index(int j, int N) {
  return (j % 2 == 0) ? j / 2 : N - 1 - (j - 1) / 2;
}
// Real component n of the packed sequence z_m = v_{2m} + i v_{2m+1}.
template <typename T>
__device__ __forceinline__ T& packed_component(Cplx<T>* z, int n) {
  return (n & 1) ? z[n >> 1].y : z[n >> 1].x;
}

// Forward DCT-II post-processing from Z (in LDS, natural order) to X (out).
// Thread tid of nth handles k = tid + q * nth (q < KQ, k <= M/2) and its
// partner M - k.  The twiddles e1[k] = e^{-2 pi i k/N} and
// e2[k] = e^{-i pi k/(2N)} are fetched by dct2_post_twiddles before the FFT.
template <int KQ>
struct PostTwiddles {
  Cplx<double> e1k[KQ], e1p[KQ], e2k[KQ], e2p[KQ];
};

template <int KQ>
__device__ __forceinline__ void dct2_post_twiddles(PostTwiddles<KQ>& t, int M,
                                                   const Cplx<double>* __restrict__ e1,
                                                   const Cplx<double>* __restrict__ e2, int tid,
                                                   int nth) {
#pragma unroll
  for (int q = 0; q < KQ; ++q) {
    const int k = tid + q * nth;
    if (k <= M / 2) {
      t.e1k[q] = e1[k];
      t.e1p[q] = e1[M - k];
      t.e2k[q] = e2[k];
      t.e2p[q] = e2[M - k];
    }
  }
}

// One k <= M/2 of the post-processing and its partner M - k; e1k = e1[k],
// e1p = e1[M - k], e2k = e2[k], e2p = e2[M - k].
template <typename T, typename Out>
__device__ __forceinline__ void dct2_post_at(const Cplx<T>* Z, int M, int k, Cplx<double> e1k,
                                             Cplx<double> e1p, Cplx<double> e2k,
                                             Cplx<double> e2p, Out out) {
  const int N = 2 * M;
  const int kp = M - k;  // partner, in [M/2, M]; Z is periodic in M
  const Cplx<T> zk = Z[k], zp = Z[kp == M ? 0 : kp];
  // V_k = s/2 - i w d / 2 with s = Z_k + conj(Z_kp), d = Z_k - conj(Z_kp)
  const Cplx<T> sk = cadd(zk, cconj(zp)), dk = csub(zk, cconj(zp));
  const Cplx<T> wdk = cmul(Cplx<T>{T(e1k.x), T(e1k.y)}, dk);
  const Cplx<T> Vk = {T(0.5) * (sk.x + wdk.y), T(0.5) * (sk.y - wdk.x)};
  const Cplx<T> wk = cmul(Cplx<T>{T(e2k.x), T(e2k.y)}, Vk);
  out(k, wk.x);
  if (k > 0) out(N - k, -wk.y);
  if (kp != k) {
    const Cplx<T> sp = cadd(zp, cconj(zk)), dp = csub(zp, cconj(zk));
    const Cplx<T> wdp = cmul(Cplx<T>{T(e1p.x), T(e1p.y)}, dp);
    const Cplx<T> Vp = {T(0.5) * (sp.x + wdp.y), T(0.5) * (sp.y - wdp.x)};
    const Cplx<T> wp = cmul(Cplx<T>{T(e2p.x), T(e2p.y)}, Vp);
    out(kp, wp.x);
    if (kp != M) out(N - kp, -wp.y);  // for kp == M, N-kp == M is already written
  }
}

template <typename T, int KQ, typename Out>
__device__ __forceinline__ void dct2_post(const Cplx<T>* Z, int M, const PostTwiddles<KQ>& t,
                                          Out out, int tid, int nth) {
#pragma unroll
  for (int q = 0; q < KQ; ++q) {
    const int k = tid + q * nth;
    if (k <= M / 2) dct2_post_at<T>(Z, M, k, t.e1k[q], t.e1p[q], t.e2k[q], t.e2p[q], out);
  }
}

// Z_k of the inverse pre-processing below from xk = X_k, xnk = X_{N-k}
// (0 for k = 0), xm = X_{M-k}, xnm = X_{N-M+k}, q2k = e2[k], q2m = e2[M-k],
// q1k = e1[k].
template <typename T>
__device__ __forceinline__ Cplx<T> dct3_pre_at(T xk, T xnk, T xm, T xnm, Cplx<double> q2k,
                                               Cplx<double> q2m, Cplx<double> q1k) {
  const Cplx<T> Vk = cmul(Cplx<T>{T(q2k.x), T(-q2k.y)}, Cplx<T>{xk, -xnk});
  const Cplx<T> Vm = cmul(Cplx<T>{T(q2m.x), T(-q2m.y)}, Cplx<T>{xm, -xnm});
  const Cplx<T> Vkm = cconj(Vm);  // V_{k+M}
  const Cplx<T> A = cadd(Vk, Vkm);
  const Cplx<T> D = csub(Vk, Vkm);
  const Cplx<T> B = cmul(D, Cplx<T>{T(q1k.x), T(-q1k.y)});
  return Cplx<T>{A.x - B.y, A.y + B.x};
}

// Inverse pre-processing: Z_k = A_k + i B_k from X (read via `in`), k in [0, M).
//   V_k = e^{+i pi k/(2N)} (X_k - i X_{N-k}),  V_{k+M} = conj(V_{M-k})
//   A_k = V_k + V_{k+M},  B_k = (V_k - V_{k+M}) e^{+2 pi i k/N}
// Thread tid of nth handles k = tid + q * nth, q < KQ.  All global reads
// (four X values and three twiddles per k) are issued before any arithmetic.
template <typename T, int KQ, typename In>
__device__ __forceinline__ void dct3_pre(Cplx<T>* Z, int M, const Cplx<double>* __restrict__ e1,
                                         const Cplx<double>* __restrict__ e2, In in, int tid,
                                         int nth) {
  const int N = 2 * M;
  T xk[KQ], xnk[KQ], xm[KQ], xnm[KQ];
  Cplx<double> q2k[KQ], q2m[KQ], q1k[KQ];
#pragma unroll
  for (int q = 0; q < KQ; ++q) {
    const int k = tid + q * nth;
    if (k < M) {
      const int mk = M - k;  // in [1, M]
      xk[q] = in(k);
      xnk[q] = (k == 0) ? T(0) : in(N - k);
      xm[q] = in(mk);
      xnm[q] = in(N - mk);  // N - mk in [M, N)
      q2k[q] = e2[k];
      q2m[q] = e2[mk];
      q1k[q] = e1[k];
    }
  }
#pragma unroll
  for (int q = 0; q < KQ; ++q) {
    const int k = tid + q * nth;
    if (k < M) Z[k] = dct3_pre_at<T>(xk[q], xnk[q], xm[q], xnm[q], q2k[q], q2m[q], q1k[q]);
  }
}

}  // namespace axs
