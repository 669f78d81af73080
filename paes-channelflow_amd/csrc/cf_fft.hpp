// cf_fft.hpp -- workgroup-cooperative complex DFTs in LDS for CDNA4 (gfx950).
//
// The channel-flow timestep needs DFTs of length 2^a (the small grid, 32..512)
// and 3*2^a (the 3/2 dealiasing grid, 48..768).  Every transform here is a
// self-sorting Stockham FFT executed by a whole workgroup on a batch of
// transforms that sits contiguously in LDS: stage s of radix R reads the R
// inputs of each butterfly into registers, applies the twiddles
// w_N^(r*k*N/(Ns*R)), does the R-point DFT, waits at a barrier, and writes the
// outputs back in place at (j/Ns)*Ns*R + k + r*Ns.  Radix 4 is used while the
// power of two allows, then one radix-2 stage, then radix 3.
//
// Sign convention: forward = exp(-2 pi i nk/N) (the sign FFTW 2.x uses for
// rfftw forward plans, filter_fftw.f90:49-74), unnormalised in both
// directions; callers apply 1/N where the reference does (ffthrz's nxym1,
// xderiv's nxm1).  Twiddles come from a table tw[m] = exp(-2 pi i m/N)
// computed on the host in double precision.
#pragma once
#include <hip/hip_runtime.h>

namespace cf {

template <typename T>
struct cx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ cx<T> cmulc(cx<T> a, cx<T> b) {
  return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename T>
__device__ __forceinline__ cx<T> cscale(cx<T> a, T s) { return {a.x * s, a.y * s}; }
// i * a
template <typename T>
__device__ __forceinline__ cx<T> cmuli(cx<T> a) { return {-a.y, a.x}; }
template <typename T>
__device__ __forceinline__ cx<T> cconj(cx<T> a) { return {a.x, -a.y}; }

constexpr int pow2_of(int n) {
  int a = 0;
  while (n % 2 == 0) { n /= 2; ++a; }
  return a;
}
constexpr int pow3_of(int n) {
  int b = 0;
  while (n % 3 == 0) { n /= 3; ++b; }
  return b;
}
constexpr int n_stages(int n) { return pow2_of(n) / 2 + pow2_of(n) % 2 + pow3_of(n); }
constexpr int radix_of(int n, int s) {
  const int q = pow2_of(n) / 2;
  if (s < q) return 4;
  s -= q;
  if (pow2_of(n) % 2) {
    if (s == 0) return 2;
    s -= 1;
  }
  return 3;
}
constexpr int span_before(int n, int s) {
  int p = 1;
  for (int i = 0; i < s; ++i) p *= radix_of(n, i);
  return p;
}

template <typename T, int R, bool INV>
__device__ __forceinline__ void butterfly(cx<T>* v) {
  if constexpr (R == 2) {
    cx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 3) {
    const T s3 = T(0.86602540378443864676372317075294);  // sqrt(3)/2
    cx<T> t1 = cadd(v[1], v[2]);
    cx<T> t2 = {v[0].x - T(0.5) * t1.x, v[0].y - T(0.5) * t1.y};
    cx<T> d = csub(v[1], v[2]);
    // forward: t3 = -i s3 d ; inverse: t3 = +i s3 d
    cx<T> t3 = INV ? cx<T>{-s3 * d.y, s3 * d.x} : cx<T>{s3 * d.y, -s3 * d.x};
    v[0] = cadd(v[0], t1);
    v[1] = cadd(t2, t3);
    v[2] = csub(t2, t3);
  } else {  // R == 4
    cx<T> a = cadd(v[0], v[2]);
    cx<T> b = csub(v[0], v[2]);
    cx<T> c = cadd(v[1], v[3]);
    cx<T> dd = csub(v[1], v[3]);
    // forward: d = -i dd ; inverse: d = +i dd
    cx<T> d = INV ? cx<T>{-dd.y, dd.x} : cx<T>{dd.y, -dd.x};
    v[0] = cadd(a, c);
    v[1] = cadd(b, d);
    v[2] = csub(a, c);
    v[3] = csub(b, d);
  }
}

// One Stockham stage over NBT transforms of length N held at a[t*N + n].
// Entry and exit are barrier-synchronised.
template <typename T, int N, bool INV, int NBT, int NT, int S>
__device__ __forceinline__ void fft_stage(cx<T>* a, const cx<T>* __restrict__ tw, int tid) {
  constexpr int R = radix_of(N, S);
  constexpr int NS = span_before(N, S);
  constexpr int NBF = N / R;
  constexpr int TOT = NBT * NBF;
  constexpr int PER = (TOT + NT - 1) / NT;
  constexpr int STEP = N / (NS * R);
  cx<T> v[PER][R];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int idx = tid + q * NT;
    if ((TOT % NT) == 0 || idx < TOT) {
      const int t = idx / NBF;
      const int j = idx - t * NBF;
      const cx<T>* src = a + t * N + j;
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = src[r * NBF];
      if constexpr (NS > 1) {
        const int k = j % NS;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const cx<T> w = tw[r * k * STEP];
          v[q][r] = INV ? cmulc(v[q][r], w) : cmul(v[q][r], w);
        }
      }
      butterfly<T, R, INV>(v[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int idx = tid + q * NT;
    if ((TOT % NT) == 0 || idx < TOT) {
      const int t = idx / NBF;
      const int j = idx - t * NBF;
      const int k = j % NS;
      cx<T>* dst = a + t * N + (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) dst[r * NS] = v[q][r];
    }
  }
  __syncthreads();
}

template <typename T, int N, bool INV, int NBT, int NT, int S = 0>
__device__ __forceinline__ void fft_run(cx<T>* a, const cx<T>* __restrict__ tw, int tid) {
  if constexpr (S < n_stages(N)) {
    fft_stage<T, N, INV, NBT, NT, S>(a, tw, tid);
    fft_run<T, N, INV, NBT, NT, S + 1>(a, tw, tid);
  }
}

}  // namespace cf
