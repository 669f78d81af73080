// cf_fft.hpp -- workgroup-cooperative complex DFTs in LDS for CDNA4 (gfx950).
//
// The channel-flow timestep needs DFTs of length 2^a (the small grid, 32..512)
// and 3*2^a (the 3/2 dealiasing grid, 48..768).  Every transform here is a
// self-sorting Stockham FFT executed by a whole workgroup on a batch of
// transforms that sits contiguously in LDS: stage s of radix R reads the R
// inputs of each butterfly into registers, applies the twiddles
// w_N^(r*k*N/(Ns*R)), does the R-point DFT, waits at a barrier, and writes the
// outputs back in place at (j/Ns)*Ns*R + k + r*Ns.  Radix 8 is used while the
// power of two allows, then radix 2/4, or radix 6 for a lone factor 2 with
// the factor 3 (384 = 8*8*6: three stages, six barriers).
//
// Sign convention: forward = exp(-2 pi i nk/N) (the sign FFTW 2.x uses for
// rfftw forward plans, filter_fftw.f90:49-74), unnormalised in both
// directions; callers apply 1/N where the reference does (ffthrz's nxym1,
// xderiv's nxm1).  Twiddles come from a table tw[m] = exp(-2 pi i m/N)
// computed on the host in double precision and staged in LDS by the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>

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
// Stage radices: 8 while at least three factors of two remain, then the rest
// of the power of two (2 or 4), fused with a factor 3 into radix 6 when a
// single 2 is left: 384 = 8*8*6, 256 = 8*8*4, 768 = 8*8*4*3, 48 = 8*6.
constexpr int n_stages(int n) {
  const int a = pow2_of(n), b = pow3_of(n);
  const int r = a % 3;
  int s = a / 3;
  if (b == 1) s += (r == 1) ? 1 : (r == 2 ? 2 : 1);
  else s += (r != 0) ? 1 : 0;
  return s;
}
constexpr int radix_of(int n, int s) {
  const int a = pow2_of(n), b = pow3_of(n);
  const int q = a / 3, r = a % 3;
  if (s < q) return 8;
  s -= q;
  if (b == 1) {
    if (r == 1) return 6;
    if (r == 2) return s == 0 ? 4 : 3;
    return 3;
  }
  return r == 1 ? 2 : 4;
}
constexpr int span_before(int n, int s) {
  int p = 1;
  for (int i = 0; i < s; ++i) p *= radix_of(n, i);
  return p;
}

// LDS layout of one transform: a pad element after every 8 so that the
// stride-R writes of a first stage spread over the banks.  Element i of a
// transform sits at pidx(i); transforms are padlen(N) elements apart.
constexpr int padlen(int n) { return n + n / 8; }
// Stride of the column tiles' transforms: at least padlen(n) and = 8 (mod 32)
// complex elements.  A tile's global rows are copied in and out with the 8
// columns of a row on consecutive lanes; a stride of 0 or 16 (mod 32) would
// put those 8 lanes on one LDS bank, this one spreads them over the banks.
constexpr int colpad(int n) { return padlen(n) + ((8 - padlen(n) % 32) + 32) % 32; }
__device__ __forceinline__ int pidx(int i) { return i + (i >> 3); }

// Element base + r * S of a padded transform: pidx(base + r S) = pidx(base)
// + r S + r S / 8 when S is a multiple of 8, and pidx(base) + r when S = 1,
// base is a multiple of 8 and r < 8.  Written this way the r-dependent part
// is a constant that the LDS instruction's offset field carries, instead of
// one address register per r.  The FP64 kernels use it (FP64 k_rows at 256^2
// 1.27 -> 1.23 ms, 161 -> 135 VGPRs); the FP32 ones measured slower with it
// (FP32 k_rows 0.638 -> 0.659 ms, the nnx 512 lean sweep 5.77 -> 6.11 ms)
// and keep pidx of the full index.
template <typename T, int S, bool BASE8 = false>
constexpr bool pofs_ok() { return sizeof(T) == 8 && (S % 8 == 0 || (S == 1 && BASE8)); }
template <int S>
constexpr int pofs(int r) { return S % 8 == 0 ? r * S + (r * S) / 8 : r * S; }

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
  } else if constexpr (R == 4) {
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
  } else if constexpr (R == 6) {
    // two 3-point DFTs (even, odd samples) joined by w6^k = exp(-+ i pi k/3)
    const T s3 = T(0.86602540378443864676372317075294);
    cx<T> a[3] = {v[0], v[2], v[4]};
    cx<T> b[3] = {v[1], v[3], v[5]};
    butterfly<T, 3, INV>(a);
    butterfly<T, 3, INV>(b);
    // w6 = 0.5 -+ i s3 ; w6^2 = -0.5 -+ i s3
    const cx<T> b1 = INV ? cx<T>{T(0.5) * b[1].x - s3 * b[1].y, T(0.5) * b[1].y + s3 * b[1].x}
                         : cx<T>{T(0.5) * b[1].x + s3 * b[1].y, T(0.5) * b[1].y - s3 * b[1].x};
    const cx<T> b2 = INV ? cx<T>{-T(0.5) * b[2].x - s3 * b[2].y, -T(0.5) * b[2].y + s3 * b[2].x}
                         : cx<T>{-T(0.5) * b[2].x + s3 * b[2].y, -T(0.5) * b[2].y - s3 * b[2].x};
    v[0] = cadd(a[0], b[0]);
    v[3] = csub(a[0], b[0]);
    v[1] = cadd(a[1], b1);
    v[4] = csub(a[1], b1);
    v[2] = cadd(a[2], b2);
    v[5] = csub(a[2], b2);
  } else {  // R == 8: two 4-point DFTs joined by w8^k
    const T h = T(0.70710678118654752440084436210485);
    cx<T> e[4] = {v[0], v[2], v[4], v[6]};
    cx<T> o[4] = {v[1], v[3], v[5], v[7]};
    butterfly<T, 4, INV>(e);
    butterfly<T, 4, INV>(o);
    const cx<T> o1 = INV ? cx<T>{h * (o[1].x - o[1].y), h * (o[1].x + o[1].y)}
                         : cx<T>{h * (o[1].x + o[1].y), h * (o[1].y - o[1].x)};
    const cx<T> o2 = INV ? cx<T>{-o[2].y, o[2].x} : cx<T>{o[2].y, -o[2].x};
    const cx<T> o3 = INV ? cx<T>{-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y)}
                         : cx<T>{h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y)};
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
  }
}

// Copy the length-n twiddle table into LDS (call before the first fft_run;
// the caller's next barrier publishes it).
template <typename T>
__device__ __forceinline__ void load_twiddles(cx<T>* dst, const cx<T>* __restrict__ src, int n,
                                              int tid, int nt) {
  for (int i = tid; i < n; i += nt) dst[i] = src[i];
}

// Ordering point for LDS traffic among the lanes of one wavefront: LDS
// instructions of a wave execute in order, so a compiler-level fence suffices
// (no s_barrier).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool WAVE>
__device__ __forceinline__ void fft_sync() {
  if constexpr (WAVE) wave_sync();
  else __syncthreads();
}

// Twiddle sources of a Stockham stage: the LDS table (w^(r k STEP)), or the
// values this thread's butterflies need, held in registers (RegTw).
template <typename T>
struct TwLds {
  static constexpr bool per_lane = false;    // any butterfly's twiddles
  const cx<T>* __restrict__ tw;
  template <int S, int STEP>
  __device__ __forceinline__ cx<T> get(int, int r, int k) const { return tw[r * k * STEP]; }
};

// The twiddles of every stage of NBT transforms of length N run by NT
// threads, for this thread's butterflies: w[S][q][r] for butterfly q of
// stage S.  A wave-owned row transform's butterflies are the same at every
// level, so a row kernel loads them once instead of once per stage.
template <typename T, int N, int NBT, int NT>
struct RegTw {
  static constexpr bool per_lane = true;     // only this thread's own butterflies
  static constexpr int NSTG = n_stages(N);
  static constexpr int PERMAX = (NBT * (N / 2) + NT - 1) / NT;
  cx<T> w[NSTG][PERMAX][8];
  template <int S = 0>
  __device__ __forceinline__ void init(const cx<T>* __restrict__ tw, int tid) {
    if constexpr (S < NSTG) {
      constexpr int R = radix_of(N, S), NS = span_before(N, S), NBF = N / R, TOT = NBT * NBF;
      constexpr int PER = (TOT + NT - 1) / NT, STEP = N / (NS * R);
      if constexpr (NS > 1) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
          const int idx = tid + q * NT;
          const int j = (idx < TOT ? idx : 0) % NBF;
#pragma unroll
          for (int r = 1; r < R; ++r) w[S][q][r] = tw[r * (j % NS) * STEP];
        }
      }
      init<S + 1>(tw, tid);
    }
  }
  template <int S, int STEP>
  __device__ __forceinline__ cx<T> get(int q, int r, int) const { return w[S][q][r]; }
};

template <typename T>
__device__ __forceinline__ TwLds<T> tw_src(const cx<T>* tw) { return TwLds<T>{tw}; }
template <typename T, int N, int NBT, int NT>
__device__ __forceinline__ const RegTw<T, N, NBT, NT>& tw_src(const RegTw<T, N, NBT, NT>& r) { return r; }

// One Stockham stage over NBT transforms of length N held at a[t*N + n],
// executed by NT threads (a whole workgroup, or one wavefront of 64 lanes
// when WAVE).  Entry and exit are synchronised at that scope.
//
// BAND > 0 marks a 3/2-grid transform whose elements BAND .. N-BAND are the
// zero pad: an inverse transform's first stage takes them as zero without
// reading them (the caller need not write them), and a forward transform's
// last stage does not store them (the caller does not read them).
template <typename T, int N, bool INV, int NBT, int NT, int S, bool WAVE, int PLD = padlen(N),
          class TWS = TwLds<T>, int BAND = 0>
__device__ __forceinline__ void fft_stage(cx<T>* a, const TWS& tws, int tid) {
  constexpr int R = radix_of(N, S);
  constexpr int NS = span_before(N, S);
  constexpr int NBF = N / R;
  constexpr int TOT = NBT * NBF;
  constexpr int PER = (TOT + NT - 1) / NT;
  constexpr int STEP = N / (NS * R);
  constexpr bool ZIN = INV && BAND > 0 && S == 0;
  constexpr bool ZOUT = !INV && BAND > 0 && S == n_stages(N) - 1;   // there NS = NBF: output j + r NS
  cx<T> v[PER][R];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int idx = tid + q * NT;
    if ((TOT % NT) == 0 || idx < TOT) {
      const int t = idx / NBF;
      const int j = idx - t * NBF;
      const cx<T>* src = a + t * PLD;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int lo = r * NBF, hi = r * NBF + NBF - 1;          // this r's input range
        if (ZIN && lo >= BAND && hi <= N - BAND) {
          v[q][r] = cx<T>{T(0), T(0)};                         // all of this r's inputs are pad
        } else if (ZIN && hi >= BAND && lo <= N - BAND) {
          // partly pad: every lane loads (an in-bounds, possibly unwritten
          // element) and the pad lanes select zero -- no exec-masked load
          const int p = j + lo;
          const cx<T> x = src[pidx(p)];
          const bool pad = p >= BAND && p <= N - BAND;
          v[q][r] = cx<T>{pad ? T(0) : x.x, pad ? T(0) : x.y};
        } else if constexpr (pofs_ok<T, NBF>()) {
          v[q][r] = (src + pidx(j))[pofs<NBF>(r)];
        } else {
          v[q][r] = src[pidx(j + r * NBF)];
        }
      }
      if constexpr (NS > 1) {
        const int k = j % NS;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const cx<T> w = tws.template get<S, STEP>(q, r, k);
          v[q][r] = INV ? cmulc(v[q][r], w) : cmul(v[q][r], w);
        }
      }
      butterfly<T, R, INV>(v[q]);
    }
  }
  fft_sync<WAVE>();
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int idx = tid + q * NT;
    if ((TOT % NT) == 0 || idx < TOT) {
      const int t = idx / NBF;
      const int j = idx - t * NBF;
      const int k = j % NS;
      cx<T>* dst = a + t * PLD;
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (ZOUT && r * NS >= BAND && r * NS + NS - 1 <= N - BAND) continue;
        if constexpr (pofs_ok<T, NS, R == 8>()) dst[pidx(base) + pofs<NS>(r)] = v[q][r];
        else dst[pidx(base + r * NS)] = v[q][r];
      }
    }
  }
  fft_sync<WAVE>();
}

// Stages S .. SE-1 (SE defaults to all of them).
template <typename T, int N, bool INV, int NBT, int NT, int S = 0, bool WAVE = false, int PLD = padlen(N),
          int BAND = 0, int SE = n_stages(N), class TWS>
__device__ __forceinline__ void fft_run_src(cx<T>* a, const TWS& tws, int tid) {
  if constexpr (S < SE) {
    fft_stage<T, N, INV, NBT, NT, S, WAVE, PLD, TWS, BAND>(a, tws, tid);
    fft_run_src<T, N, INV, NBT, NT, S + 1, WAVE, PLD, BAND, SE, TWS>(a, tws, tid);
  }
}

// Butterfly j (< N / R) of the last stage of a transform of length N: reads
// x[j + (N/R) r], r < R, applies its twiddles and the radix-R butterfly, and
// keeps the outputs X[j + (N/R) r] in v[r] instead of storing them (the stages
// before it ran with SE = n_stages(N) - 1).  A wave-owned transform with
// N = 64 R has one such butterfly per lane, j = lane.
template <typename T, int N>
constexpr int last_radix() { return radix_of(N, n_stages(N) - 1); }
template <typename T, int N, bool INV, class TWS>
__device__ __forceinline__ void fft_last_regs(const cx<T>* a, const TWS& tws, int lane,
                                              cx<T> (&v)[last_radix<T, N>()]) {
  constexpr int S = n_stages(N) - 1;
  constexpr int R = radix_of(N, S), NS = span_before(N, S), NBF = N / R, STEP = N / (NS * R);
  static_assert(NBF == NS, "the last stage");
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if constexpr (pofs_ok<T, NBF>()) v[r] = (a + pidx(lane))[pofs<NBF>(r)];
    else v[r] = a[pidx(lane + r * NBF)];
  }
#pragma unroll
  for (int r = 1; r < R; ++r) {
    const cx<T> w = tws.template get<S, STEP>(0, r, lane);
    v[r] = INV ? cmulc(v[r], w) : cmul(v[r], w);
  }
  butterfly<T, R, INV>(v);
}

// The inverse transform in transposed form.  The forward DFT matrix is
// symmetric, so for the forward stages F = S_{n-1} ... S_0 the unnormalised
// inverse is conj(F) = conj(S_0^T) ... conj(S_{n-1}^T): the stages run last
// to first, and stage S transposed gathers from stage S's output positions,
// runs the inverse butterfly, multiplies by the conjugate twiddles and
// scatters to stage S's input positions.  It needs the forward transform's
// twiddle set (the same RegTw registers), and its first stage is the forward
// transform's last one: for N = 64 R one butterfly per lane on x[j + 64 r],
// which a caller can feed from registers (fft_first_regs_t).
template <typename T, int N, int NBT, int NT, int S, bool WAVE, int PLD, class TWS>
__device__ __forceinline__ void fft_stage_t(cx<T>* a, const TWS& tws, int tid) {
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
      const int k = j % NS;
      const cx<T>* src = a + t * PLD;
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if constexpr (pofs_ok<T, NS, R == 8>()) v[q][r] = src[pidx(base) + pofs<NS>(r)];
        else v[q][r] = src[pidx(base + r * NS)];
      }
      butterfly<T, R, true>(v[q]);
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[q][r] = cmulc(v[q][r], tws.template get<S, STEP>(q, r, k));
      }
    }
  }
  fft_sync<WAVE>();
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int idx = tid + q * NT;
    if ((TOT % NT) == 0 || idx < TOT) {
      const int t = idx / NBF;
      const int j = idx - t * NBF;
      cx<T>* dst = a + t * PLD;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if constexpr (pofs_ok<T, NBF>()) dst[pidx(j) + pofs<NBF>(r)] = v[q][r];
        else dst[pidx(j + r * NBF)] = v[q][r];
      }
    }
  }
  fft_sync<WAVE>();
}

// Transposed stages S, S-1, .., 0.
template <typename T, int N, int NBT, int NT, int S, bool WAVE, int PLD, class TWS>
__device__ __forceinline__ void fft_run_t(cx<T>* a, const TWS& tws, int tid) {
  fft_stage_t<T, N, NBT, NT, S, WAVE, PLD, TWS>(a, tws, tid);
  if constexpr (S > 0) fft_run_t<T, N, NBT, NT, S - 1, WAVE, PLD, TWS>(a, tws, tid);
}

// Butterfly j (< N / R) of the first transposed stage of an inverse transform
// of length N from registers: v[r] holds x[j + (N/R) r]; inverse butterfly,
// conjugate twiddles, outputs stored at a[j + (N/R) r].  The caller orders the
// stores (wave_sync or a barrier) before fft_run_t of stages n_stages(N)-2 ..
// 0.  A wave-owned transform with N = 64 R has one such butterfly per lane.
template <typename T, int N, class TWS>
__device__ __forceinline__ void fft_first_regs_t(cx<T>* a, const TWS& tws, int lane,
                                                 cx<T> (&v)[last_radix<T, N>()]) {
  constexpr int S = n_stages(N) - 1;
  constexpr int R = radix_of(N, S), NS = span_before(N, S), NBF = N / R, STEP = N / (NS * R);
  static_assert(NBF == NS, "the last forward stage");
  butterfly<T, R, true>(v);
#pragma unroll
  for (int r = 1; r < R; ++r) v[r] = cmulc(v[r], tws.template get<S, STEP>(0, r, lane));
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if constexpr (pofs_ok<T, NBF>()) a[pidx(lane) + pofs<NBF>(r)] = v[r];
    else a[pidx(lane + r * NBF)] = v[r];
  }
}

template <typename T, int N, bool INV, int NBT, int NT, int S = 0, bool WAVE = false, int PLD = padlen(N),
          int BAND = 0>
__device__ __forceinline__ void fft_run(cx<T>* a, const cx<T>* __restrict__ tw, int tid) {
  fft_run_src<T, N, INV, NBT, NT, S, WAVE, PLD, BAND>(a, TwLds<T>{tw}, tid);
}

// NBT transforms (PLD elements apart, padded inside as pidx) spread over the
// NW wavefronts of a workgroup, each wave running its share with wave-level
// ordering only.  The caller synchronises the workgroup before (inputs
// written) and after (outputs read by others).
template <typename T, int N, bool INV, int NBT, int NW, int PLD = padlen(N), int BAND = 0>
__device__ __forceinline__ void fft_waves(cx<T>* a, const cx<T>* __restrict__ tw, int tid) {
  const int w = tid >> 6, lane = tid & 63;
  if constexpr (NBT % NW == 0) {
    constexpr int PW = NBT / NW;
    fft_run<T, N, INV, PW, 64, 0, true, PLD, BAND>(a + w * PW * PLD, tw, lane);
  } else {
    for (int t = w; t < NBT; t += NW) fft_run<T, N, INV, 1, 64, 0, true, PLD, BAND>(a + t * PLD, tw, lane);
  }
}

}  // namespace cf
