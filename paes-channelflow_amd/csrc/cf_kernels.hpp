// cf_kernels.hpp -- HIP kernels of one RK3 stage of the channel-flow LES.
//
// Data layout in HBM (level L of a field lives in slot L, slots 0..nnz+1):
//   band spectrum  S[L][j][kx]  complex, j in [0,ny) (ky = j, or j-ny above
//                  ny/2; row ny/2 is the Nyquist row and stays zero),
//                  kx in [0, nx/2)  -- the band kept by expand/compress
//                  (filter_fftw.f90:269-272,300-303).
//   x-spectral     X[L][y][kx]  complex, y a physical row of the small (ny) or
//                  3/2 (nyb) grid, kx in [0, nx/2).
//   physical       P[L][y][x]   real, the reference's u(nnx,nny,iz) order.
//
// Stage pipeline (the reference's comp1 -> comprhs -> press -> comp2):
//   k_wall       plane means of level 1 / nnz -> utau, wall-stress factors
//                (sufsim, lowerbc.f90:57-91; sufsim_upper, upperbc.f90:64-90)
//   k_col_synth  y-inverse of u,v,w,e and d/dy of each onto the 3/2 grid
//                (expand + yderiv, filter_fftw.f90:194-278)
//   k_rows       per 3/2-grid row and z-chunk: x-inverse of the 12 fields
//                needed at a level (expand + xderiv), the whole pointwise
//                tendency of comp1 (compadv1, complen, compsgs1, compsgs3,
//                wall stress, forcing), and the x-forward of the fluxes,
//                i kx F_x + H already combined (xderiv + compress)
//   k_col_anal   y-forward of the fluxes, i ky F_y added, truncation to the
//                band (compress) and rw mean removal (comp1.f90:161-177)
//   k_press      one (kx,ky) mode per lane, sweep in z: RK substep
//                (comp1.f90:126-133), Poisson RHS (comprhs.f90), Thomas
//                solve (press.F90, tridv.f90) in FP64 and, on the way back
//                up, the projection and RK update (comp2.f90)
//   k_erows      the SGS TKE update in physical space (comp2.f90:38) and
//                its x-forward for the next stage
#pragma once
#include "cf_fft.hpp"
#include "cf_mlp.hpp"

namespace cf {

// The wavefront index of thread tid as a wave-uniform (scalar) value, so that
// per-wave choices (which transform, which source row) stay in SGPRs.
__device__ __forceinline__ int wave_of(int tid) { return __builtin_amdgcn_readfirstlane(tid >> 6); }
// The same for the row transforms of length NXB, except FP64 rows of 768
// points: there the scalar index raises the register demand past the 168 of
// three waves per SIMD, and more of the transform spills.
template <typename T, int NXB>
__device__ __forceinline__ int row_wave(int tid) {
  if constexpr (sizeof(T) == 8 && NXB > 384) return tid >> 6;
  else return wave_of(tid);
}

// ---------------------------------------------------------------------------
// Compile-time shape helpers
// ---------------------------------------------------------------------------
template <typename T, int NX, int NY>
struct ColShape {
  static constexpr int MX = NX / 2;
  static constexpr int NYB = 3 * NY / 2;
  static constexpr int BYTES = 4 * NY * (int)sizeof(cx<T>);  // per column of the tile
  static constexpr int TK0 = 65536 / BYTES;
  static constexpr int TK = TK0 < MX ? TK0 : MX;             // columns per workgroup
  // k_col_synth / k_col_anal: at least one transform per wavefront of the
  // 512-thread workgroup (they pass one batch of TKC transforms through LDS)
  static constexpr int TKC = TK >= 8 ? TK : (MX < 8 ? MX : 8);
  // k_sgs_cols_in / k_sgs_cols_out (a batch of 2 TKS transforms): FP32 tiles
  // of at least 8 columns, 64 bytes of each row, though at ny = 512 the
  // 75 KB batch leaves two workgroups per CU instead of four (512^2 x 256:
  // k_sgs_cols_out 1.16 -> 0.84 ms, k_sgs_cols_in 0.87 -> 0.84 ms; learned
  // step 51.8 -> 50.7 ms, 50.9 ms with k_sgs_cols_in left at 4 columns)
  static constexpr int TKS = (sizeof(T) == 4 && TK < 8 && MX >= 8) ? 8 : TK;
  // Column stages from registers: with 64 threads per column of a TKC tile
  // and NY = 256, thread (c, j0) = (tid % TKC, tid / TKC) holds band rows
  // j0 + 64 q, q < 4.  On the 3/2 grid (384 = 64 x 6) those are stride-64
  // classes 0, 1, 4, 5 of index j0 (2 and 3 are pad, and so is class 4 at
  // j0 = 0, the Nyquist row); on the small grid (256 = 64 x 4) classes 0..3.
  // A first transposed inverse stage or a last forward stage (one butterfly
  // per index j0 of a column) then runs on exactly the thread's own values.
  static constexpr bool CREG = NY == 256 && 512 / TKC == 64;
  // Eight waves per tile.  The column kernels pass one batch of TK padded
  // transforms through LDS at a time (about 32 KB at 256^2 FP32), so four
  // workgroups share a CU.
  static constexpr int NT = 512;
};

// Column tile of workgroup x among n tiles of one plane.  Workgroups go to the
// eight XCDs round-robin by linear workgroup id; x varies fastest and n is a
// multiple of 16 wherever this remaps, so the XCD is x mod 8.  Tiles 2j and
// 2j+1 share every 128-byte line of the rows they read (each covers half);
// they go to workgroups x and x + 8, the same XCD, so the second half-line
// read hits that XCD's L2.
__device__ __forceinline__ int xcd_tile(int x, int n) {
  return (n % 16 == 0) ? ((x & ~15) | ((x & 7) << 1) | ((x >> 3) & 1)) : x;
}

template <typename T, int NX>
struct RowShape {
  static constexpr int MX = NX / 2;
  static constexpr int NXB = 3 * NX / 2;
  static constexpr int NT = NXB;         // one thread per 3/2-grid point of the row
  static constexpr int NB = 6;           // transforms per LDS batch
};

// Device-side scalars shared by the stage kernels.
struct WallScalars {
  double c_lo, c_hi;          // utau^2/vsfc of the lower and upper wall
  double utau_lo, utau_hi;
  double utau, utau_s, dudt, dvdt;
  double mu[2], mv[2];        // plane means of u, v at level 1 / nnz (without ugal)
  double vsfc[2];             // max(|<u> + ugal|, 0.01) per wall
  // tau_uw, tau_vw at the wall = A + B q + (C + D q) wind with q = u or v
  // at the wall level, wind = sqrt(u^2 + v^2): wc[wall][0: u, 1: v][A B C D]
  double wc[2][2][4];
  int flags;                  // bit 0: utau > 10, bit 1: e clamp, bit 2: non-finite
  int wind_terms;             // C or D nonzero
};

enum : int { FLAG_UTAU = 1, FLAG_ECLAMP = 2, FLAG_NONFINITE = 4 };

struct Consts {
  double dzm, dzm2, todzm, dz;
  double alk, xkmax, ceps, emin;
  double ugal, vgal;
  double dtab1, dtab2, dtabm;
  double xymb, zody, vk, zl;
  double dsl;                 // filter width (dx dy dz)^(1/3), les.F90:118
  int nz, istage, mnout, zchunk;
  int isurface;               // wall model 1..6 (lowerbc.f90:126-161)
  int imodel;                 // 3: K from e; 4: K from the face shear (k_shear)
  int sgs_momentum;           // 0: interior SGS momentum stresses off (small-grid closure)
};

// Sum of NQ per-thread values over the workgroup; results land in out[q] of
// thread 0..NQ-1 (returned through red, an LDS scratch of NQ*NT doubles).
template <int NT, int NQ>
__device__ __forceinline__ void block_sum(double (&v)[NQ], double* red, int tid) {
  if constexpr (NT % 64 == 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_xor(v[q], off, 64);
    constexpr int NW = NT / 64;
    if ((tid & 63) == 0) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) red[q * NW + tid / 64] = v[q];
    }
    __syncthreads();
    if (tid < NQ) {
      double s = 0;
      for (int w = 0; w < NW; ++w) s += red[tid * NW + w];
      red[NQ * NW + tid] = s;
    }
    __syncthreads();
    if (tid < NQ) v[0] = red[NQ * NW + tid];
  } else {
    constexpr int RP = (NT <= 64) ? 64 : (NT <= 128 ? 128 : (NT <= 256 ? 256 : (NT <= 512 ? 512 : 1024)));
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      red[q * RP + tid] = v[q];
      if (tid + NT < RP) red[q * RP + tid + NT] = 0.0;
    }
    __syncthreads();
    for (int h = RP / 2; h > 0; h >>= 1) {
      if (tid < h) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) red[q * RP + tid] += red[q * RP + tid + h];
      }
      __syncthreads();
    }
    if (tid < NQ) v[0] = red[tid * RP];
  }
  __syncthreads();
}

template <int NT, int NQ>
struct RedSize {
  static constexpr int RP = (NT <= 64) ? 64 : (NT <= 128 ? 128 : (NT <= 256 ? 256 : (NT <= 512 ? 512 : 1024)));
  static constexpr int N = (NT % 64 == 0) ? NQ * (NT / 64 + 1) : NQ * RP;
};

// ---------------------------------------------------------------------------
// Wall models (sufsim lowerbc.f90:25-165, sufsim_upper upperbc.f90:33-158)
// as coefficients of tau = A + B q + (C + D q) wind.  Wall 0 is the lower
// (level 1, signs -), wall 1 the upper (level nnz, signs +, ugal added to
// the plane mean).  The fluctuation statistics of cases 2..5 come as shifted
// plane sums st[8] over the N points of the 3/2 plane, with u' = u - <u>,
// v' = v - <v>, s' = wind - s0, s0 = |(<u>, <v>)|:
//   st = sum s', s'^2, u'^2, v'^2, s'u', s'v', u', v'.
// Undefined values of the reference: ufree (lowerbc.f90:38,63) is taken as
// 0; case 2's ufluc, vfluc (set only for isurface > 2) as defined
// (lowerbc.f90:97-98); moengv2m = <moengu^2> as the reference computes it
// (lowerbc.f90:121).
// ---------------------------------------------------------------------------
constexpr double BETAA = 0.35;                                 // parameters.f90:8
constexpr double FORCING = 0.016;                              // imposed dp/dx, comp1.f90:152

__device__ inline void wall_coeffs(WallScalars* ws, const Consts& c, int wall, const double* st,
                                   double N) {
  const double sg = wall ? 1.0 : -1.0;
  const double ugal = c.ugal, vgal = c.vgal;
  const double u1xy = ws->mu[wall] + (wall ? ugal : 0.0);
  const double v1xy = ws->mv[wall] + (wall ? vgal : 0.0);
  const double vsfc = ws->vsfc[wall];
  const double utau = wall ? ws->utau_hi : ws->utau_lo;
  const double utau2 = utau * utau;
  const double dU = ugal - u1xy, dV = vgal - v1xy;             // ufluc = u + dU
  double* A = ws->wc[wall][0];
  double* B = ws->wc[wall][1];
  for (int i = 0; i < 4; ++i) A[i] = B[i] = 0.0;
  switch (c.isurface) {
    case 1: {                                                   // Schumann
      const double k = sg * utau2 / vsfc;
      A[0] = k * ugal; A[1] = k;
      B[0] = k * vgal; B[1] = k;
      return;
    }
    case 6: {                                                   // mean-forcing balance
      const double a6u = -c.zl / 2.0 * (FORCING - ws->dudt), a6v = -c.zl / 2.0 * (-ws->dvdt);
      const double r = sg * utau * ws->utau_s / vsfc;
      A[0] = -sg * a6u + r * dU; A[1] = r;
      B[0] = -sg * a6v + r * dV; B[1] = r;
      return;
    }
    default: break;
  }
  const double mu = ws->mu[wall], mv = ws->mv[wall];
  const double s0 = sqrt(mu * mu + mv * mv);
  const double ms = st[0] / N, mss = st[1] / N, muu = st[2] / N, mvv = st[3] / N;
  const double msu = st[4] / N, msv = st[5] / N, mu_ = st[6] / N, mv_ = st[7] / N;
  double windm = s0 + ms;
  windm = windm > 0.0 ? windm : 0.0;                           // ufree
  windm = windm > 0.01 ? windm : 0.01;
  const double eu = mu + dU, ev = mv + dV;                     // ufluc = u' + eu
  const double ds = windm - s0;                                // sfluc = s' - ds
  const double ufl2 = muu + 2.0 * eu * mu_ + eu * eu;
  const double vfl2 = mvv + 2.0 * ev * mv_ + ev * ev;
  const double sufl = msu - ds * mu_ + eu * ms - eu * ds;
  const double svfl = msv - ds * mv_ + ev * ms - ev * ds;
  const double moeng2 = u1xy * u1xy * (mss - 2.0 * ds * ms + ds * ds) + 2.0 * u1xy * windm * sufl +
                        windm * windm * ufl2;
  double tep = u1xy / vsfc;
  tep = tep < 1.0 ? tep : 1.0;
  tep = tep > -1.0 ? tep : -1.0;
  const double thta = acos(tep);
  const double auuwm = sg * utau2 * cos(thta);
  const double auvwm = sg * utau2 * sin(thta) * (v1xy >= 0.0 ? 1.0 : -1.0);
  switch (c.isurface) {
    case 2: {                                                   // Moeng
      const double k = sg * utau2 / (windm * vsfc);
      A[0] = k * windm * dU; A[1] = k * windm; A[2] = k * u1xy;
      B[0] = k * windm * dV; B[1] = k * windm; B[2] = k * v1xy;
      break;
    }
    case 3: {                                                   // Wei, Schumann fluctuation
      const double qu = -sg * BETAA * utau2 / sqrt(ufl2), qv = -sg * BETAA * utau2 / sqrt(vfl2);
      A[0] = auuwm + qu * dU; A[1] = qu;
      B[0] = auvwm + qv * dV; B[1] = qv;
      break;
    }
    case 4: {                                                   // Moeng, not linearised
      const double du = windm * u1xy + sufl, dv = windm * v1xy + svfl;
      const double k = sg * utau2 / sqrt(du * du + dv * dv);
      A[2] = k * ugal; A[3] = k;
      B[2] = k * vgal; B[3] = k;
      break;
    }
    case 5: {                                                   // Wei, Moeng fluctuation
      const double q = -sg * BETAA * utau2 / sqrt(moeng2);
      A[0] = auuwm + q * windm * (dU - u1xy); A[1] = q * windm; A[2] = q * u1xy;
      B[0] = auvwm + q * windm * (dV - v1xy); B[1] = q * windm; B[2] = q * v1xy;
      break;
    }
    default: break;
  }
}

template <typename T>
__device__ __forceinline__ void wall_tau(const WallScalars* ws, int wall, T u, T v, T& tu, T& tv) {
  const double* A = ws->wc[wall][0];
  const double* B = ws->wc[wall][1];
  tu = (T)A[0] + (T)A[1] * u;
  tv = (T)B[0] + (T)B[1] * v;
  if (ws->wind_terms) {
    const T s = sqrt(u * u + v * v);
    tu += ((T)A[2] + (T)A[3] * u) * s;
    tv += ((T)B[2] + (T)B[3] * v) * s;
  }
}

// ---------------------------------------------------------------------------
// k_wall: wall-model scalars.  One workgroup of 64 threads.  As the stage's
// first kernel it also clears the stage's clamp marks (dirty[ndirty]) and the
// divz accumulators (red0[nred0]), in place of two memset dispatches.
// ---------------------------------------------------------------------------
// M0 holds the FP64 (0,0) coefficients, i.e. the plane means, of u, v, ru,
// rv: M0[q * slots + level] for q = 0..3 (M0_U .. M0_RV).
enum : int { M0_U = 0, M0_V = 1, M0_RU = 2, M0_RV = 3 };

__global__ void __launch_bounds__(64) k_wall(const double* __restrict__ M0, int slots, Consts c,
                                             WallScalars* ws, int* dirty, int ndirty, double* red0,
                                             int nred0) {
  const int nz = c.nz;
  const int tid = threadIdx.x;
  for (int i = tid; i < ndirty; i += 64) dirty[i] = 0;
  for (int i = tid; i < nred0; i += 64) red0[i] = 0.0;
  // dudt, dvdt: volume means of the previous stage's R (lowerbc.f90:70-72);
  // the plane sum of a band field is nx*ny times its (0,0) coefficient
  const double* mu = M0 + M0_U * slots;
  const double* mv = M0 + M0_V * slots;
  double su = 0.0, sv = 0.0;
  for (int z = 1 + tid; z <= nz; z += 64) {
    su += M0[M0_RU * slots + z];
    sv += M0[M0_RV * slots + z];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    su += __shfl_xor(su, off, 64);
    sv += __shfl_xor(sv, off, 64);
  }
  if (tid != 0) return;
  // plane means = the (0,0) coefficient (lowerbc.f90:57-58)
  double u1 = mu[1], v1 = mv[1];
  double vs = sqrt(u1 * u1 + v1 * v1);
  vs = vs > 0.01 ? vs : 0.01;                                   // lowerbc.f90:65
  double ut = vs * c.vk / c.zody;                               // lowerbc.f90:67
  ws->utau_lo = ut;
  ws->c_lo = ut * ut / vs;
  double un = mu[nz] + c.ugal;                                  // upperbc.f90:64-65
  double vn = mv[nz] + c.vgal;
  double vs2 = sqrt(un * un + vn * vn);
  vs2 = vs2 > 0.01 ? vs2 : 0.01;
  double ut2 = vs2 * c.vk / c.zody;
  ws->utau_hi = ut2;
  ws->c_hi = ut2 * ut2 / vs2;
  ws->utau = ut2;                                               // sufsim_upper runs last
  ws->dudt = su / nz;
  ws->dvdt = sv / nz;
  ws->utau_s = sqrt(c.zl / 2.0 * sqrt(ws->dvdt * ws->dvdt + (FORCING - ws->dudt) * (FORCING - ws->dudt)));
  if (!(ut <= 10.0) || !(ut2 <= 10.0)) ws->flags |= FLAG_UTAU;  // lowerbc.f90:77-80
  ws->mu[0] = u1; ws->mv[0] = v1;
  ws->mu[1] = mu[nz];
  ws->mv[1] = mv[nz];
  ws->vsfc[0] = vs; ws->vsfc[1] = vs2;
  ws->wind_terms = (c.isurface == 2 || c.isurface == 4 || c.isurface == 5) ? 1 : 0;
  if (c.isurface == 1 || c.isurface == 6) {                   // no plane statistics needed
    wall_coeffs(ws, c, 0, nullptr, 1.0);
    wall_coeffs(ws, c, 1, nullptr, 1.0);
  }
}

// Plane statistics of the wall-level 3/2-grid velocity for isurface 2..5:
// per row y of level 1 (wall 0) or nnz (wall 1), the sums of wall_coeffs'
// st[8] into part[(wall*NYB + y)*8 + q].  grid (NYB, 2), one thread per point.
template <typename T, int NX>
__global__ void __launch_bounds__(3 * NX / 2) k_wall_plane(const cx<T>* __restrict__ Bs, size_t bstride,
                                                           const T* __restrict__ kxp,
                                                           const cx<T>* __restrict__ twNXB, int nz,
                                                           const WallScalars* __restrict__ ws,
                                                           double* part);

// Sums of k_wall_plane's rows, then wall_coeffs for both walls.  One
// workgroup of 256 threads.
__global__ void __launch_bounds__(256) k_wall_final(const double* __restrict__ part, int nyb, double npts,
                                                    Consts c, WallScalars* ws) {
  __shared__ double red[16 * 4];
  const int tid = threadIdx.x;
  double s[16];
  for (int q = 0; q < 16; ++q) s[q] = 0.0;
  for (int y = tid; y < nyb; y += 256)
    for (int w = 0; w < 2; ++w)
      for (int q = 0; q < 8; ++q) s[w * 8 + q] += part[((size_t)w * nyb + y) * 8 + q];
  for (int q = 0; q < 16; ++q)
    for (int off = 32; off > 0; off >>= 1) s[q] += __shfl_xor(s[q], off, 64);
  if ((tid & 63) == 0)
    for (int q = 0; q < 16; ++q) red[q * 4 + (tid >> 6)] = s[q];
  __syncthreads();
  if (tid != 0) return;
  double st[16];
  for (int q = 0; q < 16; ++q) st[q] = red[q * 4] + red[q * 4 + 1] + red[q * 4 + 2] + red[q * 4 + 3];
  wall_coeffs(ws, c, 0, st, npts);
  wall_coeffs(ws, c, 1, st + 8, npts);
}

// ---------------------------------------------------------------------------
// k_col_synth: band spectrum -> 3/2-grid x-spectral rows, value and d/dy
// grid (MX/TKC, nz, 4 fields: e, w, v, u in blockIdx.z order)
//
// A workgroup holds TK band columns of one level and field in registers and
// passes the value, then the y-derivative, through one batch of TK
// transforms in LDS (zero-padded to NYB), so that four workgroups share a
// CU.  e arrives as rows of physical y: its y-forward transform runs first,
// in the same batch.
// ---------------------------------------------------------------------------
template <typename T>
struct SynthArgs {
  const cx<T>* S[3];      // U, V, W band spectra
  const cx<T>* Ex;        // e, x-spectral small grid (rows of physical y)
  cx<T>* B;               // 8 outputs: u, uy, v, vy, w, wy, e, ey ; stride bstride
  size_t bstride;         // elements per output field (all slots)
  int lean;               // 1: uy, vy, wy are not needed (lean row kernel)
  const T* kyp;           // ky of band row j (Nyquist row -> 0)
  const cx<T>* twNY;
  const cx<T>* twNYB;
  int nz;
};

template <typename T, int NX, int NY>
__global__ void __launch_bounds__(512) k_col_synth(SynthArgs<T> a) {
  using Sh = ColShape<T, NX, NY>;
  static_assert(Sh::NT == 512, "launch bound");
  constexpr int MX = Sh::MX, NYB = Sh::NYB, TK = Sh::TKC, NT = Sh::NT;
  constexpr int PLB = colpad(NYB), PLS = colpad(NY);
  constexpr int LPS = (TK * NY + NT - 1) / NT;        // band values per thread
  static_assert(PLS <= PLB, "the length-NY batch fits the length-NYB batch");
  const int f = 3 - (int)blockIdx.z;     // e first: its tiles carry three transforms, u, v, w one or two
  const int z = blockIdx.y + 1;
  if (f >= 2 && z >= a.nz) return;       // w, e live on faces 1..nz-1
  const int k0 = xcd_tile(blockIdx.x, gridDim.x) * TK;
  const int tid = threadIdx.x;
  __shared__ cx<T> buf[TK * PLB];
  __shared__ cx<T> twb[NYB];
  __shared__ cx<T> tws[NY];
  load_twiddles(twb, a.twNYB, NYB, tid, NT);
  // element idx = tid + q NT of the tile: column idx % TK, band row idx / TK
  cx<T> sv[LPS];
  if (f < 3) {
    const cx<T>* src = a.S[f] + (size_t)z * NY * MX + k0;
#pragma unroll
    for (int q = 0; q < LPS; ++q) {
      const int idx = tid + q * NT;
      if (idx < TK * NY) sv[q] = src[(size_t)(idx / TK) * MX + idx % TK];
    }
  } else {
    load_twiddles(tws, a.twNY, NY, tid, NT);
    const cx<T>* src = a.Ex + (size_t)z * NY * MX + k0;
    cx<T> ev[LPS];                       // all loads in flight before the first LDS store
#pragma unroll
    for (int q = 0; q < LPS; ++q) {
      const int i = tid + q * NT;
      if (i < TK * NY) ev[q] = src[(size_t)(i / TK) * MX + i % TK];
    }
#pragma unroll
    for (int q = 0; q < LPS; ++q) {
      const int i = tid + q * NT;
      if (i < TK * NY) buf[(i % TK) * PLS + pidx(i / TK)] = ev[q];
    }
    __syncthreads();
    const T s = T(1) / T(NY);
    if constexpr (Sh::CREG) {
      // stages before the last per owning wave, the last one on the thread
      // that keeps the band rows j0 + 64 q (classes q of the length-256 grid)
      const int w = tid >> 6, lane = tid & 63, c = tid % TK, j0 = tid / TK;
      fft_run_src<T, NY, false, TK / (NT / 64), 64, 0, true, PLS, 0, n_stages(NY) - 1>(
          buf + w * (TK / (NT / 64)) * PLS, tw_src(tws), lane);
      __syncthreads();
      cx<T> x[4];
      fft_last_regs<T, NY, false>(buf + c * PLS, tw_src(tws), j0, x);
#pragma unroll
      for (int q = 0; q < 4; ++q) sv[q] = (q == 2 && j0 == 0) ? cx<T>{0, 0} : cscale(x[q], s);
    } else {
      fft_waves<T, NY, false, TK, NT / 64, PLS>(buf, tws, tid);
      __syncthreads();
#pragma unroll
      for (int q = 0; q < LPS; ++q) {
        const int idx = tid + q * NT;
        if (idx < TK * NY) {
          const int c = idx % TK, j = idx / TK;
          sv[q] = (j == NY / 2) ? cx<T>{0, 0} : cscale(buf[c * PLS + pidx(j)], s);
        }
      }
    }
    __syncthreads();                     // buf is refilled below
  }
  cx<T>* o = a.B + (size_t)(2 * f) * a.bstride + (size_t)z * NYB * MX + k0;
  // pass d = 0: the value, d = 1: d/dy (i ky S).  Two calls rather than a
  // loop: a loop lets the compiler hoist the transform's twiddle reads out of
  // it and hold them in registers across both passes.
  auto pass = [&](int d) {
    if constexpr (Sh::CREG) {
      // the transposed inverse: its first stage (radix 6 over the stride-64
      // classes) on the thread holding the band rows, the rest per owning wave
      const int w = tid >> 6, lane = tid & 63, c = tid % TK, j0 = tid / TK;
      cx<T> x[6];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        cx<T> v = sv[q];
        if (d) {
          const T k = a.kyp[j0 + 64 * q];
          v = cx<T>{-k * v.y, k * v.x};
        }
        x[q < 2 ? q : q + 2] = v;
      }
      x[2] = x[3] = cx<T>{T(0), T(0)};
      if (j0 == 0) x[4] = cx<T>{T(0), T(0)};           // the Nyquist row is pad
      __syncthreads();                                  // twiddle table written
      fft_first_regs_t<T, NYB>(buf + c * PLB, tw_src(twb), j0, x);
      __syncthreads();
      fft_run_t<T, NYB, TK / (NT / 64), 64, n_stages(NYB) - 2, true, PLB>(buf + w * (TK / (NT / 64)) * PLB,
                                                                        tw_src(twb), lane);
    } else {
#pragma unroll
      for (int q = 0; q < LPS; ++q) {
        const int idx = tid + q * NT;
        if (idx < TK * NY) {
          const int c = idx % TK, j = idx / TK;
          if (j != NY / 2) {               // zero-pad the band into the 3/2 grid
            cx<T> v = sv[q];
            if (d) {
              const T k = a.kyp[j];
              v = cx<T>{-k * v.y, k * v.x};
            }
            buf[c * PLB + pidx(j < NY / 2 ? j : j - NY + NYB)] = v;
          }
        }
      }
      __syncthreads();                     // the zero pad NY/2 .. NYB-NY/2 is not read (BAND)
      fft_waves<T, NYB, true, TK, NT / 64, PLB, NY / 2>(buf, twb, tid);
    }
    __syncthreads();
    cx<T>* od = o + (size_t)d * a.bstride;
#pragma unroll
    for (int i = tid; i < TK * NYB; i += NT) {
      const int c = i % TK, n = i / TK;
      od[(size_t)n * MX + c] = buf[c * PLB + pidx(n)];
    }
  };
  pass(0);
  if (a.lean && f < 3) return;
  __syncthreads();                       // buf is refilled by the second pass
  pass(1);
}

// ---------------------------------------------------------------------------
// k_rows: the pointwise tendency on the 3/2 grid, one row y per workgroup,
// sliding over a chunk of levels.  grid (NYB, nchunks), one thread per point.
// ---------------------------------------------------------------------------
template <typename T>
struct RowArgs {
  const cx<T>* B;          // synthesis inputs (8 fields, see SynthArgs)
  size_t bstride;
  cx<T>* A;                // analysis outputs: Gu, Fyu, Gv, Fyv, Gw, Fyw, Ge, Fye
  size_t astride;
  const T* kxp;            // kx wavenumbers, MX entries
  const cx<T>* twNXB;
  const WallScalars* ws;
  double* part;            // divz and SGS partial sums [7][slots][NYB]
  size_t pstride;          // slots*NYB
  int* flags;
  const T* X;              // imodel 4: K at faces 0..nnz [face][NYB][NXB] (k_shear)
  int* dirty;              // [f] set where the 3/2-grid e of face f fell below emin; [0]: any
  Consts c;
};

// The expanded e of face f fell below emin at this point (comp1.f90:109,121):
// the flux corrections of k_eclamp_* run for the face.
__device__ __forceinline__ void mark_clamp(int* dirty, int f) {
  atomicOr(dirty + f, 1);
  atomicOr(dirty, 1);
}

// Row transforms of length NXB: one wavefront owns whole transforms (lanes
// exchange through LDS with wave-level ordering) when NXB is a multiple of
// 64, otherwise the workgroup shares each transform.
template <typename T, int NX>
struct RowFFT {
  static constexpr int MX = NX / 2, NXB = 3 * NX / 2, NT = NXB;
  static constexpr bool WAVE = (NXB % 64) == 0;
  static constexpr int NW = WAVE ? NXB / 64 : 1;
  static constexpr int TPW = WAVE ? (6 + NW - 1) / NW : 6;   // transforms per wave
};

// Build the full length-NXB spectrum whose inverse DFT is a + i b for two real
// rows with half spectra sa (a) and sb (b), k < MX; optional i kx factors.
// ZFILL = false leaves the zero pad MX .. NXB-MX unwritten, for a transform
// that takes it as zero (fft_run's BAND).
template <typename T, int NX, bool ZFILL = true>
__device__ __forceinline__ void fill_pair(cx<T>* z, const cx<T>* sa, bool da, const cx<T>* sb,
                                          bool db, const T* __restrict__ kxp, int t, int nt) {
  constexpr int MX = NX / 2, NXB = 3 * NX / 2;
  for (int k = t; k < MX; k += nt) {
    cx<T> av = {0, 0}, bv = {0, 0};
    if (sa) av = sa[k];
    if (sb) bv = sb[k];
    const T kx = kxp[k];
    if (da) av = cx<T>{-kx * av.y, kx * av.x};
    if (db) bv = cx<T>{-kx * bv.y, kx * bv.x};
    if (k == 0) {
      z[0] = cx<T>{av.x, bv.x};      // c2r ignores the imaginary part of the DC term
    } else {
      z[pidx(k)] = cx<T>{av.x - bv.y, av.y + bv.x};
      z[pidx(NXB - k)] = cx<T>{av.x + bv.y, bv.x - av.y};
    }
  }
  if constexpr (ZFILL)
    for (int i = t; i <= MX; i += nt) z[pidx(MX + i)] = cx<T>{0, 0};
}

// Element q (wave-uniform, not a constant) of a small local array, by
// compare-and-select: indexing it directly would place the array in scratch.
template <int NP, typename V>
__device__ __forceinline__ V pick(const V* arr, int q, V dflt) {
  V r = dflt;
#pragma unroll
  for (int j = 0; j < NP; ++j) r = (j == q) ? arr[j] : r;
  return r;
}

// Whether the synthesis of a wave-owned length-L transform of half spectra
// k < MX can start from registers: the first transposed stage (the forward
// transform's last, radix R over stride L / R) has a whole number of
// butterflies per lane (L / R a multiple of 64), and MX is a multiple of 64 so
// that every 64-element block of its inputs is all data, all pad or all
// mirrored (lane 0 aside).
template <typename T, int L, int MX>
constexpr bool synth_in_regs() {
  return (L / last_radix<T, L>()) % 64 == 0 && MX % 64 == 0 && 2 * MX <= L;
}

// NB wave-owned inverse transforms of length L (PLD elements apart) whose
// spectra Z are built from the half spectra of two real rows a, b each: lane j
// holds modes k = j + 64 kk < MX of transform t in av[t][kk], bv[t][kk] (i kx
// factors applied).  Z[k] = A[k] + i B[k] for k < MX, Z[L - k] = conj(A[k]) +
// i conj(B[k]), zero in between (the 3/2-grid pad, or the Nyquist mode of the
// small grid).  The first transposed stage (radix R over stride NBF = L / R)
// reads Z[b + NBF r] for its butterflies b = j + 64 i from registers: a
// 64-block of k below MX is the lane's own modes, one above L - MX the mirrors
// of lane 64 - j's (lane 0: its own, or zero at k = L - MX), the rest zero.
// The inverse DFTs of the Z are left in z.
template <typename T, int L, int MX, int NB, int PLD, class TWS>
__device__ __forceinline__ void synth_regs(cx<T>* z, const cx<T> (&av)[NB][MX / 64],
                                           const cx<T> (&bv)[NB][MX / 64], const TWS& tws, int lane) {
  constexpr int KPL = MX / 64, R = last_radix<T, L>(), NBF = L / R, BPL = NBF / 64;
  static_assert(synth_in_regs<T, L, MX>(), "whole butterflies per lane over whole 64-blocks");
  static_assert(BPL == 1 || !TWS::per_lane, "register twiddles hold one butterfly per lane");
  const int src = (64 - lane) & 63;
#pragma unroll
  for (int t = 0; t < NB; ++t) {
    cx<T> zo[KPL], mir[KPL];
#pragma unroll
    for (int kk = 0; kk < KPL; ++kk) {
      const cx<T> a = av[t][kk], b = bv[t][kk];
      // c2r ignores the imaginary part of the DC term
      zo[kk] = (kk == 0 && lane == 0) ? cx<T>{a.x, b.x} : cx<T>{a.x - b.y, a.y + b.x};
      mir[kk] = cx<T>{a.x + b.y, b.x - a.y};
    }
#pragma unroll
    for (int i = 0; i < BPL; ++i) {
      cx<T> v[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int k0 = 64 * i + NBF * r;              // lane j: k = k0 + j
        if (k0 + 63 < MX) {
          v[r] = zo[k0 / 64];
        } else if (k0 + 63 <= L - MX) {
          v[r] = cx<T>{T(0), T(0)};
        } else {
          const int m0 = (L - k0) / 64;               // lane j >= 1: mode 64 (m0 - 1) + 64 - j
          const cx<T> x = mir[m0 - 1];
          const cx<T> m = {__shfl(x.x, src, 64), __shfl(x.y, src, 64)};
          const cx<T> own = (64 * m0 < MX) ? mir[m0 < KPL ? m0 : 0] : cx<T>{T(0), T(0)};
          v[r] = lane != 0 ? m : own;
        }
      }
      fft_first_regs_t<T, L>(z + t * PLD, tws, lane + 64 * i, v);
    }
  }
  wave_sync();
  fft_run_t<T, L, NB, 64, n_stages(L) - 2, true, PLD>(z, tws, lane);
}

// Synthesise NP (<= 6) real row pairs; this thread's point of the two real
// rows of pair q comes back in ra[q], rb[q].  Ends with a workgroup barrier
// after which the caller must not write buf until the next anal_pairs.
// WV: each wavefront owns whole transforms (needs NXB a multiple of 64);
// otherwise the workgroup shares every transform.  PPT > 1 (NXB / PPT threads,
// at least NP wavefronts, wave-owned transforms from registers): thread t
// holds points t + p NXB / PPT, returned in ra[p NP + q], rb[p NP + q].
template <typename T, int NX, int NP, bool WV = RowFFT<T, NX>::WAVE, int PPT = 1>
__device__ __forceinline__ void synth_pairs(cx<T>* buf, const cx<T>* const* sa, const bool* da,
                                            const cx<T>* const* sb, const bool* db, T* ra, T* rb,
                                            const T* __restrict__ kxp, const cx<T>* tw, int tid) {
  using F = RowFFT<T, NX>;
  constexpr int NXB = F::NXB;
  static_assert(!WV || F::WAVE, "wave-owned transforms need whole wavefronts");
  // (FP64 rows of 768 points keep the spectrum fill: their first transposed
  // stage from registers spills more of the transform)
  constexpr bool REGS = WV && F::TPW == 1 && synth_in_regs<T, NXB, F::MX>() &&
                        (sizeof(T) == 4 || NXB / last_radix<T, NXB>() == 64);
  static_assert(PPT == 1 || REGS, "several points per thread: register synthesis only");
  if constexpr (REGS) {
    const int w = wave_of(tid), lane = tid & 63;
    if (w < NP) {
      const cx<T>* pa = pick<NP>(sa, w, (const cx<T>*)nullptr);
      const cx<T>* pb = pick<NP>(sb, w, (const cx<T>*)nullptr);
      const bool dqa = pick<NP>(da, w, false), dqb = pick<NP>(db, w, false);
      cx<T> av[1][F::MX / 64], bv[1][F::MX / 64];
#pragma unroll
      for (int kk = 0; kk < F::MX / 64; ++kk) {
        const int k = lane + 64 * kk;
        cx<T> x = pa ? pa[k] : cx<T>{T(0), T(0)};
        cx<T> y = pb ? pb[k] : cx<T>{T(0), T(0)};
        const T kx = kxp[k];
        if (dqa) x = cx<T>{-kx * x.y, kx * x.x};
        if (dqb) y = cx<T>{-kx * y.y, kx * y.x};
        av[0][kk] = x;
        bv[0][kk] = y;
      }
      synth_regs<T, NXB, F::MX, 1, padlen(NXB)>(buf + w * padlen(NXB), av, bv, tw_src(tw), lane);
    }
  } else if constexpr (WV) {
    const int w = row_wave<T, NXB>(tid), lane = tid & 63;
#pragma unroll
    for (int qq = 0; qq < F::TPW; ++qq) {
      const int q = w * F::TPW + qq;
      if (q < NP) fill_pair<T, NX, false>(buf + q * padlen(NXB), sa[q], da[q], sb[q], db[q], kxp, lane, 64);
    }
    wave_sync();
    if (w * F::TPW < NP)
      fft_run<T, NXB, true, F::TPW, 64, 0, true, padlen(NXB), F::MX>(buf + w * F::TPW * padlen(NXB), tw, lane);
  } else if constexpr (NP * F::MX == F::NT && last_radix<T, NXB>() == 3) {
    // workgroup-shared transposed inverses, one first-stage butterfly (radix 3
    // over the stride-MX classes) per thread: thread (q, j) needs Z_q[j] (mode
    // j), the pad Z_q[MX + j] and Z_q[2 MX + j], the mirror of mode MX - j
    // (zero at j = 0, the pad edge), loaded from the half spectra directly
    constexpr int MX = F::MX;
    const int q = tid / MX, j = tid - q * MX;
    const cx<T>* pa = pick<NP>(sa, q, (const cx<T>*)nullptr);
    const cx<T>* pb = pick<NP>(sb, q, (const cx<T>*)nullptr);
    const bool dqa = pick<NP>(da, q, false), dqb = pick<NP>(db, q, false);
    const int m = j == 0 ? 0 : MX - j;
    cx<T> a0 = pa ? pa[j] : cx<T>{T(0), T(0)}, b0 = pb ? pb[j] : cx<T>{T(0), T(0)};
    cx<T> am = pa ? pa[m] : cx<T>{T(0), T(0)}, bm = pb ? pb[m] : cx<T>{T(0), T(0)};
    const T kx0 = kxp[j], kxm = kxp[m];
    if (dqa) {
      a0 = cx<T>{-kx0 * a0.y, kx0 * a0.x};
      am = cx<T>{-kxm * am.y, kxm * am.x};
    }
    if (dqb) {
      b0 = cx<T>{-kx0 * b0.y, kx0 * b0.x};
      bm = cx<T>{-kxm * bm.y, kxm * bm.x};
    }
    cx<T> v[3];
    // c2r ignores the imaginary part of the DC term
    v[0] = j == 0 ? cx<T>{a0.x, b0.x} : cx<T>{a0.x - b0.y, a0.y + b0.x};
    v[1] = cx<T>{T(0), T(0)};
    v[2] = j == 0 ? cx<T>{T(0), T(0)} : cx<T>{am.x + bm.y, bm.x - am.y};
    fft_first_regs_t<T, NXB>(buf + q * padlen(NXB), tw_src(tw), j, v);
    __syncthreads();
    fft_run_t<T, NXB, NP, F::NT, n_stages(NXB) - 2, false, padlen(NXB)>(buf, tw_src(tw), tid);
  } else {
#pragma unroll
    for (int q = 0; q < NP; ++q)
      fill_pair<T, NX, false>(buf + q * padlen(NXB), sa[q], da[q], sb[q], db[q], kxp, tid, F::NT);
    __syncthreads();
    fft_run<T, NXB, true, NP, F::NT, 0, false, padlen(NXB), F::MX>(buf, tw, tid);
  }
  __syncthreads();
  if constexpr (PPT == 1) {
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      const cx<T> v = buf[q * padlen(NXB) + pidx(tid)];
      ra[q] = v.x;
      rb[q] = v.y;
    }
  } else {
#pragma unroll
    for (int p = 0; p < PPT; ++p)
#pragma unroll
      for (int q = 0; q < NP; ++q) {
        const cx<T> v = buf[q * padlen(NXB) + pidx(tid + p * (NXB / PPT))];
        ra[p * NP + q] = v.x;
        rb[p * NP + q] = v.y;
      }
  }
}

// Whether a twiddle source holds only this thread's own butterflies (RegTw).
template <class X>
constexpr bool tw_per_lane() {
  if constexpr (std::is_class<X>::value) return X::per_lane;
  else return false;
}

// Coefficients k < MX of the two real rows of transform q from the outputs of
// a wave-owned forward transform of length L = 64 R kept in registers (lane j
// holds X[j + 64 r] in v[r]), scaled by s, handed to emit(q, k, A, B).  The
// mirror X[L - k] of k = j + 64 kk is X[64 (R-1-kk) + 64 - j], held by lane
// 64 - j (k = 64 kk of lane 0: lane 0's own X[64 (R-kk)], and X[0] for k = 0).
template <typename T, int L, int MX, typename Emit>
__device__ __forceinline__ void emit_from_regs(const cx<T> (&v)[last_radix<T, L>()], int lane, T s, int q,
                                               Emit& emit) {
  constexpr int R = last_radix<T, L>(), KK = MX / 64;
  static_assert(L == 64 * R && MX % 64 == 0 && 2 * MX <= L, "one butterfly per lane, whole classes");
  const int src = (64 - lane) & 63;
  cx<T> m[KK];
#pragma unroll
  for (int kk = 0; kk < KK; ++kk) {
    const cx<T> x = v[R - 1 - kk];
    m[kk] = cx<T>{__shfl(x.x, src, 64), __shfl(x.y, src, 64)};
  }
#pragma unroll
  for (int kk = 0; kk < KK; ++kk) {
    const cx<T> zk = v[kk];
    const cx<T> zn = lane != 0 ? m[kk] : (kk == 0 ? v[0] : m[kk - 1]);
    emit(q, lane + 64 * kk, cx<T>{(zk.x + zn.x) * s, (zk.y - zn.y) * s},
         cx<T>{(zk.y + zn.y) * s, (zn.x - zk.x) * s});
  }
}

// Forward-transform NP real row pairs (this thread holds point tid of each)
// and hand coefficient k < MX of both members, scaled by 1/NXB, to
// emit(q, k, A, B).  PPT > 1 (NXB / PPT threads, at least NP wavefronts,
// wave-owned transforms with the last stage in registers): thread t holds
// points t + p NXB / PPT, in xa[p NP + q], xb[p NP + q].
template <typename T, int NX, int NP, bool WV = RowFFT<T, NX>::WAVE, int PPT = 1, typename Emit, class TWA>
__device__ __forceinline__ void anal_pairs(cx<T>* buf, const T* xa, const T* xb, const TWA& tw,
                                           int tid, Emit emit) {
  using F = RowFFT<T, NX>;
  constexpr int MX = F::MX, NXB = F::NXB;
  static_assert(!WV || F::WAVE, "wave-owned transforms need whole wavefronts");
  const T s = T(0.5) / T(NXB);
  __syncthreads();                       // every thread has read the synthesis results
  if constexpr (PPT == 1) {
#pragma unroll
    for (int q = 0; q < NP; ++q) buf[q * padlen(NXB) + pidx(tid)] = cx<T>{xa[q], xb[q]};
  } else {
#pragma unroll
    for (int p = 0; p < PPT; ++p)
#pragma unroll
      for (int q = 0; q < NP; ++q)
        buf[q * padlen(NXB) + pidx(tid + p * (NXB / PPT))] = cx<T>{xa[p * NP + q], xb[p * NP + q]};
  }
  __syncthreads();
  auto extract = [&](int q, int k) {
    const cx<T> zk = buf[q * padlen(NXB) + pidx(k)];
    const cx<T> zn = buf[q * padlen(NXB) + pidx(k == 0 ? 0 : NXB - k)];
    emit(q, k, cx<T>{(zk.x + zn.x) * s, (zk.y - zn.y) * s}, cx<T>{(zk.y + zn.y) * s, (zn.x - zk.x) * s});
  };
  constexpr int RL = last_radix<T, NXB>();
  if constexpr (WV && F::TPW == 1 && NXB == 64 * RL && MX % 64 == 0) {
    // one transform per wave whose last stage has one butterfly per lane:
    // that stage stays in registers (lane j holds X[j + 64 r]); the mirror
    // coefficient X[NXB - k] of k = j + 64 kk is X[64 (RL-1-kk) + 64 - j], held
    // by lane 64 - j (k = 64 kk of lane 0: lane 0's own X[64 (RL-kk)], and
    // X[0] for k = 0)
    const int w = wave_of(tid), lane = tid & 63;
    if (w < NP) {
      cx<T>* z = buf + w * padlen(NXB);
      fft_run_src<T, NXB, false, 1, 64, 0, true, padlen(NXB), MX, n_stages(NXB) - 1>(z, tw_src(tw), lane);
      cx<T> v[RL];
      fft_last_regs<T, NXB, false>(z, tw_src(tw), lane, v);
      emit_from_regs<T, NXB, MX>(v, lane, s, w, emit);
    }
    // the next synth_pairs only rewrites this wave's own transforms
  } else if constexpr (WV && F::TPW == 1 && sizeof(T) == 4 && RL == 3 && NXB == 3 * MX && MX % 64 == 0) {
    // FP32 rows whose last stage is radix 3 over stride MX (768 = 8.8.4.3):
    // lane j keeps the outputs of its butterflies b = j + 64 i, X[b + MX r],
    // in registers.  Coefficient k = j + 64 kk is X[k] of butterfly k; its
    // mirror X[NXB - k] = X[2 MX + (MX - k)] is output 2 of butterfly MX - k
    // = 64 (KK-1-kk) + 64 - j, held by lane 64 - j (lane 0: its own butterfly
    // 64 (KK - kk), and X[0] for k = 0).
    static_assert(!tw_per_lane<TWA>(), "any butterfly's twiddles");
    constexpr int KK = MX / 64;
    const int w = wave_of(tid), lane = tid & 63;
    if (w < NP) {
      cx<T>* z = buf + w * padlen(NXB);
      fft_run_src<T, NXB, false, 1, 64, 0, true, padlen(NXB), MX, n_stages(NXB) - 1>(z, tw_src(tw), lane);
      cx<T> x[KK][3];
#pragma unroll
      for (int i = 0; i < KK; ++i) fft_last_regs<T, NXB, false>(z, tw_src(tw), lane + 64 * i, x[i]);
      const int src = (64 - lane) & 63;
#pragma unroll
      for (int kk = 0; kk < KK; ++kk) {
        const cx<T> y = x[KK - 1 - kk][2];
        const cx<T> m = {__shfl(y.x, src, 64), __shfl(y.y, src, 64)};
        const cx<T> zk = x[kk][0];
        const cx<T> zn = lane != 0 ? m : (kk == 0 ? zk : x[KK - kk < KK ? KK - kk : 0][2]);
        emit(w, lane + 64 * kk, cx<T>{(zk.x + zn.x) * s, (zk.y - zn.y) * s},
             cx<T>{(zk.y + zn.y) * s, (zn.x - zk.x) * s});
      }
    }
    // the next synth_pairs only rewrites this wave's own transforms
  } else if constexpr (WV) {
    static_assert(PPT == 1, "several points per thread: register last stage only");
    const int w = row_wave<T, NXB>(tid), lane = tid & 63;
    if (w * F::TPW < NP) {
      fft_run_src<T, NXB, false, F::TPW, 64, 0, true, padlen(NXB), MX>(buf + w * F::TPW * padlen(NXB), tw_src(tw),
                                                                     lane);
#pragma unroll
      for (int qq = 0; qq < F::TPW; ++qq) {
        const int q = w * F::TPW + qq;
        if (q < NP)
          for (int k = lane; k < MX; k += 64) extract(q, k);
      }
    }
    // the next synth_pairs only rewrites this wave's own transforms
  } else if constexpr (PPT == 1 && NP * (MX / 2) == F::NT && last_radix<T, NXB>() == 3 && MX % 2 == 0) {
    // workgroup-shared transforms whose last stage (radix 3, MX butterflies
    // j, outputs X[j + MX r]) runs in registers: thread (q, jj) takes
    // butterflies jj and its mirror MX - jj (MX/2 for jj = 0), which hold
    // both X[k] and X[NXB - k] for k = jj and k = MX - jj (X[NXB - k] =
    // X[2 MX + (MX - k)])
    static_assert(!std::is_class<TWA>::value, "workgroup-shared transforms read the LDS table");
    fft_run_src<T, NXB, false, NP, F::NT, 0, false, padlen(NXB), MX, n_stages(NXB) - 1>(buf, tw_src(tw), tid);
    __syncthreads();
    const int q = tid / (MX / 2), jj = tid - q * (MX / 2);
    const int j2 = jj == 0 ? MX / 2 : MX - jj;
    cx<T> x1[3], x2[3];
    fft_last_regs<T, NXB, false>(buf + q * padlen(NXB), tw_src(tw), jj, x1);
    fft_last_regs<T, NXB, false>(buf + q * padlen(NXB), tw_src(tw), j2, x2);
    auto put = [&](int k, cx<T> zk, cx<T> zn) {
      emit(q, k, cx<T>{(zk.x + zn.x) * s, (zk.y - zn.y) * s}, cx<T>{(zk.y + zn.y) * s, (zn.x - zk.x) * s});
    };
    if (jj == 0) {
      put(0, x1[0], x1[0]);
      put(MX / 2, x2[0], x2[2]);
    } else {
      put(jj, x1[0], x2[2]);
      put(j2, x2[0], x1[2]);
    }
    __syncthreads();
  } else {
    static_assert(PPT == 1, "several points per thread: register last stage only");
    static_assert(!std::is_class<TWA>::value, "workgroup-shared transforms read the LDS table");
    fft_run<T, NXB, false, NP, F::NT, 0, false, padlen(NXB), MX>(buf, tw, tid);
    if (tid < MX) {
#pragma unroll
      for (int q = 0; q < NP; ++q) extract(q, tid);
    }
    __syncthreads();
  }
}

// Half spectra that a wave fills into its synthesis transforms q = w*TPW + qq
// (modes k = lane + 64 kk < MX), loaded a level ahead so that their memory
// latency overlaps the work of the current level.
template <typename T, int MX, int TPW>
struct RowPre {
  static constexpr int KPL = (MX + 63) / 64;
  cx<T> a[TPW][KPL], b[TPW][KPL];
};

template <typename T, int MX, int TPW, int NP>
__device__ __forceinline__ void row_prefetch(RowPre<T, MX, TPW>& P, const cx<T>* const* sa,
                                             const cx<T>* const* sb, int tid) {
  const int w = wave_of(tid), lane = tid & 63;
#pragma unroll
  for (int qq = 0; qq < TPW; ++qq) {
    const int q = w * TPW + qq;
    const cx<T>* pa = pick<NP>(sa, q, (const cx<T>*)nullptr);
    const cx<T>* pb = pick<NP>(sb, q, (const cx<T>*)nullptr);
#pragma unroll
    for (int kk = 0; kk < RowPre<T, MX, TPW>::KPL; ++kk) {
      const int k = lane + 64 * kk;
      P.a[qq][kk] = (pa && k < MX) ? pa[k] : cx<T>{0, 0};
      P.b[qq][kk] = (pb && k < MX) ? pb[k] : cx<T>{0, 0};
    }
  }
}

// Two levels' synthesis of the lean sweep in one round on all six wavefronts:
// wave w runs pair w % 3 of the level its prefetched half spectra belong to
// (the round's first level for waves 0-2, the second for waves 3-5) into slot
// w (w < 3) or w + 3, so the second level's rows stay in slots 6..8 until the
// next iteration reads them.  The first member of pair q takes an i kx factor
// where da[q] says so.  Ends with a workgroup barrier.
template <typename T, int L, int MX, class TWA>
__device__ __forceinline__ void synth_two(cx<T>* buf, const RowPre<T, MX, 1>& P, const bool* da,
                                          const T* __restrict__ kxp, const TWA& tw, int tid) {
  static_assert(synth_in_regs<T, L, MX>(), "register first stage");
  const int w = wave_of(tid), lane = tid & 63;
  const int q = w < 3 ? w : w - 3;
  const bool dqa = pick<3>(da, q, false);
  cx<T> av[1][MX / 64], bv[1][MX / 64];
#pragma unroll
  for (int kk = 0; kk < MX / 64; ++kk) {
    const int k = lane + 64 * kk;
    cx<T> x = P.a[0][kk];
    const T kx = kxp[k];
    if (dqa) x = cx<T>{-kx * x.y, kx * x.x};
    av[0][kk] = x;
    bv[0][kk] = P.b[0][kk];
  }
  synth_regs<T, L, MX, 1, padlen(L)>(buf + (w < 3 ? w : w + 3) * padlen(L), av, bv, tw_src(tw), lane);
  __syncthreads();
}

// synth_pairs (wave-owned transforms of length L) from prefetched half
// spectra: the same spectrum, transform and result layout.
template <typename T, int L, int MX, int TPW, int NP, class TWA>
__device__ __forceinline__ void synth_pre(cx<T>* buf, const RowPre<T, MX, TPW>& P, const bool* da,
                                          const bool* db, T* ra, T* rb, const T* __restrict__ kxp,
                                          const TWA& tw, int tid) {
  const int w = wave_of(tid), lane = tid & 63;
  if constexpr (TPW == 1 && synth_in_regs<T, L, MX>()) {
    if (w < NP) {
      const bool dqa = pick<NP>(da, w, false), dqb = pick<NP>(db, w, false);
      cx<T> av[1][MX / 64], bv[1][MX / 64];
#pragma unroll
      for (int kk = 0; kk < MX / 64; ++kk) {
        const int k = lane + 64 * kk;
        cx<T> x = P.a[0][kk], y = P.b[0][kk];
        const T kx = kxp[k];
        if (dqa) x = cx<T>{-kx * x.y, kx * x.x};
        if (dqb) y = cx<T>{-kx * y.y, kx * y.x};
        av[0][kk] = x;
        bv[0][kk] = y;
      }
      synth_regs<T, L, MX, 1, padlen(L)>(buf + w * padlen(L), av, bv, tw_src(tw), lane);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      const cx<T> v = buf[q * padlen(L) + pidx(tid)];
      ra[q] = v.x;
      rb[q] = v.y;
    }
    return;
  }
#pragma unroll
  for (int qq = 0; qq < TPW; ++qq) {
    const int q = w * TPW + qq;
    if (q < NP) {
      cx<T>* z = buf + q * padlen(L);
      const bool dqa = pick<NP>(da, q, false), dqb = pick<NP>(db, q, false);
#pragma unroll
      for (int kk = 0; kk < RowPre<T, MX, TPW>::KPL; ++kk) {
        const int k = lane + 64 * kk;
        if (k < MX) {
          cx<T> av = P.a[qq][kk], bv = P.b[qq][kk];
          const T kx = kxp[k];
          if (dqa) av = cx<T>{-kx * av.y, kx * av.x};
          if (dqb) bv = cx<T>{-kx * bv.y, kx * bv.x};
          if (k == 0) {
            z[0] = cx<T>{av.x, bv.x};      // c2r ignores the imaginary part of the DC term
          } else {
            z[pidx(k)] = cx<T>{av.x - bv.y, av.y + bv.x};
            z[pidx(L - k)] = cx<T>{av.x + bv.y, bv.x - av.y};
          }
        }
      }
    }
  }
  wave_sync();
  if (w * TPW < NP)
    fft_run_src<T, L, true, TPW, 64, 0, true, padlen(L), MX>(buf + w * TPW * padlen(L), tw_src(tw), lane);
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    const cx<T> v = buf[q * padlen(L) + pidx(tid)];
    ra[q] = v.x;
    rb[q] = v.y;
  }
}

template <typename T, int NX>
__global__ void __launch_bounds__(3 * NX / 2) k_wall_plane(const cx<T>* __restrict__ Bs, size_t bstride,
                                                           const T* __restrict__ kxp,
                                                           const cx<T>* __restrict__ twNXB, int nz,
                                                           const WallScalars* __restrict__ ws,
                                                           double* part) {
  using F = RowFFT<T, NX>;
  constexpr int MX = F::MX, NXB = F::NXB, NT = F::NT, NYB = NXB;
  __shared__ cx<T> buf[6 * padlen(NXB)];
  __shared__ cx<T> tw[NXB];
  __shared__ double red[RedSize<NT, 8>::N];
  const int tid = threadIdx.x, y = blockIdx.x, wall = blockIdx.y;
  const int L = wall ? nz : 1;
  load_twiddles(tw, twNXB, NXB, tid, NT);
  __syncthreads();
  const size_t row = (size_t)L * NYB * MX + (size_t)y * MX;
  const cx<T>* sa[1] = {Bs + row};                      // u (field 0)
  const cx<T>* sb[1] = {Bs + 2 * bstride + row};        // v (field 2)
  const bool da[1] = {false}, db[1] = {false};
  T ra[1], rb[1];
  synth_pairs<T, NX, 1>(buf, sa, da, sb, db, ra, rb, kxp, tw, tid);
  const double mu = ws->mu[wall], mv = ws->mv[wall];
  const double s0 = sqrt(mu * mu + mv * mv);
  const T uT = ra[0], vT = rb[0];
  const double sp = (double)sqrt(uT * uT + vT * vT) - s0;   // lowerbc.f90:60 in field precision
  const double up = (double)uT - mu, vp = (double)vT - mv;
  double q[8] = {sp, sp * sp, up * up, vp * vp, sp * up, sp * vp, up, vp};
  block_sum<NT, 8>(q, red, tid);
  if (tid < 8) part[((size_t)wall * NYB + y) * 8 + tid] = q[0];
}

// ---------------------------------------------------------------------------
// The comp1 clamp of the expanded SGS TKE (comp1.f90:109,121).  Where e on
// the 3/2 grid falls below emin the reference carries on with max(e, emin):
// its K, advection and dissipation use the clamped values (so do the row
// kernels), and ex, ey are the 3/2-grid derivatives of the clamped field
// (derivs2.f90:20-21).  The row kernels differentiate the band-limited e, so
// for every marked face f the derivatives of the clamp residual
// d = max(e, emin) - e are added to the e-fluxes of the levels that read
// ex, ey of face f (level f; also nnz-1 for f = nnz-2, derivs2.f90:9,20-21):
//   F_x^e += 2 K dx(d),  F_y^e += 2 K dy(d)   (compsgs3.f90:20-35),
// with K at the level's face.  dx, dy are xderiv/yderiv on the 3/2 grid
// (k = 0 and the Nyquist mode dropped, filter_fftw.f90:138-248).
//   k_eclamp_rows  d and dx(d)                  grid NYB, one thread per point
//   k_eclamp_cols  dy(d), in place of d         grid NXB/16, 256 threads
//   k_eclamp_flux  the corrections, x-forward, added to the G_e and F_y^e
//                  rows of the analysis buffer   grid NYB
// Each pass returns at once when no face is marked.
// ---------------------------------------------------------------------------
template <typename T>
struct EclampArgs {
  const cx<T>* B;          // synthesis rows (e: field 6)
  size_t bstride;
  cx<T>* A;                // analysis rows (G_e: field 6, F_y^e: field 7)
  size_t astride;
  const int* dirty;
  T* DX;                   // dx(d) [face][NYB][NXB]
  T* DY;                   // d, then dy(d)
  const T* X;              // imodel 4: K at the faces
  const T* kxp;
  const cx<T>* twNXB;
  double kxb, kyb;         // 2 pi / xl, 2 pi / yl
  Consts c;
};

// Wavenumber of FFT index k of length n (k = 0 and n/2 dropped).
template <typename T>
__device__ __forceinline__ T deriv_k(int k, int n, double unit) {
  if (k > 0 && k < n / 2) return (T)(unit * k);
  if (k > n / 2) return (T)(unit * (k - n));
  return T(0);
}

template <typename T, int NX>
__global__ void __launch_bounds__(3 * NX / 2) k_eclamp_rows(EclampArgs<T> a) {
  using F = RowFFT<T, NX>;
  constexpr int MX = F::MX, NXB = F::NXB, NT = F::NT, NYB = NXB;
  if (!a.dirty[0]) return;
  __shared__ cx<T> buf[padlen(NXB)];
  __shared__ cx<T> tw[NXB];
  const int tid = threadIdx.x, y = blockIdx.x, nz = a.c.nz;
  const T emin = (T)a.c.emin;
  const size_t pp = (size_t)NYB * NXB;
  load_twiddles(tw, a.twNXB, NXB, tid, NT);
  for (int f = 1; f <= nz - 2; ++f) {
    if (!a.dirty[f]) continue;
    __syncthreads();
    const cx<T>* src = a.B + 6 * a.bstride + (size_t)f * NYB * MX + (size_t)y * MX;
    fill_pair<T, NX>(buf, src, false, nullptr, false, a.kxp, tid, NT);
    __syncthreads();
    fft_run<T, NXB, true, 1, NT>(buf, tw, tid);
    const T e = buf[pidx(tid)].x;
    const T d = (e < emin ? emin : e) - e;
    __syncthreads();
    buf[pidx(tid)] = cx<T>{d, T(0)};
    __syncthreads();
    fft_run<T, NXB, false, 1, NT>(buf, tw, tid);
    {
      const T kk = deriv_k<T>(tid, NXB, a.kxb);
      const cx<T> z = buf[pidx(tid)];
      buf[pidx(tid)] = cx<T>{-kk * z.y, kk * z.x};
    }
    __syncthreads();
    fft_run<T, NXB, true, 1, NT>(buf, tw, tid);
    const size_t o = (size_t)f * pp + (size_t)y * NXB + tid;
    a.DX[o] = buf[pidx(tid)].x / T(NXB);
    a.DY[o] = d;
  }
}

template <typename T, int NX>
__global__ void __launch_bounds__(256) k_eclamp_cols(EclampArgs<T> a) {
  constexpr int NXB = 3 * NX / 2, NYB = NXB, P = 8, NT = 256, PL = colpad(NYB);
  static_assert(NXB % (2 * P) == 0, "column pairs per workgroup");
  if (!a.dirty[0]) return;
  __shared__ cx<T> buf[P * PL];
  __shared__ cx<T> tw[NYB];
  const int tid = threadIdx.x, x0 = blockIdx.x * 2 * P, nz = a.c.nz;
  const size_t pp = (size_t)NYB * NXB;
  load_twiddles(tw, a.twNXB, NYB, tid, NT);                  // NYB == NXB
  for (int f = 1; f <= nz - 2; ++f) {
    if (!a.dirty[f]) continue;
    T* D = a.DY + (size_t)f * pp;
    __syncthreads();
    for (int i = tid; i < P * NYB; i += NT) {                  // columns x0+2q, x0+2q+1 as one complex
      const int q = i % P, yy = i / P;
      const T* r = D + (size_t)yy * NXB + x0 + 2 * q;
      buf[q * PL + pidx(yy)] = cx<T>{r[0], r[1]};
    }
    __syncthreads();
    fft_waves<T, NYB, false, P, NT / 64, PL>(buf, tw, tid);
    __syncthreads();
    for (int i = tid; i < P * NYB; i += NT) {
      const int q = i / NYB, k = i % NYB;
      const T kk = deriv_k<T>(k, NYB, a.kyb);
      const cx<T> z = buf[q * PL + pidx(k)];
      buf[q * PL + pidx(k)] = cx<T>{-kk * z.y, kk * z.x};
    }
    __syncthreads();
    fft_waves<T, NYB, true, P, NT / 64, PL>(buf, tw, tid);
    __syncthreads();
    const T s = T(1) / T(NYB);
    for (int i = tid; i < P * NYB; i += NT) {
      const int q = i % P, yy = i / P;
      const cx<T> v = buf[q * PL + pidx(yy)];
      T* r = D + (size_t)yy * NXB + x0 + 2 * q;
      r[0] = v.x * s;
      r[1] = v.y * s;
    }
  }
}

template <typename T, int NX>
__global__ void __launch_bounds__(3 * NX / 2) k_eclamp_flux(EclampArgs<T> a) {
  using F = RowFFT<T, NX>;
  constexpr int MX = F::MX, NXB = F::NXB, NT = F::NT, NYB = NXB;
  if (!a.dirty[0]) return;
  __shared__ cx<T> buf[padlen(NXB)];
  __shared__ cx<T> tw[NXB];
  const int tid = threadIdx.x, y = blockIdx.x, nz = a.c.nz;
  const T emin = (T)a.c.emin, alk = (T)a.c.alk, xkmax = (T)a.c.xkmax;
  const size_t pp = (size_t)NYB * NXB;
  const size_t plane = (size_t)NYB * MX;
  load_twiddles(tw, a.twNXB, NXB, tid, NT);
  for (int iz = 1; iz <= nz - 1; ++iz) {
    const int ie = (iz == nz - 1) ? nz - 2 : iz;
    if (ie < 1 || !a.dirty[ie]) continue;
    __syncthreads();
    T K;
    if (a.c.imodel == 4) {
      K = a.X[(size_t)iz * pp + (size_t)y * NXB + tid];
    } else {                                                    // complen.f90:20-21 on the clamped e
      const cx<T>* src = a.B + 6 * a.bstride + (size_t)iz * plane + (size_t)y * MX;
      fill_pair<T, NX>(buf, src, false, nullptr, false, a.kxp, tid, NT);
      __syncthreads();
      fft_run<T, NXB, true, 1, NT>(buf, tw, tid);
      T e = buf[pidx(tid)].x;
      e = e > emin ? e : emin;
      K = T(0.1) * alk * sqrt(e);
      K = K < xkmax ? K : xkmax;
      __syncthreads();
    }
    const size_t o = (size_t)ie * pp + (size_t)y * NXB + tid;
    buf[pidx(tid)] = cx<T>{T(2) * K * a.DX[o], T(2) * K * a.DY[o]};
    __syncthreads();
    fft_run<T, NXB, false, 1, NT>(buf, tw, tid);
    if (tid < MX) {                                             // as anal_pairs
      const T s = T(0.5) / T(NXB);
      const int k = tid;
      const cx<T> zk = buf[pidx(k)];
      const cx<T> zn = buf[pidx(k == 0 ? 0 : NXB - k)];
      const cx<T> fx = {(zk.x + zn.x) * s, (zk.y - zn.y) * s};
      const cx<T> fy = {(zk.y + zn.y) * s, (zn.x - zk.x) * s};
      const T kx = a.kxp[k];
      cx<T>* ge = a.A + 6 * a.astride + (size_t)iz * plane + (size_t)y * MX + k;
      cx<T>* fe = a.A + 7 * a.astride + (size_t)iz * plane + (size_t)y * MX + k;
      const cx<T> g = *ge, h = *fe;
      *ge = cx<T>{g.x - kx * fx.y, g.y + kx * fx.x};
      *fe = cx<T>{h.x + fy.x, h.y + fy.y};
    }
  }
}

// imodel 4 (complen.f90:13-17,22-23): the eddy diffusivity at faces 0..nnz
// from the face shear
//   S(f) = G(f) + G(f+1) + (uz(f) + wx(f))^2 + (vz(f) + wy(f))^2,
//   G(L) = ux^2 + vy^2 + wz^2 + 0.5 (uy + vx)^2 at level L,
// with the ghost values of derivsbot/derivs (derivs1.f90:29-38,
// derivs2.f90:22-32): G(0) = G(nnz+1) = 0, uz(0) = uz(1),
// uz(nnz) = uz(nnz-1), wx = wy = 0 at faces 0 and nnz; wz(L) =
// (w(L) - w(L-1))/dz with w = 0 at faces 0 and nnz.
// K = min((0.1 dsl)^2 sqrt(S), xkmax) into X[f][y][x].
// grid (NYB, chunks of c.zchunk faces), one thread per point.  A chunk
// starts one level below its first face, so every face it writes sees the
// levels on both sides.
template <typename T, int NX>
__global__ void __launch_bounds__(3 * NX / 2) k_shear(const cx<T>* __restrict__ Bs, size_t bstride,
                                                      const T* __restrict__ kxp,
                                                      const cx<T>* __restrict__ twNXB, Consts c, T* X) {
  using F = RowFFT<T, NX>;
  constexpr int MX = F::MX, NXB = F::NXB, NT = F::NT, NYB = NXB;
  __shared__ cx<T> buf[6 * padlen(NXB)];
  __shared__ cx<T> tw[NXB];
  const int tid = threadIdx.x, y = blockIdx.x;
  const int nz = c.nz;
  const int f0 = blockIdx.y * c.zchunk;
  const int f1 = (f0 + c.zchunk < nz + 1) ? f0 + c.zchunk : nz + 1;
  load_twiddles(tw, twNXB, NXB, tid, NT);
  __syncthreads();
  const size_t plane = (size_t)NYB * MX;
  auto row = [&](int fld, int L) -> const cx<T>* {
    return Bs + (size_t)fld * bstride + (size_t)L * plane + (size_t)y * MX;
  };
  auto face = [&](int fld, int L) -> const cx<T>* { return (L >= 1 && L <= nz - 1) ? row(fld, L) : nullptr; };
  const T dzm = (T)c.dzm;
  const T kc = (T)((0.1 * c.dsl) * (0.1 * c.dsl)), xkmax = (T)c.xkmax;
  T* Xo = X + (size_t)y * NXB + tid;
  const size_t pp = (size_t)NYB * NXB;
  auto emit = [&](int f, T S) {
    if (f >= f0 && f < f1) {
      const T k = kc * sqrt(S);
      Xo[(size_t)f * pp] = k < xkmax ? k : xkmax;
    }
  };
  enum { FU = 0, FUY, FV, FVY, FW, FWY };
  const int Ls = f0 - 1 > 1 ? f0 - 1 : 1;
  const int Le = f1 < nz ? f1 : nz;
  T Gp = 0, up = 0, vp = 0, wxp = 0, wyp = 0, wprev = 0;   // previous level
  for (int L = Ls; L <= Le; ++L) {
    T ra[5], rb[5];
    const cx<T>* sa[5] = {row(FU, L), row(FUY, L), row(FU, L), face(FW, L), face(FWY, L)};
    const cx<T>* sb[5] = {row(FV, L), row(FVY, L), row(FV, L), face(FW, L), nullptr};
    const bool da[5] = {true, false, false, false, false};
    const bool db[5] = {true, false, false, true, false};
    synth_pairs<T, NX, 5>(buf, sa, da, sb, db, ra, rb, kxp, tw, tid);
    const T ux = ra[0], vx = rb[0], uy = ra[1], vy = rb[1], u = ra[2], v = rb[2];
    const T w = ra[3], wx = rb[3], wy = ra[4];
    const T wz = (w - wprev) * dzm;
    const T sxy = uy + vx;
    const T G = ux * ux + vy * vy + wz * wz + T(0.5) * sxy * sxy;
    if (L >= 2) {
      const T uz = (u - up) * dzm, vz = (v - vp) * dzm;         // face L-1
      const T a1 = uz + wxp, a2 = vz + wyp;
      emit(L - 1, Gp + G + a1 * a1 + a2 * a2);
      if (L == 2) emit(0, Gp + uz * uz + vz * vz);
      if (L == nz) emit(nz, G + uz * uz + vz * vz);
    }
    Gp = G; up = u; vp = v; wxp = wx; wyp = wy; wprev = w;
    __syncthreads();                                            // buf is rewritten next level
  }
}

// LEAN: the learned closure on the nnx x nny grid supplies the interior SGS
// momentum stresses (Consts::sgs_momentum == 0) and k_divz_band the
// divergence, so the velocity gradients are not synthesised: three pairs per
// level instead of six (u, v; w, e; ex, ey).
// PRE: the synthesis half spectra are loaded a level ahead.  Measured and
// tuned for FP32 at nnx = 256, where it is held at four waves per SIMD (128
// VGPRs; the lean sweep spills 24 bytes per lane, the full one none; left
// alone the compiler takes 133-148 and three waves, which measured slower:
// lean 0.507 -> 0.662 ms).  Other shapes keep the direct loads: there the
// extra registers spill or cost a wave per SIMD.
template <typename T, int NX>
constexpr bool rows_prefetch() { return sizeof(T) == 4 && NX == 256; }

template <typename T, int NX, bool LEAN = false>
__global__ void __launch_bounds__(3 * NX / 2) __attribute__((amdgpu_waves_per_eu(rows_prefetch<T, NX>() ? 4 : 1)))
k_rows(RowArgs<T> a) {
  using F = RowFFT<T, NX>;
  // Wave-owned transforms where rows are whole wavefronts (the FP32 lean
  // sweep at nnx = 512 is k_rows_lean2, six wavefronts of two points each:
  // here three synthesis pairs would leave three quarters of twelve idle).
  constexpr bool WV = F::WAVE;
  constexpr int MX = F::MX, NXB = F::NXB, NT = F::NT;
  constexpr int NYB = 3 * NX / 2;  // square grids: nny == nnx
  // The lean prefetching sweep synthesises two levels per round (synth_two):
  // three more slots hold the second level's rows.
  constexpr bool TWO = LEAN && WV && rows_prefetch<T, NX>() && synth_in_regs<T, NXB, MX>();
  __shared__ cx<T> buf[(TWO ? 9 : 6) * padlen(NXB)];
  __shared__ cx<T> tw[NXB];
  __shared__ double red[RedSize<NT, 7>::N];
  const int tid = threadIdx.x;
  const int y = blockIdx.x;
  const Consts& c = a.c;
  const int nz = c.nz;
  const int z0 = 1 + blockIdx.y * c.zchunk;
  int z1 = z0 + c.zchunk;
  if (z1 > nz + 1) z1 = nz + 1;
  const size_t plane = (size_t)NYB * MX;
  const size_t rowoff = (size_t)y * MX;
  auto row = [&](int fld, int L) -> const cx<T>* {
    return a.B + (size_t)fld * a.bstride + (size_t)L * plane + rowoff;
  };
  auto clampF = [&](int F) { return F <= 0 ? 1 : (F >= nz ? nz - 1 : F); };
  enum { FU = 0, FUY, FV, FVY, FW, FWY, FE, FEY };
  load_twiddles(tw, a.twNXB, NXB, tid, NT);
  __syncthreads();                       // the table is read by every wave's transforms
  // The lean FP32 sweep at nnx 256 holds each lane's twiddles of its
  // wave-owned transform in registers, loaded once per kernel instead of once
  // per transform stage (12 values; lean sweep 0.632 -> 0.612 ms).  The full
  // sweep keeps the LDS table: the registers would spill there.
  constexpr bool RT = LEAN && WV && rows_prefetch<T, NX>();
  RegTw<T, NXB, F::TPW, 64> rtw;
  if constexpr (RT) rtw.init(tw, tid & 63);
  const auto& twx = [&]() -> const auto& {
    if constexpr (RT) return rtw;
    else return tw;
  }();

  const T emin = (T)c.emin;
  const T alk = (T)c.alk, xkmax = (T)c.xkmax;
  const T dzm = (T)c.dzm, todzm = (T)c.todzm, dzm2 = (T)c.dzm2, ceps = (T)c.ceps;
  const T ugal = (T)c.ugal, vgal = (T)c.vgal;
  const bool mn = c.mnout != 0;
  const bool sgs_mom = !LEAN && c.sgs_momentum != 0;

  auto xkk_of = [&](T e) {
    T v = T(0.1) * alk * sqrt(e);
    return v < xkmax ? v : xkmax;   // complen.f90:21
  };

  // window registers (levels iz-1, iz, iz+1 of u, v; faces of w, e)
  T um, u0, up, vm, v0, vp, wm, w0, wp, em, e0, ep, wxm, wym;
  {
    // initial window: u,v at z0-1, z0 ; w,e at z0-1, z0 ; wx,wy at face z0-1
    const cx<T>* sa[5];
    const cx<T>* sb[5];
    const bool da[5] = {false, false, false, false, true};
    const bool db[5] = {false, false, false, false, false};
    T ra[5], rb[5];
    const int L0 = z0 - 1, L1 = z0;
    sa[0] = (L0 >= 1 && L0 <= nz) ? row(FU, L0) : nullptr;
    sb[0] = (L0 >= 1 && L0 <= nz) ? row(FV, L0) : nullptr;
    sa[1] = row(FU, L1);
    sb[1] = row(FV, L1);
    sa[2] = (L0 >= 1 && L0 <= nz - 1) ? row(FW, L0) : nullptr;
    sb[2] = row(FE, clampF(L0));
    sa[3] = (L1 >= 1 && L1 <= nz - 1) ? row(FW, L1) : nullptr;
    sb[3] = row(FE, clampF(L1));
    sa[4] = (L0 >= 1 && L0 <= nz - 1) ? row(FW, L0) : nullptr;
    sb[4] = (L0 >= 1 && L0 <= nz - 1) ? row(FWY, L0) : nullptr;
    synth_pairs<T, NX, 5, WV>(buf, sa, da, sb, db, ra, rb, a.kxp, tw, tid);
    um = ra[0]; vm = rb[0];
    u0 = ra[1]; v0 = rb[1];
    wm = ra[2]; w0 = ra[3];
    if (rb[2] < emin) mark_clamp(a.dirty, clampF(L0));
    if (rb[3] < emin) mark_clamp(a.dirty, clampF(L1));
    em = rb[2] > emin ? rb[2] : emin;
    e0 = rb[3] > emin ? rb[3] : emin;
    wxm = ra[4]; wym = rb[4];
    __syncthreads();                     // all reads done before waves refill buf
  }

  // the six synthesis pairs of level iz: u,v at iz+1; w,e at face iz+1;
  // (du/dx, du/dy), (dv/dx, dv/dy), (dw/dx, dw/dy) at iz; (de/dx, de/dy) at
  // face iz -- except at iz = nnz-1, where ex, ey are not refreshed
  // (derivs2.f90:9,20-21) and face nnz-2 is reused.  The first member of pairs
  // 2..5 takes an i kx factor.
  auto level_srcs = [&](int iz, const cx<T>** sa, const cx<T>** sb) {
    const int Lp = iz + 1;
    sa[0] = (Lp <= nz) ? row(FU, Lp) : nullptr;
    sb[0] = (Lp <= nz) ? row(FV, Lp) : nullptr;
    sa[1] = (Lp <= nz - 1) ? row(FW, Lp) : nullptr;
    sb[1] = row(FE, clampF(Lp));
    sa[2] = row(FU, iz);
    sb[2] = row(FUY, iz);
    sa[3] = row(FV, iz);
    sb[3] = row(FVY, iz);
    sa[4] = (iz <= nz - 1) ? row(FW, iz) : nullptr;
    sb[4] = (iz <= nz - 1) ? row(FWY, iz) : nullptr;
    const int ie = (iz == nz - 1) ? nz - 2 : iz;
    sa[5] = (iz <= nz - 1) ? row(FE, ie) : nullptr;
    sb[5] = (iz <= nz - 1) ? row(FEY, ie) : nullptr;
  };
  // wave-owned row transforms take their half spectra a level ahead
  constexpr bool PRE = WV && rows_prefetch<T, NX>();
  constexpr int NPS = LEAN ? 3 : 6;                 // synthesis pairs per level
  RowPre<T, MX, F::TPW> P;
  auto prefetch = [&](int iz) {
    const cx<T>* sa[6];
    const cx<T>* sb[6];
    level_srcs(iz, sa, sb);
    if constexpr (LEAN) {                           // (u, v) (w, e) (ex, ey)
      sa[2] = sa[5];
      sb[2] = sb[5];
    }
    if constexpr (TWO) {                            // and those of level iz + 1 as pairs 3..5
      const cx<T>* ta[6];
      const cx<T>* tb[6];
      if (iz + 1 < z1) {
        level_srcs(iz + 1, ta, tb);
        ta[2] = ta[5];
        tb[2] = tb[5];
      } else {
        for (int q = 0; q < 3; ++q) ta[q] = tb[q] = nullptr;
      }
      for (int q = 0; q < 3; ++q) {
        sa[3 + q] = ta[q];
        sb[3 + q] = tb[q];
      }
      row_prefetch<T, MX, F::TPW, 6>(P, sa, sb, tid);
    } else {
      row_prefetch<T, MX, F::TPW, NPS>(P, sa, sb, tid);
    }
  };
  if constexpr (PRE) prefetch(z0);
  for (int iz = z0; iz < z1; ++iz) {
    T ux, uy, vx, vy, wx0, wy0, ex, ey;
    if constexpr (LEAN) {
      T ra[3], rb[3];
      const bool da[3] = {false, false, true};
      const bool db[3] = {false, false, false};
      if constexpr (TWO) {
        // a synthesis round every other level; the second level's rows wait
        // in slots 6..8
        const bool first = ((iz - z0) & 1) == 0;
        if (first) {
          synth_two<T, NXB, MX>(buf, P, da, a.kxp, twx, tid);
          if (iz + 2 < z1) prefetch(iz + 2);
        }
        const int base = first ? 0 : 6;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const cx<T> v = buf[(base + q) * padlen(NXB) + pidx(tid)];
          ra[q] = v.x;
          rb[q] = v.y;
        }
      } else if constexpr (PRE) {
        synth_pre<T, NXB, MX, F::TPW, 3>(buf, P, da, db, ra, rb, a.kxp, twx, tid);
        if (iz + 1 < z1) prefetch(iz + 1);
      } else {
        const cx<T>* sa[6];
        const cx<T>* sb[6];
        level_srcs(iz, sa, sb);
        sa[2] = sa[5];
        sb[2] = sb[5];
        synth_pairs<T, NX, 3, WV>(buf, sa, da, sb, db, ra, rb, a.kxp, tw, tid);
      }
      up = ra[0]; vp = rb[0];
      wp = ra[1];
      if (rb[1] < emin) mark_clamp(a.dirty, clampF(iz + 1));
      ep = rb[1] > emin ? rb[1] : emin;
      ux = uy = vx = vy = wx0 = wy0 = T(0);
      ex = ra[2]; ey = rb[2];
    } else {
      T ra[6], rb[6];
      const bool da[6] = {false, false, true, true, true, true};
      const bool db[6] = {false, false, false, false, false, false};
      if constexpr (PRE) {
        synth_pre<T, NXB, MX, F::TPW, 6>(buf, P, da, db, ra, rb, a.kxp, twx, tid);
        if (iz + 1 < z1) prefetch(iz + 1);
      } else {
        const cx<T>* sa[6];
        const cx<T>* sb[6];
        level_srcs(iz, sa, sb);
        synth_pairs<T, NX, 6, WV>(buf, sa, da, sb, db, ra, rb, a.kxp, tw, tid);
      }
      up = ra[0]; vp = rb[0];
      wp = ra[1];
      if (rb[1] < emin) mark_clamp(a.dirty, clampF(iz + 1));
      ep = rb[1] > emin ? rb[1] : emin;
      ux = ra[2]; uy = rb[2];
      vx = ra[3]; vy = rb[3];
      wx0 = ra[4]; wy0 = rb[4];
      ex = ra[5]; ey = rb[5];
    }
    // ---- pointwise tendency (compadv1, complen, compsgs1, compsgs3, sufsim)
    // R = d/dx Fx + d/dy Fy + H per component; the forward pairs are
    // (Fx_u,H_u) (Fy_u=Fx_v,H_v) (Fx_w,H_w) (Fy_v,Fy_w) (Fx_e,H_e) (Fy_e,0)
    T fa[6], fb[6];
    double sums[7];
    const bool top = (iz == nz);
    {
      T xm, x0, xp;
      if (c.imodel == 4) {                                        // K from k_shear
        const T* Xr = a.X + (size_t)y * NXB + tid;
        const size_t pp = (size_t)NYB * NXB;
        xm = Xr[(size_t)(iz - 1) * pp];
        x0 = Xr[(size_t)iz * pp];
        xp = top ? x0 : Xr[(size_t)(iz + 1) * pp];             // unused at the top
      } else {
        xm = xkk_of(em); x0 = xkk_of(e0); xp = xkk_of(ep);
      }
      const T xs = xm + x0;
      T tuu = -xs * ux;                                           // compsgs1.f90:41-43
      T tvv = -xs * vy;
      T tuv = -xs * (uy + vx) * T(0.5);
      const T wz0 = (w0 - wm) * dzm;                              // derivs: wz at level iz
      const T wzp = (wp - w0) * dzm;
      T tww0 = -xs * wz0;                                         // compsgs1.f90:26,48
      T twwp = -(x0 + xp) * wzp;
      T tuw0, tvw0, tuwm, tvwm;
      if (top) {                                                  // sufsim_upper
        wall_tau(a.ws, 1, u0, v0, tuw0, tvw0);
      } else {                                                    // compsgs1.f90:49-50
        tuw0 = -x0 * ((up - u0) * dzm + wx0);
        tvw0 = -x0 * ((vp - v0) * dzm + wy0);
      }
      if (iz == 1) {                                              // sufsim
        wall_tau(a.ws, 0, u0, v0, tuwm, tvwm);
      } else {
        tuwm = -xm * ((u0 - um) * dzm + wxm);
        tvwm = -xm * ((v0 - vm) * dzm + wym);
      }
      if (!sgs_mom) {            // the small-grid closure supplies them; the walls stay here
        tuu = tvv = tuv = tww0 = twwp = T(0);
        if (!top) tuw0 = tvw0 = T(0);
        if (iz != 1) tuwm = tvwm = T(0);
      }
      // compadv1.f90:19-51 and compsgs3.f90:40-59
      fa[0] = -(u0 * u0 + tuu);
      // the imposed gradient 0.016 (comp1.f90:152) is a (0,0)-mode constant: k_press adds it in FP64
      fb[0] = -((up + u0) * w0 - (um + u0) * wm) * todzm - (tuw0 - tuwm) * dzm;
      fa[1] = -(u0 * v0 + tuv);
      fb[1] = -((vp + v0) * w0 - (vm + v0) * wm) * todzm - (tvw0 - tvwm) * dzm;
      fa[3] = -(v0 * v0 + tvv);
      T Fxw = 0, Fyw = 0, Hw = 0, Fxe = 0, Fye = 0, He = 0;
      if (!top) {
        Fxw = -(T(0.5) * (u0 + up) * w0 + tuw0);
        Fyw = -(T(0.5) * (v0 + vp) * w0 + tvw0);
        const T wa = T(0.5) * (wp + w0), wb = T(0.5) * (w0 + wm);
        Hw = -(wa * wa - wb * wb) * dzm - (twwp - tww0) * dzm;
        // SGS TKE: advection (compadv1.f90:59-64), diffusion and dissipation
        // (compsgs3.f90:20-35)
        Fxe = -T(0.5) * e0 * (u0 + up) + T(2) * x0 * ex;
        Fye = -T(0.5) * e0 * (v0 + vp) + T(2) * x0 * ey;
        He = -(wp * ep - wm * em) * todzm
             + ((xp + x0) * (ep - e0) - (x0 + xm) * (e0 - em)) * dzm2 - ceps * e0 * sqrt(e0);
      }
      fa[2] = Fxw; fb[2] = Hw;
      fb[3] = Fyw;
      fa[4] = Fxe; fb[4] = He;
      fa[5] = Fye; fb[5] = T(0);
      sums[0] = 0.0;                   // divz (comp1.f90:145) comes from k_divz_band
      sums[1] = (double)tuu; sums[2] = (double)tvv; sums[3] = (double)tuv;
      sums[4] = (double)tww0; sums[5] = (double)tuw0; sums[6] = (double)tvw0;
    }
    // partial sums of the SGS stress means, when means are due (slot 0, divz,
    // stays unused: k_divz_band computes it from the band spectra)
    if constexpr (WV) {
      const int nq = mn ? 7 : 0;
      for (int q = 1; q < nq; ++q) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) sums[q] += __shfl_xor(sums[q], off, 64);
      }
      // the second level of a two-level round has no synthesis barrier
      // between the previous level's reading of the partials and this write
      if constexpr (TWO)
        if (mn && ((iz - z0) & 1)) __syncthreads();
      if ((tid & 63) == 0)
        for (int q = 1; q < nq; ++q) red[q * F::NW + (tid >> 6)] = sums[q];
    } else {
      if (mn) {
        block_sum<NT, 7>(sums, red, tid);
        if (tid < 7) a.part[(size_t)tid * a.pstride + (size_t)iz * NYB + y] = sums[0];
      }
    }
    {
      cx<T>* A = a.A;
      const size_t o = (size_t)iz * plane + rowoff;
      const size_t as = a.astride;
      const T* kxp = a.kxp;
      auto emit = [&](int q, int k, cx<T> fA, cx<T> fB) {
        const cx<T> ikx = {0, kxp[k]};
        switch (q) {
          case 0: A[0 * as + o + k] = cadd(cmul(ikx, fA), fB); break;          // Gu
          case 1: A[1 * as + o + k] = fA;                                       // Fyu
                  A[2 * as + o + k] = cadd(cmul(ikx, fA), fB); break;          // Gv
          case 2: if (!top) A[4 * as + o + k] = cadd(cmul(ikx, fA), fB); break; // Gw
          case 3: A[3 * as + o + k] = fA;                                       // Fyv
                  if (!top) A[5 * as + o + k] = fB; break;                     // Fyw
          case 4: if (!top) A[6 * as + o + k] = cadd(cmul(ikx, fA), fB); break; // Ge
          default: if (!top) A[7 * as + o + k] = fA; break;                    // Fye
        }
      };
      anal_pairs<T, NX, 6, WV>(buf, fa, fb, twx, tid, emit);
    }
    if constexpr (WV) {
      // wave partials were published by the barriers inside anal_pairs
      const int nq = mn ? 7 : 0;
      if (tid >= 1 && tid < nq) {
        double sq = 0;
        for (int w = 0; w < F::NW; ++w) sq += red[tid * F::NW + w];
        a.part[(size_t)tid * a.pstride + (size_t)iz * NYB + y] = sq;
      }
    }
    // slide the window
    um = u0; u0 = up;
    vm = v0; v0 = vp;
    wm = w0; w0 = wp;
    em = e0; e0 = ep;
    wxm = wx0; wym = wy0;
  }
}

// ---------------------------------------------------------------------------
// k_rows_lean2: the lean sweep of k_rows<T, NX, true> (the small-grid closure
// supplies the SGS momentum stresses; the walls and the TKE equation stay
// here) for FP32 rows of 768 points, two points per thread: six wavefronts
// per row, each owning whole transforms -- the synthesis pairs of two levels
// per round (synth_two) and one analysis pair per level -- where k_rows'
// twelve wavefronts would share every transform or leave half of them idle
// (5.76 -> 5.54 ms per launch at 512^2 x 256).  Same arguments, grid (NYB,
// z-chunks) and outputs as k_rows<T, NX, true>.
// ---------------------------------------------------------------------------
template <typename T, int NX>
__global__ void __launch_bounds__(3 * NX / 4) __attribute__((amdgpu_waves_per_eu(3)))
k_rows_lean2(RowArgs<T> a) {
  using F = RowFFT<T, NX>;
  constexpr int MX = F::MX, NXB = F::NXB, NT = NXB / 2, NW = NT / 64;
  constexpr int NYB = 3 * NX / 2;  // square grids: nny == nnx
  static_assert(sizeof(T) == 4 && NT % 64 == 0 && NW == 6 && synth_in_regs<T, NXB, MX>(),
                "six wave-owned FP32 transforms per level, first stage from registers");
  // buf: the first level's synthesis rows in slots 0..2, the second's in 6..8,
  // the analysis pairs in 0..5.  red: two sets of wave partials, alternating
  // by level (no barrier separates one level's reading of them from the next
  // level's writing).
  constexpr int RN = 7 * NW;
  __shared__ cx<T> buf[9 * padlen(NXB)];
  __shared__ cx<T> tw[NXB];
  __shared__ double red[2 * RN];
  const int tid = threadIdx.x;
  const int y = blockIdx.x;
  const Consts& c = a.c;
  const int nz = c.nz;
  const int z0 = 1 + blockIdx.y * c.zchunk;
  int z1 = z0 + c.zchunk;
  if (z1 > nz + 1) z1 = nz + 1;
  const size_t plane = (size_t)NYB * MX;
  const size_t rowoff = (size_t)y * MX;
  auto row = [&](int fld, int L) -> const cx<T>* {
    return a.B + (size_t)fld * a.bstride + (size_t)L * plane + rowoff;
  };
  auto clampF = [&](int F) { return F <= 0 ? 1 : (F >= nz ? nz - 1 : F); };
  enum { FU = 0, FUY, FV, FVY, FW, FWY, FE, FEY };
  load_twiddles(tw, a.twNXB, NXB, tid, NT);
  __syncthreads();                       // the table is read by every wave's transforms

  const T emin = (T)c.emin;
  const T alk = (T)c.alk, xkmax = (T)c.xkmax;
  const T dzm = (T)c.dzm, todzm = (T)c.todzm, dzm2 = (T)c.dzm2, ceps = (T)c.ceps;
  const bool mn = c.mnout != 0;
  auto xkk_of = [&](T e) {
    T v = T(0.1) * alk * sqrt(e);
    return v < xkmax ? v : xkmax;   // complen.f90:21
  };

  // window registers of the thread's points tid and tid + NT (levels iz-1,
  // iz, iz+1 of u, v; faces of w, e)
  T um[2], u0[2], up[2], vm[2], v0[2], vp[2], wm[2], w0[2], wp[2], em[2], e0[2], ep[2];
  {
    // initial window: u,v at z0-1, z0 ; w,e at faces z0-1, z0
    const cx<T>* sa[4];
    const cx<T>* sb[4];
    const bool da[4] = {false, false, false, false};
    const bool db[4] = {false, false, false, false};
    T ra[8], rb[8];
    const int L0 = z0 - 1, L1 = z0;
    sa[0] = (L0 >= 1 && L0 <= nz) ? row(FU, L0) : nullptr;
    sb[0] = (L0 >= 1 && L0 <= nz) ? row(FV, L0) : nullptr;
    sa[1] = row(FU, L1);
    sb[1] = row(FV, L1);
    sa[2] = (L0 >= 1 && L0 <= nz - 1) ? row(FW, L0) : nullptr;
    sb[2] = row(FE, clampF(L0));
    sa[3] = (L1 >= 1 && L1 <= nz - 1) ? row(FW, L1) : nullptr;
    sb[3] = row(FE, clampF(L1));
    synth_pairs<T, NX, 4, true, 2>(buf, sa, da, sb, db, ra, rb, a.kxp, tw, tid);
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      um[p] = ra[4 * p]; vm[p] = rb[4 * p];
      u0[p] = ra[4 * p + 1]; v0[p] = rb[4 * p + 1];
      wm[p] = ra[4 * p + 2]; w0[p] = ra[4 * p + 3];
      if (rb[4 * p + 2] < emin) mark_clamp(a.dirty, clampF(L0));
      if (rb[4 * p + 3] < emin) mark_clamp(a.dirty, clampF(L1));
      em[p] = rb[4 * p + 2] > emin ? rb[4 * p + 2] : emin;
      e0[p] = rb[4 * p + 3] > emin ? rb[4 * p + 3] : emin;
    }
    __syncthreads();                     // all reads done before waves refill buf
  }

  // the three synthesis pairs of level iz: (u, v) at iz+1, (w, e) at face
  // iz+1, (de/dx, de/dy) at face iz -- face nnz-2 at iz = nnz-1, where ex, ey
  // are not refreshed (derivs2.f90:9,20-21)
  auto level_srcs = [&](int iz, const cx<T>** sa, const cx<T>** sb) {
    const int Lp = iz + 1;
    sa[0] = (Lp <= nz) ? row(FU, Lp) : nullptr;
    sb[0] = (Lp <= nz) ? row(FV, Lp) : nullptr;
    sa[1] = (Lp <= nz - 1) ? row(FW, Lp) : nullptr;
    sb[1] = row(FE, clampF(Lp));
    const int ie = (iz == nz - 1) ? nz - 2 : iz;
    sa[2] = (iz <= nz - 1) ? row(FE, ie) : nullptr;
    sb[2] = (iz <= nz - 1) ? row(FEY, ie) : nullptr;
  };
  // the half spectra of a round: level iz's pairs for waves 0..2, level
  // iz+1's for waves 3..5.  They are loaded at the round's start: held in
  // registers a round ahead (as the nnx = 256 sweep does) they push 38
  // registers of the transforms into scratch (measured 6.40 against 5.54 ms
  // per launch at 512^2 x 256).
  RowPre<T, MX, 1> P;
  auto load_round = [&](int iz) {
    const cx<T>* sa[6];
    const cx<T>* sb[6];
    level_srcs(iz, sa, sb);
    if (iz + 1 < z1) {
      level_srcs(iz + 1, sa + 3, sb + 3);
    } else {
      for (int q = 3; q < 6; ++q) sa[q] = sb[q] = nullptr;
    }
    row_prefetch<T, MX, 1, 6>(P, sa, sb, tid);
  };
  const bool dsyn[3] = {false, false, true};
  for (int iz = z0; iz < z1; ++iz) {
    const bool first = ((iz - z0) & 1) == 0;
    if (first) {
      load_round(iz);
      synth_two<T, NXB, MX>(buf, P, dsyn, a.kxp, tw, tid);
    }
    const int base = first ? 0 : 6;
    T ex[2], ey[2];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int x = tid + p * NT;
      const cx<T> q0 = buf[base * padlen(NXB) + pidx(x)];
      const cx<T> q1 = buf[(base + 1) * padlen(NXB) + pidx(x)];
      const cx<T> q2 = buf[(base + 2) * padlen(NXB) + pidx(x)];
      up[p] = q0.x; vp[p] = q0.y;
      wp[p] = q1.x;
      if (q1.y < emin) mark_clamp(a.dirty, clampF(iz + 1));
      ep[p] = q1.y > emin ? q1.y : emin;
      ex[p] = q2.x; ey[p] = q2.y;
    }
    // ---- pointwise tendency (compadv1, complen, compsgs3, sufsim): the
    // forward pairs (Fx_u,H_u) (Fy_u=Fx_v,H_v) (Fx_w,H_w) (Fy_v,Fy_w)
    // (Fx_e,H_e) (Fy_e,0) of point p in fa[6p + q], fb[6p + q]
    T fa[12], fb[12];
    double sums[7] = {0, 0, 0, 0, 0, 0, 0};
    const bool top = (iz == nz);
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      T xm, x0, xp;
      if (c.imodel == 4) {                                        // K from k_shear
        const T* Xr = a.X + (size_t)y * NXB + tid + p * NT;
        const size_t pp = (size_t)NYB * NXB;
        xm = Xr[(size_t)(iz - 1) * pp];
        x0 = Xr[(size_t)iz * pp];
        xp = top ? x0 : Xr[(size_t)(iz + 1) * pp];             // unused at the top
      } else {
        xm = xkk_of(em[p]); x0 = xkk_of(e0[p]); xp = xkk_of(ep[p]);
      }
      // the SGS momentum stresses come from the small-grid closure except
      // at the walls (sufsim_upper at the top face, sufsim below level 1)
      T tuw0 = 0, tvw0 = 0, tuwm = 0, tvwm = 0;
      if (top) wall_tau(a.ws, 1, u0[p], v0[p], tuw0, tvw0);
      if (iz == 1) wall_tau(a.ws, 0, u0[p], v0[p], tuwm, tvwm);
      T* fap = fa + 6 * p;
      T* fbp = fb + 6 * p;
      // compadv1.f90:19-51 and compsgs3.f90:40-59
      fap[0] = -(u0[p] * u0[p]);
      // the imposed gradient 0.016 (comp1.f90:152) is a (0,0)-mode constant: k_press adds it in FP64
      fbp[0] = -((up[p] + u0[p]) * w0[p] - (um[p] + u0[p]) * wm[p]) * todzm - (tuw0 - tuwm) * dzm;
      fap[1] = -(u0[p] * v0[p]);
      fbp[1] = -((vp[p] + v0[p]) * w0[p] - (vm[p] + v0[p]) * wm[p]) * todzm - (tvw0 - tvwm) * dzm;
      fap[3] = -(v0[p] * v0[p]);
      T Fxw = 0, Fyw = 0, Hw = 0, Fxe = 0, Fye = 0, He = 0;
      if (!top) {
        Fxw = -(T(0.5) * (u0[p] + up[p]) * w0[p]);
        Fyw = -(T(0.5) * (v0[p] + vp[p]) * w0[p]);
        const T wa = T(0.5) * (wp[p] + w0[p]), wb = T(0.5) * (w0[p] + wm[p]);
        Hw = -(wa * wa - wb * wb) * dzm;
        // SGS TKE: advection (compadv1.f90:59-64), diffusion and dissipation
        // (compsgs3.f90:20-35)
        Fxe = -T(0.5) * e0[p] * (u0[p] + up[p]) + T(2) * x0 * ex[p];
        Fye = -T(0.5) * e0[p] * (v0[p] + vp[p]) + T(2) * x0 * ey[p];
        He = -(wp[p] * ep[p] - wm[p] * em[p]) * todzm +
             ((xp + x0) * (ep[p] - e0[p]) - (x0 + xm) * (e0[p] - em[p])) * dzm2 - ceps * e0[p] * sqrt(e0[p]);
      }
      fap[2] = Fxw; fbp[2] = Hw;
      fbp[3] = Fyw;
      fap[4] = Fxe; fbp[4] = He;
      fap[5] = Fye; fbp[5] = T(0);
      // SGS stress means uusb .. vwsb: here only the top wall's
      sums[5] += (double)tuw0;
      sums[6] += (double)tvw0;
    }
    const int rpar = (iz - z0) & 1 ? RN : 0;
    {
      const int nq = mn ? 7 : 0;
      for (int q = 1; q < nq; ++q) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) sums[q] += __shfl_xor(sums[q], off, 64);
      }
      if ((tid & 63) == 0)
        for (int q = 1; q < nq; ++q) red[rpar + q * NW + (tid >> 6)] = sums[q];
    }
    {
      cx<T>* A = a.A;
      const size_t o = (size_t)iz * plane + rowoff;
      const size_t as = a.astride;
      const T* kxp = a.kxp;
      auto emit = [&](int q, int k, cx<T> fA, cx<T> fB) {
        const cx<T> ikx = {0, kxp[k]};
        switch (q) {
          case 0: A[0 * as + o + k] = cadd(cmul(ikx, fA), fB); break;          // Gu
          case 1: A[1 * as + o + k] = fA;                                       // Fyu
                  A[2 * as + o + k] = cadd(cmul(ikx, fA), fB); break;          // Gv
          case 2: if (!top) A[4 * as + o + k] = cadd(cmul(ikx, fA), fB); break; // Gw
          case 3: A[3 * as + o + k] = fA;                                       // Fyv
                  if (!top) A[5 * as + o + k] = fB; break;                     // Fyw
          case 4: if (!top) A[6 * as + o + k] = cadd(cmul(ikx, fA), fB); break; // Ge
          default: if (!top) A[7 * as + o + k] = fA; break;                    // Fye
        }
      };
      anal_pairs<T, NX, 6, true, 2>(buf, fa, fb, tw, tid, emit);
    }
    {
      // wave partials were published by the barriers inside anal_pairs
      const int nq = mn ? 7 : 0;
      if (tid >= 1 && tid < nq) {
        double sq = 0;
        for (int w = 0; w < NW; ++w) sq += red[rpar + tid * NW + w];
        a.part[(size_t)tid * a.pstride + (size_t)iz * NYB + y] = sq;
      }
    }
    // slide the window
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      um[p] = u0[p]; u0[p] = up[p];
      vm[p] = v0[p]; v0[p] = vp[p];
      wm[p] = w0[p]; w0[p] = wp[p];
      em[p] = e0[p]; e0[p] = ep[p];
    }
  }
}

// ---------------------------------------------------------------------------
// k_rows_mlp: k_rows with the learned closure in place of the compsgs1
// stresses (no reference counterpart; insertion point compsgs1.f90:1-53).
//
// The network runs at every 3/2-grid cell centre.  Its inputs:
//   - ux uy vx vy at the level;
//   - uz, vz = (u(k+1)-u(k-1))/(2dz);
//   - wx, wy = the mean of the two faces;
//   - wz = (w(k)-w(k-1))/dz;
//   - dsl and the wall distance.
// A cell uses the deviatoric network stress when the prediction is realizable
// and dissipative, and the TKE stress otherwise.  A face takes the mean of its
// two cells' tau13/tau23 when both cells passed, and the TKE face stress
// otherwise.  The wall faces keep the wall model.
//
// The face iz and the w-flux at face iz need cell iz+1, so the window runs one
// level ahead of k_rows: iteration iz synthesises u, v at iz+2, the horizontal
// gradients and wx, wy at iz+1, and evaluates the network at cell iz+1.  Each
// chunk starts with one warm-up iteration (level z0-1) that emits nothing.
// Whole wavefronts are MFMA operands, so NXB must be a multiple of 64.
// ---------------------------------------------------------------------------
template <typename T, int NX>
__device__ __forceinline__ void rows_mlp_body(const RowArgs<T>& a, const mlp::Net* __restrict__ gnet,
                                              unsigned long long* counts) {
  using F = RowFFT<T, NX>;
  static_assert(F::WAVE, "the learned closure needs 3/2-grid rows of whole wavefronts");
  constexpr int MX = F::MX, NXB = F::NXB, NT = F::NT;
  constexpr int NYB = 3 * NX / 2;
  __shared__ cx<T> buf[6 * padlen(NXB)];
  __shared__ cx<T> tw[NXB];
  __shared__ double red[RedSize<NT, 7>::N];
  __shared__ mlp::Net net;
  const int tid = threadIdx.x, lane = tid & 63;
  const int y = blockIdx.x;
  const Consts& c = a.c;
  const int nz = c.nz;
  const int z0 = 1 + blockIdx.y * c.zchunk;
  int z1 = z0 + c.zchunk;
  if (z1 > nz + 1) z1 = nz + 1;
  const size_t plane = (size_t)NYB * MX;
  const size_t rowoff = (size_t)y * MX;
  auto row = [&](int fld, int L) -> const cx<T>* {
    return a.B + (size_t)fld * a.bstride + (size_t)L * plane + rowoff;
  };
  auto clampF = [&](int F) { return F <= 0 ? 1 : (F >= nz ? nz - 1 : F); };
  auto lvl = [&](int fld, int L) -> const cx<T>* { return (L >= 1 && L <= nz) ? row(fld, L) : nullptr; };
  auto face = [&](int fld, int L) -> const cx<T>* { return (L >= 1 && L <= nz - 1) ? row(fld, L) : nullptr; };
  enum { FU = 0, FUY, FV, FVY, FW, FWY, FE, FEY };
  load_twiddles(tw, a.twNXB, NXB, tid, NT);
  for (int q = tid; q < (int)(sizeof(mlp::Net) / 16); q += NT) ((uint4*)&net)[q] = ((const uint4*)gnet)[q];
  __syncthreads();                       // the tables are read by every wave

  const T emin = (T)c.emin;
  const T alk = (T)c.alk, xkmax = (T)c.xkmax;
  const T dzm = (T)c.dzm, todzm = (T)c.todzm, dzm2 = (T)c.dzm2, ceps = (T)c.ceps;
  const T ugal = (T)c.ugal, vgal = (T)c.vgal;
  const float dslf = (float)c.dsl, dzf = (float)c.dz, zlf = (float)c.zl;
  const bool mn = c.mnout != 0;
  unsigned long long nacc = 0;

  auto xkk_of = [&](T e) {
    T v = T(0.1) * alk * sqrt(e);
    return v < xkmax ? v : xkmax;   // complen.f90:21
  };
  // the network at a cell with gradients g -> stress t, accepted or not
  auto closure = [&](int k, const float (&g)[9], float (&t)[6]) -> bool {
    const float zc = ((float)k - 0.5f) * dzf;
    const float f[mlp::NIN] = {g[0], g[1], g[2], g[3], g[4], g[5], g[6], g[7], g[8], dslf,
                               fminf(zc, zlf - zc)};
    mlp::infer(&net, f, t, lane);
    return mlp::accept(t, g) && k >= 1 && k <= nz;
  };

  // window: u, v at iz-1..iz+2; w, e at faces iz-1..iz+1; horizontal
  // gradients at iz, iz+1; wx, wy at faces iz-1..iz+1
  T um, u0, up, upp, vm, v0, vp, vpp, wm, w0, wp, em, e0, ep;
  T ux0, uy0, vx0, vy0, ux1, uy1, vx1, vy1;
  T wxm, wym, wx0, wy0, wxp, wyp;
  float t0[6];
  bool ok0;
  T fuw = 0, fvw = 0;                   // final tau13, tau23 at face iz-1
  const int zs = z0 - 1;
  {
    const cx<T>* sa[6];
    const cx<T>* sb[6];
    const bool da[6] = {false, false, false, false, false, true};
    const bool db[6] = {false, false, false, false, false, false};
    T ra[6], rb[6];
    sa[0] = lvl(FU, zs - 1); sb[0] = lvl(FV, zs - 1);
    sa[1] = lvl(FU, zs);     sb[1] = lvl(FV, zs);
    sa[2] = lvl(FU, zs + 1); sb[2] = lvl(FV, zs + 1);
    sa[3] = face(FW, zs - 1); sb[3] = row(FE, clampF(zs - 1));
    sa[4] = face(FW, zs);     sb[4] = row(FE, clampF(zs));
    sa[5] = face(FW, zs - 1); sb[5] = face(FWY, zs - 1);
    synth_pairs<T, NX, 6>(buf, sa, da, sb, db, ra, rb, a.kxp, tw, tid);
    um = ra[0]; vm = rb[0];
    u0 = ra[1]; v0 = rb[1];
    up = ra[2]; vp = rb[2];
    wm = ra[3]; w0 = ra[4];
    if (rb[3] < emin) mark_clamp(a.dirty, clampF(zs - 1));
    if (rb[4] < emin) mark_clamp(a.dirty, clampF(zs));
    em = rb[3] > emin ? rb[3] : emin;
    e0 = rb[4] > emin ? rb[4] : emin;
    wxm = ra[5]; wym = rb[5];
    __syncthreads();
    const bool da2[3] = {true, true, true};
    const bool db2[3] = {false, false, false};
    sa[0] = lvl(FU, zs);  sb[0] = lvl(FUY, zs);
    sa[1] = lvl(FV, zs);  sb[1] = lvl(FVY, zs);
    sa[2] = face(FW, zs); sb[2] = face(FWY, zs);
    synth_pairs<T, NX, 3>(buf, sa, da2, sb, db2, ra, rb, a.kxp, tw, tid);
    ux0 = ra[0]; uy0 = rb[0];
    vx0 = ra[1]; vy0 = rb[1];
    wx0 = ra[2]; wy0 = rb[2];
    __syncthreads();
    const float g[9] = {(float)ux0, (float)uy0, (float)((up - um) * todzm), (float)vx0, (float)vy0,
                        (float)((vp - vm) * todzm), (float)(T(0.5) * (wxm + wx0)),
                        (float)(T(0.5) * (wym + wy0)), (float)((w0 - wm) * dzm)};
    ok0 = closure(zs, g, t0);
  }

  // the six synthesis pairs of iteration iz: u, v at iz+2; w, e at face
  // iz+1; the horizontal gradients of u, v at iz+1 and of w at face iz+1;
  // ex, ey at face iz (at iz = nnz-1 face nnz-2 again, derivs2.f90:9,20-21)
  auto level_srcs = [&](int iz, const cx<T>** sa, const cx<T>** sb) {
    sa[0] = lvl(FU, iz + 2);  sb[0] = lvl(FV, iz + 2);
    sa[1] = face(FW, iz + 1); sb[1] = row(FE, clampF(iz + 1));
    sa[2] = lvl(FU, iz + 1);  sb[2] = lvl(FUY, iz + 1);
    sa[3] = lvl(FV, iz + 1);  sb[3] = lvl(FVY, iz + 1);
    sa[4] = face(FW, iz + 1); sb[4] = face(FWY, iz + 1);
    const int ie = (iz == nz - 1) ? nz - 2 : iz;
    sa[5] = (iz >= 1 && iz <= nz - 1) ? row(FE, ie) : nullptr;
    sb[5] = (iz >= 1 && iz <= nz - 1) ? row(FEY, ie) : nullptr;
  };
  // the half spectra a level ahead, as in k_rows
  constexpr bool PRE = rows_prefetch<T, NX>();
  RowPre<T, MX, F::TPW> P;
  auto prefetch = [&](int iz) {
    const cx<T>* sa[6];
    const cx<T>* sb[6];
    level_srcs(iz, sa, sb);
    row_prefetch<T, MX, F::TPW, 6>(P, sa, sb, tid);
  };
  if constexpr (PRE) prefetch(zs);
  for (int iz = zs; iz < z1; ++iz) {
    T ex, ey;
    {
      T ra[6], rb[6];
      const bool da[6] = {false, false, true, true, true, true};
      const bool db[6] = {false, false, false, false, false, false};
      if constexpr (PRE) {
        synth_pre<T, NXB, MX, F::TPW, 6>(buf, P, da, db, ra, rb, a.kxp, tw, tid);
        if (iz + 1 < z1) prefetch(iz + 1);
      } else {
        const cx<T>* sa[6];
        const cx<T>* sb[6];
        level_srcs(iz, sa, sb);
        synth_pairs<T, NX, 6>(buf, sa, da, sb, db, ra, rb, a.kxp, tw, tid);
      }
      upp = ra[0]; vpp = rb[0];
      wp = ra[1];
      if (rb[1] < emin) mark_clamp(a.dirty, clampF(iz + 1));
      ep = rb[1] > emin ? rb[1] : emin;
      ux1 = ra[2]; uy1 = rb[2];
      vx1 = ra[3]; vy1 = rb[3];
      wxp = ra[4]; wyp = rb[4];
      ex = ra[5]; ey = rb[5];
    }
    float t1[6];
    bool ok1;
    {
      const float g[9] = {(float)ux1, (float)uy1, (float)((upp - u0) * todzm), (float)vx1, (float)vy1,
                          (float)((vpp - v0) * todzm), (float)(T(0.5) * (wx0 + wxp)),
                          (float)(T(0.5) * (wy0 + wyp)), (float)((wp - w0) * dzm)};
      ok1 = closure(iz + 1, g, t1);
    }
    const bool top = (iz == nz);
    const T xm = xkk_of(em), x0 = xkk_of(e0), xp = xkk_of(ep);
    // final face-iz stresses (carried to the next level as face iz-1)
    T tuw0, tvw0;
    if (top) {                                                    // sufsim_upper
      wall_tau(a.ws, 1, u0, v0, tuw0, tvw0);
    } else if (ok0 && ok1) {
      tuw0 = (T)(0.5f * (t0[4] + t1[4]));
      tvw0 = (T)(0.5f * (t0[5] + t1[5]));
    } else {                                                      // compsgs1.f90:49-50
      tuw0 = -x0 * ((up - u0) * dzm + wx0);
      tvw0 = -x0 * ((vp - v0) * dzm + wy0);
    }
    if (iz >= z0) {
      T fa[6], fb[6];
      double sums[7];
      {
        const T xs = xm + x0;
        T tuu = -xs * ux0;                                        // compsgs1.f90:41-43
        T tvv = -xs * vy0;
        T tuv = -xs * (uy0 + vx0) * T(0.5);
        const T wz0 = (w0 - wm) * dzm;
        const T wzp = (wp - w0) * dzm;
        T tww0 = -xs * wz0;                                       // compsgs1.f90:26,48
        T twwp = -(x0 + xp) * wzp;
        if (ok0) {
          const float tr = (t0[0] + t0[1] + t0[2]) * (1.0f / 3.0f);
          tuu = (T)(t0[0] - tr);
          tvv = (T)(t0[1] - tr);
          tuv = (T)t0[3];
          tww0 = (T)(t0[2] - tr);
        }
        if (ok1) twwp = (T)(t1[2] - (t1[0] + t1[1] + t1[2]) * (1.0f / 3.0f));
        T tuwm, tvwm;
        if (iz == 1) {                                            // sufsim
          wall_tau(a.ws, 0, u0, v0, tuwm, tvwm);
        } else {
          tuwm = fuw;
          tvwm = fvw;
        }
        // compadv1.f90:19-51 and compsgs3.f90:40-59
        fa[0] = -(u0 * u0 + tuu);
        // the imposed gradient 0.016 (comp1.f90:152) is a (0,0)-mode constant: k_press adds it in FP64
        fb[0] = -((up + u0) * w0 - (um + u0) * wm) * todzm - (tuw0 - tuwm) * dzm;
        fa[1] = -(u0 * v0 + tuv);
        fb[1] = -((vp + v0) * w0 - (vm + v0) * wm) * todzm - (tvw0 - tvwm) * dzm;
        fa[3] = -(v0 * v0 + tvv);
        T Fxw = 0, Fyw = 0, Hw = 0, Fxe = 0, Fye = 0, He = 0;
        if (!top) {
          Fxw = -(T(0.5) * (u0 + up) * w0 + tuw0);
          Fyw = -(T(0.5) * (v0 + vp) * w0 + tvw0);
          const T wa = T(0.5) * (wp + w0), wb = T(0.5) * (w0 + wm);
          Hw = -(wa * wa - wb * wb) * dzm - (twwp - tww0) * dzm;
          Fxe = -T(0.5) * e0 * (u0 + up) + T(2) * x0 * ex;
          Fye = -T(0.5) * e0 * (v0 + vp) + T(2) * x0 * ey;
          He = -(wp * ep - wm * em) * todzm
               + ((xp + x0) * (ep - e0) - (x0 + xm) * (e0 - em)) * dzm2 - ceps * e0 * sqrt(e0);
        }
        fa[2] = Fxw; fb[2] = Hw;
        fb[3] = Fyw;
        fa[4] = Fxe; fb[4] = He;
        fa[5] = Fye; fb[5] = T(0);
        sums[0] = 0.0;                 // divz (comp1.f90:145) comes from k_divz_band
        sums[1] = (double)tuu; sums[2] = (double)tvv; sums[3] = (double)tuv;
        sums[4] = (double)tww0; sums[5] = (double)tuw0; sums[6] = (double)tvw0;
      }
      const unsigned long long bal = __ballot(ok0);
      if (lane == 0) nacc += (unsigned long long)__popcll(bal);
      {
        const int nq = mn ? 7 : 0;                 // the SGS means; divz from k_divz_band
        for (int q = 1; q < nq; ++q) {
#pragma unroll
          for (int off = 32; off > 0; off >>= 1) sums[q] += __shfl_xor(sums[q], off, 64);
        }
        if (lane == 0)
          for (int q = 1; q < nq; ++q) red[q * F::NW + (tid >> 6)] = sums[q];
      }
      {
        cx<T>* A = a.A;
        const size_t o = (size_t)iz * plane + rowoff;
        const size_t as = a.astride;
        const T* kxp = a.kxp;
        auto emit = [&](int q, int k, cx<T> fA, cx<T> fB) {
          const cx<T> ikx = {0, kxp[k]};
          switch (q) {
            case 0: A[0 * as + o + k] = cadd(cmul(ikx, fA), fB); break;          // Gu
            case 1: A[1 * as + o + k] = fA;                                       // Fyu
                    A[2 * as + o + k] = cadd(cmul(ikx, fA), fB); break;          // Gv
            case 2: if (!top) A[4 * as + o + k] = cadd(cmul(ikx, fA), fB); break; // Gw
            case 3: A[3 * as + o + k] = fA;                                       // Fyv
                    if (!top) A[5 * as + o + k] = fB; break;                     // Fyw
            case 4: if (!top) A[6 * as + o + k] = cadd(cmul(ikx, fA), fB); break; // Ge
            default: if (!top) A[7 * as + o + k] = fA; break;                    // Fye
          }
        };
        anal_pairs<T, NX, 6>(buf, fa, fb, tw, tid, emit);
      }
      {
        const int nq = mn ? 7 : 0;
        if (tid >= 1 && tid < nq) {
          double sq = 0;
          for (int w = 0; w < F::NW; ++w) sq += red[tid * F::NW + w];
          a.part[(size_t)tid * a.pstride + (size_t)iz * NYB + y] = sq;
        }
      }
    } else {
      __syncthreads();                     // every thread has read the synthesis results
    }
    // slide the window
    um = u0; u0 = up; up = upp;
    vm = v0; v0 = vp; vp = vpp;
    wm = w0; w0 = wp;
    em = e0; e0 = ep;
    ux0 = ux1; uy0 = uy1; vx0 = vx1; vy0 = vy1;
    wxm = wx0; wym = wy0; wx0 = wxp; wy0 = wyp;
#pragma unroll
    for (int q = 0; q < 6; ++q) t0[q] = t1[q];
    ok0 = ok1;
    fuw = tuw0; fvw = tvw0;
  }
  if (lane == 0) {
    atomicAdd(counts, nacc);
    atomicAdd(counts + 1, 64ull * (unsigned long long)(z1 - z0));
  }
}

// (Forcing three waves per SIMD costs spills and measured no faster: the
// kernel is bound by the network's VALU/transcendental work, not latency.)
template <typename T, int NX>
__global__ void __launch_bounds__(3 * NX / 2) k_rows_mlp(RowArgs<T> a, const mlp::Net* __restrict__ gnet,
                                                         unsigned long long* counts) {
  rows_mlp_body<T, NX>(a, gnet, counts);
}

// ---------------------------------------------------------------------------
// Learned closure on the nnx x nny grid (no reference counterpart).
//
// The network and its TKE fallback run at the cell centres of the small grid.
// The inputs and the fallback formulas are those of k_rows_mlp, with K from
// the physical e at the faces (compsgs1.f90:41-50).  The row kernel emits
// the stress-divergence contributions to the tendencies as x-spectral rows:
//   G_u = i kx F^x + H^ and F^y, per component, with
//   u: Fx = -t11, Fy = -t12, H = -(t13(iz) - t13(iz-1))/dz;
//   v: Fx = -t12, Fy = -t22, H = -(t23(iz) - t23(iz-1))/dz;
//   w (face iz): Fx = -t13(iz), Fy = -t23(iz), H = -(t33(iz+1) - t33(iz))/dz.
// The wall faces 0 and nnz carry no stress here: the 3/2-grid sweep keeps
// the wall model (k_rows with sgs_momentum == 0).  k_sgs_cols_out then adds
// G + i ky F^y, y-transformed, to the comp1 tendencies.
// ---------------------------------------------------------------------------
template <typename T>
struct SgsRowArgs {
  const cx<T>* S;          // u, v, w, uy, vy, wy as x-spectral small rows [6][slot][NY][MX]
  size_t sstride;
  cx<T>* O;                // G_u, Fy_u, G_v, Fy_v, G_w, Fy_w [6][slot][NY][MX]
  const T* e;              // physical e [slot][NY][NX]
  const T* kxp;
  const cx<T>* twNX;
  double* sred;            // SGS means [6][slots] (added atomically) when mnout
  int slots;
  unsigned long long* counts;
  Consts c;
};

// Half spectra (k < MX) of two real rows -> the length-L spectrum whose
// inverse DFT is a + i b; modes MX .. L-MX are zero.
template <typename T, int L, int MX>
__device__ __forceinline__ void fill_pair_len(cx<T>* z, const cx<T>* sa, bool da, const cx<T>* sb,
                                              bool db, const T* __restrict__ kxp, int t, int nt) {
  for (int k = t; k < MX; k += nt) {
    cx<T> av = {0, 0}, bv = {0, 0};
    if (sa) av = sa[k];
    if (sb) bv = sb[k];
    const T kx = kxp[k];
    if (da) av = cx<T>{-kx * av.y, kx * av.x};
    if (db) bv = cx<T>{-kx * bv.y, kx * bv.x};
    if (k == 0) {
      z[0] = cx<T>{av.x, bv.x};
    } else {
      z[pidx(k)] = cx<T>{av.x - bv.y, av.y + bv.x};
      z[pidx(L - k)] = cx<T>{av.x + bv.y, bv.x - av.y};
    }
  }
  for (int i = MX + t; i <= L - MX; i += nt) z[pidx(i)] = cx<T>{0, 0};
}

template <typename T, int NX>
struct SmallRowFFT {
  static constexpr int L = NX, MX = NX / 2, NT = NX;
  static constexpr int NW = NX / 64;
  static constexpr int TPW = (6 + NW - 1) / NW;
};

template <typename T, int NX, int NP>
__device__ __forceinline__ void synth_small(cx<T>* buf, const cx<T>* const* sa, const bool* da,
                                            const cx<T>* const* sb, const bool* db, T* ra, T* rb,
                                            const T* __restrict__ kxp, const cx<T>* tw, int tid) {
  using F = SmallRowFFT<T, NX>;
  const int w = wave_of(tid), lane = tid & 63;
  if constexpr (synth_in_regs<T, NX, F::MX>()) {
    if (w * F::TPW < NP) {
      cx<T> av[F::TPW][F::MX / 64], bv[F::TPW][F::MX / 64];
#pragma unroll
      for (int qq = 0; qq < F::TPW; ++qq) {
        const int q = w * F::TPW + qq;
        const cx<T>* pa = pick<NP>(sa, q, (const cx<T>*)nullptr);
        const cx<T>* pb = pick<NP>(sb, q, (const cx<T>*)nullptr);
        const bool dqa = pick<NP>(da, q, false), dqb = pick<NP>(db, q, false);
#pragma unroll
        for (int kk = 0; kk < F::MX / 64; ++kk) {
          const int k = lane + 64 * kk;
          cx<T> x = pa ? pa[k] : cx<T>{T(0), T(0)};
          cx<T> y = pb ? pb[k] : cx<T>{T(0), T(0)};
          const T kx = kxp[k];
          if (dqa) x = cx<T>{-kx * x.y, kx * x.x};
          if (dqb) y = cx<T>{-kx * y.y, kx * y.x};
          av[qq][kk] = x;
          bv[qq][kk] = y;
        }
      }
      synth_regs<T, NX, F::MX, F::TPW, padlen(NX)>(buf + w * F::TPW * padlen(NX), av, bv, tw_src(tw), lane);
    }
  } else {
#pragma unroll
    for (int qq = 0; qq < F::TPW; ++qq) {
      const int q = w * F::TPW + qq;
      if (q < NP) fill_pair_len<T, NX, F::MX>(buf + q * padlen(NX), sa[q], da[q], sb[q], db[q], kxp, lane, 64);
    }
    wave_sync();
    if (w * F::TPW < NP) fft_run<T, NX, true, F::TPW, 64, 0, true>(buf + w * F::TPW * padlen(NX), tw, lane);
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    const cx<T> v = buf[q * padlen(NX) + pidx(tid)];
    ra[q] = v.x;
    rb[q] = v.y;
  }
}

// The NP forward transforms go TPWA to a wave (NP <= NW * TPWA): with four
// pairs and four waves, one each.
// The half spectra this lane fills into the synthesis transforms
// q = w*TPW + qq (modes k = lane + 64 kk), loaded a level ahead so that their
// memory latency overlaps the network.
template <typename T, int NX>
struct SmallPre {
  static constexpr int TPW = SmallRowFFT<T, NX>::TPW, KPL = (NX / 2 + 63) / 64;
  cx<T> a[TPW][KPL], b[TPW][KPL];
};

template <typename T, int NX, int NP>
__device__ __forceinline__ void small_prefetch(SmallPre<T, NX>& P, const cx<T>* const* sa,
                                               const cx<T>* const* sb, int tid) {
  using F = SmallRowFFT<T, NX>;
  const int w = wave_of(tid), lane = tid & 63;
#pragma unroll
  for (int qq = 0; qq < F::TPW; ++qq) {
    const int q = w * F::TPW + qq;
    const cx<T>* pa = pick<NP>(sa, q, (const cx<T>*)nullptr);
    const cx<T>* pb = pick<NP>(sb, q, (const cx<T>*)nullptr);
#pragma unroll
    for (int kk = 0; kk < SmallPre<T, NX>::KPL; ++kk) {
      const int k = lane + 64 * kk;
      P.a[qq][kk] = (pa && k < F::MX) ? pa[k] : cx<T>{0, 0};
      P.b[qq][kk] = (pb && k < F::MX) ? pb[k] : cx<T>{0, 0};
    }
  }
}

// synth_small from prefetched half spectra.
template <typename T, int NX, int NP>
__device__ __forceinline__ void synth_small_pre(cx<T>* buf, const SmallPre<T, NX>& P, const bool* da,
                                                const bool* db, T* ra, T* rb, const T* __restrict__ kxp,
                                                const cx<T>* tw, int tid) {
  using F = SmallRowFFT<T, NX>;
  constexpr int L = NX, MX = F::MX;
  const int w = wave_of(tid), lane = tid & 63;
  if constexpr (synth_in_regs<T, L, MX>()) {
    if (w * F::TPW < NP) {
      cx<T> av[F::TPW][MX / 64], bv[F::TPW][MX / 64];
#pragma unroll
      for (int qq = 0; qq < F::TPW; ++qq) {
        const int q = w * F::TPW + qq;
        const bool dqa = pick<NP>(da, q, false), dqb = pick<NP>(db, q, false);
#pragma unroll
        for (int kk = 0; kk < MX / 64; ++kk) {
          const int k = lane + 64 * kk;
          cx<T> x = P.a[qq][kk], y = P.b[qq][kk];
          const T kx = kxp[k];
          if (dqa) x = cx<T>{-kx * x.y, kx * x.x};
          if (dqb) y = cx<T>{-kx * y.y, kx * y.x};
          av[qq][kk] = x;
          bv[qq][kk] = y;
        }
      }
      synth_regs<T, L, MX, F::TPW, padlen(NX)>(buf + w * F::TPW * padlen(NX), av, bv, tw_src(tw), lane);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      const cx<T> v = buf[q * padlen(NX) + pidx(tid)];
      ra[q] = v.x;
      rb[q] = v.y;
    }
    return;
  }
#pragma unroll
  for (int qq = 0; qq < F::TPW; ++qq) {
    const int q = w * F::TPW + qq;
    if (q < NP) {
      cx<T>* z = buf + q * padlen(NX);
      const bool dqa = pick<NP>(da, q, false), dqb = pick<NP>(db, q, false);
#pragma unroll
      for (int kk = 0; kk < SmallPre<T, NX>::KPL; ++kk) {
        const int k = lane + 64 * kk;
        if (k < MX) {
          cx<T> av = P.a[qq][kk], bv = P.b[qq][kk];
          const T kx = kxp[k];
          if (dqa) av = cx<T>{-kx * av.y, kx * av.x};
          if (dqb) bv = cx<T>{-kx * bv.y, kx * bv.x};
          if (k == 0) {
            z[0] = cx<T>{av.x, bv.x};
          } else {
            z[pidx(k)] = cx<T>{av.x - bv.y, av.y + bv.x};
            z[pidx(L - k)] = cx<T>{av.x + bv.y, bv.x - av.y};
          }
        }
      }
      if (lane == 0) z[pidx(MX)] = cx<T>{0, 0};
    }
  }
  wave_sync();
  if (w * F::TPW < NP) fft_run<T, NX, true, F::TPW, 64, 0, true>(buf + w * F::TPW * padlen(NX), tw, lane);
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    const cx<T> v = buf[q * padlen(NX) + pidx(tid)];
    ra[q] = v.x;
    rb[q] = v.y;
  }
}

template <typename T, int NX, int NP, typename Emit>
__device__ __forceinline__ void anal_small(cx<T>* buf, const T* xa, const T* xb, const cx<T>* tw, int tid,
                                           Emit emit) {
  using F = SmallRowFFT<T, NX>;
  constexpr int MX = F::MX;
  constexpr int TPWA = (NP + F::NW - 1) / F::NW;
  static_assert(F::NW * TPWA <= 6 || NP <= F::NW, "transforms stay inside the 6-slot buffer");
  const T s = T(0.5) / T(NX);
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NP; ++q) buf[q * padlen(NX) + pidx(tid)] = cx<T>{xa[q], xb[q]};
  __syncthreads();
  const int w = wave_of(tid), lane = tid & 63;
  constexpr int RL = last_radix<T, NX>();
  if constexpr (TPWA == 1 && NX == 64 * RL && MX % 64 == 0) {
    // the last stage stays in registers, as in anal_pairs
    if (w < NP) {
      cx<T>* z = buf + w * padlen(NX);
      fft_run_src<T, NX, false, 1, 64, 0, true, padlen(NX), 0, n_stages(NX) - 1>(z, tw_src(tw), lane);
      cx<T> v[RL];
      fft_last_regs<T, NX, false>(z, tw_src(tw), lane, v);
      emit_from_regs<T, NX, MX>(v, lane, s, w, emit);
    }
  } else if (w * TPWA < NP) {
    fft_run<T, NX, false, TPWA, 64, 0, true>(buf + w * TPWA * padlen(NX), tw, lane);
#pragma unroll
    for (int qq = 0; qq < TPWA; ++qq) {
      const int q = w * TPWA + qq;
      if (q < NP)
        for (int k = lane; k < MX; k += 64) {
          const cx<T> zk = buf[q * padlen(NX) + pidx(k)];
          const cx<T> zn = buf[q * padlen(NX) + pidx(k == 0 ? 0 : NX - k)];
          emit(q, k, cx<T>{(zk.x + zn.x) * s, (zk.y - zn.y) * s}, cx<T>{(zk.y + zn.y) * s, (zn.x - zk.x) * s});
        }
    }
  }
}

// Two waves per SIMD: left to itself the compiler takes 261 registers for a
// 256-thread block (one wave per SIMD, MFMA and LDS latencies exposed); three
// waves spill.
template <typename T, int NX>
__global__ void __launch_bounds__(NX) __attribute__((amdgpu_waves_per_eu(2)))
k_sgs_rows(SgsRowArgs<T> a, const mlp::Net* __restrict__ gnet) {
  using F = SmallRowFFT<T, NX>;
  static_assert(NX % 64 == 0, "whole wavefronts of points");
  constexpr int MX = F::MX, NT = F::NT, NY = NX;
  __shared__ cx<T> buf[6 * padlen(NX)];
  __shared__ cx<T> tw[NX];
  __shared__ mlp::Net net;
  __shared__ double red[6 * (NX / 64)];
  const int tid = threadIdx.x, lane = tid & 63;
  const int y = blockIdx.x;
  const Consts& c = a.c;
  const int nz = c.nz;
  const int z0 = 1 + blockIdx.y * c.zchunk;
  int z1 = z0 + c.zchunk;
  if (z1 > nz + 1) z1 = nz + 1;
  const size_t plane = (size_t)NY * MX;
  const size_t rowoff = (size_t)y * MX;
  const size_t pplane = (size_t)NY * NX;
  enum { SU = 0, SV, SW, SUY, SVY, SWY };
  auto row = [&](int fld, int L) -> const cx<T>* {
    return a.S + (size_t)fld * a.sstride + (size_t)L * plane + rowoff;
  };
  auto lvl = [&](int fld, int L) -> const cx<T>* { return (L >= 1 && L <= nz) ? row(fld, L) : nullptr; };
  auto face = [&](int fld, int L) -> const cx<T>* { return (L >= 1 && L <= nz - 1) ? row(fld, L) : nullptr; };
  // e at face f, with e(f <= 0) = e(1) and e(f >= nnz) = e(nnz-1) (lower/upper)
  auto e_at = [&](int f) -> T {
    const int g = f <= 0 ? 1 : (f >= nz ? nz - 1 : f);
    return a.e[(size_t)g * pplane + (size_t)y * NX + tid];
  };
  load_twiddles(tw, a.twNX, NX, tid, NT);
  for (int q = tid; q < (int)(sizeof(mlp::Net) / 16); q += NT) ((uint4*)&net)[q] = ((const uint4*)gnet)[q];
  __syncthreads();                       // the tables are read by every wave

  const T alk = (T)c.alk, xkmax = (T)c.xkmax;
  const T dzm = (T)c.dzm, todzm = (T)c.todzm;
  const float dslf = (float)c.dsl, dzf = (float)c.dz, zlf = (float)c.zl;
  const bool mn = c.mnout != 0;
  unsigned long long nacc = 0;
  auto xkk_of = [&](T e) {
    T v = T(0.1) * alk * sqrt(e);
    return v < xkmax ? v : xkmax;   // complen.f90:21
  };
  auto closure = [&](int k, const float (&g)[9], float (&t)[6]) -> bool {
    const float zc = ((float)k - 0.5f) * dzf;
    const float f[mlp::NIN] = {g[0], g[1], g[2], g[3], g[4], g[5], g[6], g[7], g[8], dslf,
                               fminf(zc, zlf - zc)};
    mlp::infer(&net, f, t, lane);
    return mlp::accept(t, g) && k >= 1 && k <= nz;
  };

  T um, u0, up, upp, vm, v0, vp, vpp, wm, w0, wp, em, e0, ep;
  T ux0, uy0, vx0, vy0, ux1, uy1, vx1, vy1;
  T wxm, wym, wx0, wy0, wxp, wyp;
  float t0[6];
  bool ok0;
  T fuw = 0, fvw = 0;                       // final t13, t23 at face iz-1
  const int zs = z0 - 1;
  {
    const cx<T>* sa[6];
    const cx<T>* sb[6];
    const bool da[6] = {false, false, false, false, false, false};
    const bool db[6] = {false, false, false, true, true, false};
    T ra[6], rb[6];
    sa[0] = lvl(SU, zs - 1); sb[0] = lvl(SV, zs - 1);
    sa[1] = lvl(SU, zs);     sb[1] = lvl(SV, zs);
    sa[2] = lvl(SU, zs + 1); sb[2] = lvl(SV, zs + 1);
    sa[3] = face(SW, zs - 1); sb[3] = face(SW, zs - 1);     // w, wx at face zs-1
    sa[4] = face(SW, zs);     sb[4] = face(SW, zs);         // w, wx at face zs
    sa[5] = face(SWY, zs - 1); sb[5] = face(SWY, zs);       // wy at faces zs-1, zs
    synth_small<T, NX, 6>(buf, sa, da, sb, db, ra, rb, a.kxp, tw, tid);
    um = ra[0]; vm = rb[0];
    u0 = ra[1]; v0 = rb[1];
    up = ra[2]; vp = rb[2];
    wm = ra[3]; wxm = rb[3];
    w0 = ra[4]; wx0 = rb[4];
    wym = ra[5]; wy0 = rb[5];
    __syncthreads();
    const bool da2[2] = {true, false};
    const bool db2[2] = {true, false};
    sa[0] = lvl(SU, zs);  sb[0] = lvl(SV, zs);               // ux, vx
    sa[1] = lvl(SUY, zs); sb[1] = lvl(SVY, zs);              // uy, vy
    synth_small<T, NX, 2>(buf, sa, da2, sb, db2, ra, rb, a.kxp, tw, tid);
    ux0 = ra[0]; vx0 = rb[0];
    uy0 = ra[1]; vy0 = rb[1];
    __syncthreads();
    em = e_at(zs - 1);
    e0 = e_at(zs);
    const float g[9] = {(float)ux0, (float)uy0, (float)((up - um) * todzm), (float)vx0, (float)vy0,
                        (float)((vp - vm) * todzm), (float)(T(0.5) * (wxm + wx0)),
                        (float)(T(0.5) * (wym + wy0)), (float)((w0 - wm) * dzm)};
    ok0 = closure(zs, g, t0);
  }

  // the synthesis sources of iteration iz: u, v at level iz+2; ux, vx, uy, vy
  // at iz+1; w, wx, wy at face iz+1
  auto prefetch = [&](SmallPre<T, NX>& P, int iz) {
    const cx<T>* sa[5];
    const cx<T>* sb[5];
    sa[0] = lvl(SU, iz + 2);  sb[0] = lvl(SV, iz + 2);
    sa[1] = lvl(SU, iz + 1);  sb[1] = lvl(SV, iz + 1);      // ux, vx
    sa[2] = lvl(SUY, iz + 1); sb[2] = lvl(SVY, iz + 1);     // uy, vy
    sa[3] = face(SW, iz + 1); sb[3] = face(SW, iz + 1);     // w, wx
    sa[4] = face(SWY, iz + 1); sb[4] = nullptr;             // wy
    small_prefetch<T, NX, 5>(P, sa, sb, tid);
  };
  SmallPre<T, NX> P;
  prefetch(P, zs);
  T epre = e_at(zs + 1);
  for (int iz = zs; iz < z1; ++iz) {
    {
      T ra[5], rb[5];
      const bool da[5] = {false, true, false, false, false};
      const bool db[5] = {false, true, false, true, false};
      synth_small_pre<T, NX, 5>(buf, P, da, db, ra, rb, a.kxp, tw, tid);
      upp = ra[0]; vpp = rb[0];
      ux1 = ra[1]; vx1 = rb[1];
      uy1 = ra[2]; vy1 = rb[2];
      wp = ra[3]; wxp = rb[3];
      wyp = ra[4];
    }
    ep = epre;
    if (iz + 1 < z1) {                     // next level's loads fly during the network
      prefetch(P, iz + 1);
      epre = e_at(iz + 2);
    }
    float t1[6];
    bool ok1;
    {
      const float g[9] = {(float)ux1, (float)uy1, (float)((upp - u0) * todzm), (float)vx1, (float)vy1,
                          (float)((vpp - v0) * todzm), (float)(T(0.5) * (wx0 + wxp)),
                          (float)(T(0.5) * (wy0 + wyp)), (float)((wp - w0) * dzm)};
      ok1 = closure(iz + 1, g, t1);
    }
    const bool top = (iz == nz);
    const T xm = xkk_of(em), x0 = xkk_of(e0), xp = xkk_of(ep);
    // final stresses of face iz (the wall faces carry none here)
    T f13 = 0, f23 = 0;
    if (iz >= 1 && iz <= nz - 1) {
      if (ok0 && ok1) {
        f13 = (T)(0.5f * (t0[4] + t1[4]));
        f23 = (T)(0.5f * (t0[5] + t1[5]));
      } else {                                                    // compsgs1.f90:49-50
        f13 = -x0 * ((up - u0) * dzm + wx0);
        f23 = -x0 * ((vp - v0) * dzm + wy0);
      }
    }
    if (iz >= z0) {
      const T xs = xm + x0;
      T c11 = -xs * ux0, c22 = -xs * vy0, c12 = -xs * (uy0 + vx0) * T(0.5);   // compsgs1.f90:41-43
      T c33 = -xs * (w0 - wm) * dzm;
      T c33p = -(x0 + xp) * (wp - w0) * dzm;
      if (ok0) {
        const float tr = (t0[0] + t0[1] + t0[2]) * (1.0f / 3.0f);
        c11 = (T)(t0[0] - tr);
        c22 = (T)(t0[1] - tr);
        c12 = (T)t0[3];
        c33 = (T)(t0[2] - tr);
      }
      if (ok1) c33p = (T)(t1[2] - (t1[0] + t1[1] + t1[2]) * (1.0f / 3.0f));
      const T fm13 = (iz == 1) ? T(0) : fuw, fm23 = (iz == 1) ? T(0) : fvw;
      T fa[4], fb[4];
      fa[0] = -c11; fb[0] = -(f13 - fm13) * dzm;                  // (Fx_u, H_u)
      fa[1] = -c12; fb[1] = -(f23 - fm23) * dzm;                  // (Fy_u = Fx_v, H_v)
      fa[2] = top ? T(0) : -f13;                                  // (Fx_w, H_w) at face iz
      fb[2] = top ? T(0) : -(c33p - c33) * dzm;
      fa[3] = -c22; fb[3] = top ? T(0) : -f23;                    // (Fy_v, Fy_w)
      const unsigned long long bal = __ballot(ok0);
      if (lane == 0) nacc += (unsigned long long)__popcll(bal);
      if (mn) {
        double s[6] = {(double)c11, (double)c22, (double)c12, (double)c33, (double)f13, (double)f23};
#pragma unroll
        for (int q = 0; q < 6; ++q)
#pragma unroll
          for (int off = 32; off > 0; off >>= 1) s[q] += __shfl_xor(s[q], off, 64);
        if (lane == 0)
          for (int q = 0; q < 6; ++q) red[q * (NX / 64) + (tid >> 6)] = s[q];
      }
      cx<T>* O = a.O;
      const size_t o = (size_t)iz * plane + rowoff;
      const size_t os = a.sstride;
      const T* kxp = a.kxp;
      auto emit = [&](int q, int k, cx<T> fA, cx<T> fB) {
        const cx<T> ikx = {0, kxp[k]};
        switch (q) {
          case 0: O[0 * os + o + k] = cadd(cmul(ikx, fA), fB); break;          // G_u
          case 1: O[1 * os + o + k] = fA;                                       // Fy_u
                  O[2 * os + o + k] = cadd(cmul(ikx, fA), fB); break;          // G_v
          case 2: O[4 * os + o + k] = cadd(cmul(ikx, fA), fB); break;          // G_w
          default: O[3 * os + o + k] = fA;                                      // Fy_v
                   O[5 * os + o + k] = fB; break;                               // Fy_w
        }
      };
      anal_small<T, NX, 4>(buf, fa, fb, tw, tid, emit);
      __syncthreads();                     // wave partials published; buf free for the next level
      if (mn && tid < 6) {
        double sq = 0;
        for (int w = 0; w < NX / 64; ++w) sq += red[tid * (NX / 64) + w];
        atomicAdd(&a.sred[(size_t)tid * a.slots + iz], sq / ((double)NX * (double)NY));
      }
    } else {
      __syncthreads();
    }
    um = u0; u0 = up; up = upp;
    vm = v0; v0 = vp; vp = vpp;
    wm = w0; w0 = wp;
    em = e0; e0 = ep;
    ux0 = ux1; uy0 = uy1; vx0 = vx1; vy0 = vy1;
    wxm = wx0; wym = wy0; wx0 = wxp; wy0 = wyp;
#pragma unroll
    for (int q = 0; q < 6; ++q) t0[q] = t1[q];
    ok0 = ok1;
    fuw = f13; fvw = f23;
  }
  if (lane == 0) {
    atomicAdd(a.counts, nacc);
    atomicAdd(a.counts + 1, 64ull * (unsigned long long)(z1 - z0));
  }
}

// Small-grid closure, columns in: U, V, W band spectra -> u, v, w and uy, vy,
// wy as x-spectral small rows (the y-inverse of a tile of TK columns and of
// its i ky multiple in one batch).  grid (MX/TK, nz, 3 fields) x 512
template <typename T>
struct SgsColsInArgs {
  const cx<T>* U[3];      // U, V, W band spectra
  cx<T>* S;               // outputs u, v, w, uy, vy, wy [6][slot][NY][MX]
  size_t sstride;
  const T* kyp;
  const cx<T>* twNY;
  int nz;
};

template <typename T, int NX, int NY>
__global__ void __launch_bounds__(512) k_sgs_cols_in(SgsColsInArgs<T> a) {
  using Sh = ColShape<T, NX, NY>;
  constexpr int MX = Sh::MX, TK = Sh::TKS, NT = 512, PL = colpad(NY);
  const int f = blockIdx.z, z = blockIdx.y + 1;
  if (f == 2 && z >= a.nz) return;                 // w lives on faces 1..nz-1
  const int k0 = xcd_tile(blockIdx.x, gridDim.x) * TK, tid = threadIdx.x;
  __shared__ cx<T> buf[2 * TK * PL];
  __shared__ cx<T> tw[NY];
  load_twiddles(tw, a.twNY, NY, tid, NT);
  const cx<T>* src = a.U[f] + (size_t)z * NY * MX + k0;
  if constexpr (Sh::CREG && TK == Sh::TKC) {
    // transposed inverses: the first stage (radix 4 over the classes j0 + 64 q,
    // this thread's band rows) of the value and of d/dy here, the rest per
    // owning wave (two transforms each)
    const int c = tid % TK, j0 = tid / TK, w = wave_of(tid), lane = tid & 63;
    cx<T> x[4], xd[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int j = j0 + 64 * q;
      cx<T> v = src[(size_t)j * MX + c];
      if (j == NY / 2) v = cx<T>{0, 0};
      const T k = a.kyp[j];
      x[q] = v;
      xd[q] = cx<T>{-k * v.y, k * v.x};
    }
    __syncthreads();                                  // twiddle table written
    fft_first_regs_t<T, NY>(buf + c * PL, tw_src(tw), j0, x);
    fft_first_regs_t<T, NY>(buf + (TK + c) * PL, tw_src(tw), j0, xd);
    __syncthreads();
    fft_run_t<T, NY, 2 * TK / (NT / 64), 64, n_stages(NY) - 2, true, PL>(buf + w * (2 * TK / (NT / 64)) * PL,
                                                                        tw_src(tw), lane);
  } else {
    for (int i = tid; i < TK * NY; i += NT) {
      const int c = i % TK, j = i / TK;
      cx<T> v = src[(size_t)j * MX + c];
      if (j == NY / 2) v = cx<T>{0, 0};
      const T k = a.kyp[j];
      buf[c * PL + pidx(j)] = v;
      buf[(TK + c) * PL + pidx(j)] = cx<T>{-k * v.y, k * v.x};
    }
    __syncthreads();
    fft_waves<T, NY, true, 2 * TK, NT / 64, PL>(buf, tw, tid);
  }
  __syncthreads();
  const size_t off = (size_t)z * NY * MX + k0;
  cx<T>* o0 = a.S + (size_t)f * a.sstride + off;
  cx<T>* o1 = a.S + (size_t)(3 + f) * a.sstride + off;
  for (int i = tid; i < 2 * TK * NY; i += NT) {
    const int c = i % TK, r = i / TK;
    const int y = r % NY, d = r / NY;
    (d ? o1 : o0)[(size_t)y * MX + c] = buf[(d * TK + c) * PL + pidx(y)];
  }
}

// Small-grid closure, columns out: the y-forward of the G and F^y rows of one
// component, then R += G + i ky F^y on the band (Nyquist row untouched; for w
// the (0,0) mode too, the plane-mean removal of comp1.f90:172).
// grid (MX/TK, nz, 3 components) x 512
template <typename T>
struct SgsColsOutArgs {
  const cx<T>* O;         // G_u, Fy_u, G_v, Fy_v, G_w, Fy_w [6][slot][NY][MX]
  size_t ostride;
  cx<T>* R[3];            // ru, rv, rw band spectra
  const T* kyp;
  const cx<T>* twNY;
  int nz;
};

template <typename T, int NX, int NY>
__global__ void __launch_bounds__(512) k_sgs_cols_out(SgsColsOutArgs<T> a) {
  using Sh = ColShape<T, NX, NY>;
  constexpr int MX = Sh::MX, TK = Sh::TKS, NT = 512, PL = colpad(NY);
  const int cmp = blockIdx.z, z = blockIdx.y + 1;
  if (cmp == 2 && z >= a.nz) return;               // rw(nnz) = 0
  const int k0 = xcd_tile(blockIdx.x, gridDim.x) * TK, tid = threadIdx.x;
  __shared__ cx<T> buf[2 * TK * PL];
  __shared__ cx<T> tw[NY];
  load_twiddles(tw, a.twNY, NY, tid, NT);
  const size_t off = (size_t)z * NY * MX + k0;
  const cx<T>* g = a.O + (size_t)(2 * cmp) * a.ostride + off;
  const cx<T>* fy = g + a.ostride;
  cx<T>* R = a.R[cmp] + off;
  // every load of the tile, and of the band values this thread updates, is
  // issued before the first is used: their latencies overlap
  constexpr int LPT = (2 * TK * NY + NT - 1) / NT, OPT = (TK * NY + NT - 1) / NT;
  cx<T> lv[LPT], rv[OPT];
#pragma unroll
  for (int q = 0; q < LPT; ++q) {
    const int i = tid + q * NT;
    if (i < 2 * TK * NY) {
      const int c = i % TK, r = i / TK;
      const int y = r % NY, d = r / NY;
      lv[q] = (d ? fy : g)[(size_t)y * MX + c];
    }
  }
#pragma unroll
  for (int q = 0; q < OPT; ++q) {
    const int i = tid + q * NT;
    if (i < TK * NY) rv[q] = R[(size_t)(i / TK) * MX + i % TK];
  }
#pragma unroll
  for (int q = 0; q < LPT; ++q) {
    const int i = tid + q * NT;
    if (i < 2 * TK * NY) {
      const int c = i % TK, r = i / TK;
      const int y = r % NY, d = r / NY;
      buf[(d * TK + c) * PL + pidx(y)] = lv[q];
    }
  }
  __syncthreads();
  fft_waves<T, NY, false, 2 * TK, NT / 64, PL>(buf, tw, tid);
  __syncthreads();
  const T s = T(1) / T(NY);
#pragma unroll
  for (int q = 0; q < OPT; ++q) {
    const int i = tid + q * NT;
    if (i >= TK * NY) continue;
    const int c = i % TK, j = i / TK;
    if (j == NY / 2 || (cmp == 2 && j == 0 && k0 + c == 0)) continue;
    const cx<T> gv = buf[c * PL + pidx(j)], fv = buf[(TK + c) * PL + pidx(j)];
    const T k = a.kyp[j];
    R[(size_t)j * MX + c] = cx<T>{rv[q].x + (gv.x - k * fv.y) * s, rv[q].y + (gv.y + k * fv.x) * s};
  }
}

// ---------------------------------------------------------------------------
// k_col_anal: y-forward of the flux rows, band truncation, R = G + i ky Fy.
// One tile per workgroup: TK kx columns of one level and component; tiles
// kx fastest, then level, then component (e, w, v, u: the heaviest first).
//
// G and F^y pass one after the other through one batch of TK transforms in
// LDS, so that four workgroups share a CU; both tiles are loaded into
// registers up front, so F^y's reads are in flight during G's transform.
// For e the band result takes a y-inverse (to x-spectral rows of physical y)
// in the same batch.
// ---------------------------------------------------------------------------
template <typename T>
struct AnalArgs {
  const cx<T>* A;         // 8 analysis rows fields (Gu, Fyu, ...)
  size_t astride;
  cx<T>* R[3];            // outputs: band spectra of ru, rv, rw (comp1 result)
  cx<T>* Tre;             // output: re as x-spectral small rows
  const T* kyp;
  const cx<T>* twNY;
  const cx<T>* twNYB;
  int nz;
};

// FP32 at ny = 512: at most 128 VGPRs, four waves per SIMD, so that the two
// workgroups the LDS allows share a CU (left alone the compiler takes 148-162
// and one workgroup per CU: 1.83 -> 1.23 ms per launch at 512^2 x 256).
template <typename T, int NX, int NY>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && NY == 512 ? 4 : 1)))
k_col_anal(AnalArgs<T> a) {
  using Sh = ColShape<T, NX, NY>;
  static_assert(Sh::NT == 512, "launch bound");
  constexpr int MX = Sh::MX, NYB = Sh::NYB, TK = Sh::TKC, NT = Sh::NT;
  constexpr int PLB = colpad(NYB), PLS = colpad(NY);
  constexpr int NKT = MX / TK;
  constexpr int LPT = (TK * NYB + NT - 1) / NT;      // tile elements per thread
  constexpr int OPT = (TK * NY + NT - 1) / NT;       // band outputs per thread
  constexpr bool FULL = (TK * NYB) % NT == 0;
  static_assert(PLS <= PLB, "the length-NY batch fits the length-NYB batch");
  const int tid = threadIdx.x;
  __shared__ cx<T> buf[TK * PLB];
  __shared__ cx<T> twb[NYB];
  __shared__ cx<T> tws[NY];
  const int nz = a.nz;
  const size_t splane = (size_t)NY * MX;
  const int t = blockIdx.x;
  if (t >= NKT * nz * 4) return;
  const int kt = xcd_tile(t % NKT, NKT), rest = t / NKT;
  const int z = rest % nz + 1, cmp = 3 - rest / nz;    // e first: three transforms, u, v, w two
  const int k0 = kt * TK;
  if (cmp >= 2 && z == nz) {                 // rw(nnz) = re(nnz) = 0 (comp1.f90:174-177)
    cx<T>* o = (cmp == 2 ? a.R[2] : a.Tre) + (size_t)z * splane + k0;
    for (int i = tid; i < TK * NY; i += NT) o[(size_t)(i / TK) * MX + (i % TK)] = cx<T>{0, 0};
    return;
  }
  load_twiddles(twb, a.twNYB, NYB, tid, NT);
  if (cmp == 3) load_twiddles(tws, a.twNY, NY, tid, NT);
  const cx<T>* g = a.A + (size_t)(2 * cmp) * a.astride + (size_t)z * NYB * MX + k0;
  const cx<T>* fy = g + a.astride;
  cx<T> lg[LPT], lf[LPT];
#pragma unroll
  for (int i = 0; i < LPT; ++i) {
    const int idx = tid + i * NT;
    if (FULL || idx < TK * NYB) {
      const int c = idx % TK, n = idx / TK;
      lg[i] = g[(size_t)n * MX + c];
      lf[i] = fy[(size_t)n * MX + c];
    }
  }
  auto band_row = [](int j) { return (j < NY / 2) ? j : j - NY + NYB; };
#pragma unroll
  for (int i = 0; i < LPT; ++i) {
    const int idx = tid + i * NT;
    if (FULL || idx < TK * NYB) buf[(idx % TK) * PLB + pidx(idx / TK)] = lg[i];
  }
  __syncthreads();
  fft_waves<T, NYB, false, TK, NT / 64, PLB>(buf, twb, tid);
  __syncthreads();
  cx<T> v[OPT];                              // G, then the band result
#pragma unroll
  for (int q = 0; q < OPT; ++q) {
    const int i = tid + q * NT;
    v[q] = cx<T>{0, 0};
    if (i < TK * NY) {
      const int c = i % TK, j = i / TK;
      if (j != NY / 2) v[q] = buf[c * PLB + pidx(band_row(j))];
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < LPT; ++i) {
    const int idx = tid + i * NT;
    if (FULL || idx < TK * NYB) buf[(idx % TK) * PLB + pidx(idx / TK)] = lf[i];
  }
  __syncthreads();
  fft_waves<T, NYB, false, TK, NT / 64, PLB>(buf, twb, tid);
  __syncthreads();
  const T s = T(1) / T(NYB);
#pragma unroll
  for (int q = 0; q < OPT; ++q) {
    const int i = tid + q * NT;
    if (i < TK * NY) {
      const int c = i % TK, j = i / TK;
      if (j != NY / 2) {
        const cx<T> fv = buf[c * PLB + pidx(band_row(j))];
        const T k = a.kyp[j];
        v[q] = cx<T>{(v[q].x - k * fv.y) * s, (v[q].y + k * fv.x) * s};
        if (cmp == 2 && j == 0 && k0 + c == 0) v[q] = cx<T>{0, 0};   // rw mean removal (comp1.f90:172)
      }
    }
  }
  if (cmp < 3) {
    cx<T>* R = a.R[cmp] + (size_t)z * splane + k0;
#pragma unroll
    for (int q = 0; q < OPT; ++q) {
      const int i = tid + q * NT;
      if (i < TK * NY) R[(size_t)(i / TK) * MX + i % TK] = v[q];
    }
    return;
  }
  // re: the y-inverse of the band result to rows of physical y
  __syncthreads();
#pragma unroll
  for (int q = 0; q < OPT; ++q) {
    const int i = tid + q * NT;
    if (i < TK * NY) buf[(i % TK) * PLS + pidx(i / TK)] = v[q];
  }
  __syncthreads();
  fft_waves<T, NY, true, TK, NT / 64, PLS>(buf, tws, tid);
  __syncthreads();
  cx<T>* o = a.Tre + (size_t)z * splane + k0;
#pragma unroll
  for (int i = tid; i < TK * NY; i += NT) {
    const int c = i % TK, y = i / TK;
    o[(size_t)y * MX + c] = buf[c * PLS + pidx(y)];
  }
}

// ---------------------------------------------------------------------------
// k_press: substep + Poisson RHS + tridiagonal z-solve (FP64) + projection and
// RK update (comprhs.f90:1-37, press.F90:1-115, tridv.f90:1-55, comp2.f90:1-45).
// One lane per band mode (kx, j).
//
// The reciprocal pivots d_z depend on k^2 and dz only: the forward sweep
// recomputes them and the backward sweep reads the identical values from the
// table k_press_pivots writes once (piv[slot][mode]).
//
// Mode (0,0) is left to the grid's last workgroup: its pressure is pinned to
// 0 (press.F90:81-88), so its levels are independent.  Its u, v, ru, rv live
// in FP64 in M0 (the band holds them rounded to T), and the imposed pressure
// gradient 0.016 (comp1.f90:152) is added to its ru there, in FP64.
// grid ceil(NY MX / 64) + 1 workgroups of 64.
// ---------------------------------------------------------------------------
template <typename T>
struct PressArgs {
  cx<T>* U[3];            // U, V, W band spectra (in: stage start, out: new)
  cx<T>* Ro[3];           // R registers (in: previous stage final, out: this stage final)
  const cx<T>* Rn[3];     // comp1 tendencies of this stage (without the 0.016 forcing)
  cx<double>* P;          // pressure band spectrum (out)
  const double* piv;      // reciprocal pivots [slots][modes] (k_press_pivots)
  cx<double>* sr;         // scratch: forward-eliminated rhs [slots][modes]
  double* M0;             // FP64 (0,0) coefficients [4][slots]: u, v, ru, rv
  const double* kxd;      // kx wavenumbers (MX)
  const double* kyd;      // ky wavenumbers per band row (NY)
  int mx, ny, nz, istage, slots;
  double dz, dzm, dtab1, dtab2, dtabm;
};

// Reciprocal pivots of the Thomas elimination (tridv.f90:34-43) for every
// mode and level: d_1 = 1/diag_1, d_z = 1/(diag_z - d_z-1).  One thread per mode.
__global__ void __launch_bounds__(64) k_press_pivots(double* piv, const double* kxd, const double* kyd,
                                                     int mx, int ny, int nz, double dz) {
  const int modes = mx * ny;
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= modes) return;
  const int j = m / mx, kx = m - j * mx;
  const double kxp = kxd[kx], kyp = kyd[j];
  const double dz2 = dz * dz;
  const double k2 = kxp * kxp + kyp * kyp;      // xk2 (init.f90:33-42)
  double d_prev = 0.0;
  for (int z = 1; z <= nz; ++z) {
    const double diag = -dz2 * k2 - ((z == 1 || z == nz) ? 1.0 : 2.0);
    const double d = (z == 1) ? 1.0 / diag : 1.0 / (diag - d_prev);
    piv[(size_t)z * modes + m] = d;
    d_prev = d;
  }
}

// The mode (0,0): substep, the forcing, the RK update in FP64; w, rw, p = 0.
template <typename T>
__device__ void press_zero_mode(const PressArgs<T>& a) {
  const size_t pl = (size_t)a.mx * a.ny;
  double* mu = a.M0 + M0_U * a.slots;
  double* mv = a.M0 + M0_V * a.slots;
  double* mru = a.M0 + M0_RU * a.slots;
  double* mrv = a.M0 + M0_RV * a.slots;
  const bool sub = (a.istage != 1);
  for (int z = 1 + threadIdx.x; z <= a.nz; z += blockDim.x) {
    const size_t o = (size_t)z * pl;
    double u = mu[z], v = mv[z];
    if (sub) {                                   // comp1.f90:126-128
      u -= a.dtab2 * mru[z];
      v -= a.dtab2 * mrv[z];
    }
    const double ru = (double)a.Rn[0][o].x + FORCING;   // comp1.f90:152
    const double rv = (double)a.Rn[1][o].x;
    u += a.dtab1 * ru;                           // comp2.f90:34-35
    v += a.dtab1 * rv;
    mu[z] = u; mv[z] = v; mru[z] = ru; mrv[z] = rv;
    a.U[0][o] = cx<T>{(T)u, T(0)};
    a.U[1][o] = cx<T>{(T)v, T(0)};
    a.U[2][o] = cx<T>{T(0), T(0)};               // the w and rw plane means are removed (comp2.f90:30,39)
    a.Ro[0][o] = cx<T>{(T)ru, T(0)};
    a.Ro[1][o] = cx<T>{(T)rv, T(0)};
    a.Ro[2][o] = cx<T>{T(0), T(0)};
    a.P[o] = cx<double>{0.0, 0.0};
  }
}

template <typename T>
__device__ __forceinline__ cx<double> widen(const cx<T>& v) {
  return cx<double>{(double)v.x, (double)v.y};
}
template <typename T>
__device__ __forceinline__ void st2(cx<T>* p, cx<double> v) {
  *p = cx<T>{(T)v.x, (T)v.y};
}

// The z sweeps are latency bound (one dependent step per level), so each
// thread loads the inputs of ZU levels before running their ZU recurrence
// steps: ZU levels of loads are in flight per memory round trip.  FP32
// inputs take ZU = 8, FP64 inputs 4.
template <typename T>
__global__ void __launch_bounds__(64) k_press(PressArgs<T> a) {
  constexpr int ZU = sizeof(T) == 4 ? 8 : 4;
  if (blockIdx.x == gridDim.x - 1) {
    press_zero_mode(a);
    return;
  }
  const int modes = a.mx * a.ny;
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= modes || m == 0) return;             // (0,0): the last workgroup
  const int j = m / a.mx, kx = m - j * a.mx;
  if (j == a.ny / 2) return;                    // Nyquist row: always zero
  const double kxp = a.kxd[kx], kyp = a.kyd[j];
  const double dz2 = a.dz * a.dz;
  const double k2 = kxp * kxp + kyp * kyp;      // xk2 (init.f90:33-42)
  const int nz = a.nz;
  const size_t pl = (size_t)modes;
  const bool sub = (a.istage != 1);
  cx<T>* __restrict__ U0 = a.U[0];
  cx<T>* __restrict__ U1 = a.U[1];
  cx<T>* __restrict__ U2 = a.U[2];
  cx<T>* __restrict__ R0 = a.Ro[0];
  cx<T>* __restrict__ R1 = a.Ro[1];
  cx<T>* __restrict__ R2 = a.Ro[2];
  const cx<T>* __restrict__ N0 = a.Rn[0];
  const cx<T>* __restrict__ N1 = a.Rn[1];
  const cx<T>* __restrict__ N2 = a.Rn[2];
  cx<double>* __restrict__ sr = a.sr;
  // forward sweep (comprhs.f90:17-35, press.F90:44-76, tridv.f90:34-43)
  cx<double> wm_prev = {0, 0}, rw_prev = {0, 0}, r_prev = {0, 0};
  double d_prev = 0.0;
  for (int zb = 1; zb <= nz; zb += ZU) {
    cx<T> lu[ZU], lv[ZU], lw[ZU], lru[ZU], lrv[ZU], lrw[ZU], lou[ZU], lov[ZU], low[ZU];
#pragma unroll
    for (int q = 0; q < ZU; ++q) {
      const int z = zb + q;
      if (z <= nz) {
        const size_t o = (size_t)z * pl + m;
        lu[q] = U0[o]; lv[q] = U1[o]; lw[q] = U2[o];
        lru[q] = N0[o]; lrv[q] = N1[o]; lrw[q] = N2[o];
        if (sub) { lou[q] = R0[o]; lov[q] = R1[o]; low[q] = R2[o]; }
      }
    }
#pragma unroll
    for (int q = 0; q < ZU; ++q) {
      const int z = zb + q;
      if (z > nz) break;
      const size_t o = (size_t)z * pl + m;
      cx<double> u = widen(lu[q]), v = widen(lv[q]), w = widen(lw[q]);
      if (sub) {                                 // comp1.f90:126-133
        u = csub(u, cscale(widen(lou[q]), a.dtab2));
        v = csub(v, cscale(widen(lov[q]), a.dtab2));
        if (z != nz) w = csub(w, cscale(widen(low[q]), a.dtab2));
        st2(U0 + o, u);
        st2(U1 + o, v);
        st2(U2 + o, w);
      }
      const cx<double> rw = widen(lrw[q]);
      const cx<double> fx = cadd(widen(lru[q]), cscale(u, a.dtabm));
      const cx<double> fy = cadd(widen(lrv[q]), cscale(v, a.dtabm));
      cx<double> rhs = cadd(cmuli(cscale(fx, kxp)), cmuli(cscale(fy, kyp)));
      const cx<double> dw = cadd(cscale(csub(w, wm_prev), a.dtabm), csub(rw, rw_prev));
      rhs = cadd(rhs, cscale(dw, a.dzm));
      wm_prev = w;
      rw_prev = rw;
      const double diag = -dz2 * k2 - ((z == 1 || z == nz) ? 1.0 : 2.0);
      cx<double> r = cscale(rhs, dz2);
      double d;
      if (z == 1) {
        d = 1.0 / diag;
      } else {
        d = 1.0 / (diag - d_prev);               // a(i-1) = b(i) = 1
        r = csub(r, cscale(r_prev, d_prev));
      }
      sr[o] = r;
      d_prev = d;
      r_prev = r;
    }
  }
  // backward sweep (tridv.f90:44-53) fused with comp2 (comp2.f90:17-40)
  cx<double> p_next = {0, 0};
  for (int zt = nz; zt >= 1; zt -= ZU) {
    double ld[ZU];
    cx<double> lr[ZU];
    cx<T> lu[ZU], lv[ZU], lw[ZU], lru[ZU], lrv[ZU], lrw[ZU];
#pragma unroll
    for (int q = 0; q < ZU; ++q) {
      const int z = zt - q;
      if (z >= 1) {
        const size_t o = (size_t)z * pl + m;
        ld[q] = a.piv[o]; lr[q] = sr[o];
        lu[q] = U0[o]; lv[q] = U1[o]; lw[q] = U2[o];
        lru[q] = N0[o]; lrv[q] = N1[o]; lrw[q] = N2[o];
      }
    }
#pragma unroll
    for (int q = 0; q < ZU; ++q) {
      const int z = zt - q;
      if (z < 1) break;
      const size_t o = (size_t)z * pl + m;
      const cx<double> p = (z == nz) ? cscale(lr[q], ld[q]) : cscale(csub(lr[q], p_next), ld[q]);
      const cx<double> ru = csub(widen(lru[q]), cmuli(cscale(p, kxp)));   // comp2.f90:18-19
      const cx<double> rv = csub(widen(lrv[q]), cmuli(cscale(p, kyp)));   // comp2.f90:23-24
      cx<double> rw = {0, 0}, wn = {0, 0};
      if (z != nz) {
        rw = csub(widen(lrw[q]), cscale(csub(p_next, p), a.dzm));        // comp2.f90:29
        wn = cadd(widen(lw[q]), cscale(rw, a.dtab1));                     // comp2.f90:37
      }
      st2(U0 + o, cadd(widen(lu[q]), cscale(ru, a.dtab1)));              // comp2.f90:34-35
      st2(U1 + o, cadd(widen(lv[q]), cscale(rv, a.dtab1)));
      st2(U2 + o, wn);
      st2(R0 + o, ru);
      st2(R1 + o, rv);
      st2(R2 + o, rw);
      a.P[o] = p;
      p_next = p;
    }
  }
}

// ---------------------------------------------------------------------------
// k_press_part: the arithmetic of k_press with each mode's z-column split into
// chunks of C levels, one thread per (mode, chunk), the chunk held in
// registers: every input is read once and every output written once (no
// eliminated-rhs scratch, no second read of U, V, W and Rn, no pivot table).
//
// The tridiagonal solve (press.F90:44-76, tridv.f90:34-53) is the partition
// (spike) method.  The last level of every chunk but the last is an
// interface; a chunk's interior levels solve A_c p = r_c - I_l e_first -
// I_r e_last, so p = y - I_l g - I_r h with A_c y = r_c, A_c g = e_first,
// A_c h = e_last.  The row of interface c then couples I_{c-1}, I_c, I_{c+1}:
//   -g_c(last) I_{c-1} + (b - h_c(last) - g_{c+1}(first)) I_c
//     - h_{c+1}(first) I_{c+1} = r_c(interface) - y_c(last) - y_{c+1}(first),
// a diagonally dominant tridiagonal system of nc-1 rows that one thread per
// mode solves in LDS by elimination.  Each chunk then back-substitutes its
// interior with the interface values folded into the eliminated rhs.  The
// solution differs from the serial sweep's by rounding: the interface system
// inherits the diagonal dominance of -dz^2 k^2 - 2 (k^2 >= 0.05 for every mode
// but (0,0), checked on the host).
//
// grid: ceil(modes / 16) + 1 workgroups of 16 NP threads, thread = chunk * 16
// + mode (a wave's loads are four 16-mode runs of one level); the last
// workgroup takes mode (0,0).  Needs nc = ceil(nz / C) <= NP.
// ---------------------------------------------------------------------------
template <typename T, int C, int NP, int MW>
__global__ void __launch_bounds__(MW * NP) __attribute__((amdgpu_waves_per_eu(2)))
k_press_part(PressArgs<T> a) {
  if (blockIdx.x == gridDim.x - 1) {
    press_zero_mode(a);
    return;
  }
  // per chunk and mode: y at the first and last interior level (complex), g
  // and h there, the interface row's rhs (complex)
  enum { YF = 0, YL = 2, GF = 4, GL = 5, HF = 6, HL = 7, RI = 8, NQ = 10 };
  __shared__ double post[NP][NQ][MW];
  __shared__ double beta[NP][MW];       // reciprocal pivots of the interface system
  __shared__ cx<double> itf[NP][MW];
  const int tid = threadIdx.x, mi = tid % MW, c = tid / MW;
  const int modes = a.mx * a.ny;
  const int m = blockIdx.x * MW + mi;
  const int j = m / a.mx, kx = m - j * a.mx;
  const bool act = m < modes && m != 0 && j != a.ny / 2;   // (0,0) and the Nyquist row excluded
  const int nz = a.nz;
  const int nc = (nz + C - 1) / C;
  const int lo = c * C + 1;
  const int n = (lo + C - 1 <= nz ? lo + C - 1 : nz) - lo + 1;   // levels of this chunk
  const bool live = act && c < nc;
  const bool right = c < nc - 1;                                  // its last level is an interface
  const int ni = right ? n - 1 : n;                               // interior levels
  double kxp = 0.0, kyp = 0.0;
  if (act) {
    kxp = a.kxd[kx];
    kyp = a.kyd[j];
  }
  const double dz2 = a.dz * a.dz;
  const double k2 = kxp * kxp + kyp * kyp;       // xk2 (init.f90:33-42)
  const size_t pl = (size_t)modes;
  const bool sub = (a.istage != 1);

  // loads: the chunk's levels, and w, rw of the level below it.  They are
  // unconditional, from a valid mode and level (a level past the chunk reads
  // level nz, an inactive thread mode 0), so that every load of the chunk is
  // in flight before the first is used; the values outside are zeroed after.
  cx<T> lu[C], lv[C], lw[C], ln0[C], ln1[C], ln2[C], ou[C], ov[C], ow[C];
  const cx<T> zt = {T(0), T(0)};
  const int mc = m < modes ? m : 0;
  const int lb = lo > 1 ? (lo - 1 <= nz ? lo - 1 : nz) : 1;
#pragma unroll
  for (int q = 0; q < C; ++q) {
    const size_t o = (size_t)(lo + q <= nz ? lo + q : nz) * pl + mc;
    lu[q] = a.U[0][o]; lv[q] = a.U[1][o]; lw[q] = a.U[2][o];
    ln0[q] = a.Rn[0][o]; ln1[q] = a.Rn[1][o]; ln2[q] = a.Rn[2][o];
  }
  cx<T> wb = a.U[2][(size_t)lb * pl + mc], nb = a.Rn[2][(size_t)lb * pl + mc], ob = zt;
  if (sub) {
#pragma unroll
    for (int q = 0; q < C; ++q) {
      const size_t o = (size_t)(lo + q <= nz ? lo + q : nz) * pl + mc;
      ou[q] = a.Ro[0][o]; ov[q] = a.Ro[1][o]; ow[q] = a.Ro[2][o];
    }
    ob = a.Ro[2][(size_t)lb * pl + mc];
  } else {
#pragma unroll
    for (int q = 0; q < C; ++q) ou[q] = ov[q] = ow[q] = zt;
  }
#pragma unroll
  for (int q = 0; q < C; ++q) {
    if (!(live && q < n)) lu[q] = lv[q] = lw[q] = ln0[q] = ln1[q] = ln2[q] = ou[q] = ov[q] = ow[q] = zt;
  }
  if (!(live && lo > 1)) wb = nb = ob = zt;

  // substep, rhs and the interior elimination (comp1.f90:126-133,
  // comprhs.f90:17-35, tridv.f90:34-43); g's rhs is e_first
  cx<double> wm_prev = {0, 0}, rw_prev = {0, 0};
  if (lo > 1) {
    wm_prev = widen(wb);
    if (sub) wm_prev = csub(wm_prev, cscale(widen(ob), a.dtab2));   // level lo-1 < nz
    rw_prev = widen(nb);
  }
  cx<double> ry[C], rif = {0, 0};
  double rg[C], dd[C];
  double d_prev = 0.0, rg_prev = 0.0;
  cx<double> ry_prev = {0, 0};
#pragma unroll
  for (int q = 0; q < C; ++q) {
    ry[q] = {0, 0};
    rg[q] = 0.0;
    dd[q] = 0.0;
    if (live && q < n) {
      const int z = lo + q;
      cx<double> u = widen(lu[q]), v = widen(lv[q]), w = widen(lw[q]);
      if (sub) {
        u = csub(u, cscale(widen(ou[q]), a.dtab2));
        v = csub(v, cscale(widen(ov[q]), a.dtab2));
        if (z != nz) w = csub(w, cscale(widen(ow[q]), a.dtab2));
        lu[q] = cx<T>{(T)u.x, (T)u.y};
        lv[q] = cx<T>{(T)v.x, (T)v.y};
        lw[q] = cx<T>{(T)w.x, (T)w.y};
      }
      const cx<double> rw = widen(ln2[q]);
      const cx<double> fx = cadd(widen(ln0[q]), cscale(u, a.dtabm));
      const cx<double> fy = cadd(widen(ln1[q]), cscale(v, a.dtabm));
      cx<double> rhs = cadd(cmuli(cscale(fx, kxp)), cmuli(cscale(fy, kyp)));
      const cx<double> dw = cadd(cscale(csub(w, wm_prev), a.dtabm), csub(rw, rw_prev));
      rhs = cadd(rhs, cscale(dw, a.dzm));
      wm_prev = w;
      rw_prev = rw;
      const cx<double> r = cscale(rhs, dz2);
      if (q < ni) {
        const double diag = -dz2 * k2 - ((z == 1 || z == nz) ? 1.0 : 2.0);
        if (q == 0) {
          dd[q] = 1.0 / diag;
          ry[q] = r;
          rg[q] = 1.0;
        } else {
          dd[q] = 1.0 / (diag - d_prev);         // sub- and super-diagonal are 1
          ry[q] = csub(r, cscale(ry_prev, d_prev));
          rg[q] = -d_prev * rg_prev;
        }
        d_prev = dd[q];
        ry_prev = ry[q];
        rg_prev = rg[q];
      } else {
        rif = r;                                 // q == ni: the interface level
      }
    }
  }

  // y, g, h at the interior's first and last level (tridv.f90:44-53)
  {
    cx<double> y = {0, 0}, yl = {0, 0};
    double g = 0.0, h = 0.0, gl = 0.0, hl = 0.0;
#pragma unroll
    for (int q = C - 1; q >= 0; --q) {
      if (live && q < ni) {
        if (q == ni - 1) {
          y = cscale(ry[q], dd[q]);
          g = dd[q] * rg[q];
          h = dd[q];
          yl = y;
          gl = g;
          hl = h;
        } else {
          y = cscale(csub(ry[q], y), dd[q]);
          g = dd[q] * (rg[q] - g);
          h = -dd[q] * h;
        }
      }
    }
    if (c == 0) g = gl = 0.0;                    // no coupling below level 1
    if (!right) h = hl = 0.0;                    // none above level nz
    post[c][YF][mi] = y.x;
    post[c][YF + 1][mi] = y.y;
    post[c][YL][mi] = yl.x;
    post[c][YL + 1][mi] = yl.y;
    post[c][GF][mi] = g;
    post[c][GL][mi] = gl;
    post[c][HF][mi] = h;
    post[c][HL][mi] = hl;
    post[c][RI][mi] = rif.x;
    post[c][RI + 1][mi] = rif.y;
  }
  __syncthreads();

  // the interface system, one thread per mode (loops unrolled over the
  // largest size so that the LDS reads of later rows are issued early)
  if (c == 0 && act) {
    const double bif = -dz2 * k2 - 2.0;           // interfaces lie strictly inside 1..nz
    double rbprev = 0.0, supprev = 0.0;
    cx<double> qprev = {0, 0};
#pragma unroll
    for (int ci = 0; ci < NP - 1; ++ci) {
      if (ci < nc - 1) {
        const double diag = bif - post[ci][HL][mi] - post[ci + 1][GF][mi];
        cx<double> q = {post[ci][RI][mi] - post[ci][YL][mi] - post[ci + 1][YF][mi],
                        post[ci][RI + 1][mi] - post[ci][YL + 1][mi] - post[ci + 1][YF + 1][mi]};
        double bt = diag;
        if (ci > 0) {
          const double f = -post[ci][GL][mi] * rbprev;   // row ci's coefficient of I_{ci-1} / pivot
          bt = diag - f * supprev;
          q = csub(q, cscale(qprev, f));
        }
        const double rb = 1.0 / bt;
        beta[ci][mi] = rb;
        itf[ci][mi] = q;
        rbprev = rb;
        qprev = q;
        supprev = -post[ci + 1][HF][mi];          // row ci's coefficient of I_{ci+1}
      }
    }
    cx<double> inext = {0, 0};
#pragma unroll
    for (int ci = NP - 2; ci >= 0; --ci) {
      if (ci < nc - 1) {
        cx<double> q = itf[ci][mi];
        if (ci < nc - 2) q = csub(q, cscale(inext, -post[ci + 1][HF][mi]));
        inext = cscale(q, beta[ci][mi]);
        itf[ci][mi] = inext;
      }
    }
  }
  __syncthreads();

  // back-substitution fused with the projection and the RK update
  // (tridv.f90:44-53, comp2.f90:17-40)
  if (live) {
    const cx<double> il = (c > 0) ? itf[c - 1][mi] : cx<double>{0, 0};
    const cx<double> ir = right ? itf[c][mi] : cx<double>{0, 0};
    cx<double> p_next = {0, 0};
    if (right) {                                  // p at level lo + n, the next chunk's first
      const cx<double> i2 = (c + 1 < nc - 1) ? itf[c + 1][mi] : cx<double>{0, 0};
      const cx<double> yf1 = {post[c + 1][YF][mi], post[c + 1][YF + 1][mi]};
      p_next = csub(csub(yf1, cscale(ir, post[c + 1][GF][mi])), cscale(i2, post[c + 1][HF][mi]));
    }
#pragma unroll
    for (int q = C - 1; q >= 0; --q) {
      if (q < n) {
        const int z = lo + q;
        cx<double> p;
        if (q == ni) {
          p = ir;                                 // the interface level
        } else {
          const cx<double> rp = csub(ry[q], cscale(il, rg[q]));
          p = (q == ni - 1) ? cscale(csub(rp, ir), dd[q]) : cscale(csub(rp, p_next), dd[q]);
        }
        const cx<double> ru = csub(widen(ln0[q]), cmuli(cscale(p, kxp)));   // comp2.f90:18-19
        const cx<double> rv = csub(widen(ln1[q]), cmuli(cscale(p, kyp)));   // comp2.f90:23-24
        cx<double> rw = {0, 0}, wn = {0, 0};
        if (z != nz) {
          rw = csub(widen(ln2[q]), cscale(csub(p_next, p), a.dzm));        // comp2.f90:29
          wn = cadd(widen(lw[q]), cscale(rw, a.dtab1));                     // comp2.f90:37
        }
        const size_t o = (size_t)z * pl + m;
        st2(a.U[0] + o, cadd(widen(lu[q]), cscale(ru, a.dtab1)));          // comp2.f90:34-35
        st2(a.U[1] + o, cadd(widen(lv[q]), cscale(rv, a.dtab1)));
        st2(a.U[2] + o, wn);
        st2(a.Ro[0] + o, ru);
        st2(a.Ro[1] + o, rv);
        st2(a.Ro[2] + o, rw);
        a.P[o] = p;
        p_next = p;
      }
    }
  }
}

// Load-time means: the band (0,0) coefficients of u and v take the FP64
// plane means M0 (rounded to T, imaginary part 0); one thread per level.
template <typename T>
__global__ void k_set_mean(cx<T>* U, cx<T>* V, const double* __restrict__ M0, int slots, size_t plane,
                           int nz) {
  const int z = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (z > nz) return;
  U[(size_t)z * plane] = cx<T>{(T)M0[M0_U * slots + z], T(0)};
  V[(size_t)z * plane] = cx<T>{(T)M0[M0_V * slots + z], T(0)};
}

// Physical levels 1..nz in out (nxy points each) of a field synthesised from
// a band whose (0,0) coefficient m0 holds in FP64: add m0[z] - band(0,0) and
// the constant add to every point of level z.  grid (nz), 256 threads.
template <typename T>
__global__ void __launch_bounds__(256) k_shift_levels(double* out, const double* __restrict__ m0,
                                                      const cx<T>* __restrict__ band, size_t plane,
                                                      size_t nxy, double add) {
  const int z = 1 + blockIdx.x;
  double d = add;
  if (m0) d += m0[z] - (double)band[(size_t)z * plane].x;
  double* o = out + (size_t)(z - 1) * nxy;
  for (size_t i = threadIdx.x; i < nxy; i += 256) o[i] += d;
}

// ---------------------------------------------------------------------------
// k_erows: SGS TKE update on the small grid and its x-forward transform.
// grid (NY/2, nz); each workgroup handles rows y and y + NY/2 of level z.
//   mode 0: re <- x-inverse(Tre); e <- max(e - dtab2 re_old + dtab1 re, emin)
//           (comp1.f90:131, comp2.f90:38); Ex <- x-forward(e)
//   mode 1: Ex <- x-forward(e) only (after set_state)
// ---------------------------------------------------------------------------
template <typename T>
struct ERowArgs {
  T* e;                   // physical e [slots][NY][NX]
  T* re;                  // physical re
  const cx<T>* Tre;       // x-spectral re of this stage
  cx<T>* Ex;              // x-spectral e (out)
  const cx<T>* twNX;
  int* flags;
  int nz, istage, mode;
  double dtab1, dtab2, emin;
};

template <typename T, int NX>
__global__ void __launch_bounds__(NX / 2) k_erows(ERowArgs<T> a) {
  constexpr int MX = NX / 2, NT = MX, NY = NX;
  __shared__ cx<T> buf[padlen(NX)];
  __shared__ cx<T> tw[NX];
  const int tid = threadIdx.x;
  const int y0 = blockIdx.x, y1 = blockIdx.x + NY / 2;
  const int z = blockIdx.y + 1;
  const int nz = a.nz;
  const size_t plane = (size_t)NY * NX, splane = (size_t)NY * MX;
  T* e = a.e + (size_t)z * plane;
  T* re = a.re + (size_t)z * plane;
  if (a.mode == 0 && z == nz) {                // re(nnz) = 0 (comp1.f90:176)
    for (int p = 0; p < 2; ++p) {
      re[(size_t)y0 * NX + tid + p * NT] = 0;
      re[(size_t)y1 * NX + tid + p * NT] = 0;
    }
    return;
  }
  if (a.mode != 0 && z >= nz) return;
  load_twiddles(tw, a.twNX, NX, tid, NT);
  T ea[2], eb[2];
  if (a.mode == 0) {
    // e and re_old of this thread's points, loaded before the transform so
    // that their latency overlaps it
    T eo0[2], eo1[2], ro0[2] = {T(0), T(0)}, ro1[2] = {T(0), T(0)};
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int x = tid + p * NT;
      eo0[p] = e[(size_t)y0 * NX + x];
      eo1[p] = e[(size_t)y1 * NX + x];
      if (a.istage != 1) {
        ro0[p] = re[(size_t)y0 * NX + x];
        ro1[p] = re[(size_t)y1 * NX + x];
      }
    }
    // synthesise re on rows y0 (real part) and y1 (imaginary part)
    const cx<T>* s0 = a.Tre + (size_t)z * splane + (size_t)y0 * MX;
    const cx<T>* s1 = a.Tre + (size_t)z * splane + (size_t)y1 * MX;
    const int k = tid;
    const cx<T> av = s0[k], bv = s1[k];
    if (k == 0) {
      buf[0] = cx<T>{av.x, bv.x};
      buf[pidx(MX)] = cx<T>{0, 0};
    } else {
      buf[pidx(k)] = cx<T>{av.x - bv.y, av.y + bv.x};
      buf[pidx(NX - k)] = cx<T>{av.x + bv.y, bv.x - av.y};
    }
    __syncthreads();
    fft_run<T, NX, true, 1, NT>(buf, tw, tid);
    int bad = 0;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int x = tid + p * NT;
      const cx<T> r = buf[pidx(x)];
      T* pe0 = e + (size_t)y0 * NX + x;
      T* pe1 = e + (size_t)y1 * NX + x;
      T* pr0 = re + (size_t)y0 * NX + x;
      T* pr1 = re + (size_t)y1 * NX + x;
      T e0 = eo0[p], e1 = eo1[p];
      if (z == nz - 1) {                       // upper: e(nnz) = e(nnz-1) (upperbc.f90:20-24)
        e[(size_t)plane + (size_t)y0 * NX + x] = e0;
        e[(size_t)plane + (size_t)y1 * NX + x] = e1;
      }
      if (a.istage != 1) {
        e0 = e0 - (T)a.dtab2 * ro0[p];
        e1 = e1 - (T)a.dtab2 * ro1[p];
      }
      e0 = e0 + (T)a.dtab1 * r.x;
      e1 = e1 + (T)a.dtab1 * r.y;
      if (!isfinite(e0) || !isfinite(e1)) bad = 1;   // before the clamp, which would turn a NaN into emin
      e0 = e0 > (T)a.emin ? e0 : (T)a.emin;
      e1 = e1 > (T)a.emin ? e1 : (T)a.emin;
      *pe0 = e0; *pe1 = e1;
      *pr0 = r.x; *pr1 = r.y;
      ea[p] = e0; eb[p] = e1;
    }
    if (bad) atomicOr(a.flags, FLAG_NONFINITE);
    __syncthreads();
  } else {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      ea[p] = e[(size_t)y0 * NX + tid + p * NT];
      eb[p] = e[(size_t)y1 * NX + tid + p * NT];
    }
  }
  // x-forward of the two rows
#pragma unroll
  for (int p = 0; p < 2; ++p) buf[pidx(tid + p * NT)] = cx<T>{ea[p], eb[p]};
  __syncthreads();
  fft_run<T, NX, false, 1, NT>(buf, tw, tid);
  const T s = T(0.5) / T(NX);
  const int k = tid;
  const cx<T> zk = buf[pidx(k)];
  const cx<T> zn = buf[pidx(k == 0 ? 0 : NX - k)];
  cx<T>* o0 = a.Ex + (size_t)z * splane + (size_t)y0 * MX;
  cx<T>* o1 = a.Ex + (size_t)z * splane + (size_t)y1 * MX;
  o0[k] = cx<T>{(zk.x + zn.x) * s, (zk.y - zn.y) * s};
  o1[k] = cx<T>{(zk.y + zn.y) * s, (zn.x - zk.x) * s};
}

// ---------------------------------------------------------------------------
// Small-grid conversions used at set/get time (not on the step path)
// ---------------------------------------------------------------------------
// physical double rows -> x-spectral rows (x 1/NX, Nyquist dropped); grid (NY/2, nl)
template <typename T, int NX>
__global__ void __launch_bounds__(NX / 2) k_rows_fwd(const double* __restrict__ in, cx<T>* out,
                                                      const cx<T>* twg, int l0) {
  constexpr int MX = NX / 2, NT = MX, NY = NX;
  __shared__ cx<T> buf[padlen(NX)];
  __shared__ cx<T> tw[NX];
  const int tid = threadIdx.x;
  load_twiddles(tw, twg, NX, tid, NT);
  const int y0 = blockIdx.x, y1 = y0 + NY / 2;
  const int L = l0 + blockIdx.y;
  const double* p = in + (size_t)blockIdx.y * NY * NX;
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int x = tid + q * NT;
    buf[pidx(x)] = cx<T>{(T)p[(size_t)y0 * NX + x], (T)p[(size_t)y1 * NX + x]};
  }
  __syncthreads();
  fft_run<T, NX, false, 1, NT>(buf, tw, tid);
  const T s = T(0.5) / T(NX);
  const int k = tid;
  const cx<T> zk = buf[pidx(k)], zn = buf[pidx(k == 0 ? 0 : NX - k)];
  cx<T>* o = out + (size_t)L * NY * MX;
  o[(size_t)y0 * MX + k] = cx<T>{(zk.x + zn.x) * s, (zk.y - zn.y) * s};
  o[(size_t)y1 * MX + k] = cx<T>{(zk.y + zn.y) * s, (zn.x - zk.x) * s};
}

// x-spectral rows -> physical double rows (unnormalised c2r); grid (NY/2, nl)
template <typename T, int NX>
__global__ void __launch_bounds__(NX / 2) k_rows_inv(const cx<T>* __restrict__ in, double* out,
                                                      const cx<T>* twg, int l0) {
  constexpr int MX = NX / 2, NT = MX, NY = NX;
  __shared__ cx<T> buf[padlen(NX)];
  __shared__ cx<T> tw[NX];
  const int tid = threadIdx.x;
  load_twiddles(tw, twg, NX, tid, NT);
  const int y0 = blockIdx.x, y1 = y0 + NY / 2;
  const int L = l0 + blockIdx.y;
  const cx<T>* s0 = in + (size_t)L * NY * MX + (size_t)y0 * MX;
  const cx<T>* s1 = in + (size_t)L * NY * MX + (size_t)y1 * MX;
  const int k = tid;
  const cx<T> av = s0[k], bv = s1[k];
  if (k == 0) {
    buf[0] = cx<T>{av.x, bv.x};
    buf[pidx(MX)] = cx<T>{0, 0};
  } else {
    buf[pidx(k)] = cx<T>{av.x - bv.y, av.y + bv.x};
    buf[pidx(NX - k)] = cx<T>{av.x + bv.y, bv.x - av.y};
  }
  __syncthreads();
  fft_run<T, NX, true, 1, NT>(buf, tw, tid);
  double* p = out + (size_t)blockIdx.y * NY * NX;
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int x = tid + q * NT;
    p[(size_t)y0 * NX + x] = (double)buf[pidx(x)].x;
    p[(size_t)y1 * NX + x] = (double)buf[pidx(x)].y;
  }
}

// x-spectral small rows <-> band spectrum along y; grid (MX/TK, nl)
template <typename T, int NX, int NY, bool INV>
__global__ void __launch_bounds__(256) k_cols_small(const cx<T>* __restrict__ in, cx<T>* out,
                                                     const cx<T>* twg, int l0) {
  constexpr int MX = NX / 2, NT = 256;
  constexpr int TK = ColShape<T, NX, NY>::TK;
  __shared__ cx<T> buf[TK * colpad(NY)];
  __shared__ cx<T> tw[NY];
  const int tid = threadIdx.x;
  load_twiddles(tw, twg, NY, tid, NT);
  const int k0 = xcd_tile(blockIdx.x, gridDim.x) * TK;
  const int L = l0 + blockIdx.y;
  const cx<T>* src = in + (size_t)L * NY * MX + k0;
#pragma unroll
  for (int i = tid; i < TK * NY; i += NT) {
    const int c = i % TK, j = i / TK;
    cx<T> v = src[(size_t)j * MX + c];
    if (INV && j == NY / 2) v = cx<T>{0, 0};
    buf[c * colpad(NY) + pidx(j)] = v;
  }
  __syncthreads();
  fft_waves<T, NY, INV, TK, NT / 64, colpad(NY)>(buf, tw, tid);
  __syncthreads();
  cx<T>* dst = out + (size_t)L * NY * MX + k0;
  const T s = INV ? T(1) : T(1) / T(NY);
#pragma unroll
  for (int i = tid; i < TK * NY; i += NT) {
    const int c = i % TK, j = i / TK;
    cx<T> v = cscale(buf[c * colpad(NY) + pidx(j)], s);
    if (!INV && j == NY / 2) v = cx<T>{0, 0};
    dst[(size_t)j * MX + c] = v;
  }
}

// sums of the per-row partials: divz (comp1.f90:145) and the SGS means
// (compsgs3.f90:71-78).  grid (slots), block 256
__global__ void k_reduce_part(const double* __restrict__ part, size_t pstride, int nyb, int q0, int nq,
                              double scale, double* out, int slots) {
  __shared__ double red[256];
  const int L = blockIdx.x;
  for (int q = q0; q < nq; ++q) {
    double s = 0;
    for (int y = threadIdx.x; y < nyb; y += 256) s += part[(size_t)q * pstride + (size_t)L * nyb + y];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[(size_t)q * slots + L] = red[0] * scale;
    __syncthreads();
  }
}

// divz (comp1.f90:145) from the band spectra: the 3/2-grid plane mean of
// (ux + vy + wz)^2 is, by Parseval, the weighted power of
// i kx U + i ky V + (W(iz) - W(iz-1))/dz (weight 2 for kx > 0), with
// w = 0 at faces 0 and nnz.  grid (nz, parts): each workgroup sums a slice
// of the plane into out[iz] (cleared beforehand) with an atomic add.
template <typename T>
__global__ void __launch_bounds__(256) k_divz_band(const cx<T>* __restrict__ U, const cx<T>* __restrict__ V,
                                                   const cx<T>* __restrict__ W, const T* __restrict__ kxp,
                                                   const T* __restrict__ kyp, int mx, int ny, int nz,
                                                   double dzm, double* out) {
  __shared__ double red[256];
  const int iz = blockIdx.x + 1, tid = threadIdx.x;
  const size_t plane = (size_t)mx * ny;
  const cx<T>* u = U + (size_t)iz * plane;
  const cx<T>* v = V + (size_t)iz * plane;
  const cx<T>* w = W + (size_t)iz * plane;
  const cx<T>* wm = W + (size_t)(iz - 1) * plane;
  double s = 0.0;
  // w of face nz and of face 0 are read (valid levels) and replaced by zero,
  // so that the loads carry no branch and several iterations' loads overlap
  const bool hw = iz < nz, hm = iz > 1;
#pragma unroll 4
  for (size_t m = (size_t)blockIdx.y * 256 + tid; m < plane; m += (size_t)gridDim.y * 256) {
    const int k = (int)(m % mx), j = (int)(m / mx);
    const double kx = (double)kxp[k], ky = (double)kyp[j];
    const cx<T> uu = u[m], vv = v[m], ww = w[m], wb = wm[m];
    const double wr = hw ? (double)ww.x : 0.0, wi = hw ? (double)ww.y : 0.0;
    const double mr = hm ? (double)wb.x : 0.0, mi = hm ? (double)wb.y : 0.0;
    const double dr = -kx * (double)uu.y - ky * (double)vv.y + (wr - mr) * dzm;
    const double di = kx * (double)uu.x + ky * (double)vv.x + (wi - mi) * dzm;
    s += (k == 0 ? 1.0 : 2.0) * (dr * dr + di * di);
  }
  red[tid] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (tid < h) red[tid] += red[tid + h];
    __syncthreads();
  }
  if (tid == 0) atomicAdd(out + iz, red[0]);
}

// i kx (dir 0) or i ky (dir 1) times a band spectrum, slots 1..nz.
template <typename T>
__global__ void __launch_bounds__(256) k_band_deriv(const cx<T>* __restrict__ in, cx<T>* out,
                                                    const T* __restrict__ kxp, const T* __restrict__ kyp,
                                                    int mx, int ny, int nz, int dir) {
  const size_t plane = (size_t)mx * ny;
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= plane * nz) return;
  const size_t o = plane + i;                   // slot 1 onwards
  const int k = (int)(i % mx), j = (int)((i / mx) % ny);
  const T kk = dir == 0 ? kxp[k] : kyp[j];
  const cx<T> v = in[o];
  out[o] = cx<T>{-kk * v.y, kk * v.x};
}

template <typename A, typename B>
__global__ void k_cast(const A* __restrict__ in, B* out, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (B)in[i];
}

}  // namespace cf
