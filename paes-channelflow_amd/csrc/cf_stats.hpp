// cf_stats.hpp -- the parity observables on the device: the mean profiles of
// compmn (compmn.F90:1-64) and the x-y spectra of calc_spectra
// (calc_spectra.f90:139-243), computed from the band spectra the state lives in.
//
// The fields are band-limited, so plane means and second moments follow
// exactly from their DFT coefficients (Parseval).  The coefficients are those
// of ffthrz, forward-scaled by 1/(nx ny) like calc_spectra's fft_out/nxy.  A
// half-spectrum coefficient with kx > 0 stands for itself and its Hermitian
// partner (weight 2); kx = 0 has weight 1.  No transform and no host copy of
// the fields is needed.
#pragma once
#include <hip/hip_runtime.h>

#include "cf_fft.hpp"

namespace cf {

constexpr int NMEAN = 10;   // uxym vxym wxym uu vv ww uw vw uv engsbz

// compmn per level iz (blockIdx.x + 1):
// - ul = ((u_iz - <u_iz>) + (u_iz+1 - <u_iz+1>))/2, with u(nnz+1) = 0
//   (upperbc.f90:12-14), and vl likewise;
// - w at face iz;
// - uu = <ul ul>, vv, ww = <w w>, uw = <ul w>, vw, uv;
// - the level means of u, v, w;
// - engsbz = <e>, from the physical e (which carries e(nnz) = e(nnz-1)).
// The means of u and v come from their FP64 copies m0u, m0v (per level).
template <typename T>
__global__ void __launch_bounds__(256) k_compmn(const cx<T>* __restrict__ U, const cx<T>* __restrict__ V,
                                                const cx<T>* __restrict__ W, const T* __restrict__ E,
                                                const double* __restrict__ m0u, const double* __restrict__ m0v,
                                                int mx, int ny, int nz, double* out) {
  __shared__ double red[6][256];
  const int iz = blockIdx.x + 1, tid = threadIdx.x;
  const size_t plane = (size_t)mx * ny;
  const cx<T>* u0 = U + (size_t)iz * plane;
  const cx<T>* v0 = V + (size_t)iz * plane;
  const cx<T>* w0 = W + (size_t)iz * plane;
  const bool top = iz == nz;
  const cx<T>* u1 = U + (size_t)(iz + 1) * plane;
  const cx<T>* v1 = V + (size_t)(iz + 1) * plane;
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (size_t m = tid; m < plane; m += 256) {
    const int k = (int)(m % mx);
    const double wt = k == 0 ? 1.0 : 2.0;
    double ur = 0.5 * (double)u0[m].x, ui = 0.5 * (double)u0[m].y;
    double vr = 0.5 * (double)v0[m].x, vi = 0.5 * (double)v0[m].y;
    if (!top) {
      ur += 0.5 * (double)u1[m].x; ui += 0.5 * (double)u1[m].y;
      vr += 0.5 * (double)v1[m].x; vi += 0.5 * (double)v1[m].y;
    }
    if (m == 0) { ur = ui = vr = vi = 0.0; }          // plane means removed
    const double wr = (double)w0[m].x, wi = (double)w0[m].y;
    s[0] += wt * (ur * ur + ui * ui);
    s[1] += wt * (vr * vr + vi * vi);
    s[2] += wt * (wr * wr + wi * wi);
    s[3] += wt * (ur * wr + ui * wi);
    s[4] += wt * (vr * wr + vi * wi);
    s[5] += wt * (ur * vr + ui * vi);
  }
#pragma unroll
  for (int q = 0; q < 6; ++q) red[q][tid] = s[q];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (tid < h)
#pragma unroll
      for (int q = 0; q < 6; ++q) red[q][tid] += red[q][tid + h];
    __syncthreads();
  }
  // engsbz: plane mean of the physical e (nx = 2 mx points per row)
  __shared__ double esum[256];
  double es = 0.0;
  const size_t pplane = (size_t)2 * mx * ny;
  for (size_t m = tid; m < pplane; m += 256) es += (double)E[(size_t)iz * pplane + m];
  esum[tid] = es;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (tid < h) esum[tid] += esum[tid + h];
    __syncthreads();
  }
  if (tid == 0) {
    double* o = out + (size_t)(iz - 1) * NMEAN;
    o[0] = m0u[iz];
    o[1] = m0v[iz];
    o[2] = (double)w0[0].x;
    for (int q = 0; q < 6; ++q) o[3 + q] = red[q][0];
    o[9] = esum[0] / (double)pplane;
  }
}

// calc_spectra for one snapshot: grid (levels, 3 fields u v w).  Level zl
// (1-based, < nnz) averages levels zl, zl+1 of u and v with their plane means
// removed; w is taken at face zl (calc_spectra.f90:147,175,202).  Per (level,
// field) it writes nb = nx/2+1 ring sums of weighted |c|^2 (ring
// r = nint(|k|) of the integer wavenumbers), then l1 (the kx = 0 column, its
// first nx/2 entries counted twice), l2 (the ky = 0 row) and var.
template <typename T>
__global__ void __launch_bounds__(256) k_spectra(const cx<T>* __restrict__ U, const cx<T>* __restrict__ V,
                                                 const cx<T>* __restrict__ W, const int* __restrict__ zl,
                                                 int mx, int ny, int nx, double* out) {
  extern __shared__ double bins[];                    // nb + 3
  const int nb = nx / 2 + 1;
  const int li = blockIdx.x, f = blockIdx.y, tid = threadIdx.x;
  for (int r = tid; r < nb + 3; r += 256) bins[r] = 0.0;
  __syncthreads();
  const size_t plane = (size_t)mx * ny;
  const int z = zl[li];
  const cx<T>* a = (f == 0 ? U : (f == 1 ? V : W)) + (size_t)z * plane;
  const cx<T>* b = a + plane;
  for (size_t m = tid; m < plane; m += 256) {
    const int k = (int)(m % mx), j = (int)(m / mx);
    double cr = (double)a[m].x, ci = (double)a[m].y;
    if (f < 2) {
      cr = 0.5 * (cr + (double)b[m].x);
      ci = 0.5 * (ci + (double)b[m].y);
      if (m == 0) cr = ci = 0.0;
    }
    const double p = cr * cr + ci * ci;
    const int kyw = j <= nx / 2 ? j : j - nx;            // waveN of calc_spectra.f90:105-112
    const int r = (int)rint(sqrt((double)(k * k + kyw * kyw)));
    const double wt = k == 0 ? 1.0 : 2.0;
    if (r <= nx / 2) atomicAdd(&bins[r], wt * p);
    if (k == 0) atomicAdd(&bins[nb], p * (j < nx / 2 ? 2.0 : 1.0));
    if (j == 0) atomicAdd(&bins[nb + 1], p);
    atomicAdd(&bins[nb + 2], wt * p);
  }
  __syncthreads();
  double* o = out + ((size_t)li * 3 + f) * (nb + 3);
  for (int r = tid; r < nb + 3; r += 256) o[r] = bins[r];
}

// calc_spectra's normalisation on the device, for one snapshot
// (calc_spectra.f90:238-257 with fcounter = 1): in holds k_spectra's ring
// sums per (level, field), [nlev][3][nb+3]; out receives hats[3][nlev][nb],
// var[3][nlev] and lscales[6][nlev] (l_f1, l_f2 per field).  The ring
// populations nCount of one (field, level) are counted here over the r2c half
// plane kx = 0..nx/2 (weight 2 for 0 < kx < nx/2), every ky
// (calc_spectra.f90:150-166).  One workgroup of 256; nx <= 512.
__global__ void __launch_bounds__(256) k_spectra_finish(const double* __restrict__ in, int nlev, int nx,
                                                        int ny, double pi, double* __restrict__ out) {
  __shared__ double cnt[512 / 2 + 1];
  const int nb = nx / 2 + 1, tid = threadIdx.x;
  for (int r = tid; r < nb; r += 256) cnt[r] = 0.0;
  __syncthreads();
  for (int i = tid; i < nb * ny; i += 256) {
    const int kx = i % nb, j = i / nb;
    const int ky = j <= nx / 2 ? j : j - nx;
    const int r = (int)rint(sqrt((double)(kx * kx + ky * ky)));
    if (r <= nx / 2) atomicAdd(&cnt[r], (kx >= 1 && kx <= nx / 2 - 1) ? 2.0 : 1.0);
  }
  __syncthreads();
  double* hats = out;
  double* var = out + (size_t)3 * nlev * nb;
  double* ls = var + (size_t)3 * nlev;
  for (int i = tid; i < 3 * nlev * nb; i += 256) {
    const int r = i % nb, fl = i / nb, l = fl % nlev, f = fl / nlev;
    const double* b = in + ((size_t)l * 3 + f) * (nb + 3);
    // the r = 0 bin is not normalised (calc_spectra.f90:239)
    hats[i] = r == 0 ? b[0] : b[r] * pi * ((r + 0.5) * (r + 0.5) - (r - 0.5) * (r - 0.5)) / cnt[r];
  }
  for (int i = tid; i < 3 * nlev; i += 256) {
    const int l = i % nlev, f = i / nlev;
    const double* b = in + ((size_t)l * 3 + f) * (nb + 3);
    const double v = b[nb + 2];
    var[i] = v;
    ls[(size_t)(2 * f) * nlev + l] = b[nb] * pi / v;           // calc_spectra.f90:252-257
    ls[(size_t)(2 * f + 1) * nlev + l] = b[nb + 1] * pi / v;
  }
}

}  // namespace cf
