// spectra.hip -- spectral diagnostics of a run's velocity field.
//
// The reference computes them in IDL from .raw frames:
//   tools/spec.pro:25-31  u1^ = fft(u1), u2^ = fft(u2) (forward, 1/N
//                         normalised), E = (|u1^|^2 + |u2^|^2) / 2
//   tools/spec.pro:58-63  omega^ = i k1 u2^ - i k2 u1^ with the Nyquist
//                         wavenumbers zeroed, Z = |omega^|^2 / 2
//   tools/getk.pro:18-33  integer wavenumbers [0, 1, ..., -2, -1] per axis
//   tools/oned.pro:18-44  shell sums over bins [i + 1/2, i + 3/2) of |k|,
//                         bin 0 = (0, 3/2), the last bin closed, for
//                         i < n = ceil(min(n1, n2) / 2 + 0.99)
//   tools/vis.pro:57-63   vorticity field = real part of ifft(omega^)
// Here they run on the device, on the run's own state, with no frame IO:
//   1. pack: z = u1 + i u2 gathered from the AoS state (one complex field
//      carries both real components);
//   2. one single-precision complex 2-D FFT of z (hipFFT);
//   3. shells: for every wavevector k, u1^(k) = (z^(k) + conj z^(-k)) / 2 and
//      u2^(k) = (z^(k) - conj z^(-k)) / 2i, binned by |k| into FP64 shell
//      sums held in LDS per workgroup, added in workgroup order;
//   4. vorticity: omega^ from z^ as in 3, an inverse FFT, the real part.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>
#include <algorithm>
#include <cmath>

#include "fg2_layout.h"
#include "launch.h"

namespace fg2 {

struct SpectralWork {
  int n1 = 0, n2 = 0, nbins = 0, grid = 0;
  float2 *z = nullptr, *w = nullptr;
  double *partial = nullptr;              // per workgroup: E[nbins], Z[nbins]
  unsigned int *pcount = nullptr;         // per workgroup: cells per shell
  double *sums = nullptr;                 // E[nbins], Z[nbins]
  unsigned long long *counts = nullptr;   // [nbins]
  hipfftHandle plan = 0;
  bool planned = false;
};

namespace {

// Integer wavenumber of index q on an axis of n points (tools/getk.pro:24-25).
__device__ __forceinline__ int wavenumber(int q, int n) { return q < n - n / 2 ? q : q - n; }

__global__ __launch_bounds__(256)
void pack_kernel(const float *__restrict__ u, const long s, const int n2, const int nv,
                 float2 *__restrict__ z)
{
  const int j = blockIdx.x * 256 + threadIdx.x;
  const int i = blockIdx.y;
  if(j >= n2) return;
  const float *c = u + (long)i * s + (long)j * nv;
  z[(long)i * n2 + j] = make_float2(c[1], c[2]);
}

// u1^, u2^ at k from z^(k) and z^(-k), unnormalised.
__device__ __forceinline__ void split(float2 a, float2 b, float2 &u1, float2 &u2)
{
  // u1 = (a + conj b) / 2, u2 = (a - conj b) / (2i) = -i (a - conj b) / 2
  u1 = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y - b.y));
  const float rx = 0.5f * (a.x - b.x), ry = 0.5f * (a.y + b.y);
  u2 = make_float2(ry, -rx);
}

// Bin of |k| (tools/oned.pro:21-36), or -1 outside every bin.  |k| is the
// FP32 square root of an exact integer, compared with the FP32 bin edges
// i + 1/2, as the IDL float arithmetic does.
__device__ __forceinline__ int shell_of(int k1, int k2, int nbins)
{
  const int kk = k1 * k1 + k2 * k2;
  if(kk == 0) return -1;
  const float k = __fsqrt_rn((float)kk);
  int b = (int)floorf(k - 0.5f);
  if(b < 0) b = 0;                                   // (0, 3/2) is bin 0
  if(b == nbins && k == (float)nbins + 0.5f) b = nbins - 1;   // last bin is closed
  return b < nbins ? b : -1;
}

// Each workgroup bins a grid-stride share of the wavevectors into LDS and
// writes its shell sums to its own row of `partial`; shell_reduce_kernel then
// adds the rows in workgroup order, so the result does not depend on which
// workgroup finished first.
__global__ __launch_bounds__(256)
void shell_kernel(const float2 *__restrict__ zh, const int n1, const int n2, const int nbins,
                  const double scale, double *__restrict__ partial,
                  unsigned int *__restrict__ pcount)
{
  extern __shared__ double lds[];
  double *sE = lds, *sZ = lds + nbins;
  unsigned int *sC = reinterpret_cast<unsigned int *>(lds + 2 * nbins);
  for(int b = threadIdx.x; b < nbins; b += blockDim.x) { sE[b] = 0.0; sZ[b] = 0.0; sC[b] = 0; }
  __syncthreads();
  const long n = (long)n1 * n2;
  for(long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    const int i = (int)(q / n2), j = (int)(q - (long)i * n2);
    const int k1 = wavenumber(i, n1), k2 = wavenumber(j, n2);
    const int b = shell_of(k1, k2, nbins);
    if(b < 0) continue;
    const int mi = i ? n1 - i : 0, mj = j ? n2 - j : 0;
    const float2 a = zh[q], c = zh[(long)mi * n2 + mj];
    float2 u1, u2;
    split(a, c, u1, u2);
    const double e = 0.5 * ((double)u1.x * u1.x + (double)u1.y * u1.y +
                            (double)u2.x * u2.x + (double)u2.y * u2.y);
    // omega^ = i k1 u2^ - i k2 u1^, Nyquist wavenumbers zeroed (tools/getk.pro:26-29)
    const float q1 = (n1 % 2 == 0 && i == n1 / 2) ? 0.0f : (float)k1;
    const float q2 = (n2 % 2 == 0 && j == n2 / 2) ? 0.0f : (float)k2;
    const float wr = -(q1 * u2.y - q2 * u1.y), wi = q1 * u2.x - q2 * u1.x;
    const double zz = 0.5 * ((double)wr * wr + (double)wi * wi);
    atomicAdd(&sE[b], e * scale);
    atomicAdd(&sZ[b], zz * scale);
    atomicAdd(&sC[b], 1u);
  }
  __syncthreads();
  double *pE = partial + (long)blockIdx.x * 2 * nbins;
  unsigned int *pC = pcount + (long)blockIdx.x * nbins;
  for(int b = threadIdx.x; b < nbins; b += blockDim.x) {
    pE[b] = sE[b];
    pE[nbins + b] = sZ[b];
    pC[b] = sC[b];
  }
}

// 16 shells per workgroup, 16 slices of the workgroup rows per shell: slice
// t adds rows t, t + 16, ... in order, then slice 0 adds the 16 slice sums in
// order, so the sum order is fixed.
__global__ __launch_bounds__(256)
void shell_reduce_kernel(const double *__restrict__ partial, const unsigned int *__restrict__ pcount,
                         const int groups, const int nbins, double *__restrict__ sums,
                         unsigned long long *__restrict__ counts)
{
  __shared__ double lE[16][16], lZ[16][16];
  __shared__ unsigned long long lC[16][16];
  const int t = threadIdx.x >> 4, c = threadIdx.x & 15;
  const int b = blockIdx.x * 16 + c;
  double e = 0.0, z = 0.0;
  unsigned long long n = 0;
  if(b < nbins)
    for(int g = t; g < groups; g += 16) {
      e += partial[(long)g * 2 * nbins + b];
      z += partial[(long)g * 2 * nbins + nbins + b];
      n += pcount[(long)g * nbins + b];
    }
  lE[t][c] = e; lZ[t][c] = z; lC[t][c] = n;
  __syncthreads();
  if(t == 0 && b < nbins) {
    for(int k = 1; k < 16; ++k) { e += lE[k][c]; z += lZ[k][c]; n += lC[k][c]; }
    sums[b] = e;
    sums[nbins + b] = z;
    counts[b] = n;
  }
}

__global__ __launch_bounds__(256)
void vort_kernel(const float2 *__restrict__ zh, float2 *__restrict__ w, const int n1, const int n2,
                 const float scale)
{
  const int j = blockIdx.x * 256 + threadIdx.x;
  const int i = blockIdx.y;
  if(j >= n2) return;
  const int mi = i ? n1 - i : 0, mj = j ? n2 - j : 0;
  const long q = (long)i * n2 + j;
  float2 u1, u2;
  split(zh[q], zh[(long)mi * n2 + mj], u1, u2);
  const float q1 = (n1 % 2 == 0 && i == n1 / 2) ? 0.0f : (float)wavenumber(i, n1);
  const float q2 = (n2 % 2 == 0 && j == n2 / 2) ? 0.0f : (float)wavenumber(j, n2);
  w[q] = make_float2(-(q1 * u2.y - q2 * u1.y) * scale, (q1 * u2.x - q2 * u1.x) * scale);
}

__global__ __launch_bounds__(256)
void real_kernel(const float2 *__restrict__ w, float *__restrict__ out, const long n)
{
  const long q = (long)blockIdx.x * 256 + threadIdx.x;
  if(q < n) out[q] = w[q].x;
}

int fft(SpectralWork *W, float2 *data, int direction, hipStream_t st)
{
  if(!W->planned) {
    if(hipfftPlan2d(&W->plan, W->n1, W->n2, HIPFFT_C2C) != HIPFFT_SUCCESS) return SPECTRA_FFT;
    W->planned = true;
  }
  if(hipfftSetStream(W->plan, st) != HIPFFT_SUCCESS) return SPECTRA_FFT;
  if(hipfftExecC2C(W->plan, (hipfftComplex *)data, (hipfftComplex *)data, direction) != HIPFFT_SUCCESS)
    return SPECTRA_FFT;
  return SPECTRA_OK;
}

int forward(SpectralWork *W, const float *u, long s, int nv, hipStream_t st)
{
  pack_kernel<<<dim3((W->n2 + 255) / 256, W->n1), 256, 0, st>>>(u, s, W->n2, nv, W->z);
  if(hipGetLastError() != hipSuccess) return SPECTRA_HIP;
  return fft(W, W->z, HIPFFT_FORWARD, st);
}

} // namespace

int spectral_bins(int n1, int n2)
{
  // n = ceil(kG), kG = min(n1, n2) / 2 + 0.99 (tools/oned.pro:20-21)
  return (int)std::ceil(std::min(n1, n2) / 2 + 0.99);
}

int spectral_create(SpectralWork **out, int n1, int n2)
{
  *out = nullptr;
  SpectralWork *W = new SpectralWork;
  W->n1 = n1; W->n2 = n2; W->nbins = spectral_bins(n1, n2);
  int dev = 0, cus = 256;
  if(hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const size_t n = (size_t)n1 * n2;
  W->grid = (int)std::max<size_t>(1, std::min<size_t>(2 * (size_t)cus, (n + 255) / 256));
  const size_t nb = W->nbins;
  hipError_t e = hipMalloc((void **)&W->z, n * sizeof(float2));
  if(e == hipSuccess) e = hipMalloc((void **)&W->w, n * sizeof(float2));
  if(e == hipSuccess) e = hipMalloc((void **)&W->partial, W->grid * 2 * nb * sizeof(double));
  if(e == hipSuccess) e = hipMalloc((void **)&W->pcount, W->grid * nb * sizeof(unsigned int));
  if(e == hipSuccess) e = hipMalloc((void **)&W->sums, 2 * (size_t)W->nbins * sizeof(double));
  if(e == hipSuccess) e = hipMalloc((void **)&W->counts, (size_t)W->nbins * sizeof(unsigned long long));
  if(e != hipSuccess) { spectral_free(W); return SPECTRA_NOMEM; }
  *out = W;
  return SPECTRA_OK;
}

void spectral_free(SpectralWork *W)
{
  if(!W) return;
  if(W->planned) hipfftDestroy(W->plan);
  if(W->z) (void)hipFree(W->z);
  if(W->w) (void)hipFree(W->w);
  if(W->partial) (void)hipFree(W->partial);
  if(W->pcount) (void)hipFree(W->pcount);
  if(W->sums) (void)hipFree(W->sums);
  if(W->counts) (void)hipFree(W->counts);
  delete W;
}

int spectral_shells(SpectralWork *W, const float *u, long s, int nv, double *E, double *Z,
                    unsigned long long *counts, hipStream_t st)
{
  const int nb = W->nbins;
  const size_t lds = 2 * (size_t)nb * sizeof(double) + (size_t)nb * sizeof(unsigned int);
  if(lds > 160 * 1024) return SPECTRA_SHAPE;
  if(int r = forward(W, u, s, nv, st)) return r;
  const double n = (double)W->n1 * W->n2;
  shell_kernel<<<W->grid, 256, lds, st>>>(W->z, W->n1, W->n2, nb, 1.0 / (n * n), W->partial,
                                          W->pcount);
  if(hipGetLastError() != hipSuccess) return SPECTRA_HIP;
  shell_reduce_kernel<<<(nb + 15) / 16, 256, 0, st>>>(W->partial, W->pcount, W->grid, nb,
                                                      W->sums, W->counts);
  if(hipGetLastError() != hipSuccess) return SPECTRA_HIP;
  if(E && hipMemcpyAsync(E, W->sums, nb * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess)
    return SPECTRA_HIP;
  if(Z && hipMemcpyAsync(Z, W->sums + nb, nb * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess)
    return SPECTRA_HIP;
  if(counts && hipMemcpyAsync(counts, W->counts, nb * sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, st) != hipSuccess)
    return SPECTRA_HIP;
  return hipStreamSynchronize(st) == hipSuccess ? SPECTRA_OK : SPECTRA_HIP;
}

int spectral_vorticity(SpectralWork *W, const float *u, long s, int nv, float *host,
                       hipStream_t st)
{
  if(int r = forward(W, u, s, nv, st)) return r;
  const float scale = (float)(1.0 / ((double)W->n1 * W->n2));   // the forward fft's 1/N
  vort_kernel<<<dim3((W->n2 + 255) / 256, W->n1), 256, 0, st>>>(W->z, W->w, W->n1, W->n2, scale);
  if(hipGetLastError() != hipSuccess) return SPECTRA_HIP;
  if(int r = fft(W, W->w, HIPFFT_BACKWARD, st)) return r;       // unnormalised inverse
  const long n = (long)W->n1 * W->n2;
  float *out = reinterpret_cast<float *>(W->z);                  // z is free again
  real_kernel<<<(n + 255) / 256, 256, 0, st>>>(W->w, out, n);
  if(hipGetLastError() != hipSuccess) return SPECTRA_HIP;
  if(host && hipMemcpyAsync(host, out, n * sizeof(float), hipMemcpyDeviceToHost, st) != hipSuccess)
    return SPECTRA_HIP;
  return hipStreamSynchronize(st) == hipSuccess ? SPECTRA_OK : SPECTRA_HIP;
}

} // namespace fg2
