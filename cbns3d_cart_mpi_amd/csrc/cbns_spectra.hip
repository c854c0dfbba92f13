// Shell spectra of the velocity field on the device: kinetic energy,
// enstrophy and dilatation, the diagnostics of the decaying-turbulence
// workload.  The reference has none (it writes VTK files for
// post-processing); these follow the host definition the turbulence
// workload's tests use: spectral derivatives on the periodic [-pi, pi)^3
// grid, integer wavenumbers in the order of numpy's fftfreq, and shell s
// holding the modes with rint(|k|) = s.
//
//  1. k_gather_velocity: u = (rho u) / rho, v, w of every interior cell of
//     every block into three contiguous whole-domain fields [k][j][i];
//  2. one batched hipFFT real-to-complex 3-D transform of the three fields
//     (the half spectrum kx = 0 .. N0/2; every other mode is the conjugate
//     of one held, so a held mode with 0 < kx < N0/2 stands for two);
//  3. k_shell_sums: 16 workgroups per shell, each over a fixed range of the
//     (kz, ky) lines of the half spectrum, walk on each line the kx range
//     whose |k| rounds to the shell
//     (an exact integer interval, no search), summing
//     0.5 |u_hat|^2, 0.5 |k x u_hat|^2 and 0.5 |k . u_hat|^2 of each held
//     mode and of the conjugate partner it stands for (amplitudes divided by
//     the cell count; the partner's wave vector with numpy's labels, which
//     are not the held mode's negation at a Nyquist index) with a fixed
//     per-thread order and a fixed tree; k_shell_final adds the 16 partial
//     sums of a shell in order: the result does not depend on scheduling.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include "cbns_kernels.h"

namespace cbns {

__global__ __launch_bounds__(256) void k_gather_velocity(const BlockDesc* __restrict__ blocks, int b, int which,
                                                          int N0, int N1, long long ntot, double* __restrict__ vel) {
  const BlockDesc& B = blocks[b];
  const long long n = (long long)B.n[0] * B.n[1] * B.n[2];
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % B.n[0]);
  const int j = (int)((t / B.n[0]) % B.n[1]);
  const int k = (int)(t / ((long long)B.n[0] * B.n[1]));
  const long long c = (i + NG) + (long long)B.g[0] * ((j + NG) + (long long)B.g[1] * (k + NG));
  const double* Q = B.Q[which];
  const double rho = Q[c];
  const long long o = (B.o[0] + i) + (long long)N0 * ((B.o[1] + j) + (long long)N1 * (B.o[2] + k));
  for (int a = 0; a < 3; a++) vel[a * ntot + o] = Q[(a + 1) * B.sg + c] / rho;
}

// numpy fftfreq order: index m of n points -> wavenumber
__device__ __forceinline__ int wavenumber(int m, int n) { return m <= (n - 1) / 2 ? m : m - n; }

// Largest integer whose square is at most v (v >= 0), exact.
__device__ __forceinline__ long long isqrt_floor(long long v) {
  long long r = (long long)sqrt((double)v);
  while (r * r > v) r--;
  while ((r + 1) * (r + 1) <= v) r++;
  return r;
}

constexpr int SPEC_TPB = 256;
constexpr int SPEC_CHUNKS = 16;  // workgroups per shell, each over a fixed range of (kz, ky) lines

// part[(a nshell + s) SPEC_CHUNKS + c], a = 0 kinetic energy, 1 enstrophy,
// 2 dilatation: the sums of shell s = blockIdx.x over chunk c = blockIdx.y
// of the (kz, ky) lines.  hat: the three transforms, each [N2][N1][N0/2+1].
__global__ __launch_bounds__(SPEC_TPB) void k_shell_sums(const double2* __restrict__ hat, int N0, int N1, int N2,
                                                         double inv2, int nshell, double* __restrict__ part) {
  const int s = blockIdx.x, c = blockIdx.y;
  const int H = N0 / 2 + 1;
  const long long hs = (long long)N2 * N1 * H;  // one transform
  const long long nl = (long long)N2 * N1;
  const long long l0 = nl * c / SPEC_CHUNKS, l1 = nl * (c + 1) / SPEC_CHUNKS;
  // rint(sqrt(k2)) = s  <=>  s^2 - s + 1 <= k2 <= s^2 + s  (k2 an integer, no ties)
  const long long lo = s == 0 ? 0 : (long long)s * s - s + 1, hi = (long long)s * s + s;
  double e[3] = {0.0, 0.0, 0.0};
  for (long long line = l0 + threadIdx.x; line < l1; line += SPEC_TPB) {
    const int jz = (int)(line / N1), jy = (int)(line % N1);
    const long long kz = wavenumber(jz, N2), ky = wavenumber(jy, N1);
    const long long r2 = kz * kz + ky * ky;
    if (r2 > hi) continue;
    long long m0 = lo - r2 > 0 ? isqrt_floor(lo - r2 - 1) + 1 : 0;  // smallest m with m^2 >= lo - r2
    long long m1 = isqrt_floor(hi - r2);
    if (m1 > H - 1) m1 = H - 1;
    // the conjugate partner (-kx, -ky, -kz) of a held mode carries numpy's
    // labels: -ky, except at a Nyquist index, which fftfreq labels -N/2 on
    // both sides (so the partner's k . u_hat and k x u_hat differ there)
    const double pky = (N1 % 2 == 0 && 2 * jy == N1) ? (double)ky : -(double)ky;
    const double pkz = (N2 % 2 == 0 && 2 * jz == N2) ? (double)kz : -(double)kz;
    const long long base = (jz * (long long)N1 + jy) * H;
    for (long long m = m0; m <= m1; m++) {
      const bool nyq = (N0 % 2 == 0) && m == N0 / 2;
      const bool two = !(m == 0 || nyq);  // the held mode stands for its partner too
      const double kx = nyq ? -(double)(N0 / 2) : (double)m;
      const double2 u = hat[base + m], v = hat[hs + base + m], q = hat[2 * hs + base + m];
      const double h = 0.5 * inv2;
      e[0] += (two ? 2.0 : 1.0) * h * (u.x * u.x + u.y * u.y + v.x * v.x + v.y * v.y + q.x * q.x + q.y * q.y);
      // |k x u_hat|^2 and |k . u_hat|^2 of wave vector (a, b, c) on this mode's
      // amplitudes (the partner's amplitudes are their conjugates, which
      // change no magnitude; the factor i neither)
      const auto cross_dot = [&](double a, double b, double c, double& cr, double& dt) {
        const double cx_r = b * q.x - c * v.x, cx_i = b * q.y - c * v.y;
        const double cy_r = c * u.x - a * q.x, cy_i = c * u.y - a * q.y;
        const double cz_r = a * v.x - b * u.x, cz_i = a * v.y - b * u.y;
        cr = cx_r * cx_r + cx_i * cx_i + cy_r * cy_r + cy_i * cy_i + cz_r * cz_r + cz_i * cz_i;
        const double d_r = a * u.x + b * v.x + c * q.x, d_i = a * u.y + b * v.y + c * q.y;
        dt = d_r * d_r + d_i * d_i;
      };
      double cr, dt;
      cross_dot(kx, (double)ky, (double)kz, cr, dt);
      if (two) {
        double cr2, dt2;
        cross_dot(-kx, pky, pkz, cr2, dt2);
        cr += cr2;
        dt += dt2;
      }
      e[1] += h * cr;
      e[2] += h * dt;
    }
  }
  __shared__ double red[3][SPEC_TPB];
  for (int a = 0; a < 3; a++) red[a][threadIdx.x] = e[a];
  __syncthreads();
  for (int h = SPEC_TPB / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
      for (int a = 0; a < 3; a++) red[a][threadIdx.x] += red[a][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int a = 0; a < 3; a++) part[((long long)a * nshell + s) * SPEC_CHUNKS + c] = red[a][0];
}

// out[a nshell + s]: the chunks of shell s summed in chunk order.
__global__ __launch_bounds__(256) void k_shell_final(const double* __restrict__ part, int n, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // i = a nshell + s
  if (i >= n) return;
  double x = 0.0;
  for (int c = 0; c < SPEC_CHUNKS; c++) x += part[(long long)i * SPEC_CHUNKS + c];
  out[i] = x;
}

// Device buffers and the FFT plan of one context's spectra.
struct Spectra {
  int N[3] = {0, 0, 0};
  double* vel = nullptr;   // 3 x N0 N1 N2
  double2* hat = nullptr;  // 3 x N2 N1 (N0/2+1)
  double* out = nullptr;   // 3 x nshell, then 3 x nshell x SPEC_CHUNKS partial sums
  int out_cap = 0;
  hipfftHandle plan = 0;
  bool have_plan = false;
};

Spectra* spectra_create() { return new Spectra(); }

void spectra_destroy(Spectra* sp) {
  if (!sp) return;
  if (sp->have_plan) (void)hipfftDestroy(sp->plan);
  if (sp->vel) (void)hipFree(sp->vel);
  if (sp->hat) (void)hipFree(sp->hat);
  if (sp->out) (void)hipFree(sp->out);
  delete sp;
}

// 0 ok, 1 allocation failed, 2 HIP or hipFFT error, 3 a transform past
// hipFFT's 32-bit sizes (more than 2^31 - 1 cells).
int spectra_run(Spectra* sp, const BlockDesc* dblocks, const BlockDesc* hblocks, const Domain& dom, int which,
                int nshell, double* host_out, hipStream_t s) {
  const int N0 = dom.N[0], N1 = dom.N[1], N2 = dom.N[2];
  const long long ntot = (long long)N0 * N1 * N2;
  const long long nhat = (long long)N2 * N1 * (N0 / 2 + 1);
  if (ntot > 0x7fffffffLL) return 3;
  if (!sp->have_plan) {  // buffers kept from a failed earlier attempt are reused
    if (!sp->vel && hipMalloc(&sp->vel, sizeof(double) * 3 * ntot) != hipSuccess) {
      sp->vel = nullptr;
      return 1;
    }
    if (!sp->hat && hipMalloc(&sp->hat, sizeof(double2) * 3 * nhat) != hipSuccess) {
      sp->hat = nullptr;
      return 1;
    }
    int n[3] = {N2, N1, N0};
    if (hipfftPlanMany(&sp->plan, 3, n, nullptr, 1, (int)ntot, nullptr, 1, (int)nhat, HIPFFT_D2Z, 3) !=
        HIPFFT_SUCCESS)
      return 2;
    sp->have_plan = true;
    for (int a = 0; a < 3; a++) sp->N[a] = dom.N[a];
  }
  if (nshell > sp->out_cap) {
    if (sp->out) (void)hipFree(sp->out);
    sp->out = nullptr;
    sp->out_cap = 0;
    if (hipMalloc(&sp->out, sizeof(double) * 3 * nshell * (1 + SPEC_CHUNKS)) != hipSuccess) {
      sp->out = nullptr;
      return 1;
    }
    sp->out_cap = nshell;
  }
  for (int b = 0; b < dom.nb; b++) {
    const long long n = (long long)hblocks[b].n[0] * hblocks[b].n[1] * hblocks[b].n[2];
    hipLaunchKernelGGL(k_gather_velocity, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dblocks, b, which, N0,
                       N1, ntot, sp->vel);
  }
  if (hipGetLastError() != hipSuccess) return 2;
  if (hipfftSetStream(sp->plan, s) != HIPFFT_SUCCESS) return 2;
  if (hipfftExecD2Z(sp->plan, (hipfftDoubleReal*)sp->vel, (hipfftDoubleComplex*)sp->hat) != HIPFFT_SUCCESS) return 2;
  const double inv = 1.0 / (double)ntot;
  double* part = sp->out + 3 * (size_t)nshell;
  hipLaunchKernelGGL(k_shell_sums, dim3(nshell, SPEC_CHUNKS), dim3(SPEC_TPB), 0, s, sp->hat, N0, N1, N2, inv * inv,
                     nshell, part);
  hipLaunchKernelGGL(k_shell_final, dim3((3 * nshell + 255) / 256), dim3(256), 0, s, part, 3 * nshell, sp->out);
  if (hipGetLastError() != hipSuccess) return 2;
  if (hipMemcpyAsync(host_out, sp->out, sizeof(double) * 3 * nshell, hipMemcpyDeviceToHost, s) != hipSuccess) return 2;
  if (hipStreamSynchronize(s) != hipSuccess) return 2;
  return 0;
}

}  // namespace cbns
