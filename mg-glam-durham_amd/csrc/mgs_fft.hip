// mgs_fft.hip -- grid-level pieces around the solve:
//   * seeded lognormal density (the synthetic input of every benchmark
//     configuration; no reference counterpart): Philox-4x32-10 white noise,
//     hipFFT D2Z, k-space filter, Z2D, normalisation to sigma_g, exp map;
//   * the fifth-force potential psi and its central-difference gradient
//     (POTENT_FOR_MG, PMP2main.f90:884-950, and MOVE's stencil :565-594);
//   * the power spectrum with POWERfft5's options: NGP or linear (CIC)
//     k-space bins, monopole or quadrupole, and its log bins
//     (PMP2mod_power.f90:323-456);
//   * the PM step's potential: POTENTfft5's discrete-Green's-function
//     Poisson solve (PMP2main.f90:705-867) and POTENT_FOR_MG's addition of
//     the scalar field (:869-935), into the handle's REAL*4 potential.

#include <hipfft/hipfft.h>

#include <cmath>
#include <vector>

#include "mgs_device.h"
#include "mgs_internal.h"
#include "mgs_kernels.h"

namespace mgs {

struct FftState {
    int N = 0;
    hipfftHandle fwd = 0, inv = 0;
    bool have_fwd = false, have_inv = false;
    double* real = nullptr;          // N^3 lexicographic
    hipfftDoubleComplex* cplx = nullptr;  // N*N*(N/2+1)
    double* bins = nullptr;          // 6 (N/2 + 1): power, count, sum|m| of the linear and the log bins
    double* binpart = nullptr;       // BIN_BLOCKS such histograms, one per block
    int* logbin = nullptr;           // log bin of every |m|^2 = 0 .. 3 (N/2)^2
};

// blocks of the binning kernel: each accumulates a private histogram in LDS
static const int BIN_BLOCKS = 512;

static FftState* fft_of(mgs_handle* h) { return (FftState*)h->fft; }

static int fft_ensure(mgs_handle* h)
{
    if (!h->fft) h->fft = new FftState();
    FftState* f = fft_of(h);
    const int N = h->p.ngrid;
    if (f->N == N && f->real) return 0;
    f->N = N;
    const size_t n3 = (size_t)N * N * N;
    const size_t nc = (size_t)N * N * (N / 2 + 1);
    if (hipMalloc((void**)&f->real, n3 * sizeof(double)) != hipSuccess) return set_err(h, MGS_E_NOMEM, "FFT buffer");
    if (hipMalloc((void**)&f->cplx, nc * sizeof(hipfftDoubleComplex)) != hipSuccess)
        return set_err(h, MGS_E_NOMEM, "FFT buffer");
    if (hipMalloc((void**)&f->bins, 3 * (size_t)(N + 2) * sizeof(double)) != hipSuccess)
        return set_err(h, MGS_E_NOMEM, "FFT bins");
    if (hipMalloc((void**)&f->binpart, (size_t)BIN_BLOCKS * 3 * (N + 2) * sizeof(double)) != hipSuccess)
        return set_err(h, MGS_E_NOMEM, "FFT bins");
    if (hipfftPlan3d(&f->fwd, N, N, N, HIPFFT_D2Z) != HIPFFT_SUCCESS) return set_err(h, MGS_E_HIP, "hipfftPlan3d D2Z");
    f->have_fwd = true;
    if (hipfftPlan3d(&f->inv, N, N, N, HIPFFT_Z2D) != HIPFFT_SUCCESS) return set_err(h, MGS_E_HIP, "hipfftPlan3d Z2D");
    f->have_inv = true;
    hipfftSetStream(f->fwd, h->s);
    hipfftSetStream(f->inv, h->s);
    return 0;
}

void fft_release(mgs_handle* h)
{
    FftState* f = fft_of(h);
    if (!f) return;
    if (f->have_fwd) hipfftDestroy(f->fwd);
    if (f->have_inv) hipfftDestroy(f->inv);
    if (f->real) (void)hipFree(f->real);
    if (f->cplx) (void)hipFree(f->cplx);
    if (f->bins) (void)hipFree(f->bins);
    if (f->binpart) (void)hipFree(f->binpart);
    if (f->logbin) (void)hipFree(f->logbin);
    delete f;
    h->fft = nullptr;
}

// ---------------------------------------------------------------------------
// Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void philox4x32(uint32_t c[4], uint32_t k0, uint32_t k1)
{
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)c[0] * 0xD2511F53ull;
        const uint64_t p1 = (uint64_t)c[2] * 0xCD9E8D57ull;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n1 = lo1, n2 = hi0 ^ c[3] ^ k1, n3 = lo0;
        c[0] = n0;
        c[1] = n1;
        c[2] = n2;
        c[3] = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// one standard normal per cell: counter (index lo, index hi, 0, 0), key = seed;
// Box-Muller on two 53-bit uniforms, cosine branch
__global__ __launch_bounds__(256) void k_white_noise(double* __restrict__ w, size_t n3, uint32_t k0, uint32_t k1)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += stride) {
        uint32_t c[4] = {(uint32_t)i, (uint32_t)(i >> 32), 0u, 0u};
        philox4x32(c, k0, k1);
        const double u1 = ((double)(c[0] >> 5) * 67108864.0 + (double)(c[1] >> 6) + 1.0) / 9007199254740992.0;
        const double u2 = ((double)(c[2] >> 5) * 67108864.0 + (double)(c[3] >> 6)) / 9007199254740992.0;
        w[i] = sqrt(-2.0 * log(u1)) * cos(2.0 * 3.14159265358979323846 * u2);
    }
}

__device__ __forceinline__ int fold(int a, int N) { return a <= N / 2 ? a : a - N; }

// multiply the half spectrum by k^(slope/2) exp(-k^2 / (4 kc^2)), k = 0 -> 0
__global__ __launch_bounds__(256) void k_filter(hipfftDoubleComplex* __restrict__ c, int N, double kf, double kc,
                                                double slope)
{
    const int nh = N / 2 + 1;
    const size_t total = (size_t)N * N * nh;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += stride) {
        const int i = (int)(q % nh);
        const size_t r = q / nh;
        const int j = (int)(r % N), k = (int)(r / N);
        const double mx = i, my = fold(j, N), mz = fold(k, N);
        const double kk = kf * sqrt(mx * mx + my * my + mz * mz);
        double a = 0.0;
        if (kk > 0.0) a = pow(kk, 0.5 * slope) * exp(-kk * kk / (4.0 * kc * kc));
        c[q].x *= a;
        c[q].y *= a;
    }
}

// block sums of g and of (g - mu)^2
__global__ __launch_bounds__(256) void k_moment(const double* __restrict__ g, size_t n3, double mu, int power,
                                                double* __restrict__ partials)
{
    double acc = 0.0;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += stride) {
        const double v = g[i] - mu;
        acc += power == 1 ? v : v * v;
    }
    const double t = block_sum(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// delta = exp((g - mu) * scale - sigma^2/2) - 1, written into the split layout
__global__ __launch_bounds__(256) void k_lognormal(const double* __restrict__ g, double* __restrict__ rho, Geo geo,
                                                   double mu, double scale, double half_var)
{
    const size_t n3 = 2 * geo.half;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n3; q += stride) {
        const int i = (int)(q % geo.n);
        const size_t r = q / geo.n;
        const int j = (int)(r % geo.n), k = (int)(r / geo.n);
        rho[split_index(geo, i, j, k)] = exp((g[q] - mu) * scale - half_var) - 1.0;
    }
}

// fifth-force potential at one cell from the split field
__device__ __forceinline__ double psi_at(const double* __restrict__ u, const Geo& g, int i, int j, int k, int model,
                                         int frn, double fct, double inv_ubg)
{
    const double v = u[split_index(g, i, j, k)];
    if (model == MGS_MODEL_DGP) return 0.5 * v;
    const double x = v * inv_ubg;
    const double w = frn == 0 ? x : (frn == 1 ? x * x : x * x * x);
    return -fct * w;
}

template <typename T>
__global__ __launch_bounds__(256) void k_fifth_force(const double* __restrict__ u, T* __restrict__ psi,
                                                     T* __restrict__ gx, T* __restrict__ gy, T* __restrict__ gz, Geo g,
                                                     int model, int frn, double fct, double inv_ubg)
{
    const size_t n3 = 2 * g.half;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n3; q += stride) {
        const int i = (int)(q % g.n);
        const size_t r = q / g.n;
        const int j = (int)(r % g.n), k = (int)(r / g.n);
        const int n = g.n;
        if (psi) psi[q] = (T)psi_at(u, g, i, j, k, model, frn, fct, inv_ubg);
        if (gx)
            gx[q] = (T)(0.5 * (psi_at(u, g, wrap_inc(i, n), j, k, model, frn, fct, inv_ubg) -
                               psi_at(u, g, wrap_dec(i, n), j, k, model, frn, fct, inv_ubg)));
        if (gy)
            gy[q] = (T)(0.5 * (psi_at(u, g, i, wrap_inc(j, n), k, model, frn, fct, inv_ubg) -
                               psi_at(u, g, i, wrap_dec(j, n), k, model, frn, fct, inv_ubg)));
        if (gz)
            gz[q] = (T)(0.5 * (psi_at(u, g, i, j, wrap_inc(k, n), model, frn, fct, inv_ubg) -
                               psi_at(u, g, i, j, wrap_dec(k, n), model, frn, fct, inv_ubg)));
    }
}

// POWERfft5's binning (PMP2mod_power.f90:359-410) of the half spectrum (x
// fastest, N/2+1 of it; an entry stands for itself and, but at i = 0 and
// i = N/2, its conjugate mirror), one private LDS histogram per block written
// to part[block]: power, mode count and sum of |m| for the linear bins 1..N/2
// (hist[0 .. nb)), then the same for the log bins (hist[3 nb .. 6 nb)).  |m|,
// the bin boundaries and the weights are the reference's REAL*4 values (RK =
// sqrt(REAL(|m|^2)) correctly rounded, as the compiled reference's is); the
// log bin of each |m|^2 comes from a table the host fills with the C
// library's log10f, the function the compiled reference calls.  The
// quadrupole's line of sight is z: the reference fills its bins after the
// I <-> K transposition (:307-316), so its MI is the z wavenumber.
// LOGB: also fill the log bins (NGP only; skipped when no log output is asked for)
template <int IPOL, int ICIC, bool LOGB>
__global__ __launch_bounds__(256) void k_bin_power(const hipfftDoubleComplex* __restrict__ c, int N, double norm,
                                                   const int* __restrict__ logbin, double* __restrict__ part, int nb)
{
    extern __shared__ double hist[];
    for (int t = threadIdx.x; t < 6 * nb; t += blockDim.x) hist[t] = 0.0;
    __syncthreads();
    const int nh = N / 2 + 1;
    const size_t total = (size_t)N * N * nh;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += stride) {
        const int i = (int)(q % nh);
        const size_t r = q / nh;
        const int j = (int)(r % N), k = (int)(r / N);
        const int my = fold(j, N), mz = fold(k, N);
        const int n2 = i * i + my * my + mz * mz;
        if (n2 == 0) continue;  // RK < 0.5: the mean
        const float rk = sqrtf((float)n2);
        const double w = (i == 0 || 2 * i == N) ? 1.0 : 2.0;  // conjugate half
        double amp = (c[q].x * c[q].x + c[q].y * c[q].y) * norm;
        if (IPOL) {
            const float cosmu2 = (float)(mz * mz) / (rk * rk);
            amp *= (double)(1.5f * cosmu2 - 0.5f);
        }
        if (ICIC) {
            int mp = (int)(rk + 0.0001f);
            if (mp < 1) mp = 1;
            const int mp2 = mp + 1 < N ? mp + 1 : N;
            const float w1 = (float)mp - rk + 1.0f;
            const float w2 = 1.0f - w1;
            const float d1 = (float)mp, d2 = d1 + 1.0f;
            if (mp < nb) {
                atomicAdd(&hist[mp], w * amp * (double)w1);
                atomicAdd(&hist[nb + mp], w * (double)w1);
                atomicAdd(&hist[2 * nb + mp], w * (double)(d1 * w1));
            }
            if (mp2 < nb) {
                atomicAdd(&hist[mp2], w * amp * (double)w2);
                atomicAdd(&hist[nb + mp2], w * (double)w2);
                atomicAdd(&hist[2 * nb + mp2], w * (double)(d2 * w2));
            }
        } else {
            int mp = (int)(rk + 0.5f);
            if (mp < 1) mp = 1;
            if (mp < nb) {
                atomicAdd(&hist[mp], w * amp);
                atomicAdd(&hist[nb + mp], w);
                atomicAdd(&hist[2 * nb + mp], w * (double)rk);
            }
            const int ml = LOGB ? logbin[n2] : nb;
            if (LOGB && ml < nb) {
                if (!IPOL) atomicAdd(&hist[3 * nb + ml], w * amp);
                atomicAdd(&hist[4 * nb + ml], w);
                atomicAdd(&hist[5 * nb + ml], w * (double)rk);
            }
        }
    }
    __syncthreads();
    double* out = part + (size_t)blockIdx.x * 6 * nb;
    for (int t = threadIdx.x; t < 6 * nb; t += blockDim.x) out[t] = hist[t];
}

// bins[t] = sum over blocks of part[block][t], in block order
__global__ __launch_bounds__(256) void k_bin_reduce(const double* __restrict__ part, int nblk, int len,
                                                    double* __restrict__ bins)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= len) return;
    double s = 0.0;
    for (int b = 0; b < nblk; b++) s += part[(size_t)b * len + t];
    bins[t] = s;
}

// POTENTfft5's Green's function, PMP2main.f90:736-742, 805-816: the discrete
// Laplacian's eigenvalue A = (G(ky) - 6 + G(kx)) + G(kz), A = 1 for the mean
// mode (|A| < 1e-7).  G is the reference's GREENf table over FFTPACK5's
// half-complex wavenumbers i = 0 .. N/2: G(0) = 2, G(N/2) = -2 and
// G(i) = 2cos(P16 i/N) between, with the reference's P16 (PMP2main.f90:711, a
// REAL*8 parameter set from the default REAL literal 6.28318530718, so it
// holds that literal's REAL*4 rounding); a complex wavenumber k > N/2 is the
// half-complex i = N - k.  The half spectrum (x fastest, N/2+1 of it) is
// multiplied by trfi / A and by 1/N^3, hipFFT's transforms being
// unnormalised where FFTPACK5's forward one divides by N per dimension.
__device__ __forceinline__ double green_g(int k, int N)
{
    const double P16 = (double)6.28318530718f;
    const int i = k <= N / 2 ? k : N - k;
    if (2 * i == N) return -2.0;
    return 2.0 * cos(P16 * i / N);
}

__global__ __launch_bounds__(256) void k_green(hipfftDoubleComplex* __restrict__ c, int N, double trfi,
                                               double inv_n3)
{
    const size_t nh = (size_t)(N / 2 + 1);
    const size_t total = nh * N * N;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += stride) {
        const int kx = (int)(q % nh);
        const size_t r = q / nh;
        const int ky = (int)(r % N), kz = (int)(r / N);
        const double gx = green_g(kx, N), gy = green_g(ky, N), gz = green_g(kz, N);
        double A = gy - 6.0;
        A = gx + A;
        A = A + gz;
        if (fabs(A) < 1.0e-7) A = 1.0;
        hipfftDoubleComplex v = c[q];
        v.x = v.x * trfi / A * inv_n3;
        v.y = v.y * trfi / A * inv_n3;
        c[q] = v;
    }
}

__global__ __launch_bounds__(256) void k_to_f32(const double* __restrict__ a, float* __restrict__ b, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) b[q] = (float)a[q];
}

__global__ __launch_bounds__(256) void k_to_f64(const float* __restrict__ a, double* __restrict__ b, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) b[q] = (double)a[q];
}

// POTENT_FOR_MG, PMP2main.f90:869-935: the REAL*4 potential (x fastest) plus
// the scalar field's contribution at each cell, summed in FP64 and rounded
// once, as FI = FI + sf does there
__global__ __launch_bounds__(256) void k_add_psi(float* __restrict__ phi, const double* __restrict__ u, Geo g,
                                                 int model, int frn, double fct, double inv_ubg)
{
    const size_t n3 = 2 * g.half;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n3; q += stride) {
        const int i = (int)(q % g.n);
        const size_t r = q / g.n;
        const int j = (int)(r % g.n), k = (int)(r / g.n);
        phi[q] = (float)((double)phi[q] + psi_at(u, g, i, j, k, model, frn, fct, inv_ubg));
    }
}

static dim3 grid_for(size_t count)
{
    size_t b = (count + 255) / 256;
    if (b > 16384) b = 16384;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}

static int sum_blocks(mgs_handle* h, double* partials, int nblk, double* out)
{
    launch_sum_partials(partials, nblk, h->dsum, h->s);
    if (hipMemcpyAsync(h->hsum, h->dsum, sizeof(double), hipMemcpyDeviceToHost, h->s) != hipSuccess ||
        hipStreamSynchronize(h->s) != hipSuccess)
        return set_err(h, MGS_E_HIP, "reduction copy");
    *out = h->hsum[0];
    return 0;
}

}  // namespace mgs

using namespace mgs;

extern "C" {

int mgs_make_lognormal_density(mgs_handle* h, uint64_t seed, double sigma_g, double slope, double kcut_frac)
{
    if (!h || !(sigma_g > 0.0) || !(kcut_frac > 0.0)) return set_err(h, MGS_E_ARG, "lognormal parameters");
    int rc = fft_ensure(h);
    if (rc) return rc;
    FftState* f = fft_of(h);
    const int N = h->p.ngrid;
    const size_t n3 = (size_t)N * N * N;
    hipLaunchKernelGGL(k_white_noise, grid_for(n3), dim3(256), 0, h->s, f->real, n3, (uint32_t)(seed & 0xFFFFFFFFu),
                       (uint32_t)(seed >> 32));
    if (hipfftExecD2Z(f->fwd, f->real, f->cplx) != HIPFFT_SUCCESS) return set_err(h, MGS_E_HIP, "hipfftExecD2Z");
    const double kf = 2.0 * M_PI / h->p.box;
    const double kc = kcut_frac * M_PI * N / h->p.box;
    hipLaunchKernelGGL(k_filter, grid_for((size_t)N * N * (N / 2 + 1)), dim3(256), 0, h->s, f->cplx, N, kf, kc, slope);
    if (hipfftExecZ2D(f->inv, f->cplx, f->real) != HIPFFT_SUCCESS) return set_err(h, MGS_E_HIP, "hipfftExecZ2D");
    const dim3 gb = grid_for(n3);
    double s1 = 0.0, s2 = 0.0;
    hipLaunchKernelGGL(k_moment, gb, dim3(256), 0, h->s, f->real, n3, 0.0, 1, h->partials);
    if ((rc = sum_blocks(h, h->partials, (int)gb.x, &s1))) return rc;
    const double mu = s1 / (double)n3;
    hipLaunchKernelGGL(k_moment, gb, dim3(256), 0, h->s, f->real, n3, mu, 2, h->partials);
    if ((rc = sum_blocks(h, h->partials, (int)gb.x, &s2))) return rc;
    const double sd = std::sqrt(s2 / (double)n3);
    if (!(sd > 0.0)) return set_err(h, MGS_E_DOMAIN, "degenerate Gaussian field");
    hipLaunchKernelGGL(k_lognormal, gb, dim3(256), 0, h->s, f->real, h->L[0].rho, make_geo(N), mu, sigma_g / sd,
                       0.5 * sigma_g * sigma_g);
    if (hipStreamSynchronize(h->s) != hipSuccess) return set_err(h, MGS_E_HIP, "lognormal density");
    h->has_density = true;
    h->rho_cached = false;
    return 0;
}

int mgs_get_density(mgs_handle* h, void* dst, int32_t dtype, int32_t where)
{
    if (!h || !dst) return set_err(h, MGS_E_ARG, "NULL argument");
    const int N = h->p.ngrid;
    const size_t bytes = (size_t)N * N * N * (dtype == MGS_F32 ? 4 : 8);
    void* out = dst;
    if (where == MGS_HOST) {
        int rc = ensure_staging(h, bytes);
        if (rc) return rc;
        out = h->staging;
    }
    if (dtype == MGS_F32)
        launch_unpack_f64_to_f32(h->L[0].rho, (float*)out, N, 1.0, h->s);
    else
        launch_unpack_f64_to_f64(h->L[0].rho, (double*)out, N, 1.0, h->s);
    if (where == MGS_HOST && hipMemcpyAsync(dst, out, bytes, hipMemcpyDeviceToHost, h->s) != hipSuccess)
        return set_err(h, MGS_E_HIP, "density copy");
    if (hipStreamSynchronize(h->s) != hipSuccess) return set_err(h, MGS_E_HIP, "density copy");
    return 0;
}

int mgs_fifth_force(mgs_handle* h, void* psi, void* gx, void* gy, void* gz, int32_t dtype, int32_t where)
{
    if (!h) return set_err(h, MGS_E_ARG, "NULL handle");
    if (dtype != MGS_F32 && dtype != MGS_F64) return set_err(h, MGS_E_ARG, "dtype");
    const int N = h->p.ngrid;
    const size_t n3 = (size_t)N * N * N;
    const size_t es = dtype == MGS_F32 ? 4 : 8;
    void* outs[4] = {psi, gx, gy, gz};
    void* dev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::vector<void*> tmp;
    for (int a = 0; a < 4; a++) {
        if (!outs[a]) continue;
        if (where == MGS_DEVICE) {
            dev[a] = outs[a];
        } else {
            void* q = nullptr;
            if (hipMalloc(&q, n3 * es) != hipSuccess) {
                for (void* t : tmp) (void)hipFree(t);
                return set_err(h, MGS_E_NOMEM, "fifth-force staging");
            }
            tmp.push_back(q);
            dev[a] = q;
        }
    }
    double ubg = 1.0;
    const double fbg = fr_background(h->p, &ubg);
    const double ctilde = 2.99792458e3 * (double)N / h->p.box;
    const double fct = ctilde * ctilde * 0.5 * fbg;
    prof_begin(h, MGS_KC_UPDATE, 0.0);
    if (dtype == MGS_F32)
        hipLaunchKernelGGL((k_fifth_force<float>), grid_for(n3), dim3(256), 0, h->s, h->L[0].u, (float*)dev[0],
                           (float*)dev[1], (float*)dev[2], (float*)dev[3], make_geo(N), h->p.model, h->p.fr_n, fct,
                           1.0 / ubg);
    else
        hipLaunchKernelGGL((k_fifth_force<double>), grid_for(n3), dim3(256), 0, h->s, h->L[0].u, (double*)dev[0],
                           (double*)dev[1], (double*)dev[2], (double*)dev[3], make_geo(N), h->p.model, h->p.fr_n,
                           fct, 1.0 / ubg);
    prof_end(h);
    int rc = 0;
    if (where == MGS_HOST)
        for (int a = 0; a < 4; a++)
            if (outs[a] && hipMemcpyAsync(outs[a], dev[a], n3 * es, hipMemcpyDeviceToHost, h->s) != hipSuccess)
                rc = set_err(h, MGS_E_HIP, "fifth-force copy");
    if (hipStreamSynchronize(h->s) != hipSuccess) rc = set_err(h, MGS_E_HIP, "fifth force");
    for (void* t : tmp) (void)hipFree(t);
    return rc;
}

// the log bin (PMP2mod_power.f90:396, dLog = 0.02 a REAL*4 parameter, :12)
// of every |m|^2 = 0 .. n2max, in the reference's REAL*4 arithmetic with the C
// library's log10f; entry 0 is never used
static std::vector<int> log_bin_table(int n2max)
{
    std::vector<int> t((size_t)n2max + 1, 1);
    const float dlog = 0.02f;
    for (int n2 = 1; n2 <= n2max; n2++) {
        const float rk = std::sqrt((float)n2);
        const volatile float v = std::log10(rk) / dlog + 1000.0f;  // rounded to REAL*4 as stored
        const int b = (int)v - 1000;
        t[n2] = b < 1 ? 1 : b;
    }
    return t;
}

// which: 0 density delta, 1 field u, 2 fifth-force potential psi
int mgs_power_spectrum_ex(mgs_handle* h, int32_t which, int32_t ipol, int32_t icic, double* k, double* pk,
                          double* count, double* klog, double* pklog, double* countlog)
{
    if (!h || !k || !pk) return set_err(h, MGS_E_ARG, "NULL argument");
    if (which < 0 || which > 2) return set_err(h, MGS_E_ARG, "which must be 0 (delta), 1 (u) or 2 (psi)");
    if (ipol != 0 && ipol != 1) return set_err(h, MGS_E_ARG, "ipol must be 0 (monopole) or 1 (quadrupole)");
    if (icic != 0 && icic != 1) return set_err(h, MGS_E_ARG, "icic must be 0 (NGP bins) or 1 (linear bins)");
    int rc = fft_ensure(h);
    if (rc) return rc;
    FftState* f = fft_of(h);
    const int N = h->p.ngrid;
    const size_t n3 = (size_t)N * N * N;
    const Geo g = make_geo(N);
    const int n2max = 3 * (N / 2) * (N / 2);
    const bool logb = !icic && (klog || pklog || countlog);  // log bins asked for
    if (logb && !f->logbin) {
        const std::vector<int> t = log_bin_table(n2max);
        if (hipMalloc((void**)&f->logbin, t.size() * sizeof(int)) != hipSuccess)
            return set_err(h, MGS_E_NOMEM, "log bin table");
        if (hipMemcpyAsync(f->logbin, t.data(), t.size() * sizeof(int), hipMemcpyHostToDevice, h->s) != hipSuccess ||
            hipStreamSynchronize(h->s) != hipSuccess)
            return set_err(h, MGS_E_HIP, "log bin table copy");
    }
    if (which == 0)
        launch_unpack_f64_to_f64(h->L[0].rho, f->real, N, 1.0, h->s);
    else if (which == 1)
        launch_unpack_f64_to_f64(h->L[0].u, f->real, N, 1.0, h->s);
    else {
        double ubg = 1.0;
        const double fbg = fr_background(h->p, &ubg);
        const double ctilde = 2.99792458e3 * (double)N / h->p.box;
        hipLaunchKernelGGL((k_fifth_force<double>), grid_for(n3), dim3(256), 0, h->s, h->L[0].u, f->real,
                           (double*)nullptr, (double*)nullptr, (double*)nullptr, g, h->p.model, h->p.fr_n,
                           ctilde * ctilde * 0.5 * fbg, 1.0 / ubg);
    }
    if (hipfftExecD2Z(f->fwd, f->real, f->cplx) != HIPFFT_SUCCESS) return set_err(h, MGS_E_HIP, "hipfftExecD2Z");
    const int nb = N / 2 + 1;  // bins 1..N/2 are returned
    const double norm = 1.0 / ((double)n3 * (double)n3);
    const size_t lds = 6 * (size_t)nb * sizeof(double);
    const dim3 G(BIN_BLOCKS), B(256);
    if (icic)
        hipLaunchKernelGGL((ipol ? k_bin_power<1, 1, false> : k_bin_power<0, 1, false>), G, B, lds, h->s, f->cplx, N,
                           norm, (const int*)f->logbin, f->binpart, nb);
    else if (logb)
        hipLaunchKernelGGL((ipol ? k_bin_power<1, 0, true> : k_bin_power<0, 0, true>), G, B, lds, h->s, f->cplx, N,
                           norm, (const int*)f->logbin, f->binpart, nb);
    else
        hipLaunchKernelGGL((ipol ? k_bin_power<1, 0, false> : k_bin_power<0, 0, false>), G, B, lds, h->s, f->cplx, N,
                           norm, (const int*)f->logbin, f->binpart, nb);
    hipLaunchKernelGGL(k_bin_reduce, dim3((6 * nb + 255) / 256), dim3(256), 0, h->s, (const double*)f->binpart,
                       BIN_BLOCKS, 6 * nb, f->bins);
    std::vector<double> hb(6 * (size_t)nb);
    if (hipMemcpyAsync(hb.data(), f->bins, hb.size() * sizeof(double), hipMemcpyDeviceToHost, h->s) != hipSuccess ||
        hipStreamSynchronize(h->s) != hipSuccess)
        return set_err(h, MGS_E_HIP, "bins copy");
    // PMP2mod_power.f90:420-456: the quadrupole times (2L+1)/2; a bin with no
    // (positive) count keeps zeros
    const double box = h->p.box, scale = 2.0 * M_PI / box, pfac = ipol ? 2.5 : 1.0;
    auto emit = [&](const double* hs, double* kk, double* pp, double* cc) {
        for (int b = 1; b <= N / 2; b++) {
            const double c = hs[nb + b];
            if (kk) kk[b - 1] = c > 0 ? hs[2 * nb + b] / c * scale : 0.0;
            if (pp) pp[b - 1] = c > 0 ? box * box * box * pfac * hs[b] / c : 0.0;
            if (cc) cc[b - 1] = c > 0 ? c : 0.0;
        }
    };
    emit(hb.data(), k, pk, count);
    if (icic) {
        for (int b = 0; b < N / 2; b++) {
            if (klog) klog[b] = 0.0;
            if (pklog) pklog[b] = 0.0;
            if (countlog) countlog[b] = 0.0;
        }
    } else {
        // the log power takes the monopole only (:398-402)
        std::vector<double> lg(hb.begin() + 3 * nb, hb.end());
        emit(lg.data(), klog, pklog, countlog);
        if (ipol && pklog)
            for (int b = 0; b < N / 2; b++) pklog[b] = 0.0;
    }
    return 0;
}

int mgs_power_spectrum(mgs_handle* h, int32_t which, double* k, double* pk, double* count)
{
    return mgs_power_spectrum_ex(h, which, 0, 0, k, pk, count, nullptr, nullptr, nullptr);
}

int mgs_potential(mgs_handle* h, double trfi)
{
    if (!h) return MGS_E_ARG;
    if (!std::isfinite(trfi)) return set_err(h, MGS_E_ARG, "trfi not finite");
    if (!h->has_density) return set_err(h, MGS_E_STATE, "no density set");
    int rc = fft_ensure(h);
    if (rc) return rc;
    float* phi = pm_potential(h);
    if (!phi) return set_err(h, MGS_E_NOMEM, "potential buffer");
    FftState* f = fft_of(h);
    const int N = h->p.ngrid;
    const double n3 = (double)N * N * N;
    prof_begin(h, MGS_KC_PM, 8.0 * n3 * 2.0 + 12.0 * n3);
    launch_unpack_f64_to_f64(h->L[0].rho, f->real, N, 1.0, h->s);
    if (hipfftExecD2Z(f->fwd, f->real, f->cplx) != HIPFFT_SUCCESS) return set_err(h, MGS_E_HIP, "hipfftExecD2Z");
    hipLaunchKernelGGL(k_green, grid_for((size_t)N * N * (N / 2 + 1)), dim3(256), 0, h->s, f->cplx, N, trfi, 1.0 / n3);
    if (hipfftExecZ2D(f->inv, f->cplx, f->real) != HIPFFT_SUCCESS) return set_err(h, MGS_E_HIP, "hipfftExecZ2D");
    // FI is REAL*4 between the reference's transform passes and after them
    hipLaunchKernelGGL(k_to_f32, grid_for((size_t)n3), dim3(256), 0, h->s, f->real, phi, (size_t)n3);
    prof_end(h);
    if (hipStreamSynchronize(h->s) != hipSuccess) return set_err(h, MGS_E_HIP, "potential");
    return 0;
}

int mgs_potential_add_fifth_force(mgs_handle* h)
{
    if (!h) return MGS_E_ARG;
    float* phi = pm_potential(h);
    if (!phi) return set_err(h, MGS_E_NOMEM, "potential buffer");
    const int N = h->p.ngrid;
    const size_t n3 = (size_t)N * N * N;
    double ubg = 1.0;
    const double fbg = fr_background(h->p, &ubg);
    const double ctilde = 2.99792458e3 * (double)N / h->p.box;
    prof_begin(h, MGS_KC_PM, 16.0 * (double)n3);
    hipLaunchKernelGGL(k_add_psi, grid_for(n3), dim3(256), 0, h->s, phi, h->L[0].u, make_geo(N), h->p.model,
                       h->p.fr_n, ctilde * ctilde * 0.5 * fbg, 1.0 / ubg);
    prof_end(h);
    if (hipStreamSynchronize(h->s) != hipSuccess) return set_err(h, MGS_E_HIP, "fifth-force potential");
    return 0;
}

int mgs_set_potential(mgs_handle* h, const void* src, int32_t dtype, int32_t where)
{
    if (!h || !src) return set_err(h, MGS_E_ARG, "NULL argument");
    if (dtype != MGS_F32 && dtype != MGS_F64) return set_err(h, MGS_E_ARG, "dtype");
    if (where != MGS_HOST && where != MGS_DEVICE) return set_err(h, MGS_E_ARG, "where");
    float* phi = pm_potential(h);
    if (!phi) return set_err(h, MGS_E_NOMEM, "potential buffer");
    const int N = h->p.ngrid;
    const size_t n3 = (size_t)N * N * N;
    const size_t bytes = n3 * (dtype == MGS_F32 ? 4 : 8);
    if (dtype == MGS_F32) {
        if (hipMemcpyAsync(phi, src, bytes, where == MGS_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice,
                           h->s) != hipSuccess)
            return set_err(h, MGS_E_HIP, "potential upload");
    } else {
        const void* d = src;
        if (where == MGS_HOST) {
            int rc = ensure_staging(h, bytes);
            if (rc) return rc;
            if (hipMemcpyAsync(h->staging, src, bytes, hipMemcpyHostToDevice, h->s) != hipSuccess)
                return set_err(h, MGS_E_HIP, "potential upload");
            d = h->staging;
        }
        hipLaunchKernelGGL(k_to_f32, grid_for(n3), dim3(256), 0, h->s, (const double*)d, phi, n3);
    }
    if (hipStreamSynchronize(h->s) != hipSuccess) return set_err(h, MGS_E_HIP, "potential upload");
    return 0;
}

int mgs_get_potential(mgs_handle* h, void* dst, int32_t dtype, int32_t where)
{
    if (!h || !dst) return set_err(h, MGS_E_ARG, "NULL argument");
    if (dtype != MGS_F32 && dtype != MGS_F64) return set_err(h, MGS_E_ARG, "dtype");
    if (where != MGS_HOST && where != MGS_DEVICE) return set_err(h, MGS_E_ARG, "where");
    if (!h->pm) return set_err(h, MGS_E_STATE, "no potential: call mgs_potential first");
    float* phi = pm_potential(h);
    const int N = h->p.ngrid;
    const size_t n3 = (size_t)N * N * N;
    const size_t bytes = n3 * (dtype == MGS_F32 ? 4 : 8);
    void* out = dst;
    if (where == MGS_HOST) {
        int rc = ensure_staging(h, bytes);
        if (rc) return rc;
        out = h->staging;
    }
    if (dtype == MGS_F64)
        hipLaunchKernelGGL(k_to_f64, grid_for(n3), dim3(256), 0, h->s, phi, (double*)out, n3);
    else if (hipMemcpyAsync(out, phi, bytes, hipMemcpyDeviceToDevice, h->s) != hipSuccess)
        return set_err(h, MGS_E_HIP, "potential copy");
    if (where == MGS_HOST && hipMemcpyAsync(dst, out, bytes, hipMemcpyDeviceToHost, h->s) != hipSuccess)
        return set_err(h, MGS_E_HIP, "potential copy");
    if (hipStreamSynchronize(h->s) != hipSuccess) return set_err(h, MGS_E_HIP, "potential copy");
    return 0;
}

}  // extern "C"
