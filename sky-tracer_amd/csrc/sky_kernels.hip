// sky_kernels.hip -- HIP kernels of the spectral sky path tracer (gfx950).
//
//   k_simple        one thread per (pixel, lambda), all samples in a loop:
//                   the direct analogue of sky_trace.wgsl:778-816.
//   k_film_to_xyz   per-pixel sum_band film * cmf * weight.
//   k_optical_depth deterministic optical depth, 16-point Gauss-Legendre per
//                   piece of the shell walk (FP64).
// The wave-pool integrator lives in sky_wavepool.hip.
#include "sky_internal.hpp"
#include "sky_trace_common.hpp"

namespace sky {

__global__ __launch_bounds__(256) void k_simple(DevScene s)
{
    uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t total = s.width * s.height * s.nlam;
    if (idx >= total) return;
    uint32_t pixel = idx / s.nlam, band = idx - pixel * s.nlam;
    uint32_t x = pixel % s.width, y = pixel / s.width;
    float sum = 0.0f, sum2 = 0.0f;
    bool wd = false;
    for (uint32_t i = 0; i < s.spp; ++i) {
        float L = trace_path(s, x, y, s.sample_offset + i, band, wd);
        sum += L;
        sum2 += L * L;
    }
    float inv = 1.0f / (float)s.spp;
    if (s.film) s.film[idx] = fmaxf(sum * inv, 0.0f);  // integrator.rs:307-310
    if (s.film_m2) s.film_m2[idx] = sum2 * inv;
    if (wd) atomicOr(s.diag, 1u);
}

__global__ __launch_bounds__(256) void k_film_to_xyz(DevScene s)
{
    uint32_t pixel = blockIdx.x * blockDim.x + threadIdx.x;
    if (pixel >= s.width * s.height) return;
    const float *f = s.film + (size_t)pixel * s.nlam;
    float X = 0.0f, Y = 0.0f, Z = 0.0f;
    for (uint32_t b = 0; b < s.nlam; ++b) {
        float v = f[b];
        X += v * s.lam[b].wx;
        Y += v * s.lam[b].wy;
        Z += v * s.lam[b].wz;
    }
    s.xyz[3 * pixel + 0] = X;
    s.xyz[3 * pixel + 1] = Y;
    s.xyz[3 * pixel + 2] = Z;
}

hipError_t launch_simple(const DevScene &s, hipStream_t stream, LaunchShape &shape)
{
    uint32_t total = s.width * s.height * s.nlam;
    shape.block = 256;
    shape.grid = (total + 255) / 256;
    shape.items = total;
    hipLaunchKernelGGL(k_simple, dim3(shape.grid), dim3(shape.block), 0, stream, s);
    return hipGetLastError();
}

hipError_t launch_film_to_xyz(const DevScene &s, hipStream_t stream)
{
    uint32_t npx = s.width * s.height;
    hipLaunchKernelGGL(k_film_to_xyz, dim3((npx + 255) / 256), dim3(256), 0, stream, s);
    return hipGetLastError();
}

// ------------------------------------------------------- optical depth

struct OdLambda {
    double br, bo, bas, baa;
};

__device__ double od_sigma(const DevScene &s, uint32_t band, const OdLambda &l, double hr, double ha, double hc,
                           double hw, double b, double q, double t)
{
    double sv = b + t;
    double w = q + sv * sv;
    double h = w / (sqrt(s.R * s.R + w) + s.R);
    if (h < 0.0 || h > s.top) return 0.0;
    if (s.tab) {  // tabulated medium: the same breakpoint table as the tracer, lerped in FP64
        uint32_t k = tab_interval(s, (float)h);
        while (k + 2 < s.tab_k && h >= (double)s.tab_z[k + 1]) ++k;
        while (k > 0 && h < (double)s.tab_z[k]) --k;
        const double z0 = s.tab_z[k], z1 = s.tab_z[k + 1];
        double t = (h - z0) / (z1 - z0);
        t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
        const float4 a = s.tab[band * s.tab_k + k], c = s.tab[band * s.tab_k + k + 1];
        return (double)a.x + ((double)c.x - a.x) * t + (double)a.y + ((double)c.y - a.y) * t + (double)a.z +
               ((double)c.z - a.z) * t + (double)a.w + ((double)c.w - a.w) * t;
    }
    double ro = 1.0 - fabs(h - hc) / hw;
    ro = ro < 0.0 ? 0.0 : ro;
    return l.br * exp(-h / hr) + (l.bas + l.baa) * exp(-h / ha) + l.bo * ro;
}

// lamd: per wavelength [br, bo, bas, baa] (km^-1, FP64), then
// [H_R, H_A, h_c, h_w], then the 16-point Gauss-Legendre rule (8 positive
// nodes, 8 weights).  rays: n x 7 (origin, direction, t_max).
__global__ void k_optical_depth(DevScene s, const double *rays, uint32_t n, const double *lamd,
                                double *tau)
{
    uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * s.nlam) return;
    uint32_t ray = idx / s.nlam, band = idx - ray * s.nlam;
    const double *rp = rays + 7 * (size_t)ray;
    OdLambda l{lamd[4 * band], lamd[4 * band + 1], lamd[4 * band + 2], lamd[4 * band + 3]};
    const double *prof = lamd + 4 * s.nlam;
    double hr = prof[0], ha = prof[1], hc = prof[2], hw = prof[3];
    const double *gx = prof + 4, *gw = prof + 12;
    double dx = rp[3], dy = rp[4], dz = rp[5];
    double inv = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
    dx *= inv;
    dy *= inv;
    dz *= inv;
    double b = rp[0] * dx + rp[1] * dy + rp[2] * dz;
    double q = rp[0] * rp[0] + rp[1] * rp[1] + rp[2] * rp[2] - b * b - s.R * s.R;
    double t_max = rp[6];
    // Cuts: every crossing of the majorant shells and of the ozone kinks,
    // walked in order of t; each piece gets 16-point Gauss-Legendre, split so
    // that no piece exceeds 8 km.
    double kinks[3] = {hc - hw, hc, hc + hw};
    double total = 0.0;
    double t = 0.0;
    while (t < t_max) {
        double next = t_max;
        const int extra = s.tab ? (int)s.tab_k : 3;  // the table's breakpoints, or the ozone tent's kinks
        for (int k = 0; k <= (int)s.L + extra; ++k) {
            double hk = k <= (int)s.L ? k * s.dz : (s.tab ? (double)s.tab_z[k - (int)s.L - 1] : kinks[k - (int)s.L - 1]);
            double disc = hk * (2.0 * s.R + hk) - q;
            if (disc <= 0.0) continue;
            double sq = sqrt(disc);
            double t0 = -sq - b, t1 = sq - b;
            if (t0 > t * (1.0 + 1e-15) + 1e-12 && t0 < next) next = t0;
            if (t1 > t * (1.0 + 1e-15) + 1e-12 && t1 < next) next = t1;
        }
        if (-b > t + 1e-12 && -b < next) next = -b;
        int m = (int)ceil((next - t) / 8.0);
        m = m < 1 ? 1 : (m > 4096 ? 4096 : m);
        for (int j = 0; j < m; ++j) {
            double a = t + (next - t) * j / m, e = t + (next - t) * (j + 1) / m;
            double c = 0.5 * (a + e), hwid = 0.5 * (e - a), acc = 0.0;
            for (int i = 0; i < 8; ++i)
                acc += gw[i] * (od_sigma(s, band, l, hr, ha, hc, hw, b, q, c - hwid * gx[i]) +
                                od_sigma(s, band, l, hr, ha, hc, hw, b, q, c + hwid * gx[i]));
            total += acc * hwid;
        }
        t = next;
    }
    tau[idx] = total;
}

hipError_t launch_optical_depth(const DevScene &s, const double *rays, uint32_t n,
                                const double *dev_lambda_d, double *tau, hipStream_t stream)
{
    uint32_t total = n * s.nlam;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(k_optical_depth, dim3((total + 127) / 128), dim3(128), 0, stream, s, rays, n,
                       dev_lambda_d, tau);
    return hipGetLastError();
}

}  // namespace sky
