// Fused L1 + SSIM photometric loss and its gradient for gfx950, one launch per view.
//
// Semantics of cuda_backend::fused_l1_ssim_loss (photara/src/splat/fused_ssim.cu:373-560, itself
// fused-ssim's fusedl1ssim_loss{,_backward}): 11x11 Gaussian window (sigma 1.5, zero padding
// outside the image), C1 = 0.01^2, C2 = 0.03^2, per channel
//     loss = norm * sum over valid-window pixels q of  w (1 - SSIM(q)) + (1 - w) |x(q) - y(q)|,
//     norm = photometric_weight / (3 (W - 10) (H - 10)), valid window = [5, W-5) x [5, H-5),
// and dL/dx accumulated from the per-pixel partials dSSIM/d(mu1, sigma1^2, sigma12) through the
// same window (their halo outside the valid window is zero).
//
// MI355X layout: one 256-thread workgroup per 16x16 output tile computes forward and backward in
// one pass, with no derivative maps in HBM. For each channel it stages the 36x36 image window
// (tile + 10 px) of both images in LDS, runs the separable moment convolution on the 26x26
// region (tile + 5 px) whose SSIM partials the tile's gradient needs, then the separable
// backward convolution down to the tile. The moments of the 5 px ring are recomputed by the
// neighbouring tiles (2.6x the moment arithmetic of a two-pass scheme) in exchange for reading
// two images and writing one gradient instead of round-tripping three partial maps per channel.
// The loss is summed over each tile's own valid pixels into one partial per workgroup, reduced
// in a fixed order by a second one-block kernel.
#include <mutex>

#include "launch.h"

namespace pht {

// Normalised Gaussian, sigma = 1.5, radius 5: w_i = exp(-(i-5)^2 / 4.5) / sum_j exp(-(j-5)^2 / 4.5),
// rounded to float (the window of fused_ssim.cu:21-24).
__constant__ float c_ssim_w[11] = {0.001028380123898387f, 0.0075987582094967365f, 0.036000773310661316f,
                                   0.10936068743467331f,  0.21300552785396576f,   0.26601171493530273f,
                                   0.21300552785396576f,  0.10936068743467331f,   0.036000773310661316f,
                                   0.0075987582094967365f, 0.001028380123898387f};

constexpr int kSsimTile = 16;
constexpr int kSsimHalo = 5;
constexpr int kSsimIn = kSsimTile + 4 * kSsimHalo;  // 36: image window
constexpr int kSsimMid = kSsimTile + 2 * kSsimHalo; // 26: region of SSIM partials

__device__ __forceinline__ float ssim_px(const float* img, const float* mask, int x, int y, int W, int H, int64_t plane) {
    if (x < 0 || y < 0 || x >= W || y >= H) return 0.0f;
    const int64_t p = int64_t(y) * W + x;
    const float v = img[plane + p];
    return mask ? v * mask[p] : v;
}

__global__ __launch_bounds__(256) void k_l1_ssim(const float* __restrict__ color, const float* __restrict__ target,
                                                  const float* __restrict__ mask, float* __restrict__ grad,
                                                  int accumulate, float* __restrict__ partials, int W, int H,
                                                  float ssim_w, float norm) {
    __shared__ float sx[kSsimIn][kSsimIn + 1];
    __shared__ float sy[kSsimIn][kSsimIn + 1];
    __shared__ float hm[5][kSsimIn][kSsimMid];   // horizontal moment sums: x, x^2, y, y^2, xy
    __shared__ float dm[3][kSsimMid][kSsimMid + 1]; // chained partials wrt mu1, sigma1^2, sigma12
    __shared__ float hd[3][kSsimMid][kSsimTile];
    __shared__ float s_red[4];
    const int t = threadIdx.x;
    const int X0 = blockIdx.x * kSsimTile, Y0 = blockIdx.y * kSsimTile;
    const int64_t P = int64_t(W) * H;
    const float C1 = 0.01f * 0.01f, C2 = 0.03f * 0.03f;
    const float chain_ssim = -ssim_w * norm;
    float loss = 0.0f;
    for (int ch = 0; ch < 3; ++ch) {
        const int64_t plane = ch * P;
        for (int i = t; i < kSsimIn * kSsimIn; i += 256) {
            const int r = i / kSsimIn, c = i % kSsimIn;
            const int x = X0 - 2 * kSsimHalo + c, y = Y0 - 2 * kSsimHalo + r;
            sx[r][c] = ssim_px(color, mask, x, y, W, H, plane);
            sy[r][c] = ssim_px(target, mask, x, y, W, H, plane);
        }
        __syncthreads();
        // horizontal moments: rows of the 36-row window, columns of the 26-wide region
        for (int i = t; i < kSsimIn * kSsimMid; i += 256) {
            const int r = i / kSsimMid, c = i % kSsimMid;
            float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f, s4 = 0.0f;
#pragma unroll
            for (int d = 0; d < 11; ++d) {
                const float w = c_ssim_w[d];
                const float a = sx[r][c + d], b = sy[r][c + d];
                s0 += a * w;
                s1 += a * a * w;
                s2 += b * w;
                s3 += b * b * w;
                s4 += a * b * w;
            }
            hm[0][r][c] = s0;
            hm[1][r][c] = s1;
            hm[2][r][c] = s2;
            hm[3][r][c] = s3;
            hm[4][r][c] = s4;
        }
        __syncthreads();
        // vertical moments, SSIM and its chained partials on the 26x26 region
        for (int i = t; i < kSsimMid * kSsimMid; i += 256) {
            const int r = i / kSsimMid, c = i % kSsimMid;
            float m0 = 0.0f, m1 = 0.0f, m2 = 0.0f, m3 = 0.0f, m4 = 0.0f;
#pragma unroll
            for (int d = 0; d < 11; ++d) {
                const float w = c_ssim_w[d];
                m0 += hm[0][r + d][c] * w;
                m1 += hm[1][r + d][c] * w;
                m2 += hm[2][r + d][c] * w;
                m3 += hm[3][r + d][c] * w;
                m4 += hm[4][r + d][c] * w;
            }
            const int qx = X0 - kSsimHalo + c, qy = Y0 - kSsimHalo + r;
            const bool valid = qx >= kSsimHalo && qx < W - kSsimHalo && qy >= kSsimHalo && qy < H - kSsimHalo;
            float d0 = 0.0f, d1 = 0.0f, d2 = 0.0f;
            if (valid) {
                const float mu1 = m0, mu2 = m2;
                const float mu1_sq = mu1 * mu1, mu2_sq = mu2 * mu2;
                const float sigma1_sq = m1 - mu1_sq, sigma2_sq = m3 - mu2_sq, sigma12 = m4 - mu1 * mu2;
                const float A = mu1_sq + mu2_sq + C1;
                const float B = sigma1_sq + sigma2_sq + C2;
                const float Cc = 2.0f * mu1 * mu2 + C1;
                const float D = 2.0f * sigma12 + C2;
                const float ab = A * B;
                d0 = ((2.0f * mu2 * D) / ab - (2.0f * mu2 * Cc) / ab - (2.0f * mu1 * Cc * D) / (A * A * B) +
                      (2.0f * mu1 * Cc * D) / (A * B * B)) * chain_ssim;
                d1 = (-Cc * D / (A * B * B)) * chain_ssim;
                d2 = (2.0f * Cc / ab) * chain_ssim;
                const bool core = r >= kSsimHalo && r < kSsimHalo + kSsimTile && c >= kSsimHalo &&
                                  c < kSsimHalo + kSsimTile;
                if (core && partials) {
                    const float ssim = Cc * D / ab;
                    const float l1 = fabsf(sx[r + kSsimHalo][c + kSsimHalo] - sy[r + kSsimHalo][c + kSsimHalo]);
                    loss += norm * (ssim_w * (1.0f - ssim) + (1.0f - ssim_w) * l1);
                }
            }
            dm[0][r][c] = d0;
            dm[1][r][c] = d1;
            dm[2][r][c] = d2;
        }
        __syncthreads();
        // backward: horizontal pass to the tile's 16 columns, then vertical to its 16 rows
        for (int i = t; i < kSsimMid * kSsimTile; i += 256) {
            const int r = i / kSsimTile, c = i % kSsimTile;
            float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
#pragma unroll
            for (int d = 0; d < 11; ++d) {
                const float w = c_ssim_w[d];
                a0 += dm[0][r][c + d] * w;
                a1 += dm[1][r][c + d] * w;
                a2 += dm[2][r][c + d] * w;
            }
            hd[0][r][c] = a0;
            hd[1][r][c] = a1;
            hd[2][r][c] = a2;
        }
        __syncthreads();
        {
            const int r = t / kSsimTile, c = t % kSsimTile;
            const int px = X0 + c, py = Y0 + r;
            if (px < W && py < H) {
                float g0 = 0.0f, g1 = 0.0f, g2 = 0.0f;
#pragma unroll
                for (int d = 0; d < 11; ++d) {
                    const float w = c_ssim_w[d];
                    g0 += hd[0][r + d][c] * w;
                    g1 += hd[1][r + d][c] * w;
                    g2 += hd[2][r + d][c] * w;
                }
                const float x = sx[r + 2 * kSsimHalo][c + 2 * kSsimHalo];
                const float y = sy[r + 2 * kSsimHalo][c + 2 * kSsimHalo];
                const bool valid = px >= kSsimHalo && px < W - kSsimHalo && py >= kSsimHalo && py < H - kSsimHalo;
                const float chain = valid ? norm : 0.0f;
                const float sign = x == y ? 0.0f : copysignf(1.0f, x - y);
                const int64_t p = int64_t(py) * W + px;
                const float m = mask ? mask[p] : 1.0f;
                const float gv = (g0 + 2.0f * x * g1 + y * g2 + (1.0f - ssim_w) * sign * chain) * m;
                grad[plane + p] = accumulate ? grad[plane + p] + gv : gv;
            }
        }
        __syncthreads();
    }
    if (partials) {
        loss = wave_sum_to_last(loss);
        if (lane_id() == 63) s_red[t >> 6] = loss;
        __syncthreads();
        if (t == 0) partials[blockIdx.y * gridDim.x + blockIdx.x] = (s_red[0] + s_red[1]) + (s_red[2] + s_red[3]);
    }
}

// Fixed-order sum of the per-tile partials, added to *loss_sum.
__global__ __launch_bounds__(256) void k_sum_partials(const float* partials, int n, float* loss_sum) {
    __shared__ float s_red[4];
    float acc = 0.0f;
    for (int i = threadIdx.x; i < n; i += 256) acc += partials[i];
    acc = wave_sum_to_last(acc);
    if (lane_id() == 63) s_red[threadIdx.x >> 6] = acc;
    __syncthreads();
    // atomic: the views of different view groups are summed from concurrent streams (within one
    // stream the kernels, and so the additions, still run in launch order)
    if (threadIdx.x == 0) atomic_add_f32(loss_sum, (s_red[0] + s_red[1]) + (s_red[2] + s_red[3]));
}

size_t l1_ssim_scratch_floats(int width, int height) {
    return size_t((width + kSsimTile - 1) / kSsimTile) * size_t((height + kSsimTile - 1) / kSsimTile);
}

void launch_l1_ssim(const float* color, const float* target, const float* mask, float* grad, int accumulate,
                    float* loss_sum, float* scratch, int width, int height, float ssim_weight,
                    float photometric_weight, hipStream_t st) {
    const float norm = photometric_weight / float(3 * (width - 2 * kSsimHalo) * (height - 2 * kSsimHalo));
    const dim3 grid((width + kSsimTile - 1) / kSsimTile, (height + kSsimTile - 1) / kSsimTile);
    float* partials = loss_sum ? scratch : nullptr;
    hipLaunchKernelGGL(k_l1_ssim, grid, dim3(256), 0, st, color, target, mask, grad, accumulate, partials, width,
                       height, ssim_weight, norm);
    if (loss_sum)
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, partials, int(grid.x * grid.y), loss_sum);
}

// ---------------------------------------------------------------- no-allocation variant
// The per-tile partials of pht_l1_ssim_loss_noalloc: a library-owned device array, one per GPU.
// Host code never passes this name to a kernel (on the host it is only a stand-in for the
// symbol): its device address is looked up once per device with hipGetSymbolAddress.
static __device__ __align__(256) float scratch[PHT_NOALLOC_SSIM_TILES];

constexpr int kMaxDevices = 64;

float* l1_ssim_static_scratch(hipError_t* err) {
    static float* ptr[kMaxDevices] = {};
    static hipError_t status[kMaxDevices] = {};
    static std::once_flag once[kMaxDevices];
    int dev = -1;
    *err = hipGetDevice(&dev);
    if (*err != hipSuccess) return nullptr;
    if (dev < 0 || dev >= kMaxDevices) {
        *err = hipErrorInvalidDevice;
        return nullptr;
    }
    // with `dev` current: each GPU holds its own copy of the array at its own address
    std::call_once(once[dev], [&] {
        void* p = nullptr;
        status[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        ptr[dev] = status[dev] == hipSuccess ? static_cast<float*>(p) : nullptr;
    });
    *err = status[dev];
    return ptr[dev];
}

} // namespace pht
