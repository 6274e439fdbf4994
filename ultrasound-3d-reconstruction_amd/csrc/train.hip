// Training-step kernels around the renderers, so a whole step (render, loss, adjoint, Adam) can be captured in
// one hipGraph without host synchronisation: per-frame L1 + global-SSIM loss with its gradient, and Adam.
#include <algorithm>
#include <cmath>

#include "common.h"

namespace us3d {

namespace {

constexpr int kLossThreads = 256;
constexpr int kLossBlocksPerFrame = 64;

// Pass 1: per (frame, block) partial sums in FP64 of |p-g|, p, g, p^2, g^2, p*g over a strided slice of the
// frame's pixels. Partials are written, not accumulated, so the result does not depend on block order.
__global__ __launch_bounds__(kLossThreads) void loss_partial_kernel(const float* __restrict__ pred,
                                                                    const float* __restrict__ gt,
                                                                    const int32_t* __restrict__ frame_base,
                                                                    double* __restrict__ partial) {
    const int f = blockIdx.y;
    const int b = blockIdx.x;
    const int64_t p0 = frame_base[f], p1 = frame_base[f + 1];
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t i = p0 + (int64_t)b * kLossThreads + threadIdx.x; i < p1; i += (int64_t)kLossThreads * gridDim.x) {
        const double x = pred[i], y = gt[i];
        s[0] += fabs(x - y);
        s[1] += x;
        s[2] += y;
        s[3] += x * x;
        s[4] += y * y;
        s[5] += x * y;
    }
    __shared__ double red[6][kLossThreads / kWave];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        double v = s[k];
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
        if (lane == 0) red[k][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        double v = 0;
        for (int w = 0; w < kLossThreads / kWave; ++w) v += red[threadIdx.x][w];
        partial[((int64_t)f * gridDim.x + b) * 8 + threadIdx.x] = v;
    }
}

// Frame statistics (the six sums over the frame's P pixels) -> loss and the frame's gradient coefficients
// d/dx_i = cs * sign(x_i - y_i) + ca + cb * x_i + cc * y_i.
__device__ void finish_frame(const double (&s)[6], double P, float w_l1, float w_ssim, float scale, float* loss,
                             double* coeff) {
    const double c1 = 0.01 * 0.01, c2 = 0.03 * 0.03;
    const double mx = s[1] / P, my = s[2] / P;
    const double vx = s[3] / P - mx * mx, vy = s[4] / P - my * my, vxy = s[5] / P - mx * my;
    const double A = 2 * mx * my + c1, B = 2 * vxy + c2, C = mx * mx + my * my + c1, D = vx + vy + c2;
    const double ssim = (A * B) / (C * D);
    *loss = (float)(w_l1 * s[0] / P + w_ssim * (1.0 - ssim));
    // d ssim / d x_i = ssim * [2 my/(A P) + 2 (y_i - my)/(B P) - 2 mx/(C P) - 2 (x_i - mx)/(D P)]
    const double k = -w_ssim * ssim * scale / P;
    coeff[0] = w_l1 * scale / P;
    coeff[1] = k * (2 * my / A - 2 * my / B - 2 * mx / C + 2 * mx / D);
    coeff[2] = k * (-2.0 / D);
    coeff[3] = k * (2.0 / B);
}

// Pass 2 (one thread per frame): the frame's block partials summed in block order, then finish_frame.
__global__ void loss_finish_kernel(const int32_t* __restrict__ frame_base, int n_frames, int n_blocks, float w_l1,
                                   float w_ssim, float scale, const double* __restrict__ partial,
                                   float* __restrict__ loss, double* __restrict__ coeff) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n_frames) return;
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int b = 0; b < n_blocks; ++b)
        for (int k = 0; k < 6; ++k) s[k] += partial[((int64_t)f * n_blocks + b) * 8 + k];
    finish_frame(s, (double)(frame_base[f + 1] - frame_base[f]), w_l1, w_ssim, scale, loss + f, coeff + 4 * f);
}

// The same from the per-(tile, wavefront) sums of a fused forward (kLossSums doubles each): one workgroup per
// frame, thread i sums the entries i, i + 256, ... of the frame's run, then a fixed tree over the threads, so the
// result is bitwise reproducible.
__global__ __launch_bounds__(kLossThreads) void loss_finish_tiles_kernel(
    const double* __restrict__ part, const int32_t* __restrict__ frame_tile_base,
    const int32_t* __restrict__ frame_pixel_base, int waves_per_tile, float w_l1, float w_ssim, float scale,
    float* __restrict__ loss, double* __restrict__ coeff) {
    const int f = blockIdx.x;
    const int64_t e0 = (int64_t)frame_tile_base[f] * waves_per_tile, e1 = (int64_t)frame_tile_base[f + 1] * waves_per_tile;
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t e = e0 + threadIdx.x; e < e1; e += kLossThreads)
#pragma unroll
        for (int k = 0; k < 6; ++k) s[k] += part[e * kLossSums + k];
    __shared__ double red[6][kLossThreads];
#pragma unroll
    for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = s[k];
    for (int h = kLossThreads / 2; h > 0; h >>= 1) {
        __syncthreads();
        if ((int)threadIdx.x < h)
#pragma unroll
            for (int k = 0; k < 6; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + h];
    }
    if (threadIdx.x == 0) {
        double t[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) t[k] = red[k][0];
        finish_frame(t, (double)(frame_pixel_base[f + 1] - frame_pixel_base[f]), w_l1, w_ssim, scale, loss + f,
                     coeff + 4 * f);
    }
}

// Pass 3: d/dx_i for every pixel; blockIdx.y is the frame, so each block reads its frame's four coefficients once.
__global__ __launch_bounds__(kLossThreads) void loss_grad_kernel(const float* __restrict__ pred,
                                                                 const float* __restrict__ gt,
                                                                 const int32_t* __restrict__ frame_base,
                                                                 const double* __restrict__ coeff,
                                                                 float* __restrict__ d_pred) {
    const int f = blockIdx.y;
    const int64_t p0 = frame_base[f], p1 = frame_base[f + 1];
    const double cs = coeff[4 * f + 0], ca = coeff[4 * f + 1], cx = coeff[4 * f + 2], cy = coeff[4 * f + 3];
    for (int64_t i = p0 + (int64_t)blockIdx.x * kLossThreads + threadIdx.x; i < p1;
         i += (int64_t)kLossThreads * gridDim.x) {
        const double x = pred[i], y = gt[i];
        const double sg = (x > y) ? 1.0 : ((x < y) ? -1.0 : 0.0);
        d_pred[i] = (float)(cs * sg + ca + cx * x + cy * y);
    }
}

// torch.optim.Adam (no weight decay, no amsgrad): m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2;
// p -= (lr / (1 - b1^t)) * m / (sqrt(v) / sqrt(1 - b2^t) + eps), t = step + 1. Skipped when *skip != 0.
__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, int64_t n, const int32_t* __restrict__ step,
                            const int32_t* __restrict__ skip, float lr, float b1, float b2, float eps) {
    if (skip && *skip) return;
    const int t = *step + 1;
    const float bc1 = 1.0f - powf(b1, (float)t);
    const float bc2 = 1.0f - powf(b2, (float)t);
    const float step_size = lr / bc1;
    const float bc2_sqrt = sqrtf(bc2);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= step_size * mi / (sqrtf(vi) / bc2_sqrt + eps);
    }
}

__global__ void step_increment_kernel(int32_t* step, const int32_t* skip) {
    if (!(skip && *skip)) *step += 1;
}

}  // namespace

int image_loss_from_tiles(const double* part, const int32_t* frame_tile_base, const int32_t* frame_pixel_base,
                          int n_frames, int waves_per_tile, float w_l1, float w_ssim, float scale, float* loss,
                          double* coeff, hipStream_t stream) {
    if (n_frames <= 0) return US3D_OK;
    hipLaunchKernelGGL(loss_finish_tiles_kernel, dim3(n_frames), dim3(kLossThreads), 0, stream, part,
                       frame_tile_base, frame_pixel_base, waves_per_tile, w_l1, w_ssim, scale, loss, coeff);
    return check_launch("loss_finish_tiles_kernel");
}

}  // namespace us3d

using namespace us3d;

extern "C" {

int64_t us3d_image_loss_tmp_bytes(int32_t n_frames) {
    return (int64_t)n_frames * (kLossBlocksPerFrame * 8 + 4) * (int64_t)sizeof(double);
}

int us3d_image_loss(const float* pred, const float* gt, const int32_t* frame_pixel_base, int32_t n_frames,
                    float w_l1, float w_ssim, float scale, float* loss, float* d_pred, double* tmp, void* stream_) {
    if (!pred || !gt || !frame_pixel_base || !loss || !d_pred || !tmp || n_frames <= 0)
        return fail(US3D_ERR_ARGUMENT, "image_loss: bad argument");
    if (n_frames > 65535) return fail(US3D_ERR_ARGUMENT, "image_loss: at most 65535 frames per call");
    hipStream_t stream = as_stream(stream_);
    double* partial = tmp;
    double* coeff = tmp + (int64_t)n_frames * kLossBlocksPerFrame * 8;
    {
        StageTimer timer_("loss_partial_kernel", stream);
        hipLaunchKernelGGL(loss_partial_kernel, dim3(kLossBlocksPerFrame, n_frames), dim3(kLossThreads), 0, stream,
                           pred, gt, frame_pixel_base, partial);
    }
    int rc = check_launch("loss_partial_kernel");
    if (rc) return rc;
    hipLaunchKernelGGL(loss_finish_kernel, dim3((n_frames + 63) / 64), dim3(64), 0, stream, frame_pixel_base, n_frames,
                       kLossBlocksPerFrame, w_l1, w_ssim, scale, partial, loss, coeff);
    if ((rc = check_launch("loss_finish_kernel"))) return rc;
    {
        StageTimer timer_("loss_grad_kernel", stream);
        hipLaunchKernelGGL(loss_grad_kernel, dim3(kLossBlocksPerFrame, n_frames), dim3(kLossThreads), 0, stream, pred,
                           gt, frame_pixel_base, coeff, d_pred);
    }
    return check_launch("loss_grad_kernel");
}

int us3d_adam_step(float* param, const float* grad, float* exp_avg, float* exp_avg_sq, int64_t n, int32_t* step_dev,
                   const int32_t* skip_dev, float lr, float beta1, float beta2, float eps, void* stream_) {
    if (!param || !grad || !exp_avg || !exp_avg_sq || !step_dev || n < 0) return fail(US3D_ERR_ARGUMENT, "adam: bad argument");
    hipStream_t stream = as_stream(stream_);
    if (n > 0) {
        const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
        hipLaunchKernelGGL(adam_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, param, grad, exp_avg, exp_avg_sq,
                           n, (const int32_t*)step_dev, skip_dev, lr, beta1, beta2, eps);
        int rc = check_launch("adam_kernel");
        if (rc) return rc;
    }
    return US3D_OK;
}

int us3d_adam_advance(int32_t* step_dev, const int32_t* skip_dev, void* stream_) {
    if (!step_dev) return fail(US3D_ERR_ARGUMENT, "adam_advance: step is NULL");
    hipLaunchKernelGGL(step_increment_kernel, dim3(1), dim3(1), 0, as_stream(stream_), step_dev, skip_dev);
    return check_launch("step_increment_kernel");
}

}  // extern "C"
