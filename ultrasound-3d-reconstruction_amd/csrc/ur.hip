// UltraG-Ray B-mode ray renderer on gfx950: mapper preprocess, scanline forward, thread-per-Gaussian adjoint
// and per-Gaussian reduction. The instance stage (count, emit, per-ray LDS sort) is shared with the slice
// splat: a (frame, Gaussian) pair carries the ray range as the tile rectangle (0, min_ray, 1, max_ray + 1) of a
// frame whose "tiles" are its scanlines.
//
// Conventions for the quaternion/covariance helpers the reference includes but does not ship
// (ops/ultragray_renderer/src/shader/forward.cu:1 includes covariance.cuh, absent from its tree):
// q = (w, x, y, z), normalised inside; R row-major with the principal axes as columns (the reference
// shader computes o_local = R^T (o - mu), shared.cuh:39-47); Sigma^-1 = R diag(1/s^2) R^T. The VJPs below
// are the exact derivatives of these definitions.
#include <algorithm>

#include "common.h"

namespace us3d {

namespace {

constexpr float kXi = 0.7745966692f;
constexpr float kW0 = 0.5555555556f;
constexpr float kW1 = 0.8888888889f;

us3d_layout ur_layout(int64_t n_gauss, int64_t n_pairs, int64_t n_tiles, int64_t k_cap, int64_t v_cap) {
    return ws_layout(n_gauss, n_pairs, n_tiles, k_cap, v_cap, 64, 64, 0, 0);
}

us3d_layout ur_layout(const us3d_plan* plan) {
    return plan ? ur_layout(plan->n_gauss, plan->n_pairs, plan->n_tiles, plan->k_capacity, plan->v_capacity)
                : us3d_layout{};
}

// One record per Gaussian: (mu, raw attenuation | scales, m | quaternion | SH). With `log_scales` the scales
// array holds the training model's log-scales and gets its activation max(exp(x), 1e-5) here
// (pipeline_god_file.py:360-361); m has bit c set where exp(x_c) >= 1e-5, i.e. where the gradient passes the
// clamp (torch's clamp_min passes it where x >= min). A NaN log-scale stays NaN, as under clamp_min.
__global__ void ur_pack_kernel(int n, const float* __restrict__ means, const float* __restrict__ scales,
                               const float* __restrict__ quats, const float* __restrict__ atten,
                               const float* __restrict__ sh, int log_scales, float4* __restrict__ rec) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int64_t G = g;
    float sc[3];
    int pass = 7;
    if (log_scales) {
        pass = 0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float e = expf(scales[3 * G + c]);
            sc[c] = (e < 1e-5f) ? 1e-5f : e;
            pass |= (e >= 1e-5f ? 1 : 0) << c;
        }
    } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) sc[c] = scales[3 * G + c];
    }
    rec[4 * G + 0] = make_float4(means[3 * G], means[3 * G + 1], means[3 * G + 2], atten[G]);
    rec[4 * G + 1] = make_float4(sc[0], sc[1], sc[2], (float)pass);
    rec[4 * G + 2] = make_float4(quats[4 * G], quats[4 * G + 1], quats[4 * G + 2], quats[4 * G + 3]);
    rec[4 * G + 3] = make_float4(sh[4 * G], sh[4 * G + 1], sh[4 * G + 2], sh[4 * G + 3]);
}

// kernel_project_and_cull, ops/ultragray_renderer/src/mapper/core.cu:3-68: per (frame, Gaussian) the inclusive
// scanline range [floor(c - r/ps), ceil(c + r/ps)] clipped to [0, W-1] with r = k * max(s); culled when
// z + r < near or |elev| > r + margin (and, defined here, when the projection is not finite). Depth key =
// (uint32)(max(0, z) * 1000) (core.cu:86), saturating. Returns the ray count (0: culled) with the range as the
// rectangle (0, lo, 1, hi + 1); the cull and the setup pass call it on the same inputs.
__device__ __forceinline__ int ur_pair_rays(const float4& r0, const float4& r1, const us3d_frame& fr, float k_sigma,
                                            float elev_margin, int4& rect, uint32_t& dk) {
    const float dx = r0.x - fr.origin[0], dy = r0.y - fr.origin[1], dz = r0.z - fr.origin[2];
    const float x = dot3(dx, dy, dz, fr.axis_lat[0], fr.axis_lat[1], fr.axis_lat[2]);
    const float y = dot3(dx, dy, dz, fr.axis_ele[0], fr.axis_ele[1], fr.axis_ele[2]);
    const float z = dot3(dx, dy, dz, fr.axis_ax[0], fr.axis_ax[1], fr.axis_ax[2]);
    const float ms = fmaxf(r1.x, fmaxf(r1.y, r1.z));
    const float r = ms * k_sigma;
    if (!(z + r >= fr.near_mm && fabsf(y) <= r + elev_margin)) return 0;
    const float c = x / fr.delta_lat + fr.col_offset;
    const float rp = r / fr.delta_lat;
    const float f0 = floorf(c - rp), f1 = ceilf(c + rp);
    if (!(isfinite(f0) && isfinite(f1))) return 0;
    const int lo = max(0, (int)fminf(fmaxf(f0, -1073741824.0f), 1073741824.0f));
    const int hi = min(fr.width - 1, (int)fminf(fmaxf(f1, -1073741824.0f), 1073741824.0f));
    if (!(lo <= hi)) return 0;
    rect = make_int4(0, lo, 1, hi + 1);
    const float q = fmaxf(0.0f, z) * 1000.0f;
    dk = q >= 4294967296.0f ? 0xffffffffu : (uint32_t)q;
    return hi - lo + 1;
}

// Cull pass: thread per dense (frame, Gaussian) pair; marks the pairs that reach at least one scanline.
__global__ __launch_bounds__(256) void ur_cull_kernel(int64_t n_pairs, const float4* __restrict__ rec,
                                                      const us3d_frame* __restrict__ frames,
                                                      const us3d_sweep* __restrict__ sweeps, int n_sweeps,
                                                      float k_sigma, float elev_margin, uint64_t* __restrict__ pair_bits,
                                                      int32_t* __restrict__ word_cnt, uint32_t* __restrict__ fm_bits) {
    const int64_t pair = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int count = 0, s = 0, f = 0, gl = 0;
    if (pair < n_pairs) {
        s = sweep_of_pair(sweeps, n_sweeps, pair);
        const us3d_sweep sw = sweeps[s];
        const int local = (int)(pair - sw.pair_base);  // pair indices are < 2^31 (check_plan)
        gl = local / sw.f_count;
        const int g = sw.g_begin + gl;
        f = sw.f_begin + (local - gl * sw.f_count);
        int4 rect;
        uint32_t dk;
        count = ur_pair_rays(rec[4 * (int64_t)g + 0], rec[4 * (int64_t)g + 1], frames[f], k_sigma, elev_margin, rect,
                             dk);
    }
    mark_pair(count > 0, pair, n_pairs, sweeps, s, f, gl, pair_bits, word_cnt, fm_bits);
}

// Setup pass: thread per visible pair v < v_capacity (entries past the visible count, or every entry when the
// visible pairs overflowed, get count 0): ray count, ray range, depth key; rays counted by range ends.
__global__ __launch_bounds__(256) void ur_setup_kernel(int64_t v_cap, const int32_t* __restrict__ vis_pair,
                                                       const float4* __restrict__ rec,
                                                       const us3d_frame* __restrict__ frames,
                                                       const us3d_sweep* __restrict__ sweeps, int n_sweeps,
                                                       float k_sigma, float elev_margin,
                                                       const us3d_stats* __restrict__ stats,
                                                       int32_t* __restrict__ pair_count, int4* __restrict__ pair_rect,
                                                       uint32_t* __restrict__ pair_dk,
                                                       int32_t* __restrict__ tile_count) {
    const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (v >= v_cap) return;
    if (stats->overflow || v >= stats->n_visible) {
        pair_count[v] = 0;
        return;
    }
    const int64_t pair = vis_pair[v];
    const us3d_sweep sw = sweeps[sweep_of_pair(sweeps, n_sweeps, pair)];
    const int local = (int)(pair - sw.pair_base);
    const int gl = local / sw.f_count;
    const int g = sw.g_begin + gl;
    const us3d_frame& fr = frames[sw.f_begin + (local - gl * sw.f_count)];
    int4 rect = make_int4(0, 0, 0, 0);
    uint32_t dk = 0;
    const int count = ur_pair_rays(rec[4 * (int64_t)g + 0], rec[4 * (int64_t)g + 1], fr, k_sigma, elev_margin, rect, dk);
    pair_count[v] = count;
    if (count > 0) {  // always, for a listed pair
        pair_rect[v] = rect;
        pair_dk[v] = dk;
        add_rect_corners(tile_count, fr.tile_base, fr.tiles_x, fr.tiles_y, rect);
    }
}

// ---------------------------------------------------------------- per (ray, Gaussian) setup

__device__ __forceinline__ void quat_rot(float4 q, float R[9]) {
    const float n = sqrtf(fmaxf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w, 1e-12f));
    const float w = q.x / n, x = q.y / n, y = q.z / n, z = q.w / n;
    R[0] = 1.f - 2.f * (y * y + z * z); R[1] = 2.f * (x * y - w * z);       R[2] = 2.f * (x * z + w * y);
    R[3] = 2.f * (x * y + w * z);       R[4] = 1.f - 2.f * (x * x + z * z); R[5] = 2.f * (y * z - w * x);
    R[6] = 2.f * (x * z - w * y);       R[7] = 2.f * (y * z + w * x);       R[8] = 1.f - 2.f * (x * x + y * y);
}

// dL/dq of the unnormalised quaternion from dL/dR (through the normalisation).
__device__ __forceinline__ float4 quat_vjp(float4 q, const float dR[9]) {
    const float n = sqrtf(fmaxf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w, 1e-12f));
    const float w = q.x / n, x = q.y / n, y = q.z / n, z = q.w / n;
    const float gw = 2.f * (-z * dR[1] + y * dR[2] + z * dR[3] - x * dR[5] - y * dR[6] + x * dR[7]);
    const float gx = 2.f * (y * dR[1] + z * dR[2] + y * dR[3] - 2.f * x * dR[4] - w * dR[5] + z * dR[6] + w * dR[7] -
                            2.f * x * dR[8]);
    const float gy = 2.f * (-2.f * y * dR[0] + x * dR[1] + w * dR[2] + x * dR[3] + z * dR[5] - w * dR[6] + z * dR[7] -
                            2.f * y * dR[8]);
    const float gz = 2.f * (-2.f * z * dR[0] - w * dR[1] + x * dR[2] + w * dR[3] - 2.f * z * dR[4] + y * dR[5] +
                            x * dR[6] + y * dR[7]);
    const float dot = gw * w + gx * x + gy * y + gz * z;
    return make_float4((gw - dot * w) / n, (gx - dot * x) / n, (gy - dot * y) / n, (gz - dot * z) / n);
}

// 3-point Gauss-Legendre approximation of the ray integral over [ts, te] in whitened coordinates
// (compute_ray_gaussian_integral, shared.cuh:27-85).
__device__ __forceinline__ float gl3(const float og[3], const float dg[3], float ts, float te) {
    const float half = 0.5f * (te - ts);
    const float c = 0.5f * (te + ts);
    const float xi[3] = {-kXi, 0.0f, kXi};
    const float wq[3] = {kW0, kW1, kW0};
    float acc = 0.f;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float t = __builtin_fmaf(half, xi[i], c);
        const float x0 = __builtin_fmaf(dg[0], t, og[0]);
        const float x1 = __builtin_fmaf(dg[1], t, og[1]);
        const float x2 = __builtin_fmaf(dg[2], t, og[2]);
        acc = __builtin_fmaf(wq[i], __expf(-0.5f * dot3(x0, x1, x2, x0, x1, x2)), acc);
    }
    return half * acc;
}

// Everything about one Gaussian that is fixed along one ray.
struct RayGauss {
    float tg, rd, tc, tau;      // centre depth along the ray, echo radius, GL3 window centre, sigmoid(raw)
    float og[3], dg[3];         // whitened ray origin / direction
    float ts, b, f_final, e_final;  // window start/end, transmittance and exp(-psi) of the complete window
    float mu[3], I;             // mean, SH intensity along the ray direction
    float S00, S01, S02, S11, S12, S22;  // Sigma^-1
    float dkf;                  // sorted depth (key / 1000)
};

// Sigma^-1 = R diag(1/s^2) R^T as (S00, S01, S02, S11, S12, S22).
__device__ __forceinline__ void sigma_inv6(const float R[9], const float s[3], float S[6]) {
    const float d0 = 1.0f / (s[0] * s[0]), d1 = 1.0f / (s[1] * s[1]), d2 = 1.0f / (s[2] * s[2]);
    auto sij = [&](int i, int j) { return R[3 * i] * d0 * R[3 * j] + R[3 * i + 1] * d1 * R[3 * j + 1] + R[3 * i + 2] * d2 * R[3 * j + 2]; };
    S[0] = sij(0, 0); S[1] = sij(0, 1); S[2] = sij(0, 2);
    S[3] = sij(1, 1); S[4] = sij(1, 2); S[5] = sij(2, 2);
}

__device__ __forceinline__ void ray_gauss(const float4* __restrict__ rec, int g, uint32_t dkey, const float ro[3],
                                          const float rdir[3], float k_sigma, float Rw, RayGauss& G, float R[9]) {
    const float4 r0 = rec[4 * (int64_t)g + 0];
    const float4 r1 = rec[4 * (int64_t)g + 1];
    const float4 r2 = rec[4 * (int64_t)g + 2];
    const float4 r3 = rec[4 * (int64_t)g + 3];
    G.mu[0] = r0.x; G.mu[1] = r0.y; G.mu[2] = r0.z;
    G.tg = dot3(r0.x - ro[0], r0.y - ro[1], r0.z - ro[2], rdir[0], rdir[1], rdir[2]);
    G.rd = fmaxf(r1.x, fmaxf(r1.y, r1.z)) * k_sigma;
    quat_rot(r2, R);
    G.tau = 1.0f / (1.0f + __expf(-r0.w));
    const float diff[3] = {ro[0] - r0.x, ro[1] - r0.y, ro[2] - r0.z};
    const float s[3] = {r1.x, r1.y, r1.z};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float ol = __builtin_fmaf(R[6 + c], diff[2], __builtin_fmaf(R[3 + c], diff[1], R[c] * diff[0]));
        const float dl = __builtin_fmaf(R[6 + c], rdir[2], __builtin_fmaf(R[3 + c], rdir[1], R[c] * rdir[0]));
        G.og[c] = ol / s[c];
        G.dg[c] = dl / s[c];
    }
    const float dn2 = dot3(G.dg[0], G.dg[1], G.dg[2], G.dg[0], G.dg[1], G.dg[2]);
    G.tc = -dot3(G.og[0], G.og[1], G.og[2], G.dg[0], G.dg[1], G.dg[2]) / (dn2 + 1e-8f);
    G.ts = fmaxf(0.0f, G.tc - Rw);
    G.b = G.tc + Rw;
    float psi = 0.f;
    if (Rw > 0.0f && G.b > 0.0f && G.b > G.ts) psi = gl3(G.og, G.dg, G.ts, G.b);
    G.e_final = __expf(-psi);
    G.f_final = __builtin_fmaf(1.0f - G.tau, G.e_final, G.tau);
    G.I = r3.x + r3.y * rdir[0] + r3.z * rdir[1] + r3.w * rdir[2];
    float S6[6];
    sigma_inv6(R, s, S6);
    G.S00 = S6[0]; G.S01 = S6[1]; G.S02 = S6[2];
    G.S11 = S6[3]; G.S12 = S6[4]; G.S22 = S6[5];
    G.dkf = (float)dkey * 0.001f;
}

struct RayFlags {
    float k_sigma, Rw, local_max, grad_threshold;
    int local_only;
};

// Which terms of pixel (depth tpx) this Gaussian enters (forward.cu:76-96).
template <typename Hit>
__device__ __forceinline__ void ray_terms(const Hit& G, float tpx, const RayFlags& F, bool& trans, bool& echo) {
    bool inside = true;
    if (F.local_only) inside = (G.dkf >= tpx - F.local_max) && !(G.dkf > tpx + F.local_max);
    trans = inside && (G.tg < tpx);
    if (F.local_only) trans = trans && ((tpx - G.tg) <= G.rd + F.Rw);
    echo = inside && (fabsf(G.tg - tpx) <= G.rd);
}

// GL3 window of pixel depth l: returns false when psi = 0 (empty window or l <= 0).
template <typename Hit>
__device__ __forceinline__ bool ray_window(const Hit& G, float l, float Rw, float& ts, float& te) {
    if (!(l > 0.0f)) return false;
    ts = Rw > 0.0f ? G.ts : 0.0f;
    te = Rw > 0.0f ? fminf(l, G.b) : l;
    return te > ts;
}

__device__ __forceinline__ float echo_weight(const RayGauss& G, float px, float py, float pz, float& dx, float& dy,
                                             float& dz, float& v0, float& v1, float& v2) {
    dx = px - G.mu[0];
    dy = py - G.mu[1];
    dz = pz - G.mu[2];
    v0 = __builtin_fmaf(G.S02, dz, __builtin_fmaf(G.S01, dy, G.S00 * dx));
    v1 = __builtin_fmaf(G.S12, dz, __builtin_fmaf(G.S11, dy, G.S01 * dx));
    v2 = __builtin_fmaf(G.S22, dz, __builtin_fmaf(G.S12, dy, G.S02 * dx));
    return __expf(-0.5f * dot3(v0, v1, v2, dx, dy, dz));
}

// Origin (with the elevational offset) and direction of scanline x of a frame. With ray_o == NULL the ray is
// the linear probe's: origin + lateral * a_lat with lateral = (x - c) * d_lat, direction a_ax, evaluated with
// the FP32 operations of NpzUltrasoundDataset._ray_origins/_ray_dirs (pipeline_god_file.py:320-333), so it
// equals those tensors bit for bit.
__device__ __forceinline__ void ray_origin(const us3d_frame& fr, int x, const float* __restrict__ ray_o,
                                           const float* __restrict__ ray_d, const float* __restrict__ ray_off,
                                           float ro[3], float rd[3]) {
    const int ray = fr.ray_base + x;
    const float delta = ray_off ? ray_off[ray] : 0.0f;
    const float lat = ((float)x - fr.col_offset) * fr.delta_lat;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float o = ray_o ? ray_o[3 * (int64_t)ray + k] : fr.origin[k] + lat * fr.axis_lat[k];
        ro[k] = __builtin_fmaf(fr.axis_ele[k], delta, o);
        rd[k] = ray_d ? ray_d[3 * (int64_t)ray + k] : fr.axis_ax[k];
    }
}

// Position of sample y of scanline x: from the (W, H, 3) pixel array, or with pixels == NULL the linear
// probe's grid point (pixel_position, the pipeline's _pixel_positions grid permuted to scanline-major).
__device__ __forceinline__ void ray_sample(const us3d_frame& fr, const float* __restrict__ pixels, int x, int y,
                                           float& px, float& py, float& pz) {
    if (pixels) {
        const float* p = pixels + 3 * ((int64_t)fr.pixel_base + (int64_t)x * fr.height + y);
        px = p[0]; py = p[1]; pz = p[2];
    } else {
        pixel_position(fr, nullptr, x, y, px, py, pz);
    }
}

// LDS image of a RayGauss (24 floats).
struct RGPack {
    float4 a, b, c, d, e, f;
};

__device__ __forceinline__ RGPack pack_rg(const RayGauss& G) {
    RGPack p;
    p.a = make_float4(G.tg, G.rd, G.tau, G.f_final);
    p.b = make_float4(G.og[0], G.og[1], G.og[2], G.ts);
    p.c = make_float4(G.dg[0], G.dg[1], G.dg[2], G.b);
    p.d = make_float4(G.mu[0], G.mu[1], G.mu[2], G.I);
    p.e = make_float4(G.S00, G.S01, G.S02, G.S11);
    p.f = make_float4(G.S12, G.S22, G.dkf, G.e_final);
    return p;
}

// Per hit: the 1-D echo quadratic along a grid scanline, d^T Sigma^-1 d = Qc + 2 c1 tau + c2 tau^2 with d(y) = P(y)
// - mu, P(y) = B + y e (B the sample-0 position, e the step), tau = y - J about the closest approach J, set up in
// FP64; returns A = d(J) and (Qc, 2 c1, c2, J).
__device__ __forceinline__ void ur_echo_line(const float mu[3], const float S[6], const float B[3], const float e[3],
                                             float A[3], float& Qc, float& c1x2, float& c2f, float& J) {
    const double b0 = (double)B[0] - mu[0], b1 = (double)B[1] - mu[1], b2 = (double)B[2] - mu[2];
    const double S00 = S[0], S01 = S[1], S02 = S[2], S11 = S[3], S12 = S[4], S22 = S[5];
    const double e0 = e[0], e1 = e[1], e2 = e[2];
    const double se0 = S00 * e0 + S01 * e1 + S02 * e2, se1 = S01 * e0 + S11 * e1 + S12 * e2,
                 se2 = S02 * e0 + S12 * e1 + S22 * e2;
    const double c2 = e0 * se0 + e1 * se1 + e2 * se2;
    const double c1 = b0 * se0 + b1 * se1 + b2 * se2;
    double j = c2 > 0.0 ? -c1 / c2 : 0.0;
    j = fmin(fmax(j, -1.0e6), 1.0e6);
    const double a0 = b0 + j * e0, a1 = b1 + j * e1, a2 = b2 + j * e2;
    const double q = a0 * (S00 * a0 + S01 * a1 + S02 * a2) + a1 * (S01 * a0 + S11 * a1 + S12 * a2) +
                     a2 * (S02 * a0 + S12 * a1 + S22 * a2);
    A[0] = (float)a0; A[1] = (float)a1; A[2] = (float)a2;
    Qc = (float)fmax(q, 0.0);
    c1x2 = (float)(2.0 * (c1 + c2 * j));
    c2f = (float)c2;
    J = (float)j;
}

// What the forward needs of one hit when the samples are the linear probe's grid (no explicit sample arrays): along
// the ray both exponents are one-dimensional quadratics. The GL3 integrand |og + dg t|^2 = a (t - tc)^2 + qmin about
// the closest approach tc (a = |dg|^2, qmin the whitened distance there), and the echo form
// d^T Sigma^-1 d = Qc + 2 c1 tau + c2 tau^2 in tau = y - J over the sample index y (ur_echo_line, the same form the
// adjoint differentiates). Each is a couple of multiply-adds per sample where the general form needs 3-vectors.
struct RayLine {
    float tg, rd, tau, f_final;
    float ts, b, tc, I;
    float a, qmin, dkf, unused;
    float Qc, c1x2, c2, J;
};

struct RLPack {
    float4 p, q, r, s;
};

__device__ __forceinline__ RLPack pack_line(const RayLine& L) {
    return RLPack{make_float4(L.tg, L.rd, L.tau, L.f_final), make_float4(L.ts, L.b, L.tc, L.I),
                  make_float4(L.a, L.qmin, L.dkf, 0.f), make_float4(L.Qc, L.c1x2, L.c2, L.J)};
}

__device__ __forceinline__ void unpack_line(const RLPack& k, RayLine& L) {
    L.tg = k.p.x; L.rd = k.p.y; L.tau = k.p.z; L.f_final = k.p.w;
    L.ts = k.q.x; L.b = k.q.y; L.tc = k.q.z; L.I = k.q.w;
    L.a = k.r.x; L.qmin = k.r.y; L.dkf = k.r.z; L.unused = 0.f;
    L.Qc = k.s.x; L.c1x2 = k.s.y; L.c2 = k.s.z; L.J = k.s.w;
}

// GL3 of the window [ts, te] in the line form: the same nodes and weights as gl3.
__device__ __forceinline__ float gl3_line(const RayLine& L, float ts, float te) {
    const float half = 0.5f * (te - ts);
    const float u0 = 0.5f * (te + ts) - L.tc;
    const float xi[3] = {-kXi, 0.0f, kXi};
    const float wq[3] = {kW0, kW1, kW0};
    float acc = 0.f;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float u = __builtin_fmaf(half, xi[i], u0);
        acc = __builtin_fmaf(wq[i], __expf(-0.5f * __builtin_fmaf(L.a * u, u, L.qmin)), acc);
    }
    return half * acc;
}

__device__ __forceinline__ void unpack_rg(const RGPack& p, RayGauss& G) {
    G.tg = p.a.x; G.rd = p.a.y; G.tau = p.a.z; G.f_final = p.a.w;
    G.og[0] = p.b.x; G.og[1] = p.b.y; G.og[2] = p.b.z; G.ts = p.b.w;
    G.dg[0] = p.c.x; G.dg[1] = p.c.y; G.dg[2] = p.c.z; G.b = p.c.w;
    G.mu[0] = p.d.x; G.mu[1] = p.d.y; G.mu[2] = p.d.z; G.I = p.d.w;
    G.S00 = p.e.x; G.S01 = p.e.y; G.S02 = p.e.z; G.S11 = p.e.w;
    G.S12 = p.f.x; G.S22 = p.f.y; G.dkf = p.f.z; G.e_final = p.f.w;
    G.tc = 0.f;
}

// ---------------------------------------------------------------- forward
// One workgroup per (frame, scanline). Threads own depth samples (kPPT per thread per chunk); the ray's sorted
// hit list is staged through LDS in batches with everything that is fixed along the ray precomputed once
// per (ray, Gaussian) -- rotation, whitened origin/direction, window, Sigma^-1, SH intensity and the
// transmittance of the complete window, which every sample beyond the window reuses (the reference
// recomputes all of it for every (sample, hit)). Per sample: T = prod over hits in front of
// tau + (1 - tau) e^-psi, echo = (1 - e^-S) sum(I w) / (sum w + 1e-6), B = T * echo (forward.cu:22-119).
// T and the echo sums are accumulated in FP64; aux keeps (T, E = num / (den + 1e-6), den) so the adjoint can
// form dL/dw = dNum (I - E) + dS without cancelling dNum I against dNum E.
//
// With the transmission window on and the chunk's sample depths nondecreasing, every staged hit gets the
// chunk-sample range [jlo, jhi) outside which its terms are fixed: before jlo it is neither in front nor in
// echo range, from jhi on it is in front with its window finished and out of echo range (factor f_final). The
// range is found by binary search with the same comparisons as the per-sample rules. A wavefront covers 64
// consecutive samples per k; hits finished for all 64 enter as one FP64 product over the wavefront's 64 hits
// (a fixed butterfly, the same on every lane), and only hits whose range meets the 64 samples are evaluated
// per lane, walked lowest-first from ballots with scalar bit scans. Otherwise every hit is evaluated for
// every sample.

constexpr int kRayThreads = 256;
constexpr int kPPT = 4;
constexpr int kRaySamples = kRayThreads * kPPT;  // samples of one forward chunk

// First index in [lo, hi) whose element satisfies pred (pred false then true along the range), else hi.
template <typename Pred>
__device__ __forceinline__ int first_true(int lo, int hi, Pred pred) {
    while (lo < hi) {
        const int m = (lo + hi) >> 1;
        if (pred(m)) hi = m; else lo = m + 1;
    }
    return lo;
}

// Product over the wavefront; the butterfly pairs the same operands on every lane, so all lanes get the same
// bits.
__device__ __forceinline__ double wave_prod(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v *= __shfl_xor(v, off, kWave);
    return v;
}

// Terms of hit G at one sample under the per-sample rules (forward.cu:76-96).
__device__ __forceinline__ void ray_accumulate(const RayGauss& G, const RayFlags& F, float px, float py, float pz,
                                               float tpx, double& T, double& num, double& den) {
    bool trans, echo;
    ray_terms(G, tpx, F, trans, echo);
    if (trans) {
        if (F.Rw > 0.0f && tpx >= G.b) {
            T *= (double)G.f_final;
        } else {
            float ts, te, psi = 0.f;
            if (ray_window(G, tpx, F.Rw, ts, te)) psi = gl3(G.og, G.dg, ts, te);
            T *= (double)__builtin_fmaf(1.0f - G.tau, __expf(-psi), G.tau);
        }
    }
    if (echo) {
        float dx, dy, dz, v0, v1, v2;
        const float w = echo_weight(G, px, py, pz, dx, dy, dz, v0, v1, v2);
        num = __builtin_fma((double)G.I, (double)w, num);
        den += (double)w;
    }
}

// ray_accumulate on the probe's grid: sample index yf, the line forms of the window integral and the echo.
__device__ __forceinline__ void ray_accumulate_line(const RayLine& G, const RayFlags& F, float yf, float tpx,
                                                    double& T, double& num, double& den) {
    bool trans, echo;
    ray_terms(G, tpx, F, trans, echo);
    if (trans) {
        if (F.Rw > 0.0f && tpx >= G.b) {
            T *= (double)G.f_final;
        } else {
            float ts, te, psi = 0.f;
            if (ray_window(G, tpx, F.Rw, ts, te)) psi = gl3_line(G, ts, te);
            T *= (double)__builtin_fmaf(1.0f - G.tau, __expf(-psi), G.tau);
        }
    }
    if (echo) {
        const float tq = yf - G.J;
        const float w = __expf(-0.5f * __builtin_fmaf(__builtin_fmaf(G.c2, tq, G.c1x2), tq, G.Qc));
        num = __builtin_fma((double)G.I, (double)w, num);
        den += (double)w;
    }
}

// Four waves per SIMD (at most 128 VGPRs) measured faster than the compiler's three, despite a small spill:
// the kernel's per-hit LDS reads and branches need the extra waves to hide their latency.
// kGrid (pixels == NULL): the samples are the linear probe's grid and the hits are staged as line forms (RayLine).
template <bool kGrid>
__global__ __launch_bounds__(kRayThreads) __attribute__((amdgpu_waves_per_eu(4, 8))) void ur_forward_kernel(
    const float4* __restrict__ rec, const int32_t* __restrict__ tile_off, const uint64_t* __restrict__ seg_key,
    const int32_t* __restrict__ pair_g, const us3d_frame* __restrict__ frames,
    const int32_t* __restrict__ frame_tile_base, int n_frames,
    const float* __restrict__ pixels, const float* __restrict__ ray_o, const float* __restrict__ ray_d,
    const float* __restrict__ ray_off, RayFlags F, const us3d_stats* __restrict__ stats, float* __restrict__ image,
    float4* __restrict__ aux) {
    __shared__ RGPack s_g[kGrid ? 1 : kRayThreads];
    __shared__ RLPack s_l[kGrid ? kRayThreads : 1];
    __shared__ float s_tp[kRaySamples];  // sample depths of the chunk
    __shared__ int2 s_rng[kRayThreads];  // per staged hit: [jlo, jhi) in chunk samples
    if (stats->overflow) return;
    const int t = blockIdx.x;
    const int f = upper_index(frame_tile_base, n_frames, t);
    const us3d_frame& fr = frames[f];
    const int x = t - fr.tile_base;
    const int W = fr.width, H = fr.height;
    const int s0 = tile_off[t], s1 = tile_off[t + 1];
    float ro[3], rd[3];
    ray_origin(fr, x, ray_o, ray_d, ray_off, ro, rd);
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    const bool window_mode = F.Rw > 0.0f && !F.local_only;
    float B0[3] = {0.f, 0.f, 0.f}, estep[3] = {0.f, 0.f, 0.f};  // grid: sample 0 and the step between samples
    if constexpr (kGrid) {
        pixel_position(fr, nullptr, x, 0, B0[0], B0[1], B0[2]);
#pragma unroll
        for (int k = 0; k < 3; ++k) estep[k] = fr.delta_ax * fr.axis_ax[k];
    }

    for (int c0 = 0; c0 < H; c0 += kRaySamples) {
        const int ns = min(kRaySamples, H - c0);
        float px[kPPT], py[kPPT], pz[kPPT], tpx[kPPT];
        double T[kPPT], num[kPPT], den[kPPT];
#pragma unroll
        for (int k = 0; k < kPPT; ++k) {
            const int y = c0 + tid + k * kRayThreads;
            px[k] = py[k] = pz[k] = tpx[k] = 0.f;
            if (y < H) {
                ray_sample(fr, pixels, x, y, px[k], py[k], pz[k]);
                tpx[k] = fmaxf(0.0f, dot3(px[k] - ro[0], py[k] - ro[1], pz[k] - ro[2], rd[0], rd[1], rd[2]));
            }
            s_tp[tid + k * kRayThreads] = tpx[k];
            T[k] = 1.0; num[k] = 0.0; den[k] = 0.0;
        }
        __syncthreads();
        bool mono = true;
#pragma unroll
        for (int q = 0; q < kPPT; ++q) {
            const int j = tid * kPPT + q;
            if (j + 1 < ns) mono = mono && (s_tp[j] <= s_tp[j + 1]);
        }
        const bool fast = window_mode && __syncthreads_and(mono) != 0;
        for (int base = s0; base < s1; base += kRayThreads) {
            __syncthreads();
            const int n = min(kRayThreads, s1 - base);
            if (tid < n) {
                const uint64_t key = seg_key[base + tid];  // depth key << 32 | the hit's visible pair
                RayGauss G;
                float R[9];
                ray_gauss(rec, pair_g[(uint32_t)key], (uint32_t)(key >> 32), ro, rd, F.k_sigma, F.Rw, G, R);
                if constexpr (kGrid) {
                    RayLine Lh;
                    Lh.tg = G.tg; Lh.rd = G.rd; Lh.tau = G.tau; Lh.f_final = G.f_final;
                    Lh.ts = G.ts; Lh.b = G.b; Lh.tc = G.tc; Lh.I = G.I; Lh.dkf = G.dkf; Lh.unused = 0.f;
                    Lh.a = dot3(G.dg[0], G.dg[1], G.dg[2], G.dg[0], G.dg[1], G.dg[2]);
                    const float m0 = __builtin_fmaf(G.dg[0], G.tc, G.og[0]);
                    const float m1 = __builtin_fmaf(G.dg[1], G.tc, G.og[1]);
                    const float m2 = __builtin_fmaf(G.dg[2], G.tc, G.og[2]);
                    Lh.qmin = dot3(m0, m1, m2, m0, m1, m2);
                    const float S6[6] = {G.S00, G.S01, G.S02, G.S11, G.S12, G.S22};
                    float A[3];
                    ur_echo_line(G.mu, S6, B0, estep, A, Lh.Qc, Lh.c1x2, Lh.c2, Lh.J);
                    s_l[tid] = pack_line(Lh);
                } else {
                    s_g[tid] = pack_rg(G);
                }
                if (fast) {
                    const int jt = first_true(0, ns, [&](int m) { return G.tg < s_tp[m]; });
                    const int jb = first_true(0, ns, [&](int m) { return s_tp[m] >= G.b; });
                    const int je0 = first_true(0, ns, [&](int m) { return (G.tg - s_tp[m]) <= G.rd; });
                    const int je1 = first_true(je0, ns, [&](int m) { return (G.tg - s_tp[m]) < -G.rd; });
                    s_rng[tid] = make_int2(min(jt, je0), max(max(jt, jb), je1));
                }
            }
            __syncthreads();
            if (fast) {
                for (int jb0 = 0; jb0 < n; jb0 += kWave) {
                    const int h = jb0 + lane;
                    const bool hv = h < n;
                    const int2 rg = hv ? s_rng[h] : make_int2(ns, ns);
                    const float ff = hv ? (kGrid ? s_l[h].p.w : s_g[h].a.w) : 1.0f;  // f_final
                    uint64_t bk[kPPT];
                    uint64_t any = 0;
#pragma unroll
                    for (int k = 0; k < kPPT; ++k) {
                        const int y0 = wave * kWave + k * kRayThreads;  // the wavefront's first sample for k
                        const bool fin = hv && y0 < ns && rg.y <= y0;
                        const bool ovl = hv && y0 < ns && rg.y > y0 && rg.x < y0 + kWave;
                        if (__ballot(fin) != 0) T[k] *= wave_prod(fin ? (double)ff : 1.0);
                        bk[k] = __ballot(ovl);
                        any |= bk[k];
                    }
                    while (any) {
                        const int bit = (int)__builtin_ctzll(any);
                        any &= any - 1;
                        if constexpr (kGrid) {
                            RayLine G;
                            unpack_line(s_l[jb0 + bit], G);
#pragma unroll
                            for (int k = 0; k < kPPT; ++k) {
                                if (!((bk[k] >> bit) & 1ull)) continue;  // scalar test
                                ray_accumulate_line(G, F, (float)(c0 + tid + k * kRayThreads), tpx[k], T[k], num[k],
                                                    den[k]);
                            }
                        } else {
                            RayGauss G;
                            unpack_rg(s_g[jb0 + bit], G);
#pragma unroll
                            for (int k = 0; k < kPPT; ++k) {
                                if (!((bk[k] >> bit) & 1ull)) continue;  // scalar test
                                ray_accumulate(G, F, px[k], py[k], pz[k], tpx[k], T[k], num[k], den[k]);
                            }
                        }
                    }
                }
            } else {
                for (int j = 0; j < n; ++j) {
                    if constexpr (kGrid) {
                        RayLine G;
                        unpack_line(s_l[j], G);
#pragma unroll
                        for (int k = 0; k < kPPT; ++k)
                            ray_accumulate_line(G, F, (float)(c0 + tid + k * kRayThreads), tpx[k], T[k], num[k],
                                                den[k]);
                    } else {
                        RayGauss G;
                        unpack_rg(s_g[j], G);
#pragma unroll
                        for (int k = 0; k < kPPT; ++k)
                            ray_accumulate(G, F, px[k], py[k], pz[k], tpx[k], T[k], num[k], den[k]);
                    }
                }
            }
        }
#pragma unroll
        for (int k = 0; k < kPPT; ++k) {
            const int y = c0 + tid + k * kRayThreads;
            if (y >= H) continue;
            const int64_t o = (int64_t)fr.pixel_base + (int64_t)y * W + x;
            double B = 0.0, E = 0.0;
            if (s1 > s0) {
                const double gamma = 1.0 - exp(-den[k]);
                E = num[k] / (den[k] + 1e-6);
                B = T[k] * (gamma * E);
            }
            image[o] = (float)B;
            aux[(int64_t)fr.pixel_base + (int64_t)x * H + y] = make_float4((float)T[k], (float)E, (float)den[k], 0.f);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- adjoint
// One workgroup per (frame, scanline); each thread owns one hit (Gaussian) of the ray and walks the ray's
// samples (per-sample adjoint coefficients staged in LDS, read as broadcasts), accumulating in registers:
//   transmission: dL/dT_i = (dL/dT T) / T_i when T_i > 1e-6 (backward.cu:273-274); d raw tau and d psi;
//     the GL3 VJP holds the window endpoints fixed (backward.cu:57-72); samples past the window share one
//     psi, so their d psi is summed and pushed through the quadrature nodes once;
//   echo: dL/dw = dNum I + dDen + dS; d mu, d Sigma^-1 accumulated as sums over samples;
// then mapped once per (ray, Gaussian) to (mu, s, q, raw tau, sh) and written at the instance's emit index.
// Samples with |dL/dB| < grad_threshold contribute nothing (backward.cu:170).
// With the transmission window on and sample depths nondecreasing along the ray (checked per chunk), the
// samples where a hit is in front and its window finished form a suffix of the chunk; their terms are
// cT / f_final times constants, so they enter through one FP64 suffix sum of cT, and each thread walks only
// the samples of its echo range and open window (located by binary search with the same comparisons as the
// per-sample rules). Otherwise every thread walks the whole chunk.
// kGrid (pixels == NULL): the scanline's samples are the linear probe's grid points P(y) = P(0) + y e, e =
// delta_ax a_ax, so along the ray the echo quadratic form is one-dimensional: with d(y) = P(y) - mu,
// d^T Sigma^-1 d = Qc + 2 c1 tau + c2 tau^2 in tau = y - J about the closest approach J (set up in FP64 per hit;
// c1 is then the FP64 residual, ~0), two multiply-adds per sample instead of a 3-vector difference and a 3x3
// product, and the echo sums sum(dQ d) and sum(dQ d d^T) follow from three moments sum(dQ tau^k), k = 0..2,
// with d = A + tau e, A = d(J) (ur_echo_moments_to_sums). The decisions (which terms a sample enters) are those
// of the general form: they use the sample depths only.

constexpr int kRayChunk = 1024;
constexpr int kRayChunkPerThread = kRayChunk / kRayThreads;
// Grid, transmission window on, monotone depths, one staged chunk ("itemized"): a hit's open window [jt, jf) and its
// echo range [je0, je1) are cut into segments of kWinSeg and kEchoSeg samples, and each wavefront walks its 64
// hits' segments 64 at a time, one segment per lane, so the lanes stay busy however unequal the hits' ranges are
// (a thread walking its own hit's ranges kept 20 % of the lanes busy on the open windows of the config-3 scene,
// 58 % on the echo ranges). The owning lane's parameters reach a segment's lane by cross-lane reads, and a hit's
// per-segment sums return to it through a fixed-order segmented lane reduction: no LDS, deterministic.
constexpr int kWinSeg = 8;
constexpr int kEchoSeg = 32;

// Wave-local segment numbering: lane l owns nseg segments; incl is the inclusive scan of nseg over the wave.
__device__ __forceinline__ int wave_incl_scan(int nseg) {
    const int lane = threadIdx.x & (kWave - 1);
    int incl = nseg;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const int v = __shfl_up(incl, off, kWave);
        if (lane >= off) incl += v;
    }
    return incl;
}

// Lane owning segment k: the first lane whose inclusive count exceeds k (k < the wave's total).
__device__ __forceinline__ int wave_seg_owner(int incl, int k) {
    int owner = 0;
#pragma unroll
    for (int step = kWave / 2; step >= 1; step >>= 1)
        if (__shfl(incl, owner + step - 1, kWave) <= k) owner += step;
    return owner;
}

// Sums of NS values over runs of adjacent lanes with the same id; afterwards the first lane of each run holds the
// run's total (runs of at most 64 lanes; the tree order is fixed by the lane positions).
template <int NS>
__device__ __forceinline__ void wave_run_sums(int id, float (&a)[NS]) {
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        // every lane takes part in the cross-lane reads (a read from a lane that skipped one is undefined)
        const int oid = __shfl_down(id, off, kWave);
        const bool same = (lane + off < kWave) && oid == id;
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const float v = __shfl_down(a[k], off, kWave);
            a[k] = same ? a[k] + v : a[k];
        }
    }
}

__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = min(v, __shfl_xor(v, off, kWave));
    return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = max(v, __shfl_xor(v, off, kWave));
    return v;
}

// Three waves per SIMD (at most 168 VGPRs): the echo-item phase would otherwise take the grid variant to 175 and two.
template <bool kGrid>
__global__ __launch_bounds__(kRayThreads) __attribute__((amdgpu_waves_per_eu(3, 8))) void ur_backward_kernel(
    const float4* __restrict__ rec, const int32_t* __restrict__ tile_off, const int32_t* __restrict__ pair_g,
    const uint64_t* __restrict__ seg_key, const int32_t* __restrict__ pair_off, const int4* __restrict__ pair_rect,
    const us3d_frame* __restrict__ frames, const int32_t* __restrict__ frame_tile_base, int n_frames,
    const float* __restrict__ pixels, const float* __restrict__ ray_o, const float* __restrict__ ray_d,
    const float* __restrict__ ray_off, RayFlags F,
    const float4* __restrict__ aux, const float* __restrict__ d_image, const us3d_stats* __restrict__ stats,
    float4* __restrict__ partial) {
    // the general form needs the samples' positions; on the probe's grid only their depths are read
    __shared__ float4 s_p[kGrid ? 1 : kRayChunk];   // px, py, pz, tpx
    __shared__ float s_tp[kGrid ? kRayChunk : 1];   // tpx
    __shared__ float4 s_c[kRayChunk];  // cT, dNum, dS, E
    __shared__ double s_S[kRayChunk + 1];  // s_S[j] = sum of cT over samples j.. of the chunk
    __shared__ double s_wsum[kRayThreads / kWave];
    auto tp_at = [&](int j) -> float {
        if constexpr (kGrid) return s_tp[j]; else return s_p[j].w;
    };
    auto p_at = [&](int j) -> float4 {
        if constexpr (kGrid) return make_float4(0.f, 0.f, 0.f, s_tp[j]); else return s_p[j];
    };
    if (stats->overflow) return;
    const bool window_mode = F.Rw > 0.0f && !F.local_only;
    const int t = blockIdx.x;
    const int s0 = tile_off[t], s1 = tile_off[t + 1];
    if (s0 == s1) return;
    const int f = upper_index(frame_tile_base, n_frames, t);
    const us3d_frame& fr = frames[f];
    const int x = t - fr.tile_base;
    const int W = fr.width, H = fr.height;
    float ro[3], rd[3];
    ray_origin(fr, x, ray_o, ray_d, ray_off, ro, rd);
    const int tid = threadIdx.x;
    float B0[3] = {0.f, 0.f, 0.f}, estep[3] = {0.f, 0.f, 0.f};  // grid: sample 0 and the step between samples
    if constexpr (kGrid) {
        pixel_position(fr, nullptr, x, 0, B0[0], B0[1], B0[2]);
#pragma unroll
        for (int k = 0; k < 3; ++k) estep[k] = fr.delta_ax * fr.axis_ax[k];
    }

    // Stage samples [c0, c0 + n) of the ray: positions and depths, the per-sample adjoint coefficients, and the
    // suffix sums of cT. Returns whether the sample depths are nondecreasing over the chunk.
    auto stage = [&](int c0, int n) -> bool {
        __syncthreads();
        for (int j = tid; j < n; j += kRayThreads) {
            const int y = c0 + j;
            const int64_t r = (int64_t)fr.pixel_base + (int64_t)x * H + y;  // ray-major sample index
            float p[3];
            ray_sample(fr, pixels, x, y, p[0], p[1], p[2]);
            const float tp = fmaxf(0.0f, dot3(p[0] - ro[0], p[1] - ro[1], p[2] - ro[2], rd[0], rd[1], rd[2]));
            const float gB = d_image[(int64_t)fr.pixel_base + (int64_t)y * W + x];
            float cT = 0.f, dNum = 0.f, dS = 0.f, E = 0.f;
            if (!(fabsf(gB) < F.grad_threshold)) {
                const float4 a = aux[r];  // T, E = num / (den + 1e-6), den
                const float ex = __expf(-a.z);
                const float gamma = 1.0f - ex;
                E = a.y;
                const float dE = gB * a.x;           // dL/dE_val
                const float dEraw = dE * gamma;
                dS = dE * E * ex;                      // through gamma = 1 - e^-S
                dNum = dEraw / (a.z + 1e-6f);          // dL/dw = dNum (I - E) + dS (= dNum I + dDen + dS)
                cT = gB * (gamma * E) * a.x;           // dL/dT * T
            }
            if constexpr (kGrid) s_tp[j] = tp; else s_p[j] = make_float4(p[0], p[1], p[2], tp);
            s_c[j] = make_float4(cT, dNum, dS, E);
        }
        __syncthreads();
        // monotone sample depths, and the suffix sums of cT (thread t owns samples 4t .. 4t+3)
        bool mono = true;
        double loc[kRayChunkPerThread];
        double tot = 0.0;
#pragma unroll
        for (int q = kRayChunkPerThread - 1; q >= 0; --q) {
            const int j = tid * kRayChunkPerThread + q;
            if (j < n) {
                tot += (double)s_c[j].x;
                if (j + 1 < n) mono = mono && (tp_at(j) <= tp_at(j + 1));
            }
            loc[q] = tot;
        }
        double suf = tot;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            const double o = __shfl_down(suf, off, kWave);
            if ((tid & (kWave - 1)) + off < kWave) suf += o;
        }
        if ((tid & (kWave - 1)) == 0) s_wsum[tid / kWave] = suf;
        mono = __syncthreads_and(mono) != 0;
        double after = suf - tot;
        for (int w = tid / kWave + 1; w < kRayThreads / kWave; ++w) after += s_wsum[w];
#pragma unroll
        for (int q = 0; q < kRayChunkPerThread; ++q) {
            const int j = tid * kRayChunkPerThread + q;
            if (j < n) s_S[j] = loc[q] + after;
        }
        if (tid == 0) s_S[n] = 0.0;
        __syncthreads();
        return mono;
    };
    // A ray of at most one chunk is staged once for all its hits; longer rays restage every chunk per batch.
    const bool one_chunk = H <= kRayChunk;
    bool mono = one_chunk ? stage(0, H) : true;

    for (int base = s0; base < s1; base += kRayThreads) {
        const int i = base + tid;
        const bool active = i < s1;
        RayGauss G;
        float R[9];
        const uint64_t key = seg_key[active ? i : s0];  // depth key << 32 | the hit's visible pair
        const int hv = (int)(uint32_t)key;
        ray_gauss(rec, pair_g[hv], (uint32_t)(key >> 32), ro, rd, F.k_sigma, F.Rw, G, R);
        float dog[3] = {0.f, 0.f, 0.f}, ddg[3] = {0.f, 0.f, 0.f};
        float fin = 0.f, dtau = 0.f, dsh = 0.f;
        float u[3] = {0.f, 0.f, 0.f};
        float g00 = 0.f, g01 = 0.f, g02 = 0.f, g11 = 0.f, g12 = 0.f, g22 = 0.f;
        float eQc = 0.f, ec1 = 0.f, ec2 = 0.f, eJ = 0.f;  // grid: the echo line (A is formed after the loop)
        float M0 = 0.f, M1 = 0.f, M2 = 0.f;             // grid: sum(dQ tau^k)
        if constexpr (kGrid) {
            const float mu0[3] = {G.mu[0], G.mu[1], G.mu[2]};
            const float S0[6] = {G.S00, G.S01, G.S02, G.S11, G.S12, G.S22};
            float A0[3];
            ur_echo_line(mu0, S0, B0, estep, A0, eQc, ec1, ec2, eJ);
        }

        // echo terms as items (see kEchoSeg): workgroup-uniform
        const bool itemized = kGrid && one_chunk && window_mode && mono;
        int jt0 = 0, jf0 = 0, je0 = 0, je1 = 0;  // this hit's open window and echo range (itemized)
        for (int c0 = 0; c0 < H; c0 += kRayChunk) {
            const int n = min(kRayChunk, H - c0);
            if (!one_chunk) mono = stage(c0, n);
            const bool fast = window_mode && mono;
            // this thread's sample range [lo, hi) and the first finished sample jf (suffix handled by s_S)
            int lo = n, hi = 0, jf = n;
            if (active) {
                if (fast) {
                    const int jt = first_true(0, n, [&](int m) { return G.tg < tp_at(m); });
                    const int jb = first_true(0, n, [&](int m) { return tp_at(m) >= G.b; });
                    jf = max(jt, jb);
                    jt0 = jt;
                    jf0 = jf;
                    je0 = first_true(0, n, [&](int m) { return (G.tg - tp_at(m)) <= G.rd; });
                    je1 = first_true(je0, n, [&](int m) { return (G.tg - tp_at(m)) < -G.rd; });
                    lo = itemized ? n : min(jt, je0);  // itemized: both ranges run as segments below
                    hi = itemized ? 0 : max(jf, je1);
                    const double S = s_S[jf];
                    if (S != 0.0 && G.f_final > 1e-6f) {
                        const float dTs = (float)(S / (double)G.f_final);
                        dtau = __builtin_fmaf(dTs, 1.0f - G.e_final, dtau);
                        fin -= dTs * (1.0f - G.tau) * G.e_final;
                    }
                } else {
                    lo = 0;
                    hi = n;
                }
            }
            const int wlo = wave_min_i(lo), whi = wave_max_i(hi);
            // the next sample's position and coefficients are read before this sample's work, so the LDS
            // latency overlaps it
            float4 pn = p_at(min(wlo, n - 1)), cn = s_c[min(wlo, n - 1)];
            for (int j = wlo; j < whi; ++j) {
                const float4 pp = pn, cc = cn;
                const int jn = (j + 1 < whi) ? j + 1 : j;
                pn = p_at(jn);
                cn = s_c[jn];
                if (j < lo || j >= hi) continue;
                bool trans, echo;
                ray_terms(G, pp.w, F, trans, echo);
                if (j >= jf) trans = false;  // in the finished suffix: counted through s_S
                if (!trans && !echo) continue;
                if (trans) {
                    const bool finished = F.Rw > 0.0f && pp.w >= G.b;
                    if (finished) {
                        const float ep = G.e_final, Ti = G.f_final;
                        const float dTi = Ti > 1e-6f ? cc.x * __builtin_amdgcn_rcpf(Ti) : 0.f;
                        dtau = __builtin_fmaf(dTi, 1.0f - ep, dtau);
                        fin += -dTi * (1.0f - G.tau) * ep;
                    } else {
                        float ts = 0.f, te = 0.f;
                        if (ray_window(G, pp.w, F.Rw, ts, te)) {
                            // the GL3 nodes and their exponentials serve both psi (as gl3 evaluates it) and
                            // its VJP, with the window endpoints held fixed (backward.cu:57-72)
                            const float half = 0.5f * (te - ts), c = 0.5f * (te + ts);
                            const float xi[3] = {-kXi, 0.0f, kXi};
                            const float wq[3] = {kW0, kW1, kW0};
                            float tq[3], x0[3], x1[3], x2[3], eq[3];
                            float acc = 0.f;
#pragma unroll
                            for (int q = 0; q < 3; ++q) {
                                tq[q] = __builtin_fmaf(half, xi[q], c);
                                x0[q] = __builtin_fmaf(G.dg[0], tq[q], G.og[0]);
                                x1[q] = __builtin_fmaf(G.dg[1], tq[q], G.og[1]);
                                x2[q] = __builtin_fmaf(G.dg[2], tq[q], G.og[2]);
                                eq[q] = __expf(-0.5f * dot3(x0[q], x1[q], x2[q], x0[q], x1[q], x2[q]));
                                acc = __builtin_fmaf(wq[q], eq[q], acc);
                            }
                            const float ep = __expf(-(half * acc));
                            const float Ti = __builtin_fmaf(1.0f - G.tau, ep, G.tau);
                            const float dTi = Ti > 1e-6f ? cc.x * __builtin_amdgcn_rcpf(Ti) : 0.f;
                            dtau = __builtin_fmaf(dTi, 1.0f - ep, dtau);
                            const float coeff = -dTi * (1.0f - G.tau) * ep * half;  // d psi * half
#pragma unroll
                            for (int q = 0; q < 3; ++q) {
                                const float a = -coeff * wq[q] * eq[q];
                                dog[0] = __builtin_fmaf(a, x0[q], dog[0]);
                                dog[1] = __builtin_fmaf(a, x1[q], dog[1]);
                                dog[2] = __builtin_fmaf(a, x2[q], dog[2]);
                                const float at = a * tq[q];
                                ddg[0] = __builtin_fmaf(at, x0[q], ddg[0]);
                                ddg[1] = __builtin_fmaf(at, x1[q], ddg[1]);
                                ddg[2] = __builtin_fmaf(at, x2[q], ddg[2]);
                            }
                        }
                        // an empty window (depth <= 0 or before the window starts) has psi = 0, T_i = 1 and
                        // contributes nothing to d raw tau or the window VJP
                    }
                }
                if (echo) {
                    if constexpr (kGrid) {
                        const float tq = (float)(c0 + j) - eJ;
                        const float w = __expf(-0.5f * __builtin_fmaf(__builtin_fmaf(ec2, tq, ec1), tq, eQc));
                        dsh = __builtin_fmaf(cc.y, w, dsh);
                        const float dQ = __builtin_fmaf(cc.y, G.I - cc.w, cc.z) * w;  // -2 dL/dQ
                        M0 += dQ;
                        M1 = __builtin_fmaf(dQ, tq, M1);
                        M2 = __builtin_fmaf(dQ * tq, tq, M2);
                    } else {
                        float dx, dy, dz, v0, v1, v2;
                        const float w = echo_weight(G, pp.x, pp.y, pp.z, dx, dy, dz, v0, v1, v2);
                        dsh = __builtin_fmaf(cc.y, w, dsh);
                        const float dw = __builtin_fmaf(cc.y, G.I - cc.w, cc.z);
                        const float dQ = -0.5f * dw * w;
                        u[0] = __builtin_fmaf(dQ, dx, u[0]);
                        u[1] = __builtin_fmaf(dQ, dy, u[1]);
                        u[2] = __builtin_fmaf(dQ, dz, u[2]);
                        const float qx = dQ * dx, qy = dQ * dy;
                        g00 = __builtin_fmaf(qx, dx, g00);
                        g01 = __builtin_fmaf(qx, dy, g01);
                        g02 = __builtin_fmaf(qx, dz, g02);
                        g11 = __builtin_fmaf(qy, dy, g11);
                        g12 = __builtin_fmaf(qy, dz, g12);
                        g22 = __builtin_fmaf(dQ * dz, dz, g22);
                    }
                }
            }
        }
        if (itemized) {
            const int lane = tid & (kWave - 1);
            // open windows: per sample the GL3 transmission of the window [ts, l] and its VJP with the endpoints
            // held fixed (backward.cu:57-72), as in the per-lane loop above; samples in [jt, jf) are in front of
            // the hit and before its window end, so te = l
            {
                const int ns = (active && jf0 > jt0) ? (jf0 - jt0 + kWinSeg - 1) / kWinSeg : 0;
                const int incl = wave_incl_scan(ns);
                const int excl = incl - ns;
                const int total = __shfl(incl, kWave - 1, kWave);
                for (int base = 0; base < total; base += kWave) {
                    const int k = base + lane;
                    const int o = wave_seg_owner(incl, min(k, total - 1));
                    const float og0 = __shfl(G.og[0], o, kWave), og1 = __shfl(G.og[1], o, kWave),
                                og2 = __shfl(G.og[2], o, kWave);
                    const float dg0 = __shfl(G.dg[0], o, kWave), dg1 = __shfl(G.dg[1], o, kWave),
                                dg2 = __shfl(G.dg[2], o, kWave);
                    const float ts = __shfl(G.ts, o, kWave), otau = __shfl(G.tau, o, kWave);
                    const int oa = __shfl(jt0, o, kWave), ob = __shfl(jf0, o, kWave), oex = __shfl(excl, o, kWave);
                    float acc7[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};  // dog, ddg, dtau
                    if (k < total) {
                        const int j0 = oa + (k - oex) * kWinSeg, j1 = min(j0 + kWinSeg, ob);
                        for (int j = j0; j < j1; ++j) {
                            const float l = tp_at(j), cT = s_c[j].x;
                            if (!(l > 0.0f && l > ts)) continue;  // empty window: psi = 0, T_i = 1
                            const float half = 0.5f * (l - ts), c = 0.5f * (l + ts);
                            const float xi[3] = {-kXi, 0.0f, kXi};
                            const float wq[3] = {kW0, kW1, kW0};
                            float tq[3], x0[3], x1[3], x2[3], eq[3];
                            float acc = 0.f;
#pragma unroll
                            for (int q = 0; q < 3; ++q) {
                                tq[q] = __builtin_fmaf(half, xi[q], c);
                                x0[q] = __builtin_fmaf(dg0, tq[q], og0);
                                x1[q] = __builtin_fmaf(dg1, tq[q], og1);
                                x2[q] = __builtin_fmaf(dg2, tq[q], og2);
                                eq[q] = __expf(-0.5f * dot3(x0[q], x1[q], x2[q], x0[q], x1[q], x2[q]));
                                acc = __builtin_fmaf(wq[q], eq[q], acc);
                            }
                            const float ep = __expf(-(half * acc));
                            const float Ti = __builtin_fmaf(1.0f - otau, ep, otau);
                            const float dTi = Ti > 1e-6f ? cT * __builtin_amdgcn_rcpf(Ti) : 0.f;
                            acc7[6] = __builtin_fmaf(dTi, 1.0f - ep, acc7[6]);
                            const float coeff = -dTi * (1.0f - otau) * ep * half;  // d psi * half
#pragma unroll
                            for (int q = 0; q < 3; ++q) {
                                const float aq = -coeff * wq[q] * eq[q];
                                acc7[0] = __builtin_fmaf(aq, x0[q], acc7[0]);
                                acc7[1] = __builtin_fmaf(aq, x1[q], acc7[1]);
                                acc7[2] = __builtin_fmaf(aq, x2[q], acc7[2]);
                                const float at = aq * tq[q];
                                acc7[3] = __builtin_fmaf(at, x0[q], acc7[3]);
                                acc7[4] = __builtin_fmaf(at, x1[q], acc7[4]);
                                acc7[5] = __builtin_fmaf(at, x2[q], acc7[5]);
                            }
                        }
                    }
                    wave_run_sums<7>(k < total ? o : -1 - lane, acc7);
                    // this lane's hit: its segments in this pass start at lane max(excl, base) - base
                    const bool mine = ns > 0 && excl < base + kWave && excl + ns > base;
                    const int head = max(excl - base, 0) & (kWave - 1);
                    float got[7];
#pragma unroll
                    for (int m = 0; m < 7; ++m) got[m] = __shfl(acc7[m], head, kWave);
                    if (mine) {
                        dog[0] += got[0]; dog[1] += got[1]; dog[2] += got[2];
                        ddg[0] += got[3]; ddg[1] += got[4]; ddg[2] += got[5];
                        dtau += got[6];
                    }
                }
            }
            // echo ranges: the one-dimensional echo quadratic along the grid scanline (ur_echo_line)
            {
                const int ns = (active && je1 > je0) ? (je1 - je0 + kEchoSeg - 1) / kEchoSeg : 0;
                const int incl = wave_incl_scan(ns);
                const int excl = incl - ns;
                const int total = __shfl(incl, kWave - 1, kWave);
                for (int base = 0; base < total; base += kWave) {
                    const int k = base + lane;
                    const int o = wave_seg_owner(incl, min(k, total - 1));
                    const float qc = __shfl(eQc, o, kWave), q1 = __shfl(ec1, o, kWave), q2 = __shfl(ec2, o, kWave),
                                qj = __shfl(eJ, o, kWave), qI = __shfl(G.I, o, kWave);
                    const int oa = __shfl(je0, o, kWave), ob = __shfl(je1, o, kWave), oex = __shfl(excl, o, kWave);
                    float acc4[4] = {0.f, 0.f, 0.f, 0.f};  // dsh, sum dQ, sum dQ tau, sum dQ tau^2
                    if (k < total) {
                        const int j0 = oa + (k - oex) * kEchoSeg, j1 = min(j0 + kEchoSeg, ob);
                        for (int j = j0; j < j1; ++j) {
                            const float4 cc = s_c[j];
                            const float tq = (float)j - qj;
                            const float w = __expf(-0.5f * __builtin_fmaf(__builtin_fmaf(q2, tq, q1), tq, qc));
                            acc4[0] = __builtin_fmaf(cc.y, w, acc4[0]);
                            const float dQ = __builtin_fmaf(cc.y, qI - cc.w, cc.z) * w;  // -2 dL/dQ
                            acc4[1] += dQ;
                            acc4[2] = __builtin_fmaf(dQ, tq, acc4[2]);
                            acc4[3] = __builtin_fmaf(dQ * tq, tq, acc4[3]);
                        }
                    }
                    wave_run_sums<4>(k < total ? o : -1 - lane, acc4);
                    const bool mine = ns > 0 && excl < base + kWave && excl + ns > base;
                    const int head = max(excl - base, 0) & (kWave - 1);
                    float got[4];
#pragma unroll
                    for (int m = 0; m < 4; ++m) got[m] = __shfl(acc4[m], head, kWave);
                    if (mine) {
                        dsh += got[0];
                        M0 += got[1];
                        M1 += got[2];
                        M2 += got[3];
                    }
                }
            }
        }
        if (!active) continue;
        // the record's rotation, Sigma^-1 and mean, formed again here rather than kept in registers through the
        // sample loop (the same operations as ray_gauss, so the same bits)
        const int gi = pair_g[hv];
        const float4 r0 = rec[4 * (int64_t)gi + 0];
        const float4 r1 = rec[4 * (int64_t)gi + 1];
        const float4 r2 = rec[4 * (int64_t)gi + 2];
        const float s[3] = {r1.x, r1.y, r1.z};
        const float mu[3] = {r0.x, r0.y, r0.z};
        float Rm[9], Sm[6];
        quat_rot(r2, Rm);
        sigma_inv6(Rm, s, Sm);
        if constexpr (kGrid) {
            float eA[3], qc_, c1_, c2_, j_;
            ur_echo_line(mu, Sm, B0, estep, eA, qc_, c1_, c2_, j_);
            // sum(dQ d) and sum(dQ d d^T) with d = A + tau e and the factor -1/2 of dL/dQ
            const float m0 = -0.5f * M0, m1 = -0.5f * M1, m2 = -0.5f * M2;
#pragma unroll
            for (int k = 0; k < 3; ++k) u[k] = __builtin_fmaf(m1, estep[k], m0 * eA[k]);
            auto gab = [&](int a, int b) {
                return __builtin_fmaf(m2 * estep[a], estep[b],
                                      __builtin_fmaf(m1, __builtin_fmaf(eA[a], estep[b], estep[a] * eA[b]),
                                                     m0 * eA[a] * eA[b]));
            };
            g00 = gab(0, 0); g01 = gab(0, 1); g02 = gab(0, 2);
            g11 = gab(1, 1); g12 = gab(1, 2); g22 = gab(2, 2);
        }
        // samples past the window: one VJP with the summed d psi
        if (fin != 0.0f && G.b > G.ts && G.b > 0.0f) {
            const float ts = G.ts, te = G.b;
            const float half = 0.5f * (te - ts), c = 0.5f * (te + ts);
            const float coeff = fin * half;
            const float xi[3] = {-kXi, 0.0f, kXi};
            const float wq[3] = {kW0, kW1, kW0};
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const float tq = __builtin_fmaf(half, xi[q], c);
                const float x0 = __builtin_fmaf(G.dg[0], tq, G.og[0]);
                const float x1 = __builtin_fmaf(G.dg[1], tq, G.og[1]);
                const float x2 = __builtin_fmaf(G.dg[2], tq, G.og[2]);
                const float a = -coeff * wq[q] * __expf(-0.5f * dot3(x0, x1, x2, x0, x1, x2));
                dog[0] += a * x0; dog[1] += a * x1; dog[2] += a * x2;
                ddg[0] += a * tq * x0; ddg[1] += a * tq * x1; ddg[2] += a * tq * x2;
            }
        }
        const float diff[3] = {ro[0] - mu[0], ro[1] - mu[1], ro[2] - mu[2]};
        float dmu[3], ds[3], dR[9];
        float dol[3], ddl[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            dol[c] = dog[c] / s[c];
            ddl[c] = ddg[c] / s[c];
            ds[c] = dog[c] * (-G.og[c] / s[c]) + ddg[c] * (-G.dg[c] / s[c]);
        }
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            dmu[r] = -(Rm[3 * r] * dol[0] + Rm[3 * r + 1] * dol[1] + Rm[3 * r + 2] * dol[2]);
#pragma unroll
            for (int c = 0; c < 3; ++c) dR[3 * r + c] = dol[c] * diff[r] + ddl[c] * rd[r];
        }
        // echo: d mu = -2 Sigma^-1 u; Sigma^-1 = R D R^T with D = diag(1/s^2): dR += 2 Gm R D (Gm symmetric),
        // ds_k += (R^T Gm R)_kk * (-2 / s_k^3)
        dmu[0] += -2.0f * (Sm[0] * u[0] + Sm[1] * u[1] + Sm[2] * u[2]);
        dmu[1] += -2.0f * (Sm[1] * u[0] + Sm[3] * u[1] + Sm[4] * u[2]);
        dmu[2] += -2.0f * (Sm[2] * u[0] + Sm[4] * u[1] + Sm[5] * u[2]);
        const float Gm[9] = {g00, g01, g02, g01, g11, g12, g02, g12, g22};
        const float D[3] = {1.0f / (s[0] * s[0]), 1.0f / (s[1] * s[1]), 1.0f / (s[2] * s[2])};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            float gr[3];  // Gm * column k of R
#pragma unroll
            for (int a = 0; a < 3; ++a) gr[a] = Gm[3 * a] * Rm[k] + Gm[3 * a + 1] * Rm[3 + k] + Gm[3 * a + 2] * Rm[6 + k];
#pragma unroll
            for (int a = 0; a < 3; ++a) dR[3 * a + k] += 2.0f * gr[a] * D[k];
            const float rgr = Rm[k] * gr[0] + Rm[3 + k] * gr[1] + Rm[6 + k] * gr[2];
            ds[k] += rgr * (-2.0f / (s[k] * s[k] * s[k]));
        }
        const float4 dq = quat_vjp(r2, dR);
        const float dt = dtau * G.tau * (1.0f - G.tau);
        const int64_t e = emit_index(pair_off, pair_rect, hv, 0, x);  // scanline x of the pair's range
        partial[4 * e + 0] = make_float4(dmu[0], dmu[1], dmu[2], ds[0]);
        partial[4 * e + 1] = make_float4(ds[1], ds[2], dt, 0.f);
        partial[4 * e + 2] = dq;
        partial[4 * e + 3] = make_float4(dsh, dsh * rd[0], dsh * rd[1], dsh * rd[2]);
    }
}

__global__ __launch_bounds__(256) void ur_reduce_kernel(int n_gauss, const us3d_sweep* __restrict__ sweeps,
                                                        int n_sweeps, const uint64_t* __restrict__ pair_bits,
                                                        const int32_t* __restrict__ word_off,
                                                        const int32_t* __restrict__ pair_off,
                                                        const float4* __restrict__ partial,
                                                        const float4* __restrict__ rec, int log_scales,
                                                        const us3d_stats* __restrict__ stats, float* __restrict__ d_means,
                                                        float* __restrict__ d_scales, float* __restrict__ d_quats,
                                                        float* __restrict__ d_atten, float* __restrict__ d_sh) {
    const int g = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    if (g >= n_gauss || stats->overflow) return;
    const us3d_sweep sw = sweep_for_reduce(sweeps, n_sweeps, g);
    const int64_t p0 = sw.pair_base + (int64_t)(g - sw.g_begin) * sw.f_count;
    // the Gaussian's visible pairs are v in [rank(p0), rank(p0 + f_count)); their instances are consecutive
    const int e0 = pair_off[pair_rank(pair_bits, word_off, p0)];
    const int e1 = pair_off[pair_rank(pair_bits, word_off, p0 + sw.f_count)];
    float acc[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) acc[k] = 0.f;
    for (int e = e0 + lane; e < e1; e += kWave) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const float4 a = partial[4 * (int64_t)e + v];
            acc[4 * v] += a.x; acc[4 * v + 1] += a.y; acc[4 * v + 2] += a.z; acc[4 * v + 3] += a.w;
        }
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) acc[k] = wave_sum(acc[k]);
    if (lane == 0) {
        const int64_t G = g;
        d_means[3 * G] = acc[0]; d_means[3 * G + 1] = acc[1]; d_means[3 * G + 2] = acc[2];
        if (log_scales) {  // through max(exp(x), 1e-5): d x = d s * exp(x) where the clamp passes, else 0
            const float4 r1 = rec[4 * G + 1];
            const int pass = (int)r1.w;
            d_scales[3 * G] = (pass & 1) ? acc[3] * r1.x : 0.f;
            d_scales[3 * G + 1] = (pass & 2) ? acc[4] * r1.y : 0.f;
            d_scales[3 * G + 2] = (pass & 4) ? acc[5] * r1.z : 0.f;
        } else {
            d_scales[3 * G] = acc[3]; d_scales[3 * G + 1] = acc[4]; d_scales[3 * G + 2] = acc[5];
        }
        d_atten[G] = acc[6];
        d_quats[4 * G] = acc[8]; d_quats[4 * G + 1] = acc[9]; d_quats[4 * G + 2] = acc[10]; d_quats[4 * G + 3] = acc[11];
        d_sh[4 * G] = acc[12]; d_sh[4 * G + 1] = acc[13]; d_sh[4 * G + 2] = acc[14]; d_sh[4 * G + 3] = acc[15];
    }
}

int check_ur_settings(const us3d_ur_settings* s) {
    if (!s) return fail(US3D_ERR_ARGUMENT, "settings is NULL");
    if (!(s->k_sigma > 0.0f)) return fail(US3D_ERR_ARGUMENT, "k_sigma must be > 0");
    if (!(s->quadrature_window_radius_mm >= 0.0f)) return fail(US3D_ERR_ARGUMENT, "quadrature_window_radius_mm must be >= 0");
    if (!(s->local_transmission_max_window_mm > 0.0f)) return fail(US3D_ERR_ARGUMENT, "local_transmission_max_window_mm must be > 0");
    return US3D_OK;
}

RayFlags flags_of(const us3d_ur_settings* s) {
    RayFlags F;
    F.k_sigma = s->k_sigma;
    F.Rw = s->quadrature_window_radius_mm;
    F.local_max = s->local_transmission_max_window_mm;
    F.grad_threshold = s->grad_threshold;
    F.local_only = s->local_transmission_window_only;
    return F;
}

}  // namespace

}  // namespace us3d

using namespace us3d;

extern "C" {

int64_t us3d_ur_workspace_bytes(int64_t n_gauss, int64_t n_pairs, int64_t n_rays, int64_t k_capacity,
                                int64_t v_capacity) {
    return ur_layout(n_gauss, n_pairs, n_rays, k_capacity, v_capacity).total;
}

int us3d_ur_workspace_layout(int64_t n_gauss, int64_t n_pairs, int64_t n_rays, int64_t k_capacity,
                             int64_t v_capacity, us3d_layout* out) {
    if (!out) return fail(US3D_ERR_ARGUMENT, "out is NULL");
    *out = ur_layout(n_gauss, n_pairs, n_rays, k_capacity, v_capacity);
    return US3D_OK;
}

int us3d_ur_map(const us3d_plan* plan, const float* means, const float* scales, const float* quats,
                const float* attenuations, const float* sh, const us3d_ur_settings* settings, void* stream_) {
    const us3d_layout L = ur_layout(plan);
    int rc = check_plan(plan, L.total);
    if (rc) return rc;
    if ((rc = check_ur_settings(settings))) return rc;
    if (plan->n_gauss > 0 && (!means || !scales || !quats || !attenuations || !sh))
        return fail(US3D_ERR_ARGUMENT, "Gaussian parameter pointer is NULL");
    hipStream_t stream = as_stream(stream_);
    float4* rec = ws_at<float4>(plan, L.records);
    if ((rc = begin_pairs(plan, L, stream))) return rc;
    if (plan->n_gauss > 0) {
        hipLaunchKernelGGL(ur_pack_kernel, dim3(blocks_for(plan->n_gauss, 256)), dim3(256), 0, stream, plan->n_gauss,
                           means, scales, quats, attenuations, sh, settings->scales_are_log ? 1 : 0, rec);
        if ((rc = check_launch("ur_pack_kernel"))) return rc;
    }
    if (plan->n_pairs > 0) {
        {
            StageTimer timer_("ur_cull_kernel", stream);
            hipLaunchKernelGGL(ur_cull_kernel, dim3(blocks_for(plan->n_pairs, 256)), dim3(256), 0, stream,
                               plan->n_pairs, rec, plan->frames, plan->sweeps, plan->n_sweeps, settings->k_sigma,
                               settings->elev_cutoff_margin_mm, ws_at<uint64_t>(plan, L.pair_bits),
                               ws_at<int32_t>(plan, L.word_cnt), ws_at<uint32_t>(plan, L.fm_bits));
        }
        if ((rc = check_launch("ur_cull_kernel"))) return rc;
        if ((rc = rank_visible_pairs(plan, L, stream))) return rc;
        {
            StageTimer timer_("ur_setup_kernel", stream);
            hipLaunchKernelGGL(ur_setup_kernel, dim3(blocks_for(plan->v_capacity, 256)), dim3(256), 0, stream,
                               plan->v_capacity, ws_at<int32_t>(plan, L.vis_pair), rec, plan->frames, plan->sweeps,
                               plan->n_sweeps, settings->k_sigma, settings->elev_cutoff_margin_mm,
                               ws_at<us3d_stats>(plan, L.stats), ws_at<int32_t>(plan, L.pair_count),
                               ws_at<int4>(plan, L.pair_rect), ws_at<uint32_t>(plan, L.pair_dk),
                               ws_at<int32_t>(plan, L.tile_count));
        }
        if ((rc = check_launch("ur_setup_kernel"))) return rc;
    }
    return build_instances(plan, L, stream);
}

int us3d_ur_forward(const us3d_plan* plan, const float* pixels, const float* ray_origins, const float* ray_dirs,
                    const float* ray_offsets, const us3d_ur_settings* settings, float* image, float* aux,
                    void* stream_) {
    const us3d_layout L = ur_layout(plan);
    int rc = check_plan(plan, L.total);
    if (rc) return rc;
    if ((rc = check_ur_settings(settings))) return rc;
    if (!image || !aux) return fail(US3D_ERR_ARGUMENT, "ur_forward: NULL array");
    {
        StageTimer timer_("ur_forward_kernel", as_stream(stream_));
        // the line forms when the samples are the linear probe's grid (see ur_forward_kernel)
        auto kernel = pixels ? ur_forward_kernel<false> : ur_forward_kernel<true>;
        hipLaunchKernelGGL(kernel, dim3(plan->n_tiles), dim3(kRayThreads), 0, as_stream(stream_),
                           ws_at<float4>(plan, L.records), ws_at<int32_t>(plan, L.tile_off),
                           ws_at<uint64_t>(plan, L.seg_key), ws_at<int32_t>(plan, L.pair_g),
                           plan->frames, plan->frame_tile_base, plan->n_frames, pixels, ray_origins, ray_dirs, ray_offsets,
                           flags_of(settings), ws_at<us3d_stats>(plan, L.stats), image, reinterpret_cast<float4*>(aux));
    }
    return check_launch("ur_forward_kernel");
}

int us3d_ur_backward(const us3d_plan* plan, const float* pixels, const float* ray_origins, const float* ray_dirs,
                     const float* ray_offsets, const us3d_ur_settings* settings, const float* aux, const float* d_image,
                     float* d_means, float* d_scales, float* d_quats, float* d_attenuations, float* d_sh,
                     void* stream_) {
    const us3d_layout L = ur_layout(plan);
    int rc = check_plan(plan, L.total);
    if (rc) return rc;
    if ((rc = check_ur_settings(settings))) return rc;
    if (!aux || !d_image || !d_means || !d_scales || !d_quats || !d_attenuations || !d_sh)
        return fail(US3D_ERR_ARGUMENT, "ur_backward: NULL array");
    hipStream_t stream = as_stream(stream_);
    {
        StageTimer timer_("ur_backward_kernel", stream);
        // the grid form of the echo sums when the samples are the linear probe's grid (see ur_backward_kernel)
        auto kernel = pixels ? ur_backward_kernel<false> : ur_backward_kernel<true>;
        hipLaunchKernelGGL(kernel, dim3(plan->n_tiles), dim3(kRayThreads), 0, stream,
                           ws_at<float4>(plan, L.records), ws_at<int32_t>(plan, L.tile_off),
                           ws_at<int32_t>(plan, L.pair_g), ws_at<uint64_t>(plan, L.seg_key),
                           ws_at<int32_t>(plan, L.pair_off),
                           ws_at<int4>(plan, L.pair_rect), plan->frames, plan->frame_tile_base, plan->n_frames, pixels,
                           ray_origins, ray_dirs, ray_offsets, flags_of(settings), reinterpret_cast<const float4*>(aux),
                           d_image, ws_at<us3d_stats>(plan, L.stats), ws_at<float4>(plan, L.partial));
    }
    if ((rc = check_launch("ur_backward_kernel"))) return rc;
    if (plan->n_gauss > 0) {
        {
            StageTimer timer_("ur_reduce_kernel", stream);
            hipLaunchKernelGGL(ur_reduce_kernel, dim3(blocks_for(plan->n_gauss, 4)), dim3(256), 0, stream, plan->n_gauss,
                               plan->sweeps, plan->n_sweeps, ws_at<uint64_t>(plan, L.pair_bits),
                               ws_at<int32_t>(plan, L.word_off), ws_at<int32_t>(plan, L.pair_off),
                               ws_at<float4>(plan, L.partial),
                               ws_at<float4>(plan, L.records), settings->scales_are_log ? 1 : 0,
                               ws_at<us3d_stats>(plan, L.stats), d_means, d_scales, d_quats, d_attenuations, d_sh);
        }
        if ((rc = check_launch("ur_reduce_kernel"))) return rc;
    }
    return US3D_OK;
}

}  // extern "C"
