// GaussianPile "us_focus" slice renderer: a second slice model on the same tile pipeline as the UltraGauss
// splat, with an additive weight. Follows the reference's CUDA path for shader weight mode US_FOCUS with the
// tile mapper and additive compositing, configured as its facade configures it
// (ops/gaussianplie_renderer/gaussianplie_renderer/autograd.py:17-55): constant elevational focus width
// sigma_z, no lateral focus, det_gamma = 0.5, no inflation.
//
// Per (frame, Gaussian) pair, the footprint (include/shader/us_focus_device.cuh:176-300): probe-frame mean
// mu_c and covariance Sc (world covariance as plane_cond2d_device.cuh:48-69 forms it, R diag(s^2) R with the
// rotation's row-major entries, then A^T Sigma_w A and a 1e-6 diagonal jitter); Lc = Sc^-1,
// Se = (Lc + e_ele e_ele^T / sigma_z^2)^-1, mu_e = Se Lc mu_c; the 2-D mean mu_e[0:2], covariance Se[0:2,0:2]
// (symmetrised) and its inverse; k-sigma box radii; base weight o * exp(-mu_ele^2 / 2 (Sc22 + sigma_z^2 +
// det_eps)) / sqrt(det2 + det_eps). Per pixel of a tile (shader/forward.cu:94-111,114-210): inside the box,
// quad = d^T inv2 d >= 0, w = b exp(-quad / 2) >= weight_min, image += intensity * w, summed in FP32 in list
// order. The mapper (mapper/core_tile.cu:25-170) does no elevation cull in this mode; each tile lists its
// Gaussians in index order (depth key 0).
//
// The footprint is evaluated with the reference's FP32 operation order (no fused multiply-adds; the library
// is built with -ffp-contract=off), so its decisions are those of the CPU restatement bit for bit. A pair
// whose base weight is below weight_min is not listed: every weight of it is below weight_min. Per pixel the
// weight is b exp2(e) with e = dl l + da q, (l, q) = C inv2 (dl, da), C = -log2(e)/2 stored pre-multiplied
// into the inverse covariance: three fused multiply-adds and one hardware exp2 instead of exp(-quad/2)
// (the same value to a few FP32 ulps; the CPU restatement follows this order).
//
// The adjoint is the exact derivative of this forward for every parameter: per (tile, Gaussian) the
// record-space gradients (base weight, intensity, 2-D mean, inverse covariance) without atomics, then per
// Gaussian over its pairs the vector-Jacobian product of the footprint (two 3x3 inversions, the quaternion
// and its normalisation).

#include "common.h"

namespace us3d {

// Per visible pair, kFootF4 float4s in pair_sect: its footprint (two float4) and its emit base (first emit index,
// rectangle x0, y0, width; written once the counts are scanned), read by the adjoint from the same record.
constexpr int kFootF4 = 3;

static us3d_layout gp_layout(int64_t n_gauss, int64_t n_pairs, int64_t n_tiles, int64_t k_cap, int64_t v_cap) {
    return ws_layout(n_gauss, n_pairs, n_tiles, k_cap, v_cap, 48, 32, 0, 16 * kFootF4);
}

static us3d_layout gp_layout(const us3d_plan* plan) {
    return plan ? gp_layout(plan->n_gauss, plan->n_pairs, plan->n_tiles, plan->k_capacity, plan->v_capacity)
                : us3d_layout{};
}

template <typename T>
static T* gp_at(const us3d_plan* plan, int64_t off) {
    return reinterpret_cast<T*>(reinterpret_cast<char*>(plan->workspace) + off);
}

// The pixel weight b exp(-quad/2) is evaluated as b exp2(e), e = C quad, C = -log2(e)/2 folded into the stored
// inverse covariance: one hardware exp2 per (pixel, Gaussian). Its decisions (e <= 0, weight >= weight_min) are
// made on the same FP32 numbers by the forward and the adjoint.
constexpr float kGpExp2Scale = -0.72134752044448170f;
constexpr float kGpLn2 = 0.69314718055994531f;

__device__ __forceinline__ int gp_clamp_to_int(float x) {
    return (int)fminf(fmaxf(x, -1073741824.0f), 1073741824.0f);
}

// Record per Gaussian: (mu, opacity), (scale, intensity), (quaternion w, x, y, z). With raw activations the
// opacity is sigmoid(x) and each scale max(exp(x), 1e-5) (pipeline_god_file.py:360-370).
__device__ __forceinline__ float gp_scale_act(float x) { return fmaxf(expf(x), 1e-5f); }

__global__ void gp_pack_kernel(int n, const float* __restrict__ means, const float* __restrict__ scales,
                               const float* __restrict__ quats, const float* __restrict__ opacity,
                               const float* __restrict__ intensity, int raw, float4* __restrict__ rec) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const float* m = means + 3 * (int64_t)g;
    const float* s = scales + 3 * (int64_t)g;
    const float* q = quats + 4 * (int64_t)g;
    const float o = raw ? 1.0f / (1.0f + expf(-opacity[g])) : opacity[g];
    const float s0 = raw ? gp_scale_act(s[0]) : s[0], s1 = raw ? gp_scale_act(s[1]) : s[1];
    const float s2 = raw ? gp_scale_act(s[2]) : s[2];
    rec[3 * (int64_t)g + 0] = make_float4(m[0], m[1], m[2], o);
    rec[3 * (int64_t)g + 1] = make_float4(s0, s1, s2, intensity ? intensity[g] : 1.0f);
    rec[3 * (int64_t)g + 2] = make_float4(q[0], q[1], q[2], q[3]);
}

struct GPFoot {
    float m0, m1, i00, i01, i11, r_lat, r_ax, b;
    bool ok;
};

// What the adjoint needs besides the footprint.
struct GPInter {
    float muc[3];
    float qn[4], nq, qq;
    float R[9];
    float s[3];
    bool s_act[3];
    float Lc[9], Se[9], t[3];
    float s00, s01, s11, det2, inv_det2;
    float sr, norm, denom, denom_raw, o;
};

__device__ __forceinline__ float gp_dot3(float ax, float ay, float az, float bx, float by, float bz) {
    return ax * bx + ay * by + az * bz;
}

// Cofactor inverse of a general 3x3 matrix (row-major); returns the determinant.
__device__ __forceinline__ float gp_inv3(const float M[9], float Mi[9]) {
    const float a00 = M[0], a01 = M[1], a02 = M[2], a10 = M[3], a11 = M[4], a12 = M[5];
    const float a20 = M[6], a21 = M[7], a22 = M[8];
    const float c00 = a11 * a22 - a12 * a21;
    const float c01 = a02 * a21 - a01 * a22;
    const float c02 = a01 * a12 - a02 * a11;
    const float c10 = a12 * a20 - a10 * a22;
    const float c11 = a00 * a22 - a02 * a20;
    const float c12 = a02 * a10 - a00 * a12;
    const float c20 = a10 * a21 - a11 * a20;
    const float c21 = a01 * a20 - a00 * a21;
    const float c22 = a00 * a11 - a01 * a10;
    const float det = a00 * c00 + a01 * c10 + a02 * c20;
    const float inv_det = 1.f / det;
    Mi[0] = c00 * inv_det; Mi[1] = c01 * inv_det; Mi[2] = c02 * inv_det;
    Mi[3] = c10 * inv_det; Mi[4] = c11 * inv_det; Mi[5] = c12 * inv_det;
    Mi[6] = c20 * inv_det; Mi[7] = c21 * inv_det; Mi[8] = c22 * inv_det;
    return det;
}

template <bool kInter>
__device__ __forceinline__ void gp_footprint(float4 r0, float4 r1, float4 r2, const us3d_frame& fr, float k_sigma,
                                             float sigma_ele, float det_eps, GPFoot& F, GPInter* I,
                                             float early_weight_min = 0.f) {
    const float* al = fr.axis_lat;
    const float* aa = fr.axis_ax;
    const float* ae = fr.axis_ele;
    const float d0 = r0.x - fr.origin[0], d1 = r0.y - fr.origin[1], d2 = r0.z - fr.origin[2];
    const float mu_lat = gp_dot3(d0, d1, d2, al[0], al[1], al[2]);
    const float mu_ax = gp_dot3(d0, d1, d2, aa[0], aa[1], aa[2]);
    const float mu_ele = gp_dot3(d0, d1, d2, ae[0], ae[1], ae[2]);
    // quaternion (w, x, y, z) -> rotation (shared.cuh:16-34), normalised with a correctly rounded 1/sqrt
    const float qq = r2.x * r2.x + r2.y * r2.y + r2.z * r2.z + r2.w * r2.w;
    const float nq = 1.f / sqrtf(fmaxf(qq, 1e-12f));
    const float w = r2.x * nq, x = r2.y * nq, y = r2.z * nq, z = r2.w * nq;
    const float xx = x * x, yy = y * y, zz = z * z;
    const float xy = x * y, xz = x * z, yz = y * z;
    const float wx = w * x, wy = w * y, wz = w * z;
    float R[9];
    R[0] = 1.f - 2.f * (yy + zz); R[1] = 2.f * (xy - wz); R[2] = 2.f * (xz + wy);
    R[3] = 2.f * (xy + wz); R[4] = 1.f - 2.f * (xx + zz); R[5] = 2.f * (yz - wx);
    R[6] = 2.f * (xz - wy); R[7] = 2.f * (yz + wx); R[8] = 1.f - 2.f * (xx + yy);
    const float s[3] = {fmaxf(r1.x, 1e-6f), fmaxf(r1.y, 1e-6f), fmaxf(r1.z, 1e-6f)};
    float M[9], Sw[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) M[k] = R[k] * s[k % 3] * s[k % 3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Sw[i * 3 + j] = M[i * 3 + 0] * R[j] + M[i * 3 + 1] * R[3 + j] + M[i * 3 + 2] * R[6 + j];
    const float* ax3[3] = {al, aa, ae};
    float col[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int r = 0; r < 3; ++r)
            col[a][r] = Sw[r * 3 + 0] * ax3[a][0] + Sw[r * 3 + 1] * ax3[a][1] + Sw[r * 3 + 2] * ax3[a][2];
    float Sc[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            Sc[i * 3 + j] = ax3[i][0] * col[j][0] + ax3[i][1] * col[j][1] + ax3[i][2] * col[j][2];
    Sc[0] += 1e-6f;
    Sc[4] += 1e-6f;
    Sc[8] += 1e-6f;
    const float sig2 = sigma_ele * sigma_ele;
    const float denom_raw = Sc[8] + sig2 + det_eps;
    const float denom = fmaxf(denom_raw, 1e-12f);
    const float sr = expf(-0.5f * mu_ele * mu_ele / denom);
    // Exact early cull (early_weight_min > 0): a valid footprint has det2 > det_eps, so its normaliser
    // 1/sqrt(det2 + det_eps) is below 1/sqrt(2 det_eps) (the 0.1 % margin covers FP32 rounding); when even that
    // bound gives a base weight below weight_min, the pair would not be listed and the inversions are skipped.
    if (early_weight_min > 0.f && r0.w * sr * (1.001f / sqrtf(2.f * det_eps)) < early_weight_min) {
        F.ok = false;
        F.b = 0.f;
        return;
    }
    float Lc[9], Le[9], Se[9];
    const float det_c = gp_inv3(Sc, Lc);
#pragma unroll
    for (int k = 0; k < 9; ++k) Le[k] = Lc[k];
    Le[8] = Le[8] + 1.f / fmaxf(sig2, 1e-12f);
    const float det_e = gp_inv3(Le, Se);
    const float t0 = Lc[0] * mu_lat + Lc[1] * mu_ax + Lc[2] * mu_ele;
    const float t1 = Lc[3] * mu_lat + Lc[4] * mu_ax + Lc[5] * mu_ele;
    const float t2 = Lc[6] * mu_lat + Lc[7] * mu_ax + Lc[8] * mu_ele;
    F.m0 = Se[0] * t0 + Se[1] * t1 + Se[2] * t2;
    F.m1 = Se[3] * t0 + Se[4] * t1 + Se[5] * t2;
    const float s00 = Se[0], s01 = 0.5f * (Se[1] + Se[3]), s11 = Se[4];
    const float det2 = s00 * s11 - s01 * s01;
    const float inv_det2 = 1.f / det2;
    F.i00 = s11 * inv_det2;
    F.i01 = -s01 * inv_det2;
    F.i11 = s00 * inv_det2;
    const float cross = fabsf(s01);
    F.r_lat = k_sigma * sqrtf(fmaxf(s00 + cross, 1e-12f));
    F.r_ax = k_sigma * sqrtf(fmaxf(s11 + cross, 1e-12f));
    const float norm = 1.f / sqrtf(det2 + det_eps);
    F.b = r0.w * sr * norm;
    F.ok = det_c > 1e-12f && isfinite(det_c) && det_e > 1e-12f && isfinite(det_e) && det2 > det_eps &&
           isfinite(det2) && isfinite(F.b);
    if constexpr (kInter) {
        I->muc[0] = mu_lat; I->muc[1] = mu_ax; I->muc[2] = mu_ele;
        I->qn[0] = w; I->qn[1] = x; I->qn[2] = y; I->qn[3] = z;
        I->nq = nq; I->qq = qq;
#pragma unroll
        for (int k = 0; k < 9; ++k) { I->R[k] = R[k]; I->Lc[k] = Lc[k]; I->Se[k] = Se[k]; }
        I->s[0] = s[0]; I->s[1] = s[1]; I->s[2] = s[2];
        I->s_act[0] = r1.x > 1e-6f; I->s_act[1] = r1.y > 1e-6f; I->s_act[2] = r1.z > 1e-6f;
        I->t[0] = t0; I->t[1] = t1; I->t[2] = t2;
        I->s00 = s00; I->s01 = s01; I->s11 = s11; I->det2 = det2; I->inv_det2 = inv_det2;
        I->sr = sr; I->norm = norm; I->denom = denom; I->denom_raw = denom_raw; I->o = r0.w;
    }
}

// ---------------------------------------------------------------- mapper
// Tile rectangle of one (frame, Gaussian) pair from its footprint (core_tile.cu:138-163); returns the instance
// count, 0 when the footprint is invalid, its peak weight is below weight_min or its box misses the frame. The cull
// and the setup pass call it on the same inputs, so they agree on every pair.
__device__ __forceinline__ int gp_pair_tiles(const GPFoot& F, const us3d_frame& fr, const us3d_gp_settings& st,
                                             int4& rect) {
    if (!(F.ok && F.b >= st.weight_min)) return 0;
    const float u = F.m0 / fr.delta_lat + fr.col_offset;
    const float v = (F.m1 - fr.near_mm) / fr.delta_ax;
    const float ru = fmaxf(1.0f, F.r_lat / fmaxf(1e-8f, fr.delta_lat));
    const float rv = fmaxf(1.0f, F.r_ax / fmaxf(1e-8f, fr.delta_ax));
    const float tsf = (float)st.tile_size;
    const float fx0 = floorf((u - ru) / tsf), fy0 = floorf((v - rv) / tsf);
    const float fx1 = ceilf((u + ru) / tsf), fy1 = ceilf((v + rv) / tsf);
    if (!(isfinite(fx0) && isfinite(fy0) && isfinite(fx1) && isfinite(fy1))) return 0;
    const int x0 = max(0, gp_clamp_to_int(fx0)), y0 = max(0, gp_clamp_to_int(fy0));
    const int x1 = min(fr.tiles_x, gp_clamp_to_int(fx1)), y1 = min(fr.tiles_y, gp_clamp_to_int(fy1));
    if (!(x0 < x1 && y0 < y1)) return 0;
    rect = make_int4(x0, y0, x1, y1);
    return (x1 - x0) * (y1 - y0);
}

// A pair gp_footprint's early cull rejects whatever the Gaussian's orientation: the footprint's elevational variance
// a_ele^T Sigma a_ele is at most max(s)^2 (+ the 1e-6 it adds), so the base weight is at most
// o exp(-ele^2 / (2 dmax)) 1.001 / sqrt(2 det_eps), dmax = max(s)^2 + 1e-6 + sigma_ele^2 + det_eps. The elevation is
// formed as gp_footprint forms it; dmax takes 0.1 % more for the rounding of the projected variance and the bound
// another 0.1 %, so a pair rejected here is rejected there. Costs the elevation and one exp, where the footprint
// builds the rotation and the projected covariance.
__device__ __forceinline__ bool gp_far_from_plane(const float4& r0, const float4& r1, const us3d_frame& fr,
                                                  const us3d_gp_settings& st) {
    const float* ae = fr.axis_ele;
    const float d0 = r0.x - fr.origin[0], d1 = r0.y - fr.origin[1], d2 = r0.z - fr.origin[2];
    const float ele = gp_dot3(d0, d1, d2, ae[0], ae[1], ae[2]);
    const float smax = fmaxf(fmaxf(fmaxf(r1.x, r1.y), r1.z), 1e-6f);
    const float dmax = smax * smax * 1.001f + 1e-6f + st.sigma_ele * st.sigma_ele + st.det_eps;
    const float bound = r0.w * expf(-0.5f * ele * ele / dmax) * (1.001f / sqrtf(2.f * st.det_eps));
    return bound * 1.001f < st.weight_min;  // NaN: false, the footprint decides
}

// Cull pass: thread per dense (frame, Gaussian) pair; marks the pairs with at least one tile.
__global__ __launch_bounds__(256) void gp_cull_kernel(int64_t n_pairs, const float4* __restrict__ rec,
                                                      const us3d_frame* __restrict__ frames,
                                                      const us3d_sweep* __restrict__ sweeps, int n_sweeps,
                                                      us3d_gp_settings st, uint64_t* __restrict__ pair_bits,
                                                      int32_t* __restrict__ word_cnt, uint32_t* __restrict__ fm_bits) {
    const int64_t pair = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int count = 0, s = 0, f = 0, gl = 0;
    if (pair < n_pairs) {
        s = sweep_of_pair(sweeps, n_sweeps, pair);
        const us3d_sweep sw = sweeps[s];
        const int local = (int)(pair - sw.pair_base);
        gl = local / sw.f_count;
        const int g = sw.g_begin + gl;
        f = sw.f_begin + (local - gl * sw.f_count);
        const us3d_frame& fr = frames[f];
        const float4 r0 = rec[3 * (int64_t)g + 0], r1 = rec[3 * (int64_t)g + 1];
        if (!gp_far_from_plane(r0, r1, fr, st)) {  // most pairs of a freehand sweep leave here
            GPFoot F;
            gp_footprint<false>(r0, r1, rec[3 * (int64_t)g + 2], fr, st.k_sigma, st.sigma_ele, st.det_eps, F, nullptr,
                                st.weight_min);
            int4 rect;
            count = gp_pair_tiles(F, fr, st, rect);
        }
    }
    mark_pair(count > 0, pair, n_pairs, sweeps, s, f, gl, pair_bits, word_cnt, fm_bits);
}

// Setup pass: thread per visible pair v < v_capacity (entries past the visible count, or every entry when the
// visible pairs overflowed, get count 0): count, tile rectangle, the footprint record, tiles counted by corners.
__global__ __launch_bounds__(256) void gp_setup_kernel(
    int64_t v_cap, const int32_t* __restrict__ vis_pair, const float4* __restrict__ rec,
    const us3d_frame* __restrict__ frames, const us3d_sweep* __restrict__ sweeps, int n_sweeps, us3d_gp_settings st,
    const us3d_stats* __restrict__ stats, int32_t* __restrict__ pair_count, int4* __restrict__ pair_rect,
    uint32_t* __restrict__ pair_dk, float4* __restrict__ pair_foot, int32_t* __restrict__ tile_count) {
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
    GPFoot F;
    gp_footprint<false>(rec[3 * (int64_t)g + 0], rec[3 * (int64_t)g + 1], rec[3 * (int64_t)g + 2], fr, st.k_sigma,
                        st.sigma_ele, st.det_eps, F, nullptr, st.weight_min);
    int4 rect = make_int4(0, 0, 0, 0);
    const int count = gp_pair_tiles(F, fr, st, rect);
    pair_count[v] = count;
    if (count > 0) {  // always, for a listed pair
        pair_rect[v] = rect;
        pair_dk[v] = 0u;
        pair_foot[kFootF4 * v + 0] = make_float4(F.m0, F.m1, F.i00 * kGpExp2Scale, F.i01 * kGpExp2Scale);
        pair_foot[kFootF4 * v + 1] = make_float4(F.i11 * kGpExp2Scale, F.r_lat, F.r_ax, F.b);
        add_rect_corners(tile_count, fr.tile_base, fr.tiles_x, fr.tiles_y, rect);
    }
}

// ---------------------------------------------------------------- forward
// One workgroup per (frame, tile); a thread owns GpFwd<TS>::kPix pixels of one column (rows kRowStep apart),
// so the instance's LDS reads and its lateral offset are shared by them. The tile's list is staged through
// LDS in batches of one instance per thread.
template <int TS>
struct GpFwd {
    static constexpr int kPix = (TS == 8) ? 1 : 4;
    static constexpr int kThreads = TS * TS / kPix;
    static constexpr int kRowStep = TS / kPix;
};

template <int TS>
__global__ __launch_bounds__(GpFwd<TS>::kThreads) void gp_forward_kernel(
    const float4* __restrict__ rec, const float4* __restrict__ pair_foot, const int32_t* __restrict__ tile_off,
    const int32_t* __restrict__ pair_g, const uint64_t* __restrict__ seg_key, const us3d_frame* __restrict__ frames,
    const int32_t* __restrict__ frame_tile_base, int n_frames, float weight_min,
    const us3d_stats* __restrict__ stats, float* __restrict__ image, int n_tiles, int32_t* __restrict__ queue_heads,
    LossFuse loss) {
    using S = GpFwd<TS>;
    constexpr int NT = S::kThreads;
    constexpr int NP = S::kPix;
    static_assert(NT % kWave == 0, "the fused loss sums whole wavefronts");
    __shared__ float4 s_f0[NT], s_f1[NT];
    __shared__ float s_I[NT];
    __shared__ int s_tile;
    if (stats->overflow) return;
    auto item = [&](int t) {
        const int f = upper_index(frame_tile_base, n_frames, t);
        const us3d_frame& fr = frames[f];
        const int lt = t - fr.tile_base;
        const int tx = lt % fr.tiles_x, ty = lt / fr.tiles_x;
        const int tid = threadIdx.x;
        const int ix = tx * TS + tid % TS, iy0 = ty * TS + tid / TS;
        const float xl = ((float)ix - fr.col_offset) * fr.delta_lat;
        float xa[NP], acc[NP];
#pragma unroll
        for (int r = 0; r < NP; ++r) {
            xa[r] = (float)(iy0 + r * S::kRowStep) * fr.delta_ax + fr.near_mm;
            acc[r] = 0.f;
        }
        const int s0 = tile_off[t], s1 = tile_off[t + 1];
        for (int base = s0; base < s1; base += NT) {
            const int n = min(NT, s1 - base);
            __syncthreads();
            if (tid < n) {
                const int64_t p = (uint32_t)seg_key[base + tid];  // the instance's visible pair
                const int g = pair_g[p];
                s_f0[tid] = pair_foot[kFootF4 * p + 0];
                s_f1[tid] = pair_foot[kFootF4 * p + 1];
                s_I[tid] = rec[3 * (int64_t)g + 1].w;
            }
            __syncthreads();
            for (int k = 0; k < n; ++k) {
                const float4 a = s_f0[k], c = s_f1[k];  // (m0, m1, C i00, C i01), (C i11, r_lat, r_ax, b)
                const float I = s_I[k];
                const float dl = xl - a.x;
                const bool lat_in = fabsf(dl) <= c.y;
#pragma unroll
                for (int r = 0; r < NP; ++r) {
                    const float da = xa[r] - a.y;
                    const float l = __builtin_fmaf(a.z, dl, a.w * da);
                    const float q = __builtin_fmaf(a.w, dl, c.x * da);
                    const float e = __builtin_fmaf(dl, l, da * q);
                    const float wgt = c.w * __builtin_amdgcn_exp2f(e);
                    // box, quad >= 0 (a NaN fails it), weight >= weight_min (an infinite quad gives weight 0)
                    const bool in = lat_in && fabsf(da) <= c.z && e <= 0.f && wgt >= weight_min;
                    acc[r] += in ? I * wgt : 0.f;
                }
            }
        }
        // the loss sums of the thread's pixels (fused training step): at most four pixels in FP32, then FP64
        // across the wavefront and the frame; the targets are loaded before the image stores
        float gtv[NP];
#pragma unroll
        for (int r = 0; r < NP; ++r) {
            const int iy = iy0 + r * S::kRowStep;
            gtv[r] = (loss.gt && ix < fr.width && iy < fr.height)
                         ? loss.gt[(int64_t)fr.pixel_base + (int64_t)iy * fr.width + ix] : 0.f;
        }
        float ls[kLossSums] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int r = 0; r < NP; ++r) {
            const int iy = iy0 + r * S::kRowStep;
            if (ix < fr.width && iy < fr.height) {
                image[(int64_t)fr.pixel_base + (int64_t)iy * fr.width + ix] = acc[r];
                const float x = acc[r], y = gtv[r];
                ls[0] += fabsf(x - y);
                ls[1] += x;
                ls[2] += y;
                ls[3] = __builtin_fmaf(x, x, ls[3]);
                ls[4] = __builtin_fmaf(y, y, ls[4]);
                ls[5] = __builtin_fmaf(x, y, ls[5]);
            }
        }
        if (loss.gt) {  // uniform: one entry per (tile, wavefront)
            const int64_t e = (int64_t)t * (NT / kWave) + tid / kWave;
#pragma unroll
            for (int j = 0; j < kLossSums; ++j) {
                const double v = wave_sum_f64((double)ls[j]);
                if ((tid & (kWave - 1)) == j) loss.part[e * kLossSums + j] = v;
            }
        }
    };
    if (!queue_heads) {  // one workgroup per tile
        item(remap_xcd(blockIdx.x, (int)gridDim.x));
        return;
    }
    int q = blockIdx.x % kXcds;  // persistent workgroups over the per-XCD tile queues
    for (int t = next_queue_tile(queue_heads, q, n_tiles, &s_tile); t >= 0;
         t = next_queue_tile(queue_heads, q, n_tiles, &s_tile))
        item(t);
}

// ---------------------------------------------------------------- adjoint
// Record-space gradients per (tile, Gaussian). The work item is one half of the tile's rows for one instance:
// the two items of an instance sit on adjacent lanes, each walks its rows of the instance's box and the two
// sums are added with one lane swap (a fixed order), then the even lane writes the instance's partial at its
// emit index (no atomics). With t = dL/dimage * w summed over the pixels, the intensity I factors out:
//   dL/dI = sum t, dL/db = I sum t / b, dL/dinv2 = -I/2 sum t (dl^2, 2 dl da, da^2),
//   dL/dmu2 = -2 ln2 I sum t (l, q) with (l, q) = C inv2 d, the pre-scaled half-gradient of e = C quad.
template <int TS>
__global__ __launch_bounds__(TS * TS) void gp_backward_kernel(
    const float4* __restrict__ rec, const float4* __restrict__ pair_foot, const int32_t* __restrict__ tile_off,
    const int32_t* __restrict__ pair_g, const uint64_t* __restrict__ seg_key, const us3d_frame* __restrict__ frames,
    const int32_t* __restrict__ frame_tile_base, int n_frames, float weight_min, const float* __restrict__ d_image,
    const us3d_stats* __restrict__ stats, float4* __restrict__ partial, int n_tiles,
    int32_t* __restrict__ queue_heads, const float* __restrict__ image, LossFuse loss) {
    constexpr int NT = TS * TS;
    constexpr int HALF = TS / 2;
    static_assert(TS % 4 == 0, "rows of the staged gradient are read as float4");
    __shared__ __align__(16) float s_g[NT];  // dL/dimage of the tile's pixels
    __shared__ __align__(16) float s_xl[TS];  // lateral position of each column
    __shared__ int s_tile;
    if (stats->overflow) return;
    auto item = [&](int t) {
        const int s0 = tile_off[t], s1 = tile_off[t + 1];
        if (s0 == s1) return;
        const int f = upper_index(frame_tile_base, n_frames, t);
        const us3d_frame& fr = frames[f];
        const int lt = t - fr.tile_base;
        const int tx = lt % fr.tiles_x, ty = lt / fr.tiles_x;
        const int tid = threadIdx.x;
        const int cx0 = tx * TS, cy0 = ty * TS;
        {
            const int ix = cx0 + tid % TS, iy = cy0 + tid / TS;
            float gi = 0.f;
            if (ix < fr.width && iy < fr.height) {
                const int64_t p = (int64_t)fr.pixel_base + (int64_t)iy * fr.width + ix;
                if (loss.coeff) {  // uniform: dL/dx of the fused image loss, x the image the forward wrote
                    const double* cf = loss.coeff + 4 * f;
                    const double xi = image[p], yi = loss.gt[p];
                    const double sg = (xi > yi) ? 1.0 : ((xi < yi) ? -1.0 : 0.0);
                    gi = (float)(cf[0] * sg + cf[1] + cf[2] * xi + cf[3] * yi);
                } else {
                    gi = d_image[p];
                }
            }
            s_g[tid] = gi;
            if (tid < TS) s_xl[tid] = ((float)(cx0 + tid) - fr.col_offset) * fr.delta_lat;
        }
        __syncthreads();
        const int n_items = 2 * (s1 - s0);
        // NT is even, so both items of an instance run in the same iteration, on lanes tid and tid ^ 1
        for (int it = tid; it < n_items; it += NT) {
            const int i = s0 + (it >> 1);
            const int half = it & 1;
            const int pv = (int)(uint32_t)seg_key[i];  // the instance's visible pair
            const int g = pair_g[pv];
            const float4 a = pair_foot[kFootF4 * (int64_t)pv + 0], c = pair_foot[kFootF4 * (int64_t)pv + 1];  // (m0, m1, C i00, C i01), (C i11, r_lat, r_ax, b)
            // every lane walks all pixels of its half tile in the same order (uniform loops, columns unrolled); the
            // box test excludes the pixels outside the instance's box and pixels outside the frame have gradient 0
            // da is constant along a row, so per row only sum t, sum t dl and sum t dl^2 are accumulated per pixel;
            // the sums over da (and the pre-scaled half-gradients l, q, which are linear in dl and da) follow
            // from them once per row
            float sT = 0.f, sA = 0.f, sB = 0.f, sC = 0.f, sD = 0.f, sE = 0.f;  // sD = sum t dl, sE = sum t da
#pragma unroll 1
            for (int r = 0; r < HALF; ++r) {
                const int yy = half * HALF + r;
                const float xa_ = (float)(cy0 + yy) * fr.delta_ax + fr.near_mm;
                const float da = xa_ - a.y;
                const bool row_in = fabsf(da) <= c.z;
                const float c01da = a.w * da, c11da = c.x * da;
                const float4* grow = reinterpret_cast<const float4*>(s_g + yy * TS);
                float r0 = 0.f, r1 = 0.f, r2 = 0.f;
#pragma unroll 2
                for (int x4 = 0; x4 < TS / 4; ++x4) {
                    const float4 go4 = grow[x4];
                    const float4 xl4 = reinterpret_cast<const float4*>(s_xl)[x4];
                    const float gov[4] = {go4.x, go4.y, go4.z, go4.w};
                    const float xlv[4] = {xl4.x, xl4.y, xl4.z, xl4.w};
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const float dl = xlv[j] - a.x;
                        const float l = __builtin_fmaf(a.z, dl, c01da);
                        const float q = __builtin_fmaf(a.w, dl, c11da);
                        const float e = __builtin_fmaf(dl, l, da * q);
                        const float wgt = c.w * __builtin_amdgcn_exp2f(e);
                        const bool in = row_in && fabsf(dl) <= c.y && e <= 0.f && wgt >= weight_min;
                        const float tw = in ? gov[j] * wgt : 0.f;
                        r0 += tw;
                        const float td = tw * dl;
                        r1 += td;
                        r2 = __builtin_fmaf(td, dl, r2);
                    }
                }
                sT += r0;
                sA += r2;
                sD += r1;
                sE = __builtin_fmaf(da, r0, sE);
                sB = __builtin_fmaf(da, r1, sB);
                sC = __builtin_fmaf(da * da, r0, sC);
            }
            sT += __shfl_xor(sT, 1, kWave);
            sA += __shfl_xor(sA, 1, kWave);
            sB += __shfl_xor(sB, 1, kWave);
            sC += __shfl_xor(sC, 1, kWave);
            sD += __shfl_xor(sD, 1, kWave);
            sE += __shfl_xor(sE, 1, kWave);
            // sum t l = C i00 sum t dl + C i01 sum t da, sum t q = C i01 sum t dl + C i11 sum t da
            const float sP = __builtin_fmaf(a.z, sD, a.w * sE);
            const float sQ = __builtin_fmaf(a.w, sD, c.x * sE);
            if (half == 0) {
                const float I = rec[3 * (int64_t)g + 1].w;
                const float mI = -2.f * kGpLn2 * I;
                const int64_t ei = emit_index_rec(pair_foot[kFootF4 * (int64_t)pv + 2], tx, ty);
                partial[2 * ei + 0] = make_float4(I * sT / c.w, sT, mI * sP, mI * sQ);
                partial[2 * ei + 1] = make_float4(-0.5f * I * sA, -I * sB, -0.5f * I * sC, 0.f);
            }
        }
    };
    if (!queue_heads) {  // one workgroup per tile
        item(remap_xcd(blockIdx.x, (int)gridDim.x));
        return;
    }
    int q = blockIdx.x % kXcds;  // persistent workgroups over the per-XCD tile queues
    for (int t = next_queue_tile(queue_heads, q, n_tiles, &s_tile); t >= 0;
         t = next_queue_tile(queue_heads, q, n_tiles, &s_tile))
        item(t);
}

// Vector-Jacobian product of the footprint: record-space gradients of one pair -> parameter gradients.
// acc: d mu (3), d scale (3), d quat (4), d opacity.
__device__ __forceinline__ void gp_footprint_vjp(const GPInter& I, const us3d_frame& fr, float gb, float gm0,
                                                 float gm1, float gi00, float gi01, float gi11, float acc[11]) {
    // base weight b = o sr norm
    const float g_sr = gb * I.o * I.norm;
    const float g_norm = gb * I.o * I.sr;
    acc[10] += gb * I.sr * I.norm;
    float g_det2 = g_norm * -0.5f * I.norm * I.norm * I.norm;
    // inverse of the 2x2 covariance
    const float g_invdet = gi00 * I.s11 - gi01 * I.s01 + gi11 * I.s00;
    float g_s11 = gi00 * I.inv_det2, g_s01 = -gi01 * I.inv_det2, g_s00 = gi11 * I.inv_det2;
    g_det2 += -g_invdet * I.inv_det2 * I.inv_det2;
    g_s00 += g_det2 * I.s11;
    g_s11 += g_det2 * I.s00;
    g_s01 += -2.f * I.s01 * g_det2;
    // slice response sr = exp(-mu_ele^2 / (2 denom))
    float g_muc[3] = {0.f, 0.f, 0.f};
    const float mue = I.muc[2];
    g_muc[2] += g_sr * I.sr * (-mue / I.denom);
    const float g_denom = g_sr * I.sr * (0.5f * mue * mue / (I.denom * I.denom));
    // Se -> (2-D covariance, 2-D mean)
    float gSe[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    gSe[0] += g_s00;
    gSe[1] += 0.5f * g_s01;
    gSe[3] += 0.5f * g_s01;
    gSe[4] += g_s11;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        gSe[j] += gm0 * I.t[j];
        gSe[3 + j] += gm1 * I.t[j];
    }
    float g_t[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) g_t[j] = I.Se[j] * gm0 + I.Se[3 + j] * gm1;
    // t = Lc mu_c
    float gLc[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) gLc[i * 3 + j] = g_t[i] * I.muc[j];
#pragma unroll
    for (int j = 0; j < 3; ++j) g_muc[j] += I.Lc[j] * g_t[0] + I.Lc[3 + j] * g_t[1] + I.Lc[6 + j] * g_t[2];
    // Se = Le^-1: gLe = -Se^T gSe Se^T; Le = Lc + const
    {
        float T[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)  // T = Se^T gSe
                T[i * 3 + j] = I.Se[i] * gSe[j] + I.Se[3 + i] * gSe[3 + j] + I.Se[6 + i] * gSe[6 + j];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)  // gLc -= T Se^T
                gLc[i * 3 + j] -= T[i * 3 + 0] * I.Se[j * 3 + 0] + T[i * 3 + 1] * I.Se[j * 3 + 1] + T[i * 3 + 2] * I.Se[j * 3 + 2];
    }
    // Lc = Sc^-1: gSc = -Lc^T gLc Lc^T
    float gSc[9];
    {
        float T[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                T[i * 3 + j] = I.Lc[i] * gLc[j] + I.Lc[3 + i] * gLc[3 + j] + I.Lc[6 + i] * gLc[6 + j];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                gSc[i * 3 + j] = -(T[i * 3 + 0] * I.Lc[j * 3 + 0] + T[i * 3 + 1] * I.Lc[j * 3 + 1] + T[i * 3 + 2] * I.Lc[j * 3 + 2]);
    }
    if (I.denom_raw > 1e-12f) gSc[8] += g_denom;
    // Sc = A^T Sw A: gSw = sum_ij gSc[i][j] a_i a_j^T
    const float* ax3[3] = {fr.axis_lat, fr.axis_ax, fr.axis_ele};
    float gSw[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const float gij = gSc[i * 3 + j];
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) gSw[r * 3 + c] += gij * ax3[i][r] * ax3[j][c];
        }
    // Sw = M R with M[i][k] = R[i][k] s_k^2
    float gR[9];
    float gs2[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 9; ++k) gR[k] = 0.f;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            // gM[i][k] = sum_j gSw[i][j] R[k][j]
            const float gM = gSw[i * 3 + 0] * I.R[k * 3 + 0] + gSw[i * 3 + 1] * I.R[k * 3 + 1] + gSw[i * 3 + 2] * I.R[k * 3 + 2];
            const float s2 = I.s[k] * I.s[k];
            gR[i * 3 + k] += gM * s2;
            gs2[k] += gM * I.R[i * 3 + k];
        }
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            // gR[k][j] += sum_i M[i][k] gSw[i][j]
            float sum = 0.f;
#pragma unroll
            for (int i = 0; i < 3; ++i) sum += I.R[i * 3 + k] * I.s[k] * I.s[k] * gSw[i * 3 + j];
            gR[k * 3 + j] += sum;
        }
#pragma unroll
    for (int k = 0; k < 3; ++k) acc[3 + k] += I.s_act[k] ? 2.f * I.s[k] * gs2[k] : 0.f;
    // rotation entries -> normalised quaternion
    const float w = I.qn[0], x = I.qn[1], y = I.qn[2], z = I.qn[3];
    const float gw = -2.f * z * gR[1] + 2.f * y * gR[2] + 2.f * z * gR[3] - 2.f * x * gR[5] - 2.f * y * gR[6] + 2.f * x * gR[7];
    const float gx = 2.f * y * gR[1] + 2.f * z * gR[2] + 2.f * y * gR[3] - 4.f * x * gR[4] - 2.f * w * gR[5] +
                     2.f * z * gR[6] + 2.f * w * gR[7] - 4.f * x * gR[8];
    const float gy = -4.f * y * gR[0] + 2.f * x * gR[1] + 2.f * w * gR[2] + 2.f * x * gR[3] + 2.f * z * gR[5] -
                     2.f * w * gR[6] + 2.f * z * gR[7] - 4.f * y * gR[8];
    const float gz = -4.f * z * gR[0] - 2.f * w * gR[1] + 2.f * x * gR[2] + 2.f * w * gR[3] - 4.f * z * gR[4] +
                     2.f * y * gR[5] + 2.f * x * gR[6] + 2.f * y * gR[7];
    if (I.qq > 1e-12f) {
        const float dotq = w * gw + x * gx + y * gy + z * gz;
        acc[6] += I.nq * (gw - w * dotq);
        acc[7] += I.nq * (gx - x * dotq);
        acc[8] += I.nq * (gy - y * dotq);
        acc[9] += I.nq * (gz - z * dotq);
    } else {
        acc[6] += I.nq * gw;
        acc[7] += I.nq * gx;
        acc[8] += I.nq * gy;
        acc[9] += I.nq * gz;
    }
    // mu_c = A^T (mu - origin)
#pragma unroll
    for (int r = 0; r < 3; ++r) acc[r] += g_muc[0] * ax3[0][r] + g_muc[1] * ax3[1][r] + g_muc[2] * ax3[2][r];
}

// Per (frame, Gaussian) pair, each parameter gradient is sanitised before the pairs are summed, as the reference
// sanitises the gradients of every render call (one frame) per Gaussian (us_focus_sanitize_grad, us_focus.cu:129-137,
// applied at :694-752, with kGradClamp = 1e3 of plane_cond2d_device.cuh:12): a non-finite value becomes 0, a
// magnitude above 1e3 is clamped to +-1e3. The raw-activation chain rules come after, as torch autograd applies them
// to the reference's returned gradients.
constexpr float kGpGradClamp = 1e3f;

__device__ __forceinline__ float gp_sanitize_grad(float g) {
    if (!isfinite(g)) return 0.f;
    return fminf(fmaxf(g, -kGpGradClamp), kGpGradClamp);
}

// kGpReduceLanes lanes per Gaussian (four Gaussians per wavefront: a Gaussian of a freehand sweep has about ten
// visible pairs, so a wavefront per Gaussian left most lanes idle in the footprint's vector-Jacobian product). The
// lanes take its visible pairs, v in [rank(p0), rank(p0 + f_count)) in frame order, kGpReduceLanes at a time; each
// sums the pair's instance partials in emit order, applies the footprint's vector-Jacobian product and sanitises the
// pair's gradients; a fixed butterfly over the Gaussian's lanes sums them.
constexpr int kGpReduceLanes = 16;

__global__ __launch_bounds__(256, 4) void gp_reduce_kernel(int n_gauss, const us3d_sweep* __restrict__ sweeps,
                                                        int n_sweeps, const us3d_frame* __restrict__ frames,
                                                        us3d_gp_settings st, const uint64_t* __restrict__ pair_bits,
                                                        const int32_t* __restrict__ word_off,
                                                        const int32_t* __restrict__ vis_pair,
                                                        const int32_t* __restrict__ pair_count,
                                                        const int32_t* __restrict__ pair_off,
                                                        const float4* __restrict__ partial,
                                                        const float4* __restrict__ rec,
                                                        const us3d_stats* __restrict__ stats,
                                                        float* __restrict__ d_means, float* __restrict__ d_scales,
                                                        float* __restrict__ d_quats, float* __restrict__ d_opacity,
                                                        float* __restrict__ d_intensity) {
    const int g = (blockIdx.x * blockDim.x + threadIdx.x) / kGpReduceLanes;
    const int lane = threadIdx.x & (kGpReduceLanes - 1);
    if (g >= n_gauss || stats->overflow) return;  // uniform over a Gaussian's lanes
    const us3d_sweep sw = sweep_for_reduce(sweeps, n_sweeps, g);
    const int64_t p0 = sw.pair_base + (int64_t)(g - sw.g_begin) * sw.f_count;
    const int v0 = pair_rank(pair_bits, word_off, p0), v1 = pair_rank(pair_bits, word_off, p0 + sw.f_count);
    const float4 r0 = rec[3 * (int64_t)g + 0], r1 = rec[3 * (int64_t)g + 1], r2 = rec[3 * (int64_t)g + 2];
    float acc[11];
#pragma unroll
    for (int k = 0; k < 11; ++k) acc[k] = 0.f;
    float accI = 0.f;
    for (int v = v0 + lane; v < v1; v += kGpReduceLanes) {
        const int cnt = pair_count[v];
        const int e0 = pair_off[v];
        float gb = 0.f, gI = 0.f, gm0 = 0.f, gm1 = 0.f, g00 = 0.f, g01 = 0.f, g11 = 0.f;
#pragma unroll 4
        for (int e = e0; e < e0 + cnt; ++e) {
            const float4 a = partial[2 * (int64_t)e + 0], b = partial[2 * (int64_t)e + 1];
            gb += a.x; gI += a.y; gm0 += a.z; gm1 += a.w;
            g00 += b.x; g01 += b.y; g11 += b.z;
        }
        accI += gp_sanitize_grad(gI);
        const us3d_frame& fr = frames[sw.f_begin + (int)(vis_pair[v] - p0)];
        GPFoot F;
        GPInter I;
        gp_footprint<true>(r0, r1, r2, fr, st.k_sigma, st.sigma_ele, st.det_eps, F, &I);
        float pacc[11];
#pragma unroll
        for (int k = 0; k < 11; ++k) pacc[k] = 0.f;
        gp_footprint_vjp(I, fr, gb, gm0, gm1, g00, g01, g11, pacc);
#pragma unroll
        for (int k = 0; k < 11; ++k) acc[k] += gp_sanitize_grad(pacc[k]);
    }
    for (int off = kGpReduceLanes / 2; off >= 1; off >>= 1) {  // inside the Gaussian's lanes
#pragma unroll
        for (int k = 0; k < 11; ++k) acc[k] += __shfl_xor(acc[k], off, kWave);
        accI += __shfl_xor(accI, off, kWave);
    }
    if (lane == 0) {
        d_means[3 * (int64_t)g + 0] = acc[0];
        d_means[3 * (int64_t)g + 1] = acc[1];
        d_means[3 * (int64_t)g + 2] = acc[2];
        if (st.raw_activations) {
            // chain of max(exp(x), 1e-5): d x = d s * s where the clamp is inactive; of the sigmoid: o (1 - o)
            const float sv[3] = {r1.x, r1.y, r1.z};
#pragma unroll
            for (int k = 0; k < 3; ++k) acc[3 + k] = sv[k] > 1e-5f ? acc[3 + k] * sv[k] : 0.f;
            acc[10] = acc[10] * (r0.w * (1.f - r0.w));
        }
        d_scales[3 * (int64_t)g + 0] = acc[3];
        d_scales[3 * (int64_t)g + 1] = acc[4];
        d_scales[3 * (int64_t)g + 2] = acc[5];
        d_quats[4 * (int64_t)g + 0] = acc[6];
        d_quats[4 * (int64_t)g + 1] = acc[7];
        d_quats[4 * (int64_t)g + 2] = acc[8];
        d_quats[4 * (int64_t)g + 3] = acc[9];
        d_opacity[g] = acc[10];
        if (d_intensity) d_intensity[g] = accI;
    }
}

static int gp_check_settings(const us3d_gp_settings* s) {
    if (!s) return fail(US3D_ERR_ARGUMENT, "settings is NULL");
    if (!(s->tile_size == 8 || s->tile_size == 16 || s->tile_size == 32))
        return fail(US3D_ERR_ARGUMENT, "tile_size must be 8, 16 or 32");
    if (!(s->k_sigma > 0.0f)) return fail(US3D_ERR_ARGUMENT, "k_sigma must be > 0");
    if (!(s->sigma_ele > 0.0f)) return fail(US3D_ERR_ARGUMENT, "sigma_ele must be > 0");
    if (!(s->det_eps > 0.0f)) return fail(US3D_ERR_ARGUMENT, "det_eps must be > 0");
    // the reference floors it at 1e-12 (max(threshold, 1e-12)); at 0 a listed pair whose base weight underflowed
    // to 0 would make the adjoint divide 0 by 0
    if (!(s->weight_min > 0.0f)) return fail(US3D_ERR_ARGUMENT, "weight_min must be > 0");
    if (!(s->work_queue == 0 || s->work_queue == 1)) return fail(US3D_ERR_ARGUMENT, "work_queue must be 0 or 1");
    return US3D_OK;
}

}  // namespace us3d

using namespace us3d;

extern "C" {

int64_t us3d_gp_workspace_bytes(int64_t n_gauss, int64_t n_pairs, int64_t n_tiles, int64_t k_capacity,
                                int64_t v_capacity) {
    return gp_layout(n_gauss, n_pairs, n_tiles, k_capacity, v_capacity).total;
}

int us3d_gp_workspace_layout(int64_t n_gauss, int64_t n_pairs, int64_t n_tiles, int64_t k_capacity,
                             int64_t v_capacity, us3d_layout* out) {
    if (!out) return fail(US3D_ERR_ARGUMENT, "layout output is NULL");
    *out = gp_layout(n_gauss, n_pairs, n_tiles, k_capacity, v_capacity);
    return US3D_OK;
}

int us3d_gp_map(const us3d_plan* plan, const float* means, const float* scales, const float* quats,
                const float* opacity, const float* intensity, const us3d_gp_settings* settings, void* stream_) {
    const us3d_layout L = gp_layout(plan);
    int rc = check_plan(plan, L.total);
    if (rc) return rc;
    if ((rc = gp_check_settings(settings))) return rc;
    if (plan->n_gauss > 0 && (!means || !scales || !quats || !opacity))
        return fail(US3D_ERR_ARGUMENT, "Gaussian parameter pointer is NULL");
    hipStream_t stream = as_stream(stream_);
    float4* rec = gp_at<float4>(plan, L.records);
    if ((rc = begin_pairs(plan, L, stream))) return rc;
    if (plan->n_gauss > 0) {
        hipLaunchKernelGGL(gp_pack_kernel, dim3(blocks_for(plan->n_gauss, 256)), dim3(256), 0, stream, plan->n_gauss,
                           means, scales, quats, opacity, intensity, settings->raw_activations, rec);
        if ((rc = check_launch("gp_pack_kernel"))) return rc;
    }
    const int64_t np = plan->n_pairs;
    if (np > 0) {
        {
            StageTimer timer_("gp_cull_kernel", stream);
            hipLaunchKernelGGL(gp_cull_kernel, dim3(blocks_for(np, 256)), dim3(256), 0, stream, np, rec, plan->frames,
                               plan->sweeps, plan->n_sweeps, *settings, gp_at<uint64_t>(plan, L.pair_bits),
                               gp_at<int32_t>(plan, L.word_cnt), gp_at<uint32_t>(plan, L.fm_bits));
        }
        if ((rc = check_launch("gp_cull_kernel"))) return rc;
        if ((rc = rank_visible_pairs(plan, L, stream))) return rc;
        {
            StageTimer timer_("gp_setup_kernel", stream);
            hipLaunchKernelGGL(gp_setup_kernel, dim3(blocks_for(plan->v_capacity, 256)), dim3(256), 0, stream,
                               plan->v_capacity, gp_at<int32_t>(plan, L.vis_pair), rec, plan->frames, plan->sweeps,
                               plan->n_sweeps, *settings, gp_at<us3d_stats>(plan, L.stats),
                               gp_at<int32_t>(plan, L.pair_count), gp_at<int4>(plan, L.pair_rect),
                               gp_at<uint32_t>(plan, L.pair_dk), gp_at<float4>(plan, L.pair_sect),
                               gp_at<int32_t>(plan, L.tile_count));
        }
        if ((rc = check_launch("gp_setup_kernel"))) return rc;
    }
    return build_instances(plan, L, stream, kFootF4, 2);
}

}  // extern "C"

namespace {

int gp_forward_impl(const us3d_plan* plan, const us3d_gp_settings* settings, float* image, const LossFuse& loss,
                    void* stream_) {
    const us3d_layout L = gp_layout(plan);
    int rc = check_plan(plan, L.total);
    if (rc) return rc;
    if ((rc = gp_check_settings(settings))) return rc;
    if (!image) return fail(US3D_ERR_ARGUMENT, "forward: image is NULL");
    if (plan->n_tiles <= 0) return US3D_OK;
    hipStream_t stream = as_stream(stream_);
    int32_t* heads = settings->work_queue ? gp_at<int32_t>(plan, L.queue) : nullptr;
#define GP_FWD(TS)                                                                                                    \
    hipLaunchKernelGGL((gp_forward_kernel<TS>),                                                                       \
                       dim3(tile_grid((const void*)gp_forward_kernel<TS>, GpFwd<TS>::kThreads, plan, heads, stream)), \
                       dim3(GpFwd<TS>::kThreads), 0, stream,                                                          \
                       gp_at<float4>(plan, L.records), gp_at<float4>(plan, L.pair_sect),                             \
                       gp_at<int32_t>(plan, L.tile_off), gp_at<int32_t>(plan, L.pair_g),                             \
                       gp_at<uint64_t>(plan, L.seg_key), plan->frames, plan->frame_tile_base, plan->n_frames,        \
                       settings->weight_min,                                                                          \
                       gp_at<us3d_stats>(plan, L.stats), image, plan->n_tiles, heads, loss)
    {
        StageTimer timer_("gp_forward_kernel", stream);
        switch (settings->tile_size) {
            case 8: GP_FWD(8); break;
            case 16: GP_FWD(16); break;
            default: GP_FWD(32); break;
        }
    }
#undef GP_FWD
    return check_launch("gp_forward_kernel");
}

int gp_backward_impl(const us3d_plan* plan, const us3d_gp_settings* settings, const float* d_image,
                     float* d_means, float* d_scales, float* d_quats, float* d_opacity, float* d_intensity,
                     const float* image, const LossFuse& loss, void* stream_) {
    const us3d_layout L = gp_layout(plan);
    int rc = check_plan(plan, L.total);
    if (rc) return rc;
    if ((rc = gp_check_settings(settings))) return rc;
    if (!(d_image || (loss.coeff && image && loss.gt)) || !d_means || !d_scales || !d_quats || !d_opacity)
        return fail(US3D_ERR_ARGUMENT, "backward: NULL array");
    hipStream_t stream = as_stream(stream_);
    if (plan->n_tiles > 0) {
        int32_t* heads = settings->work_queue ? gp_at<int32_t>(plan, L.queue) : nullptr;
#define GP_BWD(TS)                                                                                                    \
    hipLaunchKernelGGL((gp_backward_kernel<TS>),                                                                      \
                       dim3(tile_grid((const void*)gp_backward_kernel<TS>, TS * TS, plan, heads, stream)),            \
                       dim3(TS * TS), 0, stream,                                                                      \
                       gp_at<float4>(plan, L.records), gp_at<float4>(plan, L.pair_sect),                             \
                       gp_at<int32_t>(plan, L.tile_off), gp_at<int32_t>(plan, L.pair_g),                             \
                       gp_at<uint64_t>(plan, L.seg_key), plan->frames, plan->frame_tile_base, plan->n_frames,        \
                       settings->weight_min, d_image, gp_at<us3d_stats>(plan, L.stats),                               \
                       gp_at<float4>(plan, L.partial), plan->n_tiles, heads, image, loss)
        {
            StageTimer timer_("gp_backward_kernel", stream);
            switch (settings->tile_size) {
                case 8: GP_BWD(8); break;
                case 16: GP_BWD(16); break;
                default: GP_BWD(32); break;
            }
        }
#undef GP_BWD
        if ((rc = check_launch("gp_backward_kernel"))) return rc;
    }
    if (plan->n_gauss > 0) {
        {
            StageTimer timer_("gp_reduce_kernel", stream);
            hipLaunchKernelGGL(gp_reduce_kernel, dim3(blocks_for(plan->n_gauss, 256 / kGpReduceLanes)), dim3(256), 0,
                               stream, plan->n_gauss, plan->sweeps, plan->n_sweeps, plan->frames, *settings,
                               gp_at<uint64_t>(plan, L.pair_bits), gp_at<int32_t>(plan, L.word_off),
                               gp_at<int32_t>(plan, L.vis_pair), gp_at<int32_t>(plan, L.pair_count),
                               gp_at<int32_t>(plan, L.pair_off),
                               gp_at<float4>(plan, L.partial), gp_at<float4>(plan, L.records),
                               gp_at<us3d_stats>(plan, L.stats), d_means, d_scales, d_quats, d_opacity, d_intensity);
        }
        if ((rc = check_launch("gp_reduce_kernel"))) return rc;
    }
    return US3D_OK;
}

// Per-(tile, wavefront) loss sums of a fused forward: up to four wavefronts per tile (32-pixel tiles).
int64_t gp_loss_part_entries(const us3d_plan* plan) { return (int64_t)plan->n_tiles * (GpFwd<32>::kThreads / kWave); }

}  // namespace

extern "C" {

int us3d_gp_forward(const us3d_plan* plan, const us3d_gp_settings* settings, float* image, void* stream_) {
    return gp_forward_impl(plan, settings, image, LossFuse{nullptr, nullptr, nullptr}, stream_);
}

int us3d_gp_backward(const us3d_plan* plan, const us3d_gp_settings* settings, const float* d_image,
                     float* d_means, float* d_scales, float* d_quats, float* d_opacity, float* d_intensity,
                     void* stream_) {
    if (!d_image) return fail(US3D_ERR_ARGUMENT, "backward: NULL array");
    return gp_backward_impl(plan, settings, d_image, d_means, d_scales, d_quats, d_opacity, d_intensity, nullptr,
                            LossFuse{nullptr, nullptr, nullptr}, stream_);
}

int64_t us3d_gp_loss_tmp_bytes(const us3d_plan* plan) {
    if (!plan) return 0;
    return (gp_loss_part_entries(plan) * kLossSums + 4 * (int64_t)plan->n_frames) * (int64_t)sizeof(double);
}

int us3d_gp_forward_loss(const us3d_plan* plan, const us3d_gp_settings* settings, float* image, const float* gt,
                         float w_l1, float w_ssim, float scale, float* loss, double* tmp, void* stream_) {
    if (!plan || !settings || !gt || !loss || !tmp) return fail(US3D_ERR_ARGUMENT, "forward_loss: NULL argument");
    int rc = gp_forward_impl(plan, settings, image, LossFuse{gt, tmp, nullptr}, stream_);
    if (rc) return rc;
    const int waves = settings->tile_size == 32 ? GpFwd<32>::kThreads / kWave
                                                : (settings->tile_size == 16 ? GpFwd<16>::kThreads / kWave
                                                                             : GpFwd<8>::kThreads / kWave);
    return image_loss_from_tiles(tmp, plan->frame_tile_base, plan->frame_pixel_base, plan->n_frames, waves, w_l1,
                                 w_ssim, scale, loss, tmp + gp_loss_part_entries(plan) * kLossSums,
                                 as_stream(stream_));
}

int us3d_gp_backward_loss(const us3d_plan* plan, const us3d_gp_settings* settings, const float* image,
                          const float* gt, const double* tmp, float* d_means, float* d_scales, float* d_quats,
                          float* d_opacity, float* d_intensity, void* stream_) {
    if (!plan || !image || !gt || !tmp) return fail(US3D_ERR_ARGUMENT, "backward_loss: NULL argument");
    return gp_backward_impl(plan, settings, nullptr, d_means, d_scales, d_quats, d_opacity, d_intensity, image,
                            LossFuse{gt, nullptr, tmp + gp_loss_part_entries(plan) * kLossSums}, stream_);
}

}  // extern "C"
