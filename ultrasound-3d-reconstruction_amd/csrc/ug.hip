// UltraGauss slice splat on gfx950: mapper, per-tile forward splat, per-tile adjoint, per-Gaussian
// gradient reduction and the per-pixel candidate lists.
//
// Data layout in HBM (one batch = frames of one or more sweeps):
//   records   f32x12 per Gaussian: (mu.xyz, opacity | intensity, l11, l12, l13 | l22, l23, l33, r_max),
//             l_ii = raw_ii^2 + beta, packed once per call so every gather is three aligned 16 B loads.
//   pairs     (frame, Gaussian), numbered Gaussian-major per sweep; the visible ones (a few percent) ranked
//             into a compact index v with count, tile rect, depth key and section record (mapper.hip).
//   instances (tile, Gaussian) in emit order e (pair order, then tile rows, then columns) -- the
//             reference's emit order for one frame -- bucketed by global tile, each bucket sorted by
//             the 64-bit key (depth key << 32 | Gaussian), value = v (e follows from v and the tile).
//   partial   f32x12 per instance, written at e by the adjoint, so each Gaussian's partials across all
//             frames are contiguous and reduce in a fixed order.
#include <algorithm>
#include <mutex>
#include <type_traits>

#include "common.h"

namespace us3d {

// ---------------------------------------------------------------- workspace layout
// Per visible pair, kSectF4 float4s in pair_sect: the pair's part of the sub-block test (s0), the plane form of q
// (s1, s2; pair_section) with the pair's Gaussian in s2.w, and its emit base (s3: first emit index, rectangle x0, y0
// and width, written once the counts are scanned), so the forward (s0-s2) and the adjoint (s1-s3) each read one
// 64 B record, which also names the Gaussian record to gather.
constexpr int kSectF4 = 4;

static us3d_layout ug_layout(int64_t n_gauss, int64_t n_pairs, int64_t n_tiles, int64_t k_cap, int64_t v_cap) {
    return ws_layout(n_gauss, n_pairs, n_tiles, k_cap, v_cap, 48, 48, 2, 16 * kSectF4);
}

static us3d_layout ug_layout(const us3d_plan* plan) {
    return plan ? ug_layout(plan->n_gauss, plan->n_pairs, plan->n_tiles, plan->k_capacity, plan->v_capacity)
                : us3d_layout{};
}

// ---------------------------------------------------------------- mapper kernels

// Packs one record per Gaussian. The spare last word holds r_max = sqrt(chi2) ||M^-1||_F * 1.01 (M = the upper
// triangular factor with q = |M d|^2): aT Sigma a = |M^-T a|^2 <= ||M^-1||_F^2 for any unit a, so r_max bounds
// every marginal radius and a pair with |elevation| > r_max is culled without computing them (NaN for a
// degenerate factor, which then takes the exact path). With `logits`, opacity and intensity are the training
// model's logits and get its sigmoid here (pipeline_god_file.py:369-373), so no separate activation pass runs.
__device__ __forceinline__ float sigmoidf(float x) { return 1.0f / (1.0f + expf(-x)); }

__global__ void ug_pack_kernel(int n, const float* __restrict__ means, const float* __restrict__ tril,
                               const float* __restrict__ opacity, const float* __restrict__ intensity, float beta,
                               float sqrt_chi2, int logits, float4* __restrict__ rec) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const float* t = tril + (int64_t)g * 6;
    const float* m = means + (int64_t)g * 3;
    const float l11 = __builtin_fmaf(t[0], t[0], beta), l12 = t[1], l13 = t[2];
    const float l22 = __builtin_fmaf(t[3], t[3], beta), l23 = t[4], l33 = __builtin_fmaf(t[5], t[5], beta);
    const double i11 = 1.0 / l11, i22 = 1.0 / l22, i33 = 1.0 / l33;
    const double i12 = -(double)l12 * i11 * i22, i23 = -(double)l23 * i22 * i33;
    const double i13 = ((double)l12 * l23 - (double)l13 * l22) * i11 * i22 * i33;
    const double fro2 = i11 * i11 + i22 * i22 + i33 * i33 + i12 * i12 + i23 * i23 + i13 * i13;
    const float r_max = (float)((double)sqrt_chi2 * sqrt(fro2) * 1.01);
    const float o = logits ? sigmoidf(opacity[g]) : opacity[g];
    const float c = logits ? sigmoidf(intensity[g]) : intensity[g];
    rec[3 * (int64_t)g + 0] = make_float4(m[0], m[1], m[2], o);
    rec[3 * (int64_t)g + 1] = make_float4(c, l11, l12, l13);
    rec[3 * (int64_t)g + 2] = make_float4(l22, l23, l33, r_max);
}

__device__ __forceinline__ int clamp_to_int(float x) {
    return (int)fminf(fmaxf(x, -1073741824.0f), 1073741824.0f);
}

// Depth key (uint32)((v_pix + 500) * 1000) of core.cu:134, with the out-of-range cases defined:
// negative or NaN -> 0, >= 2^32 -> 0xffffffff.
__device__ __forceinline__ uint32_t depth_key(float v_pix) {
    const float x = (v_pix + 500.0f) * 1000.0f;
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)x;
}

// Tile rectangle [fx0, fx1) x [fy0, fy1) (unclipped, as floats) and radial row coordinate v of a Gaussian in a
// fan frame, from the polar extent of its in-plane marginal box [lat +- r_lat] x [ax +- r_ax] about the apex:
// radius from the box's distance to the apex to its farthest corner; angle t = atan2(lateral, axial) from the
// box's extreme corners (a convex region clear of the apex and of the half-line behind it has its angular
// extremes at vertices), or every angle when the box holds the apex or crosses that half-line. Any pixel
// within the box lies in the rectangle; one pixel of margin on each side covers the rounding of the grid.
__device__ __forceinline__ void fan_tile_bounds(const us3d_frame& fr, float lat, float ax, float r_lat, float r_ax,
                                                float tsf, float& fx0, float& fy0, float& fx1, float& fy1,
                                                float& v) {
    const float L0 = lat - r_lat, L1 = lat + r_lat, A0 = ax - r_ax, A1 = ax + r_ax;
    const float dl = L0 > 0.0f ? L0 : (L1 < 0.0f ? -L1 : 0.0f);
    const float da = A0 > 0.0f ? A0 : (A1 < 0.0f ? -A1 : 0.0f);
    const float rmin = sqrtf(dl * dl + da * da);
    const float rmax = sqrtf(fmaxf(L0 * L0, L1 * L1) + fmaxf(A0 * A0, A1 * A1));
    float t0 = -3.14159274f, t1 = 3.14159274f;
    const bool apex = dl == 0.0f && da == 0.0f;
    const bool behind = A0 < 0.0f && L0 <= 0.0f && L1 >= 0.0f;
    if (!apex && !behind) {
        const float c0 = atan2f(L0, A0), c1 = atan2f(L0, A1), c2 = atan2f(L1, A0), c3 = atan2f(L1, A1);
        t0 = fminf(fminf(c0, c1), fminf(c2, c3));
        t1 = fmaxf(fmaxf(c0, c1), fmaxf(c2, c3));
    }
    const float u0 = t0 / fr.delta_lat + fr.col_offset, u1 = t1 / fr.delta_lat + fr.col_offset;
    const float v0 = (rmin - fr.near_mm) / fr.delta_ax, v1 = (rmax - fr.near_mm) / fr.delta_ax;
    fx0 = floorf((u0 - 1.0f) / tsf);
    fx1 = ceilf((u1 + 1.0f) / tsf);
    fy0 = floorf((v0 - 1.0f) / tsf);
    fy1 = ceilf((v1 + 1.0f) / tsf);
    v = (sqrtf(lat * lat + ax * ax) - fr.near_mm) / fr.delta_ax;
}

// Per (frame, Gaussian) pair: probe-frame projection, chi2 marginal radii, elevation cull, tile rectangle
// and depth key (compute_probe_bbox_radii + kernel_project_gauss_and_compute_tile_rects,
// ops/ultragauss_renderer/src/mapper/core.cu:9-116). Non-finite projections are culled.
// Per-pair part of the sub-block test (the tests themselves are described with the splat kernels): computed
// once per visible (frame, Gaussian) pair by the preprocess, read by the forward for each of the pair's
// instances. With orthonormal probe axes and a regular section (kk >= 0):
//   s0 = (u_lat, u_ax, kk, t_x): in-plane centre drift per unit elevation, elevational Schur complement, and
//        t_x = M01 / M11;
//   s1 = (p, q, s, t_y): the in-plane precision M = [[M00 M01] [M01 M11]] as f(x, y) = (p x + q y)^2 + (s y)^2
//        (p = sqrt(M00), q = M01 / p, s = sqrt(det M / M00)), and t_y = M01 / M00.
// Otherwise kk = -1 and s0.xy = (wl, wa), the padded marginal half-widths (infinite for a degenerate factor,
// some l_ii <= 1e-10, so every sub-block). The section setup is in FP64 (no cancellation in det or the Schur
// complement).
__device__ __forceinline__ bool axes_orthonormal(const us3d_frame& fr) {
    const float* a = fr.axis_lat;
    const float* b = fr.axis_ax;
    const float* c = fr.axis_ele;
    const float ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    const float ac = a[0] * c[0] + a[1] * c[1] + a[2] * c[2];
    const float bc = b[0] * c[0] + b[1] * c[1] + b[2] * c[2];
    const float aa = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
    const float bb = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
    const float cc = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
    return fabsf(ab) < 1e-5f && fabsf(ac) < 1e-5f && fabsf(bc) < 1e-5f && fabsf(aa - 1.f) < 2e-5f &&
           fabsf(bb - 1.f) < 2e-5f && fabsf(cc - 1.f) < 2e-5f;
}

__device__ __forceinline__ float sub_pad(float lat_g, float ax_g) {
    return 1e-3f + 4e-6f * (fabsf(lat_g) + fabsf(ax_g));
}

// Plane form of q for the linear probe's pixel grid (every frame geometry with an affine grid, whatever the axes):
// the pixel at probe coordinates (l, a) sits at o + l a_lat + a a_ax, so with (l, a) = (lat_g + x, ax_g + y),
// L^T (p - mu) = c0 + x A + y B with A = L^T a_lat, B = L^T a_ax, c0 = L^T g0, g0 = (o - mu) + lat_g a_lat +
// ax_g a_ax, and q = |c0 + x A + y B|^2, a 2-D quadratic. Written about its minimum (x*, y*) with value K:
//   q = (p (l - cx) + qq (a - cy))^2 + (s (a - cy))^2 + K,   cx = lat_g + x*, cy = ax_g + y*,
// p = sqrt(M00), qq = M01 / p, s = sqrt(det M / M00), M = [[A.A, A.B], [A.B, B.B]] -- a sum of squares, so the
// evaluation per pixel is two subtractions and five multiply-adds without cancellation. Set up in FP64 per
// visible pair (the preprocess), evaluated in FP32 per (pixel, instance) by the splat kernels. With det M <= 0
// (only for parallel axes or a singular factor) the minimum is not unique: x* = y* = 0, K = |c0|^2, s = 0.
// Returns s1 = (p, qq, s, t_y = M01 / M00) and s2 = (cx, cy, K, 0); the setup pass stores the Gaussian in s2.w.
__device__ __forceinline__ void pair_plane_form(double L11, double L12, double L13, double L22, double L23,
                                                double L33, float mx, float my, float mz, float lat_g, float ax_g,
                                                const us3d_frame& fr, double& M00, double& M01, double& M11,
                                                float4& s1, float4& s2) {
    const float* a = fr.axis_lat;
    const float* c = fr.axis_ax;
    const double A0 = a[0] * L11 + a[1] * L12 + a[2] * L13, A1 = a[1] * L22 + a[2] * L23, A2 = a[2] * L33;
    const double B0 = c[0] * L11 + c[1] * L12 + c[2] * L13, B1 = c[1] * L22 + c[2] * L23, B2 = c[2] * L33;
    M00 = A0 * A0 + A1 * A1 + A2 * A2;
    M01 = A0 * B0 + A1 * B1 + A2 * B2;
    M11 = B0 * B0 + B1 * B1 + B2 * B2;
    const double lg = lat_g, ag = ax_g;
    const double g0 = ((double)fr.origin[0] - (double)mx) + lg * a[0] + ag * c[0];
    const double g1 = ((double)fr.origin[1] - (double)my) + lg * a[1] + ag * c[1];
    const double g2 = ((double)fr.origin[2] - (double)mz) + lg * a[2] + ag * c[2];
    const double c0 = L11 * g0 + L12 * g1 + L13 * g2, c1 = L22 * g1 + L23 * g2, c2 = L33 * g2;
    const double h0 = A0 * c0 + A1 * c1 + A2 * c2, h1 = B0 * c0 + B1 * c1 + B2 * c2;
    const double cc = c0 * c0 + c1 * c1 + c2 * c2;
    const double det = M00 * M11 - M01 * M01;
    double xs = 0.0, ys = 0.0, K = cc, pp = sqrt(fmax(M00, 0.0)), qq = 0.0, ss = 0.0;
    if (M00 > 0.0 && det > 0.0) {
        xs = -(M11 * h0 - M01 * h1) / det;
        ys = -(M00 * h1 - M01 * h0) / det;
        K = fmax(0.0, cc + h0 * xs + h1 * ys);
        qq = M01 / pp;
        ss = sqrt(det / M00);
    } else if (M00 > 0.0) {
        qq = M01 / pp;
    }
    s1 = make_float4((float)pp, (float)qq, (float)ss, M00 > 0.0 ? (float)(M01 / M00) : 0.f);
    s2 = make_float4((float)(lg + xs), (float)(ag + ys), (float)K, 0.f);
}

__device__ __forceinline__ void pair_section(float l11, float l12, float l13, float l22, float l23, float l33,
                                             float mx, float my, float mz, float lat_g, float ax_g, float r_lat,
                                             float r_ax, const us3d_frame& fr, float4& s0, float4& s1, float4& s2) {
    const double L11 = l11, L12 = l12, L13 = l13, L22 = l22, L23 = l23, L33 = l33;
    double M00, M01, M11;
    pair_plane_form(L11, L12, L13, L22, L23, L33, mx, my, mz, lat_g, ax_g, fr, M00, M01, M11, s1, s2);
    if (!(l11 > 1e-10f && l22 > 1e-10f && l33 > 1e-10f)) {
        s0 = make_float4(INFINITY, INFINITY, -1.f, 0.f);
        return;
    }
    const float pad = sub_pad(lat_g, ax_g);
    s0 = make_float4(r_lat * 1.001f + pad, r_ax * 1.001f + pad, -1.f, 0.f);
    if (!axes_orthonormal(fr)) return;
    // rows of A = R^T L, L = [[l11 0 0] [l12 l22 0] [l13 l23 l33]] so that q = |L^T d|^2
    const float* a = fr.axis_lat;
    const float* c = fr.axis_ax;
    const float* e = fr.axis_ele;
    const double A0 = a[0] * L11 + a[1] * L12 + a[2] * L13, A1 = a[1] * L22 + a[2] * L23, A2 = a[2] * L33;
    const double B0 = c[0] * L11 + c[1] * L12 + c[2] * L13, B1 = c[1] * L22 + c[2] * L23, B2 = c[2] * L33;
    const double C0 = e[0] * L11 + e[1] * L12 + e[2] * L13, C1 = e[1] * L22 + e[2] * L23, C2 = e[2] * L33;
    const double b0 = A0 * C0 + A1 * C1 + A2 * C2;
    const double b1 = B0 * C0 + B1 * C1 + B2 * C2;
    const double cz = C0 * C0 + C1 * C1 + C2 * C2;
    const double det = M00 * M11 - M01 * M01;
    if (!(det > 1e-8 * M00 * M11)) return;
    const double w0 = (M11 * b0 - M01 * b1) / det, w1 = (M00 * b1 - M01 * b0) / det;
    const double kk = fmax(0.0, (cz - (b0 * w0 + b1 * w1)) * (1.0 - 1e-5));
    s0 = make_float4((float)(-w0), (float)(-w1), (float)kk, (float)(M01 / M11));
}

// Tiles of one (frame, Gaussian) pair (core.cu:9-116 restated): a pair farther than r_max from the plane is culled
// before its radii are computed (most pairs of a freehand sweep); otherwise the chi2 marginal radii, the elevation
// cull, the tile rectangle (linear grid, or the polar extent of a fan frame) and the depth key. Returns the instance
// count (0: culled); the cull and the setup pass call it on the same inputs, so they agree on every pair.
struct UGPairTiles {
    int4 rect;
    uint32_t dk;
    float lat, ax, r_lat, r_ax;
};

__device__ __forceinline__ int ug_pair_tiles(const float4& r0, const float4& r1, const float4& r2, const us3d_frame& fr,
                                             float sqrt_chi2, int ts, int sort_by_depth, UGPairTiles& out) {
    const float l11 = r1.y, l12 = r1.z, l13 = r1.w, l22 = r2.x, l23 = r2.y, l33 = r2.z;
    const float vx = r0.x - fr.origin[0], vy = r0.y - fr.origin[1], vz = r0.z - fr.origin[2];
    const float ele = dot3(vx, vy, vz, fr.axis_ele[0], fr.axis_ele[1], fr.axis_ele[2]);
    // beyond every marginal radius (see ug_pack_kernel); NaN takes the exact path. Most pairs leave here, before the
    // frame's other axes are read.
    if (fabsf(ele) > r2.w) return 0;
    // the elevation cull first: a pair it removes never computes its in-plane radii
    const float r_ele = marginal_radius(l11, l12, l13, l22, l23, l33, fr.axis_ele[0], fr.axis_ele[1], fr.axis_ele[2], sqrt_chi2);
    if (!(fabsf(ele) <= r_ele)) return 0;
    const float lat = dot3(vx, vy, vz, fr.axis_lat[0], fr.axis_lat[1], fr.axis_lat[2]);
    const float ax = dot3(vx, vy, vz, fr.axis_ax[0], fr.axis_ax[1], fr.axis_ax[2]);
    const float r_lat = marginal_radius(l11, l12, l13, l22, l23, l33, fr.axis_lat[0], fr.axis_lat[1], fr.axis_lat[2], sqrt_chi2);
    const float r_ax = marginal_radius(l11, l12, l13, l22, l23, l33, fr.axis_ax[0], fr.axis_ax[1], fr.axis_ax[2], sqrt_chi2);
    const float tsf = (float)ts;
    float fx0, fy0, fx1, fy1, v;
    if (fr.geometry == 1) {
        fan_tile_bounds(fr, lat, ax, r_lat, r_ax, tsf, fx0, fy0, fx1, fy1, v);
    } else {
        const float u = lat / fr.delta_lat + fr.col_offset;
        v = (ax - fr.near_mm) / fr.delta_ax;
        const float ru = fmaxf(1.0f, r_lat / fr.delta_lat);
        const float rv = fmaxf(1.0f, r_ax / fr.delta_ax);
        fx0 = floorf((u - ru) / tsf);
        fy0 = floorf((v - rv) / tsf);
        fx1 = ceilf((u + ru) / tsf);
        fy1 = ceilf((v + rv) / tsf);
    }
    if (!(isfinite(fx0) && isfinite(fy0) && isfinite(fx1) && isfinite(fy1))) return 0;
    const int x0 = max(0, clamp_to_int(fx0)), y0 = max(0, clamp_to_int(fy0));
    const int x1 = min(fr.tiles_x, clamp_to_int(fx1)), y1 = min(fr.tiles_y, clamp_to_int(fy1));
    if (!(x0 < x1 && y0 < y1)) return 0;
    out.rect = make_int4(x0, y0, x1, y1);
    out.dk = sort_by_depth ? depth_key(v) : 0u;
    out.lat = lat;
    out.ax = ax;
    out.r_lat = r_lat;
    out.r_ax = r_ax;
    return (x1 - x0) * (y1 - y0);
}

// Cull pass: thread per dense (frame, Gaussian) pair, numbered Gaussian-major inside a sweep; marks the visible ones.
__global__ __launch_bounds__(256) void ug_cull_kernel(int64_t n_pairs, const float4* __restrict__ rec,
                                                      const us3d_frame* __restrict__ frames,
                                                      const us3d_sweep* __restrict__ sweeps, int n_sweeps,
                                                      float sqrt_chi2, int ts, uint64_t* __restrict__ pair_bits,
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
        UGPairTiles pt;
        count = ug_pair_tiles(rec[3 * (int64_t)g + 0], rec[3 * (int64_t)g + 1], rec[3 * (int64_t)g + 2], frames[f],
                              sqrt_chi2, ts, 0, pt);
    }
    mark_pair(count > 0, pair, n_pairs, sweeps, s, f, gl, pair_bits, word_cnt, fm_bits);
}

// Setup pass: thread per visible pair v < v_capacity (entries past the visible count, or every entry when the visible
// pairs overflowed, get count 0): count, tile rectangle, depth key, the pair's part of the sub-block test and the
// plane form of q (pair_section), and its tiles counted by rectangle corners.
__global__ __launch_bounds__(256) void ug_setup_kernel(
    int64_t v_cap, const int32_t* __restrict__ vis_pair, const float4* __restrict__ rec,
    const us3d_frame* __restrict__ frames, const us3d_sweep* __restrict__ sweeps, int n_sweeps, float sqrt_chi2,
    int ts, int sort_by_depth, const us3d_stats* __restrict__ stats, int32_t* __restrict__ pair_count,
    int4* __restrict__ pair_rect, uint32_t* __restrict__ pair_dk, float4* __restrict__ pair_sect,
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
    const float4 r0 = rec[3 * (int64_t)g + 0];
    const float4 r1 = rec[3 * (int64_t)g + 1];
    const float4 r2 = rec[3 * (int64_t)g + 2];
    UGPairTiles pt;
    const int count = ug_pair_tiles(r0, r1, r2, fr, sqrt_chi2, ts, sort_by_depth, pt);
    pair_count[v] = count;
    if (count > 0) {  // always, for a listed pair: the cull made the same decision
        pair_rect[v] = pt.rect;
        pair_dk[v] = pt.dk;
        float4 s0, s1, s2;
        pair_section(r1.y, r1.z, r1.w, r2.x, r2.y, r2.z, r0.x, r0.y, r0.z, pt.lat, pt.ax, pt.r_lat, pt.r_ax, fr, s0,
                     s1, s2);
        pair_sect[kSectF4 * v + 0] = s0;
        pair_sect[kSectF4 * v + 1] = s1;
        pair_sect[kSectF4 * v + 2] = make_float4(s2.x, s2.y, s2.z, __int_as_float(g));  // .w: the pair's Gaussian
        add_rect_corners(tile_count, fr.tile_base, fr.tiles_x, fr.tiles_y, pt.rect);
    }
}

// ---------------------------------------------------------------- splat geometry
// A tile of TS x TS pixels is split into 8x8 sub-blocks; one wavefront owns a sub-block at a time, so a
// test of a Gaussian against a sub-block is wavefront-uniform. Each Gaussian loaded into LDS carries a
// bit mask of the sub-blocks that can hold a pixel with q <= chi2. Two conservative tests, both widened for
// rounding, are ANDed:
//   marginal -- such a pixel satisfies |(p - mu).a| <= r_a along both in-plane axes (r_a is the ellipsoid's
//               marginal extent), valid for any pixel positions;
//   section  -- with orthonormal probe axes, write q in probe coordinates (lat, ax, ele) with precision
//               P = (R^T L)(R^T L)^T, R = [a_lat a_ax a_ele]. At elevation offset d (pixel minus mean) the set
//               q <= chi2 is the ellipse (x - u d)^T M (x - u d) <= chi2 - k d^2, M = P[:2,:2],
//               u = -M^-1 P[:2,2], k = P[2,2] - P[2,:2] M^-1 P[:2,2]; over the sub-block's elevation range
//               its bounding box is [u d_lo, u d_hi] widened by sqrt(rem M^-1_ii), rem = chi2 - k d*^2 at the
//               d* of the range closest to 0. Skipping a sub-block never drops a contributing pair.

template <int TS>
struct TileShape {
    static constexpr int kSubPerRow = TS / 8;
    static constexpr int kSub = kSubPerRow * kSubPerRow;
    static constexpr int kThreads = (TS >= 16) ? 256 : 64;
    static constexpr int kWaves = kThreads / kWave;
    static constexpr int kSubPerWave = kSub / kWaves;  // pixels per thread
    // Minimum waves per SIMD asked of the compiler for the adjoint on 16-pixel tiles (at most 128 VGPRs): its LDS
    // allows four work-groups of 256 threads per CU at the default batch, and left to itself the compiler takes
    // more registers than that occupancy leaves (144 for the plane form); 1 leaves the choice to the compiler.
    static constexpr int kBwdMinWavesPerSimd = (TS == 16) ? 4 : 1;
};

// Launch shape of the forward splat: a wavefront owns kSubPerWave sub-blocks (wave w: w, w + kWaves, ...). A
// 16-pixel tile is one wavefront holding all four sub-blocks, so every Gaussian record read from LDS serves up to
// four pixels per lane; a 32-pixel tile is four wavefronts of four sub-blocks each. kDefaultBatch: the Gaussians
// staged per LDS batch unless the settings choose.
template <int TS>
struct FwdShape {
    static constexpr int kSubPerRow = TS / 8;
    static constexpr int kSub = kSubPerRow * kSubPerRow;
    static constexpr int kThreads = (TS == 32) ? 256 : 64;
    static constexpr int kWaves = kThreads / kWave;
    static constexpr int kSubPerWave = kSub / kWaves;
    static constexpr int kDefaultBatch = (TS == 8) ? 64 : (TS == 16 ? 128 : 256);
    // Minimum waves per SIMD asked of the compiler: for 16-pixel tiles five (at most 96 VGPRs, no spill); config 2
    // measured 2.14 ms against 2.26 at six (80 VGPRs, spilling) and 2.37 at four; 1 leaves the choice to the
    // compiler.
    static constexpr int kMinWavesPerSimd = (TS == 16) ? 5 : 1;
};

struct SubBounds {
    float lat_lo, lat_hi, ax_lo, ax_hi, ele_lo, ele_hi;
};

// Pixel of sub-block sb handled by `lane`.
template <int TS>
__device__ __forceinline__ void sub_pixel(int tx, int ty, int sb, int lane, int& px, int& py) {
    px = tx * TS + (sb % TileShape<TS>::kSubPerRow) * 8 + (lane & 7);
    py = ty * TS + (sb / TileShape<TS>::kSubPerRow) * 8 + (lane >> 3);
}

// Lateral/axial/elevational extent of the valid pixels of a sub-block (wave reduction), widened for rounding.
__device__ __forceinline__ SubBounds sub_bounds(bool valid, float lat, float ax, float ele) {
    SubBounds b;
    b.lat_lo = wave_min(valid ? lat : 3.0e38f);
    b.lat_hi = wave_max(valid ? lat : -3.0e38f);
    b.ax_lo = wave_min(valid ? ax : 3.0e38f);
    b.ax_hi = wave_max(valid ? ax : -3.0e38f);
    b.ele_lo = wave_min(valid ? ele : 3.0e38f);
    b.ele_hi = wave_max(valid ? ele : -3.0e38f);
    const float pad = 1e-3f + 4e-6f * fmaxf(fmaxf(fabsf(b.lat_lo), fabsf(b.lat_hi)), fmaxf(fabsf(b.ax_lo), fabsf(b.ax_hi)));
    if (b.lat_lo <= b.lat_hi) {
        b.lat_lo -= pad; b.lat_hi += pad; b.ax_lo -= pad; b.ax_hi += pad;
        b.ele_lo -= pad; b.ele_hi += pad;
    }
    return b;
}

// Sub-block mask of one instance from its pair's setup (pair_section); chi2w = sqrt_chi2^2 * 1.002.
// Section test: at elevation offset d the pixels with q <= chi2 form the ellipse f(x - c(d)) <= chi2 - kk d^2,
// c(d) = (lat_g, ax_g) + (u_lat, u_ax) d. Over the sub-block's elevation range all of them lie in the ellipse of
// the largest level (d* = the offset closest to 0) moved along the centre segment, so the sub-block can hold
// such a pixel only if that ellipse meets its rectangle shifted by the segment (and widened by the pad). The
// minimum of the convex f over a rectangle not containing its centre lies on the edge x = x_n or the edge
// y = y_n nearest the centre (x_n, y_n = 0 clamped into the rectangle; a minimiser on a far edge or corner would
// make f <= 0 there by the optimality conditions), and along an edge at the clamped 1-D minimiser, so
// min(f(x_n, y_a), f(x_b, y_n)) is the exact minimum; it is compared with the level widened by 0.4 %.
template <int NSUB>
__device__ __forceinline__ uint32_t sub_mask(const SubBounds* __restrict__ sb, float4 s0, float4 s1, float lat_g,
                                             float ax_g, float ele_g, float chi2w) {
    const float kk = s0.z;
    uint32_t m = 0;
    if (kk < 0.0f) {  // marginal box only
#pragma unroll
        for (int s = 0; s < NSUB; ++s) {
            const bool hit = (lat_g + s0.x >= sb[s].lat_lo) && (lat_g - s0.x <= sb[s].lat_hi) &&
                             (ax_g + s0.y >= sb[s].ax_lo) && (ax_g - s0.y <= sb[s].ax_hi);
            m |= (hit ? 1u : 0u) << s;
        }
        return m;
    }
    const float u_lat = s0.x, u_ax = s0.y, t_x = s0.w;
    const float p = s1.x, q = s1.y, sy = s1.z, t_y = s1.w;
    const float pad = sub_pad(lat_g, ax_g);
#pragma unroll
    for (int s = 0; s < NSUB; ++s) {
        const float d_lo = sb[s].ele_lo - ele_g, d_hi = sb[s].ele_hi - ele_g;
        const float d0 = fminf(fmaxf(0.0f, d_lo), d_hi);
        const float rem = chi2w - kk * d0 * d0;
        const float cl0 = lat_g + u_lat * d_lo, cl1 = lat_g + u_lat * d_hi;
        const float ca0 = ax_g + u_ax * d_lo, ca1 = ax_g + u_ax * d_hi;
        const float x0 = sb[s].lat_lo - fmaxf(cl0, cl1) - pad, x1 = sb[s].lat_hi - fminf(cl0, cl1) + pad;
        const float y0 = sb[s].ax_lo - fmaxf(ca0, ca1) - pad, y1 = sb[s].ax_hi - fminf(ca0, ca1) + pad;
        const float xn = fminf(fmaxf(0.0f, x0), x1), yn = fminf(fmaxf(0.0f, y0), y1);
        const float ya = fminf(fmaxf(-t_x * xn, y0), y1);
        const float xb = fminf(fmaxf(-t_y * yn, x0), x1);
        const float u1 = p * xn + q * ya, v1 = sy * ya;
        const float u2 = p * xb + q * yn, v2 = sy * yn;
        const float f = fminf(u1 * u1 + v1 * v1, u2 * u2 + v2 * v2);
        const bool hit = sb[s].lat_lo <= sb[s].lat_hi && rem > 0.0f && f <= rem * 1.004f;
        m |= (hit ? 1u : 0u) << s;
    }
    return m;
}

__device__ __forceinline__ int frame_of_tile(const int32_t* __restrict__ frame_tile_base, int n_frames, int t) {
    return upper_index(frame_tile_base, n_frames, t);
}

// ---------------------------------------------------------------- forward splat
// One workgroup per (frame, tile) work item. Pixel-parallel: each thread accumulates its pixel's sums over
// the tile's sorted list, which is staged through LDS in batches of kBatch Gaussians (records gathered
// with 16 B loads, read back as wavefront broadcasts; kBatch is the work-group's thread count unless the settings
// choose 64, 128 or 256, and the sums do not depend on it). Follows kernel_pixel_ultragauss_forward,
// ops/ultragauss_renderer/src/shader/forward.cu:17-76: alpha_hat = o * exp(-q/2) for q <= chi2 and
// alpha_hat > alpha_min; A = sum alpha_hat + bg; image = clamp((sum alpha_hat c + bg I_bg) / max(A,1e-8));
// depth = sum alpha_hat elev / max(sum alpha_hat, 1e-8). The per-pixel sums are FP32 in list order, as in the
// reference kernel (its accumulators are floats): per evaluated pair one add and one multiply-add with no
// conversions. Against an FP64 sum that is an error of a few 1e-7 relative in the image, and the adjoint's
// coefficient dL/dalpha_hat = dN (c - image) inherits it only where c is close to the image, where the term is
// small.

// Plane form of q (pair_plane_form) at a pixel with probe coordinates (lat, ax), given x = lat - cx: with that
// subtraction, two subtractions and five multiply-adds in one fixed order, the same in the forward and the adjoint
// so that their decisions are the same bits. The splat kernels pass the coefficients scaled (kQScaleSqrt,
// kQScale), which gives q' = q log2(e)/2.
__device__ __forceinline__ float plane_q_x(float x, float ax, float cy, float K, float p, float qq, float s) {
    const float y = ax - cy;
    const float u = __builtin_fmaf(p, x, qq * y);
    const float v = s * y;
    return __builtin_fmaf(u, u, __builtin_fmaf(v, v, K));
}

// Probe coordinates (lat, ax) of pixel (ix, iy) of a linear-probe frame, with the FP32 operations of
// pixel_position (so the plane form sees the lateral and axial values the world position is built from).
__device__ __forceinline__ void grid_lat_ax(const us3d_frame& fr, int ix, int iy, float& lat, float& ax) {
    lat = ((float)ix - fr.col_offset) * fr.delta_lat;
    ax = fr.near_mm + (float)iy * fr.delta_ax;
}

// LDS of one forward workgroup.
template <int TS, int kBatch>
struct FwdLds {
    float4 r0[kBatch], r1[kBatch], r2[kBatch];
    uint32_t mask[kBatch];
    SubBounds sb[FwdShape<TS>::kSub];
};

// The forward of one (frame, tile) work item; kPlaneBody selects the form of q (see ug_forward_kernel); the
// tile's list is staged kBatch Gaussians at a time.
template <int TS, bool kDepth, bool kPlaneBody, int kBatch>
__device__ __forceinline__ void ug_forward_tile(
    FwdLds<TS, kBatch>& L, const float4* __restrict__ rec, const int32_t* __restrict__ pair_g,
    const uint64_t* __restrict__ seg_key, const us3d_frame& fr, int tx, int ty, int s0, int s1,
    const float* __restrict__ pixels,
    const float4* __restrict__ pair_sect, float chi2, float sqrt_chi2, float alpha_min, float bg_alpha,
    float bg_intensity, int alpha_excludes_bg, float* __restrict__ out_image, float* __restrict__ out_alpha,
    float* __restrict__ out_depth, float4* __restrict__ out_aux, uint16_t* __restrict__ inst_mask,
    const LossFuse& loss) {
    using S = FwdShape<TS>;
    constexpr int NT = S::kThreads;
    constexpr int PPT = S::kSubPerWave;
    constexpr int NC = kPlaneBody ? 2 : 3;  // per-pixel coordinates held: (lat, ax) or the world position
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    const float ox = fr.origin[0], oy = fr.origin[1], oz = fr.origin[2];
    const float chi2w = sqrt_chi2 * sqrt_chi2 * 1.002f;
    const float thr = kPlaneBody ? chi2 * kQScale : chi2;  // the q test (scaled plane form: q')

    // A lane's pixels share their column: sub-block sb = wave + k kWaves sits in column sb % kSubPerRow, and
    // kWaves is a multiple of kSubPerRow whenever a lane holds more than one pixel; so with the plane form the
    // lateral offset lat - cx is formed once per Gaussian for all of them.
    // With one wavefront per tile a lane's pixels fall in kSubPerRow columns, pixel k in column k % kSubPerRow.
    constexpr int kCols = (PPT == 1 || S::kWaves % S::kSubPerRow == 0) ? 1 : S::kSubPerRow;
    static_assert(kCols == 1 || S::kWaves == 1, "a lane's pixels share their column, or one wave holds the tile");
    float pc[NC][PPT];
    bool valid[PPT];
    int pidx[PPT];
    float S_[PPT], N_[PPT], D_[PPT];
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const int sb = wave + k * S::kWaves;
        int ix, iy;
        sub_pixel<TS>(tx, ty, sb, lane, ix, iy);
        valid[k] = ix < fr.width && iy < fr.height;
        pidx[k] = fr.pixel_base + iy * fr.width + ix;
        float x = 0.f, y = 0.f, z = 0.f;
        if (valid[k]) pixel_position(fr, pixels, ix, iy, x, y, z);
        const float dx = x - ox, dy = y - oy, dz = z - oz;
        const float lat = dot3(dx, dy, dz, fr.axis_lat[0], fr.axis_lat[1], fr.axis_lat[2]);
        const float ax = dot3(dx, dy, dz, fr.axis_ax[0], fr.axis_ax[1], fr.axis_ax[2]);
        const float ele = dot3(dx, dy, dz, fr.axis_ele[0], fr.axis_ele[1], fr.axis_ele[2]);
        const SubBounds b = sub_bounds(valid[k], lat, ax, ele);
        if (lane == 0) L.sb[sb] = b;
        if constexpr (kPlaneBody) {
            grid_lat_ax(fr, ix, iy, pc[0][k], pc[1][k]);
        } else {
            pc[0][k] = x;
            pc[1][k] = y;
            pc[NC - 1][k] = z;
        }
        S_[k] = N_[k] = D_[k] = 0.f;
    }
    for (int base = s0; base < s1; base += kBatch) {
        __syncthreads();
        const int n = min(kBatch, s1 - base);
        for (int j = tid; j < n; j += NT) {
            const int64_t pv = (uint32_t)seg_key[base + j];  // the instance's visible pair
            const float4 ps2 = pair_sect[kSectF4 * pv + 2];   // plane form (cx, cy, K) and the pair's Gaussian
            const int g = __float_as_int(ps2.w);
            const float4 r0 = rec[3 * (int64_t)g + 0];
            const float4 r1 = rec[3 * (int64_t)g + 1];
            const float4 r2 = rec[3 * (int64_t)g + 2];
            const float vx = r0.x - ox, vy = r0.y - oy, vz = r0.z - oz;
            const float e = dot3(vx, vy, vz, fr.axis_ele[0], fr.axis_ele[1], fr.axis_ele[2]);
            const float lat_g = dot3(vx, vy, vz, fr.axis_lat[0], fr.axis_lat[1], fr.axis_lat[2]);
            const float ax_g = dot3(vx, vy, vz, fr.axis_ax[0], fr.axis_ax[1], fr.axis_ax[2]);
            const float4 ps0 = pair_sect[kSectF4 * pv + 0], ps1 = pair_sect[kSectF4 * pv + 1];
            const uint32_t m = sub_mask<S::kSub>(L.sb, ps0, ps1, lat_g, ax_g, e, chi2w);
            L.mask[j] = m;
            inst_mask[base + j] = (uint16_t)m;  // the adjoint reads it instead of computing it again
            if constexpr (kPlaneBody) {
                L.r0[j] = make_float4(ps2.x, ps2.y, ps2.z * kQScale, r0.w);  // cx, cy, K', opacity
                L.r1[j] = make_float4(ps1.x * kQScaleSqrt, ps1.y * kQScaleSqrt, ps1.z * kQScaleSqrt,
                                      r1.x);  // p', qq', s', intensity
                L.r2[j] = make_float4(e, 0.f, 0.f, 0.f);
            } else {
                L.r0[j] = r0;                                  // mu, opacity
                L.r1[j] = make_float4(r1.x, e, r1.y, r1.z);   // intensity, elev, l11, l12
                L.r2[j] = make_float4(r1.w, r2.x, r2.y, r2.z); // l13, l22, l23, l33
            }
        }
        __syncthreads();
        // The Gaussians that touch this wave's sub-blocks, 64 list entries at a time: ballots of their mask
        // bits give wave-uniform 64-bit sets; their union is walked lowest-first with scalar bit scans, each
        // record is read from LDS once and evaluated for every sub-block whose bit is set. Skipped entries
        // cost nothing and each pixel still sums its Gaussians in list order.
        for (int jb = 0; jb < n; jb += kWave) {
            const uint32_t mm = (jb + lane < n) ? L.mask[jb + lane] : 0u;
            uint64_t bk[PPT];
            uint64_t any = 0;
#pragma unroll
            for (int k = 0; k < PPT; ++k) {
                bk[k] = __ballot((mm >> (wave + k * S::kWaves)) & 1u);
                any |= bk[k];
            }

            while (any) {
                const int bit = (int)__builtin_ctzll(any);
                any &= any - 1;
                const int j = jb + bit;
                const float4 a = L.r0[j];
                const float4 b = L.r1[j];
                const float4 c = L.r2[j];
                const float cf = kPlaneBody ? b.w : b.x;
                const float ef = kPlaneBody ? c.x : b.y;
                float xg[kCols];  // lat - cx of the lane's columns
#pragma unroll
                for (int cc = 0; cc < kCols; ++cc) xg[cc] = kPlaneBody ? pc[0][cc] - a.x : 0.f;
#pragma unroll
                for (int k = 0; k < PPT; ++k) {
                    if (!((bk[k] >> bit) & 1ull)) continue;  // scalar test
                    float q;
                    if constexpr (kPlaneBody) {
                        q = plane_q_x(xg[k % kCols], pc[1][k], a.y, a.z, b.x, b.y, b.z);
                    } else {
                        float y0, y1, y2;
                        q = quad_form(b.z, b.w, c.x, c.y, c.z, c.w, pc[0][k] - a.x, pc[1][k] - a.y,
                                      pc[NC - 1][k] - a.z, y0, y1, y2);
                    }
                    if (q <= thr) {
                        const float ah = a.w * (kPlaneBody ? exp2_neg(q) : exp_neg_half(q));
                        if (ah > alpha_min) {
                            S_[k] += ah;
                            N_[k] = __builtin_fmaf(ah, cf, N_[k]);
                            if constexpr (kDepth) D_[k] = __builtin_fmaf(ah, ef, D_[k]);
                        }
                    }
                }
            }
        }
    }

    // the loss sums of the lane's pixels (fused training step): at most four pixels in FP32, then FP64 across the
    // wavefront and the frame
    float ls[kLossSums] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float gtv[PPT];  // their targets, loaded before the image is formed so the loads overlap it
#pragma unroll
    for (int k = 0; k < PPT; ++k) gtv[k] = (loss.gt && valid[k]) ? loss.gt[pidx[k]] : 0.f;
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        if (!valid[k]) continue;
        const float A = S_[k] + bg_alpha;
        const float num = __builtin_fmaf(bg_alpha, bg_intensity, N_[k]);
        const float img = num / fmaxf(A, 1e-8f);
        const float pred = fmaxf(0.f, fminf(1.f, img));
        out_image[pidx[k]] = pred;
        out_alpha[pidx[k]] = alpha_excludes_bg ? S_[k] : A;
        if constexpr (kDepth) out_depth[pidx[k]] = D_[k] / fmaxf(S_[k], 1e-8f);
        out_aux[pidx[k]] = make_float4(S_[k], img, kDepth ? D_[k] : 0.f, 0.f);
        if (loss.gt) {
            const float x = pred, y = gtv[k];
            ls[0] += fabsf(x - y);
            ls[1] += x;
            ls[2] += y;
            ls[3] = __builtin_fmaf(x, x, ls[3]);
            ls[4] = __builtin_fmaf(y, y, ls[4]);
            ls[5] = __builtin_fmaf(x, y, ls[5]);
        }
    }
    if (loss.gt) {  // uniform: one entry per (tile, wavefront), the lanes' sums combined in a fixed order
        const int64_t e = (int64_t)(fr.tile_base + ty * fr.tiles_x + tx) * S::kWaves + wave;
#pragma unroll
        for (int j = 0; j < kLossSums; ++j) {
            const double v = wave_sum_f64((double)ls[j]);
            if (lane == j) loss.part[e * kLossSums + j] = v;
        }
    }
}

// kDepth: the depth output is requested (out_depth != NULL); without it the depth numerator is not accumulated
// and aux holds 0 in its place. kPlane: the pixels are the frames' own grids (pixels == NULL); then the tiles of
// linear-probe frames evaluate q in the plane form of their pairs (pair_plane_form, read from pair_sect) from the
// pixels' probe coordinates, and fan-probe tiles take the general form q = |L^T (p - mu)|^2 from the pixels'
// world positions, as every tile does with explicit pixels.
template <int TS, bool kDepth, bool kPlane, int kBatch>
__global__ __launch_bounds__(FwdShape<TS>::kThreads) __attribute__((amdgpu_waves_per_eu(FwdShape<TS>::kMinWavesPerSimd, 8)))
void ug_forward_kernel(
    const float4* __restrict__ rec, const int32_t* __restrict__ tile_off, const int32_t* __restrict__ pair_g,
    const uint64_t* __restrict__ seg_key, const us3d_frame* __restrict__ frames,
    const int32_t* __restrict__ frame_tile_base, int n_frames, const float* __restrict__ pixels,
    const float4* __restrict__ pair_sect, float chi2, float sqrt_chi2, float alpha_min, float bg_alpha,
    float bg_intensity, int alpha_excludes_bg, const us3d_stats* __restrict__ stats, float* __restrict__ out_image,
    float* __restrict__ out_alpha, float* __restrict__ out_depth, float4* __restrict__ out_aux,
    uint16_t* __restrict__ inst_mask, int n_tiles, int32_t* __restrict__ queue_heads, LossFuse loss) {
    __shared__ FwdLds<TS, kBatch> lds;
    __shared__ int s_tile;
    if (stats->overflow) return;
    auto item = [&](int t) {
        const int f = frame_of_tile(frame_tile_base, n_frames, t);
        const us3d_frame& fr = frames[f];
        const int lt = t - fr.tile_base;
        const int tx = lt % fr.tiles_x, ty = lt / fr.tiles_x;
        const int s0 = tile_off[t], s1 = tile_off[t + 1];
        if (kPlane && fr.geometry == 0)  // workgroup-uniform
            ug_forward_tile<TS, kDepth, true, kBatch>(lds, rec, pair_g, seg_key, fr, tx, ty, s0, s1, pixels, pair_sect,
                                                      chi2, sqrt_chi2, alpha_min, bg_alpha, bg_intensity,
                                                      alpha_excludes_bg, out_image, out_alpha, out_depth, out_aux,
                                                      inst_mask, loss);
        else
            ug_forward_tile<TS, kDepth, false, kBatch>(lds, rec, pair_g, seg_key, fr, tx, ty, s0, s1, pixels,
                                                       pair_sect, chi2, sqrt_chi2, alpha_min, bg_alpha, bg_intensity,
                                                       alpha_excludes_bg, out_image, out_alpha, out_depth, out_aux,
                                                       inst_mask, loss);
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
// One workgroup per (frame, tile). Per-pixel adjoint coefficients are computed once from the stored sums
// and staged in LDS (sub-block-major). The work unit is an ITEM: one Gaussian of the tile's list paired with
// one 8x8 sub-block its mask keeps. A batch of up to kBatch Gaussians is loaded, its items are listed
// Gaussian-major (a Gaussian's items are consecutive, in sub-block order), and the wavefronts take chunks of
// 64 items: each lane loops over its item's 64 pixels (LDS reads with at most one address per sub-block),
// accumulating the 12 gradient sums in registers with no atomics. The lanes of one Gaussian are adjacent, so
// a segmented lane reduction in a fixed tree order combines them; a Gaussian cut by a chunk boundary keeps
// its two halves in LDS and they are added after the pass. Every step is order-fixed: bitwise reproducible.
// Formulas follow kernel_pixel_ultragauss_backward, ops/ultragauss_renderer/src/shader/backward.cu:17-131,
// with q = ||L^T d||^2:
//   dN = gI/A', v = num/A' (unclamped image), dX = gA - gD*Dn/S'^2, dD = gD/S'
//   (A' = max(A,1e-8), S' = max(S,1e-8)); dah = dN*(c - v) + dX + dD*elev (= dN*c + dA + ..., written so
//   that c - v is formed from two FP32 numbers instead of cancelling dN*c against dN*v);
//   dc += dN*ah; do += dah*w; m = 2 dq = -dah*ah
//   dmu = -L * sum(m y) + sum(dD*ah) * a_ele;  dl_ij = sum(m * y_j * d_i)
// Pairs are gated exactly as in the forward (q <= chi2 and ah > alpha_min), so the adjoint is the exact
// derivative of the forward. Partials go to the instance's emit index.

// float4 stride of a staged sub-block in LDS: 65 puts the sub-blocks that one instruction reads in
// different banks.
constexpr int kSubStride = 65;

// The 12 sums of one Gaussian: do, dc, sum(dD ah), sum(m y_0..2), dl11, dl12, dl13, dl22, dl23, dl33.
struct GradSums {
    float v[12];
};

__device__ __forceinline__ void write_partial(float4* __restrict__ partial, int64_t ei, const GradSums& a,
                                              const float4& r1, const float4& r2, const us3d_frame& fr) {
    const float l11 = r1.y, l12 = r1.z, l13 = r1.w, l22 = r2.x, l23 = r2.y, l33 = r2.z;
    const float aed = a.v[2], Y0 = a.v[3], Y1 = a.v[4], Y2 = a.v[5];
    const float dmx = __builtin_fmaf(aed, fr.axis_ele[0], -(l11 * Y0));
    const float dmy = __builtin_fmaf(aed, fr.axis_ele[1], -__builtin_fmaf(l22, Y1, l12 * Y0));
    const float dmz = __builtin_fmaf(aed, fr.axis_ele[2], -__builtin_fmaf(l33, Y2, __builtin_fmaf(l23, Y1, l13 * Y0)));
    partial[3 * ei + 0] = make_float4(a.v[0], a.v[1], dmx, dmy);
    partial[3 * ei + 1] = make_float4(dmz, a.v[6], a.v[7], a.v[8]);
    partial[3 * ei + 2] = make_float4(a.v[9], a.v[10], a.v[11], 0.f);
}

// Plane form (linear-probe tiles with the frames' own grids): an item's lane walks its sub-block row by row. Along
// a row the axial offset is fixed, so q = (p x + qq y)^2 + (s y)^2 + K costs one subtraction and two multiply-adds
// per pixel (the same operations as plane_q_x, so the decisions equal the forward's), and the gradient needs only
// six moments of m = -dah ah over the pixel offsets (X, Y) from the tile centre: with d = p - mu and y = L^T d
// affine in (X, Y), d = d_c + X U + Y V, sum(m y_k) and sum(m y_k d_i) follow from sum(m), sum(m X), sum(m Y),
// sum(m X^2), sum(m X Y), sum(m Y^2) once per instance (plane_sums_to_grad). Per pixel the moments cost three
// multiply-adds (row sums of m, m X, m X^2; the Y terms are added per row).
constexpr int kPlaneSums = 9;  // do, dc, sum(dD ah), M0, MX, MY, MXX, MXY, MYY

// The twelve record-space sums (GradSums) of an instance from its plane-form moments about the tile centre.
__device__ __forceinline__ GradSums plane_sums_to_grad(const GradSums& a, const float4& r0, const float4& r1,
                                                       const float4& r2, const us3d_frame& fr, int TS, int tx,
                                                       int ty) {
    const float l11 = r1.y, l12 = r1.z, l13 = r1.w, l22 = r2.x, l23 = r2.y, l33 = r2.z;
    const float half = 0.5f * (float)(TS - 1);
    const float lat_c = ((float)(tx * TS) + half - fr.col_offset) * fr.delta_lat;
    const float ax_c = fr.near_mm + ((float)(ty * TS) + half) * fr.delta_ax;
    float dc[3], U[3], V[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        dc[i] = __builtin_fmaf(ax_c, fr.axis_ax[i], __builtin_fmaf(lat_c, fr.axis_lat[i], fr.origin[i] - (&r0.x)[i]));
        U[i] = fr.delta_lat * fr.axis_lat[i];
        V[i] = fr.delta_ax * fr.axis_ax[i];
    }
    auto lt = [&](const float v[3], float y[3]) {  // y = L^T v
        y[0] = __builtin_fmaf(l13, v[2], __builtin_fmaf(l12, v[1], l11 * v[0]));
        y[1] = __builtin_fmaf(l23, v[2], l22 * v[1]);
        y[2] = l33 * v[2];
    };
    float yc[3], yU[3], yV[3];
    lt(dc, yc);
    lt(U, yU);
    lt(V, yV);
    const float M0 = a.v[3], MX = a.v[4], MY = a.v[5], MXX = a.v[6], MXY = a.v[7], MYY = a.v[8];
    GradSums o;
    o.v[0] = a.v[0];
    o.v[1] = a.v[1];
    o.v[2] = a.v[2];
#pragma unroll
    for (int k = 0; k < 3; ++k) o.v[3 + k] = __builtin_fmaf(yV[k], MY, __builtin_fmaf(yU[k], MX, yc[k] * M0));
    // sum(m y_k d_i) for (k, i) = (0,0) (0,1) (0,2) (1,1) (1,2) (2,2)
    const int kk[6] = {0, 0, 0, 1, 1, 2}, ii[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int e = 0; e < 6; ++e) {
        const int k = kk[e], i = ii[e];
        float g = yc[k] * dc[i] * M0;
        g = __builtin_fmaf(__builtin_fmaf(yc[k], U[i], yU[k] * dc[i]), MX, g);
        g = __builtin_fmaf(__builtin_fmaf(yc[k], V[i], yV[k] * dc[i]), MY, g);
        g = __builtin_fmaf(yU[k] * U[i], MXX, g);
        g = __builtin_fmaf(__builtin_fmaf(yU[k], V[i], yV[k] * U[i]), MXY, g);
        g = __builtin_fmaf(yV[k] * V[i], MYY, g);
        o.v[6 + e] = g;
    }
    return o;
}

// Plane-form staging inside the adjoint's pixel arrays: per sub-block kPlanePixStride float2 of (dN, v) (or
// (dN, dN v)) in p and of (dX, dD) in c (66: 16-byte aligned rows, sub-blocks in different banks), then from
// float4 NSUB * kPlaneCoordBase of c, per sub-block 8 column lateral and 8 row axial coordinates.
constexpr int kPlanePixStride = 66;
constexpr int kPlaneCoordBase = 33;
static_assert(kPlanePixStride * 8 <= kSubStride * 16 && (kPlaneCoordBase + 4) <= kSubStride, "plane staging fits");

// LDS of one adjoint workgroup staging kBatch Gaussians at a time (the plane form's two extra record arrays only
// where a tile can take it).
template <int TS, bool kPlane, int kBatch>
struct BwdLds {
    static constexpr int NSUB = TileShape<TS>::kSub;
    static constexpr int MAXI = kBatch * NSUB;                  // items of one batch, at most
    static constexpr int MAXC = (MAXI + kWave - 1) / kWave;     // chunks of one batch, at most
    float4 p[NSUB * kSubStride];  // general form: px, py, pz, dN; plane form: lat, ax, dN, v
    float4 c[NSUB * kSubStride];  // general form: v, dX, dD; plane form: dX, dD
    float4 g0[kBatch], g1[kBatch], g2[kBatch];                     // the batch's Gaussian records
    // the scaled plane form: (cx, cy, K', o), (p', qq', s', c)
    float4 g3[kPlane ? kBatch : 1], g4[kPlane ? kBatch : 1];
    int32_t ei[kBatch];                                            // emit index of each instance of the batch
    uint16_t item[MAXI];
    int wtot[TileShape<TS>::kWaves];
    GradSums pend[MAXC];   // head part of a Gaussian that continues into the next chunk
    GradSums carry[MAXC];  // tail part of a Gaussian that began in the previous chunk
    int pend_slot[MAXC];
};

// The adjoint of one (frame, tile) work item; kPlaneBody selects the form of q (see ug_backward_kernel). The list
// is staged kBatch Gaussians at a time, slot j of a batch by thread j mod kThreads.
template <int TS, bool kFull, bool kPlaneBody, bool kPlane, int kBatch>
__device__ __forceinline__ void ug_backward_tile(
    BwdLds<TS, kPlane, kBatch>& L, const float4* __restrict__ rec, const int32_t* __restrict__ pair_g,
    const uint64_t* __restrict__ seg_key, const us3d_frame& fr, int tx, int ty, int s0, int s1,
    const float* __restrict__ pixels,
    const float4* __restrict__ pair_sect, float chi2, float alpha_min,
    float bg_alpha, const float4* __restrict__ aux, const float* __restrict__ d_image,
    const float* __restrict__ d_alpha, const float* __restrict__ d_depth, float4* __restrict__ partial,
    const uint16_t* __restrict__ inst_mask, const float* __restrict__ loss_gt, const double* __restrict__ loss_coeff) {
    using S = TileShape<TS>;
    constexpr int NT = S::kThreads;
    constexpr int PPT = S::kSubPerWave;
    constexpr int NSUB = S::kSub;
    constexpr int NW = S::kWaves;
    constexpr int MAXC = BwdLds<TS, kPlane, kBatch>::MAXC;
    constexpr int SPT = (kBatch + NT - 1) / NT;  // batch slots per thread
    static_assert(kPlane || !kPlaneBody, "the plane form needs the plane record arrays");
    constexpr int NS = kPlaneBody ? kPlaneSums : 12;
    static_assert(kBatch <= 256 && NSUB <= 16, "item code packs (slot << 4) | sub-block into 12 bits");
    static_assert(FwdShape<TS>::kSub == NSUB, "the forward's instance masks number the same sub-blocks");
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    const float ox = fr.origin[0], oy = fr.origin[1], oz = fr.origin[2];
    const float chi2s = chi2 * kQScale;  // the plane form's q' test
    const float wmin_half = alpha_free_wmin_half(chi2s);

#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const int sb = wave + k * S::kWaves;
        int ix, iy;
        sub_pixel<TS>(tx, ty, sb, lane, ix, iy);
        const bool valid = ix < fr.width && iy < fr.height;
        float x = 0.f, y = 0.f, z = 0.f, dN = 0.f, dX = 0.f, dD = 0.f, vimg = 0.f;
        if (valid) {
            const int64_t p = (int64_t)fr.pixel_base + iy * fr.width + ix;
            if constexpr (!kPlaneBody) pixel_position(fr, pixels, ix, iy, x, y, z);
            // sum alpha_hat, unclamped image, sum alpha_hat * elev (the third only with alpha or depth gradients)
            float4 sm;
            if constexpr (kFull) {
                sm = aux[p];
            } else {
                const float2 h = reinterpret_cast<const float2*>(aux)[2 * p];
                sm = make_float4(h.x, h.y, 0.f, 0.f);
            }
            const float A = sm.x + bg_alpha;
            const float inv = 1.0f / fmaxf(A, 1e-8f);
            float gi;
            if (loss_coeff) {  // uniform: dL/dx of the fused image loss, x the clamped image the forward wrote
                const double xi = fmaxf(0.f, fminf(1.f, sm.y)), yi = loss_gt[p];
                const double sg = (xi > yi) ? 1.0 : ((xi < yi) ? -1.0 : 0.0);
                gi = (float)(loss_coeff[0] * sg + loss_coeff[1] + loss_coeff[2] * xi + loss_coeff[3] * yi);
            } else {
                gi = d_image[p];
            }
            const float ga = (kFull && d_alpha) ? d_alpha[p] : 0.f;
            const float gd = (kFull && d_depth) ? d_depth[p] : 0.f;
            const float invS = 1.0f / fmaxf(sm.x, 1e-8f);
            dN = gi * inv;
            vimg = sm.y;
            dD = gd * invS;
            dX = ga - gd * sm.z * invS * invS;
        }
        if constexpr (kPlaneBody) {
            // per pixel (dN, v) and (dX, dD), or (dN, dN v) without alpha and depth gradients (then dL/dah =
            // dN c - dN v is one multiply-add per pair); per sub-block the lateral coordinate of each column and the
            // axial coordinate of each row (every pixel of the sub-block: dN = 0 outside the frame)
            grid_lat_ax(fr, ix, iy, x, y);
            reinterpret_cast<float2*>(L.p)[sb * kPlanePixStride + lane] = make_float2(dN, kFull ? vimg : dN * vimg);
            reinterpret_cast<float2*>(L.c)[sb * kPlanePixStride + lane] = make_float2(dX, dD);
            float* la = reinterpret_cast<float*>(L.c + NSUB * kPlaneCoordBase) + sb * 16;
            if (lane < 8) la[lane] = x;
            if ((lane & 7) == 0) la[8 + (lane >> 3)] = y;
        } else {
            L.p[sb * kSubStride + lane] = make_float4(x, y, z, dN);
            L.c[sb * kSubStride + lane] = make_float4(vimg, dX, dD, 0.f);
        }
    }

    for (int base = s0; base < s1; base += kBatch) {
        const int n = min(kBatch, s1 - base);
        __syncthreads();  // staged pixels are in place; the previous batch is done with the batch arrays
        uint32_t m[SPT];
        int pc = 0;
#pragma unroll
        for (int i = 0; i < SPT; ++i) {
            const int j = tid + i * NT;
            m[i] = 0;
            if (j < n) {
                const int pv = (int)(uint32_t)seg_key[base + j];  // the instance's visible pair
                const float4 ps2 = pair_sect[kSectF4 * (int64_t)pv + 2];  // plane form and the pair's Gaussian
                const int g = __float_as_int(ps2.w);
                const float4 q0 = rec[3 * (int64_t)g + 0];
                const float4 q1 = rec[3 * (int64_t)g + 1];
                const float4 q2 = rec[3 * (int64_t)g + 2];
                const int64_t ei = emit_index_rec(pair_sect[kSectF4 * (int64_t)pv + 3], tx, ty);
                m[i] = inst_mask[base + j];  // the sub-block mask the forward left for this instance
                L.g0[j] = q0;
                L.g1[j] = q1;
                L.g2[j] = q2;
                L.ei[j] = (int32_t)ei;
                if constexpr (kPlaneBody) {
                    const float4 ps1 = pair_sect[kSectF4 * (int64_t)pv + 1];
                    L.g3[j] = make_float4(ps2.x, ps2.y, ps2.z * kQScale, q0.w);  // cx, cy, K', opacity
                    L.g4[j] = make_float4(ps1.x * kQScaleSqrt, ps1.y * kQScaleSqrt, ps1.z * kQScaleSqrt,
                                          q1.x);  // p', qq', s', intensity
                }
                if (m[i] == 0) {  // touches no pixel of the tile: its partial is zero
                    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
                    partial[3 * ei + 0] = z4;
                    partial[3 * ei + 1] = z4;
                    partial[3 * ei + 2] = z4;
                }
            }
            pc += __popc(m[i]);
        }
        for (int c = tid; c < MAXC; c += NT) L.pend_slot[c] = -1;
        // exclusive scan of the per-thread item counts over the batch; a thread lists its slots' items in slot
        // order, so each Gaussian's items are consecutive
        int incl = pc;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            const int v = __shfl_up(incl, off, kWave);
            if (lane >= off) incl += v;
        }
        if (lane == kWave - 1) L.wtot[wave] = incl;
        __syncthreads();
        int pos = incl - pc, total = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const int tw = L.wtot[w];
            pos += (w < wave) ? tw : 0;
            total += tw;
        }
#pragma unroll
        for (int i = 0; i < SPT; ++i)
            for (uint32_t mm = m[i]; mm; mm &= mm - 1)
                L.item[pos++] = (uint16_t)(((tid + i * NT) << 4) | __builtin_ctz(mm));
        __syncthreads();

        const int nchunks = (total + kWave - 1) / kWave;
        for (int c = wave; c < nchunks; c += NW) {
            const int it = c * kWave + lane;
            const bool valid = it < total;
            const int item = valid ? (int)L.item[it] : 0;
            const int slot = valid ? (item >> 4) : -1 - lane;  // unique negative ids for empty lanes
            const int sb = item & 15;
            const int prev_slot = (valid && it > 0) ? ((int)L.item[it - 1] >> 4) : -1000;
            const int next_slot = (it + 1 < total) ? ((int)L.item[it + 1] >> 4) : -1000;
            const bool first = valid && prev_slot != slot;  // first item of its Gaussian
            const bool last = valid && next_slot != slot;   // last item of its Gaussian

            GradSums a;
#pragma unroll
            for (int k = 0; k < 12; ++k) a.v[k] = 0.f;
            // wave-uniform: every item's opacity passes alpha_free, so no pixel can fail the alpha_hat test alone
            const float o_item = L.g0[valid ? slot : 0].w;
            const bool chunk_alpha_free = __ballot(valid && !alpha_free(o_item, wmin_half, alpha_min)) == 0;
            if (valid) {
                const float4 r0 = L.g0[slot];
                const float e = dot3(r0.x - ox, r0.y - oy, r0.z - oz, fr.axis_ele[0], fr.axis_ele[1], fr.axis_ele[2]);
                const float4* __restrict__ sp = L.p + sb * kSubStride;
                const float4* __restrict__ sc = L.c + sb * kSubStride;
                if constexpr (kPlaneBody) {
                    const float4 A = L.g3[slot];  // cx, cy, K', o
                    const float4 B = L.g4[slot];  // p', qq', s', c
                    const float* la = reinterpret_cast<const float*>(L.c + NSUB * kPlaneCoordBase) + sb * 16;
                    const float4 la0 = *reinterpret_cast<const float4*>(la), la1 = *reinterpret_cast<const float4*>(la + 4);
                    // lateral offsets from the minimum, once per item (the subtraction before plane_q_x, same bits)
                    const float xs[8] = {la0.x - A.x, la0.y - A.x, la0.z - A.x, la0.w - A.x,
                                         la1.x - A.x, la1.y - A.x, la1.z - A.x, la1.w - A.x};
                    // (dN, v) or (dN, dN v) of the sub-block's pixels, two per float4, a row = four float4 read
                    // together at the start of the row, so the LDS latency is paid once per row
                    const float4* __restrict__ dv = reinterpret_cast<const float4*>(reinterpret_cast<const float2*>(L.p) +
                                                                                    sb * kPlanePixStride);
                    const float2* __restrict__ dxd = reinterpret_cast<const float2*>(L.c) + sb * kPlanePixStride;
                    // With wp = w at a passing pixel and 0 elsewhere and t = dL/dah * wp, every sum is one of
                    // t's or wp's: do = sum t, dc = o sum dN wp, sum(dD ah) = o sum dD wp, and m = -dL/dah ah =
                    // -o t, so the moments of m are -o times those of t (applied once per item below).
                    float acw = 0.f, aedw = 0.f;
                    float T0 = 0.f, TX = 0.f, TY = 0.f, TXX = 0.f, TXY = 0.f, TYY = 0.f;
                    // the sub-block's 64 pixels, with or without the alpha_hat test (the same sums either way)
                    auto walk = [&](auto alpha_test) {
                    constexpr bool kAlphaTest = decltype(alpha_test)::value;
#pragma unroll
                    for (int r = 0; r < 8; ++r) {
                        float4 cur[4];
#pragma unroll
                        for (int i = 0; i < 4; ++i) cur[i] = dv[4 * r + i];
                        const float y = la[8 + r] - A.y;
                        const float qy = B.y * y;
                        const float v = B.z * y;
                        const float vv = __builtin_fmaf(v, v, A.z);
                        float tt[8];  // t of the row's pixels
#pragma unroll
                        for (int jx = 0; jx < 8; ++jx) {
                            const float dN = (jx & 1) ? cur[jx >> 1].z : cur[jx >> 1].x;
                            const float vi = (jx & 1) ? cur[jx >> 1].w : cur[jx >> 1].y;  // v, or dN v (!kFull)
                            // branch-free: a pixel that fails the forward's tests (q <= chi2, ah > alpha_min)
                            // enters every sum through wp = 0; the coefficients of a listed pair and the pixel
                            // coefficients are finite, so the masked terms are exact zeros
                            const float u = __builtin_fmaf(B.x, xs[jx], qy);
                            const float q = __builtin_fmaf(u, u, vv);
                            const float w = exp2_neg(q);
                            const bool pass = (q <= chi2s) && (!kAlphaTest || A.w * w > alpha_min);
                            const float wp = pass ? w : 0.f;
                            float da;
                            if constexpr (kFull) {
                                const float2 cc = dxd[8 * r + jx];
                                da = __builtin_fmaf(dN, B.w - vi, __builtin_fmaf(cc.y, e, cc.x));
                                aedw = __builtin_fmaf(cc.y, wp, aedw);
                            } else {
                                da = __builtin_fmaf(dN, B.w, -vi);
                            }
                            acw = __builtin_fmaf(dN, wp, acw);
                            tt[jx] = da * wp;
                        }
                        // the row's moments in X = jx - 3.5 from the symmetric columns jx and 7 - jx: with
                        // s = t_jx + t_(7-jx), d = t_(7-jx) - t_jx and Xh = 3.5 - jx, sum t = sum s,
                        // sum t X = sum d Xh and sum t X^2 = sum s Xh^2 (five operations per two pixels)
                        float R0 = tt[0] + tt[7];
                        float RX = (tt[7] - tt[0]) * 3.5f;
                        float RXX = R0 * 12.25f;
#pragma unroll
                        for (int h = 1; h < 4; ++h) {
                            const float sp = tt[h] + tt[7 - h], dp = tt[7 - h] - tt[h];
                            const float Xh = 3.5f - (float)h;  // constants once the loop is unrolled
                            R0 += sp;
                            RX = __builtin_fmaf(dp, Xh, RX);
                            RXX = __builtin_fmaf(sp, Xh * Xh, RXX);
                        }
                        const float Y = (float)r - 3.5f;
                        T0 += R0;
                        TX += RX;
                        TXX += RXX;
                        TY = __builtin_fmaf(R0, Y, TY);
                        TYY = __builtin_fmaf(R0, Y * Y, TYY);
                        TXY = __builtin_fmaf(RX, Y, TXY);
                    }
                    };
                    if (chunk_alpha_free)
                        walk(std::false_type{});
                    else
                        walk(std::true_type{});
                    const float mo = -A.w;
                    const float M0 = mo * T0, MX = mo * TX, MY = mo * TY, MXX = mo * TXX, MXY = mo * TXY,
                                MYY = mo * TYY;
                    // moments about the sub-block centre -> about the tile centre (offset (ox_, oy_) pixels)
                    const float ox_ = (float)((sb % S::kSubPerRow) * 8) + 3.5f - 0.5f * (float)(TS - 1);
                    const float oy_ = (float)((sb / S::kSubPerRow) * 8) + 3.5f - 0.5f * (float)(TS - 1);
                    a.v[0] = T0; a.v[1] = A.w * acw; a.v[2] = A.w * aedw; a.v[3] = M0;
                    a.v[6] = __builtin_fmaf(2.f * ox_, MX, __builtin_fmaf(ox_ * ox_, M0, MXX));
                    a.v[7] = __builtin_fmaf(ox_, MY, __builtin_fmaf(oy_, MX, __builtin_fmaf(ox_ * oy_, M0, MXY)));
                    a.v[8] = __builtin_fmaf(2.f * oy_, MY, __builtin_fmaf(oy_ * oy_, M0, MYY));
                    a.v[4] = __builtin_fmaf(ox_, M0, MX);
                    a.v[5] = __builtin_fmaf(oy_, M0, MY);
                } else {
                    const float4 r1 = L.g1[slot], r2 = L.g2[slot];
                    const float mx = r0.x, my = r0.y, mz = r0.z, o = r0.w, c = r1.x;
                    const float l11 = r1.y, l12 = r1.z, l13 = r1.w, l22 = r2.x, l23 = r2.y, l33 = r2.z;
                    float ao = 0.f, ac = 0.f, aed = 0.f, Y0 = 0.f, Y1 = 0.f, Y2 = 0.f;
                    float g11 = 0.f, g12 = 0.f, g13 = 0.f, g22 = 0.f, g23 = 0.f, g33 = 0.f;
                    // the next pixel's position and coefficients are read before this pixel's work, so the LDS
                    // latency overlaps it (the branches below would otherwise expose it on every pixel)
                    float4 pn = sp[0], cn = sc[0];
                    for (int j = 0; j < kWave; ++j) {
                        const float4 pp = pn, cc = cn;
                        pn = sp[(j + 1) & (kWave - 1)];
                        cn = sc[(j + 1) & (kWave - 1)];
                        const float dx = pp.x - mx, dy = pp.y - my, dz = pp.z - mz;
                        float y0, y1, y2;
                        const float q = quad_form(l11, l12, l13, l22, l23, l33, dx, dy, dz, y0, y1, y2);
                        if (q <= chi2) {
                            const float w = exp_neg_half(q);
                            const float ah = o * w;
                            if (ah > alpha_min) {
                                // dL/dah = dN c + dA + ... = dN (c - v) + dX + dD elev: the (c - v) form avoids the
                                // cancellation of dN c against dN v
                                const float da = kFull ? __builtin_fmaf(pp.w, c - cc.x, __builtin_fmaf(cc.z, e, cc.y))
                                                       : pp.w * (c - cc.x);
                                ac = __builtin_fmaf(pp.w, ah, ac);
                                ao = __builtin_fmaf(da, w, ao);
                                if constexpr (kFull) aed = __builtin_fmaf(cc.z, ah, aed);
                                const float mm = -da * ah;
                                const float m0 = mm * y0, m1 = mm * y1, m2 = mm * y2;
                                Y0 += m0; Y1 += m1; Y2 += m2;
                                g11 = __builtin_fmaf(m0, dx, g11);
                                g12 = __builtin_fmaf(m0, dy, g12);
                                g13 = __builtin_fmaf(m0, dz, g13);
                                g22 = __builtin_fmaf(m1, dy, g22);
                                g23 = __builtin_fmaf(m1, dz, g23);
                                g33 = __builtin_fmaf(m2, dz, g33);
                            }
                        }
                    }
                    a.v[0] = ao; a.v[1] = ac; a.v[2] = aed; a.v[3] = Y0; a.v[4] = Y1; a.v[5] = Y2;
                    a.v[6] = g11; a.v[7] = g12; a.v[8] = g13; a.v[9] = g22; a.v[10] = g23; a.v[11] = g33;
                }
            }
            // segmented reduction: a Gaussian's lanes are adjacent; after the steps the first lane of each run
            // holds the run's sum (tree order fixed by the lane positions)
#pragma unroll
            for (int off = 1; off < NSUB; off <<= 1) {
                const int oslot = __shfl_down(slot, off, kWave);
                const bool same = (lane + off < kWave) && oslot == slot;
#pragma unroll
                for (int k = 0; k < NS; ++k) {
                    const float v = __shfl_down(a.v[k], off, kWave);
                    a.v[k] = same ? a.v[k] + v : a.v[k];
                }
            }
            const int up_slot = __shfl_up(slot, 1, kWave);
            const bool head = valid && (lane == 0 || up_slot != slot);
            const int slot63 = __shfl(slot, kWave - 1, kWave);
            const bool last63 = __shfl(last ? 1 : 0, kWave - 1, kWave) != 0;
            const bool to_next = (slot == slot63) && !last63;  // run reaches lane 63, Gaussian goes on
            const bool from_prev = (lane == 0) && valid && !first;
            if (head) {
                if (from_prev) {
                    L.carry[c] = a;
                } else if (to_next) {
                    L.pend[c] = a;
                    L.pend_slot[c] = slot;
                } else {
                    if constexpr (kPlaneBody)
                        a = plane_sums_to_grad(a, L.g0[slot], L.g1[slot], L.g2[slot], fr, TS, tx, ty);
                    write_partial(partial, L.ei[slot], a, L.g1[slot], L.g2[slot], fr);
                }
            }
        }
        __syncthreads();
        // Gaussians cut by a chunk boundary: head part + tail part
        for (int c = tid; c + 1 < nchunks; c += NT) {
            const int slot = L.pend_slot[c];
            if (slot < 0) continue;
            GradSums a = L.pend[c];
            const GradSums& b = L.carry[c + 1];
#pragma unroll
            for (int k = 0; k < NS; ++k) a.v[k] += b.v[k];
            if constexpr (kPlaneBody) a = plane_sums_to_grad(a, L.g0[slot], L.g1[slot], L.g2[slot], fr, TS, tx, ty);
            write_partial(partial, L.ei[slot], a, L.g1[slot], L.g2[slot], fr);
        }
    }
}

// kFull: alpha or depth gradients are given (d_alpha or d_depth != NULL); without them dL/dalpha_hat is
// dN (c - v) alone and the depth term is not accumulated. kPlane: the pixels are the frames' own grids (pixels ==
// NULL); linear-probe tiles then take the plane form of q (as in the forward) and fan tiles the general form.
template <int TS, bool kFull, bool kPlane, int kBatch>
__global__ __launch_bounds__(TileShape<TS>::kThreads)
__attribute__((amdgpu_waves_per_eu(TileShape<TS>::kBwdMinWavesPerSimd, 8))) void ug_backward_kernel(
    const float4* __restrict__ rec, const int32_t* __restrict__ tile_off, const int32_t* __restrict__ pair_g,
    const uint64_t* __restrict__ seg_key, const us3d_frame* __restrict__ frames,
    const int32_t* __restrict__ frame_tile_base, int n_frames,
    const float* __restrict__ pixels, const float4* __restrict__ pair_sect, float chi2, float alpha_min,
    float bg_alpha, const float4* __restrict__ aux,
    const float* __restrict__ d_image, const float* __restrict__ d_alpha, const float* __restrict__ d_depth,
    const us3d_stats* __restrict__ stats, float4* __restrict__ partial, const uint16_t* __restrict__ inst_mask,
    int n_tiles, int32_t* __restrict__ queue_heads, LossFuse loss) {
    __shared__ BwdLds<TS, kPlane, kBatch> lds;
    __shared__ int s_tile;
    if (stats->overflow) return;
    auto item = [&](int t) {
        const int f = frame_of_tile(frame_tile_base, n_frames, t);
        const us3d_frame& fr = frames[f];
        const int lt = t - fr.tile_base;
        const int tx = lt % fr.tiles_x, ty = lt / fr.tiles_x;
        const int s0 = tile_off[t], s1 = tile_off[t + 1];
        if (s0 == s1) return;  // workgroup-uniform
        const double* coeff = loss.coeff ? loss.coeff + 4 * f : nullptr;
        if (kPlane && fr.geometry == 0)  // workgroup-uniform
            ug_backward_tile<TS, kFull, kPlane, kPlane, kBatch>(lds, rec, pair_g, seg_key, fr, tx,
                                                                ty, s0, s1, pixels, pair_sect, chi2, alpha_min, bg_alpha,
                                                                aux, d_image, d_alpha, d_depth, partial, inst_mask,
                                                                loss.gt, coeff);
        else
            ug_backward_tile<TS, kFull, false, kPlane, kBatch>(lds, rec, pair_g, seg_key, fr, tx,
                                                               ty, s0, s1, pixels, pair_sect, chi2, alpha_min, bg_alpha,
                                                               aux, d_image, d_alpha, d_depth, partial, inst_mask,
                                                               loss.gt, coeff);
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

// One wavefront per Gaussian: lanes stride over the Gaussian's instances (contiguous in emit order across
// all frames of its sweep), then a fixed butterfly reduction -- bitwise reproducible. Applies the raw
// parameter chain d raw_ii = dl_ii * 2 raw_ii (backward.cu:124-129).
__global__ __launch_bounds__(256) void ug_reduce_kernel(int n_gauss, const us3d_sweep* __restrict__ sweeps,
                                                        int n_sweeps, const uint64_t* __restrict__ pair_bits,
                                                        const int32_t* __restrict__ word_off,
                                                        const int32_t* __restrict__ pair_off,
                                                        const float4* __restrict__ partial,
                                                        const float* __restrict__ tril,
                                                        const float4* __restrict__ rec, int logits,
                                                        const us3d_stats* __restrict__ stats,
                                                        float* __restrict__ d_means, float* __restrict__ d_tril,
                                                        float* __restrict__ d_opacity,
                                                        float* __restrict__ d_intensity) {
    const int g = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    if (g >= n_gauss || stats->overflow) return;
    const us3d_sweep sw = sweep_for_reduce(sweeps, n_sweeps, g);
    const int64_t p0 = sw.pair_base + (int64_t)(g - sw.g_begin) * sw.f_count;
    // the Gaussian's visible pairs are v in [rank(p0), rank(p0 + f_count)); their instances are consecutive
    const int e0 = pair_off[pair_rank(pair_bits, word_off, p0)];
    const int e1 = pair_off[pair_rank(pair_bits, word_off, p0 + sw.f_count)];
    float acc[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[k] = 0.f;
    for (int e = e0 + lane; e < e1; e += kWave) {
        const float4 a = partial[3 * (int64_t)e + 0];
        const float4 b = partial[3 * (int64_t)e + 1];
        const float4 c = partial[3 * (int64_t)e + 2];
        acc[0] += a.x; acc[1] += a.y; acc[2] += a.z; acc[3] += a.w;
        acc[4] += b.x; acc[5] += b.y; acc[6] += b.z; acc[7] += b.w;
        acc[8] += c.x; acc[9] += c.y; acc[10] += c.z;
    }
#pragma unroll
    for (int k = 0; k < 11; ++k) acc[k] = wave_sum(acc[k]);
    if (lane == 0) {
        const float* t = tril + (int64_t)g * 6;
        if (logits) {  // through the sigmoids of the pack: d x = d s * s (1 - s)
            const float o = rec[3 * (int64_t)g + 0].w, c = rec[3 * (int64_t)g + 1].x;
            d_opacity[g] = acc[0] * (o * (1.0f - o));
            d_intensity[g] = acc[1] * (c * (1.0f - c));
        } else {
            d_opacity[g] = acc[0];
            d_intensity[g] = acc[1];
        }
        d_means[3 * (int64_t)g + 0] = acc[2];
        d_means[3 * (int64_t)g + 1] = acc[3];
        d_means[3 * (int64_t)g + 2] = acc[4];
        d_tril[6 * (int64_t)g + 0] = acc[5] * (2.0f * t[0]);
        d_tril[6 * (int64_t)g + 1] = acc[6];
        d_tril[6 * (int64_t)g + 2] = acc[7];
        d_tril[6 * (int64_t)g + 3] = acc[8] * (2.0f * t[3]);
        d_tril[6 * (int64_t)g + 4] = acc[9];
        d_tril[6 * (int64_t)g + 5] = acc[10] * (2.0f * t[5]);
    }
}

// ---------------------------------------------------------------- per-pixel candidate lists
// Thread per pixel over its tile's sorted list, keeping g iff |(p-mu).a_lat| <= r_lat and
// |(p-mu).a_ax| <= r_ax (pixel_hits_gaussian_probe_window, core.cu:166-182; count/fill core.cu:184-287).
__global__ void ug_pixel_lists_kernel(int64_t n_pixels, const us3d_frame* __restrict__ frames, int n_frames,
                                      const int32_t* __restrict__ frame_pixel_base, const us3d_sweep* __restrict__ sweeps,
                                      const float4* __restrict__ rec, const int32_t* __restrict__ tile_off,
                                      const uint64_t* __restrict__ seg_key, const int32_t* __restrict__ pair_g,
                                      const float* __restrict__ pixels,
                                      float sqrt_chi2, int ts, int pass, int32_t* __restrict__ counts,
                                      const int32_t* __restrict__ offsets, int32_t* __restrict__ indices,
                                      const us3d_stats* __restrict__ stats) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= n_pixels || stats->overflow) return;
    const int f = upper_index(frame_pixel_base, n_frames, (int)p);
    const us3d_frame& fr = frames[f];
    const int lp = (int)(p - fr.pixel_base);
    const int ix = lp % fr.width, iy = lp / fr.width;
    const int t = fr.tile_base + (iy / ts) * fr.tiles_x + ix / ts;
    const int s0 = tile_off[t], s1 = tile_off[t + 1];
    float x, y, z;
    pixel_position(fr, pixels, ix, iy, x, y, z);
    const int g_begin = sweeps[fr.sweep].g_begin;
    int c = 0;
    int w = pass ? offsets[p] : 0;
    for (int i = s0; i < s1; ++i) {
        const int g = pair_g[(uint32_t)seg_key[i]];
        const float4 r0 = rec[3 * (int64_t)g + 0];
        const float4 r1 = rec[3 * (int64_t)g + 1];
        const float4 r2 = rec[3 * (int64_t)g + 2];
        const float r_lat = marginal_radius(r1.y, r1.z, r1.w, r2.x, r2.y, r2.z, fr.axis_lat[0], fr.axis_lat[1], fr.axis_lat[2], sqrt_chi2);
        const float r_ax = marginal_radius(r1.y, r1.z, r1.w, r2.x, r2.y, r2.z, fr.axis_ax[0], fr.axis_ax[1], fr.axis_ax[2], sqrt_chi2);
        const float dx = x - r0.x, dy = y - r0.y, dz = z - r0.z;
        const float dl = fabsf(dot3(dx, dy, dz, fr.axis_lat[0], fr.axis_lat[1], fr.axis_lat[2]));
        const float da = fabsf(dot3(dx, dy, dz, fr.axis_ax[0], fr.axis_ax[1], fr.axis_ax[2]));
        if (dl <= r_lat && da <= r_ax) {
            if (pass) indices[w++] = g - g_begin;
            ++c;
        }
    }
    if (!pass) counts[p] = c;
}

static int check_settings(const us3d_ug_settings* s) {
    if (!s) return fail(US3D_ERR_ARGUMENT, "settings is NULL");
    if (!(s->tile_size == 8 || s->tile_size == 16 || s->tile_size == 32))
        return fail(US3D_ERR_ARGUMENT, "tile_size must be 8, 16 or 32");
    if (!(s->chi2_threshold > 0.0f)) return fail(US3D_ERR_ARGUMENT, "chi2_threshold must be > 0");
    if (!(s->diag_beta > 0.0f)) return fail(US3D_ERR_ARGUMENT, "diag_beta must be > 0");
    for (const int b : {s->fwd_lds_batch, s->bwd_lds_batch})
        if (!(b == 0 || b == 64 || b == 128 || b == 256))
            return fail(US3D_ERR_ARGUMENT, "fwd_lds_batch and bwd_lds_batch must be 0, 64, 128 or 256");
    if (!(s->work_queue == 0 || s->work_queue == 1)) return fail(US3D_ERR_ARGUMENT, "work_queue must be 0 or 1");
    return US3D_OK;
}

// Gaussians a splat kernel stages in LDS per batch: the settings' choice, or the kernel's thread count.
static int lds_batch(int chosen, int threads) { return chosen ? chosen : threads; }

template <typename T>
static T* at(const us3d_plan* plan, int64_t off) {
    return ws_at<T>(plan, off);
}

// ---------------------------------------------------------------- static scratch (no-allocation render)
// us3d_ug_render_static takes its workspace from this array instead of the caller. Every GPU the library runs on
// holds its own copy until the process ends. In host code `scratch` names a host-side stand-in, not a device
// address: the device address comes from hipGetSymbolAddress, once per device (static_scratch_ptr), and only that
// pointer is passed to kernels.
static __device__ __align__(256) unsigned char scratch[US3D_STATIC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;

static int static_scratch_ptr(void** out) {
    static std::mutex mu;
    static void* ptrs[kMaxDevices] = {nullptr};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return fail(US3D_ERR_HIP, "render_static: no current device");
    std::lock_guard<std::mutex> lock(mu);
    if (!ptrs[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(US3D_ERR_HIP, std::string("render_static: hipGetSymbolAddress(scratch) failed: ") +
                                          hipGetErrorString(e));
        ptrs[dev] = p;
    }
    *out = ptrs[dev];
    return US3D_OK;
}

// Visible-pair capacity of the static render for instance capacity k: every visible pair has an instance, so
// min(n_pairs, k) overflows only when the instances do.
static int64_t static_v_capacity(int64_t n_pairs, int64_t k) { return std::max<int64_t>(1, std::min(n_pairs, k)); }

// Largest instance capacity whose workspace fits `bytes` (0 if none does); the layout grows with k.
static int64_t static_k_capacity(int64_t n_gauss, int64_t n_pairs, int64_t n_tiles, int64_t bytes) {
    auto total = [&](int64_t k) { return ug_layout(n_gauss, n_pairs, n_tiles, k, static_v_capacity(n_pairs, k)).total; };
    if (total(1) > bytes) return 0;
    int64_t lo = 1, hi = (int64_t(1) << 31) - 1;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo + 1) / 2;
        if (total(mid) <= bytes) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ void copy_overflow_kernel(const us3d_stats* stats, int32_t* out) { *out = stats->overflow; }

}  // namespace us3d

using namespace us3d;

extern "C" {

int64_t us3d_ug_workspace_bytes(int64_t n_gauss, int64_t n_pairs, int64_t n_tiles, int64_t k_capacity,
                                int64_t v_capacity) {
    return ug_layout(n_gauss, n_pairs, n_tiles, k_capacity, v_capacity).total;
}

int us3d_ug_workspace_layout(int64_t n_gauss, int64_t n_pairs, int64_t n_tiles, int64_t k_capacity,
                             int64_t v_capacity, us3d_layout* out) {
    if (!out) return fail(US3D_ERR_ARGUMENT, "out is NULL");
    *out = ug_layout(n_gauss, n_pairs, n_tiles, k_capacity, v_capacity);
    return US3D_OK;
}

int us3d_ug_map(const us3d_plan* plan, const float* means, const float* tril, const float* opacity,
                const float* intensity, const us3d_ug_settings* settings, void* stream_) {
    const us3d_layout L = ug_layout(plan);
    int rc = check_plan(plan, L.total);
    if (rc) return rc;
    rc = check_settings(settings);
    if (rc) return rc;
    if (plan->n_gauss > 0 && (!means || !tril || !opacity || !intensity))
        return fail(US3D_ERR_ARGUMENT, "Gaussian parameter pointer is NULL");
    hipStream_t stream = as_stream(stream_);
    float4* rec = at<float4>(plan, L.records);
    const float sqrt_chi2 = sqrtf(settings->chi2_threshold);

    if ((rc = begin_pairs(plan, L, stream))) return rc;
    if (plan->n_gauss > 0) {
        hipLaunchKernelGGL(ug_pack_kernel, dim3(blocks_for(plan->n_gauss, 256)), dim3(256), 0, stream, plan->n_gauss,
                           means, tril, opacity, intensity, settings->diag_beta, sqrt_chi2,
                           settings->raw_activations ? 1 : 0, rec);
        if ((rc = check_launch("ug_pack_kernel"))) return rc;
    }
    const int64_t np = plan->n_pairs;
    if (np > 0) {
        {
            StageTimer timer_("ug_cull_kernel", stream);
            hipLaunchKernelGGL(ug_cull_kernel, dim3(blocks_for(np, 256)), dim3(256), 0, stream, np, rec, plan->frames,
                               plan->sweeps, plan->n_sweeps, sqrt_chi2, settings->tile_size,
                               at<uint64_t>(plan, L.pair_bits), at<int32_t>(plan, L.word_cnt),
                               at<uint32_t>(plan, L.fm_bits));
        }
        if ((rc = check_launch("ug_cull_kernel"))) return rc;
        if ((rc = rank_visible_pairs(plan, L, stream))) return rc;
        {
            StageTimer timer_("ug_setup_kernel", stream);
            hipLaunchKernelGGL(ug_setup_kernel, dim3(blocks_for(plan->v_capacity, 256)), dim3(256), 0, stream,
                               plan->v_capacity, at<int32_t>(plan, L.vis_pair), rec, plan->frames, plan->sweeps,
                               plan->n_sweeps, sqrt_chi2, settings->tile_size, settings->sort_by_depth,
                               at<us3d_stats>(plan, L.stats), at<int32_t>(plan, L.pair_count), at<int4>(plan, L.pair_rect),
                               at<uint32_t>(plan, L.pair_dk), at<float4>(plan, L.pair_sect),
                               at<int32_t>(plan, L.tile_count));
        }
        if ((rc = check_launch("ug_setup_kernel"))) return rc;
    }
    return build_instances(plan, L, stream, kSectF4, 3);
}

}  // extern "C"

namespace {

int ug_forward_impl(const us3d_plan* plan, const float* pixels, const us3d_ug_settings* settings, float* image,
                    float* alpha, float* depth, float* aux, const LossFuse& loss, void* stream_) {
    const us3d_layout L = ug_layout(plan);
    int rc = check_plan(plan, L.total);
    if (rc) return rc;
    if ((rc = check_settings(settings))) return rc;
    if (!image || !alpha || !aux) return fail(US3D_ERR_ARGUMENT, "forward: NULL array");
    hipStream_t stream = as_stream(stream_);
    const float sqrt_chi2 = sqrtf(settings->chi2_threshold);
    int32_t* heads = settings->work_queue ? at<int32_t>(plan, L.queue) : nullptr;
#define UG_FWD_B(TS, DEPTH, PLANE, B)                                                                                 \
    hipLaunchKernelGGL((ug_forward_kernel<TS, DEPTH, PLANE, B>),                                                      \
                       dim3(tile_grid((const void*)ug_forward_kernel<TS, DEPTH, PLANE, B>, FwdShape<TS>::kThreads,    \
                                      plan, heads, stream)),                                                          \
                       dim3(FwdShape<TS>::kThreads), 0, stream,                                                       \
                       at<float4>(plan, L.records), at<int32_t>(plan, L.tile_off), at<int32_t>(plan, L.pair_g),       \
                       at<uint64_t>(plan, L.seg_key), plan->frames, plan->frame_tile_base, plan->n_frames, pixels,     \
                       at<float4>(plan, L.pair_sect), settings->chi2_threshold, sqrt_chi2, settings->alpha_min,      \
                       settings->bg_alpha, settings->bg_intensity, settings->alpha_excludes_bg,                       \
                       at<us3d_stats>(plan, L.stats), image, alpha, depth,                                            \
                       reinterpret_cast<float4*>(aux), at<uint16_t>(plan, L.inst_mask), plan->n_tiles, heads, loss)
#define UG_FWD_(TS, DEPTH, PLANE)                                                                                     \
    do {                                                                                                              \
        switch (lds_batch(settings->fwd_lds_batch, FwdShape<TS>::kDefaultBatch)) {                                    \
            case 64: UG_FWD_B(TS, DEPTH, PLANE, 64); break;                                                           \
            case 128: UG_FWD_B(TS, DEPTH, PLANE, 128); break;                                                         \
            default: UG_FWD_B(TS, DEPTH, PLANE, 256); break;                                                          \
        }                                                                                                             \
    } while (0)
#define UG_FWD(TS, DEPTH)                                                                                             \
    do {                                                                                                              \
        if (pixels) UG_FWD_(TS, DEPTH, false); else UG_FWD_(TS, DEPTH, true);                                         \
    } while (0)
    {
        StageTimer timer_("ug_forward_kernel", stream);
        switch (settings->tile_size) {
            case 8: if (depth) UG_FWD(8, true); else UG_FWD(8, false); break;
            case 16: if (depth) UG_FWD(16, true); else UG_FWD(16, false); break;
            default: if (depth) UG_FWD(32, true); else UG_FWD(32, false); break;
        }
    }
#undef UG_FWD
#undef UG_FWD_
#undef UG_FWD_B
    return check_launch("ug_forward_kernel");
}

int ug_backward_impl(const us3d_plan* plan, const float* pixels, const float* tril, const us3d_ug_settings* settings,
                     const float* aux, const float* d_image, const float* d_alpha, const float* d_depth,
                     float* d_means, float* d_tril, float* d_opacity, float* d_intensity, const LossFuse& loss,
                     void* stream_) {
    const us3d_layout L = ug_layout(plan);
    int rc = check_plan(plan, L.total);
    if (rc) return rc;
    if ((rc = check_settings(settings))) return rc;
    if (!tril || !aux || !(d_image || loss.coeff) || !d_means || !d_tril || !d_opacity || !d_intensity)
        return fail(US3D_ERR_ARGUMENT, "backward: NULL array");
    hipStream_t stream = as_stream(stream_);
    int32_t* heads = settings->work_queue ? at<int32_t>(plan, L.queue) : nullptr;
#define UG_BWD_B(TS, FULL, PLANE, B)                                                                                  \
    hipLaunchKernelGGL((ug_backward_kernel<TS, FULL, PLANE, B>),                                                      \
                       dim3(tile_grid((const void*)ug_backward_kernel<TS, FULL, PLANE, B>, TileShape<TS>::kThreads,   \
                                      plan, heads, stream)),                                                          \
                       dim3(TileShape<TS>::kThreads), 0, stream, at<float4>(plan, L.records),                         \
                       at<int32_t>(plan, L.tile_off),                                                                 \
                       at<int32_t>(plan, L.pair_g), at<uint64_t>(plan, L.seg_key), plan->frames,                      \
                       plan->frame_tile_base, plan->n_frames, pixels,                                                 \
                       at<float4>(plan, L.pair_sect),                                                                 \
                       settings->chi2_threshold, settings->alpha_min, settings->bg_alpha,                             \
                       reinterpret_cast<const float4*>(aux), d_image, d_alpha, d_depth,                               \
                       at<us3d_stats>(plan, L.stats), at<float4>(plan, L.partial), at<uint16_t>(plan, L.inst_mask),   \
                       plan->n_tiles, heads, loss)
#define UG_BWD_(TS, FULL, PLANE)                                                                                      \
    do {                                                                                                              \
        switch (lds_batch(settings->bwd_lds_batch, TileShape<TS>::kThreads)) {                                                       \
            case 64: UG_BWD_B(TS, FULL, PLANE, 64); break;                                                            \
            case 128: UG_BWD_B(TS, FULL, PLANE, 128); break;                                                          \
            default: UG_BWD_B(TS, FULL, PLANE, 256); break;                                                           \
        }                                                                                                             \
    } while (0)
#define UG_BWD(TS, FULL)                                                                                              \
    do {                                                                                                              \
        if (pixels) UG_BWD_(TS, FULL, false); else UG_BWD_(TS, FULL, true);                                           \
    } while (0)
    const bool full = d_alpha || d_depth;
    {
        StageTimer timer_("ug_backward_kernel", stream);
        switch (settings->tile_size) {
            case 8: if (full) UG_BWD(8, true); else UG_BWD(8, false); break;
            case 16: if (full) UG_BWD(16, true); else UG_BWD(16, false); break;
            default: if (full) UG_BWD(32, true); else UG_BWD(32, false); break;
        }
    }
#undef UG_BWD
#undef UG_BWD_
#undef UG_BWD_B
    if ((rc = check_launch("ug_backward_kernel"))) return rc;
    if (plan->n_gauss > 0) {
        {
            StageTimer timer_("ug_reduce_kernel", stream);
            hipLaunchKernelGGL(ug_reduce_kernel, dim3(blocks_for(plan->n_gauss, 4)), dim3(256), 0, stream, plan->n_gauss,
                               plan->sweeps, plan->n_sweeps, at<uint64_t>(plan, L.pair_bits),
                               at<int32_t>(plan, L.word_off), at<int32_t>(plan, L.pair_off), at<float4>(plan, L.partial),
                               tril, at<float4>(plan, L.records), settings->raw_activations ? 1 : 0,
                               at<us3d_stats>(plan, L.stats), d_means, d_tril, d_opacity, d_intensity);
        }
        if ((rc = check_launch("ug_reduce_kernel"))) return rc;
    }
    return US3D_OK;
}

// Per-(tile, wavefront) loss sums of a fused forward: one wavefront per tile below 32-pixel tiles, four at 32.
int64_t loss_part_entries(const us3d_plan* plan) { return (int64_t)plan->n_tiles * FwdShape<32>::kWaves; }

}  // namespace

extern "C" {

int us3d_ug_forward(const us3d_plan* plan, const float* pixels, const us3d_ug_settings* settings, float* image,
                    float* alpha, float* depth, float* aux, void* stream_) {
    return ug_forward_impl(plan, pixels, settings, image, alpha, depth, aux, LossFuse{nullptr, nullptr, nullptr},
                           stream_);
}

int us3d_ug_backward(const us3d_plan* plan, const float* pixels, const float* tril, const us3d_ug_settings* settings,
                     const float* aux, const float* d_image, const float* d_alpha, const float* d_depth,
                     float* d_means, float* d_tril, float* d_opacity, float* d_intensity, void* stream_) {
    if (!d_image) return fail(US3D_ERR_ARGUMENT, "backward: NULL array");
    return ug_backward_impl(plan, pixels, tril, settings, aux, d_image, d_alpha, d_depth, d_means, d_tril, d_opacity,
                            d_intensity, LossFuse{nullptr, nullptr, nullptr}, stream_);
}

int64_t us3d_ug_loss_tmp_bytes(const us3d_plan* plan) {
    if (!plan) return 0;
    return (loss_part_entries(plan) * kLossSums + 4 * (int64_t)plan->n_frames) * (int64_t)sizeof(double);
}

int us3d_ug_forward_loss(const us3d_plan* plan, const float* pixels, const us3d_ug_settings* settings, float* image,
                         float* alpha, float* depth, float* aux, const float* gt, float w_l1, float w_ssim,
                         float scale, float* loss, double* tmp, void* stream_) {
    if (!plan || !settings || !gt || !loss || !tmp) return fail(US3D_ERR_ARGUMENT, "forward_loss: NULL argument");
    int rc = ug_forward_impl(plan, pixels, settings, image, alpha, depth, aux, LossFuse{gt, tmp, nullptr}, stream_);
    if (rc) return rc;
    const int waves = settings->tile_size == 32 ? FwdShape<32>::kWaves
                                                : (settings->tile_size == 16 ? FwdShape<16>::kWaves : FwdShape<8>::kWaves);
    return image_loss_from_tiles(tmp, plan->frame_tile_base, plan->frame_pixel_base, plan->n_frames, waves, w_l1,
                                 w_ssim, scale, loss, tmp + loss_part_entries(plan) * kLossSums, as_stream(stream_));
}

int us3d_ug_backward_loss(const us3d_plan* plan, const float* pixels, const float* tril,
                          const us3d_ug_settings* settings, const float* aux, const float* gt, const double* tmp,
                          float* d_means, float* d_tril, float* d_opacity, float* d_intensity, void* stream_) {
    if (!plan || !gt || !tmp) return fail(US3D_ERR_ARGUMENT, "backward_loss: NULL argument");
    return ug_backward_impl(plan, pixels, tril, settings, aux, nullptr, nullptr, nullptr, d_means, d_tril, d_opacity,
                            d_intensity, LossFuse{gt, nullptr, tmp + loss_part_entries(plan) * kLossSums}, stream_);
}

int us3d_ug_pixel_lists(const us3d_plan* plan, const float* pixels, const us3d_ug_settings* settings, int pass,
                        int32_t* counts, const int32_t* offsets, int32_t* indices, void* stream_) {
    const us3d_layout L = ug_layout(plan);
    int rc = check_plan(plan, L.total);
    if (rc) return rc;
    if ((rc = check_settings(settings))) return rc;
    if ((pass == 0 && !counts) || (pass != 0 && (!offsets || !indices)))
        return fail(US3D_ERR_ARGUMENT, "pixel_lists: NULL array");
    if (!plan->frame_pixel_base) return fail(US3D_ERR_ARGUMENT, "pixel_lists: frame_pixel_base is NULL");
    const int32_t* frame_pixel_base = plan->frame_pixel_base;
    hipStream_t stream = as_stream(stream_);
    {
        StageTimer timer_("ug_pixel_lists_kernel", stream);
        hipLaunchKernelGGL(ug_pixel_lists_kernel, dim3(blocks_for(plan->n_pixels, 256)), dim3(256), 0, stream,
                           plan->n_pixels, plan->frames, plan->n_frames, frame_pixel_base, plan->sweeps,
                           at<float4>(plan, L.records), at<int32_t>(plan, L.tile_off), at<uint64_t>(plan, L.seg_key),
                           at<int32_t>(plan, L.pair_g), pixels,
                           sqrtf(settings->chi2_threshold), settings->tile_size, pass != 0, counts, offsets, indices,
                           at<us3d_stats>(plan, L.stats));
    }
    return check_launch("ug_pixel_lists_kernel");
}

int us3d_ug_render_static(const us3d_plan* plan_in, const float* means, const float* tril, const float* opacity,
                          const float* intensity, const float* pixels, const us3d_ug_settings* settings, float* image,
                          float* alpha, float* depth, float* aux, int32_t* overflow_dev, void* stream_) {
    if (!plan_in) return fail(US3D_ERR_ARGUMENT, "render_static: plan is NULL");
    us3d_plan plan = *plan_in;
    const int64_t k_cap = static_k_capacity(plan.n_gauss, plan.n_pairs, plan.n_tiles, US3D_STATIC_SCRATCH_BYTES);
    if (k_cap < 1)
        return fail(US3D_ERR_CAPACITY, "render_static: the batch's workspace does not fit the " +
                                           std::to_string(US3D_STATIC_SCRATCH_BYTES >> 20) +
                                           " MiB static scratch (use us3d_ug_map / us3d_ug_forward with a workspace)");
    void* ws = nullptr;
    int rc = static_scratch_ptr(&ws);
    if (rc) return rc;
    plan.workspace = ws;
    plan.workspace_bytes = US3D_STATIC_SCRATCH_BYTES;
    plan.k_capacity = k_cap;
    plan.v_capacity = static_v_capacity(plan.n_pairs, k_cap);
    if ((rc = us3d_ug_map(&plan, means, tril, opacity, intensity, settings, stream_))) return rc;
    if ((rc = us3d_ug_forward(&plan, pixels, settings, image, alpha, depth, aux, stream_))) return rc;
    if (overflow_dev) {
        const us3d_layout L = ug_layout(&plan);
        hipLaunchKernelGGL(copy_overflow_kernel, dim3(1), dim3(1), 0, as_stream(stream_), at<us3d_stats>(&plan, L.stats),
                           overflow_dev);
        if ((rc = check_launch("copy_overflow_kernel"))) return rc;
    }
    return US3D_OK;
}

}  // extern "C"
