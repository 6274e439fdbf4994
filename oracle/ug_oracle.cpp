// TEST INFRASTRUCTURE ONLY -- the oracle. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load this library, and only as the checker / CPU baseline; the product never calls it.
//
// CPU restatement of the UltraGauss slice renderer of the reference, one frame at a time, following:
//   mapper   ops/ultragauss_renderer/src/mapper/core.cu:9-164 and src/mapper/wrapper_mapper.cu:35-206
//   pixels   ops/ultragauss_renderer/src/mapper/core.cu:166-287
//   forward  ops/ultragauss_renderer/src/shader/forward.cu:17-76 (+ include/shader/shared.cuh:4-52)
//   backward ops/ultragauss_renderer/src/shader/backward.cu:17-131
//
// FP32 arithmetic is canonical: the reference was built with nvcc --use_fast_math, so its own rounding is
// unspecified; this restatement fixes one IEEE order of operations (explicit fmaf, correctly rounded
// division and sqrt, no contraction -- build with -ffp-contract=off) so tile rectangles, keys and the
// q <= chi2 decisions can be compared bit for bit. Every template also runs in double as the FP64 oracle.
//
// Deviations, defined here because the reference leaves them undefined:
//   - Gaussians whose projection or rectangle bounds are not finite (NaN/Inf parameters) are culled;
//   - the depth key (uint32)((v_pix + 500) * 1000) maps negative/NaN to 0 and >= 2^32 to 0xffffffff.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

namespace {

template <typename T> inline T fma_(T a, T b, T c);
template <> inline float fma_<float>(float a, float b, float c) { return std::fmaf(a, b, c); }
template <> inline double fma_<double>(double a, double b, double c) { return std::fma(a, b, c); }

template <typename T> inline T dot3(T ax, T ay, T az, T bx, T by, T bz) { return fma_(az, bz, fma_(ay, by, ax * bx)); }

// compute_probe_bbox_radii, core.cu:9-44: y = L^-1 a by forward substitution, r = max(1e-3, sqrt(chi2)*|y|).
template <typename T>
T marginal_radius(const T l[6], const float a[3], T sqrt_chi2) {
    const T y0 = T(a[0]) / std::max(l[0], T(1e-10f));
    const T y1 = fma_(-l[1], y0, T(a[1])) / std::max(l[3], T(1e-10f));
    const T y2 = fma_(-l[4], y1, fma_(-l[2], y0, T(a[2]))) / std::max(l[5], T(1e-10f));
    const T s = std::max(fma_(y2, y2, fma_(y1, y1, y0 * y0)), T(1e-12f));
    return std::max(T(1e-3f), sqrt_chi2 * std::sqrt(s));
}

// l11, l12, l13, l22, l23, l33 with l_ii = raw_ii^2 + beta (core.cu:20-25, shared.cuh:9-15).
template <typename T, typename In>
void tril_factor(const In* raw, T beta, T l[6]) {
    l[0] = fma_(T(raw[0]), T(raw[0]), beta);
    l[1] = T(raw[1]);
    l[2] = T(raw[2]);
    l[3] = fma_(T(raw[3]), T(raw[3]), beta);
    l[4] = T(raw[4]);
    l[5] = fma_(T(raw[5]), T(raw[5]), beta);
}

inline uint32_t depth_key(float v_pix) {
    const float x = (v_pix + 500.0f) * 1000.0f;
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)x;
}

inline int clamp_to_int(float x) { return (int)std::min(std::max(x, -1073741824.0f), 1073741824.0f); }

// q = ||L^T d||^2 = d^T (L L^T) d, with y = L^T d.
template <typename T>
T quad_form(const T l[6], T dx, T dy, T dz, T y[3]) {
    y[0] = fma_(l[2], dz, fma_(l[1], dy, l[0] * dx));
    y[1] = fma_(l[4], dz, l[3] * dy);
    y[2] = l[5] * dz;
    return fma_(y[2], y[2], fma_(y[1], y[1], y[0] * y[0]));
}

struct Frame {
    float o[3], lat[3], ax[3], ele[3];
};

Frame make_frame(const float* origin, const float* a_lat, const float* a_ax, const float* a_ele) {
    Frame f;
    std::memcpy(f.o, origin, 12);
    std::memcpy(f.lat, a_lat, 12);
    std::memcpy(f.ax, a_ax, 12);
    std::memcpy(f.ele, a_ele, 12);
    return f;
}

template <typename T>
T exp_(T x) { return std::exp(x); }

// The q <= chi2 cut evaluated in the canonical FP32 arithmetic on the FP32 values of the inputs. The
// pair set is part of the renderer's semantics: the FP64 evaluation uses the same pairs and only measures the
// arithmetic error of the values (deciding the cut in double would flip pairs that lie within rounding of
// the chi2 boundary).
template <typename In>
bool keep_pair(const In* raw, float beta, In px, In py, In pz, In mx, In my, In mz, float chi2) {
    float l[6];
    const float r32[6] = {(float)raw[0], (float)raw[1], (float)raw[2], (float)raw[3], (float)raw[4], (float)raw[5]};
    tril_factor<float>(r32, beta, l);
    float y[3];
    const float q = quad_form<float>(l, (float)px - (float)mx, (float)py - (float)my, (float)pz - (float)mz, y);
    return !(q > chi2);
}

// Plane form of q for a frame's own linear-probe pixel grid, restating the kernels' pair setup (ug.hip,
// pair_plane_form: FP64 from the FP32 record values) and evaluation (plane_q: FP32, fixed order). The pixel at probe
// coordinates (l, a) = (lat_g + x, ax_g + y) has L^T (p - mu) = c0 + x A + y B (A = L^T a_lat, B = L^T a_ax, c0 =
// L^T g0, g0 = (o - mu) + lat_g a_lat + ax_g a_ax), so q = (p (l - cx) + qq (a - cy))^2 + (s (a - cy))^2 + K about
// its minimum. This is the renderer's canonical FP32 q for frames rendered on their own grid; explicit pixel
// positions keep q = |L^T (p - mu)|^2 (quad_form).
struct Plane {
    float cx, cy, K, p, qq, s;
};

Plane plane_form(const float* mu, const float l[6], const Frame& fr) {
    const float vx = mu[0] - fr.o[0], vy = mu[1] - fr.o[1], vz = mu[2] - fr.o[2];
    const float lat_g = dot3(vx, vy, vz, fr.lat[0], fr.lat[1], fr.lat[2]);
    const float ax_g = dot3(vx, vy, vz, fr.ax[0], fr.ax[1], fr.ax[2]);
    const double L11 = l[0], L12 = l[1], L13 = l[2], L22 = l[3], L23 = l[4], L33 = l[5];
    const float* a = fr.lat;
    const float* c = fr.ax;
    const double A0 = a[0] * L11 + a[1] * L12 + a[2] * L13, A1 = a[1] * L22 + a[2] * L23, A2 = a[2] * L33;
    const double B0 = c[0] * L11 + c[1] * L12 + c[2] * L13, B1 = c[1] * L22 + c[2] * L23, B2 = c[2] * L33;
    const double M00 = A0 * A0 + A1 * A1 + A2 * A2;
    const double M01 = A0 * B0 + A1 * B1 + A2 * B2;
    const double M11 = B0 * B0 + B1 * B1 + B2 * B2;
    const double lg = lat_g, ag = ax_g;
    const double g0 = ((double)fr.o[0] - (double)mu[0]) + lg * a[0] + ag * c[0];
    const double g1 = ((double)fr.o[1] - (double)mu[1]) + lg * a[1] + ag * c[1];
    const double g2 = ((double)fr.o[2] - (double)mu[2]) + lg * a[2] + ag * c[2];
    const double c0 = L11 * g0 + L12 * g1 + L13 * g2, c1 = L22 * g1 + L23 * g2, c2 = L33 * g2;
    const double h0 = A0 * c0 + A1 * c1 + A2 * c2, h1 = B0 * c0 + B1 * c1 + B2 * c2;
    const double cc = c0 * c0 + c1 * c1 + c2 * c2;
    const double det = M00 * M11 - M01 * M01;
    double xs = 0.0, ys = 0.0, K = cc, pp = std::sqrt(std::max(M00, 0.0)), qq = 0.0, ss = 0.0;
    if (M00 > 0.0 && det > 0.0) {
        xs = -(M11 * h0 - M01 * h1) / det;
        ys = -(M00 * h1 - M01 * h0) / det;
        K = std::max(0.0, cc + h0 * xs + h1 * ys);
        qq = M01 / pp;
        ss = std::sqrt(det / M00);
    } else if (M00 > 0.0) {
        qq = M01 / pp;
    }
    return Plane{(float)(lg + xs), (float)(ag + ys), (float)K, (float)pp, (float)qq, (float)ss};
}

// The pixel grid of a frame rendered on its own grid: pixel index p -> probe coordinates, as the kernels form them.
struct Grid {
    int W;
    float col_offset, dlat, near_mm, dax;
};

inline void grid_lat_ax(const Grid& gr, int64_t p, float& lat, float& ax) {
    const int ix = (int)(p % gr.W), iy = (int)(p / gr.W);
    lat = ((float)ix - gr.col_offset) * gr.dlat;
    ax = gr.near_mm + (float)iy * gr.dax;
}

// The kernels evaluate the plane form scaled by log2(e)/2 (ug.hip, the splat kernels' staging and plane_q_x):
// p, qq, s times kQScaleSqrt and K times kQScale in FP32, q' = (p' x + qq' y)^2 + (s' y)^2 + K', the test q' <=
// chi2 * kQScale and the weight exp2(-q').
constexpr float kQScale = 0.721347510814666748046875f;      // log2(e)/2 rounded to FP32
constexpr float kQScaleSqrt = 0.849321782588958740234375f;  // the FP32 square root of kQScale

inline float plane_q(const Plane& P, float lat, float ax) {
    const float x = lat - P.cx, y = ax - P.cy;
    const float ps = P.p * kQScaleSqrt, qqs = P.qq * kQScaleSqrt, ss = P.s * kQScaleSqrt, Ks = P.K * kQScale;
    const float u = std::fmaf(ps, x, qqs * y);
    const float v = ss * y;
    return std::fmaf(u, u, std::fmaf(v, v, Ks));
}

}  // namespace

extern "C" {

// The plane form (cx, cy, K, p, qq, s) of every Gaussian for one frame (out[6 G]).
int ug_oracle_plane_form(int G, const float* means, const float* tril, const float* origin, const float* a_lat,
                         const float* a_ax, const float* a_ele, float beta, float* out) {
    const Frame fr = make_frame(origin, a_lat, a_ax, a_ele);
    for (int g = 0; g < G; ++g) {
        float l[6];
        tril_factor<float>(tril + 6 * (int64_t)g, beta, l);
        const Plane P = plane_form(means + 3 * (int64_t)g, l, fr);
        const float v[6] = {P.cx, P.cy, P.K, P.p, P.qq, P.s};
        std::memcpy(out + 6 * (int64_t)g, v, sizeof(v));
    }
    return 0;
}

// kernel_project_gauss_and_compute_tile_rects (core.cu:46-116) for every Gaussian of one frame.
// Outputs count[G], rect[G*4] (min_tx, min_ty, max_tx, max_ty; zeros when culled), dk[G].
int ug_oracle_preprocess(int G, const float* means, const float* tril, const float* origin, const float* a_lat,
                         const float* a_ax, const float* a_ele, int H, int W, float dlat, float dax, float near_mm,
                         float col_offset, float chi2, float beta, int sort_by_depth, int ts, int32_t* count,
                         int32_t* rect, uint32_t* dk) {
    const Frame fr = make_frame(origin, a_lat, a_ax, a_ele);
    const float sq = std::sqrt(chi2);
    const int tiles_x = (W + ts - 1) / ts, tiles_y = (H + ts - 1) / ts;
    for (int g = 0; g < G; ++g) {
        float l[6];
        tril_factor<float>(tril + 6 * (int64_t)g, beta, l);
        const float vx = means[3 * g] - fr.o[0], vy = means[3 * g + 1] - fr.o[1], vz = means[3 * g + 2] - fr.o[2];
        const float lat = dot3(vx, vy, vz, fr.lat[0], fr.lat[1], fr.lat[2]);
        const float ele = dot3(vx, vy, vz, fr.ele[0], fr.ele[1], fr.ele[2]);
        const float ax = dot3(vx, vy, vz, fr.ax[0], fr.ax[1], fr.ax[2]);
        const float r_lat = marginal_radius<float>(l, fr.lat, sq);
        const float r_ax = marginal_radius<float>(l, fr.ax, sq);
        const float r_ele = marginal_radius<float>(l, fr.ele, sq);
        count[g] = 0;
        dk[g] = 0;
        for (int k = 0; k < 4; ++k) rect[4 * g + k] = 0;
        if (!(std::fabs(ele) <= r_ele)) continue;
        const float u = lat / dlat + col_offset;
        const float v = (ax - near_mm) / dax;
        const float ru = std::max(1.0f, r_lat / dlat);
        const float rv = std::max(1.0f, r_ax / dax);
        const float tsf = (float)ts;
        const float fx0 = std::floor((u - ru) / tsf), fy0 = std::floor((v - rv) / tsf);
        const float fx1 = std::ceil((u + ru) / tsf), fy1 = std::ceil((v + rv) / tsf);
        if (!(std::isfinite(fx0) && std::isfinite(fy0) && std::isfinite(fx1) && std::isfinite(fy1))) continue;
        const int x0 = std::max(0, clamp_to_int(fx0)), y0 = std::max(0, clamp_to_int(fy0));
        const int x1 = std::min(tiles_x, clamp_to_int(fx1)), y1 = std::min(tiles_y, clamp_to_int(fy1));
        if (!(x0 < x1 && y0 < y1)) continue;
        count[g] = (x1 - x0) * (y1 - y0);
        rect[4 * g + 0] = x0;
        rect[4 * g + 1] = y0;
        rect[4 * g + 2] = x1;
        rect[4 * g + 3] = y1;
        dk[g] = sort_by_depth ? depth_key(v) : 0u;
    }
    return 0;
}

// Emit (core.cu:118-143), stable LSD radix sort on bits [begin_bit, 64) (wrapper_mapper.cu:134-146) and
// tile ranges (core.cu:145-164). keys/vals have K = sum(count) entries; start/end have tiles entries.
// The sort is an 8-bit LSD radix sort with counting passes -- stable by construction, like CUB's.
int ug_oracle_emit_sort(int G, const int32_t* count, const int32_t* rect, const uint32_t* dk, int H, int W, int ts,
                        int sort_by_depth, uint64_t* keys, int32_t* vals, int32_t* tile_start, int32_t* tile_end) {
    const int tiles_x = (W + ts - 1) / ts, tiles_y = (H + ts - 1) / ts;
    int64_t e = 0;
    for (int g = 0; g < G; ++g) {
        if (count[g] <= 0) continue;
        for (int y = rect[4 * g + 1]; y < rect[4 * g + 3]; ++y)
            for (int x = rect[4 * g + 0]; x < rect[4 * g + 2]; ++x) {
                const uint64_t tile = (uint64_t)(y * tiles_x + x);
                keys[e] = (tile << 32) | (uint64_t)(sort_by_depth ? dk[g] : 0u);
                vals[e] = g;
                ++e;
            }
    }
    const int64_t K = e;
    std::vector<uint64_t> k2(K);
    std::vector<int32_t> v2(K);
    const int begin_bit = sort_by_depth ? 0 : 32;
    for (int shift = begin_bit; shift < 64; shift += 8) {
        int64_t hist[257] = {0};
        for (int64_t i = 0; i < K; ++i) hist[((keys[i] >> shift) & 0xff) + 1]++;
        for (int b = 0; b < 256; ++b) hist[b + 1] += hist[b];
        for (int64_t i = 0; i < K; ++i) {
            const int64_t pos = hist[(keys[i] >> shift) & 0xff]++;
            k2[pos] = keys[i];
            v2[pos] = vals[i];
        }
        std::copy(k2.begin(), k2.end(), keys);
        std::copy(v2.begin(), v2.end(), vals);
    }
    const int n_tiles = tiles_x * tiles_y;
    for (int t = 0; t < n_tiles; ++t) tile_start[t] = tile_end[t] = -1;
    for (int64_t i = 0; i < K; ++i) {
        const uint32_t cur = (uint32_t)(keys[i] >> 32);
        if (i == 0) {
            tile_start[cur] = 0;
        } else {
            const uint32_t prev = (uint32_t)(keys[i - 1] >> 32);
            if (cur != prev) {
                tile_end[prev] = (int32_t)i;
                tile_start[cur] = (int32_t)i;
            }
        }
        if (i == K - 1) tile_end[cur] = (int32_t)K;
    }
    return 0;
}

// Per-pixel candidate window test (core.cu:166-287). pass 0 writes counts[P]; pass 1 fills indices using
// offsets[P]. Pixel (x, y) of the H x W frame is pixels[(y*W + x)*3 ..].
int ug_oracle_pixel_lists(int G, const float* means, const float* tril, const float* pixels, int H, int W,
                          const float* a_lat, const float* a_ax, float chi2, float beta, int ts,
                          const int32_t* sorted_vals, const int32_t* tile_start, const int32_t* tile_end, int pass,
                          int32_t* counts, const int32_t* offsets, int32_t* indices) {
    (void)G;
    const float sq = std::sqrt(chi2);
    float lat[3], ax[3];
    std::memcpy(lat, a_lat, 12);
    std::memcpy(ax, a_ax, 12);
    const int tiles_x = (W + ts - 1) / ts;
    for (int py = 0; py < H; ++py)
        for (int px = 0; px < W; ++px) {
            const int p = py * W + px;
            const int t = (py / ts) * tiles_x + px / ts;
            const int s = tile_start[t], e = tile_end[t];
            int c = 0;
            int w = pass ? offsets[p] : 0;
            if (s != -1 && s < e) {
                for (int i = s; i < e; ++i) {
                    const int g = sorted_vals[i];
                    float l[6];
                    tril_factor<float>(tril + 6 * (int64_t)g, beta, l);
                    const float r_lat = marginal_radius<float>(l, lat, sq);
                    const float r_ax = marginal_radius<float>(l, ax, sq);
                    const float dx = pixels[3 * p] - means[3 * g];
                    const float dy = pixels[3 * p + 1] - means[3 * g + 1];
                    const float dz = pixels[3 * p + 2] - means[3 * g + 2];
                    const float dl = std::fabs(dot3(dx, dy, dz, lat[0], lat[1], lat[2]));
                    const float da = std::fabs(dot3(dx, dy, dz, ax[0], ax[1], ax[2]));
                    if (dl <= r_lat && da <= r_ax) {
                        if (pass) indices[w++] = g;
                        ++c;
                    }
                }
            }
            if (!pass) counts[p] = c;
        }
    return 0;
}

}  // extern "C"

namespace {

// Forward (forward.cu:17-76) for one frame over per-pixel candidate lists.
// With `plane` (one Plane per Gaussian) the frame is rendered on its own grid `grid`: the decisions are the scaled
// plane form's (q' <= chi2 kQScale), and so is the FP32 weight exp2(-q'); the FP64 weight is exp(-q/2) with q =
// |L^T (p - mu)|^2 from the given pixels.
template <typename T>
void forward_impl(int G, const T* means, const T* tril, const T* opac, const T* inten,
                  const T* pixels, int64_t P, const float* origin, const float* a_ele, const int32_t* counts,
                  const int32_t* offsets, const int32_t* indices, float chi2, float beta, float alpha_min, float bg,
                  float bgI, int excl, T* image, T* alpha, T* depth, T* sums, const Plane* plane, Grid grid) {
    (void)G;
    for (int64_t p = 0; p < P; ++p) {
        T S = 0, N = 0, D = 0;
        const T px = pixels[3 * p], py = pixels[3 * p + 1], pz = pixels[3 * p + 2];
        float glat = 0.f, gax = 0.f;
        if (plane) grid_lat_ax(grid, p, glat, gax);
        for (int k = 0; k < counts[p]; ++k) {
            const int g = indices[offsets[p] + k];
            T l[6];
            tril_factor<T>(tril + 6 * (int64_t)g, T(beta), l);
            const T mx = means[3 * g], my = means[3 * g + 1], mz = means[3 * g + 2];
            T y[3];
            T q = quad_form<T>(l, px - mx, py - my, pz - mz, y);
            T w = 0;
            if (plane) {
                const float q32 = plane_q(plane[g], glat, gax);  // q' = q log2(e)/2
                if (q32 > chi2 * kQScale) continue;
                w = (sizeof(T) == sizeof(float)) ? T(std::exp2(-q32)) : exp_(T(-0.5) * q);
            } else if (!keep_pair<T>(tril + 6 * (int64_t)g, beta, px, py, pz, mx, my, mz, chi2)) {
                continue;
            } else {
                w = exp_(T(-0.5) * q);
            }
            const T ah = T(opac[g]) * w;
            if (!(ah > T(alpha_min))) continue;
            const T e = dot3<T>(mx - T(origin[0]), my - T(origin[1]), mz - T(origin[2]), T(a_ele[0]), T(a_ele[1]), T(a_ele[2]));
            S += ah;
            N = fma_(ah, T(inten[g]), N);
            D = fma_(ah, e, D);
        }
        const T A = S + T(bg);
        const T num = fma_(T(bg), T(bgI), N);
        const T img = num / std::max(A, T(1e-8f));
        image[p] = std::max(T(0), std::min(T(1), img));
        alpha[p] = excl ? S : A;
        depth[p] = D / std::max(S, T(1e-8f));
        if (sums) {
            sums[3 * p] = S;
            sums[3 * p + 1] = N;
            sums[3 * p + 2] = D;
        }
    }
}

// Backward (backward.cu:17-131): the reference's 9-entry Lambda expansion and its dl formulas
// (backward.cu:106-129), recomputing all per-pixel sums. gate = 1 applies the forward's
// alpha_hat > alpha_min test to pairs (the exact adjoint); gate = 0 reproduces the reference, whose
// backward omits that test.
template <typename T>
void backward_impl(int G, const T* means, const T* tril, const T* opac, const T* inten,
                   const T* pixels, int64_t P, const float* origin, const float* a_ele, const int32_t* counts,
                   const int32_t* offsets, const int32_t* indices, float chi2, float beta, float alpha_min, float bg,
                   float bgI, int gate, const float* gI, const float* gA, const float* gD, T* d_means, T* d_tril,
                   T* d_opac, T* d_inten, const Plane* plane, Grid grid) {
    std::fill(d_means, d_means + 3 * (int64_t)G, T(0));
    std::fill(d_tril, d_tril + 6 * (int64_t)G, T(0));
    std::fill(d_opac, d_opac + G, T(0));
    std::fill(d_inten, d_inten + G, T(0));
    const T ex = a_ele[0], ey = a_ele[1], ez = a_ele[2];
    for (int64_t p = 0; p < P; ++p) {
        const T px = pixels[3 * p], py = pixels[3 * p + 1], pz = pixels[3 * p + 2];
        float glat = 0.f, gax = 0.f;
        if (plane) grid_lat_ax(grid, p, glat, gax);
        T S = 0, N = 0, D = 0;
        auto weight = [&](int g, T& w, T& dx, T& dy, T& dz, T lam[9], T l[6]) -> bool {
            tril_factor<T>(tril + 6 * (int64_t)g, T(beta), l);
            lam[0] = l[0] * l[0];
            lam[1] = l[0] * l[1];
            lam[2] = l[0] * l[2];
            lam[3] = lam[1];
            lam[4] = l[1] * l[1] + l[3] * l[3];
            lam[5] = l[1] * l[2] + l[3] * l[4];
            lam[6] = lam[2];
            lam[7] = lam[5];
            lam[8] = l[2] * l[2] + l[4] * l[4] + l[5] * l[5];
            dx = px - T(means[3 * g]);
            dy = py - T(means[3 * g + 1]);
            dz = pz - T(means[3 * g + 2]);
            T y[3];
            T q = quad_form<T>(l, dx, dy, dz, y);
            if (plane) {
                const float q32 = plane_q(plane[g], glat, gax);  // q' = q log2(e)/2
                if (q32 > chi2 * kQScale) return false;
                w = (sizeof(T) == sizeof(float)) ? T(std::exp2(-q32)) : exp_(T(-0.5) * q);
                return true;
            }
            if (!keep_pair<T>(tril + 6 * (int64_t)g, beta, px, py, pz, T(means[3 * g]), T(means[3 * g + 1]),
                              T(means[3 * g + 2]), chi2)) {
                return false;
            }
            w = exp_(T(-0.5) * q);
            return true;
        };
        for (int k = 0; k < counts[p]; ++k) {
            const int g = indices[offsets[p] + k];
            T w, dx, dy, dz, lam[9], l[6];
            if (!weight(g, w, dx, dy, dz, lam, l)) continue;
            const T ah = T(opac[g]) * w;
            if (!(ah > T(alpha_min))) continue;
            const T e = (T(means[3 * g]) - T(origin[0])) * ex + (T(means[3 * g + 1]) - T(origin[1])) * ey +
                        (T(means[3 * g + 2]) - T(origin[2])) * ez;
            S += ah;
            N += ah * T(inten[g]);
            D += ah * e;
        }
        const T A = S + T(bg);
        const T inv = T(1) / std::max(A, T(1e-8f));
        const T num = N + T(bg) * T(bgI);
        const T dN = T(gI[p]) * inv;
        const T dA = (gA ? T(gA[p]) : T(0)) - T(gI[p]) * num * inv * inv;
        const T gd = gD ? T(gD[p]) : T(0);
        const T invS = T(1) / std::max(S, T(1e-8f));
        const T dDw = gd * invS;
        const T dSd = -gd * D * invS * invS;
        for (int k = 0; k < counts[p]; ++k) {
            const int g = indices[offsets[p] + k];
            T w, dx, dy, dz, lam[9], l[6];
            if (!weight(g, w, dx, dy, dz, lam, l)) continue;
            const T o = opac[g], c = inten[g];
            const T ah = o * w;
            if (gate && !(ah > T(alpha_min))) continue;
            const T e = (T(means[3 * g]) - T(origin[0])) * ex + (T(means[3 * g + 1]) - T(origin[1])) * ey +
                        (T(means[3 * g + 2]) - T(origin[2])) * ez;
            const T dah = dN * c + dA + dSd + dDw * e;
            d_inten[g] += dN * ah;
            d_opac[g] += dah * w;
            const T dq = dah * o * w * T(-0.5);
            const T v0 = lam[0] * dx + lam[1] * dy + lam[2] * dz;
            const T v1 = lam[3] * dx + lam[4] * dy + lam[5] * dz;
            const T v2 = lam[6] * dx + lam[7] * dy + lam[8] * dz;
            d_means[3 * g + 0] += dq * (T(-2) * v0) + dDw * ah * ex;
            d_means[3 * g + 1] += dq * (T(-2) * v1) + dDw * ah * ey;
            d_means[3 * g + 2] += dq * (T(-2) * v2) + dDw * ah * ez;
            T dS[9] = {dq * dx * dx, dq * dx * dy, dq * dx * dz, dq * dy * dx, dq * dy * dy,
                       dq * dy * dz, dq * dz * dx, dq * dz * dy, dq * dz * dz};
            const T dl11 = dS[0] * (T(2) * l[0]) + (dS[1] + dS[3]) * l[1] + (dS[2] + dS[6]) * l[2];
            const T dl12 = (dS[1] + dS[3]) * l[0] + dS[4] * (T(2) * l[1]) + (dS[5] + dS[7]) * l[2];
            const T dl13 = (dS[2] + dS[6]) * l[0] + (dS[5] + dS[7]) * l[1] + dS[8] * (T(2) * l[2]);
            const T dl22 = dS[4] * (T(2) * l[3]) + (dS[5] + dS[7]) * l[4];
            const T dl23 = (dS[5] + dS[7]) * l[3] + dS[8] * (T(2) * l[4]);
            const T dl33 = dS[8] * (T(2) * l[5]);
            const T* raw = tril + 6 * (int64_t)g;
            d_tril[6 * g + 0] += dl11 * (T(2) * T(raw[0]));
            d_tril[6 * g + 1] += dl12;
            d_tril[6 * g + 2] += dl13;
            d_tril[6 * g + 3] += dl22 * (T(2) * T(raw[3]));
            d_tril[6 * g + 4] += dl23;
            d_tril[6 * g + 5] += dl33 * (T(2) * T(raw[5]));
        }
    }
}

}  // namespace

extern "C" {

int ug_oracle_forward_f32(int G, const float* means, const float* tril, const float* opac, const float* inten,
                          const float* pixels, int64_t P, const float* origin, const float* a_ele, const int32_t* counts,
                          const int32_t* offsets, const int32_t* indices, float chi2, float beta, float alpha_min,
                          float bg, float bgI, int excl, float* image, float* alpha, float* depth, float* sums,
                          const float* plane, int W, float col_offset, float dlat, float near_mm, float dax) {
    forward_impl<float>(G, means, tril, opac, inten, pixels, P, origin, a_ele, counts, offsets, indices, chi2, beta,
                        alpha_min, bg, bgI, excl, image, alpha, depth, sums, reinterpret_cast<const Plane*>(plane),
                        Grid{W, col_offset, dlat, near_mm, dax});
    return 0;
}

int ug_oracle_forward_f64(int G, const double* means, const double* tril, const double* opac, const double* inten,
                          const double* pixels, int64_t P, const float* origin, const float* a_ele, const int32_t* counts,
                          const int32_t* offsets, const int32_t* indices, float chi2, float beta, float alpha_min,
                          float bg, float bgI, int excl, double* image, double* alpha, double* depth, double* sums,
                          const float* plane, int W, float col_offset, float dlat, float near_mm, float dax) {
    forward_impl<double>(G, means, tril, opac, inten, pixels, P, origin, a_ele, counts, offsets, indices, chi2, beta,
                         alpha_min, bg, bgI, excl, image, alpha, depth, sums, reinterpret_cast<const Plane*>(plane),
                         Grid{W, col_offset, dlat, near_mm, dax});
    return 0;
}

int ug_oracle_backward_f32(int G, const float* means, const float* tril, const float* opac, const float* inten,
                           const float* pixels, int64_t P, const float* origin, const float* a_ele,
                           const int32_t* counts, const int32_t* offsets, const int32_t* indices, float chi2,
                           float beta, float alpha_min, float bg, float bgI, int gate, const float* gI, const float* gA,
                           const float* gD, float* d_means, float* d_tril, float* d_opac, float* d_inten,
                           const float* plane, int W, float col_offset, float dlat, float near_mm, float dax) {
    backward_impl<float>(G, means, tril, opac, inten, pixels, P, origin, a_ele, counts, offsets, indices, chi2, beta,
                         alpha_min, bg, bgI, gate, gI, gA, gD, d_means, d_tril, d_opac, d_inten,
                         reinterpret_cast<const Plane*>(plane), Grid{W, col_offset, dlat, near_mm, dax});
    return 0;
}

int ug_oracle_backward_f64(int G, const double* means, const double* tril, const double* opac, const double* inten,
                           const double* pixels, int64_t P, const float* origin, const float* a_ele,
                           const int32_t* counts, const int32_t* offsets, const int32_t* indices, float chi2,
                           float beta, float alpha_min, float bg, float bgI, int gate, const float* gI, const float* gA,
                           const float* gD, double* d_means, double* d_tril, double* d_opac, double* d_inten,
                           const float* plane, int W, float col_offset, float dlat, float near_mm, float dax) {
    backward_impl<double>(G, means, tril, opac, inten, pixels, P, origin, a_ele, counts, offsets, indices, chi2, beta,
                          alpha_min, bg, bgI, gate, gI, gA, gD, d_means, d_tril, d_opac, d_inten,
                          reinterpret_cast<const Plane*>(plane), Grid{W, col_offset, dlat, near_mm, dax});
    return 0;
}

}  // extern "C"
