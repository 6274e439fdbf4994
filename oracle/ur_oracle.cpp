// TEST INFRASTRUCTURE ONLY -- the oracle. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load this library, and only as the checker / CPU baseline; the product never calls it.
//
// CPU restatement of the UltraG-Ray B-mode renderer of the reference, one frame at a time, following:
//   mapper   ops/ultragray_renderer/src/mapper/core.cu:3-129, src/mapper/wrapper_mapper.cu:7-128
//   shared   ops/ultragray_renderer/include/shader/shared.cuh:5-85 (GL3 ray integral, SH, weight)
//   forward  ops/ultragray_renderer/src/shader/forward.cu:22-119
//   backward ops/ultragray_renderer/src/shader/backward.cu:26-325
// The reference's covariance.cuh (quaternion -> rotation, Sigma^-1 from scale/rotation, and their VJPs) is
// absent from its tree. Conventions used here: q = (w, x, y, z), normalised inside; R row-major with columns
// the principal axes (the shader uses o_local = R^T (o - mu)); Sigma^-1 = R diag(1/s^2) R^T. Their VJPs are
// exact derivatives of these definitions -- parity unpinned by any reference artifact, checked by finite
// differences only.
//
// FP32 arithmetic is canonical (explicit fmaf, -ffp-contract=off); templates also run in double.
// Deviations, defined here because the reference leaves them undefined: Gaussians with non-finite
// projections are culled; the depth key saturates at 0xffffffff.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

template <typename T> inline T fma_(T a, T b, T c);
template <> inline float fma_<float>(float a, float b, float c) { return std::fmaf(a, b, c); }
template <> inline double fma_<double>(double a, double b, double c) { return std::fma(a, b, c); }
template <typename T> inline T dot3(const T a[3], const T b[3]) { return fma_(a[2], b[2], fma_(a[1], b[1], a[0] * b[0])); }

const double kXi = 0.7745966692;
const double kW0 = 0.5555555556;
const double kW1 = 0.8888888889;

template <typename T>
void quat_to_rot(const T* q4, T R[9]) {
    const T w0 = q4[0], x0 = q4[1], y0 = q4[2], z0 = q4[3];
    const T n = std::sqrt(std::max(w0 * w0 + x0 * x0 + y0 * y0 + z0 * z0, T(1e-12f)));
    const T w = w0 / n, x = x0 / n, y = y0 / n, z = z0 / n;
    R[0] = T(1) - T(2) * (y * y + z * z); R[1] = T(2) * (x * y - w * z);     R[2] = T(2) * (x * z + w * y);
    R[3] = T(2) * (x * y + w * z);     R[4] = T(1) - T(2) * (x * x + z * z); R[5] = T(2) * (y * z - w * x);
    R[6] = T(2) * (x * z - w * y);     R[7] = T(2) * (y * z + w * x);     R[8] = T(1) - T(2) * (x * x + y * y);
}

// dL/dq (unnormalised input) from dL/dR, through the normalisation.
template <typename T>
void rot_vjp_to_quat(const T* q4, const T dR[9], T dq[4]) {
    const T w0 = q4[0], x0 = q4[1], y0 = q4[2], z0 = q4[3];
    const T n = std::sqrt(std::max(w0 * w0 + x0 * x0 + y0 * y0 + z0 * z0, T(1e-12f)));
    const T w = w0 / n, x = x0 / n, y = y0 / n, z = z0 / n;
    // derivatives of R entries w.r.t. the unit quaternion
    const T gw = T(2) * (-z * dR[1] + y * dR[2] + z * dR[3] - x * dR[5] - y * dR[6] + x * dR[7]);
    const T gx = T(2) * (y * dR[1] + z * dR[2] + y * dR[3] - T(2) * x * dR[4] - w * dR[5] + z * dR[6] + w * dR[7] -
                         T(2) * x * dR[8]);
    const T gy = T(2) * (-T(2) * y * dR[0] + x * dR[1] + w * dR[2] + x * dR[3] + z * dR[5] - w * dR[6] + z * dR[7] -
                         T(2) * y * dR[8]);
    const T gz = T(2) * (-T(2) * z * dR[0] - w * dR[1] + x * dR[2] + w * dR[3] - T(2) * z * dR[4] + y * dR[5] +
                         x * dR[6] + y * dR[7]);
    const T dot = gw * w + gx * x + gy * y + gz * z;
    dq[0] = (gw - dot * w) / n;
    dq[1] = (gx - dot * x) / n;
    dq[2] = (gy - dot * y) / n;
    dq[3] = (gz - dot * z) / n;
}

template <typename T>
void sigma_inv(const T R[9], const T s[3], T Si[9]) {
    const T d[3] = {T(1) / (s[0] * s[0]), T(1) / (s[1] * s[1]), T(1) / (s[2] * s[2])};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Si[3 * i + j] = R[3 * i] * d[0] * R[3 * j] + R[3 * i + 1] * d[1] * R[3 * j + 1] +
                                                  R[3 * i + 2] * d[2] * R[3 * j + 2];
}

struct Frame {
    float o[3], lat[3], ax[3], ele[3];
};

// Mapper outputs for one frame.
struct Geometry {
    int W, H;
    const float* pixels;  // (W, H, 3)
    const float* ray_o;   // (W, 3)
    const float* ray_d;   // (W, 3)
    const float* a_ele;   // (3)
    const float* delta;   // (W) OPS offsets
};

template <typename T>
struct GL3 {
    T psi;
    T half, t[3], f[3], x[3][3], og[3], dg[3], diff[3], ol[3], dl[3];
    bool valid;
};

// compute_ray_gaussian_integral (shared.cuh:27-85). tc_frozen (optional) replaces t_center.
template <typename T>
GL3<T> ray_integral(const T ro[3], const T rd[3], const T* mu, const T* s, const T R[9], T l, T Rw,
                    const T* tc_frozen) {
    GL3<T> g{};
    g.valid = false;
    g.psi = 0;
    if (!(l > T(0))) return g;
    for (int k = 0; k < 3; ++k) g.diff[k] = ro[k] - T(mu[k]);
    for (int c = 0; c < 3; ++c) {
        g.ol[c] = fma_(R[6 + c], g.diff[2], fma_(R[3 + c], g.diff[1], R[c] * g.diff[0]));
        g.dl[c] = fma_(R[6 + c], rd[2], fma_(R[3 + c], rd[1], R[c] * rd[0]));
        g.og[c] = g.ol[c] / T(s[c]);
        g.dg[c] = g.dl[c] / T(s[c]);
    }
    const T dn2 = dot3(g.dg, g.dg);
    const T tc = tc_frozen ? *tc_frozen : -dot3(g.og, g.dg) / (dn2 + T(1e-8f));
    T ts = 0, te = l;
    if (Rw > T(0)) {
        ts = std::max(T(0), tc - Rw);
        te = std::min(l, tc + Rw);
    }
    if (!(te > ts)) return g;
    g.valid = true;
    g.half = T(0.5f) * (te - ts);
    const T c = T(0.5f) * (te + ts);
    const T xi[3] = {T(-kXi), T(0), T(kXi)};
    const T wq[3] = {T(kW0), T(kW1), T(kW0)};
    T acc = 0;
    for (int i = 0; i < 3; ++i) {
        g.t[i] = fma_(g.half, xi[i], c);
        for (int k = 0; k < 3; ++k) g.x[i][k] = fma_(g.dg[k], g.t[i], g.og[k]);
        g.f[i] = std::exp(T(-0.5f) * dot3(g.x[i], g.x[i]));
        acc = fma_(wq[i], g.f[i], acc);
    }
    g.psi = g.half * acc;
    return g;
}

template <typename T>
T t_center(const T ro[3], const T rd[3], const T* mu, const T* s, const T R[9]) {
    T og[3], dg[3];
    for (int c = 0; c < 3; ++c) {
        const T d0 = ro[0] - T(mu[0]), d1 = ro[1] - T(mu[1]), d2 = ro[2] - T(mu[2]);
        og[c] = fma_(R[6 + c], d2, fma_(R[3 + c], d1, R[c] * d0)) / T(s[c]);
        dg[c] = fma_(R[6 + c], rd[2], fma_(R[3 + c], rd[1], R[c] * rd[0])) / T(s[c]);
    }
    return -dot3(og, dg) / (dot3(dg, dg) + T(1e-8f));
}

template <typename T>
struct Params {
    int G;
    const T *means, *scales, *quats, *atten, *sh;  // (G,3), (G,3), (G,4), (G,), (G,4)
};

struct Lists {
    const int32_t *counts, *offsets, *indices;  // per ray
    const float* sorted_depths;
};

template <typename T>
void ray_setup(const Geometry& geo, int x, T ro[3], T rd[3]) {
    for (int k = 0; k < 3; ++k) {
        ro[k] = fma_(T(geo.a_ele[k]), T(geo.delta[x]), T(geo.ray_o[3 * x + k]));
        rd[k] = T(geo.ray_d[3 * x + k]);
    }
}

int lower_bound_depth(const float* d, int off, int n, float v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (d[off + mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

struct Flags {
    float k_sigma, Rw;
    int local_only;
    float local_max;
};

// Membership of a hit in the transmission / echo sums of a sample, decided in the canonical FP32 arithmetic on
// the FP32 values of the inputs (forward.cu:76-96). The set is part of the semantics: the FP64 evaluation uses
// the same decisions and only measures the arithmetic error of the values.
template <typename In>
void pair_terms32(const In* mu, const In* s, const float ro[3], const float rd[3], float tpx, const Flags& F,
                  bool& need_t, bool& need_e) {
    const float md[3] = {(float)mu[0] - ro[0], (float)mu[1] - ro[1], (float)mu[2] - ro[2]};
    const float tg = dot3(md, rd);
    const float ms = std::max((float)s[0], std::max((float)s[1], (float)s[2]));
    const float rdep = ms * F.k_sigma;
    need_t = tg < tpx;
    if (F.local_only) need_t = need_t && ((tpx - tg) <= rdep + F.Rw);
    need_e = std::fabs(tg - tpx) <= rdep;
}

float sample_depth32(const Geometry& geo, int x, int y, float ro[3], float rd[3]) {
    ray_setup<float>(geo, x, ro, rd);
    const float* pp = geo.pixels + 3 * ((int64_t)x * geo.H + y);
    const float pd[3] = {pp[0] - ro[0], pp[1] - ro[1], pp[2] - ro[2]};
    return std::max(0.0f, dot3(pd, rd));
}

// Per-pixel forward (forward.cu:45-118). Returns B; T, num, den through pointers.
template <typename T>
T pixel_forward(const Params<T>& P, const Geometry& geo, const Lists& Ls, const Flags& F, int x, int y, T* pT, T* pnum,
                T* pden, const T* tc_frozen /* (W, G) or null */) {
    const int count = Ls.counts[x];
    *pT = 1;
    *pnum = 0;
    *pden = 0;
    if (count == 0) return 0;
    const int off = Ls.offsets[x];
    T ro[3], rd[3];
    ray_setup<T>(geo, x, ro, rd);
    const float* pp = geo.pixels + 3 * ((int64_t)x * geo.H + y);
    const T p[3] = {pp[0], pp[1], pp[2]};
    const T pd[3] = {p[0] - ro[0], p[1] - ro[1], p[2] - ro[2]};
    const T tpx = std::max(T(0), dot3(pd, rd));
    T trans = 1, num = 0, den = 0, S = 0;
    float ro32[3], rd32[3];
    const float tpx32 = sample_depth32(geo, x, y, ro32, rd32);
    int begin = 0;
    float upper = 1e30f;
    if (F.local_only) {
        begin = lower_bound_depth(Ls.sorted_depths, off, count, tpx32 - F.local_max);
        upper = tpx32 + F.local_max;
    }
    for (int k = begin; k < count; ++k) {
        if (F.local_only && Ls.sorted_depths[off + k] > upper) break;
        const int g = Ls.indices[off + k];
        const T* mu = P.means + 3 * g;
        const T* s = P.scales + 3 * g;
        bool need_t, need_e;
        pair_terms32(mu, s, ro32, rd32, tpx32, F, need_t, need_e);
        if (!need_t && !need_e) continue;
        T R[9];
        quat_to_rot<T>(P.quats + 4 * g, R);
        const T tau = T(1) / (T(1) + std::exp(-T(P.atten[g])));
        if (need_t) {
            const T* tcf = tc_frozen ? tc_frozen + (int64_t)x * P.G + g : nullptr;
            const GL3<T> gl = ray_integral<T>(ro, rd, mu, s, R, tpx, T(F.Rw), tcf);
            const T Ti = fma_(T(1) - tau, std::exp(-gl.psi), tau);
            trans *= Ti;
        }
        if (need_e) {
            const T sv[3] = {s[0], s[1], s[2]};
            T Si[9];
            sigma_inv<T>(R, sv, Si);
            const T d[3] = {p[0] - T(mu[0]), p[1] - T(mu[1]), p[2] - T(mu[2])};
            const T q = (Si[0] * d[0] + Si[1] * d[1] + Si[2] * d[2]) * d[0] + (Si[3] * d[0] + Si[4] * d[1] + Si[5] * d[2]) * d[1] +
                        (Si[6] * d[0] + Si[7] * d[1] + Si[8] * d[2]) * d[2];
            const T w = std::exp(T(-0.5f) * q);
            const T* c = P.sh + 4 * g;
            const T I = T(c[0]) + T(c[1]) * rd[0] + T(c[2]) * rd[1] + T(c[3]) * rd[2];
            num += I * w;
            den += w;
            S += w;
        }
    }
    *pT = trans;
    *pnum = num;
    *pden = den;
    const T gamma = T(1) - std::exp(-S);
    const T E = num / (den + T(1e-6f));
    return trans * (gamma * E);
}

}  // namespace

extern "C" {

// kernel_project_and_cull (core.cu:3-68) for one frame: per Gaussian ray range [min_ray, max_ray] and
// quantised depth; count = 0 when culled.
int ur_oracle_preprocess(int G, const float* means, const float* scales, const float* origin, const float* a_lat,
                         const float* a_ele, const float* a_ax, int W, float width_mm, float near_mm, float col_offset,
                         float elev_margin, float k_sigma, int32_t* count, int32_t* rmin, int32_t* rmax, uint32_t* dq) {
    const float ps = width_mm / (float)W;
    for (int g = 0; g < G; ++g) {
        const float d[3] = {means[3 * g] - origin[0], means[3 * g + 1] - origin[1], means[3 * g + 2] - origin[2]};
        const float x = dot3(d, a_lat), y = dot3(d, a_ele), z = dot3(d, a_ax);
        const float ms = std::max(scales[3 * g], std::max(scales[3 * g + 1], scales[3 * g + 2]));
        const float r = ms * k_sigma;
        count[g] = 0;
        rmin[g] = rmax[g] = 0;
        dq[g] = 0;
        if (!(z + r >= near_mm)) continue;
        if (!(std::fabs(y) <= r + elev_margin)) continue;
        const float c = x / ps + col_offset;
        const float rp = r / ps;
        const float f0 = std::floor(c - rp), f1 = std::ceil(c + rp);
        if (!(std::isfinite(f0) && std::isfinite(f1))) continue;
        const int lo = std::max(0, (int)std::min(std::max(f0, -1073741824.0f), 1073741824.0f));
        const int hi = std::min(W - 1, (int)std::min(std::max(f1, -1073741824.0f), 1073741824.0f));
        if (lo > hi) continue;
        count[g] = hi - lo + 1;
        rmin[g] = lo;
        rmax[g] = hi;
        const float q = std::max(0.0f, z) * 1000.0f;
        dq[g] = q >= 4294967296.0f ? 0xffffffffu : (uint32_t)q;
    }
    return 0;
}

// Emit (core.cu:70-96), stable full 64-bit sort (wrapper_mapper.cu:88-105), CSR (core.cu:98-129,
// wrapper_mapper.cu:107-127). Outputs counts[W], offsets[W], indices[K], sorted_depths[K].
int ur_oracle_emit_sort(int G, int W, const int32_t* count, const int32_t* rmin, const int32_t* rmax,
                        const uint32_t* dq, int32_t* counts, int32_t* offsets, int32_t* indices, float* sorted_depths) {
    std::vector<uint64_t> keys;
    std::vector<int32_t> vals;
    for (int g = 0; g < G; ++g) {
        if (count[g] <= 0) continue;
        for (int r = rmin[g]; r <= rmax[g]; ++r) {
            keys.push_back(((uint64_t)r << 32) | dq[g]);
            vals.push_back(g);
        }
    }
    std::vector<int64_t> ord(keys.size());
    for (size_t i = 0; i < ord.size(); ++i) ord[i] = (int64_t)i;
    std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) { return keys[a] < keys[b]; });
    for (int x = 0; x < W; ++x) counts[x] = offsets[x] = 0;
    std::vector<int> start(W, -1), end(W, -1);
    for (size_t i = 0; i < ord.size(); ++i) {
        const uint64_t k = keys[ord[i]];
        indices[i] = vals[ord[i]];
        sorted_depths[i] = (float)(uint32_t)(k & 0xffffffffu) * 0.001f;
        const int ray = (int)(k >> 32);
        if (start[ray] < 0) start[ray] = (int)i;
        end[ray] = (int)i + 1;
    }
    for (int x = 0; x < W; ++x)
        if (start[x] >= 0) {
            offsets[x] = start[x];
            counts[x] = end[x] - start[x];
        }
    return 0;
}

#define UR_ARGS_T(PT)                                                                                        \
    int G, const PT *means, const PT *scales, const PT *quats, const PT *atten, const PT *sh, int W, int H,        \
        const float *pixels, const float *ray_o, const float *ray_d, const float *a_ele, const float *delta,      \
        const int32_t *counts, const int32_t *offsets, const int32_t *indices, const float *sorted_depths,       \
        float k_sigma, float Rw, int local_only, float local_max

// Forward over all pixels: image[y*W + x] (H, W); aux (W*H, 3) = (T, num, den) at (y*W + x).
int ur_oracle_forward_f32(UR_ARGS_T(float), float* image, float* aux) {
    Params<float> P{G, means, scales, quats, atten, sh};
    Geometry geo{W, H, pixels, ray_o, ray_d, a_ele, delta};
    Lists Ls{counts, offsets, indices, sorted_depths};
    Flags F{k_sigma, Rw, local_only, local_max};
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            float T_, n_, d_;
            image[y * W + x] = pixel_forward<float>(P, geo, Ls, F, x, y, &T_, &n_, &d_, nullptr);
            if (aux) {
                aux[3 * (y * W + x)] = T_;
                aux[3 * (y * W + x) + 1] = n_;
                aux[3 * (y * W + x) + 2] = d_;
            }
        }
    return 0;
}

// FP64 forward; tc_frozen (W, G) optional: freezes the GL3 window centres (for finite differences that match
// the reference VJP, which holds the window endpoints constant).
int ur_oracle_forward_f64(UR_ARGS_T(double), const double* tc_frozen, double* image) {
    Params<double> P{G, means, scales, quats, atten, sh};
    Geometry geo{W, H, pixels, ray_o, ray_d, a_ele, delta};
    Lists Ls{counts, offsets, indices, sorted_depths};
    Flags F{k_sigma, Rw, local_only, local_max};
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            double T_, n_, d_;
            image[y * W + x] = pixel_forward<double>(P, geo, Ls, F, x, y, &T_, &n_, &d_, tc_frozen);
        }
    return 0;
}

// Window centres t_c for every (ray, Gaussian) (W, G), FP64.
int ur_oracle_window_centres(int G, const double* means, const double* scales, const double* quats, int W,
                             const float* ray_o, const float* ray_d, const float* a_ele, const float* delta,
                             double* tc) {
    Geometry geo{W, 0, nullptr, ray_o, ray_d, a_ele, delta};
    for (int x = 0; x < W; ++x) {
        double ro[3], rd[3];
        ray_setup<double>(geo, x, ro, rd);
        for (int g = 0; g < G; ++g) {
            double R[9];
            quat_to_rot<double>(quats + 4 * g, R);
            tc[(int64_t)x * G + g] = t_center<double>(ro, rd, means + 3 * g, scales + 3 * g, R);
        }
    }
    return 0;
}

// FP64 backward (backward.cu:154-325): recompute T, E; per pair the transmission VJP through the GL3 nodes
// with the window held fixed, the echo weight VJP through Sigma^-1 -> (R, s) -> q, SH gradient along the ray
// direction. Pixels with |dL/dB| < grad_threshold are skipped and dL/dT_i uses T/T_i only when T_i > 1e-6,
// as in the reference. Outputs d_means (G,3), d_scales (G,3), d_quats (G,4), d_atten (G), d_sh (G,4).
int ur_oracle_backward_f64(UR_ARGS_T(double), const float* dL_dB, float grad_threshold, double* d_means,
                           double* d_scales, double* d_quats, double* d_atten, double* d_sh) {
    Params<double> P{G, means, scales, quats, atten, sh};
    Geometry geo{W, H, pixels, ray_o, ray_d, a_ele, delta};
    Lists Ls{counts, offsets, indices, sorted_depths};
    Flags F{k_sigma, Rw, local_only, local_max};
    typedef double T;
    std::fill(d_means, d_means + 3 * (int64_t)G, 0.0);
    std::fill(d_scales, d_scales + 3 * (int64_t)G, 0.0);
    std::fill(d_quats, d_quats + 4 * (int64_t)G, 0.0);
    std::fill(d_atten, d_atten + G, 0.0);
    std::fill(d_sh, d_sh + 4 * (int64_t)G, 0.0);
    std::vector<T> dR_all(9 * (int64_t)G, 0.0);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const T gB = dL_dB[y * W + x];
            if (std::fabs(gB) < T(grad_threshold)) continue;
            const int count = counts[x];
            if (count == 0) continue;
            const int off = offsets[x];
            T Tf, num, den;
            pixel_forward<T>(P, geo, Ls, F, x, y, &Tf, &num, &den, nullptr);
            const T S = den;
            const T gamma = T(1) - std::exp(-S);
            const T eps = 1e-6;
            const T Eraw = num / (den + eps);
            const T dT = gB * gamma * Eraw;
            const T dE = gB * Tf;
            const T dGamma = dE * Eraw;
            const T dEraw = dE * gamma;
            const T dS = dGamma * std::exp(-S);
            const T dNum = dEraw / (den + eps);
            const T dDen = -dEraw * num / ((den + eps) * (den + eps));
            T ro[3], rd[3];
            ray_setup<T>(geo, x, ro, rd);
            const float* pp = pixels + 3 * ((int64_t)x * H + y);
            const T p[3] = {pp[0], pp[1], pp[2]};
            const T pd[3] = {p[0] - ro[0], p[1] - ro[1], p[2] - ro[2]};
            const T tpx = std::max(T(0), dot3(pd, rd));
            float ro32[3], rd32[3];
            const float tpx32 = sample_depth32(geo, x, y, ro32, rd32);
            int begin = 0;
            float upper = 1e30f;
            if (local_only) {
                begin = lower_bound_depth(sorted_depths, off, count, tpx32 - local_max);
                upper = tpx32 + local_max;
            }
            for (int k = begin; k < count; ++k) {
                if (local_only && sorted_depths[off + k] > upper) break;
                const int g = indices[off + k];
                const double* mu = means + 3 * g;
                const double* s = scales + 3 * g;
                bool need_t, need_e;
                pair_terms32(mu, s, ro32, rd32, tpx32, F, need_t, need_e);
                if (!need_t && !need_e) continue;
                T R[9];
                quat_to_rot<T>(quats + 4 * g, R);
                T* dR = &dR_all[9 * (int64_t)g];
                if (need_t) {
                    const GL3<T> gl = ray_integral<T>(ro, rd, mu, s, R, tpx, T(Rw), nullptr);
                    const T tau = T(1) / (T(1) + std::exp(-T(atten[g])));
                    const T ep = std::exp(-gl.psi);
                    const T Ti = tau + (T(1) - tau) * ep;
                    const T dTi = Ti > T(1e-6) ? dT * (Tf / Ti) : T(0);
                    d_atten[g] += dTi * (T(1) - ep) * tau * (T(1) - tau);
                    const T dpsi = -dTi * (T(1) - tau) * ep;
                    if (gl.valid) {
                        const T coeff = dpsi * gl.half;
                        const T wq[3] = {kW0, kW1, kW0};
                        T dog[3] = {0, 0, 0}, ddg[3] = {0, 0, 0};
                        for (int i = 0; i < 3; ++i)
                            for (int c = 0; c < 3; ++c) {
                                dog[c] += -coeff * wq[i] * gl.f[i] * gl.x[i][c];
                                ddg[c] += -coeff * wq[i] * gl.f[i] * gl.t[i] * gl.x[i][c];
                            }
                        T dol[3], ddl[3];
                        for (int c = 0; c < 3; ++c) {
                            dol[c] = dog[c] / T(s[c]);
                            ddl[c] = ddg[c] / T(s[c]);
                            d_scales[3 * g + c] += dog[c] * (-gl.og[c] / T(s[c])) + ddg[c] * (-gl.dg[c] / T(s[c]));
                        }
                        for (int r = 0; r < 3; ++r) {
                            d_means[3 * g + r] += -(R[3 * r] * dol[0] + R[3 * r + 1] * dol[1] + R[3 * r + 2] * dol[2]);
                            for (int c = 0; c < 3; ++c) dR[3 * r + c] += dol[c] * gl.diff[r] + ddl[c] * rd[r];
                        }
                    }
                }
                if (need_e) {
                    const T sv[3] = {s[0], s[1], s[2]};
                    T Si[9];
                    sigma_inv<T>(R, sv, Si);
                    const T d[3] = {p[0] - T(mu[0]), p[1] - T(mu[1]), p[2] - T(mu[2])};
                    T v[3];
                    for (int i = 0; i < 3; ++i) v[i] = Si[3 * i] * d[0] + Si[3 * i + 1] * d[1] + Si[3 * i + 2] * d[2];
                    const T q = v[0] * d[0] + v[1] * d[1] + v[2] * d[2];
                    const T w = std::exp(T(-0.5) * q);
                    const double* c4 = sh + 4 * g;
                    const T I = T(c4[0]) + T(c4[1]) * rd[0] + T(c4[2]) * rd[1] + T(c4[3]) * rd[2];
                    d_sh[4 * g + 0] += dNum * w;
                    d_sh[4 * g + 1] += dNum * w * rd[0];
                    d_sh[4 * g + 2] += dNum * w * rd[1];
                    d_sh[4 * g + 3] += dNum * w * rd[2];
                    const T dw = dNum * I + dDen + dS;
                    const T dQ = dw * w * T(-0.5);
                    for (int i = 0; i < 3; ++i) d_means[3 * g + i] += dQ * (T(-2) * v[i]);
                    // Sigma^-1 = R D R^T, D = diag(1/s^2): dR = (G + G^T) R D, dD_k = (R^T G R)_kk
                    T Gm[9];
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j) Gm[3 * i + j] = dQ * d[i] * d[j];
                    const T Dk[3] = {T(1) / (sv[0] * sv[0]), T(1) / (sv[1] * sv[1]), T(1) / (sv[2] * sv[2])};
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) {
                            T acc = 0;
                            for (int j = 0; j < 3; ++j) acc += (Gm[3 * a + j] + Gm[3 * j + a]) * R[3 * j + b];
                            dR[3 * a + b] += acc * Dk[b];
                        }
                    for (int k2 = 0; k2 < 3; ++k2) {
                        T acc = 0;
                        for (int i = 0; i < 3; ++i)
                            for (int j = 0; j < 3; ++j) acc += R[3 * i + k2] * Gm[3 * i + j] * R[3 * j + k2];
                        d_scales[3 * g + k2] += acc * (T(-2) / (sv[k2] * sv[k2] * sv[k2]));
                    }
                }
            }
        }
    for (int g = 0; g < G; ++g) {
        T dq[4];
        rot_vjp_to_quat<T>(quats + 4 * g, &dR_all[9 * (int64_t)g], dq);
        for (int k = 0; k < 4; ++k) d_quats[4 * g + k] = dq[k];
    }
    return 0;
}

}  // extern "C"
