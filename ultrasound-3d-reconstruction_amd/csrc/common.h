// Internal helpers shared by the gfx950 kernels of libus3d.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "us3d.h"

namespace us3d {

constexpr int kWave = 64;

// ---------------------------------------------------------------- host-side error plumbing
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int check_launch(const char* what);

// Opt-in per-kernel timing (us3d_stage_timing): when enabled, a hipEvent pair brackets every kernel launch of
// the library on its stream; us3d_stage_timing_read sums the elapsed times per kernel name.
bool timing_enabled();
void timing_mark(const char* name, hipStream_t stream, bool begin);

struct StageTimer {
    const char* name;
    hipStream_t stream;
    StageTimer(const char* n, hipStream_t s) : name(n), stream(s) {
        if (timing_enabled()) timing_mark(name, stream, true);
    }
    ~StageTimer() {
        if (timing_enabled()) timing_mark(name, stream, false);
    }
};

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int64_t align_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

// ---------------------------------------------------------------- device math
// All library code is compiled with -ffp-contract=off: every FMA below is explicit, so the CPU
// restatement that performs the same operations reproduces the results bit for bit.
__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    return __builtin_fmaf(az, bz, __builtin_fmaf(ay, by, ax * bx));
}

// Marginal chi2 radius of the ellipsoid {d : d^T L L^T d <= chi2} along unit axis a, by forward
// substitution y = L^-1 a (L lower triangular, rows (l11), (l12 l22), (l13 l23 l33)).
// Follows compute_probe_bbox_radii, ops/ultragauss_renderer/src/mapper/core.cu:9-44.
__device__ __forceinline__ float marginal_radius(float l11, float l12, float l13, float l22, float l23,
                                                 float l33, float ax, float ay, float az, float sqrt_chi2) {
    const float y0 = ax / fmaxf(l11, 1e-10f);
    const float y1 = __builtin_fmaf(-l12, y0, ay) / fmaxf(l22, 1e-10f);
    const float y2 = __builtin_fmaf(-l23, y1, __builtin_fmaf(-l13, y0, az)) / fmaxf(l33, 1e-10f);
    const float s = fmaxf(__builtin_fmaf(y2, y2, __builtin_fmaf(y1, y1, y0 * y0)), 1e-12f);
    return fmaxf(1e-3f, sqrt_chi2 * sqrtf(s));
}

// q = || L^T d ||^2 (= d^T Lambda d with Lambda = L L^T); y = L^T d is returned for the adjoint.
__device__ __forceinline__ float quad_form(float l11, float l12, float l13, float l22, float l23, float l33,
                                          float dx, float dy, float dz, float& y0, float& y1, float& y2) {
    y0 = __builtin_fmaf(l13, dz, __builtin_fmaf(l12, dy, l11 * dx));
    y1 = __builtin_fmaf(l23, dz, l22 * dy);
    y2 = l33 * dz;
    return __builtin_fmaf(y2, y2, __builtin_fmaf(y1, y1, y0 * y0));
}

// World position of pixel (ix, iy) of a frame. With pixels == NULL the position is the frame's grid. Linear
// probe: the grid of NpzUltrasoundDataset._pixel_positions (pipeline_god_file.py:299-306), evaluated with the
// same FP32 operations in the same order (lateral = (u - c) * d_lat; axial = near + v * d_ax;
// p = (o + lateral * a_lat) + axial * a_ax), so it equals that tensor bit for bit. Fan probe: angle
// t = (u - c) * d_lat, radius r = near + v * d_ax, lateral = r sin t, axial = r cos t, then the same sum.
__device__ __forceinline__ void pixel_position(const us3d_frame& fr, const float* __restrict__ pixels, int ix, int iy,
                                               float& x, float& y, float& z) {
    if (pixels) {
        const int64_t p = (int64_t)fr.pixel_base + (int64_t)iy * fr.width + ix;
        x = pixels[3 * p + 0];
        y = pixels[3 * p + 1];
        z = pixels[3 * p + 2];
    } else {
        float lat = ((float)ix - fr.col_offset) * fr.delta_lat;
        float ax = fr.near_mm + (float)iy * fr.delta_ax;
        if (fr.geometry == 1) {
            float s, c;
            sincosf(lat, &s, &c);
            const float r = ax;
            lat = r * s;
            ax = r * c;
        }
        x = (fr.origin[0] + lat * fr.axis_lat[0]) + ax * fr.axis_ax[0];
        y = (fr.origin[1] + lat * fr.axis_lat[1]) + ax * fr.axis_ax[1];
        z = (fr.origin[2] + lat * fr.axis_lat[2]) + ax * fr.axis_ax[2];
    }
}

// XCD-aware work order. Workgroups are dispatched to the 8 XCDs round-robin (workgroup b to XCD b % 8), and each
// XCD has its own L2, while neighbouring tiles of a frame share most of their Gaussians. Workgroup b therefore
// takes tile remap_xcd(b): within each super-block of 8 * kXcdRun tiles, XCD r walks the run of kXcdRun
// consecutive tiles [r kXcdRun, (r + 1) kXcdRun), so the records a tile gathers are in its XCD's L2 from the tiles
// before it, and the runs stay short enough that every XCD sees the whole batch (work per tile varies along a
// sweep). Tiles past the last whole super-block keep their own index: a bijection of [0, n).
constexpr int kXcds = 8;
constexpr int kXcdRun = 32;
__device__ __forceinline__ int remap_xcd(int b, int n) {
    constexpr int kSuper = kXcds * kXcdRun;
    if (b >= (n / kSuper) * kSuper) return b;
    const int r = b % kXcds, q = b / kXcds;
    return (q / kXcdRun) * kSuper + r * kXcdRun + (q % kXcdRun);
}

// Persistent (frame, tile) work queues. Instead of one workgroup per tile, a persistent launch keeps as many
// workgroups resident as the chip holds, and they take tiles from eight queues, one per XCD: workgroup b serves
// queue b % 8 (the XCD the round-robin dispatch places it on) and takes that XCD's tiles in the remap_xcd order, so
// a queue walks runs of neighbouring tiles whose Gaussians are in the XCD's L2, and a workgroup whose tiles cost
// more simply takes fewer of them. A workgroup whose queue is exhausted moves on to the following queues; it exits
// when all eight are. The queue heads (int32 counters kQueueStride apart) are cleared by a fill kernel before each
// launch, and every tile is handed out exactly once (one atomicAdd per take), so the outputs equal the one-workgroup-
// per-tile launch's bit for bit.
constexpr int kQueueStride = 64;

// Tile of item i of XCD queue r over n tiles (the tiles remap_xcd gives workgroups b = r mod 8), -1 past its end.
__device__ __forceinline__ int queue_tile(int r, int i, int n) {
    constexpr int kSuper = kXcds * kXcdRun;
    const int nsb = n / kSuper;
    if (i < nsb * kXcdRun) return (i / kXcdRun) * kSuper + r * kXcdRun + (i % kXcdRun);
    const int64_t t = (int64_t)nsb * kSuper + r + (int64_t)kXcds * (i - nsb * kXcdRun);
    return t < n ? (int)t : -1;
}

// Next tile of a persistent workgroup (every thread calls it; -1: all queues exhausted). The first barrier also
// separates the previous tile's use of LDS from the next tile's. q: the queue thread 0 draws from (starts at
// blockIdx.x % kXcds; only thread 0's copy is used).
__device__ __forceinline__ int next_queue_tile(int32_t* __restrict__ heads, int& q, int n, int* s_slot) {
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = -1;
        for (int k = 0; k < kXcds; ++k) {
            t = queue_tile(q, atomicAdd(&heads[q * kQueueStride], 1), n);
            if (t >= 0) break;
            q = (q + 1) % kXcds;
        }
        *s_slot = t;
    }
    __syncthreads();
    return *s_slot;
}

// Binary search: largest i in [0, n) with base[i] <= v (base ascending, base[0] <= v).
__device__ __forceinline__ int upper_index(const int32_t* __restrict__ base, int n, int v) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (base[mid] <= v) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Sweep of a (frame, Gaussian) pair index (sweeps ordered by pair_base; a handful per batch).
__device__ __forceinline__ int sweep_of_pair(const us3d_sweep* __restrict__ sweeps, int n_sweeps, int64_t pair) {
    int s = 0;
    while (s + 1 < n_sweeps && sweeps[s + 1].pair_base <= pair) ++s;
    return s;
}

// Sweep owning concatenated Gaussian row g.
__device__ __forceinline__ int sweep_of_gauss(const us3d_sweep* __restrict__ sweeps, int n_sweeps, int g) {
    int s = 0;
    while (s + 1 < n_sweeps && sweeps[s + 1].g_begin <= g) ++s;
    return (g >= sweeps[s].g_begin && g - sweeps[s].g_begin < sweeps[s].g_count) ? s : -1;
}

// The sweep of Gaussian g for a per-Gaussian reduction, with f_count = 0 when no sweep of the batch holds g
// (a batch of some sweeps' frames, or ranges with gaps): such a Gaussian has no pairs, so its sums stay zero and
// pair_base + (g - g_begin) * f_count is never formed outside the pair table.
__device__ __forceinline__ us3d_sweep sweep_for_reduce(const us3d_sweep* __restrict__ sweeps, int n_sweeps, int g) {
    const int s = sweep_of_gauss(sweeps, n_sweeps, g);
    us3d_sweep sw = sweeps[s < 0 ? 0 : s];
    if (s < 0) {
        sw.g_begin = g;
        sw.f_count = 0;
    }
    return sw;
}

// exp(-q/2) exactly as __expf(-0.5f * q) evaluates it (v_exp_f32 of (-0.5 q) * log2(e)_f32), with the two
// scalings folded into one multiply: -0.5 * log2(e)_f32 is exactly representable, so q * (-0.5 log2(e))
// rounds the same real number once and the result is bit-identical, one VALU operation cheaper.
__device__ __forceinline__ float exp_neg_half(float q) {
    return __builtin_amdgcn_exp2f(q * -0.721347510814666748046875f);
}

// The plane form of q (frames rendered on their own grids) is evaluated scaled by log2(e)/2, q' = q log2(e)/2, so
// that the weight exp(-q/2) is one exp2 of -q': the splat kernels multiply p, qq, s by kQScaleSqrt and K by kQScale
// in FP32 when they stage a pair, and compare q' with chi2 * kQScale.
constexpr float kQScale = 0.721347510814666748046875f;      // log2(e)/2 rounded to FP32
constexpr float kQScaleSqrt = 0.849321782588958740234375f;  // the FP32 square root of kQScale
__device__ __forceinline__ float exp2_neg(float qs) { return __builtin_amdgcn_exp2f(-qs); }

// Opacities o for which the alpha_hat > alpha_min test cannot fail at a pair that passes the q test (scaled plane
// form: q' <= chi2' = chi2 * kQScale): o * wmin_half > alpha_min, wmin_half = exp2_neg(chi2') / 2. v_exp_f32 is
// within 1 ulp, so such a pair has alpha_hat = o * exp2_neg(q') >= o * exp2_neg(chi2') (1 - 2^-21) > alpha_min
// (and with alpha_min < 0 the test never fails). The slice adjoint leaves the test out for these Gaussians; no
// decision changes.
__device__ __forceinline__ float alpha_free_wmin_half(float chi2s) { return 0.5f * exp2_neg(chi2s); }
__device__ __forceinline__ bool alpha_free(float o, float wmin_half, float alpha_min) {
    return o * wmin_half > alpha_min;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}
__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fminf(v, __shfl_xor(v, off, kWave));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, kWave));
    return v;
}

// v of lane ^ m (m a power of two < 64, a constant once the caller is unrolled): DPP quad_perm for m = 1, 2,
// DPP row_ror:8 for m = 8 (rotation by 8 inside a 16-lane row is xor 8), ds_swizzle in bitmask mode for m = 4,
// 16 (source lane ((lane & 31) ^ m) inside each 32-lane half), ds_bpermute for m = 32. Every lane must be active.
__device__ __forceinline__ int shfl_xor_pow2(int v, int m) {
    switch (m) {
        case 1: return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
        case 2: return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
        case 4: return __builtin_amdgcn_ds_swizzle(v, 0x1F | (4 << 10));
        case 8: return __builtin_amdgcn_update_dpp(0, v, 0x128, 0xF, 0xF, false);  // row_ror:8
        case 16: return __builtin_amdgcn_ds_swizzle(v, 0x1F | (16 << 10));
        default: return __shfl_xor(v, m, kWave);
    }
}

// Sum of v over the wavefront in a fixed order, the same value in every lane: xor butterflies over the
// distances 1 .. 16 (both halves of the double through shfl_xor_pow2), then the sums of the two 32-lane halves
// read from lanes 0 and 32 and added. Every lane must be active.
__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int m = 1; m < 32; m <<= 1)
        v += __hiloint2double(shfl_xor_pow2(__double2hiint(v), m), shfl_xor_pow2(__double2loint(v), m));
    const double a = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 0),
                                      __builtin_amdgcn_readlane(__double2loint(v), 0));
    const double b = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 32),
                                      __builtin_amdgcn_readlane(__double2loint(v), 32));
    return a + b;
}

// ---------------------------------------------------------------- shared pipeline pieces
// Exclusive scan (int64 accumulation, int32 output) with the total stored to *total.
int scan_i32(const int32_t* in, int32_t* out, int64_t n, int64_t* total, void* tmp, hipStream_t stream);
int64_t scan_tmp_bytes(int64_t n);

// Per-segment sort of unique 64-bit keys (segment t is [seg_off[t], seg_off[t+1])); keys_tmp is scratch of the
// same size for segments longer than 4096.
int sort_segments(uint64_t* keys, uint64_t* keys_tmp, const int32_t* seg_off, int n_segments, us3d_stats* stats,
                  int64_t k_capacity, hipStream_t stream);

// Image loss of a training step fused into a slice splat (train.hip holds the frame statistics and their
// gradient): the forward reads the targets gt (indexed like the image) and writes, per (tile, wavefront), the six
// sums sum|x-y|, sum x, sum y, sum x^2, sum y^2, sum xy of its pixels (x the clamped image; a lane's at most four
// pixels summed in FP32, the lanes in FP64) into part; the
// finish reduces them per frame into the loss and four gradient coefficients per frame (coeff), from which the
// adjoint forms dL/dx = cs sign(x-y) + ca + cx x + cy y while it stages its pixels. NULL pointers: no loss.
struct LossFuse {
    const float* gt;
    double* part;
    const double* coeff;
};
constexpr int kLossSums = 6;
// Per-frame loss and coefficients from the per-(tile, wavefront) sums of a fused forward: frame f owns the
// entries [frame_tile_base[f] * waves_per_tile, frame_tile_base[f + 1] * waves_per_tile) of part.
int image_loss_from_tiles(const double* part, const int32_t* frame_tile_base, const int32_t* frame_pixel_base,
                          int n_frames, int waves_per_tile, float w_l1, float w_ssim, float scale, float* loss,
                          double* coeff, hipStream_t stream);

// Workspace layout shared by the renderers (record, partial, mask and per-pair record sizes differ).
us3d_layout ws_layout(int64_t n_gauss, int64_t n_pairs, int64_t n_tiles, int64_t k_cap, int64_t v_cap,
                      int64_t record_bytes, int64_t partial_bytes, int64_t mask_bytes, int64_t sect_bytes);
int64_t pair_words(int64_t n_pairs);
int64_t fm_words(int64_t n_pairs, int64_t n_tiles);

template <typename T>
inline T* ws_at(const us3d_plan* plan, int64_t off) {
    return reinterpret_cast<T*>(reinterpret_cast<char*>(plan->workspace) + off);
}

inline unsigned blocks_for(int64_t n, int threads) {
    return (unsigned)((n + threads - 1) / threads > 0 ? (n + threads - 1) / threads : 1);
}

int check_plan(const us3d_plan* plan, int64_t ws_needed);


// The map of every renderer (mapper.hip): begin_pairs resets the stats block, the tile counters and the frame-major
// mask; the renderer's cull pass marks the visible pairs (mark_pair); rank_visible_pairs ranks them, checks
// v_capacity and lists them (vis_pair, pair_g, vis_list); the renderer's setup pass writes each visible pair's count,
// tile rectangle, depth key and record and counts its tiles (add_rect_corners); build_instances scans the counts
// (instance total into stats, overflow flag), scans the tile counts into offsets, emits the instances into tile
// buckets (seg_key = depth key << 32 | the instance's visible pair) and sorts every bucket, which orders it by
// (depth key, Gaussian).
int begin_pairs(const us3d_plan* plan, const us3d_layout& L, hipStream_t stream);

// Grid of a splat launch over the plan's tiles: n_tiles workgroups, or with per-XCD work queues (heads != NULL) the
// workgroups the device holds at once, after a fill kernel has cleared the queue heads on `stream`.
int tile_grid(const void* kernel, int threads, const us3d_plan* plan, int32_t* heads, hipStream_t stream);
int rank_visible_pairs(const us3d_plan* plan, const us3d_layout& L, hipStream_t stream);
// sect_f4 / emit_slot: when sect_f4 > 0, each visible pair's record in pair_sect (sect_f4 float4s) gets its emit base
// (first emit index, rectangle x0, y0 and width, as int bits) in float4 number emit_slot once the counts are scanned,
// so an adjoint reads it from the line it reads the record from (emit_index_rec).
int build_instances(const us3d_plan* plan, const us3d_layout& L, hipStream_t stream, int sect_f4 = 0,
                    int emit_slot = 0);

// First frame-major word of frame f of sweep s: each sweep before s takes f_count * ceil(g_count / 32) words.
__device__ __forceinline__ int64_t fm_word_base(const us3d_sweep* __restrict__ sweeps, int s, int f) {
    int64_t base = 0;
    for (int i = 0; i < s; ++i) base += (int64_t)sweeps[i].f_count * ((sweeps[i].g_count + 31) >> 5);
    return base + (int64_t)(f - sweeps[s].f_begin) * ((sweeps[s].g_count + 31) >> 5);
}

// Compact index of dense pair p: the number of visible pairs before it (p may equal n_pairs: the total).
__device__ __forceinline__ int pair_rank(const uint64_t* __restrict__ bits, const int32_t* __restrict__ word_off,
                                         int64_t p) {
    const int64_t w = p >> 6;
    const int b = (int)(p & 63);
    int r = word_off[w];
    if (b) r += __popcll(bits[w] & ((1ull << b) - 1ull));
    return r;
}

// Records the visibility of dense pair `pair` in a cull pass. Every thread of the launch calls it once, in converged
// control flow, with visible = false past n_pairs: a wavefront holds 64 consecutive pairs starting at a multiple of
// 64, so its ballot is their word. A visible pair also sets its frame-major bit (frame f of sweep s, Gaussian gl).
__device__ __forceinline__ void mark_pair(bool visible, int64_t pair, int64_t n_pairs, const us3d_sweep* sweeps, int s,
                                          int f, int gl, uint64_t* __restrict__ pair_bits,
                                          int32_t* __restrict__ word_cnt, uint32_t* __restrict__ fm_bits) {
    const uint64_t bal = __ballot(visible);
    if ((threadIdx.x & (kWave - 1)) == 0 && pair < n_pairs) {
        pair_bits[pair >> 6] = bal;
        word_cnt[pair >> 6] = (int32_t)__popcll(bal);
    }
    if (visible) {
        const int64_t fm = fm_word_base(sweeps, s, f) * 32 + gl;
        atomicOr(&fm_bits[fm >> 5], 1u << (fm & 31));
    }
}

// Emit index of an instance in tile (tx, ty) from its pair's emit base (first emit index, x0, y0, width as int bits).
__device__ __forceinline__ int64_t emit_index_rec(float4 base, int tx, int ty) {
    return (int64_t)__float_as_int(base.x) + (int64_t)(ty - __float_as_int(base.z)) * __float_as_int(base.w) +
           (tx - __float_as_int(base.y));
}

// Emit index of the instance of visible pair v in tile (tx, ty) of its frame: the pair's first emit index plus the
// tile's row-major position in the pair's rectangle (scanline ty for a ray pair, whose rectangle is one column wide).
__device__ __forceinline__ int64_t emit_index(const int32_t* __restrict__ pair_off, const int4* __restrict__ pair_rect,
                                              int v, int tx, int ty) {
    const int4 r = pair_rect[v];
    return (int64_t)pair_off[v] + (int64_t)(ty - r.y) * (r.z - r.x) + (tx - r.x);
}

// Wave-cooperative walk over the tile instances of the pairs held by the 64 lanes of a wavefront. Lane l holds
// a pair with `count` instances covering `rect` = [x0,x1)x[y0,y1) of a frame whose tiles start at tile_base,
// tiles_x per row (count = 0 for no pair), plus NP payload words. The wave's instances are numbered by an
// inclusive scan of the counts; lanes take instances k = lane, lane + 64, ..., so one pair's tiles are spread
// over the lanes instead of walked serially by one. For instance j (row-major in the rectangle) of the pair
// of lane `owner`: fn(owner payload, j, global tile). Every lane of the wave must call this together.
template <int NP, typename Fn>
__device__ __forceinline__ void for_each_wave_instance(int count, int4 rect, int tile_base, int tiles_x,
                                                       const uint32_t (&payload)[NP], Fn&& fn) {
    const int lane = threadIdx.x & (kWave - 1);
    int incl = count;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const int v = __shfl_up(incl, off, kWave);
        if (lane >= off) incl += v;
    }
    const int total = __shfl(incl, kWave - 1, kWave);
    const int rw = max(rect.z - rect.x, 1);
    for (int base = 0; base < total; base += kWave) {
        const int k = base + lane;
        int owner = 0;  // first lane whose inclusive count exceeds k
#pragma unroll
        for (int step = kWave / 2; step >= 1; step >>= 1)
            if (__shfl(incl, owner + step - 1, kWave) <= k) owner += step;
        const int o_incl = __shfl(incl, owner, kWave);
        const int o_cnt = __shfl(count, owner, kWave);
        const int o_x = __shfl(rect.x, owner, kWave), o_y = __shfl(rect.y, owner, kWave);
        const int o_w = __shfl(rw, owner, kWave);
        const int o_tb = __shfl(tile_base, owner, kWave), o_tx = __shfl(tiles_x, owner, kWave);
        uint32_t o_pay[NP];
#pragma unroll
        for (int i = 0; i < NP; ++i) o_pay[i] = (uint32_t)__shfl((int)payload[i], owner, kWave);
        if (k < total) {
            const int j = k - (o_incl - o_cnt);
            const int jy = j / o_w;
            fn(o_pay, j, o_tb + (o_y + jy) * o_tx + (o_x + j - jy * o_w));
        }
    }
}

// Counts a pair's instances into its frame's tiles without one atomic per instance: the rectangle
// [r.x, r.z) x [r.y, r.w) of tiles goes into a 2-D difference array over the frame's tiles (+1 at the first
// corner, -1 at the two corners past it along one axis, +1 past both; corners beyond the frame's last column or
// row are dropped, the prefix sums never reach them). prefix_tile_counts_kernel then turns each frame's array
// into per-tile counts.
__device__ __forceinline__ void add_rect_corners(int32_t* __restrict__ tile_diff, int tile_base, int tiles_x,
                                                 int tiles_y, int4 r) {
    int32_t* a = tile_diff + tile_base;
    atomicAdd(&a[r.y * tiles_x + r.x], 1);
    if (r.z < tiles_x) atomicAdd(&a[r.y * tiles_x + r.z], -1);
    if (r.w < tiles_y) {
        atomicAdd(&a[r.w * tiles_x + r.x], -1);
        if (r.z < tiles_x) atomicAdd(&a[r.w * tiles_x + r.z], 1);
    }
}

// One workgroup per frame: inclusive 2-D prefix sum of the frame's tile difference array, in place.
__global__ void prefix_tile_counts_kernel(const us3d_frame* __restrict__ frames, int32_t* __restrict__ tile_count);

__global__ void init_stats_kernel(us3d_stats* stats);

__global__ void fill_i32_kernel(int32_t* dst, int64_t n, int32_t v);
__global__ void fill_f32_kernel(float* dst, int64_t n, float v);

}  // namespace us3d
