// Mapper stages shared by the renderers of libus3d: the workspace layout, the ranking of the visible
// (frame, Gaussian) pairs, the emit of instances into tile buckets and the bucket build that ends in the per-bucket
// sort.
//
// A batch's pairs are numbered densely, Gaussian-major per sweep (us3d_sweep.pair_base), but only a few percent of
// them are visible (at least one instance): a freehand sweep's plane passes near a Gaussian in a handful of frames.
// Records are therefore kept only for visible pairs, under a compact index v:
//   1. the renderer's cull pass (thread per dense pair, no per-pair stores) sets one bit per pair: the wavefront's
//      ballot is the word of its 64 consecutive pairs (pair_bits, with its popcount in word_cnt), and a visible pair
//      also sets its bit of the frame-major mask (fm_bits: per frame ceil(g_count / 32) words, bit gl for Gaussian
//      g_begin + gl; one uncontended atomicOr);
//   2. exclusive scans of the words' popcounts give every visible pair its rank in both orders: v, the Gaussian-major
//      rank (a Gaussian's visible pairs are consecutive, in frame order), and its slot in vis_list, the frame-major
//      list of v (a frame's visible pairs are consecutive, in Gaussian order); list_visible_kernel writes vis_pair[v]
//      (the dense index), pair_g[v] (its Gaussian) and vis_list[slot] = v;
//   3. the renderer's setup pass (thread per v) writes the pair's instance count, tile rectangle, depth key and its
//      own per-pair record, and counts the pair's tiles by rectangle corners;
//   4. build_instances scans the counts (v order, so a Gaussian's instances are consecutive across frames), emits
//      every instance into its tile's bucket as the key pair_dk[v] << 32 | v and sorts each bucket.
// Every rank is a scan over fixed positions, so the lists are deterministic; the workspace holds 0.625 B per pair for
// the masks and their scans, and the records of at most v_capacity visible pairs.
#include <algorithm>
#include <map>
#include <mutex>
#include <utility>

#include "common.h"

namespace us3d {

// ---------------------------------------------------------------- workspace layout

int64_t pair_words(int64_t n_pairs) { return (n_pairs + 63) / 64; }

// A frame takes ceil(g_count / 32) <= g_count / 32 + 1 frame-major words and holds at least one tile.
int64_t fm_words(int64_t n_pairs, int64_t n_tiles) { return (n_pairs + 31) / 32 + n_tiles; }

us3d_layout ws_layout(int64_t n_gauss, int64_t n_pairs, int64_t n_tiles, int64_t k_cap, int64_t v_cap,
                      int64_t record_bytes, int64_t partial_bytes, int64_t mask_bytes, int64_t sect_bytes) {
    us3d_layout L{};
    int64_t off = 0;
    auto take = [&](int64_t bytes) {
        const int64_t o = off;
        off += align_up(std::max<int64_t>(bytes, 0), 256);
        return o;
    };
    const int64_t nw = pair_words(n_pairs), nf = fm_words(n_pairs, n_tiles);
    L.stats = take(sizeof(us3d_stats));
    L.queue = take(kXcds * kQueueStride * 4);
    L.pair_bits = take(nw * 8);
    L.word_cnt = take(nw * 4);
    L.word_off = take((nw + 1) * 4);
    L.fm_bits = take(nf * 4);
    L.fm_cnt = take(nf * 4);
    L.fm_off = take((nf + 1) * 4);
    L.vis_pair = take(v_cap * 4);
    L.pair_count = take(v_cap * 4);
    L.pair_off = take((v_cap + 1) * 4);
    L.pair_rect = take(v_cap * 16);
    L.pair_dk = take(v_cap * 4);
    L.pair_g = take(v_cap * 4);
    L.vis_list = take(v_cap * 4);
    L.tile_count = take(n_tiles * 4);
    L.tile_off = take((n_tiles + 1) * 4);
    L.tile_cursor = take(n_tiles * 4);
    L.seg_key = take(k_cap * 8);
    L.seg_key_tmp = take(k_cap * 8);
    L.records = take(n_gauss * record_bytes);
    L.partial = take(k_cap * partial_bytes);
    L.scan_tmp = take(std::max(std::max(scan_tmp_bytes(nw + 1), scan_tmp_bytes(nf + 1)),
                               std::max(scan_tmp_bytes(v_cap + 1), scan_tmp_bytes(n_tiles + 1))));
    L.inst_mask = take(k_cap * mask_bytes);
    L.pair_sect = take(v_cap * sect_bytes);
    L.total = off;
    return L;
}

int check_plan(const us3d_plan* plan, int64_t ws_needed) {
    if (!plan) return fail(US3D_ERR_ARGUMENT, "plan is NULL");
    if (plan->n_gauss < 0 || plan->n_frames <= 0 || plan->n_sweeps <= 0 || plan->n_tiles <= 0 || plan->n_pairs < 0)
        return fail(US3D_ERR_ARGUMENT, "plan sizes out of range");
    if (plan->n_pairs >= (int64_t(1) << 31) || plan->n_pixels >= (int64_t(1) << 31))
        return fail(US3D_ERR_ARGUMENT, "n_pairs and n_pixels must be < 2^31 (split the batch)");
    if (plan->n_frames > plan->n_tiles) return fail(US3D_ERR_ARGUMENT, "every frame needs at least one tile");
    if (plan->k_capacity < 1 || plan->k_capacity >= (int64_t(1) << 31))
        return fail(US3D_ERR_ARGUMENT, "k_capacity must be in [1, 2^31)");
    if (plan->v_capacity < 1 || plan->v_capacity >= (int64_t(1) << 31))
        return fail(US3D_ERR_ARGUMENT, "v_capacity must be in [1, 2^31)");
    if (!plan->workspace || plan->workspace_bytes < ws_needed)
        return fail(US3D_ERR_ARGUMENT, "workspace missing or smaller than the *_workspace_bytes() query");
    if (!plan->frames || !plan->frame_tile_base || !plan->sweeps)
        return fail(US3D_ERR_ARGUMENT, "plan tables are NULL");
    return US3D_OK;
}

// ---------------------------------------------------------------- small kernels

__global__ void init_stats_kernel(us3d_stats* stats) {
    stats->n_instances = 0;
    stats->max_segment = 0;
    stats->overflow = 0;
    stats->n_sorted_big = 0;
    stats->reserved_i32 = 0;
    stats->n_visible = 0;
    for (int i = 0; i < 3; ++i) stats->reserved[i] = 0;  // overflow_any is the caller's: left as it is
}

// More visible pairs than v_cap: overflow (and the sticky flag); the lists are then not written.
__global__ void vis_capacity_kernel(us3d_stats* stats, int64_t v_cap) {
    const int over = stats->n_visible > v_cap ? 1 : 0;
    stats->overflow = over;
    if (over) stats->overflow_any = 1;
}

// Instance capacity: the overflow of the visible pairs stands, or K > k_cap sets it.
__global__ void capacity_kernel(us3d_stats* stats, int64_t k_cap) {
    const int over = (stats->overflow || stats->n_instances > k_cap) ? 1 : 0;
    stats->overflow = over;
    if (over) stats->overflow_any = 1;
}

// off[n] = total (or 0 when the batch overflowed), closing an (n+1)-entry offset array.
__global__ void close_offsets_kernel(int32_t* off, int64_t n, const int64_t* total, const int32_t* overflow) {
    off[n] = (overflow && *overflow) ? 0 : (int32_t)*total;
}

__global__ void popcount_words_kernel(const uint32_t* __restrict__ bits, int64_t n, int32_t* __restrict__ cnt) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        cnt[i] = __popc(bits[i]);
}

// Emit base of every visible pair into its record (see build_instances): first emit index, rectangle x0, y0, width.
__global__ void emit_base_kernel(int64_t v_cap, const us3d_stats* __restrict__ stats, const int32_t* __restrict__ pair_off,
                                 const int4* __restrict__ pair_rect, float4* __restrict__ sect, int sect_f4,
                                 int emit_slot) {
    const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (v >= v_cap || stats->overflow || v >= stats->n_visible) return;
    const int4 r = pair_rect[v];
    sect[v * sect_f4 + emit_slot] = make_float4(__int_as_float(pair_off[v]), __int_as_float(r.x), __int_as_float(r.y),
                                                __int_as_float(r.z - r.x));
}

// Thread per dense pair: a visible pair writes its dense index and its Gaussian at its compact index v and v at its
// frame-major slot.
__global__ __launch_bounds__(256) void list_visible_kernel(
    int64_t n_pairs, const us3d_sweep* __restrict__ sweeps, int n_sweeps, const uint64_t* __restrict__ pair_bits,
    const int32_t* __restrict__ word_off, const uint32_t* __restrict__ fm_bits, const int32_t* __restrict__ fm_off,
    const us3d_stats* __restrict__ stats, int32_t* __restrict__ vis_pair, int32_t* __restrict__ pair_g,
    int32_t* __restrict__ vis_list) {
    const int64_t pair = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (pair >= n_pairs || stats->overflow) return;
    const uint64_t word = pair_bits[pair >> 6];
    const int b = (int)(pair & 63);
    if (!((word >> b) & 1ull)) return;
    const int v = word_off[pair >> 6] + __popcll(word & ((1ull << b) - 1ull));
    const int s = sweep_of_pair(sweeps, n_sweeps, pair);
    const us3d_sweep sw = sweeps[s];
    const int local = (int)(pair - sw.pair_base);
    const int gl = local / sw.f_count;
    const int f = sw.f_begin + (local - gl * sw.f_count);
    const int64_t fm = fm_word_base(sweeps, s, f) * 32 + gl;
    const int slot = fm_off[fm >> 5] + __popc(fm_bits[fm >> 5] & ((1u << (fm & 31)) - 1u));
    vis_pair[v] = (int32_t)pair;
    pair_g[v] = sw.g_begin + gl;
    vis_list[slot] = v;
}

// ---------------------------------------------------------------- emit
// Write the key (depth key << 32 | visible pair v) of every instance into its tile's bucket. Workgroup w serves
// frame w / (bands * parts), band b and part p of it (w % (bands * parts) = b * parts + p): the tile rows
// [b * rows, (b + 1) * rows), rows = ceil(tiles_y / bands), and the chunks p, p + parts, ... of the frame's run of
// vis_list. From 16 frames on every frame has at least two bands, more when the batch has few frames, so the
// launch fills the chip and a batch of uneven frames has a short tail; below 16 frames parts replace bands. Bands
// split the frame's buckets (every band reads the frame's whole run), parts split the run (their buckets are
// shared). A workgroup reads kEmitPerThread entries per thread at a time, gathers the pairs whose rectangle meets its
// rows into an LDS queue (ballot compaction) and expands them, clipped to its rows, kEmitThreads at a time with the
// wave-cooperative instance walk. With one part the band's bucket cursors live in LDS, seeded with the bucket
// offsets, so a slot costs one LDS atomic instead of a returning atomic in L2; with several parts, or more than
// kEmitLdsTiles tiles in the band, slots come from the global cursors. Slot order inside a bucket is arbitrary; the
// bucket sort by the unique keys makes the final order exact.
constexpr int kEmitThreads = 1024;
constexpr int kEmitPerThread = 8;
constexpr int kEmitChunk = kEmitThreads * kEmitPerThread;
constexpr int kEmitQueue = kEmitChunk + kEmitThreads;
constexpr int kEmitLdsTiles = 8192;
constexpr int kEmitTargetGroups = 1024;

__global__ __launch_bounds__(kEmitThreads) void emit_instances_kernel(
    const us3d_frame* __restrict__ frames, const us3d_sweep* __restrict__ sweeps, const int32_t* __restrict__ fm_off,
    const int32_t* __restrict__ vis_list, const int4* __restrict__ pair_rect, const uint32_t* __restrict__ pair_dk,
    const int32_t* __restrict__ tile_off, int32_t* __restrict__ tile_cursor, uint64_t* __restrict__ seg_key, int bands,
    int parts, const us3d_stats* stats) {
    __shared__ int32_t s_cur[kEmitLdsTiles];  // next free slot of each bucket of the band
    __shared__ int32_t s_q[kEmitQueue];       // visible pairs not yet expanded
    __shared__ int s_n;
    if (stats->overflow) return;
    const int f = blockIdx.x / (bands * parts), band = (blockIdx.x % (bands * parts)) / parts;
    const int part = blockIdx.x % parts;
    const us3d_frame& fr = frames[f];
    const int rows = (fr.tiles_y + bands - 1) / bands;
    const int y_lo = band * rows, y_hi = min(fr.tiles_y, y_lo + rows);
    if (y_lo >= y_hi) return;  // workgroup-uniform: a band past the frame's last tile row
    const us3d_sweep sw = sweeps[fr.sweep];
    const int64_t fw = fm_word_base(sweeps, fr.sweep, f);
    const int lo = fm_off[fw], n_list = fm_off[fw + ((sw.g_count + 31) >> 5)] - lo;
    const int32_t* list = vis_list + lo;
    const int band_base = fr.tile_base + y_lo * fr.tiles_x;  // first tile of the band
    const int n_tiles = (y_hi - y_lo) * fr.tiles_x;
    const bool local = parts == 1 && n_tiles <= kEmitLdsTiles;  // workgroup-uniform
    const int tid = threadIdx.x, lane = tid & (kWave - 1);
    if (local)
        for (int i = tid; i < n_tiles; i += kEmitThreads) s_cur[i] = tile_off[band_base + i];
    if (tid == 0) s_n = 0;
    __syncthreads();
    const uint64_t below = (1ull << lane) - 1ull;
    // parts split the run into chunks of whole kEmitThreads strides, at most kEmitChunk entries each
    const int chunk = parts == 1 ? kEmitChunk
                                 : min(kEmitChunk, max(1, (n_list + parts * kEmitThreads - 1) / (parts * kEmitThreads)) *
                                                       kEmitThreads);
    for (int c0 = part * chunk; c0 < n_list; c0 += parts * chunk) {
        const int c1 = min(n_list, c0 + chunk);
        int pr[kEmitPerThread];
#pragma unroll
        for (int k = 0; k < kEmitPerThread; ++k) {
            const int idx = c0 + k * kEmitThreads + tid;
            pr[k] = idx < c1 ? list[idx] : -1;
            if (bands > 1 && pr[k] >= 0) {  // keep only pairs whose rectangle meets the band's rows
                const int4 r = pair_rect[pr[k]];
                if (r.w <= y_lo || r.y >= y_hi) pr[k] = -1;
            }
        }
#pragma unroll
        for (int k = 0; k < kEmitPerThread; ++k) {
            const uint64_t bal = __ballot(pr[k] >= 0);
            int wbase = 0;
            if (lane == 0 && bal) wbase = atomicAdd(&s_n, (int)__popcll(bal));
            wbase = __shfl(wbase, 0, kWave);
            if (pr[k] >= 0) s_q[wbase + (int)__popcll(bal & below)] = pr[k];
        }
        __syncthreads();
        int n = s_n;  // workgroup-uniform
        const bool last = c0 + parts * chunk >= n_list;
        while (n >= kEmitThreads || (last && n > 0)) {
            const int take = min(n, kEmitThreads);
            n -= take;  // expand the queue's last `take` entries, so the rest stays in place
            const int v = tid < take ? s_q[n + tid] : -1;
            int count = 0;
            int4 rect = make_int4(0, 0, 0, 0);
            uint32_t pay[2] = {0u, 0u};  // visible pair, depth key
            if (v >= 0) {
                const int4 full = pair_rect[v];
                rect = make_int4(full.x, max(full.y, y_lo), full.z, min(full.w, y_hi));  // clipped to the band
                count = (rect.z - rect.x) * (rect.w - rect.y);
                pay[0] = (uint32_t)v;
                pay[1] = pair_dk[v];
            }
            for_each_wave_instance<2>(count, rect, fr.tile_base, fr.tiles_x, pay,
                                      [&](const uint32_t* o, int, int tile) {
                                          const int slot = local ? atomicAdd(&s_cur[tile - band_base], 1)
                                                                 : tile_off[tile] + atomicAdd(&tile_cursor[tile], 1);
                                          seg_key[slot] = ((uint64_t)o[1] << 32) | o[0];
                                      });
        }
        __syncthreads();  // every read of the queue is done
        if (tid == 0) s_n = n;
        __syncthreads();
    }
}

// ---------------------------------------------------------------- host orchestration

// Workgroups of a splat launch: one per tile, or, with the per-XCD work queues, as many as the device holds at once
// (occupancy of the kernel times the CU count, cached per kernel and device; never more than the tiles). With the
// queues their heads are cleared first, by a kernel, so a captured graph replays correctly.
int tile_grid(const void* kernel, int threads, const us3d_plan* plan, int32_t* heads, hipStream_t stream) {
    if (!heads) return plan->n_tiles;
    static std::mutex mu;
    static std::map<std::pair<const void*, int>, int> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    int groups = 0;
    {
        std::lock_guard<std::mutex> lock(mu);
        auto it = cache.find({kernel, dev});
        if (it != cache.end()) {
            groups = it->second;
        } else {
            int per_cu = 0, cus = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess) per_cu = 1;
            if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 1;
            groups = std::max(1, per_cu) * std::max(1, cus);
            cache[{kernel, dev}] = groups;
        }
    }
    hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(256), 0, stream, heads, (int64_t)kXcds * kQueueStride, 0);
    return std::min(groups, plan->n_tiles);
}


int begin_pairs(const us3d_plan* plan, const us3d_layout& L, hipStream_t stream) {
    hipLaunchKernelGGL(init_stats_kernel, dim3(1), dim3(1), 0, stream, ws_at<us3d_stats>(plan, L.stats));
    hipLaunchKernelGGL(fill_i32_kernel, dim3(blocks_for(plan->n_tiles, 256)), dim3(256), 0, stream,
                       ws_at<int32_t>(plan, L.tile_count), (int64_t)plan->n_tiles, 0);
    const int64_t nf = fm_words(plan->n_pairs, plan->n_tiles);
    hipLaunchKernelGGL(fill_i32_kernel, dim3(std::min<unsigned>(blocks_for(nf, 256), 4096u)), dim3(256), 0, stream,
                       ws_at<int32_t>(plan, L.fm_bits), nf, 0);
    return check_launch("fill_i32_kernel");
}

int rank_visible_pairs(const us3d_plan* plan, const us3d_layout& L, hipStream_t stream) {
    int rc;
    us3d_stats* stats = ws_at<us3d_stats>(plan, L.stats);
    void* tmp = ws_at<void>(plan, L.scan_tmp);
    const int64_t np = plan->n_pairs;
    if (np == 0) return US3D_OK;
    const int64_t nw = pair_words(np), nf = fm_words(np, plan->n_tiles);
    int32_t* word_off = ws_at<int32_t>(plan, L.word_off);
    int32_t* fm_off = ws_at<int32_t>(plan, L.fm_off);
    if ((rc = scan_i32(ws_at<int32_t>(plan, L.word_cnt), word_off, nw, &stats->n_visible, tmp, stream))) return rc;
    hipLaunchKernelGGL(close_offsets_kernel, dim3(1), dim3(1), 0, stream, word_off, nw, &stats->n_visible,
                       (const int32_t*)nullptr);
    hipLaunchKernelGGL(popcount_words_kernel, dim3(std::min<unsigned>(blocks_for(nf, 256), 4096u)), dim3(256), 0,
                       stream, ws_at<uint32_t>(plan, L.fm_bits), nf, ws_at<int32_t>(plan, L.fm_cnt));
    if ((rc = check_launch("popcount_words_kernel"))) return rc;
    if ((rc = scan_i32(ws_at<int32_t>(plan, L.fm_cnt), fm_off, nf, &stats->reserved[1], tmp, stream))) return rc;
    hipLaunchKernelGGL(close_offsets_kernel, dim3(1), dim3(1), 0, stream, fm_off, nf, &stats->reserved[1],
                       (const int32_t*)nullptr);
    hipLaunchKernelGGL(vis_capacity_kernel, dim3(1), dim3(1), 0, stream, stats, plan->v_capacity);
    if ((rc = check_launch("vis_capacity_kernel"))) return rc;
    {
        StageTimer timer_("list_visible_kernel", stream);
        hipLaunchKernelGGL(list_visible_kernel, dim3(blocks_for(np, 256)), dim3(256), 0, stream, np, plan->sweeps,
                           plan->n_sweeps, ws_at<uint64_t>(plan, L.pair_bits), word_off,
                           ws_at<uint32_t>(plan, L.fm_bits), fm_off, (const us3d_stats*)stats,
                           ws_at<int32_t>(plan, L.vis_pair), ws_at<int32_t>(plan, L.pair_g),
                           ws_at<int32_t>(plan, L.vis_list));
    }
    return check_launch("list_visible_kernel");
}

int build_instances(const us3d_plan* plan, const us3d_layout& L, hipStream_t stream, int sect_f4, int emit_slot) {
    int rc;
    us3d_stats* stats = ws_at<us3d_stats>(plan, L.stats);
    int32_t* pair_off = ws_at<int32_t>(plan, L.pair_off);
    int32_t* tile_count = ws_at<int32_t>(plan, L.tile_count);
    int32_t* tile_off = ws_at<int32_t>(plan, L.tile_off);
    void* tmp = ws_at<void>(plan, L.scan_tmp);
    const int64_t np = plan->n_pairs, vc = plan->v_capacity;
    if (np > 0) {
        hipLaunchKernelGGL(prefix_tile_counts_kernel, dim3(plan->n_frames), dim3(256), 0, stream, plan->frames,
                           tile_count);
        if ((rc = check_launch("prefix_tile_counts_kernel"))) return rc;
        if ((rc = scan_i32(ws_at<int32_t>(plan, L.pair_count), pair_off, vc, &stats->n_instances, tmp, stream)))
            return rc;
    }
    hipLaunchKernelGGL(capacity_kernel, dim3(1), dim3(1), 0, stream, stats, plan->k_capacity);
    if ((rc = check_launch("capacity_kernel"))) return rc;
    if (np > 0) {
        hipLaunchKernelGGL(close_offsets_kernel, dim3(1), dim3(1), 0, stream, pair_off, vc, &stats->n_instances,
                           (const int32_t*)nullptr);
        if (sect_f4 > 0) {
            hipLaunchKernelGGL(emit_base_kernel, dim3(blocks_for(vc, 256)), dim3(256), 0, stream, vc,
                               (const us3d_stats*)stats, (const int32_t*)pair_off, ws_at<int4>(plan, L.pair_rect),
                               ws_at<float4>(plan, L.pair_sect), sect_f4, emit_slot);
            if ((rc = check_launch("emit_base_kernel"))) return rc;
        }
    } else {
        hipLaunchKernelGGL(fill_i32_kernel, dim3(blocks_for(vc + 1, 256)), dim3(256), 0, stream, pair_off, vc + 1, 0);
    }
    int64_t* tile_total = &stats->reserved[0];
    if ((rc = scan_i32(tile_count, tile_off, plan->n_tiles, tile_total, tmp, stream))) return rc;
    hipLaunchKernelGGL(close_offsets_kernel, dim3(1), dim3(1), 0, stream, tile_off, (int64_t)plan->n_tiles,
                       tile_total, (const int32_t*)&stats->overflow);
    int32_t* tile_cursor = ws_at<int32_t>(plan, L.tile_cursor);
    hipLaunchKernelGGL(fill_i32_kernel, dim3(blocks_for(plan->n_tiles, 256)), dim3(256), 0, stream, tile_cursor,
                       (int64_t)plan->n_tiles, 0);
    if ((rc = check_launch("fill_i32_kernel"))) return rc;
    if (np > 0) {
        {
            StageTimer timer_("emit_instances_kernel", stream);
            // enough workgroups for the chip when the batch has few frames: bands (LDS cursors) down to 16
            // frames, below that parts of the frame's list (global cursors), since every band rereads the list;
            // at least two bands per frame, which shortens the tail of large batches of uneven frames (multisweep,
            // 1754 frames: 6.39 -> 6.10 ms)
            const int groups = std::max(kEmitTargetGroups, 2 * plan->n_frames);
            const int split = std::min(32, (groups + plan->n_frames - 1) / plan->n_frames);
            const int bands = plan->n_frames >= 16 ? std::max(1, split) : 1;
            const int parts = plan->n_frames >= 16 ? 1 : std::max(1, split);
            hipLaunchKernelGGL(emit_instances_kernel, dim3((unsigned)(plan->n_frames * bands * parts)),
                               dim3(kEmitThreads), 0, stream, plan->frames, plan->sweeps,
                               ws_at<int32_t>(plan, L.fm_off), ws_at<int32_t>(plan, L.vis_list),
                               ws_at<int4>(plan, L.pair_rect), ws_at<uint32_t>(plan, L.pair_dk), tile_off, tile_cursor,
                               ws_at<uint64_t>(plan, L.seg_key), bands, parts, (const us3d_stats*)stats);
        }
        if ((rc = check_launch("emit_instances_kernel"))) return rc;
    }
    return sort_segments(ws_at<uint64_t>(plan, L.seg_key), ws_at<uint64_t>(plan, L.seg_key_tmp), tile_off, plan->n_tiles,
                         stats, plan->k_capacity, stream);
}

}  // namespace us3d
