// Shared pipeline pieces of libus3d: error state, device-wide exclusive scan, per-segment LDS sort,
// fill kernels.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"

namespace us3d {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
    set_error(msg);
    return code;
}

int check_launch(const char* what) {
    const hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        return fail(US3D_ERR_HIP, std::string(what) + ": " + hipGetErrorString(err));
    }
    return US3D_OK;
}

// ---------------------------------------------------------------- per-kernel timing

namespace {
struct Timing {
    std::mutex mu;
    bool enabled = false;
    std::map<std::string, std::vector<std::pair<hipEvent_t, hipEvent_t>>> events;
    std::map<std::string, hipEvent_t> open;
};
Timing& timing() {
    static Timing t;
    return t;
}
}  // namespace

bool timing_enabled() { return timing().enabled; }

void timing_mark(const char* name, hipStream_t stream, bool begin) {
    Timing& T = timing();
    std::lock_guard<std::mutex> lock(T.mu);
    hipEvent_t ev;
    if (hipEventCreate(&ev) != hipSuccess) return;
    hipEventRecord(ev, stream);
    if (begin) {
        T.open[name] = ev;
    } else {
        auto it = T.open.find(name);
        if (it == T.open.end()) {
            hipEventDestroy(ev);
            return;
        }
        T.events[name].push_back({it->second, ev});
        T.open.erase(it);
    }
}

// ---------------------------------------------------------------- fill

__global__ void fill_i32_kernel(int32_t* dst, int64_t n, int32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = v;
}

constexpr int kPrefixThreads = 256;
constexpr int kPrefixLds = 16384;  // tiles of a frame summed in LDS; larger frames are summed in place

__global__ __launch_bounds__(kPrefixThreads) void prefix_tile_counts_kernel(const us3d_frame* __restrict__ frames,
                                                                            int32_t* __restrict__ tile_count) {
    __shared__ int32_t s_a[kPrefixLds];
    const us3d_frame& fr = frames[blockIdx.x];
    const int tx = fr.tiles_x, ty = fr.tiles_y;
    const int n = tx * ty;
    int32_t* g = tile_count + fr.tile_base;
    const bool lds = n <= kPrefixLds;
    int32_t* a = lds ? s_a : g;
    if (lds) {
        for (int i = threadIdx.x; i < n; i += kPrefixThreads) s_a[i] = g[i];
        __syncthreads();
    }
    for (int r = threadIdx.x; r < ty; r += kPrefixThreads) {  // along each row
        int32_t s = 0;
        for (int x = 0; x < tx; ++x) {
            s += a[r * tx + x];
            a[r * tx + x] = s;
        }
    }
    __syncthreads();
    for (int c = threadIdx.x; c < tx; c += kPrefixThreads) {  // then down each column
        int32_t s = 0;
        for (int y = 0; y < ty; ++y) {
            s += a[y * tx + c];
            a[y * tx + c] = s;
        }
    }
    if (lds) {
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += kPrefixThreads) g[i] = s_a[i];
    }
}

__global__ void fill_f32_kernel(float* dst, int64_t n, float v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = v;
}

// ---------------------------------------------------------------- exclusive scan
// Reduce-then-scan over blocks of 2048 elements (256 threads x 8 contiguous items), int64 accumulation.
// Level 0 reads int32 counts and writes int32 offsets; block totals recurse as int64.

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanBlock = kScanThreads * kScanItems;
static_assert(kScanItems == 8, "the vector path of scan_block_kernel moves 8 int32 items as two int4");

template <typename TIn, typename TOut>
__global__ __launch_bounds__(kScanThreads) void scan_block_kernel(const TIn* __restrict__ in, TOut* __restrict__ out,
                                                                  int64_t n, int64_t* __restrict__ block_sums,
                                                                  int64_t* __restrict__ total) {
    __shared__ int64_t wave_tot[kScanThreads / kWave];
    const int tid = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * kScanBlock + (int64_t)tid * kScanItems;
    int64_t v[kScanItems];
    int64_t sum = 0;
    // int32 items of a full chunk are read as two 16 B loads (chunks start 32 B apart from a 16 B aligned base)
    bool vec_in = false;
    if constexpr (sizeof(TIn) == 4) {
        vec_in = base + kScanItems <= n && (reinterpret_cast<uintptr_t>(in) & 15) == 0;
        if (vec_in) {
            const int4 a = *reinterpret_cast<const int4*>(in + base);
            const int4 b = *reinterpret_cast<const int4*>(in + base + 4);
            v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
            v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
        }
    }
    if (!vec_in) {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const int64_t i = base + k;
            v[k] = (i < n) ? (int64_t)in[i] : 0;
        }
    }
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) sum += v[k];
    // inclusive wave scan of per-thread sums
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    int64_t incl = sum;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const int64_t o = __shfl_up(incl, off, kWave);
        if (lane >= off) incl += o;
    }
    if (lane == kWave - 1) wave_tot[wave] = incl;
    __syncthreads();
    int64_t wave_base = 0;
    for (int w = 0; w < wave; ++w) wave_base += wave_tot[w];
    int64_t run = wave_base + incl - sum;  // exclusive prefix of this thread
    bool vec_out = false;
    if constexpr (sizeof(TOut) == 4) {
        vec_out = base + kScanItems <= n && (reinterpret_cast<uintptr_t>(out) & 15) == 0;
        if (vec_out) {
            int32_t o[kScanItems];
#pragma unroll
            for (int k = 0; k < kScanItems; ++k) {
                o[k] = (int32_t)run;
                run += v[k];
            }
            *reinterpret_cast<int4*>(out + base) = make_int4(o[0], o[1], o[2], o[3]);
            *reinterpret_cast<int4*>(out + base + 4) = make_int4(o[4], o[5], o[6], o[7]);
        }
    }
    if (!vec_out) {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const int64_t i = base + k;
            if (i < n) out[i] = (TOut)run;
            run += v[k];
        }
    }
    if (tid == kScanThreads - 1) {
        int64_t block_total = 0;
        for (int w = 0; w < kScanThreads / kWave; ++w) block_total += wave_tot[w];
        if (block_sums) block_sums[blockIdx.x] = block_total;
        if (total) *total = block_total;
    }
}

template <typename TOut>
__global__ __launch_bounds__(kScanThreads) void scan_add_kernel(TOut* __restrict__ out, int64_t n,
                                                                const int64_t* __restrict__ block_off) {
    const int64_t b = blockIdx.x;
    const int64_t add = block_off[b];
    const int64_t base = b * kScanBlock;
    for (int k = threadIdx.x; k < kScanBlock; k += kScanThreads) {
        const int64_t i = base + k;
        if (i < n) out[i] = (TOut)((int64_t)out[i] + add);
    }
}

int64_t scan_tmp_bytes(int64_t n) {
    int64_t bytes = 0;
    int64_t m = n;
    while (m > kScanBlock) {
        const int64_t nb = (m + kScanBlock - 1) / kScanBlock;
        bytes += 2 * align_up(nb * (int64_t)sizeof(int64_t), 256);
        m = nb;
    }
    return bytes + 256;
}

template <typename TIn, typename TOut>
static int scan_rec(const TIn* in, TOut* out, int64_t n, int64_t* total, char* tmp, hipStream_t stream) {
    if (n <= kScanBlock) {
        hipLaunchKernelGGL((scan_block_kernel<TIn, TOut>), dim3(1), dim3(kScanThreads), 0, stream, in, out, n,
                           (int64_t*)nullptr, total);
        return check_launch("scan_block_kernel");
    }
    const int64_t nb = (n + kScanBlock - 1) / kScanBlock;
    int64_t* sums = reinterpret_cast<int64_t*>(tmp);
    int64_t* offs = reinterpret_cast<int64_t*>(tmp + align_up(nb * (int64_t)sizeof(int64_t), 256));
    char* next = tmp + 2 * align_up(nb * (int64_t)sizeof(int64_t), 256);
    hipLaunchKernelGGL((scan_block_kernel<TIn, TOut>), dim3((unsigned)nb), dim3(kScanThreads), 0, stream, in, out, n,
                       sums, (int64_t*)nullptr);
    int rc = check_launch("scan_block_kernel");
    if (rc) return rc;
    rc = scan_rec<int64_t, int64_t>(sums, offs, nb, total, next, stream);
    if (rc) return rc;
    hipLaunchKernelGGL((scan_add_kernel<TOut>), dim3((unsigned)nb), dim3(kScanThreads), 0, stream, out, n, offs);
    return check_launch("scan_add_kernel");
}

int scan_i32(const int32_t* in, int32_t* out, int64_t n, int64_t* total, void* tmp, hipStream_t stream) {
    if (n <= 0) return fail(US3D_ERR_ARGUMENT, "scan_i32: n must be positive");
    return scan_rec<int32_t, int32_t>(in, out, n, total, reinterpret_cast<char*>(tmp), stream);
}

// ---------------------------------------------------------------- per-segment sort
// A segment holds instances as 64-bit keys depth_key << 32 | v (v: the instance's visible pair), unique since v is,
// and sorted without a payload. Inside a tile every instance belongs to one frame, where v ascends with the
// Gaussian index, so the order is that of (depth key, Gaussian).
// Segments of up to kWaveSortCap keys are sorted by one wavefront each, four per workgroup: a bitonic network
// over registers (up to 8 keys per lane) with cross-lane shuffles, no LDS. Segments of up to kWaveSortCapMid
// keys get two wavefronts (two register-sorted halves, one LDS stage, a register merge). Longer segments get
// one workgroup each (a workgroup examines kLongSortSegments consecutive segments and sorts the long ones): runs
// of kWaveSortCap keys are sorted by the wavefronts in registers the same way and joined by a bitonic merge whose
// long-distance stages run in LDS and the rest in registers (kMergeCap keys); beyond kMergeCap, chunks sorted that
// way are merged in global memory (ping-pong with the tmp array), where an element's place is its index in its run
// plus its rank in the partner run, so a pass is fully parallel.

constexpr int kSortThreads = 256;
constexpr int kWaveSortCap = 512;      // segments sorted by one wavefront, <= 8 keys per lane
constexpr int kWaveSortCapMid = 1024;  // segments sorted by two wavefronts
constexpr int kMergeCap = 4096;        // segments (and chunks) sorted in one workgroup's LDS
constexpr int kWaveSortWaves = 4;
constexpr int kLongSortSegments = 4;   // segments examined by one workgroup of the long-segment sort


// Bitonic sort of 64*E keys held by one wavefront in registers: lane l holds elements
// l*E .. l*E+E-1. The network is unrolled completely (the partner distance and direction of every exchange
// are compile-time constants, so shuffles and register indices are static). Stages whose partner index
// differs only in the low log2(E) bits exchange inside a lane; the others exchange with lane ^ (j / E)
// through cross-lane shuffles. Keys must be unique.
// kDpp: shuffles through shfl_xor_pow2 (only where the network unrolls completely, so every distance is a
// constant); otherwise ds_bpermute.
template <int E, bool kDpp>
__device__ __forceinline__ void bitonic_regs(uint64_t (&k)[E], int t) {
    constexpr int N = kWave * E;
#pragma unroll
    for (int kk = 2; kk <= N; kk <<= 1) {
#pragma unroll
        for (int j = kk >> 1; j > 0; j >>= 1) {
            if (j < E) {
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const int p = e ^ j;
                    if (p > e) {
                        const bool asc = (((t * E + e) & kk) == 0);
                        if ((k[e] > k[p]) == asc) {
                            const uint64_t tk = k[e]; k[e] = k[p]; k[p] = tk;
                        }
                    }
                }
            } else {
                const int m = j / E;
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const int i = t * E + e;
                    const int klo = (int)(uint32_t)k[e], khi = (int)(uint32_t)(k[e] >> 32);
                    const uint32_t lo = (uint32_t)(kDpp ? shfl_xor_pow2(klo, m) : __shfl_xor(klo, m, kWave));
                    const uint32_t hi = (uint32_t)(kDpp ? shfl_xor_pow2(khi, m) : __shfl_xor(khi, m, kWave));
                    const uint64_t ok = ((uint64_t)hi << 32) | lo;
                    const bool asc = (i & kk) == 0;
                    const bool lower = (i & j) == 0;
                    // the lower element keeps the min when ascending, the max when descending; keys are unique
                    // (padding keys are equal, and taking an equal key changes nothing), so the max is taken
                    // exactly when the partner is not smaller, and the choice is a mask, not a branch
                    const bool take = (ok < k[e]) != (lower != asc);
                    if (take) k[e] = ok;
                }
            }
        }
    }
}

template <int E>
__device__ __forceinline__ void wave_bitonic_regs(uint64_t (&k)[E]) {
    bitonic_regs<E, (E <= 8)>(k, (int)(threadIdx.x & (kWave - 1)));
}

// The network sorts whatever it is given, so the keys are loaded lane-interleaved (element e*64 + lane: one
// 512-byte row per load) and stored in the network's layout (lane l's elements l*E .. l*E+E-1). Every lane has
// loaded all its keys before any store.
template <int E>
__device__ __forceinline__ void wave_sort_segment(uint64_t* __restrict__ keys, int n) {
    const int lane = threadIdx.x & (kWave - 1);
    uint64_t k[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int i = e * kWave + lane;
        k[e] = (i < n) ? keys[i] : ~0ull;
    }
    wave_bitonic_regs<E>(k);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int i = lane * E + e;
        if (i < n) keys[i] = k[e];
    }
}

// Longest segment into stats->max_segment: a block-wide max, then one atomic per block.
__global__ __launch_bounds__(256) void segment_max_kernel(const int32_t* __restrict__ seg_off, int n_segments,
                                                          us3d_stats* stats) {
    int m = 0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n_segments; i += gridDim.x * 256)
        m = max(m, seg_off[i + 1] - seg_off[i]);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) m = max(m, __shfl_xor(m, off, kWave));
    __shared__ int red[256 / kWave];
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 256 / kWave; ++w) m = max(m, red[w]);
        if (m > 0) atomicMax((unsigned long long*)&stats->max_segment, (unsigned long long)m);
    }
}

__global__ __launch_bounds__(kWaveSortWaves * kWave) void sort_small_segments_kernel(
    uint64_t* __restrict__ keys, const int32_t* __restrict__ seg_off, int n_segments,
    const us3d_stats* __restrict__ stats) {
    if (stats->overflow) return;
    // wave-uniform, and stated so: the segment's size then selects the register network by a scalar branch
    const int seg = blockIdx.x * kWaveSortWaves + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    if (seg >= n_segments) return;
    const int s0 = seg_off[seg];
    const int n = seg_off[seg + 1] - s0;
    if (n <= 1 || n > kWaveSortCap) return;
    if (n <= kWave) wave_sort_segment<1>(keys + s0, n);
    else if (n <= 2 * kWave) wave_sort_segment<2>(keys + s0, n);
    else if (n <= 4 * kWave) wave_sort_segment<4>(keys + s0, n);
    else wave_sort_segment<8>(keys + s0, n);
}

// Ascending bitonic merge of a bitonic sequence of 64*E keys held by a wavefront (lane l holds elements
// l*E .. l*E+E-1): the stages j = 32E .. 1 of the network's last level, in registers, exchanges across lanes
// through the same shuffles as bitonic_regs.
template <int E>
__device__ __forceinline__ void wave_bitonic_merge_regs(uint64_t (&k)[E], int t) {
    constexpr int N = kWave * E;
#pragma unroll
    for (int j = N >> 1; j > 0; j >>= 1) {
        if (j < E) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int p = e ^ j;
                if (p > e && k[e] > k[p]) {
                    const uint64_t tk = k[e]; k[e] = k[p]; k[p] = tk;
                }
            }
        } else {
            const int m = j / E;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int i = t * E + e;
                const int klo = (int)(uint32_t)k[e], khi = (int)(uint32_t)(k[e] >> 32);
                const uint32_t lo = (uint32_t)shfl_xor_pow2(klo, m);
                const uint32_t hi = (uint32_t)shfl_xor_pow2(khi, m);
                const uint64_t ok = ((uint64_t)hi << 32) | lo;
                const bool take = (ok < k[e]) != ((i & j) != 0);  // as in bitonic_regs
                if (take) k[e] = ok;
            }
        }
    }
}

// Compare-exchange of LDS entries a < b, ascending.
__device__ __forceinline__ void lds_cas(uint64_t* sk, int a, int b) {
    const uint64_t x = sk[a], y = sk[b];
    if (x > y) {
        sk[a] = y;
        sk[b] = x;
    }
}

// Segments of kWaveSortCap+1 .. kWaveSortCapMid keys, kLongSortSegments consecutive ones per two-wavefront
// workgroup: each wavefront sorts a 512-key half in registers, one stage in LDS compares element i of the first
// half with element 1023-i (the two sorted halves become bitonic halves, the smaller keys in the first), and
// each wavefront finishes its half with the register merge and writes it out.
constexpr int kMidThreads = 2 * kWave;

__global__ __launch_bounds__(kMidThreads) void sort_mid_segments_kernel(
    uint64_t* __restrict__ keys, const int32_t* __restrict__ seg_off, int n_segments,
    const us3d_stats* __restrict__ stats) {
    __shared__ uint64_t sk[kWaveSortCapMid];
    if (stats->overflow) return;
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    for (int it = 0; it < kLongSortSegments; ++it) {
        const int seg = blockIdx.x * kLongSortSegments + it;
        if (seg >= n_segments) return;  // workgroup-uniform
        const int s0 = seg_off[seg], n = seg_off[seg + 1] - s0;
        if (n <= kWaveSortCap || n > kWaveSortCapMid) continue;
        __syncthreads();  // a previous segment's reads of sk are done
        uint64_t k[8];
        const int b = wave * kWaveSortCap + lane * 8;
#pragma unroll
        for (int e = 0; e < 8; ++e) {  // the wavefront's half, loaded lane-interleaved (any order sorts)
            const int i = wave * kWaveSortCap + e * kWave + lane;
            k[e] = (i < n) ? keys[s0 + i] : ~0ull;
        }
        wave_bitonic_regs<8>(k);
#pragma unroll
        for (int e = 0; e < 8; ++e) sk[b + e] = k[e];
        __syncthreads();
        for (int q = threadIdx.x; q < kWaveSortCap; q += kMidThreads) lds_cas(sk, q, kWaveSortCapMid - 1 - q);
        __syncthreads();
#pragma unroll
        for (int e = 0; e < 8; ++e) k[e] = sk[b + e];
        wave_bitonic_merge_regs<8>(k, lane);
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int i = b + e;
            if (i < n) keys[s0 + i] = k[e];
        }
    }
}

// Sorts the keys of n (1 <= n <= kMergeCap) instances in LDS. The data is padded with ~0 keys to 1024, 2048 or
// 4096 entries; the wavefronts sort runs of kWaveSortCap ascending in registers; then every level of the
// bitonic merge doubles the sorted blocks: a first stage compares element i of a block's first half with
// element kk-1-i (two ascending runs become two bitonic halves, the smaller elements in the first), then the
// half-cleaners with partner distance >= kWaveSortCap run in LDS and the remaining ones in registers, each
// wavefront merging whole 512-entry blocks. Ends with the sorted keys in sk (the first n are the data).
__device__ __forceinline__ void lds_merge_sort(const uint64_t* __restrict__ keys, int n, uint64_t* sk) {
    constexpr int kWaves = kSortThreads / kWave;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    int padded = 2 * kWaveSortCap;
    while (padded < n) padded <<= 1;
    const int runs = padded / kWaveSortCap;
    for (int r = wave; r < runs; r += kWaves) {
        uint64_t k[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {  // the run, loaded lane-interleaved (any order sorts)
            const int i = r * kWaveSortCap + e * kWave + lane;
            k[e] = (i < n) ? keys[i] : ~0ull;
        }
        if (r * kWaveSortCap < n) wave_bitonic_regs<8>(k);  // wave-uniform
#pragma unroll
        for (int e = 0; e < 8; ++e) sk[r * kWaveSortCap + lane * 8 + e] = k[e];
    }
    __syncthreads();
    const int half = padded >> 1;  // compare-exchanges per stage
    for (int kk = 2 * kWaveSortCap; kk <= padded; kk <<= 1) {
        for (int q = tid; q < half; q += kSortThreads) {  // first stage: i against kk-1-i within the block
            const int base = (q / (kk >> 1)) * kk, off = q % (kk >> 1);
            lds_cas(sk, base + off, base + kk - 1 - off);
        }
        __syncthreads();
        for (int j = kk >> 2; j >= kWaveSortCap; j >>= 1) {
            for (int q = tid; q < half; q += kSortThreads) {
                const int lo = ((q & ~(j - 1)) << 1) | (q & (j - 1));
                lds_cas(sk, lo, lo + j);
            }
            __syncthreads();
        }
        for (int r = wave; r < runs; r += kWaves) {  // distances < kWaveSortCap inside each 512-entry block
            uint64_t k[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) k[e] = sk[r * kWaveSortCap + lane * 8 + e];
            wave_bitonic_merge_regs<8>(k, lane);
#pragma unroll
            for (int e = 0; e < 8; ++e) sk[r * kWaveSortCap + lane * 8 + e] = k[e];
        }
        __syncthreads();
    }
}

// Sorts segment [s0, s0 + n) (n > kWaveSortCapMid) with the whole workgroup.
__device__ __forceinline__ void sort_long_segment_body(uint64_t* __restrict__ keys, uint64_t* __restrict__ keys_tmp,
                                                       int s0, int n, us3d_stats* stats, uint64_t* sk) {
    uint64_t* k = keys + s0;
    if (n > kMergeCap && threadIdx.x == 0) atomicAdd(&stats->n_sorted_big, 1);
    // chunks of kMergeCap sorted in LDS (one chunk when n <= kMergeCap)
    for (int c = 0; c < n; c += kMergeCap) {
        const int m = min(kMergeCap, n - c);
        lds_merge_sort(k + c, m, sk);
        for (int i = threadIdx.x; i < m; i += kSortThreads) k[c + i] = sk[i];
        __syncthreads();
    }
    if (n <= kMergeCap) return;
    uint64_t* src_k = k;
    uint64_t* dst_k = keys_tmp + s0;
    for (int width = kMergeCap; width < n; width <<= 1) {
        for (int i = threadIdx.x; i < n; i += kSortThreads) {
            const int run = i / width;
            const int pair_start = (run & ~1) * width;
            const int partner_start = (run & 1) ? pair_start : pair_start + width;
            const int partner_len = max(0, min(width, n - partner_start));
            const uint64_t key = src_k[i];
            int lo = 0, hi = partner_len;  // count of partner keys < key
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (src_k[partner_start + mid] < key) lo = mid + 1; else hi = mid;
            }
            dst_k[pair_start + (i - run * width) + lo] = key;
        }
        __syncthreads();
        uint64_t* tk = src_k; src_k = dst_k; dst_k = tk;
    }
    if (src_k != k)
        for (int i = threadIdx.x; i < n; i += kSortThreads) k[i] = src_k[i];
}

// Segments kLongSortSegments * blockIdx.x + (0 .. kLongSortSegments-1): each one longer than kWaveSortCapMid
// is sorted by the whole workgroup.
__global__ __launch_bounds__(kSortThreads) void sort_segments_kernel(uint64_t* __restrict__ keys,
                                                                     uint64_t* __restrict__ keys_tmp,
                                                                     const int32_t* __restrict__ seg_off,
                                                                     int n_segments, us3d_stats* stats) {
    __shared__ uint64_t sk[kMergeCap];
    if (stats->overflow) return;
    for (int i = 0; i < kLongSortSegments; ++i) {
        const int seg = blockIdx.x * kLongSortSegments + i;
        if (seg >= n_segments) return;  // workgroup-uniform
        const int s0 = seg_off[seg], n = seg_off[seg + 1] - s0;
        if (n <= kWaveSortCapMid) continue;  // sorted by the wavefront kernels
        __syncthreads();  // a previous segment's last reads of sk are done
        sort_long_segment_body(keys, keys_tmp, s0, n, stats, sk);
    }
}

int sort_segments(uint64_t* keys, uint64_t* keys_tmp, const int32_t* seg_off, int n_segments, us3d_stats* stats,
                  int64_t k_capacity, hipStream_t stream) {
    (void)k_capacity;
    if (n_segments <= 0) return US3D_OK;
    hipLaunchKernelGGL(segment_max_kernel, dim3(std::min<unsigned>(blocks_for(n_segments, 256), 1024u)), dim3(256), 0,
                       stream, seg_off, n_segments, stats);
    if (int rc0 = check_launch("segment_max_kernel")) return rc0;
    {
        StageTimer timer("sort_small_segments_kernel", stream);
        hipLaunchKernelGGL(sort_small_segments_kernel, dim3(blocks_for(n_segments, kWaveSortWaves)),
                           dim3(kWaveSortWaves * kWave), 0, stream, keys, seg_off, n_segments, stats);
    }
    int rc = check_launch("sort_small_segments_kernel");
    if (rc) return rc;
    {
        StageTimer timer("sort_mid_segments_kernel", stream);
        hipLaunchKernelGGL(sort_mid_segments_kernel, dim3(blocks_for(n_segments, kLongSortSegments)),
                           dim3(kMidThreads), 0, stream, keys, seg_off, n_segments, stats);
    }
    if ((rc = check_launch("sort_mid_segments_kernel"))) return rc;
    StageTimer timer("sort_segments_kernel", stream);
    hipLaunchKernelGGL(sort_segments_kernel, dim3(blocks_for(n_segments, kLongSortSegments)), dim3(kSortThreads), 0,
                       stream, keys, keys_tmp, seg_off, n_segments, stats);
    return check_launch("sort_segments_kernel");
}

}  // namespace us3d

// ---------------------------------------------------------------- exported helpers

extern "C" {

const char* us3d_last_error(void) { return us3d::g_last_error.c_str(); }

int us3d_stage_timing(int enable) {
    us3d::timing().enabled = enable != 0;
    return US3D_OK;
}

int us3d_stage_timing_read(char* buf, int64_t bytes) {
    us3d::Timing& T = us3d::timing();
    std::lock_guard<std::mutex> lock(T.mu);
    if (!buf || bytes <= 0) return (int)(T.events.size() * 160 + 1);  // size query: consumes nothing
    std::string out;
    for (auto& kv : T.events) {
        double total = 0.0;
        for (auto& pr : kv.second) {
            float ms = 0.f;
            hipEventSynchronize(pr.second);
            if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) total += ms;
            hipEventDestroy(pr.first);
            hipEventDestroy(pr.second);
        }
        char line[256];
        snprintf(line, sizeof(line), "%s %zu %.6f\n", kv.first.c_str(), kv.second.size(), total);
        out += line;
    }
    T.events.clear();
    const size_t n = std::min<size_t>(out.size(), (size_t)bytes - 1);
    memcpy(buf, out.data(), n);
    buf[n] = 0;
    return US3D_OK;
}

int us3d_version(void) { return 3; }

int64_t us3d_scan_tmp_bytes(int64_t n) { return us3d::scan_tmp_bytes(n); }

int us3d_scan_i32(const int32_t* in, int32_t* out, int64_t n, int64_t* total_dev, void* tmp, void* stream) {
    return us3d::scan_i32(in, out, n, total_dev, tmp, us3d::as_stream(stream));
}

int us3d_fill_i32(int32_t* dst, int64_t n, int32_t value, void* stream) {
    if (n <= 0) return US3D_OK;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(us3d::fill_i32_kernel, dim3((unsigned)blocks), dim3(256), 0, us3d::as_stream(stream), dst, n,
                       value);
    return us3d::check_launch("fill_i32_kernel");
}

}  // extern "C"
