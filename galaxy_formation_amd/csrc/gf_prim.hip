// Device-wide primitives: exclusive scan of int32 and a stable LSD radix sort
// of (uint64 key, uint32 value) pairs.  Both are written for wave64: ranks
// inside a wave come from ballots, block-level offsets from LDS.
#include <atomic>

#include "gf_prim.h"
#include "gf_lookback.h"

// look-back waits that gave up, in every single-pass launch of the library
// (a diagnostic: stays 0 unless a tile's predecessor never published, which
// the ticket order rules out)
__device__ unsigned int g_lookback_timeouts = 0;

namespace {

// len (device, may be null): scan only the first min(n, *len) inputs; the
// blocks past them do no loads.
__device__ __forceinline__ int scan_len(int n, const int* len) { return len ? min(n, *len) : n; }

// One block of 1024 threads scans up to SCAN_ONE_MAX inputs (exclusive) and
// writes the total at out[n], in one launch.  Thread t holds items 4t..4t+3
// of each chunk of 4096 (16-byte loads, coalesced across the block).  The
// first round loads every chunk before the first wait and keeps only the
// waves' totals; the second reads the chunks again (unchanged until the
// stores) and writes the prefixes.  in and out must be 16-byte aligned and
// may be the same array: every store follows the last barrier, and the
// second round's loads of a thread's items precede its own stores to them.
constexpr int SCAN_ONE_CHUNKS = 16;
constexpr int SCAN_ONE_MAX = SCAN_ONE_CHUNKS * 4096;

__device__ __forceinline__ int4 scan_one_load(const int* in, int n, int i) {
    if (i + 3 < n) return *reinterpret_cast<const int4*>(in + i);
    return make_int4(i < n ? in[i] : 0, i + 1 < n ? in[i + 1] : 0, i + 2 < n ? in[i + 2] : 0, 0);
}

__global__ __launch_bounds__(1024) void k_scan_one(const int* in, int n,
                                                   const int* __restrict__ len, int* out) {
    __shared__ int wtot[SCAN_ONE_CHUNKS][16];   // per chunk, per wave: the wave's total
    __shared__ int woff[SCAN_ONE_CHUNKS][16];   // ... the prefix of the waves before it
    __shared__ int cpre[SCAN_ONE_CHUNKS + 1];   // prefix of the chunk totals
    n = scan_len(n, len);
    const int t = threadIdx.x, lane = __lane_id(), w = t >> 6;
    {
        int4 v[SCAN_ONE_CHUNKS];
#pragma unroll
        for (int c = 0; c < SCAN_ONE_CHUNKS; ++c) v[c] = scan_one_load(in, n, c * 4096 + 4 * t);
#pragma unroll
        for (int c = 0; c < SCAN_ONE_CHUNKS; ++c) {
            const int inc = gf_wave_incl_scan(v[c].x + v[c].y + v[c].z + v[c].w);
            if (lane == 63) wtot[c][w] = inc;
        }
    }
    __syncthreads();
    if (t < SCAN_ONE_CHUNKS) {
        int run = 0;
        for (int q = 0; q < 16; ++q) {
            woff[t][q] = run;
            run += wtot[t][q];
        }
        cpre[t + 1] = run;                        // chunk t's total, summed below
    }
    __syncthreads();
    if (t == 0) {
        cpre[0] = 0;
        for (int c = 1; c <= SCAN_ONE_CHUNKS; ++c) cpre[c] += cpre[c - 1];
    }
    __syncthreads();
    int4 v[SCAN_ONE_CHUNKS];
#pragma unroll
    for (int c = 0; c < SCAN_ONE_CHUNKS; ++c) v[c] = scan_one_load(in, n, c * 4096 + 4 * t);
#pragma unroll
    for (int c = 0; c < SCAN_ONE_CHUNKS; ++c) {
        const int i = c * 4096 + 4 * t;
        const int sum = v[c].x + v[c].y + v[c].z + v[c].w;
        const int base = cpre[c] + woff[c][w] + gf_wave_incl_scan(sum) - sum;
        const int4 o = make_int4(base, base + v[c].x, base + v[c].x + v[c].y,
                                 base + v[c].x + v[c].y + v[c].z);
        if (i + 3 < n) {
            *reinterpret_cast<int4*>(out + i) = o;
        } else {
            if (i < n) out[i] = o.x;
            if (i + 1 < n) out[i + 1] = o.y;
            if (i + 2 < n) out[i + 2] = o.z;
        }
    }
    if (t == 0) out[n] = cpre[SCAN_ONE_CHUNKS];
}

// ------------------------------------------------- single-pass scan
// A scan past SCAN_ONE_MAX inputs is one launch of gf_k_lookback: each
// thread loads its 4 inputs as one 16-byte load and stores its 4 prefixes
// as one.  in and out may be the same array: a thread reads its items
// before it writes them, and no other block reads them.
struct ScanOp {
    const int* in;
    int* out;
    __device__ void values(int i, int n, int (&v)[4]) const {
        const int4 x = scan_one_load(in, n, i);
        v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
    }
    __device__ void put(int i, int n, const int (&)[4], const int (&o)[4]) const {
        if (i + 3 < n) {
            *reinterpret_cast<int4*>(out + i) = make_int4(o[0], o[1], o[2], o[3]);
        } else {
            for (int q = 0; q < 3; ++q)
                if (i + q < n) out[i + q] = o[q];
        }
    }
    __device__ void total(int n, int s) const { out[n] = s; }
};

// ------------------------------------------------------------------ radix sort
// A block sorts a tile of 4096 keys per pass with 16 waves of 4 keys per lane
// (wide blocks keep a 1M-key pass at about 16 waves per CU: a tile per block
// of 4 waves left most CUs with one block).
#ifndef RS_THREADS_CFG
#define RS_THREADS_CFG 1024
#endif
constexpr int RS_THREADS = RS_THREADS_CFG;
constexpr int RS_TILE = 4096;                      // keys per block
constexpr int RS_ITEMS = RS_TILE / RS_THREADS;     // per lane
constexpr int RS_WAVES = RS_THREADS / 64;
constexpr int RS_WAVE_SPAN = 64 * RS_ITEMS;        // contiguous keys per wave
static_assert(RS_ITEMS * RS_THREADS == RS_TILE && RS_THREADS >= 256, "radix tile shape");

__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const uint64_t* __restrict__ key, int n,
                                                        int shift, int nblocks,
                                                        int* __restrict__ hist) {
    __shared__ int h[256];
    if (threadIdx.x < 256) h[threadIdx.x] = 0;
    __syncthreads();
    const long long base = (long long)blockIdx.x * RS_TILE;
#pragma unroll
    for (int k = 0; k < RS_ITEMS; ++k) {
        long long i = base + (long long)k * RS_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&h[(key[i] >> shift) & 0xFF], 1);
    }
    __syncthreads();
    if (threadIdx.x < 256) hist[threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];  // digit-major
}

// One wave per digit: the exclusive prefix of the digit's row of block
// counts (in place) and the row total (tot[d]).
__global__ __launch_bounds__(64) void k_rs_rowscan(int* __restrict__ hist, int nblocks,
                                                   int* __restrict__ tot) {
    int* row = hist + size_t(blockIdx.x) * nblocks;
    const int lane = threadIdx.x;
    const int per = (nblocks + 63) / 64;
    const int b0 = min(nblocks, lane * per), b1 = min(nblocks, b0 + per);
    int sum = 0;
    for (int b = b0; b < b1; ++b) sum += row[b];
    const int incl = gf_wave_incl_scan(sum);
    int run = incl - sum;
    for (int b = b0; b < b1; ++b) {
        const int v = row[b];
        row[b] = run;
        run += v;
    }
    if (lane == 63) tot[blockIdx.x] = incl;
}

// Stable scatter of one tile of RS_TILE keys.  Ranking: wave w owns keys
// [tile + w*SPAN, +SPAN), read as chunks of 64; lanes holding the same digit
// are ranked by lane order (8 ballots), chunks in order through a per-wave
// digit counter, waves in order through an LDS prefix.  The tile is first
// written to LDS in digit order, then copied out in that order: consecutive
// threads write consecutive slots of one digit's output run, so the global
// stores coalesce.  A digit's run starts at the keys of smaller digits
// (scan of tot) plus its keys in earlier tiles (the row prefix of hist).
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(
    const uint64_t* __restrict__ key, const uint32_t* __restrict__ val, int n, int shift,
    int nblocks, const int* __restrict__ hist, const int* __restrict__ tot,
    uint64_t* __restrict__ key_out, uint32_t* __restrict__ val_out) {
    __shared__ int cnt[RS_WAVES][256];
    __shared__ int lbase[256];                   // tile-local start of each digit
    __shared__ int gbase[256];                   // global start of the tile's run of each digit
    __shared__ int s_wtot[2][4];
    __shared__ uint64_t skey[RS_TILE];
    __shared__ uint32_t sval[RS_TILE];
    const int lane = __lane_id(), w = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < RS_WAVES * 256; i += RS_THREADS) (&cnt[0][0])[i] = 0;
    __syncthreads();

    const long long tile0 = (long long)blockIdx.x * RS_TILE;
    const long long base = tile0 + (long long)w * RS_WAVE_SPAN;
    uint64_t k_[RS_ITEMS];
    uint32_t v_[RS_ITEMS];
    int r_[RS_ITEMS];
    const uint64_t lt = gf_lanemask_lt();
#pragma unroll
    for (int c = 0; c < RS_ITEMS; ++c) {
        const long long i = base + (long long)c * 64 + lane;
        const bool ok = i < n;
        k_[c] = ok ? key[i] : 0ull;
        v_[c] = ok ? val[i] : 0u;
        const int d = int((k_[c] >> shift) & 0xFF);
        uint64_t peers = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        int pre = 0;
        if (ok) pre = cnt[w][d];
        const int rank = __popcll(peers & lt);
        const bool leader = ok && (peers & lt) == 0;
        // every lane has read the counter before any leader writes it: the
        // wave issues its LDS operations in program order
        __builtin_amdgcn_wave_barrier();
        if (leader) cnt[w][d] = pre + __popcll(peers);
        __builtin_amdgcn_wave_barrier();
        r_[c] = ok ? pre + rank : -1;
    }
    __syncthreads();
    // per digit (first 4 waves, one digit per thread): the waves' prefixes,
    // the tile's count, and both scans over the 256 digits
    int dcnt = 0, gt = 0, lincl = 0, gincl = 0;
    if (threadIdx.x < 256) {
        const int d = threadIdx.x;
#pragma unroll
        for (int ww = 0; ww < RS_WAVES; ++ww) {
            const int c = cnt[ww][d];
            cnt[ww][d] = dcnt;                   // from here on: the wave's prefix in the tile
            dcnt += c;
        }
        gt = tot[d];
        lincl = gf_wave_incl_scan(dcnt);
        gincl = gf_wave_incl_scan(gt);
        if (lane == 63) {
            s_wtot[0][w] = lincl;
            s_wtot[1][w] = gincl;
        }
    }
    __syncthreads();
    if (threadIdx.x < 256) {
        const int d = threadIdx.x;
        int lo = lincl - dcnt, go = gincl - gt;
        for (int q = 0; q < w; ++q) {
            lo += s_wtot[0][q];
            go += s_wtot[1][q];
        }
        lbase[d] = lo;
        gbase[d] = go + hist[size_t(d) * nblocks + blockIdx.x];
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < RS_ITEMS; ++c) {
        if (r_[c] >= 0) {
            const int d = int((k_[c] >> shift) & 0xFF);
            const int lp = lbase[d] + cnt[w][d] + r_[c];
            skey[lp] = k_[c];
            sval[lp] = v_[c];
        }
    }
    __syncthreads();
    const int m = int(min((long long)RS_TILE, (long long)n - tile0));
    for (int i = threadIdx.x; i < m; i += RS_THREADS) {
        const uint64_t kk = skey[i];
        const int d = int((kk >> shift) & 0xFF);
        const int dst = gbase[d] + (i - lbase[d]);
        key_out[dst] = kk;
        val_out[dst] = sval[i];
    }
}

}  // namespace

// Scan scratch: the look-back state (control words, then one status word
// per tile).
size_t gf_scan_temp_ints(int n) { return gf_lb_state_ints(n); }

int gf_scan_temp_clear(int* temp, int n) {
    GF_HIP(hipMemsetAsync(temp, 0, sizeof(int) * gf_scan_temp_ints(n), nullptr));
    GF_HIP(hipStreamSynchronize(nullptr));
    return 0;
}

// The counter's device address, looked up once per device (each device holds
// its own copy of the variable) and kept.
unsigned int* gf_prim_timeout_counter() {
    constexpr int kMaxDevices = 64;
    static std::atomic<unsigned int*> ptr[kMaxDevices];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    unsigned int* p = ptr[dev].load(std::memory_order_acquire);
    if (!p) {
        void* q = nullptr;
        if (hipGetSymbolAddress(&q, HIP_SYMBOL(g_lookback_timeouts)) != hipSuccess) return nullptr;
        p = static_cast<unsigned int*>(q);
        ptr[dev].store(p, std::memory_order_release);
    }
    return p;
}

unsigned int gf_prim_lookback_timeouts() {
    unsigned int v = 0;
    if (hipMemcpyFromSymbol(&v, HIP_SYMBOL(g_lookback_timeouts), sizeof(v), 0,
                            hipMemcpyDeviceToHost) != hipSuccess)
        return ~0u;
    return v;
}

int gf_scan_excl(const int* in, int* out, int n, int* temp, hipStream_t s, const int* len) {
    if (n <= 0) {
        GF_HIP(hipMemsetAsync(out, 0, sizeof(int), s));
        return 0;
    }
    if (((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) != 0) {
        gf_set_error("gf_scan_excl: in and out must be 16-byte aligned");
        return -1;
    }
    if (n <= SCAN_ONE_MAX) {
        hipLaunchKernelGGL(k_scan_one, dim3(1), dim3(1024), 0, s, in, n, len, out);
    } else {
        unsigned int* timeouts = gf_prim_timeout_counter();
        if (!timeouts) {
            gf_set_error("gf_scan_excl: no device address for the look-back counter");
            return -1;
        }
        hipLaunchKernelGGL(gf_k_lookback<ScanOp>, dim3(gf_lb_blocks(n)), dim3(GF_LB_THREADS), 0,
                           s, ScanOp{in, out}, n, len, reinterpret_cast<GfLbCtl*>(temp),
                           reinterpret_cast<unsigned long long*>(temp + GF_LB_CTL_INTS), timeouts);
    }
    GF_LAUNCH_CHECK();
    return 0;
}

size_t gf_sort_temp_ints(int n) { return size_t(256) * gf_div_up(n, RS_TILE) + 256; }

int gf_radix_sort_pairs(uint64_t* key, uint32_t* val, uint64_t* key_alt, uint32_t* val_alt, int n,
                        int begin_bit, int end_bit, int* temp, hipStream_t s, int* in_alt) {
    *in_alt = 0;
    if (n <= 1) return 0;
    const int nb = gf_div_up(n, RS_TILE);
    int* hist = temp;               // digit-major per-tile digit counts
    int* tot = temp + size_t(256) * nb;
    uint64_t *ka = key, *kb = key_alt;
    uint32_t *va = val, *vb = val_alt;
    for (int shift = begin_bit; shift < end_bit; shift += 8) {
        hipLaunchKernelGGL(k_rs_hist, dim3(nb), dim3(RS_THREADS), 0, s, ka, n, shift, nb, hist);
        hipLaunchKernelGGL(k_rs_rowscan, dim3(256), dim3(64), 0, s, hist, nb, tot);
        hipLaunchKernelGGL(k_rs_scatter, dim3(nb), dim3(RS_THREADS), 0, s, ka, va, n, shift, nb,
                           hist, tot, kb, vb);
        GF_LAUNCH_CHECK();
        uint64_t* tk = ka; ka = kb; kb = tk;
        uint32_t* tv = va; va = vb; vb = tv;
        *in_alt ^= 1;
    }
    return 0;
}
