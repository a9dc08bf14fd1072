// Single-pass scans and stable compactions by decoupled look-back (Merrill
// and Garland, "Single-pass parallel prefix scan with decoupled look-back",
// 2016), written for this part:
//  * a block takes its tile from a ticket (an atomic counter), so it only
//    ever waits for tiles whose blocks already run, whatever order the
//    blocks are dispatched in;
//  * it publishes its tile's total at once (an aggregate), then one wave
//    looks back over the status words of the 64 tiles before it at a time:
//    the nearest inclusive prefix, plus the aggregates after it, is the
//    tile's exclusive prefix, which it then publishes as an inclusive one;
//  * a status word is (epoch << 34 | state << 32 | value), state 1 an
//    aggregate, 2 an inclusive prefix; a word of another epoch reads as not
//    yet published, so the words are never cleared;
//  * the last block to finish starts the next launch's epoch at ticket 0.
// Status words are read and written with device-scope atomics, which reach
// past the L2 of one XCD.  The state (GfLbCtl, then one status word per
// tile) starts at zero and belongs to one stream at a time.
#pragma once
#include "gf_common.h"

struct GfLbCtl {
    unsigned long long ticket;     // epoch << 32 | next tile
    unsigned long long finished;   // blocks done with the state
};
constexpr int GF_LB_CTL_INTS = 4;  // sizeof(GfLbCtl) / sizeof(int)
constexpr int GF_LB_THREADS = 1024;
constexpr int GF_LB_TILE = 4 * GF_LB_THREADS;   // 4 items a thread
constexpr int GF_LB_SPIN_MAX = 1 << 20;

// Ints of look-back state for n items.
inline size_t gf_lb_state_ints(int n) {
    return GF_LB_CTL_INTS + 2 * size_t((n + GF_LB_TILE - 1) / GF_LB_TILE);
}

__device__ __forceinline__ unsigned long long gf_lb_load(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void gf_lb_store(unsigned long long* p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The sum of the values of the tiles before `tile` (tag: this launch's
// epoch << 34).  Called by a whole wave, every lane returns it.  Lane l reads
// the word of tile j0 - l; lanes past tile 0 read as a prefix of 0.  The
// window is used when every tile from the nearest one with a prefix (lane
// p) up to this tile has published: the values of lanes 0..p are summed,
// and without a prefix in the window the walk moves 64 tiles back.  Every
// tile publishes its aggregate before it looks back, and tile 0 publishes
// its prefix at once, so the waits end; a wait of GF_LB_SPIN_MAX polls is
// counted in *timeouts and abandoned.
__device__ inline int gf_lb_exclusive(const unsigned long long* status, int tile,
                                      unsigned long long tag, unsigned int* timeouts) {
    const int lane = __lane_id();
    int sum = 0;
    int j0 = tile - 1;
    int spin = 0;
    while (j0 >= 0) {
        const int j = j0 - lane;
        int st = 2, val = 0;
        if (j >= 0) {
            const unsigned long long v = gf_lb_load(status + j);
            st = (v >> 34) == (tag >> 34) ? int((v >> 32) & 3u) : 0;
            val = int(uint32_t(v));
        }
        const uint64_t pre = __ballot(st == 2);
        const uint64_t ready = __ballot(st != 0);
        const int p = pre ? __ffsll((unsigned long long)pre) - 1 : 63;
        const uint64_t need = p == 63 ? ~0ull : (2ull << p) - 1;
        if ((ready & need) != need) {
            if (++spin > GF_LB_SPIN_MAX) {
                if (lane == 0) atomicAdd(timeouts, 1u);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
            continue;
        }
        sum += gf_wave_sum_i(lane <= p ? val : 0);
        if (pre) break;
        j0 -= 64;
    }
    return sum;
}

// One launch over items [0, n) (n = min(n_max, *len) when len is given;
// the grid covers n_max, gf_lb_blocks): every thread takes 4 consecutive
// items, Op::values gives their values (0 past n), and Op::put receives
// them with their exclusive prefixes over all items; Op::total receives n
// and the sum, from one thread.  Only the first ceil(n / GF_LB_TILE) blocks take
// part; the others leave at once.  Op must not read what Op::put writes for
// another tile.
template <class Op>
__global__ __launch_bounds__(GF_LB_THREADS) void gf_k_lookback(Op op, int n_max,
                                                               const int* __restrict__ len,
                                                               GfLbCtl* ctl,
                                                               unsigned long long* status,
                                                               unsigned int* timeouts) {
    __shared__ unsigned long long s_ticket;
    __shared__ int s_wpre[GF_LB_THREADS / 64];
    __shared__ int s_base;
    const int t = threadIdx.x, lane = __lane_id(), w = t >> 6;
    const int n = len ? min(n_max, *len) : n_max;
    const int ntiles = (n + GF_LB_TILE - 1) / GF_LB_TILE;
    if (ntiles == 0) {
        if (blockIdx.x == 0 && t == 0) op.total(0, 0);
        return;
    }
    if (int(blockIdx.x) >= ntiles) return;
    if (t == 0) s_ticket = atomicAdd(&ctl->ticket, 1ull);
    __syncthreads();
    const uint32_t epoch = uint32_t(s_ticket >> 32);
    const int tile = int(uint32_t(s_ticket));
    const int i = tile * GF_LB_TILE + 4 * t;
    int v[4];
    op.values(i, n, v);
    const int sum = v[0] + v[1] + v[2] + v[3];
    const int inc = gf_wave_incl_scan(sum);
    if (lane == 63) s_wpre[w] = inc;
    __syncthreads();
    if (w == 0) {
        const int wt = lane < GF_LB_THREADS / 64 ? s_wpre[lane] : 0;
        const int winc = gf_wave_incl_scan(wt);
        const int total = __builtin_amdgcn_readlane(winc, GF_LB_THREADS / 64 - 1);
        if (lane < GF_LB_THREADS / 64) s_wpre[lane] = winc - wt;
        const unsigned long long tag = (unsigned long long)(epoch & 0x3fffffffu) << 34;
        if (lane == 0)
            gf_lb_store(status + tile, tag | ((tile == 0 ? 2ull : 1ull) << 32) | uint32_t(total));
        int excl = 0;
        if (tile > 0) {
            excl = gf_lb_exclusive(status, tile, tag, timeouts);
            if (lane == 0) gf_lb_store(status + tile, tag | (2ull << 32) | uint32_t(excl + total));
        }
        if (lane == 0) {
            s_base = excl;
            if (tile == ntiles - 1) op.total(n, excl + total);
        }
    }
    __syncthreads();
    const int b = s_base + s_wpre[w] + inc - sum;
    const int off[4] = {b, b + v[0], b + v[0] + v[1], b + v[0] + v[1] + v[2]};
    op.put(i, n, v, off);
    if (t == 0 && atomicAdd(&ctl->finished, 1ull) == (unsigned long long)(ntiles - 1)) {
        // the last block: no tile reads this launch's state any more
        __hip_atomic_store(&ctl->finished, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        atomicExch(&ctl->ticket, (unsigned long long)(epoch + 1u) << 32);
    }
}

// The grid of a launch over at most n_max items (one block when n_max is 0).
inline int gf_lb_blocks(int n_max) { return n_max > 0 ? (n_max + GF_LB_TILE - 1) / GF_LB_TILE : 1; }
