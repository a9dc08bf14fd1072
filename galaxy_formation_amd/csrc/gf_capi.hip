// C ABI of libgf.so: the five symbols of the reference engine
// (cpp/nbody.cpp:415-505, bound by galaxy_formation/engine.py:26-35) with the
// same arguments, plus gf_* entry points for the device-resident path, the
// SPH kernels, the learned smoothing length and parity checks.
//
// State kept between calls, as the reference keeps its global tree
// (cpp/nbody.cpp:307): the last gravity tree (for rebuild = 0 refits) and
// staging buffers.  Entry points are not re-entrant.  Host-pointer entry
// points run on the library's stream and return when their outputs are in
// host memory; *_dev entry points take device pointers and a stream (null =
// the library's stream) and return without synchronising.
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "gf_capi.h"
#include "gf_gravity.h"
#include "gf_prim.h"
#include "gf_render.h"
#include "gf_step.h"
#include "gf_xfer.h"

static char g_err[1024] = "";

void gf_set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

GfCtx g_ctx;

int gf_ctx_init() {
    if (g_ctx.inited) return 0;
    int count = 0;
    const hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
        gf_set_error("no HIP device available (%s)", hipGetErrorString(e));
        return -1;
    }
    GF_HIP(hipGetDevice(&g_ctx.device));
    GF_HIP(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
    hipDeviceProp_t prop;
    GF_HIP(hipGetDeviceProperties(&prop, g_ctx.device));
    g_ctx.n_cu = prop.multiProcessorCount;
    g_ctx.inited = 1;
    return 0;
}

// Stream of a *_dev call: the caller's; NULL is HIP's default stream (the
// one PyTorch's default stream is), so device-array callers stay ordered
// with their own work.
hipStream_t gf_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

int gf_stage_reserve(void** p, size_t* cap, size_t bytes) {
    if (bytes <= *cap && *p) return 0;
    if (*p) GF_HIP(hipFree(*p));
    *p = nullptr;
    *cap = 0;
    const size_t b = bytes ? bytes : 16;
    GF_HIP(hipMalloc(p, b));
    *cap = b;
    return 0;
}

#define STAGE(ptr, bytes) gf_stage_reserve(reinterpret_cast<void**>(&(ptr)), &ptr##_cap, (bytes))

static int check_walk_errors() {
    int oc[3] = {0, 0, 0};
    GF_HIP(hipMemcpy(oc, g_ctx.walk.ovf_count, sizeof(oc), hipMemcpyDeviceToHost));
    if (oc[1] != 0) {
        gf_set_error("tree walk: %d groups overflowed the global stack", oc[1]);
        return -1;
    }
    if (oc[2] != 0) {
        gf_set_error("tree walk: %d nodes whose children do not hold consecutive records", oc[2]);
        return -1;
    }
    return 0;
}

static int gravity_impl(int n, const double* pos, const double* mass, double* acc, double theta,
                        double eps, const uint8_t* active, int rebuild, int precision,
                        hipStream_t s) {
    GfTree& T = g_ctx.tree;
    const float th = float(theta);
    if (rebuild || !T.built || T.n != n) {
        if (gf_tree_build(T, n, pos, mass, th, s)) return -1;
    } else {
        if (gf_tree_refit(T, pos, mass, s)) return -1;
    }
    if (n == 0) return 0;
    if (gf_walk_reserve(g_ctx.walk, T.cap_nodes)) return -1;
    GF_HIP(hipMemsetAsync(g_ctx.walk.inter, 0, 2 * sizeof(unsigned long long), s));
    return gf_walk(T, g_ctx.walk, active, acc, eps, precision, s);
}

GfTestLimits& gf_test_limits() {
    static GfTestLimits limits;
    return limits;
}

extern "C" {

const char* gf_last_error(void) { return g_err; }

int gf_version(void) { return GF_ABI_VERSION; }

int gf_device_info(char* buf, int len) {
    if (gf_ctx_init()) return -1;
    hipDeviceProp_t prop;
    GF_HIP(hipGetDeviceProperties(&prop, g_ctx.device));
    snprintf(buf, size_t(len), "%s %s CUs=%d LDS/block=%zu clock_kHz=%d", prop.name,
             prop.gcnArchName, prop.multiProcessorCount, prop.sharedMemPerBlock, prop.clockRate);
    return 0;
}

int nbody_num_threads(void) {
    if (gf_ctx_init()) return -1;
    return g_ctx.n_cu;
}

int gf_set_quad_tolerance(double tol) {
    if (!(tol >= 0.0) || tol > 1.0) {
        gf_set_error("quadrupole tolerance must be in [0, 1]");
        return -1;
    }
    g_ctx.walk.quad_tol = float(tol);
    return 0;
}

double gf_get_quad_tolerance(void) { return double(g_ctx.walk.quad_tol); }

int gf_set_quad_relative_tolerance(double tol) {
    if (!(tol >= 0.0) || tol > 1.0) {
        gf_set_error("relative quadrupole tolerance must be in [0, 1]");
        return -1;
    }
    g_ctx.walk.quad_rel_tol = float(tol);
    return 0;
}

double gf_get_quad_relative_tolerance(void) { return double(g_ctx.walk.quad_rel_tol); }

int gf_set_sph_precision(int precision) {
    if (precision != 0 && precision != 1) {
        gf_set_error("SPH precision must be 0 (FP64) or 1 (FP32 pair arithmetic)");
        return -1;
    }
    g_ctx.sph.pair_f32 = precision;
    return 0;
}

int gf_get_sph_precision(void) { return g_ctx.sph.pair_f32; }

int gf_set_knn_blocks(int blocks) {
    if (blocks < 0 || blocks > 65536) {
        gf_set_error("neighbour-pass blocks must be in [0, 65536] (0 = default)");
        return -1;
    }
    g_ctx.sph.fast_blocks = blocks;
    return 0;
}

int gf_get_knn_blocks(void) { return gf_sph_fast_blocks(g_ctx.sph); }

int gf_set_knn_merge(int members) {
    if (members < 0 || members > 64) {
        gf_set_error("merged work items: members must be in [0, 64] (0 = default, 1 = none)");
        return -1;
    }
    g_ctx.sph.merge_members = members;
    return 0;
}

int gf_get_knn_merge(void) { return gf_sph_merge_members(g_ctx.sph); }

int gf_set_walk_blocks(int blocks) {
    if (blocks < 0 || blocks > 65536) {
        gf_set_error("walk blocks must be in [0, 65536] (0 = default)");
        return -1;
    }
    g_ctx.walk.walk_blocks = blocks;
    return 0;
}

int gf_get_walk_blocks(void) {
    return g_ctx.walk.walk_blocks > 0 ? g_ctx.walk.walk_blocks : gf_walk_default_blocks();
}

int gf_set_test_limits(int walk_stack, int knn_list, int moment_top) {
    if (walk_stack < 0 || walk_stack > 1024 || knn_list < 0 || knn_list > 1024 ||
        moment_top < 0 || moment_top > 2048) {
        gf_set_error("test limits: walk_stack and knn_list must be in [0, 1024], moment_top in "
                     "[0, 2048] (0 = default)");
        return -1;
    }
    gf_test_limits().walk_stack = walk_stack;
    gf_test_limits().knn_list = knn_list;
    gf_test_limits().moment_top = moment_top;
    return 0;
}

int gf_set_precision(int p) {
    if (p != GF_FP32 && p != GF_FP64 && p != GF_FP32_QUAD) {
        gf_set_error("precision must be 0 (FP32 interactions), 1 (FP64) or 2 (FP32 with "
                     "node quadrupoles)");
        return -1;
    }
    g_ctx.precision = p;
    return 0;
}

int gf_get_precision(void) { return g_ctx.precision; }

int nbody_accel_active(int n, const double* pos, const double* mass, double* acc, double theta,
                       double eps, const uint8_t* active, int rebuild) {
    if (gf_ctx_init()) return -1;
    if (n < 0) {
        gf_set_error("n must be >= 0");
        return -1;
    }
    hipStream_t s = g_ctx.stream;
    const size_t b3 = sizeof(double) * 3 * size_t(n), b1 = sizeof(double) * size_t(n);
    if (STAGE(g_ctx.d_pos, b3) || STAGE(g_ctx.d_mass, b1) || STAGE(g_ctx.d_acc, b3) ||
        STAGE(g_ctx.d_act, size_t(n)))
        return -1;
    if (n > 0) {
        if (gf_copy_h2d(g_ctx.d_pos, pos, b3, s)) return -1;
        if (gf_copy_h2d(g_ctx.d_mass, mass, b1, s)) return -1;
        if (gf_copy_h2d(g_ctx.d_acc, acc, b3, s)) return -1;
        if (active && gf_copy_h2d(g_ctx.d_act, active, size_t(n), s)) return -1;
    }
    if (gravity_impl(n, g_ctx.d_pos, g_ctx.d_mass, g_ctx.d_acc, theta, eps,
                     active ? g_ctx.d_act : nullptr, rebuild, g_ctx.precision, s))
        return -1;
    if (n > 0 && gf_copy_d2h(acc, g_ctx.d_acc, b3, s)) return -1;
    GF_HIP(hipStreamSynchronize(s));
    if (n > 0) {
        if (check_walk_errors()) return -1;
        GF_HIP(hipMemcpy(&g_ctx.last_inter, g_ctx.walk.inter, sizeof(unsigned long long),
                         hipMemcpyDeviceToHost));
    }
    return 0;
}

int nbody_accel_direct(int n, const double* pos, const double* mass, double* acc, double eps) {
    return gf_accel_direct_subset(n, pos, mass, n, nullptr, acc, eps);
}

int gf_accel_direct_subset(int n, const double* pos, const double* mass, int nt,
                           const int64_t* tgt, double* acc, double eps) {
    if (gf_ctx_init()) return -1;
    if (n <= 0 || nt <= 0) return 0;
    hipStream_t s = g_ctx.stream;
    const size_t b3 = sizeof(double) * 3 * size_t(n), b1 = sizeof(double) * size_t(n);
    const size_t bt = sizeof(double) * 3 * size_t(nt);
    if (STAGE(g_ctx.d_pos, b3) || STAGE(g_ctx.d_mass, b1) || STAGE(g_ctx.d_acc, bt) ||
        STAGE(g_ctx.d_tgt, sizeof(int64_t) * size_t(nt)) ||
        STAGE(g_ctx.d_partial, sizeof(double) * gf_direct_partial_doubles(n, nt)))
        return -1;
    if (gf_copy_h2d(g_ctx.d_pos, pos, b3, s)) return -1;
    if (gf_copy_h2d(g_ctx.d_mass, mass, b1, s)) return -1;
    if (tgt && gf_copy_h2d(g_ctx.d_tgt, tgt, sizeof(int64_t) * nt, s)) return -1;
    if (gf_direct(n, g_ctx.d_pos, g_ctx.d_mass, nt, tgt ? g_ctx.d_tgt : nullptr, g_ctx.d_acc, eps,
                  g_ctx.d_partial, s))
        return -1;
    if (gf_copy_d2h(acc, g_ctx.d_acc, bt, s)) return -1;
    GF_HIP(hipStreamSynchronize(s));
    return 0;
}

long long gf_last_interactions(void) { return (long long)g_ctx.last_inter; }

int gf_tree_info(int* nodes, int* groups, int* deep_cells) {
    if (gf_ctx_init()) return -1;
    int c[4] = {0, 0, 0, 0};
    if (g_ctx.tree.counters) GF_HIP(hipMemcpy(c, g_ctx.tree.counters, sizeof(c), hipMemcpyDeviceToHost));
    if (nodes) *nodes = c[0];
    if (groups) *groups = c[1];
    if (deep_cells) *deep_cells = c[2];
    return 0;
}

int gf_lookback_timeouts(void) {
    if (gf_ctx_init()) return -1;
    const unsigned int v = gf_prim_lookback_timeouts();
    if (v == ~0u) {
        gf_set_error("gf_lookback_timeouts: the device counter could not be read");
        return -1;
    }
    return int(v);
}

int gf_tree_export(int cap, int* start, int* count, int* next, int* leaf, int* depth,
                   double* com4, double* open_r, double* box4, int* idx, int* groups) {
    if (gf_ctx_init()) return -1;
    const GfTree& T = g_ctx.tree;
    int nn = 0, ng = 0;
    if (gf_tree_info(&nn, &ng, nullptr)) return -1;
    if (nn > cap) {
        gf_set_error("tree export: %d nodes do not fit in %d", nn, cap);
        return -1;
    }
    const size_t ni = sizeof(int) * size_t(nn);
    if (start) GF_HIP(hipMemcpy(start, T.start, ni, hipMemcpyDeviceToHost));
    if (count) GF_HIP(hipMemcpy(count, T.count, ni, hipMemcpyDeviceToHost));
    if (next) GF_HIP(hipMemcpy(next, T.next, ni, hipMemcpyDeviceToHost));
    if (leaf || depth) {
        int8_t* tmp = new int8_t[size_t(nn) + 1];
        if (leaf) {
            GF_HIP(hipMemcpy(tmp, T.leaf, size_t(nn), hipMemcpyDeviceToHost));
            for (int k = 0; k < nn; ++k) leaf[k] = tmp[k];
        }
        if (depth) {
            GF_HIP(hipMemcpy(tmp, T.depth, size_t(nn), hipMemcpyDeviceToHost));
            for (int k = 0; k < nn; ++k) depth[k] = tmp[k];
        }
        delete[] tmp;
    }
    if (com4) GF_HIP(hipMemcpy(com4, T.com, sizeof(double4) * size_t(nn), hipMemcpyDeviceToHost));
    if (open_r) GF_HIP(hipMemcpy(open_r, T.open_r, sizeof(double) * size_t(nn), hipMemcpyDeviceToHost));
    if (box4) GF_HIP(hipMemcpy(box4, T.box, sizeof(double4) * size_t(nn), hipMemcpyDeviceToHost));
    if (idx) GF_HIP(hipMemcpy(idx, T.idx, sizeof(int) * size_t(T.n), hipMemcpyDeviceToHost));
    if (groups) GF_HIP(hipMemcpy(groups, T.groups, sizeof(int) * size_t(ng), hipMemcpyDeviceToHost));
    return nn;
}

int gf_gravity_dev(int n, const double* pos, const double* mass, double* acc, double theta,
                   double eps, const uint8_t* active, int rebuild, int precision, void* stream) {
    if (gf_ctx_init()) return -1;
    if (n < 0 || (precision != GF_FP32 && precision != GF_FP64 && precision != GF_FP32_QUAD)) {
        gf_set_error("gf_gravity_dev: bad n or precision");
        return -1;
    }
    return gravity_impl(n, pos, mass, acc, theta, eps, active, rebuild, precision,
                        gf_stream(stream));
}

int gf_tree_build_dev(int n, const double* pos, const double* mass, double theta, void* stream) {
    if (gf_ctx_init()) return -1;
    if (n < 0) {
        gf_set_error("gf_tree_build_dev: n must be >= 0");
        return -1;
    }
    if (!g_ctx.tree_sorted) GF_HIP(hipEventCreateWithFlags(&g_ctx.tree_sorted, hipEventDisableTiming));
    g_ctx.tree.keep_sorted = 1;
    g_ctx.tree_pos = nullptr;
    if (gf_tree_build(g_ctx.tree, n, pos, mass, float(theta), gf_stream(stream), true,
                      g_ctx.tree_sorted))
        return -1;
    g_ctx.tree_pos = pos;
    return 0;
}

int gf_walk_dev(int n, const uint8_t* active, double* acc, double eps, int precision,
                void* stream) {
    if (gf_ctx_init()) return -1;
    if (n != g_ctx.tree.n || (n > 0 && !g_ctx.tree.built)) {
        gf_set_error("gf_walk_dev: no tree built for n = %d", n);
        return -1;
    }
    if (precision != GF_FP32 && precision != GF_FP64 && precision != GF_FP32_QUAD) {
        gf_set_error("gf_walk_dev: bad precision");
        return -1;
    }
    if (n == 0) return 0;
    hipStream_t s = gf_stream(stream);
    if (gf_walk_reserve(g_ctx.walk, g_ctx.tree.cap_nodes)) return -1;
    GF_HIP(hipMemsetAsync(g_ctx.walk.inter, 0, 2 * sizeof(unsigned long long), s));
    return gf_walk(g_ctx.tree, g_ctx.walk, active, acc, eps, precision, s);
}

int gf_walk_prepare_dev(int precision, void* stream) {
    if (gf_ctx_init()) return -1;
    if (g_ctx.tree.n > 0 && !g_ctx.tree.built) {
        gf_set_error("gf_walk_prepare_dev: no tree built");
        return -1;
    }
    if (precision != GF_FP32 && precision != GF_FP64 && precision != GF_FP32_QUAD) {
        gf_set_error("gf_walk_prepare_dev: bad precision");
        return -1;
    }
    return gf_walk_prepare(g_ctx.tree, g_ctx.walk, precision, gf_stream(stream));
}

int gf_direct_dev(int n, const double* pos, const double* mass, int nt, const int64_t* tgt,
                  double* acc, double eps, void* stream) {
    if (gf_ctx_init()) return -1;
    if (STAGE(g_ctx.d_partial, sizeof(double) * gf_direct_partial_doubles(n, nt))) return -1;
    return gf_direct(n, pos, mass, nt, tgt, acc, eps, g_ctx.d_partial, gf_stream(stream));
}

int gf_potential_dev(int n, const double* pos, const double* mass, double* phi, double eps,
                     void* stream) {
    if (gf_ctx_init()) return -1;
    if (n < 0 || (n > 0 && (!pos || !mass || !phi))) {
        gf_set_error("gf_potential_dev: bad arguments");
        return -1;
    }
    if (STAGE(g_ctx.d_partial, sizeof(double) * gf_potential_partial_doubles(n))) return -1;
    return gf_potential(n, pos, mass, phi, eps, g_ctx.d_partial, gf_stream(stream));
}

int gf_combine_dev(int n, int ng, double eps, const double* acc_grav, const double* acc_sph,
                   const double* dt_c, double* acc, double* dt_ind, double* dt_grav,
                   void* stream) {
    if (gf_ctx_init()) return -1;
    if (n < 0 || ng < 0 || ng > n || !dt_grav ||
        (n > 0 && (!acc_grav || !acc || !dt_ind)) || (ng > 0 && (!acc_sph || !dt_c))) {
        gf_set_error("gf_combine_dev: bad arguments");
        return -1;
    }
    return gf_combine(n, ng, eps, acc_grav, acc_sph, dt_c, acc, dt_ind, dt_grav,
                      gf_stream(stream));
}

int gf_step_drift_dev(int n, double dt, long long k, int rungs, double* pos, const double* vel,
                      const int64_t* rung, uint8_t* active, int* count, void* stream) {
    if (gf_ctx_init()) return -1;
    if (n < 0 || k < 0 || rungs < 1 || rungs > 62 ||
        (n > 0 && (!pos || !vel || !rung || !active || !count))) {
        gf_set_error("gf_step_drift_dev: bad arguments");
        return -1;
    }
    return gf_step_drift(n, dt, k, rungs, pos, vel, rung, active, count, gf_stream(stream));
}

int gf_step_kick_dev(int n, int ng, long long k, int rungs, double dt_max, int mode,
                     const uint8_t* active, const double* acc, const double* dudt,
                     const double* dt_ind, int64_t* rung, double* vel, double* u, double u_floor,
                     void* stream) {
    if (gf_ctx_init()) return -1;
    if (n < 0 || ng < 0 || ng > n || k < 0 || rungs < 1 || rungs > 62 || mode < 0 || mode > 2 ||
        (n > 0 && (!active || !acc || !rung || !vel || (mode > 0 && !dt_ind))) ||
        (ng > 0 && (!dudt || !u))) {
        gf_set_error("gf_step_kick_dev: bad arguments");
        return -1;
    }
    return gf_step_kick(n, ng, k, rungs, dt_max, mode, active, acc, dudt, dt_ind, rung, vel, u,
                        u_floor, gf_stream(stream));
}

int gf_interactions_dev(long long* out) {
    if (gf_ctx_init()) return -1;
    unsigned long long v = 0;
    if (g_ctx.walk.inter) GF_HIP(hipMemcpy(&v, g_ctx.walk.inter, sizeof(v), hipMemcpyDeviceToHost));
    *out = (long long)v;
    return 0;
}

int gf_node_interactions_dev(long long* out) {
    if (gf_ctx_init()) return -1;
    unsigned long long v = 0;
    if (g_ctx.walk.inter_nodes)
        GF_HIP(hipMemcpy(&v, g_ctx.walk.inter_nodes, sizeof(v), hipMemcpyDeviceToHost));
    *out = (long long)v;
    return 0;
}

int gf_sync(void* stream) {
    if (gf_ctx_init()) return -1;
    GF_HIP(hipStreamSynchronize(gf_stream(stream)));
    if (g_ctx.walk.ovf_count && check_walk_errors()) return -1;
    return 0;
}

int nbody_render(int n, const double* pos, const uint8_t* group, int ngroups, const double* rot,
                 const double* center, double scale, int W, int H, float* img) {
    if (n < 0 || ngroups < 1 || W < 1 || H < 1 || !rot || !center || !img ||
        (n > 0 && (!pos || !group))) {
        gf_set_error("nbody_render: bad sizes or null pointers (n %d, ngroups %d, %d x %d)", n,
                     ngroups, W, H);
        return -1;
    }
    for (int i = 0; i < n; ++i)          // a group outside the image planes would write past them
        if (group[i] >= ngroups) {
            gf_set_error("nbody_render: particle %d has group %d of %d", i, int(group[i]),
                         ngroups);
            return -1;
        }
    if (gf_ctx_init()) return -1;
    hipStream_t s = g_ctx.stream;
    const size_t bimg = sizeof(float) * size_t(ngroups) * W * H;
    if (STAGE(g_ctx.d_pos, sizeof(double) * 3 * size_t(n)) || STAGE(g_ctx.d_act, size_t(n)) ||
        STAGE(g_ctx.d_img, bimg) || STAGE(g_ctx.d_small, sizeof(double) * 12))
        return -1;
    if (n > 0) {
        if (gf_copy_h2d(g_ctx.d_pos, pos, sizeof(double) * 3 * size_t(n), s)) return -1;
        if (gf_copy_h2d(g_ctx.d_act, group, size_t(n), s)) return -1;
    }
    double small[12];
    memcpy(small, rot, sizeof(double) * 9);
    memcpy(small + 9, center, sizeof(double) * 3);
    if (gf_copy_h2d(g_ctx.d_small, small, sizeof(small), s)) return -1;
    double* dsm = reinterpret_cast<double*>(g_ctx.d_small);
    if (gf_render_deposit(n, g_ctx.d_pos, g_ctx.d_act, ngroups, dsm, dsm + 9, scale, W, H,
                          reinterpret_cast<float*>(g_ctx.d_img), s))
        return -1;
    if (gf_copy_d2h(img, g_ctx.d_img, bimg, s)) return -1;
    GF_HIP(hipStreamSynchronize(s));
    return 0;
}

int nbody_tonemap(int W, int H, int ngroups, const float* img, const float* glow, int gw, int gh,
                  const float* colors, float inv_ref, float K, float exposure, uint8_t* out) {
    // the glow's bilinear stencil reads two cells along each axis
    if (W < 1 || H < 1 || ngroups < 1 || gw < 2 || gh < 2 || !img || !glow || !colors || !out) {
        gf_set_error("nbody_tonemap: bad sizes or null pointers (%d x %d, %d groups, glow %d x %d)",
                     W, H, ngroups, gw, gh);
        return -1;
    }
    if (gf_ctx_init()) return -1;
    hipStream_t s = g_ctx.stream;
    const size_t bimg = sizeof(float) * size_t(ngroups) * W * H;
    const size_t bglow = sizeof(float) * size_t(ngroups) * gw * gh;
    const size_t bout = 3 * size_t(W) * H;
    if (STAGE(g_ctx.d_img, bimg) || STAGE(g_ctx.d_glow, bglow) || STAGE(g_ctx.d_out8, bout) ||
        STAGE(g_ctx.d_small, sizeof(float) * 3 * size_t(ngroups) + 96))
        return -1;
    if (gf_copy_h2d(g_ctx.d_img, img, bimg, s)) return -1;
    if (gf_copy_h2d(g_ctx.d_glow, glow, bglow, s)) return -1;
    if (gf_copy_h2d(g_ctx.d_small, colors, sizeof(float) * 3 * size_t(ngroups), s)) return -1;
    if (gf_render_tonemap(W, H, ngroups, reinterpret_cast<const float*>(g_ctx.d_img),
                          reinterpret_cast<const float*>(g_ctx.d_glow), gw, gh,
                          reinterpret_cast<const float*>(g_ctx.d_small), inv_ref, K, exposure,
                          reinterpret_cast<uint8_t*>(g_ctx.d_out8), s))
        return -1;
    if (gf_copy_d2h(out, g_ctx.d_out8, bout, s)) return -1;
    GF_HIP(hipStreamSynchronize(s));
    return 0;
}

// ------------------------------------------------------------------ SPH

int gf_sph_kernel(int n, const double* r, const double* h, double* w, double* g) {
    if (gf_ctx_init()) return -1;
    if (n <= 0) return 0;
    hipStream_t s = g_ctx.stream;
    const size_t b = sizeof(double) * size_t(n);
    if (STAGE(g_ctx.d_pos, 2 * b) || STAGE(g_ctx.d_out1, 2 * b)) return -1;
    if (gf_copy_h2d(g_ctx.d_pos, r, b, s)) return -1;
    if (gf_copy_h2d(g_ctx.d_pos + n, h, b, s)) return -1;
    if (gf_sph_kernel_eval(n, g_ctx.d_pos, g_ctx.d_pos + n, w ? g_ctx.d_out1 : nullptr,
                           g ? g_ctx.d_out1 + n : nullptr, s))
        return -1;
    if (w && gf_copy_d2h(w, g_ctx.d_out1, b, s)) return -1;
    if (g && gf_copy_d2h(g, g_ctx.d_out1 + n, b, s)) return -1;
    GF_HIP(hipStreamSynchronize(s));
    return 0;
}

int gf_sph_forces(int n, int k, const double* pos, const double* vel, const double* mass,
                  const int64_t* nb, const double* h, const double* u, double gamma, double alpha,
                  double beta, double* rho, double* acc, double* dudt, double* dt_i, double* divv) {
    if (gf_ctx_init()) return -1;
    if (n < 0 || k < 0) {
        gf_set_error("gf_sph_forces: bad n or k");
        return -1;
    }
    if (n == 0) return 0;
    hipStream_t s = g_ctx.stream;
    const size_t b1 = sizeof(double) * size_t(n), b3 = 3 * b1;
    if (STAGE(g_ctx.d_pos, b3) || STAGE(g_ctx.d_vel, b3) || STAGE(g_ctx.d_mass, b1) ||
        STAGE(g_ctx.d_h, b1) || STAGE(g_ctx.d_u, b1) || STAGE(g_ctx.d_acc, b3) ||
        STAGE(g_ctx.d_out1, 4 * b1) || STAGE(g_ctx.d_nb, sizeof(int64_t) * size_t(n) * k) ||
        STAGE(g_ctx.d_scratch, gf_sph_table_scratch_bytes(n, k)))
        return -1;
    if (gf_copy_h2d(g_ctx.d_pos, pos, b3, s)) return -1;
    if (gf_copy_h2d(g_ctx.d_vel, vel, b3, s)) return -1;
    if (gf_copy_h2d(g_ctx.d_mass, mass, b1, s)) return -1;
    if (gf_copy_h2d(g_ctx.d_h, h, b1, s)) return -1;
    if (gf_copy_h2d(g_ctx.d_u, u, b1, s)) return -1;
    if (gf_copy_h2d(g_ctx.d_nb, nb, sizeof(int64_t) * size_t(n) * k, s)) return -1;
    double* o = g_ctx.d_out1;
    if (gf_sph_forces_table(n, k, g_ctx.d_pos, g_ctx.d_vel, g_ctx.d_mass, g_ctx.d_nb, g_ctx.d_h,
                            g_ctx.d_u, gamma, alpha, beta, o, g_ctx.d_acc, o + n, o + 2 * size_t(n),
                            o + 3 * size_t(n), g_ctx.d_scratch, s))
        return -1;
    if (rho && gf_copy_d2h(rho, o, b1, s)) return -1;
    if (acc && gf_copy_d2h(acc, g_ctx.d_acc, b3, s)) return -1;
    if (dudt && gf_copy_d2h(dudt, o + n, b1, s)) return -1;
    if (dt_i && gf_copy_d2h(dt_i, o + 2 * size_t(n), b1, s)) return -1;
    if (divv && gf_copy_d2h(divv, o + 3 * size_t(n), b1, s)) return -1;
    GF_HIP(hipStreamSynchronize(s));
    return 0;
}

int gf_knn(int n, const double* pos, int k, const double* r_init, int64_t* nb, double* dist,
           double* mean_passes) {
    if (gf_ctx_init()) return -1;
    if (n <= 0) return 0;
    if (k < 1 || k > n) {
        gf_set_error("gf_knn: k = %d must be in [1, n = %d]", k, n);
        return -1;
    }
    hipStream_t s = g_ctx.stream;
    const size_t b1 = sizeof(double) * size_t(n), b3 = 3 * b1;
    if (STAGE(g_ctx.d_pos, b3) || STAGE(g_ctx.d_mass, b1) || STAGE(g_ctx.d_h, b1) ||
        STAGE(g_ctx.d_nb, sizeof(int64_t) * size_t(n) * k) ||
        STAGE(g_ctx.d_dist, sizeof(double) * size_t(n) * k))
        return -1;
    if (gf_copy_h2d(g_ctx.d_pos, pos, b3, s)) return -1;
    GF_HIP(hipMemsetAsync(g_ctx.d_mass, 0, b1, s));
    if (r_init && gf_copy_h2d(g_ctx.d_h, r_init, b1, s)) return -1;
    GfSph& S = g_ctx.sph;
    if (gf_sph_setup(S, n, g_ctx.d_pos, nullptr, g_ctx.d_mass, nullptr, s)) return -1;
    if (gf_sph_knn(S, k, r_init ? g_ctx.d_h : nullptr, false, true, 64, s)) return -1;
    g_ctx.sph_k = 0;   // the held table holds input indices: not usable for forces
    if (gf_sph_export_knn(S, k, g_ctx.d_nb, g_ctx.d_dist, s)) return -1;
    if (gf_copy_d2h(nb, g_ctx.d_nb, sizeof(int64_t) * size_t(n) * k, s)) return -1;
    if (gf_copy_d2h(dist, g_ctx.d_dist, sizeof(double) * size_t(n) * k, s)) return -1;
    GF_HIP(hipStreamSynchronize(s));
    if (mean_passes) {
        int* p = new int[size_t(n)];
        GF_HIP(hipMemcpy(p, S.passes, sizeof(int) * size_t(n), hipMemcpyDeviceToHost));
        long long t = 0;
        for (int i = 0; i < n; ++i) t += p[i];
        delete[] p;
        *mean_passes = double(t) / double(n);
    }
    return S.last_passes;
}

int gf_knn_from_candidates_host(int n, int m, const double* pos, const int64_t* cand, int k,
                                int64_t* nb, double* dist) {
    if (gf_ctx_init()) return -1;
    if (n <= 0) return 1;
    hipStream_t s = g_ctx.stream;
    const size_t b3 = sizeof(double) * 3 * size_t(n);
    if (STAGE(g_ctx.d_pos, b3) || STAGE(g_ctx.d_nb, sizeof(int64_t) * size_t(n) * (m + k)) ||
        STAGE(g_ctx.d_dist, sizeof(double) * size_t(n) * k) || STAGE(g_ctx.d_flag, sizeof(int)))
        return -1;
    int one = 1;
    if (gf_copy_h2d(g_ctx.d_pos, pos, b3, s)) return -1;
    int64_t* dcand = g_ctx.d_nb;
    int64_t* dnb = g_ctx.d_nb + size_t(n) * m;
    if (gf_copy_h2d(dcand, cand, sizeof(int64_t) * size_t(n) * m, s)) return -1;
    if (gf_copy_h2d(dnb, nb, sizeof(int64_t) * size_t(n) * k, s)) return -1;
    if (gf_copy_h2d(g_ctx.d_dist, dist, sizeof(double) * size_t(n) * k, s)) return -1;
    if (gf_copy_h2d(g_ctx.d_flag, &one, sizeof(int), s)) return -1;
    if (gf_knn_from_candidates(n, m, g_ctx.d_pos, dcand, k, dnb, g_ctx.d_dist, g_ctx.d_flag, s))
        return -1;
    int ok = 0;
    if (gf_copy_d2h(nb, dnb, sizeof(int64_t) * size_t(n) * k, s)) return -1;
    if (gf_copy_d2h(dist, g_ctx.d_dist, sizeof(double) * size_t(n) * k, s)) return -1;
    if (gf_copy_d2h(&ok, g_ctx.d_flag, sizeof(int), s)) return -1;
    GF_HIP(hipStreamSynchronize(s));
    return ok;
}

// sph.NeighbourFinder (sph.py:138-170): full search into the device cache,
// re-selection from it.

static int check_gas_list(const char* who, int n_total, int n, const int64_t* g) {
    if (n_total < 0 || n < 0 || n > n_total || (n > 0 && !g)) {
        gf_set_error("%s: %d gas particles of %d", who, n, n_total);
        return -1;
    }
    for (int i = 0; i < n; ++i)
        if (g[i] < 0 || g[i] >= n_total) {
            gf_set_error("%s: gas index %lld outside [0, %d)", who, (long long)g[i], n_total);
            return -1;
        }
    return 0;
}

int gf_nbcache_search(int n_total, int n, const int64_t* g, const double* pos, int m, int k,
                      int64_t* nb, double* dist) {
    if (check_gas_list("gf_nbcache_search", n_total, n, g)) return -1;
    if (n == 0) return 0;
    if (m < 1 || m > n || m > GF_VERLET_MAX_CAND || k < 1 || k > m) {
        gf_set_error("gf_nbcache_search: need 1 <= k = %d <= m = %d <= min(n = %d, %d)", k, m, n,
                     GF_VERLET_MAX_CAND);
        return -1;
    }
    if (gf_ctx_init()) return -1;
    hipStream_t s = g_ctx.stream;
    const size_t b1 = sizeof(double) * size_t(n), b3 = 3 * b1;
    // d_nb / d_dist: the search's n x m table, then its first k columns
    // behind it
    if (STAGE(g_ctx.d_pos, b3) || STAGE(g_ctx.d_mass, b1) || STAGE(g_ctx.d_g, sizeof(int64_t) * n) ||
        STAGE(g_ctx.d_nb, sizeof(int64_t) * size_t(n) * (m + k)) ||
        STAGE(g_ctx.d_dist, sizeof(double) * size_t(n) * (m + k)))
        return -1;
    if (gf_copy_h2d(g_ctx.d_pos, pos, b3, s)) return -1;
    if (gf_copy_h2d(g_ctx.d_g, g, sizeof(int64_t) * size_t(n), s)) return -1;
    GF_HIP(hipMemsetAsync(g_ctx.d_mass, 0, b1, s));
    GfSph& S = g_ctx.sph;
    if (gf_sph_setup(S, n, g_ctx.d_pos, nullptr, g_ctx.d_mass, nullptr, s)) return -1;
    if (gf_sph_knn(S, m, nullptr, false, true, 64, s)) return -1;
    g_ctx.sph_k = 0;   // the held table holds input indices: not usable for forces
    if (gf_sph_export_knn(S, m, g_ctx.d_nb, g_ctx.d_dist, s)) return -1;
    if (gf_verlet_store(g_ctx.verlet, n_total, n, m, g_ctx.d_g, g_ctx.d_nb, s)) return -1;
    // the first k columns of each row, compacted on the device (a pitched
    // copy into pageable host memory goes row by row)
    int64_t* knb = g_ctx.d_nb + size_t(n) * m;
    double* kdist = g_ctx.d_dist + size_t(n) * m;
    if (gf_first_columns(n, m, k, g_ctx.d_nb, g_ctx.d_dist, knb, kdist, s)) return -1;
    if (gf_copy_d2h(nb, knb, sizeof(int64_t) * size_t(n) * k, s)) return -1;
    if (gf_copy_d2h(dist, kdist, sizeof(double) * size_t(n) * k, s)) return -1;
    GF_HIP(hipStreamSynchronize(s));
    return 0;
}

int gf_nbcache_reselect(int n_total, int n, const int64_t* g, const double* pos, int k,
                        int64_t* nb, double* dist) {
    if (check_gas_list("gf_nbcache_reselect", n_total, n, g)) return -1;
    if (g_ctx.verlet.rows <= 0) return 0;
    if (n == 0) return 1;
    if (gf_ctx_init()) return -1;
    hipStream_t s = g_ctx.stream;
    const size_t b3 = sizeof(double) * 3 * size_t(n);
    if (STAGE(g_ctx.d_pos, b3) || STAGE(g_ctx.d_g, sizeof(int64_t) * n) ||
        STAGE(g_ctx.d_nb, sizeof(int64_t) * size_t(n) * k) ||
        STAGE(g_ctx.d_dist, sizeof(double) * size_t(n) * k) ||
        STAGE(g_ctx.d_scratch, sizeof(int) * size_t(n)))
        return -1;
    if (gf_copy_h2d(g_ctx.d_pos, pos, b3, s)) return -1;
    if (gf_copy_h2d(g_ctx.d_g, g, sizeof(int64_t) * size_t(n), s)) return -1;
    const int ok = gf_verlet_reselect(g_ctx.verlet, n_total, n, k, g_ctx.d_g, g_ctx.d_pos,
                                      g_ctx.d_scratch, g_ctx.d_nb, g_ctx.d_dist, s);
    if (ok <= 0) return ok;
    if (gf_copy_d2h(nb, g_ctx.d_nb, sizeof(int64_t) * size_t(n) * k, s)) return -1;
    if (gf_copy_d2h(dist, g_ctx.d_dist, sizeof(double) * size_t(n) * k, s)) return -1;
    GF_HIP(hipStreamSynchronize(s));
    return 1;
}

int gf_nbcache_info(int* n_total, int* rows, int* m) {
    if (n_total) *n_total = g_ctx.verlet.n_total;
    if (rows) *rows = g_ctx.verlet.rows;
    if (m) *m = g_ctx.verlet.m;
    return g_ctx.verlet.rows;
}

int gf_nbcache_export(int64_t* owner, int64_t* cand) {
    if (g_ctx.verlet.rows <= 0) return 0;
    if (gf_ctx_init()) return -1;
    return gf_verlet_export(g_ctx.verlet, owner, cand, g_ctx.stream);
}

// device-resident hydro: gas tree, neighbour solve, h, density, forces
int gf_hydro_setup_dev(int n, const double* pos, const double* vel, const double* mass,
                       const double* u, void* stream) {
    if (gf_ctx_init()) return -1;
    g_ctx.sph_k = 0;
    return gf_sph_setup(g_ctx.sph, n, pos, vel, mass, u, gf_stream(stream));
}

int gf_hydro_setup_from_tree_dev(int n, const double* pos, const double* vel, const double* mass,
                                 const double* u, void* stream) {
    if (gf_ctx_init()) return -1;
    g_ctx.sph_k = 0;
    if (n < 0 || n > g_ctx.tree.n) {
        gf_set_error("gf_hydro_setup_from_tree_dev: n = %d must be in [0, %d] (the particles of "
                     "the last gf_tree_build_dev)", n, g_ctx.tree.n);
        return -1;
    }
    if (n > 0 && (!g_ctx.tree.built || pos != g_ctx.tree_pos || !g_ctx.tree_sorted)) {
        gf_set_error("gf_hydro_setup_from_tree_dev: pos is not the array of the last "
                     "gf_tree_build_dev");
        return -1;
    }
    hipStream_t s = gf_stream(stream);
    if (n > 0) GF_HIP(hipStreamWaitEvent(s, g_ctx.tree_sorted, 0));
    return gf_sph_setup_from_tree(g_ctx.sph, g_ctx.tree, n, pos, vel, mass, u, s);
}

int gf_hydro_knn_dev(int k, const double* r_init, void* stream) {
    if (gf_ctx_init()) return -1;
    g_ctx.sph_k = 0;
    if (gf_sph_knn(g_ctx.sph, k, r_init, true, false, 64, gf_stream(stream))) return -1;
    g_ctx.sph_k = k;
    return g_ctx.sph.last_passes;
}

int gf_hydro_knn_begin_dev(int k, const double* r_init, int passes, void* stream) {
    if (gf_ctx_init()) return -1;
    g_ctx.sph_k = 0;
    if (passes < 0 || passes > 64) {
        gf_set_error("gf_hydro_knn_begin_dev: passes = %d must be in [0, 64]", passes);
        return -1;
    }
    if (gf_sph_knn_begin(g_ctx.sph, k, r_init, true, false, passes, gf_stream(stream))) return -1;
    g_ctx.sph_pending_k = k;
    return 0;
}

int gf_hydro_knn_more_dev(int passes, void* stream) {
    if (gf_ctx_init()) return -1;
    if (passes < 0 || passes > 64) {
        gf_set_error("gf_hydro_knn_more_dev: passes = %d must be in [0, 64]", passes);
        return -1;
    }
    return gf_sph_knn_more(g_ctx.sph, passes, gf_stream(stream));
}

int gf_hydro_knn_end_dev(void* stream) {
    if (gf_ctx_init()) return -1;
    const int k = g_ctx.sph_pending_k;
    g_ctx.sph_pending_k = 0;
    if (gf_sph_knn_end(g_ctx.sph, 64, gf_stream(stream))) return -1;
    g_ctx.sph_k = k;
    return g_ctx.sph.last_passes;
}

int gf_hydro_forces_dev(double eps, double gamma, double alpha, double beta, double* h,
                        double* rho, double* acc, double* dudt, double* dt, double* divv,
                        void* stream) {
    if (gf_ctx_init()) return -1;
    if (g_ctx.sph.tree.n > 0 && g_ctx.sph_k <= 0) {
        gf_set_error("gf_hydro_forces_dev: no neighbour solve held");
        return -1;
    }
    hipStream_t s = gf_stream(stream);
    if (gf_sph_smoothing(g_ctx.sph, g_ctx.sph_k, eps, s)) return -1;
    return gf_sph_forces_tree(g_ctx.sph, g_ctx.sph_k, gamma, alpha, beta, rho, acc, dudt, dt, divv,
                              h, s);
}

int gf_hydro_dev(int n, const double* pos, const double* vel, const double* mass, const double* u,
                 double eps, int k, const double* r_init, double gamma, double alpha, double beta,
                 double* h, double* rho, double* acc, double* dudt, double* dt, double* divv,
                 void* stream) {
    if (gf_hydro_setup_dev(n, pos, vel, mass, u, stream)) return -1;
    if (n == 0) return 0;
    const int passes = gf_hydro_knn_dev(k, r_init, stream);
    if (passes < 0) return -1;
    if (gf_hydro_forces_dev(eps, gamma, alpha, beta, h, rho, acc, dudt, dt, divv, stream)) return -1;
    return passes;
}

int gf_hydro_dk_dev(double* dk, void* stream) {
    if (gf_ctx_init()) return -1;
    if (g_ctx.sph.tree.n > 0 && g_ctx.sph_k <= 0) {
        gf_set_error("gf_hydro_dk_dev: no neighbour solve held");
        return -1;
    }
    return gf_sph_dk_out(g_ctx.sph, dk, gf_stream(stream));
}

int gf_hydro_neighbours(int64_t* nb, double* dk) {
    if (gf_ctx_init()) return -1;
    const int n = g_ctx.sph.tree.n, k = g_ctx.sph_k;
    if (n == 0) return 0;
    if (k <= 0) {
        gf_set_error("gf_hydro_neighbours: no neighbour solve held");
        return -1;
    }
    hipStream_t s = g_ctx.stream;
    if (STAGE(g_ctx.d_nb, sizeof(int64_t) * size_t(n) * k) || STAGE(g_ctx.d_out1, sizeof(double) * n))
        return -1;
    GF_HIP(hipDeviceSynchronize());
    if (gf_sph_table_out(g_ctx.sph, k, g_ctx.d_nb, s) || gf_sph_dk_out(g_ctx.sph, g_ctx.d_out1, s))
        return -1;
    if (gf_copy_d2h(nb, g_ctx.d_nb, sizeof(int64_t) * size_t(n) * k, s)) return -1;
    if (dk && gf_copy_d2h(dk, g_ctx.d_out1, sizeof(double) * n, s)) return -1;
    GF_HIP(hipStreamSynchronize(s));
    return k;
}

int gf_hydro_stats(long long* out, int len) {
    if (gf_ctx_init()) return -1;
    int nn = 0, ng = 0;
    const GfSph& S = g_ctx.sph;
    if (S.tree.counters) {
        int c[4];
        GF_HIP(hipMemcpy(c, S.tree.counters, sizeof(c), hipMemcpyDeviceToHost));
        nn = c[0];
        ng = c[1];
    }
    int items = 0;
    if (S.state && S.knn_items) GF_HIP(hipMemcpy(&items, S.state + 14, sizeof(int), hipMemcpyDeviceToHost));
    const long long v[7] = {S.last_passes, S.stat_queued, S.stat_queued2, nn, ng, S.tree.n, items};
    for (int i = 0; i < len && i < 7; ++i) out[i] = v[i];
    return 0;
}

int gf_hydro_passes(int* per_particle, long long* total) {
    if (gf_ctx_init()) return -1;
    const int n = g_ctx.sph.tree.n;
    if (n == 0) {
        if (total) *total = 0;
        return 0;
    }
    int* p = new int[size_t(n)];
    uint32_t* idx = new uint32_t[size_t(n)];
    hipError_t e1 = hipMemcpy(p, g_ctx.sph.passes, sizeof(int) * size_t(n), hipMemcpyDeviceToHost);
    hipError_t e2 = hipMemcpy(idx, g_ctx.sph.tree.idx, sizeof(uint32_t) * size_t(n), hipMemcpyDeviceToHost);
    long long t = 0;
    if (e1 == hipSuccess && e2 == hipSuccess) {
        for (int i = 0; i < n; ++i) {
            t += p[i];
            if (per_particle) per_particle[idx[i]] = p[i];
        }
    }
    delete[] p;
    delete[] idx;
    if (e1 != hipSuccess || e2 != hipSuccess) {
        gf_set_error("gf_hydro_passes: copy failed");
        return -1;
    }
    if (total) *total = t;
    return g_ctx.sph.last_passes;
}

// ------------------------------------------------------ learned h (MLP)

int gf_hydro_cell_density_dev(int k, double* out, double* out_count, void* stream) {
    if (gf_ctx_init()) return -1;
    return gf_sph_cell_density(g_ctx.sph, k, out, out_count, gf_stream(stream));
}

int gf_mlp_params(void) { return GF_MLP_PARAMS; }

int gf_mlp_forward_dev(int n, const uint16_t* x, const float* params, float* out, void* stream) {
    if (gf_ctx_init()) return -1;
    return gf_mlp_forward(n, x, params, out, gf_stream(stream));
}

int gf_mlp_grad_dev(int n, const uint16_t* x, const float* target, const float* params,
                    float tau, float* grad, float* loss, void* stream) {
    if (gf_ctx_init()) return -1;
    if (STAGE(g_ctx.d_mlp_part, sizeof(float) * gf_mlp_partial_floats(n))) return -1;
    return gf_mlp_grad(n, n, x, target, params, tau, grad, loss, g_ctx.d_mlp_part,
                       gf_stream(stream));
}

int gf_mlp_train_step_dev(int n, const uint16_t* x, const float* target, float* params, float* m,
                          float* v, int step, float lr, float beta1, float beta2, float eps,
                          float tau, float* loss, void* stream) {
    if (gf_ctx_init()) return -1;
    if (step < 1) {
        gf_set_error("gf_mlp_train_step_dev: step counts from 1");
        return -1;
    }
    if (STAGE(g_ctx.d_mlp_part, sizeof(float) * gf_mlp_partial_floats(n))) return -1;
    return gf_mlp_train_step(n, x, target, params, m, v, step, lr, beta1, beta2, eps, tau, loss,
                             g_ctx.d_mlp_part, gf_stream(stream));
}

// No gf_ctx_init here: the library context creates a stream, and this entry
// point allocates nothing.  It runs on the caller's stream and device.
int gf_mlp_train_step_noalloc_dev(int n, const uint16_t* x, const float* target, float* params,
                                  float* m, float* v, int step, float lr, float beta1,
                                  float beta2, float eps, float tau, float* loss, void* stream) {
    if (step < 1) {
        gf_set_error("gf_mlp_train_step_noalloc_dev: step counts from 1");
        return -1;
    }
    return gf_mlp_train_step_noalloc(n, x, target, params, m, v, step, lr, beta1, beta2, eps, tau,
                                     loss, gf_stream(stream));
}

int gf_mlp_features_dev(int n, int k, const double* dk_prev, const double* cell_rho,
                        const double* cell_count, const double* divv, const double* dtc_prev,
                        float dt, const float* norm, uint16_t* x, float* raw, void* stream) {
    if (gf_ctx_init()) return -1;
    return gf_mlp_features(n, k, dk_prev, cell_rho, cell_count, divv, dtc_prev, dt, norm, x, raw,
                           gf_stream(stream));
}

int gf_mlp_propose_dev(int n, const uint16_t* x, const float* params, const double* dk_prev,
                       float margin, double* r, void* stream) {
    if (gf_ctx_init()) return -1;
    return gf_mlp_propose(n, x, params, dk_prev, margin, r, gf_stream(stream));
}

// host-pointer gradient (tests)
int gf_mlp_grad_host(int n, const uint16_t* x, const float* target, const float* params,
                     float tau, float* grad, float* loss) {
    if (gf_ctx_init()) return -1;
    hipStream_t s = g_ctx.stream;
    const size_t bx = sizeof(uint16_t) * 4 * size_t(n), bt = sizeof(float) * size_t(n);
    const size_t bp = sizeof(float) * GF_MLP_PARAMS;
    const size_t need = bx + bt + 2 * bp + 64;
    if (STAGE(g_ctx.d_mlp, need) || STAGE(g_ctx.d_mlp_part, sizeof(float) * gf_mlp_partial_floats(n)))
        return -1;
    char* base = reinterpret_cast<char*>(g_ctx.d_mlp);
    float* dp = reinterpret_cast<float*>(base);
    float* dg = dp + GF_MLP_PARAMS;
    float* dl = dg + GF_MLP_PARAMS;
    float* dt = dl + 4;
    uint16_t* dx = reinterpret_cast<uint16_t*>(dt + n);
    if (gf_copy_h2d(dp, params, bp, s)) return -1;
    if (gf_copy_h2d(dt, target, bt, s)) return -1;
    if (gf_copy_h2d(dx, x, bx, s)) return -1;
    if (gf_mlp_grad(n, n, dx, dt, dp, tau, dg, dl, g_ctx.d_mlp_part, s)) return -1;
    if (gf_copy_d2h(grad, dg, bp, s)) return -1;
    if (loss && gf_copy_d2h(loss, dl, sizeof(float), s)) return -1;
    GF_HIP(hipStreamSynchronize(s));
    return 0;
}


// ------------------------------------------------------------ gas physics
// cooling.py, simulation.py:97-104, star_formation.py (gf_gas.hip).

int gf_gas_constants(double* out) {
    if (!out) {
        gf_set_error("gf_gas_constants: null output");
        return -1;
    }
    const GfGasUnits& u = gf_gas_units();
    const double v[8] = {u.msun, u.msun_g, u.m_h, u.x_h, u.rho_cgs, u.u_cgs, u.myr, u.sec_per_myr};
    memcpy(out, v, sizeof(v));
    return 0;
}

int gf_cooling_function_dev(int n, const double* T, double* lam, void* stream) {
    if (gf_ctx_init()) return -1;
    if (n < 0 || (n > 0 && (!T || !lam))) {
        gf_set_error("gf_cooling_function_dev: bad arguments");
        return -1;
    }
    return gf_cooling_function(n, T, lam, gf_stream(stream));
}

int gf_cool_dev(int n, const double* u, const double* T, const double* rho, double dt,
                const uint8_t* blocked, double u_floor, double* out, void* stream) {
    if (gf_ctx_init()) return -1;
    if (n < 0 || (n > 0 && (!u || !T || !rho || !out))) {
        gf_set_error("gf_cool_dev: bad arguments");
        return -1;
    }
    return gf_cool(n, u, T, rho, dt, blocked, u_floor, out, gf_stream(stream));
}

int gf_thermo_dev(int ng, const int64_t* g, int n_total, double* u, const double* dudt,
                  const double* rho, const double* no_cool_until, double t, double dt,
                  double u_per_K, double u_floor, void* stream) {
    if (gf_ctx_init()) return -1;
    if (ng < 0 || (ng > 0 && (!u || !dudt || !rho)) || (!g && n_total < ng)) {
        gf_set_error("gf_thermo_dev: bad arguments");
        return -1;
    }
    return gf_thermo_launch(g_ctx.gas, ng, g, n_total, u, dudt, rho, no_cool_until, t, dt,
                            u_per_K, u_floor, gf_stream(stream));
}

int gf_gas_errors_dev(void* stream) {
    if (gf_ctx_init()) return -1;
    return gf_gas_errors(g_ctx.gas, gf_stream(stream));
}

int gf_sf_candidates_dev(int ng, const int64_t* g, int n_total, const double* u,
                         const double* rho, double u_per_K, double rho_sf, double sf_eff,
                         double dt, int64_t* rows, double* p, void* stream) {
    if (gf_ctx_init()) return -1;
    if (ng < 0 || (ng > 0 && (!u || !rho || !rows || !p)) || (!g && n_total < ng)) {
        gf_set_error("gf_sf_candidates_dev: bad arguments");
        return -1;
    }
    return gf_sf_candidates(g_ctx.gas, ng, g, n_total, u, rho, u_per_K, rho_sf, sf_eff, dt, rows,
                            p, gf_stream(stream));
}

int gf_sf_select_dev(int m, const int64_t* rows, const double* p, const double* uniform,
                     int64_t* out, void* stream) {
    if (gf_ctx_init()) return -1;
    if (m < 0 || (m > 0 && (!rows || !p || !uniform || !out))) {
        gf_set_error("gf_sf_select_dev: bad arguments");
        return -1;
    }
    return gf_sf_select(g_ctx.gas, m, rows, p, uniform, out, gf_stream(stream));
}

int gf_knn_query_dev(int n_src, const double* src, int n_q, const double* q, int k, int64_t* nb,
                     double* dist, void* stream) {
    if (gf_ctx_init()) return -1;
    if (n_src < 1 || n_q < 0 || !src || (n_q > 0 && (!q || !nb))) {
        gf_set_error("gf_knn_query_dev: bad arguments");
        return -1;
    }
    return gf_knn_query(g_ctx.gas, n_src, src, n_q, q, k, nb, dist, gf_stream(stream));
}

int gf_feedback_heat_dev(int n_new, int k, const int64_t* targets, const double* new_mass,
                         double e_per_g, double until, int n_total, const double* mass, double* u,
                         double* no_cool_until, void* stream) {
    if (gf_ctx_init()) return -1;
    if (n_new < 0 || n_total < 1 ||
        (n_new > 0 && (!targets || !new_mass || !mass || !u || !no_cool_until))) {
        gf_set_error("gf_feedback_heat_dev: bad arguments");
        return -1;
    }
    return gf_feedback_heat(g_ctx.gas, n_new, k, targets, new_mass, e_per_g, until, n_total, mass,
                            u, no_cool_until, gf_stream(stream));
}

}  // extern "C"
