// Library context shared by the C ABI translation units.
#pragma once
#include "../../include/gf.h"
#include "gf_gravity.h"
#include "gf_sph.h"
#include "gf_mlp.h"

#define GF_ABI_VERSION 1

#include "gf_gas.h"
#include "gf_verlet.h"

struct GfCtx {
    int inited = 0;
    int device = 0;
    int n_cu = 0;
    hipStream_t stream = nullptr;
    int precision = GF_FP32;
    unsigned long long last_inter = 0;
    GfTree tree;
    // the last gf_tree_build_dev: its positions, and an event recorded once
    // its sorted keys are kept (gas trees built from them wait for it)
    const double* tree_pos = nullptr;
    hipEvent_t tree_sorted = nullptr;
    GfWalkScratch walk;
    GfSph sph;
    int sph_k = 0;                  // k of the held neighbour solve
    int sph_pending_k = 0;          // k of a solve begun and not yet ended
    // staging buffers of the host-pointer entry points (grown, never shrunk)
    double* d_pos = nullptr; size_t d_pos_cap = 0;
    double* d_mass = nullptr; size_t d_mass_cap = 0;
    double* d_acc = nullptr; size_t d_acc_cap = 0;
    uint8_t* d_act = nullptr; size_t d_act_cap = 0;
    int64_t* d_tgt = nullptr; size_t d_tgt_cap = 0;
    double* d_partial = nullptr; size_t d_partial_cap = 0;
    void* d_img = nullptr; size_t d_img_cap = 0;
    void* d_glow = nullptr; size_t d_glow_cap = 0;
    void* d_out8 = nullptr; size_t d_out8_cap = 0;
    void* d_small = nullptr; size_t d_small_cap = 0;
    double* d_vel = nullptr; size_t d_vel_cap = 0;
    double* d_h = nullptr; size_t d_h_cap = 0;
    double* d_u = nullptr; size_t d_u_cap = 0;
    double* d_out1 = nullptr; size_t d_out1_cap = 0;   // 5 n doubles of scalar outputs
    int64_t* d_nb = nullptr; size_t d_nb_cap = 0;
    double* d_dist = nullptr; size_t d_dist_cap = 0;
    int* d_scratch = nullptr; size_t d_scratch_cap = 0;
    int* d_flag = nullptr; size_t d_flag_cap = 0;
    float* d_mlp = nullptr; size_t d_mlp_cap = 0;
    float* d_mlp_part = nullptr; size_t d_mlp_part_cap = 0;
    GfGas gas;                      // gas-physics scratch (cooling, star formation, feedback)
    GfVerlet verlet;                // sph.NeighbourFinder's candidate cache
    int64_t* d_g = nullptr; size_t d_g_cap = 0;
};

extern GfCtx g_ctx;
int gf_ctx_init();
hipStream_t gf_stream(void* s);
int gf_stage_reserve(void** p, size_t* cap, size_t bytes);
