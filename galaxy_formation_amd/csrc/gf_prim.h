// Device-wide scan and radix sort (gf_prim.hip).
#pragma once
#include "gf_common.h"

// Ints of temporary storage gf_scan_excl needs for n inputs.  The storage
// holds the single-pass scan's state across calls: clear it once after
// allocating it (gf_scan_temp_clear, with the n it was sized for), and let
// only one stream at a time scan with it.
size_t gf_scan_temp_ints(int n);
int gf_scan_temp_clear(int* temp, int n);
// out[i] = in[0] + ... + in[i-1] for i in [0, n]; out has n + 1 entries
// (out[n] is the total).  in == out is allowed; both 16-byte aligned.  With
// len (device) given, only the first m = min(n, *len) inputs are scanned:
// out[0..m] is written, out past m is not.  One launch: one block up to
// 65,536 inputs, else a single pass with decoupled look-back.
int gf_scan_excl(const int* in, int* out, int n, int* temp, hipStream_t s,
                 const int* len = nullptr);
// Look-back waits abandoned since the library was loaded (0 unless a
// single-pass launch went wrong), and the device address of that counter on
// the current device, which single-pass launches (gf_lookback.h) take.
unsigned int gf_prim_lookback_timeouts();
unsigned int* gf_prim_timeout_counter();

// Ints of temporary storage gf_radix_sort_pairs needs for n pairs.
size_t gf_sort_temp_ints(int n);
// Stable LSD radix sort of (key, val) on key bits [begin_bit, end_bit), 8 bits
// per pass.  The sorted pairs end in (key_alt, val_alt) when *in_alt is set on
// return, else in (key, val).
int gf_radix_sort_pairs(uint64_t* key, uint32_t* val, uint64_t* key_alt, uint32_t* val_alt, int n,
                        int begin_bit, int end_bit, int* temp, hipStream_t s, int* in_alt);
