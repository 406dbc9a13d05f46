// Device MT19937 disorder streams (rng.hip) and their host-side jump tables.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace rlm {

// Jumps of a set of streams: kind 0 (none), 1 (mono words, mono < 19937: stepped) or
// 2 (poly[i * 312 ..]: t^skip mod phi, applied by correlation).
struct MtJumpTable {
    long count = 0;
    std::vector<int> kind, mono;
    std::vector<uint64_t> poly;
    void set(long n);
    void fill(long i, uint64_t skip);
};

// True once the MT19937 transition's minimal polynomial is known (computed on first use).
bool mt_jump_ready();
size_t mt_stream_lds_bytes();

// n_streams workgroups: stream b = init_by_array(seed), jump A (kindA, monoA, d_polyA),
// jump B of stream b (d_kindB[b], d_monoB[b], d_polyB + 312 b), then `words` outputs
// converted by mode (0: rlm x / y / sign sections of n_sec = words / 3 words each, x and y
// scaled by scale0 and scale1; 1: hlm, -1 + 2u scaled by scale0).
int launch_mt_streams(uint32_t seed, int kindA, int monoA, const uint64_t *d_polyA, const int *d_kindB,
                      const int *d_monoB, const uint64_t *d_polyB, int n_streams, long words, int mode, long n_sec,
                      double scale0, double scale1, double *out0, double *out1, double *out2, hipStream_t stream);

}  // namespace rlm
