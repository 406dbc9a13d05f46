// MKL-compatible MT19937 disorder streams generated on the device.
//
// The reference draws each rlm sample from vslNewStream(VSL_BRNG_MT19937, seed) after
// vslSkipAheadStream(3 * impurity_num * sample) (rlm/disorder_potential.cpp:5-17), and
// the hlm samples consecutively from one stream (hlm/main_hlm_cpu.cpp:67,79). Here
// every sample of a batch is one workgroup: it seeds the window (init_by_array), jumps
// it to the batch's first sample and then to its own sample (jump polynomials from
// mt_jump.hpp, computed on the host; they do not depend on the seed), and generates its
// words in chunks of 227 (the largest set of words whose recurrence inputs are all
// known), tempering and converting each word straight to its output slot.
#include <cstdint>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "mt_jump.hpp"
#include "rng.hpp"

namespace rlm {

namespace {

constexpr int kMtN = mtj::kN, kMtM = mtj::kM, kChunk = kMtN - kMtM;   // 227
constexpr int kSeq = mtj::kDeg + kMtN;    // words of the sequence a correlation reads
constexpr int kRing = 1024;               // ring of the generation phase (slot = index & 1023)
constexpr int kNT = 256;

// seq[0..623] holds a window; extend the sequence to seq[0 .. last] (last < kSeq)
__device__ void extend_sequence(uint32_t *seq, int last) {
    for (int c = 0; c + kMtN <= last; c += kChunk) {
        const int k = c + (int)threadIdx.x;
        if (threadIdx.x < kChunk && k + kMtN <= last) seq[k + kMtN] = mtj::next_word(seq[k], seq[k + 1], seq[k + kMtM]);
        __syncthreads();
    }
}

// seq[0..623] <- jump(seq window) for a jump of `kind` (0: none, 1: `mono` words with
// mono < 19937, 2: the polynomial g, 312 words)
__device__ void apply_jump(uint32_t *seq, int kind, int mono, const uint64_t *g) {
    if (kind == 0) return;
    if (kind == 1) {
        extend_sequence(seq, mono + kMtN - 1);
        uint32_t v[3];
#pragma unroll
        for (int r = 0; r < 3; r++) {
            const int j = (int)threadIdx.x + r * kNT;
            v[r] = j < kMtN ? seq[mono + j] : 0u;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 3; r++) {
            const int j = (int)threadIdx.x + r * kNT;
            if (j < kMtN) seq[j] = v[r];
        }
        __syncthreads();
        return;
    }
    extend_sequence(seq, kSeq - 1);
    // word j of the jumped window: XOR over the set bits i of g of seq[i + j]
    uint32_t acc[3] = {0u, 0u, 0u};
    const int j0 = (int)threadIdx.x;
    for (int w = 0; w < mtj::kPolyWords; w++) {
        uint64_t bits = g[w];   // uniform across the workgroup
        while (bits) {
            const int b = __builtin_ctzll(bits);
            bits &= bits - 1;
            const int i = 64 * w + b;
            if (i >= mtj::kDeg) break;
            acc[0] ^= seq[i + j0];
            acc[1] ^= seq[i + j0 + kNT];
            if (j0 + 2 * kNT < kMtN) acc[2] ^= seq[i + j0 + 2 * kNT];
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const int j = j0 + r * kNT;
        if (j < kMtN) seq[j] = acc[r];
    }
    __syncthreads();
}

// One workgroup per stream b: window of init_by_array(seed), jump A (common), jump B
// (per stream: kindB[b], monoB[b], polyB + b * 312), then `words` outputs, output word
// w converted by `mode`:
//   mode 0 (rlm, sections of n_sec words): w < n_sec -> out0[w] = scale0 * u,
//     w < 2 n_sec -> out1[w - n_sec] = scale1 * u, else out2[w - 2 n_sec] = +1 for an
//     even word, -1 for an odd one (rlm/disorder_potential.cpp:9-17);
//   mode 1 (hlm): out0[w] = -1 + 2 u, times scale0 unless scale0 == 1
//     (hlm/main_hlm_cpu.cpp:79, width W: scale0 = W / 2);
//   mode 2 (rlm, the single-precision GPU builds' stream): as mode 0 with x and y the
//     float values L * float(word) * 2^-32 (vsRngUniform), written negated: the GPU
//     builds' V_mn sums e^{+i phase} (rlm/potential_coeff.cu:14-15), which the FP64
//     potential kernel's e^{-i phase} gives for negated positions (exactly);
// with u = word * 2^-32 (vdRngUniform's a + (b - a) * word * 2^-32). Outputs of stream
// b start at out + b * n_sec (mode 0) or out0 + b * words (mode 1).
__global__ __launch_bounds__(kNT) void mt_stream_kernel(uint32_t seed, int kindA, int monoA,
                                                        const uint64_t *__restrict__ polyA,
                                                        const int *__restrict__ kindB, const int *__restrict__ monoB,
                                                        const uint64_t *__restrict__ polyB, long words, int mode,
                                                        long n_sec, double scale0, double scale1, double *out0,
                                                        double *out1, double *out2) {
    extern __shared__ uint32_t seq[];   // kSeq words
    const long b = blockIdx.x;
    if (threadIdx.x == 0) mtj::init_by_seed(seed, seq);
    __syncthreads();
    apply_jump(seq, kindA, monoA, polyA);
    apply_jump(seq, kindB[b], monoB[b], polyB + b * mtj::kPolyWords);
    // move the window into the ring (slot = sequence index & (kRing - 1)); the ring
    // lives past the correlation buffer's window part
    uint32_t *ring = seq + kMtN;
    {
        uint32_t v[3];
#pragma unroll
        for (int r = 0; r < 3; r++) {
            const int j = (int)threadIdx.x + r * kNT;
            v[r] = j < kMtN ? seq[j] : 0u;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 3; r++) {
            const int j = (int)threadIdx.x + r * kNT;
            if (j < kMtN) ring[j] = v[r];
        }
        __syncthreads();
    }
    const bool secs = mode != 1;
    double *o0 = secs ? out0 + b * n_sec : out0 + b * words;
    double *o1 = secs ? out1 + b * n_sec : nullptr;
    double *o2 = secs ? out2 + b * n_sec : nullptr;
    // output w is the tempered x_{624 + w}, computed from x_w, x_{w+1}, x_{w+397}
    for (long c = 0; c < words; c += kChunk) {
        const long w = c + threadIdx.x;
        if (threadIdx.x < kChunk && w < words) {
            const uint32_t x = mtj::next_word(ring[w & (kRing - 1)], ring[(w + 1) & (kRing - 1)],
                                             ring[(w + kMtM) & (kRing - 1)]);
            ring[(w + kMtN) & (kRing - 1)] = x;
            const uint32_t t = mtj::temper(x);
            const double u = (double)t * 0x1p-32;
            if (mode == 0) {
                if (w < n_sec) {
                    o0[w] = __dmul_rn(scale0, u);
                } else if (w < 2 * n_sec) {
                    o1[w - n_sec] = __dmul_rn(scale1, u);
                } else {
                    o2[w - 2 * n_sec] = (t & 1u) ? -1.0 : 1.0;
                }
            } else if (mode == 2) {
                // single-precision stream of the reference's FP32 builds (vsRngUniform:
                // L * float(word) * 2^-32 in float arithmetic), positions negated (see
                // launch_mt_streams)
                const float uf = (float)t * 0x1p-32f;
                if (w < n_sec) {
                    o0[w] = -(double)__fmul_rn((float)scale0, uf);
                } else if (w < 2 * n_sec) {
                    o1[w - n_sec] = -(double)__fmul_rn((float)scale1, uf);
                } else {
                    o2[w - 2 * n_sec] = (t & 1u) ? -1.0 : 1.0;
                }
            } else {
                const double v = __dadd_rn(-1.0, __dmul_rn(2.0, u));
                o0[w] = scale0 == 1.0 ? v : __dmul_rn(scale0, v);
            }
        }
        __syncthreads();
    }
}

}  // namespace

// ---------------------------------------------------------------- host side
void MtJumpTable::set(long n) {
    count = n;
    kind.assign(n, 0);
    mono.assign(n, 0);
    poly.assign((size_t)n * mtj::kPolyWords, 0);
}

void MtJumpTable::fill(long i, uint64_t skip) {
    if (skip == 0) {
        kind[i] = 0;
        mono[i] = 0;
    } else if (skip < (uint64_t)mtj::kDeg) {
        kind[i] = 1;
        mono[i] = (int)skip;
    } else {
        kind[i] = 2;
        mono[i] = 0;
        const mtj::Poly g = mtj::powmod_t(skip);
        std::copy(g.begin(), g.end(), poly.begin() + i * mtj::kPolyWords);
    }
}

bool mt_jump_ready() { return !mtj::charpoly().empty(); }

size_t mt_stream_lds_bytes() { return sizeof(uint32_t) * (size_t)(kSeq > kMtN + kRing ? kSeq : kMtN + kRing); }

int launch_mt_streams(uint32_t seed, int kindA, int monoA, const uint64_t *d_polyA, const int *d_kindB,
                      const int *d_monoB, const uint64_t *d_polyB, int n_streams, long words, int mode, long n_sec,
                      double scale0, double scale1, double *out0, double *out1, double *out2, hipStream_t stream) {
    if (n_streams < 1 || words < 0 || ((mode == 0 || mode == 2) && 3 * n_sec != words) || mode < 0 || mode > 2)
        return -22;
    if (kindA < 0 || kindA > 2 || (kindA == 1 && (monoA < 1 || monoA >= mtj::kDeg)) || (kindA == 2 && !d_polyA))
        return -22;
    const size_t lds = mt_stream_lds_bytes();
    static std::once_flag once;
    static int attr_rc = 0;
    std::call_once(once, [&] {
        attr_rc = hipFuncSetAttribute((const void *)mt_stream_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds) == hipSuccess ? 0 : -1;
    });
    if (attr_rc) return attr_rc;
    hipLaunchKernelGGL(mt_stream_kernel, dim3(n_streams), dim3(kNT), lds, stream, seed, kindA, monoA, d_polyA, d_kindB,
                       d_monoB, d_polyB, words, mode, n_sec, scale0, scale1, out0, out1, out2);
    RLM_HIP_CHECK(hipGetLastError());
    return 0;
}

}  // namespace rlm
