// Neural radiance cache: host side of the MLP (parameters, Adam state, training and inference).
#pragma once
#include <cstdint>
#include "nrc_mlp.h"
#include "scene.h"

namespace nr {

struct NrcConfig {
    float lr = 1e-3f, beta1 = 0.9f, beta2 = 0.99f, eps = 1e-8f;
    float ema = 0.99f;         // inference uses the bias-corrected exponential average of the weights
    int max_batch = 1 << 16;   // training samples per step
};

constexpr int TRAIN_MAX_BLOCKS = 256;  // training blocks per step (one weight-gradient partial each)

class Nrc {
public:
    NrcConfig cfg;
    int step = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    DevBuf w32, w16, m, v;     // FP32 master weights, BF16 copy (training), Adam moments
    DevBuf w16t;               // BF16 weights transposed per layer (row = input): the backward's operand
    DevBuf wema, w16inf;       // FP32 weight average, BF16 inference weights
    DevBuf part;               // training: per-block weight-gradient partials, [6][blocks][4096] fp32
    DevBuf lossp;              // training: summed loss of each 64-sample chunk
    DevBuf q_tmp, t_tmp, o_tmp;  // staging for host-buffer entry points

    Nrc(uint32_t seed, const NrcConfig& c);
    ~Nrc();
    // He-uniform initialisation from a seeded splitmix64 stream (host), padded entries zero.
    void init_weights(uint32_t seed);
    void set_weights(const float* w);   // NRC_NPARAM floats (W^T per layer)
    void get_weights(float* w);

    // Device-buffer entry points, on stream s (nullptr: the cache's own stream).
    void infer(const NrcQuery* q, int n, float4* out, hipStream_t s = nullptr);
    // Renderer form: radiance[slot] += beta * max(0, y) for queries [0, min(*count, max_n)), with
    // slot = bits of q.c.w and beta = q.d.xyz (one query per slot).
    void infer_accumulate(const NrcQuery* q, const int* count, int max_n, float4* radiance, hipStream_t s = nullptr);
    // One Adam step on n <= max_batch samples (targets rgb in .xyz).  grad_out (host, NRC_NPARAM)
    // when non-null: only the mean-loss gradient is computed and copied out, no update.
    // Returns the mean relative L2 loss when want_loss (synchronises), else 0.
    double train(const NrcQuery* q, const float4* target, int n, bool want_loss, float* grad_out = nullptr,
                 hipStream_t s = nullptr);

    // Timing on synthetic device-resident queries: ms per inference launch over n queries and, in
    // *train_ms, per training step of min(n, max_batch) samples (HIP events, `reps` launches each).
    // Trains the network: use a scratch cache.
    double bench(int n, int reps, double* train_ms);

private:
    int train_blocks(int n) const;
};

// Training targets of recorded path vertices: target = (L_path - L_before) / beta, per channel,
// where L_path = radiance[slot] at the end of the frame, L_before = lb[i] and beta = rec[i].d.
void nrc_targets(const NrcQuery* rec, const float4* lb, const float4* radiance, int n, float4* target,
                 hipStream_t s);

// Splits n records into `steps` batches of interleaved elements (batch k = records k, k + steps, ...),
// stored one after another in out / tout; batch k starts at k * (n / steps) + min(k, n % steps).
void nrc_interleave(const NrcQuery* in, const float4* tin, int n, int steps, NrcQuery* out, float4* tout,
                    hipStream_t s);

}  // namespace nr
