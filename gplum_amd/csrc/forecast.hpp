// forecast.hpp -- learned sizing of the EP-EP candidate buffer (per i-group segments).
//
// Per step, before the FP32 soft-force launch: features of every particle -> the
// forecaster network's expected candidate count -> a capacity per i-group -> segment
// offsets (clipped to the buffer that is allocated).  The force kernel fills the
// segments (LDS counter per group, no global atomic); groups whose exact count exceeds
// their segment are redone by a candidate-only pass into an exactly sized area, so
// the candidate set (and everything after it) never depends on the forecast.
// After the exact search the network takes a few Adam steps on (features, exact FP32
// candidate count) pairs sampled from this step.
#pragma once
#include "mlp.hpp"
#include "soft.hpp"

namespace gplum {

constexpr int FC_SORT_MAX = 2048;     // a group's segment holds at most this many candidates
constexpr int FC_SORT_BLOCK = 256;    // (sorted in LDS by one block)

struct FcNorm {
    float mu[MLP_NIN], inv_sd[MLP_NIN];
};

struct CandSizer {
    Forecaster net;
    FcNorm norm{{0.f, 0.f, 0.f, 0.f}, {1.f, 1.f, 1.f, 1.f}};
    bool enabled = false, normalized = false;
    bool uniform = false;             // ablation: every particle gets uniform_rate
    float uniform_rate = 0.f;         //   (the previous call's mean count)
    int train_steps = 2;              // Adam steps per soft step
    int batch = 8192;                 // samples per Adam step
    int trained_rounds = 0;           // soft steps with training so far
    int min_rounds = 1;               // forecast used once this many rounds trained
    unsigned phase = 0;
    double last_loss = 0.;
    // a fresh forecast (features, network, one training round) every predict_every soft
    // forces; in between the plan reuses the rates, which travel with the particles
    // (ParticleSoA::fc_rate: gathered with every reorder and compaction)
    int predict_every = 4;
    int since_predict = 1 << 30;
    bool lam_valid = false;
    int last_batch = 0;               // samples of the last Adam step (its loss partials stay)

    DArray<float> x, lam, xs, ys;     // features [4][n], rates, training batch
    DArray<int32_t> seg_off, seg_cap, seg_cnt, fin_off, ovf_list, ovf_pos;
    DArray<uint64_t> ovf_keys;        // the overflowed groups' candidates, sorted
    DArray<int64_t> info;             // requested, stored, n_ovf, ovf_total
    DArray<int2> pairs_seg;           // segment storage
    DArray<SoftGroup> ovf_groups;
    DArray<unsigned int> ovf_count;
    size_t seg_capacity = 0;          // slots of pairs_seg

    bool active() const { return enabled && (uniform || trained_rounds >= min_rounds); }
};

struct FcGroupView {                  // tree data the features read
    const int32_t *gstart, *gcount, *gnode, *ep_cnt;
    const double *bmin, *bmax, *rs;
};

// raw or standardised features into s.x (feature-major)
int fc_features(CandSizer &s, const FcGroupView &gv, int ng, const double *pos, const double *mass,
                const double *vdisp, double m_sun, int n, hipStream_t st);
// forecast (fresh: from s.x into rate, the particles' rate array; else rate as the
// reorders left it) -> per-group capacity -> offsets
// (clipped to s.seg_capacity); returns the CandBuf for the segment-mode launch
int fc_plan(CandSizer &s, const FcGroupView &gv, int ng, int n, float *rate, bool fresh, CandBuf *cb,
            hipStream_t st);
// after the segment-mode launch: exact counts -> final offsets of every group's
// candidates (group order) and the list of overflowed groups (on the device)
int fc_summary(CandSizer &s, int ng, hipStream_t st);
// out = {requested slots, total, n_ovf, ovf_total} of the last summary (synchronises)
int fc_read_info(CandSizer &s, int64_t out[4], hipStream_t st);
// the candidates of the groups that fit their segments as sorted keys i 2^kb + j at
// their final offsets in keys; groups whose keys would pass keys_cap are left out
int fc_sort_keys(CandSizer &s, int ng, int kb, uint64_t *keys, long long keys_cap, hipStream_t st);
// the overflowed groups' candidates (sorted[0..ovf_total), ascending) to their offsets
int fc_scatter_overflow(CandSizer &s, int n_ovf, const uint64_t *sorted, uint64_t *keys, hipStream_t st);
// s.ovf_groups[0..n_ovf) = the overflowed groups, for the candidate-only pass
int fc_overflow_groups(CandSizer &s, const SoftGroup *groups, int n_ovf, hipStream_t st);
// standardise (first call) and train on this step's exact FP32 candidate counts;
// want_loss: also evaluate the last step's loss into s.last_loss (synchronises st)
int fc_train(CandSizer &s, const int32_t *count_fp32, int n, hipStream_t st, bool want_loss);
// s.last_loss from the loss partials of the last Adam step (synchronises st)
int fc_read_loss(CandSizer &s, hipStream_t st);

}  // namespace gplum
