// Hash-grid parameter gradients by binning + LDS reduction (grid_backward.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "layout.h"
#include "nlod_common.h"

namespace nlod {

class GridBackward {
  public:
    // Plans the tiles of every grid table of the layout (host only; the plan is
    // copied to the device, synchronously, on the first run).
    explicit GridBackward(const Layout& L);
    ~GridBackward();
    // false: mixed feature counts or an oversized level; use the atomic kernel instead
    bool supported() const { return supported_; }
    // grads: the full parameter vector.  overwrite: every table entry is stored
    // (entries without contributions get 0); otherwise contributions are added.
    void run(hipStream_t s, const float* in, uint32_t N, const float* d_enc, float* grads, bool overwrite);
    // attaches the owner's epoch counter to every buffer (see DevBuf)
    void track(uint64_t* epoch) {
        for (DevBuf* b : {&d_rows_, &d_bins_, &counts_, &run_addr_, &totals_, &chunk_start_, &apply_list_, &meta_,
                          &records_, &partial_})
            b->epoch = epoch;
    }

  private:
    uint32_t queries_per_pass() const;
    uint64_t max_groups(uint32_t n) const;
    void prepare(hipStream_t s, uint32_t N);

    bool supported_ = false, uploaded_ = false;
    uint32_t F_ = 0, n_rows_ = 0, n_bins_ = 0, corners_per_query_ = 0, n_input_dims_ = 0;
    uint32_t count_rows_ = 0;  // sum over tiles of the repetitions reading the tile
    uint32_t max_reps_ = 0;    // largest number of segments sharing one table
    struct HostPlan;
    HostPlan* host_ = nullptr;
    DevBuf d_rows_, d_bins_;
    // per (tile, repetition, query block): run length (scanned in place into the
    // run's first record in the tile's order) and the run's first slot
    DevBuf counts_, run_addr_, totals_, chunk_start_, apply_list_, meta_, records_, partial_;
};

}  // namespace nlod
