// Engine: drives the HIP kernels for one validated Model on one GPU.
#pragma once

#include <atomic>
#include <string>
#include <vector>

#include "host_model.hpp"
#include "kernels.hpp"

namespace mtg {

struct RunControl {
    std::atomic<bool> stop{false};
    std::atomic<int> iter{0};
    std::atomic<int> niter{0};
    size_t memory_budget = 0;       // 0: derived from free device memory
    int batch_override = 0;         // 0: planner's choice
    int wave_pad = -1;              // STEM wave-batch row pitch = ny + pad; < 0: the default (0)
    int chains = 0;                 // STEM batch split into concurrent chains; 0: automatic
    bool use_graph = true;
    bool time_kernels = true;       // per-kernel HIP-event timing of one batch per configuration
    int autotune = 1;               // 0: default launch variants; 1: time the candidates once per
                                    // shape and keep the choice; 2: time them on every run
    int force_row = -1, force_col = -1, force_col_det = -1;   // >= 0: use this variant
};

// Launch variants chosen for one shape of the slice step.
struct PassChoice {
    int row = 0, col = 0, col_det = 0;
    std::string log;                // candidates with their times and the choice
};

struct Timing {
    double ms_total = 0;            // device time of the whole run (events on the engine stream)
    double ms_slice = 0;            // device time of the batched slice loops
    double probe_slices = 0;        // probe.slices covered by ms_slice
    double ms_row = 0, ms_col = 0;  // summed durations of timed row / column passes
    int n_row = 0, n_col = 0;       // launches timed
    double ms_tf = 0;               // transmission build (all configurations)
    std::string plan;
    std::string tune_log;           // the tuner's selection log for this run
};

RunControl &run_control();
const char *last_successful_call();   // name of the last HIP call that succeeded on this thread
Timing &last_timing();

// Full simulation into caller-owned output arrays; R = float runs the FP32
// wave path, R = double the FP64 one.
template <class R>
void run_multislice(const ModelT<R> &m, int device, mt_output_multislice *out);

// Several STEM specimens (grid sizes may differ) run together, one hipGraph
// launch per round of batches.
template <class R>
void run_multislice_group(const std::vector<const ModelT<R> *> &ms, int device,
                          const std::vector<mt_output_multislice *> &outs);

// Per-piece entries (natural, unshifted layouts).
template <class R>
void projected_potential(const ModelT<R> &m, int device, int iconf, int islice, double *V);
template <class R>
void projected_potential_sum(const ModelT<R> &m, int device, int iconf, double *V);
template <class R>
void transmission_function(const ModelT<R> &m, int device, int iconf, int islice, double *trans);
template <class R>
void incident_wave(const ModelT<R> &m, int device, double *psi0);
template <class R>
void propagate(const ModelT<R> &m, int device, const double *psi_in, double z, int space, double *psi_out);
// propagate() with all device memory in a static per-GPU scratch array
// (nx * ny at most propagate_noalloc_max_nxy()); no device allocation.
template <class R>
void propagate_noalloc(const ModelT<R> &m, int device, const double *psi_in, double z, int space, double *psi_out);
size_t propagate_noalloc_max_nxy();
// device blocks kept between runs: bytes kept on a device; wait for the
// device and free its kept blocks
size_t kept_device_bytes(int device);
void release_kept_blocks(int device);

// read + write bandwidth (GB/s) of a streaming copy of `bytes` bytes, reps times
double stream_copy_gbs(int device, size_t bytes, int reps);

} // namespace mtg
