// extern "C" entry points declared in include/multem_hip.h.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>
#include <string>

#include "../../include/multem_hip.h"
#include "engine.hpp"
#include "host_model.hpp"

namespace {

thread_local std::string g_last_error;
std::mutex g_engine_mutex;     // one engine call at a time per process

mt_status fail(mt_status c, const std::string &m)
{
    g_last_error = m;
    return c;
}

// precision of a call: 1 float, 2 double (system_conf.precision, eP_float /
// eP_double of the reference); prec_forced > 0 overrides it
int precision_of(const mt_system_conf *sys, int prec_forced)
{
    if (prec_forced > 0) return prec_forced;
    return sys->precision == 2 ? 2 : 1;
}

// Builds the model in the call's precision and runs f on it; f is generic in
// the model type (ModelT<float> or ModelT<double>).
template <class F>
mt_status guarded(const mt_system_conf *sys, const mt_input_multislice *in, F &&f, int prec_forced = 0)
{
    if (!sys || !in) return fail(MT_ERR_INVALID_ARGUMENT, "null argument");
    if (sys->device == 1) return fail(MT_ERR_UNSUPPORTED, "system_conf.device = 1 (CPU): this library runs on the GPU");
    std::lock_guard<std::mutex> lock(g_engine_mutex);
    try {
        if (precision_of(sys, prec_forced) == 2) {
            mtg::Model64 m(*in);
            f(m);
        } else {
            mtg::Model m(*in);
            f(m);
        }
        g_last_error.clear();
        return MT_OK;
    } catch (const mtg::Error &e) {
        return fail(e.code, e.msg);
    } catch (const std::bad_alloc &) {
        return fail(MT_ERR_OUT_OF_MEMORY, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(MT_ERR_DEVICE, std::string(e.what()) + " (after " + mtg::last_successful_call() + ")");
    }
}

template <class M>
void run_group_as(const mt_system_conf *sys, const mt_input_multislice *const *in, int n,
                         mt_output_multislice *out)
{
    std::vector<std::unique_ptr<M>> models;
    std::vector<const M *> mp;
    std::vector<mt_output_multislice *> op;
    for (int i = 0; i < n; ++i) {
        models.emplace_back(new M(*in[i]));
        mp.push_back(models.back().get());
        op.push_back(&out[i]);
    }
    mtg::run_multislice_group(mp, sys->gpu_device, op);
}

} // namespace

namespace mtg {
// error message of the calling thread, shared with the surrogate entries
void set_last_error(const std::string &msg) { g_last_error = msg; }
} // namespace mtg

static_assert(MT_PASS_VARIANTS == mtg::kPassVariants, "variant count of the C ABI");
static_assert(MT_NOALLOC_MAX_NXY == 1024 * 1024, "allocation-free propagation limit of the C ABI");

extern "C" {

int mt_hip_abi_version(void) { return MT_ABI_VERSION; }

const char *mt_hip_last_error(void) { return g_last_error.c_str(); }

mt_status mt_hip_output_dims(const mt_system_conf *sys, const mt_input_multislice *in, mt_output_multislice *out)
{
    if (!out) return fail(MT_ERR_INVALID_ARGUMENT, "null output");
    return guarded(sys, in, [&](const auto &m) { m.output_dims(out); });
}

static mt_status run_multislice_as(const mt_system_conf *sys, const mt_input_multislice *in,
                                   mt_output_multislice *out, int prec)
{
    if (!out) return fail(MT_ERR_INVALID_ARGUMENT, "null output");
    // progress restarts from (0, 0): a poll never sees the previous run's count
    mtg::run_control().stop = false;
    mtg::run_control().iter = 0;
    mtg::run_control().niter = 0;
    return guarded(sys, in, [&](const auto &m) { mtg::run_multislice(m, sys->gpu_device, out); }, prec);
}

mt_status mt_hip_run_multislice_f32(const mt_system_conf *sys, const mt_input_multislice *in, mt_output_multislice *out)
{
    return run_multislice_as(sys, in, out, 1);
}

mt_status mt_hip_run_multislice_f64(const mt_system_conf *sys, const mt_input_multislice *in, mt_output_multislice *out)
{
    return run_multislice_as(sys, in, out, 2);
}

mt_status mt_hip_run_multislice(const mt_system_conf *sys, const mt_input_multislice *in, mt_output_multislice *out)
{
    return run_multislice_as(sys, in, out, 0);
}

mt_status mt_hip_run_multislice_group(const mt_system_conf *sys, const mt_input_multislice *const *in, int n,
                                      mt_output_multislice *out)
{
    if (!sys || !in || !out || n < 1 || n > 16) return fail(MT_ERR_INVALID_ARGUMENT, "need 1..16 inputs and outputs");
    if (sys->device == 1) return fail(MT_ERR_UNSUPPORTED, "system_conf.device = 1 (CPU): this library runs on the GPU");
    for (int i = 0; i < n; ++i)
        if (!in[i]) return fail(MT_ERR_INVALID_ARGUMENT, "null input");
    mtg::run_control().stop = false;
    std::lock_guard<std::mutex> lock(g_engine_mutex);
    try {
        if (precision_of(sys, 0) == 2) run_group_as<mtg::Model64>(sys, in, n, out);
        else run_group_as<mtg::Model>(sys, in, n, out);
        g_last_error.clear();
        return MT_OK;
    } catch (const mtg::Error &e) {
        return fail(e.code, e.msg);
    } catch (const std::bad_alloc &) {
        return fail(MT_ERR_OUT_OF_MEMORY, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(MT_ERR_DEVICE, std::string(e.what()) + " (after " + mtg::last_successful_call() + ")");
    }
}

mt_status mt_hip_spec_slicing(const mt_system_conf *sys, const mt_input_multislice *in, int iconf, mt_slice *slices,
                              int max_slice, int *n_slice)
{
    if (!n_slice) return fail(MT_ERR_INVALID_ARGUMENT, "null n_slice");
    if (!sys || !in) return fail(MT_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lock(g_engine_mutex);
    auto fill = [&](const auto &m) {
        // the boundaries are the same for every configuration; with z
        // displacements (sub-slicing) the atom ranges are configuration iconf's,
        // into its z-sorted atoms
        using Rm = typename std::decay_t<decltype(m)>::real;
        auto sl = m.slices;
        if (m.subslicing && iconf >= 1) {
            std::vector<int> order;
            std::vector<Rm> x, y, z;
            m.config_atoms(iconf, order, x, y, z, sl);
        }
        *n_slice = (int)sl.size();
        for (int i = 0; slices && i < (int)sl.size() && i < max_slice; ++i) {
            const auto &s = sl[i];
            slices[i].z_0 = s.z_0; slices[i].z_e = s.z_e;
            slices[i].z_int_0 = s.z_int_0; slices[i].z_int_e = s.z_int_e;
            slices[i].iatom_0 = s.iatom_0; slices[i].iatom_e = s.iatom_e;
            slices[i].ithk = s.ithk;
        }
    };
    try {
        if (precision_of(sys, 0) == 2) fill(mtg::Model64(*in));
        else fill(mtg::Model(*in));
        g_last_error.clear();
        return MT_OK;
    } catch (const mtg::Error &e) {
        return fail(e.code, e.msg);
    }
}

mt_status mt_hip_projected_potential(const mt_system_conf *sys, const mt_input_multislice *in, int iconf, int islice,
                                     double *V)
{
    if (!V) return fail(MT_ERR_INVALID_ARGUMENT, "null V");
    return guarded(sys, in, [&](const auto &m) { mtg::projected_potential(m, sys->gpu_device, iconf, islice, V); });
}

mt_status mt_hip_projected_potential_sum(const mt_system_conf *sys, const mt_input_multislice *in, int iconf,
                                         double *V)
{
    if (!V) return fail(MT_ERR_INVALID_ARGUMENT, "null V");
    return guarded(sys, in, [&](const auto &m) { mtg::projected_potential_sum(m, sys->gpu_device, iconf, V); });
}

mt_status mt_hip_transmission_function(const mt_system_conf *sys, const mt_input_multislice *in, int iconf, int islice,
                                       double *trans)
{
    if (!trans) return fail(MT_ERR_INVALID_ARGUMENT, "null trans");
    return guarded(sys, in,
                   [&](const auto &m) { mtg::transmission_function(m, sys->gpu_device, iconf, islice, trans); });
}

mt_status mt_hip_incident_wave(const mt_system_conf *sys, const mt_input_multislice *in, double *psi0)
{
    if (!psi0) return fail(MT_ERR_INVALID_ARGUMENT, "null psi0");
    return guarded(sys, in, [&](const auto &m) { mtg::incident_wave(m, sys->gpu_device, psi0); });
}

mt_status mt_hip_propagate(const mt_system_conf *sys, const mt_input_multislice *in, const double *psi_in, double z,
                           int space, double *psi_out)
{
    if (!psi_in || !psi_out) return fail(MT_ERR_INVALID_ARGUMENT, "null wave");
    return guarded(sys, in,
                   [&](const auto &m) { mtg::propagate(m, sys->gpu_device, psi_in, z, space, psi_out); });
}

mt_status mt_hip_propagate_noalloc(const mt_system_conf *sys, const mt_input_multislice *in, const double *psi_in,
                                   double z, int space, double *psi_out)
{
    if (!psi_in || !psi_out) return fail(MT_ERR_INVALID_ARGUMENT, "null wave");
    return guarded(sys, in,
                   [&](const auto &m) { mtg::propagate_noalloc(m, sys->gpu_device, psi_in, z, space, psi_out); });
}

mt_status mt_hip_atom_table(int Z, double dR_min, double *R2, double *c0, double *c1, double *c2, double *c3,
                            double *scal)
{
    if (Z < 1 || Z > 103) return fail(MT_ERR_INVALID_ARGUMENT, "Z out of range 1..103");
    if (!(dR_min > 0)) return fail(MT_ERR_INVALID_ARGUMENT, "dR_min must be positive");
    mtg::AtomTable t = mtg::build_atom_table<float>(Z, (double)(float)dR_min, (double)0.015f);
    for (int i = 0; i < 128; ++i) {
        if (R2) R2[i] = t.R2[i];
        if (c0) c0[i] = t.c0[i];
        if (c1) c1[i] = t.c1[i];
        if (c2) c2[i] = t.c2[i];
        if (c3) c3[i] = t.c3[i];
    }
    if (scal) { scal[0] = t.R_min; scal[1] = t.R_max; scal[2] = t.R_tap; scal[3] = t.tap_cf; }
    g_last_error.clear();
    return MT_OK;
}

mt_status mt_hip_synchronize(int device)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(MT_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= n) return fail(MT_ERR_INVALID_ARGUMENT, "device out of range");
    if (hipSetDevice(device) != hipSuccess || hipDeviceSynchronize() != hipSuccess)
        return fail(MT_ERR_DEVICE, hipGetErrorString(hipGetLastError()));
    return MT_OK;
}

mt_status mt_hip_kept_device_bytes(int device, size_t *bytes)
{
    if (!bytes) return fail(MT_ERR_INVALID_ARGUMENT, "null argument");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(MT_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= n) return fail(MT_ERR_INVALID_ARGUMENT, "device out of range");
    std::lock_guard<std::mutex> lock(g_engine_mutex);
    *bytes = mtg::kept_device_bytes(device);
    return MT_OK;
}

mt_status mt_hip_release_device_memory(int device)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(MT_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= n) return fail(MT_ERR_INVALID_ARGUMENT, "device out of range");
    std::lock_guard<std::mutex> lock(g_engine_mutex);
    try {
        mtg::release_kept_blocks(device);
    } catch (const mtg::Error &e) {
        return fail(e.code, e.msg);
    }
    g_last_error.clear();
    return MT_OK;
}

void mt_hip_request_stop(void) { mtg::run_control().stop = true; }

void mt_hip_progress(int *iter, int *niter)
{
    if (iter) *iter = mtg::run_control().iter.load();
    if (niter) *niter = mtg::run_control().niter.load();
}

void mt_hip_set_memory_budget(size_t bytes) { mtg::run_control().memory_budget = bytes; }

mt_status mt_hip_last_plan(char *buf, size_t n)
{
    if (!buf || n == 0) return fail(MT_ERR_INVALID_ARGUMENT, "null buffer");
    std::snprintf(buf, n, "%s", mtg::last_timing().plan.c_str());
    return MT_OK;
}

mt_status mt_hip_last_timing(double *ms_slice_kernels, double *ms_total, double *probe_slices)
{
    const mtg::Timing &t = mtg::last_timing();
    if (ms_slice_kernels) *ms_slice_kernels = t.ms_slice;
    if (ms_total) *ms_total = t.ms_total;
    if (probe_slices) *probe_slices = t.probe_slices;
    return MT_OK;
}

// ---- engine tuning hooks (no counterpart in the reference) ----

// Batch width override (0: planner's choice); disables the hipGraph replay if use_graph == 0.
void mt_hip_set_batch(int batch, int use_graph)
{
    mtg::run_control().batch_override = batch;
    mtg::run_control().use_graph = use_graph != 0;
}

void mt_hip_set_wave_pad(int pad)
{
    mtg::run_control().wave_pad = pad < 0 ? -1 : pad;
}

void mt_hip_set_batch_chains(int chains)
{
    mtg::run_control().chains = chains < 0 ? 0 : chains;
}

// Slice-step launch variants: mode 0 fixed (variant 0 or the forced ones),
// 1 tuned once per shape and cached (default), 2 tuned on every run.
void mt_hip_set_autotune(int mode)
{
    mtg::run_control().autotune = mode < 0 ? 0 : (mode > 2 ? 2 : mode);
}

// Force launch variants of the row pass, the column pass and the column pass
// with detector sums (-1: leave to the tuner).  Out-of-range values are
// refused when the run launches them.
void mt_hip_set_pass_variants(int row, int col, int col_det)
{
    mtg::run_control().force_row = row;
    mtg::run_control().force_col = col;
    mtg::run_control().force_col_det = col_det;
}

// The tuner's selection log of the last run: per pass the candidates' mean
// launch times (us, '-' = not valid for the shape) and the variant chosen.
mt_status mt_hip_tune_log(char *buf, size_t n)
{
    if (!buf || n == 0) return fail(MT_ERR_INVALID_ARGUMENT, "null buffer");
    std::snprintf(buf, n, "%s", mtg::last_timing().tune_log.c_str());
    return MT_OK;
}

// Read + write bandwidth of a plain streaming copy kernel (16-byte accesses)
// over `bytes` bytes, `reps` launches, on GPU `device`: the reference for
// the achievable HBM bandwidth of the slice passes.
mt_status mt_hip_stream_copy_bandwidth(int device, size_t bytes, int reps, double *gbs)
{
    if (!gbs || bytes < 16 || bytes % 16 || reps < 1) return fail(MT_ERR_INVALID_ARGUMENT, "bytes (multiple of 16), reps >= 1");
    std::lock_guard<std::mutex> lock(g_engine_mutex);
    try {
        *gbs = mtg::stream_copy_gbs(device, bytes, reps);
        return MT_OK;
    } catch (const mtg::Error &e) {
        return fail(e.code, e.msg);
    } catch (const std::exception &e) {
        return fail(MT_ERR_DEVICE, e.what());
    }
}

// Per-kernel timing of the last run: summed row / column pass durations and launch counts.
void mt_hip_kernel_timing(double *ms_row, int *n_row, double *ms_col, int *n_col, double *ms_tf)
{
    const mtg::Timing &t = mtg::last_timing();
    if (ms_row) *ms_row = t.ms_row;
    if (n_row) *n_row = t.n_row;
    if (ms_col) *ms_col = t.ms_col;
    if (n_col) *n_col = t.n_col;
    if (ms_tf) *ms_tf = t.ms_tf;
}

} // extern "C"
