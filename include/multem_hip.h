/*
 * multem_hip.h -- C ABI of the MI355X multislice engine.
 *
 * The engine replaces the multislice path reached through MULTEM's
 * `ilc_multem` mex entry (mex_files_multem/ilc_multem.cu:520-548) and the C++
 * driver `mt::Multislice<T,dev>` (src/tem_simulation.cuh:51-97).  The input
 * structure carries the same fields, with the same names and units, as the
 * MATLAB struct that `read_input_multislice` (ilc_multem.cu:37-379) parses:
 * angles in mrad or degrees, Cs in mm, lengths in Angstrom, 1-based output
 * area.  The engine applies the reader's unit conversions and
 * `Input_Multislice::validate_parameters` (src/input_multislice.cuh:622-796).
 *
 * Arrays are column-major as in MATLAB: a ny x nx image stores pixel (ix, iy)
 * at ix*ny + iy.  Complex arrays are interleaved (re, im).  All outputs are
 * caller-owned and written in double precision (the mex layer hands MATLAB
 * doubles; ilc_multem.cu:381-481).
 *
 * Threading: one host thread per engine call.  Calls on the same GPU are
 * serialised by the library.  Errors are returned as mt_status codes; the
 * message of the last error of the calling thread is read with
 * mt_hip_last_error().
 */
#ifndef MULTEM_HIP_H
#define MULTEM_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MT_ABI_VERSION 4

typedef enum {
    MT_OK = 0,
    MT_ERR_INVALID_ARGUMENT = 1,   /* an input field is out of range */
    MT_ERR_UNSUPPORTED = 2,        /* a valid MULTEM option this engine does not provide */
    MT_ERR_OUT_OF_MEMORY = 3,      /* device or host allocation failed / budget too small */
    MT_ERR_DEVICE = 4,             /* HIP runtime error */
    MT_ERR_STOPPED = 5,            /* mt_hip_request_stop() ended the run early */
    MT_ERR_NO_DEVICE = 6           /* no GPU visible */
} mt_status;

/* system_conf (src/types.cuh:4440-4512; matlab_mex.cuh:263-288) */
typedef struct {
    int precision;     /* 1: float, 2: double */
    int device;        /* 1: CPU, 2: GPU */
    int cpu_nthread;
    int gpu_device;
} mt_system_conf;

/* Lens fields of the input struct, cond_lens_* / obj_lens_* (ilc_multem.cu:147-259).
   Units as in MATLAB: c_30, c_50 in mm; phi_* in degrees; apertures in mrad. */
typedef struct {
    double m;
    double c_10, c_12, phi_12;
    double c_21, phi_21, c_23, phi_23;
    double c_30, c_32, phi_32, c_34, phi_34;
    double c_41, phi_41, c_43, phi_43, c_45, phi_45;
    double c_50, c_52, phi_52, c_54, phi_54, c_56, phi_56;
    double inner_aper_ang, outer_aper_ang;
    double ti_a, ti_sigma, ti_beta;
    int ti_npts;
    double si_a, si_sigma, si_beta;
    int si_rad_npts, si_azm_npts;
    int zero_defocus_type;        /* 1 first, 2 middle, 3 last, 4 user defined */
    double zero_defocus_plane;
} mt_lens;

/* detector.cir(i) (ilc_multem.cu:272-290), angles in mrad */
typedef struct {
    double inner_ang;
    double outer_ang;
} mt_detector_cir;

/* The ilc_multem input struct (MATLAB multem_input.parameters, parameters.m:6-239). */
typedef struct {
    int simulation_type;       /* 11 STEM, 12 ISTEM, 21 CBED, 22 CBEI, 31 ED, 32 HRTEM, 41 PED, 42 HCTEM,
                                  51 EWFS, 52 EWRS (types.cuh:193-205) */
    int interaction_model;     /* 1 multislice, 2 phase object, 3 weak phase object */
    int potential_type;        /* 6 Lobato (the only parameterisation provided) */
    int operation_mode;
    int reverse_multislice;

    int pn_model;              /* 1 still atom, 2 absorptive, 3 frozen phonon */
    int pn_coh_contrib;
    int pn_single_conf;
    int pn_nconf;
    int pn_dim;                /* 110, 111, ... */
    int pn_seed;

    const double *spec_atoms;  /* column-major n_atoms x spec_atoms_cols: Z x y z sigma occ region charge */
    int spec_atoms_rows;
    int spec_atoms_cols;
    double spec_lx, spec_ly, spec_lz, spec_dz;

    double spec_rot_theta;     /* degrees; axis and centre at the end of the struct */

    int thick_type;            /* 1 whole specimen, 2 through thickness, 3 through slices */
    const double *thick;
    int n_thick;

    int potential_slicing;     /* 1 planes, 2 dz_Proj, 3 dz_Sub (4 auto: refused) */

    int nx, ny;                /* 2^k, 3 x 2^k or 5 x 2^k in [64, 2048] (others: MT_ERR_UNSUPPORTED) */
    int bwl;

    int iw_type;               /* 1 plane, 2 convergent, 3 user defined, 4 auto */
    const double *iw_psi;      /* interleaved complex ny x nx, iw_type 3 only */
    const double *iw_x;
    const double *iw_y;
    int n_iw_xy;

    double E_0;                /* keV */
    double theta, phi;         /* degrees */

    int illumination_model;    /* 1 coherent, 2 partial coherent, 3 TCC, 4 full integration */
    int temporal_spatial_incoh;

    mt_lens cond_lens;
    mt_lens obj_lens;

    int scanning_type;         /* 1 line, 2 area */
    int scanning_periodic;
    int scanning_square_pxs;
    int scanning_ns;
    double scanning_x0, scanning_y0, scanning_xe, scanning_ye;

    int detector_type;         /* 1 circular, 3 matrix (2, radial, integrates to zero in the reference: refused) */
    int n_detector;
    const mt_detector_cir *detector_cir;

    int output_area_ix_0, output_area_iy_0, output_area_ix_e, output_area_iy_e;  /* 1-based as in MATLAB */

    /* specimen rotation (ilc_multem.cu:98-102; applied by validate_parameters,
       src/input_multislice.cuh:707-722) by spec_rot_theta (degrees, above)
       about the axis spec_rot_u0 through the centre point: the atoms'
       geometric centre (spec_rot_center_type 1) or spec_rot_center_p (2) */
    double spec_rot_u0[3];
    int spec_rot_center_type;
    double spec_rot_center_p[3];

    /* matrix detectors (detector_type 3, detector.matrix(i).fR;
       ilc_multem.cu:317-333): n_detector sensitivity maps, each ny x nx,
       contiguous, centred (g = 0 at column nx/2, row ny/2, as the
       reference's examples build them); the engine applies fft2_shift as
       Wave_Function::set_input_data does (src/wave_function.cuh:53-61) */
    const double *detector_fR;

    /* precession electron diffraction (simulation_type 41) and hollow-cone
       illumination TEM (42): cone angle in degrees and number of azimuths
       (ped_theta, ped_nrot, hci_theta, hci_nrot; ilc_multem.cu:338-347) */
    double ped_theta;
    int ped_nrot;
    double hci_theta;
    int hci_nrot;
} mt_input_multislice;

/* Output kinds (eTEM_Output_Type, types.cuh:209-214) */
enum {
    MT_OT_IMAGE_TOT_COH = 1, MT_OT_IMAGE_TOT = 2,
    MT_OT_M2PSI_TOT_COH = 3, MT_OT_M2PSI_TOT = 4,
    MT_OT_M2PSI_TOT_PSI_COH = 5, MT_OT_PSI_COH = 6,
    MT_OT_PSI_0 = 7, MT_OT_V = 8, MT_OT_TRANS = 9
};

/* Output_Multislice (src/output_multislice.hpp:1165-1199), flattened.
   mt_hip_output_dims() fills the scalar fields; the caller then provides
   each array that output_type needs (others may be NULL):
     image_tot / image_coh : [n_thk][n_det][ny*nx]      (STEM, scan grid)
     m2psi_tot / m2psi_coh : [n_thk][ny*nx]
     psi_coh               : [n_thk][ny*nx] complex
   x (nx), y (ny) and thick (n_thk) are optional. */
typedef struct {
    int output_type;
    int nx, ny;
    int n_thk;
    int n_det;
    double dx, dy;
    double *x, *y;
    double *thick;
    double *image_tot;
    double *image_coh;
    double *m2psi_tot;
    double *m2psi_coh;
    double *psi_coh;
} mt_output_multislice;

/* One slice of the specimen as found by Slicing (src/slicing.hpp:339-403) */
typedef struct {
    double z_0, z_e;
    double z_int_0, z_int_e;
    int iatom_0, iatom_e;
    int ithk;
} mt_slice;

/* ---- library --------------------------------------------------------- */

int mt_hip_abi_version(void);

/* Message of the last failed call made by this thread ("" if none). */
const char *mt_hip_last_error(void);

/* Fills out->output_type and the sizes.  No device work.
   Replaces the sizing half of Output_Multislice::set_input_data
   (output_multislice.hpp:416-632). */
mt_status mt_hip_output_dims(const mt_system_conf *sys, const mt_input_multislice *in,
                             mt_output_multislice *out);

/* Full simulation.  Replaces run_multislice<T, e_device>
   (ilc_multem.cu:483-518) = Multislice::set_input_data + operator()
   (tem_simulation.cuh:51-97) + Output_Multislice::gather.
   sys->precision selects the _f32 / _f64 variant; the _f32 / _f64 entries
   force it. */
mt_status mt_hip_run_multislice(const mt_system_conf *sys, const mt_input_multislice *in,
                                mt_output_multislice *out);
mt_status mt_hip_run_multislice_f32(const mt_system_conf *sys, const mt_input_multislice *in,
                                    mt_output_multislice *out);
/* run_multislice<double, e_device> (ilc_multem.cu:483-518, chosen by
   system_conf.is_double_device() at ilc_multem.cu:520-546): the wave,
   transmission stack, propagator and FFTs in FP64. */
mt_status mt_hip_run_multislice_f64(const mt_system_conf *sys, const mt_input_multislice *in,
                                    mt_output_multislice *out);

/* Several STEM inputs run together: grid sizes, slice counts, scans and
   configuration counts may differ.  Each round advances every input with
   work left by one batch of probes through all its slices, and the round's
   device work (one branch per input, each with its own tuned kernels) is a
   single hipGraph launch.  out[i] is filled as mt_hip_run_multislice would
   fill it for in[i] alone (the caller provides the arrays that
   mt_hip_output_dims reports).  No counterpart in the reference, which runs
   one specimen per call (Multislice::operator(), src/tem_simulation.cuh:51-97). */
mt_status mt_hip_run_multislice_group(const mt_system_conf *sys, const mt_input_multislice *const *in, int n,
                                      mt_output_multislice *out);

/* ---- per-piece entries (same input struct) --------------------------- */

/* Slices of the specimen.  Replaces ilc_spec_slicing
   (mex_files_multem/ilc_spec_slicing.cpp:98-135).  *n_slice receives the
   count; slices (capacity max_slice, may be NULL) receives the slices.  The
   boundaries are the undisplaced specimen's; with potential_slicing = 3 and
   z displacements (pn_dim 111, 101, 011, 001) the atom ranges are those of
   configuration iconf (>= 1) into its z-sorted atoms, else iconf is
   ignored.  z_int_0 .. z_int_e is the z range whose atoms contribute to the
   slice (wider than z_0 .. z_e with potential_slicing = 3). */
mt_status mt_hip_spec_slicing(const mt_system_conf *sys, const mt_input_multislice *in,
                              int iconf, mt_slice *slices, int max_slice, int *n_slice);

/* Projected potential of slice islice (0-based) for phonon configuration
   iconf (1-based, ignored for still atoms), ny x nx real, natural layout.
   Replaces ilc_projected_potential (ilc_projected_potential.cu:136-158). */
mt_status mt_hip_projected_potential(const mt_system_conf *sys, const mt_input_multislice *in,
                                     int iconf, int islice, double *V);

/* Sum over all slices of the projected potential of configuration iconf
   (natural layout, nx*ny doubles, accumulated in FP64 in slice order): the
   potential map input of the surrogate network.  No counterpart in the
   reference (sum of ilc_projected_potential over islice). */
mt_status mt_hip_projected_potential_sum(const mt_system_conf *sys, const mt_input_multislice *in, int iconf,
                                         double *V);

/* Transmission function exp(i sigma V) of slice islice, ny x nx complex,
   natural layout.  Replaces ilc_transmission_function
   (ilc_transmission_function.cu:142-168). */
mt_status mt_hip_transmission_function(const mt_system_conf *sys, const mt_input_multislice *in,
                                       int iconf, int islice, double *trans);

/* Incident wave at (iw_x, iw_y) in real space, ny x nx complex, natural
   layout.  Replaces ilc_incident_wave (ilc_incident_wave.cu:144-169). */
mt_status mt_hip_incident_wave(const mt_system_conf *sys, const mt_input_multislice *in,
                               double *psi0);

/* Fresnel propagation of psi_in (ny x nx complex, natural layout) by z
   (Angstrom); space 1 real, 2 reciprocal output.  Replaces ilc_propagate
   (ilc_propagate.cu:115-140). */
mt_status mt_hip_propagate(const mt_system_conf *sys, const mt_input_multislice *in,
                           const double *psi_in, double z, int space, double *psi_out);

/* mt_hip_propagate without device allocation: the same arguments, meaning
   and result, bit for bit, with every device buffer of the call (staging of
   psi_in / psi_out, the wave, twiddle and propagator tables) inside one
   static device array of MT_NOALLOC_SCRATCH_BYTES per GPU, held from the
   first call on that GPU until the process ends.  nx * ny must not exceed
   MT_NOALLOC_MAX_NXY (1024^2), else MT_ERR_INVALID_ARGUMENT.  Work runs on
   the null stream.  Calls on one GPU share the array; this library
   serialises its own calls, and calls that overlap on one GPU from outside
   it would race.  No counterpart in the reference: for callers that cannot
   allocate device memory (code under HIP graph capture, tight serving
   loops, harnesses that forbid hipMalloc). */
#define MT_NOALLOC_MAX_NXY (1024 * 1024)
#define MT_NOALLOC_SCRATCH_BYTES (2u * 16u * MT_NOALLOC_MAX_NXY + 262144u)
mt_status mt_hip_propagate_noalloc(const mt_system_conf *sys, const mt_input_multislice *in,
                                   const double *psi_in, double z, int space, double *psi_out);

/* Projected-potential spline of element Z for pixel size dR_min (Angstrom),
   as the engine builds it: 128 knots each of R^2 and the cubic coefficients
   c0..c3, and scal = {R_min, R_max, R_tap, tap_cf}.  Host only.  Replaces
   Atomic_Data::To_atom_type_CPU (src/atomic_data.hpp:148-175) as exposed by
   the mex helpers. */
mt_status mt_hip_atom_table(int Z, double dR_min, double *R2, double *c0, double *c1, double *c2,
                            double *c3, double *scal);

/* ---- run control (Multislice::ext_stop_sim / ext_niter / ext_iter,
        tem_simulation.cuh:47-49) ---------------------------------------- */
void mt_hip_request_stop(void);
void mt_hip_progress(int *iter, int *niter);

/* Caps the device memory the engine plans for (bytes; 0 = 85 % of the free
   device memory).  The counterpart of Memory_Slice's free-memory probe
   (transmission_function.cuh:146-215): when one configuration's
   transmission stack does not fit beside two waves, the engine stores the
   potentials of as many slices as fit and rebuilds each slice's
   transmission -- from its potential, or from the atoms beyond those --
   just before it is used; results are bit-identical to the resident stack.
   Below one slice and one wave a run fails with MT_ERR_OUT_OF_MEMORY.  The
   last plan (mt_hip_last_plan) names the mode ("slices=rebuilt
   stored_potentials=k/n"). */
void mt_hip_set_memory_budget(size_t bytes);

/* Engine choices made for the last run (batch width, FFT schedule, tiles),
   as a NUL-terminated text line written into buf (capacity n). */
mt_status mt_hip_last_plan(char *buf, size_t n);

/* Device time (ms) of the slice-step kernels of the last run, measured with
   HIP events on the engine's stream, and the number of probe.slices they
   covered. */
mt_status mt_hip_last_timing(double *ms_slice_kernels, double *ms_total, double *probe_slices);

/* Blocks until all work the library queued on GPU `device` has finished
   (hipDeviceSynchronize on the library's HIP runtime).  No counterpart in
   the reference, whose mex entry synchronises its stream before returning. */
mt_status mt_hip_synchronize(int device);

/* Device memory kept between runs.  Device blocks a run releases are kept
   for the next run that asks for blocks of the same sizes (a simulation
   called again allocates the same buffers), and are counted as free memory
   by the memory planner; when an allocation runs out of memory the device's
   kept blocks are freed first.  mt_hip_kept_device_bytes reports the bytes
   kept on `device`; mt_hip_release_device_memory waits for the device and
   returns its kept blocks to the HIP runtime (for callers that share the GPU
   with other work between runs).  No counterpart in the reference, which
   allocates and frees its buffers in every call. */
mt_status mt_hip_kept_device_bytes(int device, size_t *bytes);
mt_status mt_hip_release_device_memory(int device);

/* ---- engine tuning hooks (no counterpart in the reference; for the
        benchmark and tests) ------------------------------------------------ */

/* Batch width (waves per batch; 0 = planner's choice) and whether the slice
   loop of a batch is replayed from a captured hipGraph. */
void mt_hip_set_batch(int batch, int use_graph);

/* Row pitch of the STEM wave batches: ny + pad complex elements (pad < 0:
   the default, 0).  A padded pitch moves the column pass's strided row
   segments to other HBM channels; results do not depend on it.  No
   counterpart in the reference (a layout choice of this engine). */
void mt_hip_set_wave_pad(int pad);

/* Concurrent chains of a STEM batch: the waves of a batch are split into
   that many groups, each running its slice loop on its own stream inside
   the batch's graph, so one group's passes fill the CUs while the other's
   drain (0: automatic, two chains for grids of 512^2 and more whose batch
   holds at least 128 MiB of waves, with the transmission stack resident;
   1: one chain; at most 4).  Results do not depend on it.  No counterpart
   in the reference. */
void mt_hip_set_batch_chains(int chains);

/* Per-kernel device time of the last run, from HIP events around the row and
   column passes of the first batch of each configuration: summed ms and
   launch counts; ms_tf is the transmission-function build time. */
void mt_hip_kernel_timing(double *ms_row, int *n_row, double *ms_col, int *n_col, double *ms_tf);

/* Read + write bandwidth (GB/s) of a streaming copy kernel (one 16-byte
   element per thread, nontemporal loads and stores) over `bytes` bytes
   (a multiple of 16), `reps` launches: the achievable-bandwidth reference
   for the slice passes on this GPU. */
mt_status mt_hip_stream_copy_bandwidth(int device, size_t bytes, int reps, double *gbs);

/* Autotuner of the slice step.  Each pass has MT_PASS_VARIANTS launch
   variants (radix schedule, rows/columns per workgroup, twiddles in registers
   or read from the cached table).  Variants with the default schedule give
   bit-identical waves; the alternative schedule rounds differently (FP32
   FFT rounding only).  mode 0: fixed variants (0, or the forced ones);
   1: each candidate is timed once per shape (grid, batch, detector
   epilogue) and the choice kept for the process (default); 2: timed on
   every run.  The choice is the lowest-numbered variant within 2 % of the
   fastest. */
#define MT_PASS_VARIANTS 11
void mt_hip_set_autotune(int mode);

/* Forces the variants of the row pass, the column pass and the last-slice
   column pass with detector sums (-1: tuner's choice).  A run whose forced
   variant is invalid for its grid fails with MT_ERR_INVALID_ARGUMENT. */
void mt_hip_set_pass_variants(int row, int col, int col_det);

/* Selection log of the last run: per pass the mean launch time of each
   candidate in us ('-': not valid for the grid) and the variant chosen. */
mt_status mt_hip_tune_log(char *buf, size_t n);

#ifdef __cplusplus
}
#endif

#endif /* MULTEM_HIP_H */
