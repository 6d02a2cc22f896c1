// Host-visible interface of the HIP kernels: argument blocks and launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "complex_types.hpp"

namespace mtg {

constexpr int kMaxDet = 8;
constexpr int kMaxTypes = 16;

// One wave in a batch: its probe position(s), its configuration's
// transmission stack, its scan index and its integration weight.
struct Member {
    int tslot;      // which transmission stack (configuration slot)
    int iscan;      // STEM scan index, -1 for padding members
    int beam0;      // first beam in the beam table
    int nbeam;      // number of beams (0: plane wave)
    double wq;      // weight of the wave's defocus / source point (1 without partial-coherence integration)
    int irot;       // tilt-azimuth table of the wave (precession / hollow cone); 0 otherwise
};

// One beam of a convergent incident wave: the position phase x*gx + y*gy
// with x = -2 pi (x_b - lx/2), and the defocus coefficient of chi,
// -pi lambda f with f the defocus referred to the first slice; FP64.
struct Beam {
    double xph, yph;
    double c10;
};

template <class R>
struct ProbeParamsT {
    double c10, c30, c50;         // -pi*lambda*f, -pi*Cs3*lambda^3/2, -pi*Cs5*lambda^5/3 (FP64);
                                  // the probe kernel takes the defocus term from each Beam
    double dgx, dgy;              // FP64 reciprocal sampling (1/lx, 1/ly)
    R dgx_f, dgy_f;               // sampling in R, for the aperture mask as the reference evaluates it
    // azimuthal aberrations (Lens::eval_m ... eval_c_52_c_54_c_56, src/types.cuh:3180-3300):
    // chi += m phi + sum_k az_c[k] g^az_p[k] sin(az_n[k] (phi - az_phi[k])), FP64
    int azim;                     // nonzero: the terms below are present
    double az_m;
    double az_c[11], az_phi[11];  // order 12, 21, 23, 32, 34, 41, 43, 45, 52, 54, 56
    double gxu, gyu;              // frequency shift of chi (tilted per-piece incident wave; 0 in simulations)
    R gxu_f, gyu_f;
    R g2_min, g2_max;             // aperture
    R norm;                       // 1/sqrt(number of aperture pixels)
};
using ProbeParams = ProbeParamsT<float>;

enum RowFlags : int {
    ROW_LOAD = 1,       // load the wave (else start from psi = 1, a plane wave)
    ROW_IFFT = 2,       // inverse transform along y first
    ROW_MUL_T = 4,      // multiply by the transmission function of slice `slice`
    ROW_FFT = 8,        // forward transform along y
    ROW_STORE = 16,
    ROW_MUL_RAMP = 32,  // multiply by ramp_x[row] * ramp_y[iy] (after the transmission)
};

template <class Z>
struct RowArgsT {
    const Z *src;
    Z *dst;
    long long member_stride;      // complex elements between members
    long long pitch;              // complex elements between rows of src / dst (0: the row length)
    int nrows;                    // rows per member (= nx)
    int n_members;
    int flags;
    const Z *tw;             // twiddles W_ny
    const Z *trans;          // transmission stacks: [tslot][n_slice][nx*ny]
    long long tslot_stride;       // complex elements between stacks
    int slice;
    const Member *members;
    int variant;                  // launch variant of the full slice step (kRowSlice flags)
    const Z *ramp_x, *ramp_y;   // ROW_MUL_RAMP tables [nrows], [row length]
    long long rot_x, rot_y;       // strides of per-azimuth ramp tables (members[].irot); 0: one table
};
using RowArgs = RowArgsT<float2>;

enum ColFlags : int {
    COL_LOAD = 1,
    COL_FFT = 2,
    COL_MUL_P = 4,      // multiply by px[ix]*py[iy] (Fresnel propagator incl. 1/(nx ny))
    COL_BWL = 8,        // multiply by the bandwidth-limit factor (and 1/(nx ny) when !COL_MUL_P)
    COL_IFFT = 16,
    COL_STORE = 32,
    COL_DET_BWL = 64,   // detector sums take one more bandwidth-limit factor (the back-propagation's)
    COL_COH = 128,      // detector passes: coh[iscan][ithk] += coh_w * psi(g) (coherent STEM image)
    COL_MUL_CTF = 256,  // multiply by ctf[ix*ny + iy] after the propagator (objective-lens image formation)
    COL_USER_WAVE = 512,  // probe passes: the wave is fpsi0(g) exp(i (x_b gx + y_b gy)) / (nx ny), first beam
    COL_ROT = 1024,     // COL_MUL_P tables per wave: px + members[].irot * rot_x, py + ... * rot_y
    COL_DET_FR = 2048,  // detector passes: matrix detectors det_fR instead of rings
};

template <class Z>
struct ColArgsT {
    using R = real_t<Z>;
    const Z *src;
    Z *dst;
    long long member_stride;
    long long pitch;              // element stride of x in src / dst (0: ncols)
    int ncols;                    // = ny (row length, element stride of x in the other tables)
    int n_members;
    int flags;
    const Z *tw;             // twiddles W_nx
    const Z *px;             // [nx]
    const Z *py;             // [ny]
    long long rot_x, rot_y;       // COL_ROT: strides of the per-azimuth propagator tables (members[].irot)
    R dgx_f, dgy_f;               // sampling in R for masks
    int nx, ny;
    R bwl_alpha, gl2_max;
    R inv_nxy;
    R p_scale;                // extra factor on px*py (1 for a slice step)
    // detector epilogue
    int n_det;
    R det_g2_min[kMaxDet], det_g2_max[kMaxDet];
    const R *det_fR;              // matrix detectors [n_det][nx*ny] natural k-space order (null: rings)
    double *det_part;             // [member][n_thk][n_det][ntile]
    int n_thk, ithk;
    const Z *ctf;            // COL_MUL_CTF table [nx*ny], natural order
    const Z *fpsi0;          // COL_USER_WAVE: FFT of the user's incident wave [nx*ny], natural order
    // coherent STEM accumulators [scan][n_thk][nx*ny], FP64
    double2 *coh;
    double coh_w;
    // probe generation
    const Member *members;
    const Beam *beams;
    ProbeParamsT<R> probe;
    int variant;                  // launch variant of the slice step / last slice
};
using ColArgs = ColArgsT<float2>;

template <class R>
struct TfArgsT {
    int nx, ny, nxh, nyh;
    R dRx, dRy;
    const vec4_t<R> *atoms;       // x, y, occ, type (bit pattern of an int in the low 32 bits of .w)
    const int2 *slice_range;      // per slice: first, last atom (inclusive); empty if first > last
    int slice0, n_slices;         // slices built by this launch
    const R *tables;              // [type][5][128]: R2, c0, c1, c2, c3
    const cplx_t<R> *rmax;        // [type]: R_max, R2_max
    int n_types;
    R w;                          // sigma factor (Vr_factor)
    int weak_phase;               // transmission 1 + i w V instead of exp(i w V)
    cplx_t<R> *trans;             // out, shifted layout: [slice][nx*ny] (relative to slice0), may be null
    R *V;                         // out, natural layout: [slice][nx*ny], may be null
    // sub-slicing: c0..c3 of each (slice, atom) pair, [pair][4][128], the
    // pairs of slice s starting at pair_base[s] in atom order; null otherwise
    const R *sub_coef;
    const int *pair_base;
};
using TfArgs = TfArgsT<float>;

// Sub-slice spline coefficients of (slice, atom) pairs: the slice's z range
// integrated through the atom's 3D potential at each of its element's 128
// knots in R^2, tapered, with cubic Hermite coefficients; computed in FP64
// and cast to R, as the element tables are built in double and cast.
template <class R>
struct SubArgsT {
    int n_pairs;
    const int2 *pairs;            // (slice, atom) of each pair
    const vec4_t<R> *atoms;       // x, y, occ, type of the configuration's atoms
    const R *z;                   // z of the configuration's atoms
    const R *slice_z;             // [slice][2]: z_0, z_e
    const R *tables;              // [type][5][128]: the knots R2 first
    const double *par;            // [type][12]: Vr terms a[5], k[5], R2_tap, tap_cf
    const double *qx, *qw;        // 128-node rule on [-1, 1]
    R *coef;                      // out: [pair][4][128]
};
template <class R>
hipError_t launch_subslice_coef(const SubArgsT<R> &a, hipStream_t s);

enum OutMode : int {
    OUT_SET_PSI = 1,       // psi_coh = O                     (single configuration)
    OUT_ADD_M2 = 2,        // m2psi_tot += w |O|^2
    OUT_ADD_PSI = 4,       // psi_coh += w O (cropped, shifted)
    OUT_ADD_PSI_FULL = 8,  // psi_full += w O (uncropped, unshifted; for m2psi_coh)
};

template <class Z>
struct OutArgsT {
    const Z *src;
    long long member_stride;
    int n_members;
    int nx, ny;
    int ix_0, iy_0, nx_o, ny_o;
    double w;
    const Member *members; // per-member weight factor wq (null: 1)
    int mode;
    double *m2;            // [nx_o*ny_o]
    double *psi;           // complex [nx_o*ny_o]
    double *psi_full;      // complex [nx*ny]
};
using OutArgs = OutArgsT<float2>;

// launchers (return hipError_t of the launch); Z = float2 or double2, R its real type
template <class Z> hipError_t launch_row_pass(int n, const RowArgsT<Z> &a, hipStream_t s);
template <class Z> hipError_t launch_col_pass(int n, const ColArgsT<Z> &a, bool probe, bool det, hipStream_t s);
template <class R> hipError_t launch_tf_build(const TfArgsT<R> &a, hipStream_t s);
// T = exp(i w V) (1 + i w V if weak_phase) of n_slices stored potentials V
// (natural layout), written fft-shifted; bit-identical to k_tf_build's T
template <class R>
hipError_t launch_trans_from_pot(const R *V, int n_slices, int nx, int ny, R w, int weak_phase, cplx_t<R> *trans,
                                 hipStream_t s);
hipError_t launch_det_reduce(const double *part, int n_members, int n_thk, int n_det, int ntile,
                             const Member *members, double w, double *image, int nscan, hipStream_t s);
template <class Z>
hipError_t launch_prop_tables(int nx, int ny, const double *z, int nz, double lambda, int mul_sign,
                              double lx, double ly, double gx0, double gy0, Z *px, Z *py, hipStream_t s);
template <class Z>
hipError_t launch_ramp_tables(int nx, int ny, double gx0, double gy0, double dRx, double dRy, double Rx_c,
                              double Ry_c, real_t<Z> scale, Z *rx, Z *ry, hipStream_t s);
template <class R>
hipError_t launch_count_aperture(int nx, int ny, R dgx, R dgy, R ux, R uy, R g2_min, R g2_max,
                                 unsigned int *count, hipStream_t s);
template <class Z> hipError_t launch_out_accum(const OutArgsT<Z> &a, hipStream_t s);
hipError_t launch_m2_from_psi(const double *psi_full, int nx, int ny, int ix_0, int iy_0, int nx_o, int ny_o,
                              double *m2, hipStream_t s);

// Objective-lens transfer function on the natural reciprocal grid
// (apply_CTF / apply_PCTF, src/cgpu_fcns.cuh:1993-2040): mode 1 coherent
// exp(i chi(g)), mode 2 partially coherent polar(E(g), chi_10,30(g));
// zero outside g2_min <= g^2 < g2_max.
template <class R>
struct CtfParamsT {
    int mode;
    double dgx, dgy;              // FP64 sampling for chi
    R dgx_f, dgy_f;               // sampling in R for the aperture test
    R g2_min, g2_max;
    double cc[5];                 // c_c_10, c_c_30, c_c_50 (indices 0..2); raw c_10 (3), c_30 (4) in A
    int azim;
    double az_m, az_c[11], az_phi[11];
    double ti, sth, lambda;       // ti_iehwgd, si_theta_c (after the temporal/spatial selection), lambda
};
template <class R>
hipError_t launch_ctf_table(int nx, int ny, const CtfParamsT<R> &p, cplx_t<R> *ctf, hipStream_t s);

// image_coh[ithk][det][iscan] = sum over the ring (or weighted by the matrix
// detector det_fR) of f |coh|^2 (f = bwl^2 if extra_bwl, else 1)
template <class R>
hipError_t launch_coh_image(const double2 *coh, int nscan, int n_thk, int n_det, int nx, int ny, R dgx, R dgy,
                            const R *det_g2_min, const R *det_g2_max, const R *det_fR, int extra_bwl, R bwl_alpha,
                            R gl2_max, double *image_coh, hipStream_t s);
hipError_t launch_stream_copy(const void *a, void *b, size_t bytes, hipStream_t s);
template <class R>
hipError_t launch_slice_sum(const R *V, int n_slices, long long nxy, double *out, hipStream_t s);
template <class Z> hipError_t launch_shift_complex(const Z *src, Z *dst, int nx, int ny, hipStream_t s);
// FP64 natural layout -> shifted wave (converted to Z), and back
template <class Z> hipError_t launch_load_shifted(const double2 *src, Z *dst, int nx, int ny, hipStream_t s);
template <class Z> hipError_t launch_store_unshifted(const Z *src, double2 *dst, int nx, int ny, hipStream_t s);

// Launch variants of the slice passes.  Variant 0 is the default; the tuner
// may pick another.  Variants differ in the radix schedule (different
// rounding), the workgroup tile (rows or columns per workgroup: identical
// results) and where twiddles live (registers or the cached table:
// identical results).
struct PassVariant {
    int sched;      // Sched<N, sched>
    int tile;       // rows (row pass) or columns (column pass) per workgroup
    int tw_table;   // 1: twiddles read from the table at each use
    int persist;    // 1: one workgroup per resident slot, looping over tiles and
                    //    loading the next tile while transforming the current one
};
constexpr int kPassVariants = 11;

// Transform lengths of the slice passes: 2^k, 3 x 2^k and 5 x 2^k in
// [64, 2048] (every one a multiple of 16).  Powers of two have every launch
// variant; the others run variant 0 only.
constexpr bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
constexpr bool fft_length_supported(int n)
{
    if (n < 64 || n > 2048) return false;
    const int m = n % 3 == 0 ? 3 : n % 5 == 0 ? 5 : 1;
    return is_pow2(n / m) && (n / m) % 16 == 0;
}
constexpr int pass_variants(int n) { return is_pow2(n) ? kPassVariants : 1; }

// elements per thread of schedule sc at length n (Sched<n, sc>::E)
constexpr int sched_elems(int n, int sc)
{
    if (n % 3 == 0) return sc == 0 ? (n <= 192 ? 12 : 24) : (n <= 192 ? 24 : 48);
    if (n % 5 == 0) return sc == 0 ? (n <= 320 ? 20 : 40) : 40;
    return sc == 0 ? (n >= 1024 ? 16 : 8) : (n >= 1024 ? 8 : 16);
}
constexpr int kLdsBytes = 160 * 1024;
// rows of length n per row-pass workgroup: 256 threads, at most 16 rows (so
// the count divides every supported length), LDS rows (n + n/8 elements of
// 16 bytes, either precision) within the 160 KiB of a CU
constexpr int row_tile0(int n)
{
    if (is_pow2(n)) return n <= 128 ? 16 : n == 256 ? 8 : n <= 1024 ? 4 : 2;
    int r = 256 / (n / sched_elems(n, 0));
    r = r > 16 ? 16 : r;
    while (r > 1 && r * (n + n / 8) * 16 > kLdsBytes) r /= 2;
    return r;
}
constexpr int col_tile0(int n) { return n > 1024 ? 8 : 16; }
// FP64 column pass: columns per workgroup that keep the double LDS tile
// (N x C x 16 bytes) at or below 128 KiB
constexpr int col_tile0_fp64(int n) { return n * 16 * 16 <= 128 * 1024 ? 16 : n * 8 * 16 <= 128 * 1024 ? 8 : 4; }

// Row pass, v: 0 default; 1 table twiddles; 2 twice the rows; 3 alternative
// schedule at the same thread count; 4 = 3 with table twiddles; 5 half the
// rows; 6 = 2 with table twiddles; 7 alternative schedule, twice the rows
// (capped at 1024 threads); 8, 9: variants 0 and 3 as persistent
// workgroups that prefetch the next tile; 10: variant 2 persistent.
constexpr PassVariant row_variant(int n, int v)
{
    const int r = row_tile0(n), alt = r * sched_elems(n, 1) / sched_elems(n, 0);
    const int t1 = n / sched_elems(n, 1);
    switch (v) {
    case 10: return PassVariant{0, 2 * r, 0, 1};
    case 9: return PassVariant{1, alt, 0, 1};
    case 8: return PassVariant{0, r, 0, 1};
    case 7: return PassVariant{1, 2 * alt * t1 > 1024 ? 1024 / t1 : 2 * alt, 0, 0};
    case 1: return PassVariant{0, r, 1, 0};
    case 2: return PassVariant{0, 2 * r, 0, 0};
    case 3: return PassVariant{1, alt, 0, 0};
    case 4: return PassVariant{1, alt, 1, 0};
    case 5: return PassVariant{0, r > 1 ? r / 2 : 1, 0, 0};
    case 6: return PassVariant{0, 2 * r, 1, 0};
    default: return PassVariant{0, r, is_pow2(n) ? 0 : 1, 0};   // 3 x 2^k, 5 x 2^k: table twiddles
    }
}
// Column pass, v: 0 default; 1 table twiddles; 2 half the columns;
// 3 alternative schedule (columns capped at 1024 threads per workgroup);
// 4 = 3 with table twiddles; 5 a quarter of the columns; 6 = 2 with table
// twiddles; 7 alternative schedule, twice the columns (capped at 1024
// threads: 256-byte row segments at 512^2); 8, 9: variants 0 and 3 as
// persistent workgroups that prefetch the next tile; 10: variant 2
// persistent.  Detector epilogues
// use only non-persistent variants with the default width, which fixes the
// layout of the detector partials.
constexpr PassVariant col_variant(int n, int v)
{
    const int c = col_tile0(n), t1 = n / sched_elems(n, 1);
    const int c1 = c * t1 > 1024 ? 1024 / t1 : c;
    switch (v) {
    case 10: return PassVariant{0, c / 2, 0, 1};
    case 9: return PassVariant{1, c1, 0, 1};
    case 8: return PassVariant{0, c, 0, 1};
    case 7: return PassVariant{1, 2 * c * t1 > 1024 ? 1024 / t1 : 2 * c, 0, 0};
    case 1: return PassVariant{0, c, 1, 0};
    case 2: return PassVariant{0, c / 2, 0, 0};
    case 3: return PassVariant{1, c1, 0, 0};
    case 4: return PassVariant{1, c1, 1, 0};
    case 5: return PassVariant{0, c / 4, 0, 0};
    case 6: return PassVariant{0, c / 2, 1, 0};
    default: return PassVariant{0, c, is_pow2(n) ? 0 : 1, 0};   // 3 x 2^k, 5 x 2^k: table twiddles
    }
}

// Whether the launch variants apply to a pass with these flags: only the
// specialised slice-step kernels (full row step, full column step, last
// column step with detector sums) have variants; any other flag set runs the
// generic kernel whatever the variant.
bool row_variants_apply(int flags);
bool col_variants_apply(int flags, bool det);

int col_tile_width(int nx);   // columns per column-pass workgroup (variant 0; the detector partial layout)
int col_tile_width_fp64(int nx);   // the same for the FP64 passes
int row_group(int ny);        // rows per row-pass workgroup (variant 0)

} // namespace mtg
