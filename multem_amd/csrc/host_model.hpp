// Host-side model of one simulation: the validated input (Input_Multislice<T>
// of the reference, T = float for the FP32 path, double for the FP64 path),
// the specimen, its slicing and the per-element spline tables.  Everything
// here runs once per call or once per phonon configuration; the per-probe
// work is on the GPU.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/multem_hip.h"

namespace mtg {

struct Error {
    mt_status code;
    std::string msg;
};

// Projected-potential spline of one element: 128 knots log-spaced in R,
// cubic Hermite in R^2 (Atomic_Data::get_atom_coef, src/atomic_data.hpp:1546-1668),
// built in double and cast to R.
template <class R>
struct AtomTableT {
    int Z = 0;
    double R_min = 0, R_max = 0, R_tap = 0, tap_cf = 0;   // as built (double)
    R R_max_f = 0, R2_max_f = 0;
    std::vector<R> R2, c0, c1, c2, c3;                     // 128 each
    // 3D potential Vr(r) = sum_i a_i exp(-k_i r) (2/(k_i r) + 1) (Lobato terms,
    // set_Lobato_neutral_coef, src/atomic_data.hpp:1440-1485): integrated over
    // each sub-slice's z range for potential_slicing = 3
    double vr_a[5] = {0, 0, 0, 0, 0}, vr_k[5] = {1, 1, 1, 1, 1};
};

template <class R>
AtomTableT<R> build_atom_table(int Z, double dR_min, double Vrl = 0.015, int nR = 128);

template <class R>
struct LensModelT {
    // MATLAB-unit fields converted as ilc_multem.cu:147-259 does
    int m = 0;
    R c_10 = 0, c_12 = 0, phi_12 = 0, c_21 = 0, phi_21 = 0, c_23 = 0, phi_23 = 0;
    R c_30 = 0, c_32 = 0, phi_32 = 0, c_34 = 0, phi_34 = 0;
    R c_41 = 0, phi_41 = 0, c_43 = 0, phi_43 = 0, c_45 = 0, phi_45 = 0;
    R c_50 = 0, c_52 = 0, phi_52 = 0, c_54 = 0, phi_54 = 0, c_56 = 0, phi_56 = 0;
    R inner_aper_ang = 0, outer_aper_ang = 0;
    int zero_defocus_type = 1;
    R zero_defocus_plane = 0;
    R lambda = 0;
    R g2_min = 0, g2_max = 0;      // aperture (in R, as the reference evaluates it)
    bool phi_required = false;
    R ti_iehwgd = 0, si_theta_c = 0;   // partial-coherence envelope (objective lens)
    // defocus and source spread functions (after Lens::set_input_data's clamps)
    R ti_a = 1, ti_sigma = 0, ti_beta = 0;
    int ti_npts = 1;
    R si_a = 1, si_sigma = 0, si_beta = 0;
    int si_rad_npts = 1, si_azm_npts = 1;
};

// One-dimensional (defocus) and two-dimensional (source) integration rules:
// offsets and weights
struct Quad1 {
    std::vector<double> x, w;
};
struct Quad2 {
    std::vector<double> x, y, w;
};

template <class R>
struct SliceInfoT {
    R z_0, z_e;
    int iatom_0, iatom_e;   // inclusive range into the z-sorted atoms; empty if iatom_0 > iatom_e
    int ithk;
    // z range whose atoms contribute to the slice: [z_0, z_e], widened by the
    // largest cut-off radius around the slice centre for potential_slicing = 3
    R z_int_0 = 0, z_int_e = 0;
};

template <class R>
struct ThickInfoT {
    R z;
    int islice;
    R z_back_prop;
    int iatom_e;    // last atom below the slice boundary that the thickness falls in
};

template <class R>
struct SpecimenT {
    // SoA, z-sorted (Spec::atoms_u); sigma and occupancy are float in the
    // reference's Atom_Data whatever T is
    std::vector<int> Z;
    std::vector<R> x, y, z;
    std::vector<float> sigma, occ;
    std::vector<int> region;
};

template <class R>
class ModelT {
public:
    using real = R;
    using LensModel = LensModelT<R>;
    using SliceInfo = SliceInfoT<R>;
    using ThickInfo = ThickInfoT<R>;
    using Specimen = SpecimenT<R>;
    using AtomTable = AtomTableT<R>;

    // Reads and validates; throws Error.
    ModelT(const mt_input_multislice &in);

    // simulation kind
    int sim_type = 0;
    bool is_stem() const { return sim_type == 11; }
    bool is_cbed() const { return sim_type == 21; }
    bool is_ewfs() const { return sim_type == 51; }
    bool is_ewrs() const { return sim_type == 52; }
    bool is_hrtem() const { return sim_type == 32; }
    bool is_istem() const { return sim_type == 12; }
    bool is_cbei() const { return sim_type == 22; }
    bool is_ed() const { return sim_type == 31; }
    bool is_ped() const { return sim_type == 41; }
    bool is_hctem() const { return sim_type == 42; }
    bool is_scanning() const { return is_stem() || is_istem(); }
    bool is_cbed_cbei() const { return is_cbed() || is_cbei(); }
    bool is_ped_hctem() const { return is_ped() || is_hctem(); }
    // objective-lens image of the exit wave (is_ISTEM_CBEI_HRTEM_HCTEM_EFTEM,
    // src/input_multislice.cuh:402-405)
    bool is_image() const { return is_istem() || is_cbei() || is_hrtem() || is_hctem(); }
    // outputs on the reciprocal grid (is_grid_FS, src/input_multislice.cuh:366-372)
    bool is_grid_fs() const { return is_cbed() || is_ed() || is_ped() || is_ewfs(); }
    // precession / hollow-cone rotations of the tilt azimuth (set_phi,
    // src/input_multislice.cuh:903-906)
    int nrot = 1;
    R rot_phi(int irot) const { return irot == 0 ? R(0) : (R(6.283185307179586476925) * (R)irot) / (R)nrot; }
    int illumination_model = 1;    // 1 coherent, 2 partially coherent (objective-lens envelope)
    int temporal_spatial_incoh = 1; // 1 both, 2 temporal only, 3 spatial only
    bool frozen() const { return pn_model == 3; }
    // 1 multislice, 2 phase object, 3 weak-phase object (after validation: a
    // thin specimen becomes a phase object, src/spec.hpp:65-77)
    int interaction_model = 1;
    bool multislice() const { return interaction_model == 1; }
    // illumination model 4: defocus (and, for CBED, source) spread integrated
    // numerically; the coherent STEM image keeps the coherent loop
    // (src/tem_simulation.cuh:115-137)
    bool full_integration() const { return illumination_model == 4 && !(is_stem() && pn_coh_contrib); }
    // cond_lens_temporal_spatial_quadratures (src/cpu_fcns.hpp:1539-1626)
    void cond_quadratures(Quad1 &qt, Quad2 &qs) const;
    // obj_lens_temporal_spatial_quadratures, defocus part (src/cpu_fcns.hpp:1503-1510)
    void obj_temporal_quadrature(Quad1 &qt) const;

    // electron / phonon
    int pn_model = 1, pn_coh_contrib = 0, pn_single_conf = 0, pn_nconf = 1, fp_iconf_0 = 1, pn_seed = 300183;
    bool pn_x = true, pn_y = true, pn_z = false;
    int mul_sign = 1;
    R E_0 = 300, lambda = 0, theta = 0, phi = 0;
    R gx0 = 0, gy0 = 0;         // tilt of the incident beam in reciprocal space (1/A)
    bool tilted() const { return gx0 != R(0) || gy0 != R(0); }
    // tilt of azimuth irot for precession / hollow cone (gx_0, gy_0 after
    // set_phi, src/input_multislice.cuh:903-927), else the fixed tilt
    R gx0_rot(int irot) const;
    R gy0_rot(int irot) const;
    // outputs multiplied by exp(-2 pi i g0.r) before the back-propagation;
    // PED keeps the pattern shifted (dp_Shift, src/wave_function.cuh:77-90)
    bool out_ramp() const { return tilted() && !is_ped(); }
    double lambda_d = 0;            // FP64 wavelength for phase arguments
    R Vr_factor = 0;            // mul_sign * gamma*lambda/(c_Potf*cos(theta)), evaluated in R
    int iw_type = 1;                // 1 plane, 2 convergent, 3 user defined
    std::vector<R> iw_x, iw_y;
    std::vector<double> iw_psi;     // iw_type 3: interleaved complex, natural layout ix*ny + iy

    // grid
    int nx = 0, ny = 0, nxh = 0, nyh = 0;
    R lx = 0, ly = 0, dRx = 0, dRy = 0, dgx = 0, dgy = 0;
    double lx_d = 0, ly_d = 0;      // box lengths as given, for FP64 phase arguments
    bool bwl = false;
    R gl2_max = 0, bwl_alpha = 0;

    LensModel cond, obj;

    // scanning
    int sc_nx = 0, sc_ny = 0;
    R sc_x0 = 0, sc_y0 = 0, sc_dRx = 0, sc_dRy = 0;
    std::vector<R> sc_x, sc_y;

    // detectors: g^2 bounds (pow(g, 2) in R), or matrix sensitivity maps
    // [n_det][nx*ny] in natural k-space order (detector_type 3)
    std::vector<R> det_g2_min, det_g2_max;
    std::vector<R> det_fR;

    // output area (0-based, half open)
    int oa_ix_0 = 0, oa_ix_e = 0, oa_iy_0 = 0, oa_iy_e = 0;

    // thicknesses requested (validated)
    std::vector<R> thick;

    // specimen
    Specimen atoms_u;                 // undisplaced, pbc-deduped, z-sorted
    std::vector<int> z_types;         // distinct atomic numbers, ascending
    std::vector<int> type_of_atom;    // index into z_types per atom of atoms_u
    std::vector<AtomTable> tables;    // per z_types entry
    std::vector<R> z_plane;
    std::vector<SliceInfo> slices;    // of the undisplaced specimen
    std::vector<ThickInfo> thk;

    int n_slice() const { return (int)slices.size(); }
    int n_conf_run() const { return pn_nconf - fp_iconf_0 + 1; }
    R phonon_weight() const;      // get_phonon_rot_weight (input_multislice.cuh:885-889)
    // propagation distance after slice is (none for the (weak) phase object)
    R dz_slice(int is) const { return multislice() ? (slices[is].z_e - slices[is].z_0) / std::cos(theta) : R(0); }
    R z_init() const { return R(0.5) * (slices[0].z_e + slices[0].z_0); }

    // Displaced positions of configuration iconf (Spec::move_atoms, src/spec.hpp:85-127);
    // returns x, y (z unchanged for plane slicing).
    void displaced(int iconf, std::vector<R> &x, std::vector<R> &y) const;
    // x, y and z of configuration iconf (z displaced only when sub-slicing with
    // pn_dim's z digit set), in atoms_u order
    void displaced(int iconf, std::vector<R> &x, std::vector<R> &y, std::vector<R> &z) const;

    // potential_slicing = 3 (dz_Sub) on a multislice run: each slice's
    // potential is the z integral over the slice of every atom within the
    // largest cut-off radius of it (set_atom_Vp with linear_Vz / cubic_poly_coef,
    // src/projected_potential.cuh:226-284, src/gpu_fcns.cuh:175-285)
    bool subslicing = false;
    // Configuration iconf's atoms for the sub-sliced potential: `order` lists
    // atoms_u indices sorted by the displaced z (stable), x/y/z are the
    // displaced positions in that order, and `sl` the slices with the atom
    // ranges of that order (Spec::move_atoms re-sorts and re-slices,
    // src/spec.hpp:119-126).  Slice boundaries are the undisplaced specimen's.
    void config_atoms(int iconf, std::vector<int> &order, std::vector<R> &x, std::vector<R> &y, std::vector<R> &z,
                      std::vector<SliceInfo> &sl) const;
    // tanh-sinh rule on [-1, 1] with 128 nodes (Quadrature type 0, src/quadrature.hpp:134-151):
    // limits in R, nodes and weights in double
    void subslice_rule(std::vector<double> &x, std::vector<double> &w) const;

    // output dims
    void output_dims(mt_output_multislice *out) const;
    void fill_axes(mt_output_multislice *out) const;

    R R_int_max = R(2.5);          // largest cut-off radius of the elements present

private:
    void slice_specimen();
    // fd_by_z over [z_0, z_e] of sorted z (src/slicing.hpp:428-452); Inc_Borders
    // widens the range by the precision's relative epsilon
    static void atoms_in(const std::vector<R> &z, R z_0, R z_e, bool inc, int &i0, int &ie);
};

using Model = ModelT<float>;      // FP32 path
using Model64 = ModelT<double>;   // FP64 path
using AtomTable = AtomTableT<float>;
using LensModel = LensModelT<float>;
using SliceInfo = SliceInfoT<float>;
using ThickInfo = ThickInfoT<float>;

} // namespace mtg
