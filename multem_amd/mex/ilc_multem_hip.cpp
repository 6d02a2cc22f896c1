// MATLAB gateway with the calling convention of MULTEM's ilc_multem
// (mex_files_multem/ilc_multem.cu:520-548):
//
//     output = ilc_multem_hip(system_conf, input_multislice)
//     output = ilc_multem_hip(input_multislice)
//
// It reads the same struct fields as read_input_multislice
// (ilc_multem.cu:37-379), runs the HIP engine through the C ABI of
// libmultem_hip.so and returns the struct that set_struct_multislice builds
// (ilc_multem.cu:381-481): dx, dy, x, y, thick and data(ithk) with
// image_tot(idet).image / image_coh (STEM), m2psi_tot / psi_coh (EWFS,
// EWRS) or m2psi_tot / m2psi_coh (CBED, HRTEM).
//
// Build (MATLAB R2017b or later, separate complex storage):
//     mex -R2017b ilc_multem_hip.cpp -I<repo>/include -L<repo>/multem_amd/lib -lmultem_hip
//
// Differences a MATLAB user sees: errors raise (mexErrMsgIdAndTxt) instead
// of printing; system_conf.device = 1 (CPU) is refused; a missing field
// reads as 0, as mx_get_scalar_field does (src/matlab_mex.cuh:118-131).
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "mex.h"
#include "multem_hip.h"

namespace {

double num(const mxArray *s, const char *name, mwIndex idx = 0)
{
    const mxArray *f = mxGetField(s, idx, name);
    return (f && !mxIsEmpty(f)) ? mxGetScalar(f) : 0.0;
}

// numeric field as a dense double vector (empty if absent)
std::vector<double> vec(const mxArray *s, const char *name)
{
    const mxArray *f = mxGetField(s, 0, name);
    if (!f || mxIsEmpty(f)) return {};
    const double *p = mxGetPr(f);
    return std::vector<double>(p, p + mxGetNumberOfElements(f));
}

void read_lens(const mxArray *s, const char *pfx, const char *si_pfx, mt_lens *L)
{
    auto f = [&](const char *p, const char *name) {
        const std::string k = std::string(p) + "_" + name;
        return num(s, k.c_str());
    };
    L->m = f(pfx, "m");
    L->c_10 = f(pfx, "c_10"); L->c_12 = f(pfx, "c_12"); L->phi_12 = f(pfx, "phi_12");
    L->c_21 = f(pfx, "c_21"); L->phi_21 = f(pfx, "phi_21"); L->c_23 = f(pfx, "c_23"); L->phi_23 = f(pfx, "phi_23");
    L->c_30 = f(pfx, "c_30"); L->c_32 = f(pfx, "c_32"); L->phi_32 = f(pfx, "phi_32");
    L->c_34 = f(pfx, "c_34"); L->phi_34 = f(pfx, "phi_34");
    L->c_41 = f(pfx, "c_41"); L->phi_41 = f(pfx, "phi_41"); L->c_43 = f(pfx, "c_43"); L->phi_43 = f(pfx, "phi_43");
    L->c_45 = f(pfx, "c_45"); L->phi_45 = f(pfx, "phi_45");
    L->c_50 = f(pfx, "c_50"); L->c_52 = f(pfx, "c_52"); L->phi_52 = f(pfx, "phi_52");
    L->c_54 = f(pfx, "c_54"); L->phi_54 = f(pfx, "phi_54"); L->c_56 = f(pfx, "c_56"); L->phi_56 = f(pfx, "phi_56");
    L->inner_aper_ang = f(pfx, "inner_aper_ang");
    L->outer_aper_ang = f(pfx, "outer_aper_ang");
    L->ti_a = f(pfx, "ti_a"); L->ti_sigma = f(pfx, "ti_sigma"); L->ti_beta = f(pfx, "ti_beta");
    L->ti_npts = (int)f(pfx, "ti_npts");
    // both lenses take the source spread from cond_lens_si_* (ilc_multem.cu:192-196, 246-250)
    L->si_a = f(si_pfx, "si_a"); L->si_sigma = f(si_pfx, "si_sigma"); L->si_beta = f(si_pfx, "si_beta");
    L->si_rad_npts = (int)f(si_pfx, "si_rad_npts"); L->si_azm_npts = (int)f(si_pfx, "si_azm_npts");
    L->zero_defocus_type = (int)f(pfx, "zero_defocus_type");
    L->zero_defocus_plane = f(pfx, "zero_defocus_plane");
}

// Host copies of the arrays the input struct points to.
struct InputArrays {
    std::vector<double> atoms, thick, iw_psi, iw_x, iw_y;
    std::vector<mt_detector_cir> det;
    std::vector<double> det_fR;
};

void read_input(const mxArray *s, mt_input_multislice *in, InputArrays &A)
{
    std::memset(in, 0, sizeof *in);
    in->simulation_type = (int)num(s, "simulation_type");
    in->interaction_model = (int)num(s, "interaction_model");
    in->potential_type = (int)num(s, "potential_type");
    in->operation_mode = (int)num(s, "operation_mode");
    in->reverse_multislice = (int)num(s, "reverse_multislice");
    in->pn_model = (int)num(s, "pn_model");
    in->pn_coh_contrib = (int)num(s, "pn_coh_contrib");
    in->pn_single_conf = (int)num(s, "pn_single_conf");
    in->pn_nconf = (int)num(s, "pn_nconf");
    in->pn_dim = (int)num(s, "pn_dim");
    in->pn_seed = (int)num(s, "pn_seed");

    const mxArray *atoms = mxGetField(s, 0, "spec_atoms");
    if (atoms && !mxIsEmpty(atoms)) {
        const double *p = mxGetPr(atoms);
        A.atoms.assign(p, p + mxGetNumberOfElements(atoms));   // column-major n x cols, as MATLAB holds it
        in->spec_atoms = A.atoms.data();
        in->spec_atoms_rows = (int)mxGetM(atoms);
        in->spec_atoms_cols = (int)mxGetN(atoms);
    }
    in->spec_lx = num(s, "spec_lx"); in->spec_ly = num(s, "spec_ly");
    in->spec_lz = num(s, "spec_lz"); in->spec_dz = num(s, "spec_dz");
    in->spec_rot_theta = num(s, "spec_rot_theta");
    {
        const std::vector<double> u0 = vec(s, "spec_rot_u0"), p0 = vec(s, "spec_rot_center_p");
        for (int k = 0; k < 3; ++k) {
            in->spec_rot_u0[k] = k < (int)u0.size() ? u0[k] : 0.0;
            in->spec_rot_center_p[k] = k < (int)p0.size() ? p0[k] : 0.0;
        }
        in->spec_rot_center_type = (int)num(s, "spec_rot_center_type");
    }

    in->thick_type = (int)num(s, "thick_type");
    A.thick = vec(s, "thick");
    in->thick = A.thick.empty() ? nullptr : A.thick.data();
    in->n_thick = (int)A.thick.size();
    in->potential_slicing = (int)num(s, "potential_slicing");
    in->nx = (int)num(s, "nx");
    in->ny = (int)num(s, "ny");
    in->bwl = (int)num(s, "bwl");

    in->iw_type = (int)num(s, "iw_type");
    const mxArray *psi = mxGetField(s, 0, "iw_psi");
    if (psi && !mxIsEmpty(psi)) {
        // interleave MATLAB's separate real / imaginary parts
        const size_t n = mxGetNumberOfElements(psi);
        const double *re = mxGetPr(psi), *im = mxIsComplex(psi) ? mxGetPi(psi) : nullptr;
        A.iw_psi.resize(2 * n);
        for (size_t i = 0; i < n; ++i) {
            A.iw_psi[2 * i] = re[i];
            A.iw_psi[2 * i + 1] = im ? im[i] : 0.0;
        }
        in->iw_psi = A.iw_psi.data();
    }
    A.iw_x = vec(s, "iw_x");
    A.iw_y = vec(s, "iw_y");
    in->iw_x = A.iw_x.empty() ? nullptr : A.iw_x.data();
    in->iw_y = A.iw_y.empty() ? nullptr : A.iw_y.data();
    in->n_iw_xy = (int)std::min(A.iw_x.size(), A.iw_y.size());

    in->E_0 = num(s, "E_0");
    in->theta = num(s, "theta");
    in->phi = num(s, "phi");
    // precession / hollow-cone fields (ilc_multem.cu:338-347)
    in->ped_theta = num(s, "ped_theta");
    in->ped_nrot = (int)num(s, "ped_nrot");
    in->hci_theta = num(s, "hci_theta");
    in->hci_nrot = (int)num(s, "hci_nrot");
    in->illumination_model = (int)num(s, "illumination_model");
    in->temporal_spatial_incoh = (int)num(s, "temporal_spatial_incoh");
    read_lens(s, "cond_lens", "cond_lens", &in->cond_lens);
    read_lens(s, "obj_lens", "cond_lens", &in->obj_lens);

    in->scanning_type = (int)num(s, "scanning_type");
    in->scanning_periodic = (int)num(s, "scanning_periodic");
    in->scanning_square_pxs = (int)num(s, "scanning_square_pxs");
    in->scanning_ns = (int)num(s, "scanning_ns");
    in->scanning_x0 = num(s, "scanning_x0"); in->scanning_y0 = num(s, "scanning_y0");
    in->scanning_xe = num(s, "scanning_xe"); in->scanning_ye = num(s, "scanning_ye");

    // detector.type and detector.cir(i).inner_ang / outer_ang (ilc_multem.cu:272-290)
    const mxArray *det = mxGetField(s, 0, "detector");
    if (det) {
        in->detector_type = (int)num(det, "type");
        const mxArray *cir = mxGetField(det, 0, "cir");
        if (cir && in->detector_type == 1) {
            const mwSize n = mxGetN(cir);
            for (mwSize i = 0; i < n; ++i) A.det.push_back(mt_detector_cir{num(cir, "inner_ang", i), num(cir, "outer_ang", i)});
        }
        // detector.matrix(i).fR, ny x nx each (ilc_multem.cu:317-333)
        const mxArray *mat = mxGetField(det, 0, "matrix");
        if (mat && in->detector_type == 3) {
            const mwSize n = mxGetN(mat);
            for (mwSize i = 0; i < n; ++i) {
                const mxArray *f = mxGetField(mat, i, "fR");
                if (!f || mxGetNumberOfElements(f) != (size_t)in->nx * in->ny)
                    mexErrMsgIdAndTxt("multem_hip:args", "detector.matrix(%d).fR must be ny x nx", (int)i + 1);
                const double *p = mxGetPr(f);
                A.det_fR.insert(A.det_fR.end(), p, p + mxGetNumberOfElements(f));
            }
            in->n_detector = (int)n;
        }
    }
    if (in->detector_type != 3) in->n_detector = (int)A.det.size();
    in->detector_cir = A.det.empty() ? nullptr : A.det.data();
    in->detector_fR = A.det_fR.empty() ? nullptr : A.det_fR.data();

    in->output_area_ix_0 = (int)num(s, "output_area_ix_0");
    in->output_area_iy_0 = (int)num(s, "output_area_iy_0");
    in->output_area_ix_e = (int)num(s, "output_area_ix_e");
    in->output_area_iy_e = (int)num(s, "output_area_iy_e");
}

mxArray *real_matrix(mwSize m, mwSize n, const double *v)
{
    mxArray *a = mxCreateDoubleMatrix(m, n, mxREAL);
    if (v) std::memcpy(mxGetPr(a), v, sizeof(double) * m * n);
    return a;
}

// interleaved complex -> MATLAB complex matrix
mxArray *complex_matrix(mwSize m, mwSize n, const double *v)
{
    mxArray *a = mxCreateDoubleMatrix(m, n, mxCOMPLEX);
    double *re = mxGetPr(a), *im = mxGetPi(a);
    for (mwSize i = 0; i < m * n; ++i) {
        re[i] = v[2 * i];
        im[i] = v[2 * i + 1];
    }
    return a;
}

void check(mt_status st)
{
    if (st != MT_OK) mexErrMsgIdAndTxt("multem_hip:run", "%s", mt_hip_last_error());
}

} // namespace

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)nlhs;
    if (nrhs < 1 || !mxIsStruct(prhs[0])) mexErrMsgIdAndTxt("multem_hip:args", "usage: ilc_multem_hip([system_conf,] input)");
    // read_system_conf (src/matlab_mex.cuh:263-288): a struct with a
    // precision field is the system configuration, the input follows
    mt_system_conf sys{1, 2, 1, 0};
    int i0 = 0;
    if (mxGetField(prhs[0], 0, "precision")) {
        sys.precision = (int)num(prhs[0], "precision");
        sys.device = (int)num(prhs[0], "device");
        sys.cpu_nthread = (int)num(prhs[0], "cpu_nthread");
        sys.gpu_device = (int)num(prhs[0], "gpu_device");
        i0 = 1;
    }
    if (nrhs <= i0 || !mxIsStruct(prhs[i0])) mexErrMsgIdAndTxt("multem_hip:args", "input_multislice struct expected");

    mt_input_multislice in;
    InputArrays A;
    read_input(prhs[i0], &in, A);

    mt_output_multislice out;
    std::memset(&out, 0, sizeof out);
    check(mt_hip_output_dims(&sys, &in, &out));
    const size_t npx = (size_t)out.nx * out.ny, nthk = (size_t)out.n_thk, ndet = (size_t)out.n_det;
    std::vector<double> x(out.nx), y(out.ny), thick(nthk), img_tot, img_coh, m2_tot, m2_coh, psi;
    out.x = x.data();
    out.y = y.data();
    out.thick = thick.data();
    const int ot = out.output_type;
    const bool stem = ot == MT_OT_IMAGE_TOT || ot == MT_OT_IMAGE_TOT_COH;
    if (stem) {
        img_tot.resize(nthk * ndet * npx);
        out.image_tot = img_tot.data();
        if (ot == MT_OT_IMAGE_TOT_COH) {
            img_coh.resize(nthk * ndet * npx);
            out.image_coh = img_coh.data();
        }
    }
    if (ot == MT_OT_M2PSI_TOT || ot == MT_OT_M2PSI_TOT_COH || ot == MT_OT_M2PSI_TOT_PSI_COH) {
        m2_tot.resize(nthk * npx);
        out.m2psi_tot = m2_tot.data();
    }
    if (ot == MT_OT_M2PSI_TOT_COH) {
        m2_coh.resize(nthk * npx);
        out.m2psi_coh = m2_coh.data();
    }
    if (ot == MT_OT_PSI_COH || ot == MT_OT_M2PSI_TOT_PSI_COH) {
        psi.resize(2 * nthk * npx);
        out.psi_coh = psi.data();
    }
    check(mt_hip_run_multislice(&sys, &in, &out));

    // set_struct_multislice (ilc_multem.cu:381-481)
    const char *top[] = {"dx", "dy", "x", "y", "thick", "data"};
    mxArray *o = mxCreateStructMatrix(1, 1, 6, top);
    mxSetField(o, 0, "dx", mxCreateDoubleScalar(out.dx));
    mxSetField(o, 0, "dy", mxCreateDoubleScalar(out.dy));
    mxSetField(o, 0, "x", real_matrix(1, out.nx, x.data()));
    mxSetField(o, 0, "y", real_matrix(1, out.ny, y.data()));
    mxSetField(o, 0, "thick", real_matrix(1, nthk, thick.data()));
    if (stem) {
        const bool coh = ot == MT_OT_IMAGE_TOT_COH;
        const char *names[] = {"image_tot", "image_coh"};
        mxArray *data = mxCreateStructMatrix(1, nthk, coh ? 2 : 1, names);
        const char *img[] = {"image"};
        for (size_t k = 0; k < nthk; ++k) {
            mxArray *dt = mxCreateStructMatrix(1, ndet, 1, img);
            mxArray *dc = coh ? mxCreateStructMatrix(1, ndet, 1, img) : nullptr;
            for (size_t d = 0; d < ndet; ++d) {
                // image(iy, ix): ny x nx, column-major = the engine's scan order
                mxSetField(dt, d, "image", real_matrix(out.ny, out.nx, &img_tot[(k * ndet + d) * npx]));
                if (dc) mxSetField(dc, d, "image", real_matrix(out.ny, out.nx, &img_coh[(k * ndet + d) * npx]));
            }
            mxSetField(data, k, "image_tot", dt);
            if (dc) mxSetField(data, k, "image_coh", dc);
        }
        mxSetField(o, 0, "data", data);
    } else if (ot == MT_OT_PSI_COH || ot == MT_OT_M2PSI_TOT_PSI_COH) {
        const bool full = ot == MT_OT_M2PSI_TOT_PSI_COH;
        const char *names_full[] = {"m2psi_tot", "psi_coh"};
        const char *names_sc[] = {"psi_coh"};
        mxArray *data = mxCreateStructMatrix(1, nthk, full ? 2 : 1, full ? names_full : names_sc);
        for (size_t k = 0; k < nthk; ++k) {
            if (full) mxSetField(data, k, "m2psi_tot", real_matrix(out.ny, out.nx, &m2_tot[k * npx]));
            mxSetField(data, k, "psi_coh", complex_matrix(out.ny, out.nx, &psi[2 * k * npx]));
        }
        mxSetField(o, 0, "data", data);
    } else {
        const bool coh = ot == MT_OT_M2PSI_TOT_COH;
        const char *names[] = {"m2psi_tot", "m2psi_coh"};
        mxArray *data = mxCreateStructMatrix(1, nthk, coh ? 2 : 1, names);
        for (size_t k = 0; k < nthk; ++k) {
            mxSetField(data, k, "m2psi_tot", real_matrix(out.ny, out.nx, &m2_tot[k * npx]));
            if (coh) mxSetField(data, k, "m2psi_coh", real_matrix(out.ny, out.nx, &m2_coh[k * npx]));
        }
        mxSetField(o, 0, "data", data);
    }
    plhs[0] = o;
}
