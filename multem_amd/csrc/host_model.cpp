// Host model: input reading and validation, spline tables, slicing, phonons.
//
// Follows Input_Multislice<T> of the reference: quantities the reference keeps
// in T are R here (float for the FP32 path, double for the FP64 path); tables
// are built in double and cast.
#include "host_model.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <numeric>
#include <random>

#include "lobato_table.inc"
#include "quadrature.hpp"

namespace mtg {

namespace {

constexpr double kPi = 3.141592653589793238463;
constexpr double kSqrt2 = 1.414213562373095048802;
constexpr double kTwoPi = 6.283185307179586476925;
constexpr double kPi2 = 9.869604401089358618834;      // c_Pi2 (src/types.cuh:88)
constexpr double kPotf = 47.877645145863056;      // c_Potf (src/types.cuh:66)
constexpr double kDegToRad = 0.01745329251994329576924;
constexpr double kMradToRad = 1.0e-3;
constexpr double kMmToAngs = 1.0e7;
constexpr double kSqrtPi = 1.772453850905516027298;   // c_Pii2 (src/types.cuh:90)

[[noreturn]] void fail(mt_status c, const std::string &m) { throw Error{c, m}; }

// isEqual<float> / isEqual<double> (src/types.cuh:562-592)
bool near_zero(float a) { return std::fabs(a) <= 1e-5f; }
bool near_zero(double a) { return std::fabs(a) <= 1e-13; }
template <class R>
bool same_value(R a, R b)
{
    // isEqual<T> (src/types.cuh:562-592)
    const R eps_abs = sizeof(R) == 4 ? R(1e-5) : R(1e-13), eps_rel = sizeof(R) == 4 ? R(1e-4) : R(1e-8);
    R d = std::fabs(a - b);
    return d <= eps_abs || d <= std::max(std::fabs(a), std::fabs(b)) * eps_rel;
}

// Polynomial approximations of I0, I1, K0, K1: Abramowitz & Stegun 9.8.1-9.8.8.
double horner(const double *c, int n, double y)
{
    double r = c[n - 1];
    for (int i = n - 2; i >= 0; --i) r = c[i] + y * r;
    return r;
}
double i0_as(double x)
{
    static const double a[] = {1.0, 3.5156229, 3.0899424, 1.2067492, 0.2659732, 0.360768e-1, 0.45813e-2};
    static const double b[] = {0.39894228, 0.1328592e-1, 0.225319e-2, -0.157565e-2, 0.916281e-2,
                               -0.2057706e-1, 0.2635537e-1, -0.1647633e-1, 0.392377e-2};
    double ax = std::fabs(x);
    if (ax < 3.75) { double t = x / 3.75; return horner(a, 7, t * t); }
    return (std::exp(ax) / std::sqrt(ax)) * horner(b, 9, 3.75 / ax);
}
double i1_as(double x)
{
    static const double a[] = {0.5, 0.87890594, 0.51498869, 0.15084934, 0.2658733e-1, 0.301532e-2, 0.32411e-3};
    double ax = std::fabs(x), r;
    if (ax < 3.75) {
        double t = x / 3.75;
        r = ax * horner(a, 7, t * t);
    } else {
        double y = 3.75 / ax;
        double tail = 0.2282967e-1 + y * (-0.2895312e-1 + y * (0.1787654e-1 - y * 0.420059e-2));
        r = 0.39894228 + y * (-0.3988024e-1 + y * (-0.362018e-2 + y * (0.163801e-2 + y * (-0.1031555e-1 + y * tail))));
        r *= std::exp(ax) / std::sqrt(ax);
    }
    return x < 0.0 ? -r : r;
}
double k0_as(double x)
{
    static const double a[] = {-0.57721566, 0.42278420, 0.23069756, 0.3488590e-1, 0.262698e-2, 0.10750e-3, 0.74e-5};
    static const double b[] = {1.25331414, -0.7832358e-1, 0.2189568e-1, -0.1062446e-1, 0.587872e-2, -0.251540e-2, 0.53208e-3};
    if (x <= 2.0) return (-std::log(x / 2.0) * i0_as(x)) + horner(a, 7, x * x / 4.0);
    return (std::exp(-x) / std::sqrt(x)) * horner(b, 7, 2.0 / x);
}
double k1_as(double x)
{
    static const double a[] = {1.0, 0.15443144, -0.67278579, -0.18156897, -0.1919402e-1, -0.110404e-2, -0.4686e-4};
    static const double b[] = {1.25331414, 0.23498619, -0.3655620e-1, 0.1504268e-1, -0.780353e-2, 0.325614e-2, -0.68245e-3};
    if (x <= 2.0) return (std::log(x / 2.0) * i1_as(x)) + (1.0 / x) * horner(a, 7, x * x / 4.0);
    return (std::exp(-x) / std::sqrt(x)) * horner(b, 7, 2.0 / x);
}

// Lobato hydrogen-like terms for the 3D potential Vr and the projected VR
// (set_Lobato_neutral_coef, src/atomic_data.hpp:1440-1485).
struct Terms { double a[5], k[5]; };
void lobato_terms(int Z, Terms &vr, Terms &VR)
{
    const double *t = mt_lobato_coef[Z - 1];
    for (int j = 0; j < 5; ++j) {
        double cl = t[2 * j], cnl = near_zero(cl) ? 0.0 : t[2 * j + 1];
        if (near_zero(cl) && near_zero(cnl)) {
            vr.a[j] = VR.a[j] = 0.0;
            vr.k[j] = VR.k[j] = 1.0;
            continue;
        }
        // same operation order as the reference, so the tables round identically
        vr.a[j] = kPotf * kPi2 * cl / std::pow(cnl, 1.5);
        VR.a[j] = 2.0 * kPotf * kPi2 * cl / std::pow(cnl, 1.5);
        vr.k[j] = VR.k[j] = kTwoPi / std::sqrt(cnl);
    }
}
double vr3(const Terms &t, double r)
{
    double ir = 1 / r, s = 0;
    for (int i = 0; i < 5; ++i) s += t.a[i] * std::exp(-t.k[i] * r) * ir * (2 / t.k[i] + r);
    return s;
}

} // namespace

template <class T>
AtomTableT<T> build_atom_table(int Z, double dR_min, double Vrl, int nR)
{
    AtomTableT<T> tb;
    tb.Z = Z;
    Terms vr, VR;
    lobato_terms(Z, vr, VR);
    const double rn_c = 1.2e-05 * std::pow((double)mt_element_table[Z - 1].A, 1.0 / 3.0);
    const double R_min = std::max(rn_c, dR_min);

    // cut-off radius where Vr = Vrl: bisection on [rn_c, 25]
    // (AtomicRadius_Cutoff + Root_Finder, src/atomic_fcns_mt.hpp:269-296, src/cgpu_fcns.cuh:262-286)
    double R_max = 0.0;
    if (!(near_zero(Vrl) || near_zero(vr.a[0]))) {
        double lo = rn_c, hi = 25.0, f_hi = vr3(vr, hi) - Vrl, x, fx;
        int it = 0;
        do {
            x = 0.5 * (lo + hi);
            fx = vr3(vr, x) - Vrl;
            if (fx * f_hi < 0) lo = x;
            else { hi = x; f_hi = fx; }
            ++it;
        } while (std::fabs(fx) > 1e-8 && it < 200);
        R_max = x;
    }
    R_max = std::max(R_max, 2.0 * R_min);
    if (near_zero(R_max)) R_max = 1.75 * mt_element_table[Z - 1].ra;

    tb.R_min = R_min;
    tb.R_max = R_max;
    tb.R_tap = R_min + 0.85 * (R_max - R_min);
    const double R2_tap = tb.R_tap * tb.R_tap;
    tb.tap_cf = (kPi / 2) / (R_max * R_max - R2_tap);

    std::vector<double> R(nR), R2(nR), y(nR), dy(nR);
    const double dlnr = std::log(R_max / R_min) / double(nR - 1);
    for (int i = 0; i < nR; ++i) {
        R[i] = R_min * std::exp(i * dlnr);
        R2[i] = std::pow(R[i], 2);
        double v = 0, dv = 0;
        for (int j = 0; j < 5; ++j) {
            double kx = VR.k[j] * R[i], k0 = k0_as(kx), k1 = k1_as(kx);
            v += VR.a[j] * (2 * k0 / VR.k[j] + R[i] * k1);
            dv += -VR.a[j] * (VR.k[j] * R[i] * k0 + 2 * k1);
        }
        double d = 0.5 * dv / R[i];       // derivative with respect to R^2
        if (R2_tap < R2[i]) {             // cosine taper (apply_tapering, src/cgpu_fcns.cuh:1528-1538)
            double arg = tb.tap_cf * (R2[i] - R2_tap);
            double c = std::cos(arg), s = std::sin(arg);
            d = d * c - tb.tap_cf * v * s;
            v *= c;
        }
        y[i] = v;
        dy[i] = d;
    }
    tb.R2.resize(nR); tb.c0.resize(nR); tb.c1.resize(nR); tb.c2.assign(nR, T(0)); tb.c3.assign(nR, T(0));
    for (int i = 0; i < nR; ++i) {
        tb.R2[i] = (T)R2[i];
        tb.c0[i] = (T)y[i];
        tb.c1[i] = (T)dy[i];
    }
    for (int i = 0; i + 1 < nR; ++i) {   // Hermite cubic on [R2_i, R2_i+1]
        double h = 1.0 / (R2[i + 1] - R2[i]);
        double m = (y[i + 1] - y[i]) * h, n = dy[i] + dy[i + 1];
        tb.c2[i] = (T)((3.0 * m - n - dy[i]) * h);
        tb.c3[i] = (T)((n - 2.0 * m) * h * h);
    }
    tb.R_max_f = (T)R_max;
    tb.R2_max_f = tb.R_max_f * tb.R_max_f;
    for (int j = 0; j < 5; ++j) { tb.vr_a[j] = vr.a[j]; tb.vr_k[j] = vr.k[j]; }
    return tb;
}

template AtomTableT<float> build_atom_table<float>(int, double, double, int);
template AtomTableT<double> build_atom_table<double>(int, double, double, int);

// ---------------------------------------------------------------------------
namespace {

template <class R>
R lambda_of(R E_0)                         // get_lambda<T> (src/cgpu_fcns.cuh:51-58)
{
    R emass = R(510.99906), hc = R(12.3984244);
    return hc / std::sqrt(E_0 * (2 * emass + E_0));
}

template <class R>
void read_lens(const mt_lens &m, LensModelT<R> &L)
{
    L.m = (int)m.m;
    L.c_10 = (R)m.c_10; L.c_12 = (R)m.c_12; L.phi_12 = (R)(m.phi_12 * kDegToRad);
    L.c_21 = (R)m.c_21; L.phi_21 = (R)(m.phi_21 * kDegToRad);
    L.c_23 = (R)m.c_23; L.phi_23 = (R)(m.phi_23 * kDegToRad);
    L.c_30 = (R)(m.c_30 * kMmToAngs); L.c_32 = (R)m.c_32; L.phi_32 = (R)(m.phi_32 * kDegToRad);
    L.c_34 = (R)m.c_34; L.phi_34 = (R)(m.phi_34 * kDegToRad);
    L.c_41 = (R)m.c_41; L.phi_41 = (R)(m.phi_41 * kDegToRad);
    L.c_43 = (R)m.c_43; L.phi_43 = (R)(m.phi_43 * kDegToRad);
    L.c_45 = (R)m.c_45; L.phi_45 = (R)(m.phi_45 * kDegToRad);
    L.c_50 = (R)(m.c_50 * kMmToAngs); L.c_52 = (R)m.c_52; L.phi_52 = (R)(m.phi_52 * kDegToRad);
    L.c_54 = (R)m.c_54; L.phi_54 = (R)(m.phi_54 * kDegToRad);
    L.c_56 = (R)m.c_56; L.phi_56 = (R)(m.phi_56 * kDegToRad);
    L.inner_aper_ang = (R)(m.inner_aper_ang * kMradToRad);
    L.outer_aper_ang = (R)(m.outer_aper_ang * kMradToRad);
    L.zero_defocus_type = m.zero_defocus_type;
    L.zero_defocus_plane = (R)m.zero_defocus_plane;
    L.ti_a = (R)m.ti_a; L.ti_sigma = (R)m.ti_sigma; L.ti_beta = (R)m.ti_beta; L.ti_npts = m.ti_npts;
    L.si_a = (R)m.si_a; L.si_sigma = (R)m.si_sigma; L.si_beta = (R)m.si_beta;
    L.si_rad_npts = m.si_rad_npts; L.si_azm_npts = m.si_azm_npts;
}

// the spread-function clamps of Lens::set_input_data (src/types.cuh:2953-2963)
template <class R>
void clamp_spreads(LensModelT<R> &L)
{
    L.ti_a = std::max(R(0), std::min(R(1), L.ti_a));
    L.ti_sigma = std::max(R(0), L.ti_sigma);
    L.ti_beta = std::max(R(0), L.ti_beta);
    L.ti_npts = std::max(1, L.ti_npts);
    L.si_a = std::max(R(0), std::min(R(1), L.si_a));
    L.si_sigma = std::max(R(0), L.si_sigma);
    L.si_beta = std::max(R(0), L.si_beta);
    L.si_rad_npts = std::max(1, L.si_rad_npts);
    L.si_azm_npts = std::max(1, L.si_azm_npts);
}

// aperture bounds of Lens::set_input_data (src/types.cuh:2924-2976)
template <class R>
void set_lens(LensModelT<R> &L, R E_0, R g_max)
{
    L.lambda = lambda_of<R>(E_0);
    L.g2_min = (near_zero(L.inner_aper_ang) || L.inner_aper_ang < 0)
                   ? R(0.0) : (R)std::pow(std::sin(L.inner_aper_ang) / L.lambda, 2);
    L.g2_max = (near_zero(L.outer_aper_ang) || L.outer_aper_ang < 0)
                   ? g_max * g_max : (R)std::pow(std::sin(L.outer_aper_ang) / L.lambda, 2);
    const R c[] = {(R)L.m, L.c_12, L.c_21, L.c_23, L.c_32, L.c_34, L.c_41, L.c_43, L.c_45, L.c_52, L.c_54, L.c_56};
    L.phi_required = false;
    for (R v : c) L.phi_required = L.phi_required || !near_zero(v);
}

template <class R>
R zero_defocus_plane(const LensModelT<R> &L, R z_min, R z_max)
{
    switch (L.zero_defocus_type) {
    case 1: return z_min;
    case 2: return R(0.5) * (z_min + z_max);
    case 3: return z_max;
    default: return L.zero_defocus_plane;
    }
}

// Plane finder of Identify_Planes (src/types.cuh:1166-1372): a 0.1 A histogram
// whose one-bin gaps inside a layer are closed, then Kahan means per layer.
template <class R>
std::vector<R> find_planes(std::vector<R> v)
{
    std::vector<R> planes;
    if (v.empty()) return planes;
    std::sort(v.begin(), v.end());
    const double dv = 0.1;
    const double v_min = v.front(), v_max = v.back();
    int nb = (int)std::ceil(std::max(v_max - v_min, dv) / dv);
    std::vector<int> h(nb, 0);
    for (R f : v) {
        double x = f;
        int ih = (int)std::floor((x - v_min) / dv);
        // move to the bin [v_min + ih dv, v_min + (ih+1) dv) that holds x
        if (x < v_min + (ih - 1) * dv) {
            for (int k = ih; k >= 0; --k) {
                double lo = v_min + (k - 1) * dv;
                if (lo <= x && x < lo + dv) { ih = k - 1; break; }
            }
        } else if (x > v_min + ih * dv) {
            for (int k = ih; k < nb; ++k) {
                double lo = v_min + k * dv;
                if (lo <= x && x < lo + dv) { ih = k; break; }
            }
        }
        ih = std::max(0, std::min(nb - 1, ih));
        h[ih]++;
    }
    while (!h.empty() && h.back() == 0) h.pop_back();
    nb = (int)h.size();
    for (int i = 1; i < nb - 1; ++i) {
        bool ahead = (i < nb - 2) ? (h[i + 2] > 0) : false;
        if (h[i] == 0 && h[i - 1] > 0 && (h[i + 1] > 0 || ahead)) h[i] = 1;
    }
    if (nb == 1) {
        R s = 0;
        for (R f : v) s += f;
        planes.push_back(s / R(v.size()));
        return planes;
    }
    std::vector<R> lim;
    for (int i = 0; i + 1 < nb; ++i)
        if (h[i] > 0 && h[i + 1] == 0) lim.push_back((R)(v.front() + (i + 1) * dv));
    lim.push_back((R)(v.back() + dv));
    R mean = v.front(), carry = 0;
    int cnt = 1;
    size_t il = 0;
    for (R f : v) {
        if (f < lim[il]) {
            R y = f - carry, t = mean + y;   // Kahan summation
            carry = (t - mean) - y;
            mean = t;
            ++cnt;
        } else {
            planes.push_back(mean / cnt);
            mean = f; carry = 0; cnt = 1; ++il;
        }
    }
    planes.push_back(mean / cnt);
    return planes;
}

// evenly spaced planes with matched borders (Identify_Planes(v_min, v_max, dv), eMB_MinMax)
template <class R>
std::vector<R> even_planes(R v_min, R v_max, R dv)
{
    const R eps = R(1e-4);
    if (std::fabs(v_max - v_min) < eps) return {R(0.5) * (v_min + v_max)};
    if (v_max < v_min) return {};
    R s = v_max - v_min;
    int nv = std::max(1, (int)std::floor(s / dv + eps)) + 1;
    R db = dv + R(0.5) * (s - (nv - 1) * dv);
    std::vector<R> p(nv);
    p[0] = v_min;
    for (int i = 1; i < nv; ++i) p[i] = p[i - 1] + ((i == 1 || i == nv - 1) ? db : dv);
    return p;
}

template <class R>
void unique_sorted(std::vector<R> &v)
{
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end(), same_value<R>), v.end());
}

template <class R>
R spacing(const std::vector<R> &p, int i, R dz)
{
    if (p.size() == 1) return dz;
    i = std::max(1, std::min(i, (int)p.size() - 1));
    return p[i] - p[i - 1];
}

template <class R>
std::vector<R> slice_bounds(const std::vector<R> &p, R dz, R z_int_min, R z_int_max, bool sub = false)
{
    if (p.size() == 1 && !sub) return {z_int_min, z_int_max};
    std::vector<R> zs(p.size() + 1);
    zs[0] = p[0] - R(0.5) * spacing(p, 0, dz);
    for (size_t i = 1; i + 1 < zs.size(); ++i) zs[i] = R(0.5) * (p[i] + p[i - 1]);
    zs.back() = zs[zs.size() - 2] + spacing(p, (int)p.size(), dz);
    if (sub) {
        // dz_Sub: slices of the neighbouring spacing out to the interaction
        // range on both sides (Slicing::get_z_slice, src/slicing.hpp:220-247)
        R dz_b = spacing(p, 1, dz);
        if (z_int_min < zs.front() - dz_b) {
            std::vector<R> top = even_planes(z_int_min, zs.front() - dz_b, spacing(p, 2, dz));
            zs.insert(zs.begin(), top.begin(), top.end());
        } else {
            zs.front() = z_int_min;
        }
        dz_b = spacing(p, (int)p.size() - 1, dz);
        if (zs.back() + dz_b < z_int_max) {
            std::vector<R> bottom = even_planes(zs.back() + dz_b, z_int_max, spacing(p, (int)p.size() - 2, dz));
            zs.insert(zs.end(), bottom.begin(), bottom.end());
        } else {
            zs.back() = z_int_max;
        }
    }
    return zs;
}

} // namespace

template <class R>
ModelT<R>::ModelT(const mt_input_multislice &in)
{
    sim_type = in.simulation_type;
    if (!(is_stem() || is_istem() || is_cbed() || is_cbei() || is_ed() || is_hrtem() || is_ped() || is_hctem() ||
          is_ewfs() || is_ewrs()))
        fail(MT_ERR_UNSUPPORTED, "simulation_type must be 11 (STEM), 12 (ISTEM), 21 (CBED), 22 (CBEI), 31 (ED), "
                                 "32 (HRTEM), 41 (PED), 42 (HCTEM), 51 (EWFS) or 52 (EWRS)");
    if (in.interaction_model < 1 || in.interaction_model > 3)
        fail(MT_ERR_INVALID_ARGUMENT, "interaction_model must be 1 (multislice), 2 (phase object) or 3 (weak phase object)");
    interaction_model = in.interaction_model;
    if (in.potential_type != 6) fail(MT_ERR_UNSUPPORTED, "only potential_type 6 (Lobato)");
    if (in.pn_model == 2) fail(MT_ERR_UNSUPPORTED, "absorptive potential model");
    if (in.pn_model < 1 || in.pn_model > 3) fail(MT_ERR_INVALID_ARGUMENT, "pn_model must be 1..3");
    // Auto (4) has no implementation in the reference (only the enum, src/types.cuh:237)
    if (!(in.potential_slicing >= 1 && in.potential_slicing <= 3))
        fail(MT_ERR_UNSUPPORTED, "potential_slicing must be 1 (planes), 2 (dz_Proj) or 3 (dz_Sub)");
    // exit waves are coherent; CBED is coherent unless fully integrated
    // (Input_Multislice::validate_parameters, src/input_multislice.cuh:672-682)
    illumination_model = in.illumination_model;
    if (is_ewfs() || is_ewrs() || (is_cbed_cbei() && illumination_model != 4)) illumination_model = 1;
    // no source spread: temporal incoherence only (src/input_multislice.cuh:746-751)
    temporal_spatial_incoh = near_zero((R)in.cond_lens.si_sigma) ? 2 : in.temporal_spatial_incoh;
    if (is_image() && illumination_model == 4 && temporal_spatial_incoh != 2)
        fail(MT_ERR_UNSUPPORTED, "objective-lens full integration with spatial incoherence: the reference's "
                                 "source quadrature is empty (src/cpu_fcns.hpp:1535) and its image is zero");
    if (is_image() && illumination_model == 3)
        fail(MT_ERR_UNSUPPORTED, "objective-lens image with illumination_model 3 (transmission cross coefficient)");

    pn_model = in.pn_model;
    pn_coh_contrib = in.pn_coh_contrib;
    pn_single_conf = in.pn_single_conf;
    pn_nconf = in.pn_nconf;
    pn_seed = in.pn_seed;
    switch (in.pn_dim) {                // FP_Dim::set (src/types.cuh:1687-1740)
    case 111: pn_x = pn_y = pn_z = true; break;
    case 110: pn_x = pn_y = true; pn_z = false; break;
    case 101: pn_x = true; pn_y = false; pn_z = true; break;
    case 11: pn_x = false; pn_y = pn_z = true; break;
    case 100: pn_x = true; pn_y = pn_z = false; break;
    case 10: pn_x = false; pn_y = true; pn_z = false; break;
    case 1: pn_x = pn_y = false; pn_z = true; break;
    default: break;
    }

    // grid (Grid_2d::set_input_data, src/types.cuh:2154-2178), periodic
    nx = in.nx; ny = in.ny;
    // transform lengths of the slice passes: 2^k, 3 x 2^k, 5 x 2^k in [64, 2048]
    auto supported = [](int n) {
        if (n < 64 || n > 2048) return false;
        const int m = n % 3 == 0 ? 3 : n % 5 == 0 ? 5 : 1, p = n / m;
        return (p & (p - 1)) == 0 && p % 16 == 0;
    };
    if (nx <= 0 || ny <= 0) fail(MT_ERR_INVALID_ARGUMENT, "nx, ny must be positive");
    if (!supported(nx) || !supported(ny))
        fail(MT_ERR_UNSUPPORTED, "nx, ny must be 2^k, 3 x 2^k or 5 x 2^k in [64, 2048]");
    nxh = nx / 2; nyh = ny / 2;
    lx = (R)in.spec_lx; ly = (R)in.spec_ly;
    lx_d = in.spec_lx; ly_d = in.spec_ly;
    if (!(lx > 0 && ly > 0)) fail(MT_ERR_INVALID_ARGUMENT, "spec_lx, spec_ly must be positive");
    dRx = lx / nx; dRy = ly / ny;
    dgx = R(1.0) / lx; dgy = R(1.0) / ly;
    bwl = in.bwl != 0;
    const R g_max = std::fmin((R)nxh * dgx, (R)nyh * dgy);
    const R gl_max = R(2.0) * g_max / R(3.0);
    gl2_max = gl_max * gl_max;
    bwl_alpha = (R)(std::log(1.0 / 1e-2 - 1.0) / ((gl_max + R(0.25)) * (gl_max + R(0.25)) - gl2_max));

    iw_type = in.iw_type;
    if (iw_type == 4) iw_type = (is_scanning() || is_cbed_cbei()) ? 2 : 1;
    if (iw_type < 1 || iw_type > 3) fail(MT_ERR_INVALID_ARGUMENT, "iw_type must be 1..4");
    for (int i = 0; i < in.n_iw_xy; ++i) { iw_x.push_back((R)in.iw_x[i]); iw_y.push_back((R)in.iw_y[i]); }
    if (!is_scanning() && iw_type == 2 && iw_x.empty()) fail(MT_ERR_INVALID_ARGUMENT, "convergent wave needs iw_x, iw_y");
    if (iw_type == 3) {
        // user-defined incident wave, ny x nx complex, placed at (iw_x[0], iw_y[0])
        // (Incident_Wave::operator(), src/incident_wave.cuh:92-110)
        if (!in.iw_psi) fail(MT_ERR_INVALID_ARGUMENT, "iw_type 3 needs iw_psi (ny x nx complex)");
        if (!is_scanning() && iw_x.empty()) fail(MT_ERR_INVALID_ARGUMENT, "iw_type 3 needs iw_x, iw_y");
        iw_psi.assign(in.iw_psi, in.iw_psi + 2 * (size_t)nx * ny);
    }

    E_0 = (R)in.E_0;
    if (!(E_0 > 0)) fail(MT_ERR_INVALID_ARGUMENT, "E_0 must be positive");
    theta = (R)(in.theta * kDegToRad);
    phi = (R)(in.phi * kDegToRad);
    // precession (PED) and hollow-cone (HCTEM) illumination: the cone angle
    // and the number of azimuths (ilc_multem.cu:338-347); phi follows the
    // rotation (validate_parameters, src/input_multislice.cuh:655-660)
    if (is_ped_hctem()) {
        theta = (R)((is_ped() ? in.ped_theta : in.hci_theta) * kDegToRad);
        nrot = std::max(1, is_ped() ? in.ped_nrot : in.hci_nrot);
        phi = R(0);
    }
    if (std::fabs(theta) >= R(0.5) * (R)kPi) fail(MT_ERR_INVALID_ARGUMENT, "theta must be below 90 degrees");
    lambda = lambda_of<R>(E_0);
    // Input_Multislice::set_incident_angle (src/input_multislice.cuh:891-895):
    // sin(theta) quantised to a multiple of lambda*dg_min
    if (theta != R(0.0)) {
        const R dgm = std::fmin(dgx, dgy);
        const R nq = std::round(std::sin(theta) / (lambda * dgm));
        theta = std::asin(nq * lambda * dgm);
    }
    // gx_0, gy_0 (src/input_multislice.cuh:919-927)
    gx0 = std::sin(theta) * std::cos(phi) / lambda;
    gy0 = std::sin(theta) * std::sin(phi) / lambda;
    {
        double e = in.E_0, emass = 510.99906, hc = 12.3984244;
        lambda_d = hc / std::sqrt(e * (2 * emass + e));
    }
    mul_sign = in.reverse_multislice ? -1 : 1;
    {
        R emass = R(510.99906);
        R gamma = 1 + E_0 / emass;
        R cpotf = (R)kPotf;
        Vr_factor = mul_sign * (gamma * lambda / (cpotf * std::cos(theta)));
    }

    read_lens(in.cond_lens, cond);
    set_lens(cond, E_0, g_max);
    read_lens(in.obj_lens, obj);
    // the objective lens takes its source spread from the cond_lens_si_* fields
    // (ilc_multem.cu:246-250)
    obj.si_a = cond.si_a; obj.si_sigma = cond.si_sigma; obj.si_beta = cond.si_beta;
    obj.si_rad_npts = cond.si_rad_npts; obj.si_azm_npts = cond.si_azm_npts;
    // integration points (src/input_multislice.cuh:753-776), then the clamps
    for (LensModel *L : {&cond, &obj}) {
        if (illumination_model != 4) L->si_rad_npts = L->si_azm_npts = L->ti_npts = 1;
        if (temporal_spatial_incoh == 2) L->si_rad_npts = L->si_azm_npts = 1;
        else if (temporal_spatial_incoh == 3) L->ti_npts = 1;
        clamp_spreads(*L);
    }
    // Lens::set_ti_sigma / set_si_sigma (src/types.cuh:2978-2989); the objective
    // lens reads si_sigma from cond_lens_si_sigma (ilc_multem.cu:247)
    obj.ti_iehwgd = (R)kSqrt2 * std::fmax(R(0.0), (R)in.obj_lens.ti_sigma);
    obj.si_theta_c = std::asin((R)kSqrt2 * std::fmax(R(0.0), (R)in.cond_lens.si_sigma) * lambda);
    set_lens(obj, E_0, g_max);

    // scanning (Scanning::set_grid, src/types.cuh:4294-4389)
    if (is_scanning()) {
        int ns = in.scanning_ns;
        bool pbc = in.scanning_periodic != 0, spxs = in.scanning_square_pxs != 0;
        if (ns <= 0) fail(MT_ERR_INVALID_ARGUMENT, "scanning_ns must be positive");
        R x0 = (R)in.scanning_x0, y0 = (R)in.scanning_y0;
        R xe = (R)in.scanning_xe, ye = (R)in.scanning_ye;
        R xu = xe - x0, yu = ye - y0;
        if (in.scanning_type == 1) {
            R ds = std::sqrt(yu * yu + xu * xu), th = std::atan2(yu, xu);
            R ct = std::cos(th);
            ct = near_zero(ct) ? R(0.0) : ct;
            R st = std::sin(th);
            sc_dRx = ds * ct / (pbc ? ns : (ns - 1));
            sc_dRy = ds * st / (pbc ? ns : (ns - 1));
            sc_nx = ns; sc_ny = 1;
            for (int i = 0; i < ns; ++i) { sc_x.push_back(x0 + i * sc_dRx); sc_y.push_back(y0 + i * sc_dRy); }
        } else {
            sc_nx = sc_ny = ns;
            if (std::fabs(xu) > std::fabs(yu)) {
                sc_dRx = xu / (pbc ? ns : (ns - 1));
                sc_dRy = std::copysign(sc_dRx, yu);
                sc_ny = (int)std::floor(yu / sc_dRy + R(1e-4) + R(0.5)) + (pbc ? 0 : 1);
                if (!spxs) sc_dRy = yu / (pbc ? sc_ny : (sc_ny - 1));
            } else {
                sc_dRy = yu / (pbc ? ns : (ns - 1));
                sc_dRx = std::copysign(sc_dRy, xu);
                sc_nx = (int)std::floor(xu / sc_dRx + R(1e-4) + R(0.5)) + (pbc ? 0 : 1);
                if (!spxs) sc_dRx = xu / (pbc ? sc_nx : (sc_nx - 1));
            }
            if (sc_nx <= 0 || sc_ny <= 0) fail(MT_ERR_INVALID_ARGUMENT, "empty scan grid");
            for (int ix = 0; ix < sc_nx; ++ix)
                for (int iy = 0; iy < sc_ny; ++iy) {
                    sc_x.push_back(x0 + ix * sc_dRx);
                    sc_y.push_back(y0 + iy * sc_dRy);
                }
        }
        sc_x0 = x0; sc_y0 = y0;
    }
    if (is_stem()) {
        // detectors (ilc_multem.cu:272-337): circular rings or matrix
        // sensitivity maps; the reference's radial detector integrates to zero
        // (integrated_intensity_over_det, src/wave_function.cuh:120-124)
        if (in.detector_type == 2)
            fail(MT_ERR_UNSUPPORTED, "radial detectors: the reference's radial detector integrates to zero");
        if (in.detector_type != 1 && in.detector_type != 3) fail(MT_ERR_INVALID_ARGUMENT, "detector_type must be 1..3");
        if (in.n_detector <= 0) fail(MT_ERR_INVALID_ARGUMENT, "STEM needs at least one detector");
        if (in.detector_type == 1) {
            if (!in.detector_cir) fail(MT_ERR_INVALID_ARGUMENT, "circular detectors need detector_cir");
            for (int i = 0; i < in.n_detector; ++i) {
                R gi = std::sin((R)(in.detector_cir[i].inner_ang * kMradToRad)) / lambda;
                R go = std::sin((R)(in.detector_cir[i].outer_ang * kMradToRad)) / lambda;
                det_g2_min.push_back((R)std::pow(gi, 2));
                det_g2_max.push_back((R)std::pow(go, 2));
            }
        } else {
            if (!in.detector_fR) fail(MT_ERR_INVALID_ARGUMENT, "matrix detectors need detector_fR");
            const size_t n = (size_t)in.n_detector * nx * ny;
            // the maps come centred (g = 0 at nx/2, ny/2, as the reference's
            // examples build them); fft2_shift takes them to the natural
            // Fourier order of the detector sums (Wave_Function::set_input_data,
            // src/wave_function.cuh:53-61)
            det_fR.resize(n);
            const int nxh = nx / 2, nyh = ny / 2;
            for (int d = 0; d < in.n_detector; ++d)
                for (int ix = 0; ix < nx; ++ix)
                    for (int iy = 0; iy < ny; ++iy) {
                        const int sx = ix < nxh ? ix + nxh : ix - nxh, sy = iy < nyh ? iy + nyh : iy - nyh;
                        det_fR[((size_t)d * nx + ix) * ny + iy] =
                            (R)in.detector_fR[((size_t)d * nx + sx) * ny + sy];
                    }
            det_g2_min.assign(in.n_detector, R(0));
            det_g2_max.assign(in.n_detector, R(0));
        }
    }

    // atoms (Atom_Data::set_atoms, src/atomic_data_mt.hpp:320-350)
    const int na = in.spec_atoms_rows, nc = in.spec_atoms_cols;
    if (na <= 0 || !in.spec_atoms || nc < 4) fail(MT_ERR_INVALID_ARGUMENT, "spec_atoms must be n x (4..8), n > 0");
    SpecimenT<R> in_atoms;
    for (int i = 0; i < na; ++i) {
        const double *A = in.spec_atoms;
        int Zi = (int)A[i];
        int charge = (int)(nc > 7 ? A[7 * na + i] : 0);
        if (Zi % 1000 < 1 || Zi % 1000 > 103) fail(MT_ERR_INVALID_ARGUMENT, "atomic number out of range 1..103");
        if (charge != 0) fail(MT_ERR_UNSUPPORTED, "ionic charges");
        in_atoms.Z.push_back(Zi);
        in_atoms.x.push_back((R)A[na + i]);
        in_atoms.y.push_back((R)A[2 * na + i]);
        in_atoms.z.push_back((R)A[3 * na + i]);
        in_atoms.sigma.push_back((float)(nc > 4 ? A[4 * na + i] : 0.085));  // float in Atom_Data
        in_atoms.occ.push_back((float)(nc > 5 ? A[5 * na + i] : 1.0));
        in_atoms.region.push_back(std::abs((int)(nc > 6 ? A[6 * na + i] : 0)));
    }
    const R spec_dz = (R)in.spec_dz;

    // validate_parameters (src/input_multislice.cuh:622-796)
    pn_seed = std::max(1, pn_seed);
    pn_nconf = frozen() ? std::max(1, pn_nconf) : 1;
    fp_iconf_0 = frozen() ? (pn_single_conf ? pn_nconf : 1) : 1;
    if (is_ewfs() || is_ewrs()) pn_coh_contrib = 1;
    // z displacements only with sub-slicing (src/input_multislice.cuh:697-705)
    subslicing = multislice() && in.potential_slicing == 3;
    if (!subslicing) pn_z = false;

    auto zminmax = [](const SpecimenT<R> &s, R &lo, R &hi) {
        auto mm = std::minmax_element(s.z.begin(), s.z.end());
        lo = *mm.first; hi = *mm.second;
    };
    R zmin, zmax;
    zminmax(in_atoms, zmin, zmax);

    // thicknesses matched to planes (Slicing::match_thickness, src/slicing.hpp:53-71)
    // validate_parameters for the (weak) phase object (src/input_multislice.cuh:686-728):
    // slicing by planes, through-slices thicknesses become through-thickness,
    // and a whole-specimen run slices by dz
    int thick_type = in.thick_type;
    int pot_sli = in.potential_slicing;
    if (!multislice()) {
        pot_sli = 1;
        if (thick_type == 3) thick_type = 2;
    }
    // specimen rotation (src/input_multislice.cuh:707-722): the whole specimen,
    // rotated by theta about the unit axis u0 through the geometric centre of
    // the atoms or a given point (rotate_atoms, src/atomic_data_mt.hpp:61-74,
    // get_rotation_matrix, src/cgpu_fcns.cuh:185-204)
    const R rot_theta = (R)(in.spec_rot_theta * kDegToRad);
    if (!near_zero(rot_theta)) {
        thick_type = 1;
        R u[3] = {(R)in.spec_rot_u0[0], (R)in.spec_rot_u0[1], (R)in.spec_rot_u0[2]};
        const R un = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        if (!(un > R(0))) fail(MT_ERR_INVALID_ARGUMENT, "spec_rot_u0 must be a nonzero vector");
        for (R &c : u) c /= un;
        R p0[3] = {(R)in.spec_rot_center_p[0], (R)in.spec_rot_center_p[1], (R)in.spec_rot_center_p[2]};
        if (in.spec_rot_center_type == 1) {
            double sx = 0, sy = 0, sz = 0;
            for (size_t i = 0; i < in_atoms.x.size(); ++i) { sx += in_atoms.x[i]; sy += in_atoms.y[i]; sz += in_atoms.z[i]; }
            const double n = (double)in_atoms.x.size();
            p0[0] = (R)(sx / n); p0[1] = (R)(sy / n); p0[2] = (R)(sz / n);
        }
        R a = R(1) - std::cos(rot_theta), b = std::sin(rot_theta);
        if (near_zero(a)) a = 0;
        if (near_zero(b)) b = 0;
        // column-major 3x3: r'_i = sum_j M[i + 3 j] (r_j - p0_j) + p0_i
        const R M[9] = {R(1) + a * (u[0] * u[0] - R(1)), u[1] * u[0] * a + u[2] * b, u[2] * u[0] * a - u[1] * b,
                        u[0] * u[1] * a - u[2] * b, R(1) + a * (u[1] * u[1] - R(1)), u[2] * u[1] * a + u[0] * b,
                        u[0] * u[2] * a + u[1] * b, u[1] * u[2] * a - u[0] * b, R(1) + a * (u[2] * u[2] - R(1))};
        for (size_t i = 0; i < in_atoms.x.size(); ++i) {
            const R x = in_atoms.x[i] - p0[0], y = in_atoms.y[i] - p0[1], z = in_atoms.z[i] - p0[2];
            in_atoms.x[i] = M[0] * x + M[3] * y + M[6] * z + p0[0];
            in_atoms.y[i] = M[1] * x + M[4] * y + M[7] * z + p0[1];
            in_atoms.z[i] = M[2] * x + M[5] * y + M[8] * z + p0[2];
        }
        zminmax(in_atoms, zmin, zmax);   // get_statistic after the rotation
    }
    // sub-slicing keeps only the whole-specimen exit (src/input_multislice.cuh:697-700)
    if (subslicing && thick_type == 2) thick_type = 1;
    // the phase object's whole specimen is sliced by dz (src/input_multislice.cuh:724-728)
    if (!multislice() && thick_type == 1) pot_sli = 2;
    if (thick_type == 1) {
        thick = {zmax};
    } else {
        if (thick_type != 2 && thick_type != 3) fail(MT_ERR_INVALID_ARGUMENT, "thick_type must be 1..3");
        std::vector<R> zc;
        for (int i = 0; i < na; ++i) if (in_atoms.region[i] == 0) zc.push_back(in_atoms.z[i]);
        std::sort(zc.begin(), zc.end());
        std::vector<R> zp = (pot_sli == 1) ? find_planes(zc) : even_planes(zc.front(), zc.back(), spec_dz);
        unique_sorted(zp);
        if (thick_type == 3) zp = slice_bounds(zp, spec_dz, zmin - R(2.5), zmax + R(2.5), subslicing);
        std::vector<R> req;
        for (int i = 0; i < in.n_thick; ++i) req.push_back((R)in.thick[i]);
        // match_vectors (src/cpu_fcns.hpp:3613-3654): each request snaps to the
        // nearest plane (ties to the lower), each plane used once
        R dv = std::fmax(R(0.25) * spec_dz, R(0.1));
        R lo = *std::min_element(zp.begin(), zp.end()) - dv, hi = *std::max_element(zp.begin(), zp.end()) + dv;
        std::vector<bool> used(zp.size(), false);
        for (R r : req) {
            if (r < lo || r > hi) continue;
            size_t best = 0;
            for (size_t k = 1; k < zp.size(); ++k)
                if (std::fabs(zp[k] - r - R(1e-4)) < std::fabs(zp[best] - r)) best = k;
            if (!used[best]) { used[best] = true; thick.push_back(zp[best]); }
        }
        if (thick.empty()) fail(MT_ERR_INVALID_ARGUMENT, "no requested thickness lies inside the specimen");
    }
    // drop atoms beyond the last thickness
    {
        R z_e = thick.back() + R(0.1);
        if (zmax > z_e) {
            R z_0 = zmin - R(0.1);
            SpecimenT<R> t;
            for (int i = 0; i < (int)in_atoms.Z.size(); ++i) {
                R z = in_atoms.z[i];
                if (z_0 < z && z < z_e) {
                    t.Z.push_back(in_atoms.Z[i]); t.x.push_back(in_atoms.x[i]); t.y.push_back(in_atoms.y[i]);
                    t.z.push_back(z); t.sigma.push_back(in_atoms.sigma[i]); t.occ.push_back(in_atoms.occ[i]);
                    t.region.push_back(in_atoms.region[i]);
                }
            }
            in_atoms = std::move(t);
            zminmax(in_atoms, zmin, zmax);
        }
    }
    cond.zero_defocus_plane = zero_defocus_plane(cond, zmin, zmax);
    cond.zero_defocus_type = 4;

    if (is_stem()) { oa_ix_0 = 0; oa_ix_e = nx; oa_iy_0 = 0; oa_iy_e = ny; }   // validate_output_area
    else {
        oa_ix_0 = in.output_area_ix_0 - 1; oa_ix_e = in.output_area_ix_e - 1;
        oa_iy_0 = in.output_area_iy_0 - 1; oa_iy_e = in.output_area_iy_e - 1;
        if (oa_ix_0 == oa_ix_e) { oa_ix_0 = 0; oa_ix_e = nx; }
        if (oa_iy_0 == oa_iy_e) { oa_iy_0 = 0; oa_iy_e = ny; }
        if (oa_ix_0 > oa_ix_e) std::swap(oa_ix_0, oa_ix_e);
        if (oa_iy_0 > oa_iy_e) std::swap(oa_iy_0, oa_iy_e);
        oa_ix_0 = std::clamp(oa_ix_0, 0, nx); oa_ix_e = std::clamp(oa_ix_e, 0, nx);
        oa_iy_0 = std::clamp(oa_iy_0, 0, ny); oa_iy_e = std::clamp(oa_iy_e, 0, ny);
    }

    // Spec::set_input_data (src/spec.hpp:44-82): periodic de-duplication,
    // stable z sort, tables of the elements present
    const R lx_b = lx - R(1e-4), ly_b = ly - R(1e-4);
    std::vector<int> keep;
    for (int i = 0; i < (int)in_atoms.Z.size(); ++i)
        if (in_atoms.x[i] < lx_b && in_atoms.y[i] < ly_b) keep.push_back(i);
    if (keep.empty()) fail(MT_ERR_INVALID_ARGUMENT, "no atoms inside the simulation box");
    std::stable_sort(keep.begin(), keep.end(), [&](int a, int b) { return in_atoms.z[a] < in_atoms.z[b]; });
    for (int i : keep) {
        atoms_u.Z.push_back(in_atoms.Z[i] % 1000); atoms_u.x.push_back(in_atoms.x[i]); atoms_u.y.push_back(in_atoms.y[i]);
        atoms_u.z.push_back(in_atoms.z[i]); atoms_u.sigma.push_back(in_atoms.sigma[i]); atoms_u.occ.push_back(in_atoms.occ[i]);
        atoms_u.region.push_back(in_atoms.region[i]);
    }
    z_types = atoms_u.Z;
    std::sort(z_types.begin(), z_types.end());
    z_types.erase(std::unique(z_types.begin(), z_types.end()), z_types.end());
    // Vrl and dR_min are T in the reference and widened for the
    // double-precision table builder (Spec::set_input_data, src/spec.hpp:51-55).
    const double dR_min = std::fmin(dRx, dRy);
    const double Vrl = (double)R(0.015);
    for (int Zt : z_types) tables.push_back(build_atom_table<R>(Zt, dR_min, Vrl));
    for (int Zi : atoms_u.Z)
        type_of_atom.push_back((int)(std::lower_bound(z_types.begin(), z_types.end(), Zi) - z_types.begin()));
    R_int_max = 0;
    for (auto &t : tables) R_int_max = std::max(R_int_max, t.R_max_f);

    // planes of the undisplaced crystal
    {
        std::vector<R> zc;
        for (size_t i = 0; i < atoms_u.z.size(); ++i) if (atoms_u.region[i] == 0) zc.push_back(atoms_u.z[i]);
        z_plane = (pot_sli == 1) ? find_planes(zc) : even_planes(zc.front(), zc.back(), spec_dz);
        unique_sorted(z_plane);
    }
    R uzmin, uzmax;
    zminmax(atoms_u, uzmin, uzmax);
    const R s_z_int = (uzmax + R_int_max) - (uzmin - R_int_max);
    // a thin specimen is a phase object (Spec::set_input_data, src/spec.hpp:65-77)
    if (interaction_model == 2 || s_z_int < R(2.0) * spec_dz || (z_plane.size() == 1 && pot_sli == 1)) {
        interaction_model = 2;
        if (thick_type == 3) thick_type = 2;
        subslicing = false;
        pn_z = false;
    }

    // slices and thicknesses (Slicing::calculate, src/slicing.hpp:93-100 / 253-403)
    // the phase object's dz is the specimen's interaction depth (src/spec.hpp:67)
    const R dz_sl = interaction_model == 2 ? s_z_int : spec_dz;
    std::vector<R> zs = slice_bounds(z_plane, dz_sl, uzmin - R_int_max, uzmax + R_int_max, subslicing);
    const std::vector<R> &az = atoms_u.z;
    auto lower = [&](R v) { return (int)(std::lower_bound(az.begin(), az.end(), v) - az.begin()); };
    for (size_t is = 0; is + 1 < zs.size(); ++is) {
        SliceInfoT<R> s{zs[is], zs[is + 1], 1, 0, -1};
        s.z_int_0 = s.z_0;
        s.z_int_e = s.z_e;
        if (subslicing) {
            // atoms within R_int_max of the slice centre (get_slicing, src/slicing.hpp:376-386)
            const R z_m = R(0.5) * (s.z_0 + s.z_e);
            s.z_int_0 = std::fmin(z_m - R_int_max, s.z_0);
            s.z_int_e = std::fmax(z_m + R_int_max, s.z_e);
        }
        atoms_in(az, s.z_int_0, s.z_int_e, subslicing, s.iatom_0, s.iatom_e);
        slices.push_back(s);
    }
    for (R t : thick) {
        ThickInfoT<R> ti{t, 0, R(0.0)};
        if (subslicing && thick_type == 1) {
            ti.islice = (int)zs.size() - 2;   // whole specimen: the last slice (src/slicing.hpp:297-298)
        } else if (thick_type == 3) {
            for (size_t i = 0; i + 1 < zs.size(); ++i) if (std::fabs(zs[i + 1] - t) < R(1e-4)) { ti.islice = (int)i; break; }
        } else {
            for (size_t i = 0; i + 1 < zs.size(); ++i) if (zs[i] < t && t <= zs[i + 1]) { ti.islice = (int)i; break; }
        }
        R ze = zs[ti.islice + 1];
        int iatom_e = (az.back() <= ze) ? (int)az.size() - 1 : lower(ze) - 1;
        if (iatom_e < 0) iatom_e = 0;
        ti.iatom_e = iatom_e;
        R zdp = zero_defocus_plane(obj, az[0], az[iatom_e]);
        if (!multislice()) ti.z_back_prop = zdp - t;   // src/slicing.hpp:326-333
        else if (thick_type == 3) ti.z_back_prop = R(0.0);
        else {
            ti.z_back_prop = zdp - R(0.5) * (zs[ti.islice] + zs[ti.islice + 1]);
            if (std::fabs(ti.z_back_prop) < R(1e-4)) ti.z_back_prop = R(0.0);
        }
        thk.push_back(ti);
    }
    if (!multislice()) {
        // one slice per thickness, from the previous thickness (or the
        // specimen's lower bound) to this one (get_slicing, src/slicing.hpp:341-355)
        slices.clear();
        const R z_int_min = uzmin - R_int_max, z_int_max = uzmax + R_int_max;
        for (size_t k = 0; k < thk.size(); ++k) {
            SliceInfoT<R> s;
            s.z_0 = k == 0 ? z_int_min : thk[k - 1].z;
            s.z_e = thk.size() == 1 ? z_int_max : thk[k].z;
            s.iatom_0 = k == 0 ? 0 : thk[k - 1].iatom_e + 1;
            s.iatom_e = thk[k].iatom_e;
            s.ithk = (int)k;
            s.z_int_0 = s.z_0;
            s.z_int_e = s.z_e;
            slices.push_back(s);
            thk[k].islice = (int)k;
        }
    }
    for (size_t k = 0; k < thk.size(); ++k) slices[thk[k].islice].ithk = (int)k;
}

template <class R>
R ModelT<R>::gx0_rot(int irot) const
{
    return is_ped_hctem() ? std::sin(theta) * std::cos(rot_phi(irot)) / lambda : gx0;
}

template <class R>
R ModelT<R>::gy0_rot(int irot) const
{
    return is_ped_hctem() ? std::sin(theta) * std::sin(rot_phi(irot)) / lambda : gy0;
}

template <class R>
R ModelT<R>::phonon_weight() const
{
    int nconf = (!frozen() || pn_single_conf) ? 1 : pn_nconf;
    return R(1.0) / R(nconf * nrot);
}

// Randn_3d<T, e_host> (src/cgpu_rand.cuh:303-402): a 64-bit Mersenne Twister
// seeded with pn_seed draws three seeds per configuration index; each axis
// then has its own generator and a R normal distribution.
template <class R>
void ModelT<R>::displaced(int iconf, std::vector<R> &x, std::vector<R> &y) const
{
    x = atoms_u.x;
    y = atoms_u.y;
    if (!frozen()) return;
    std::mt19937_64 gen_u((uint64_t)(int64_t)pn_seed);
    std::uniform_int_distribution<int> rand_u(1, INT_MAX);
    unsigned int sx = 0, sy = 0, sz = 0;
    for (int ic = 0; ic < std::max(1, iconf); ++ic) {
        sx = (unsigned int)rand_u(gen_u);
        sy = (unsigned int)rand_u(gen_u);
        sz = (unsigned int)rand_u(gen_u);
    }
    (void)sz;
    std::mt19937_64 gx(sx), gy(sy);
    std::normal_distribution<R> nx_(R(0), R(1)), ny_(R(0), R(1));
    for (size_t i = 0; i < x.size(); ++i) {
        R s = (R)atoms_u.sigma[i];
        R dx = pn_x ? s * nx_(gx) : R(0.0);
        R dy = pn_y ? s * ny_(gy) : R(0.0);
        x[i] = atoms_u.x[i] + dx;
        y[i] = atoms_u.y[i] + dy;
    }
}

template <class R>
void ModelT<R>::atoms_in(const std::vector<R> &z, R z_0, R z_e, bool inc, int &i0, int &ie)
{
    const R eps = sizeof(R) == 4 ? R(1e-4) : R(1e-8);   // Epsilon<T>::rel
    if (inc) { z_0 -= eps; z_e += eps; }
    i0 = 1; ie = 0;
    if (z.empty() || z_0 > z_e || z.back() < z_0 || z_e < z.front()) return;
    auto lower = [&](R v) { return (int)(std::lower_bound(z.begin(), z.end(), v) - z.begin()); };
    const int a = (z_0 <= z.front()) ? 0 : lower(z_0);
    const int b = (z.back() <= z_e) ? (int)z.size() - 1 : lower(z_e) - 1;
    if (!(a > b || z[b] < z_0 || z_e < z[a])) { i0 = a; ie = b; }
}

template <class R>
void ModelT<R>::displaced(int iconf, std::vector<R> &x, std::vector<R> &y, std::vector<R> &z) const
{
    displaced(iconf, x, y);
    z = atoms_u.z;
    if (!frozen() || !pn_z) return;
    // the third seed of the configuration drives z (Randn_3d, src/cgpu_rand.cuh:322-345)
    std::mt19937_64 gen_u((uint64_t)(int64_t)pn_seed);
    std::uniform_int_distribution<int> rand_u(1, INT_MAX);
    unsigned int sz = 0;
    for (int ic = 0; ic < std::max(1, iconf); ++ic) {
        (void)rand_u(gen_u);
        (void)rand_u(gen_u);
        sz = (unsigned int)rand_u(gen_u);
    }
    std::mt19937_64 gz(sz);
    std::normal_distribution<R> nz_(R(0), R(1));
    for (size_t i = 0; i < z.size(); ++i) z[i] = atoms_u.z[i] + (R)atoms_u.sigma[i] * nz_(gz);
}

template <class R>
void ModelT<R>::config_atoms(int iconf, std::vector<int> &order, std::vector<R> &x, std::vector<R> &y,
                             std::vector<R> &z, std::vector<SliceInfo> &sl) const
{
    std::vector<R> dx, dy, dz;
    displaced(iconf, dx, dy, dz);
    order.resize(dz.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    // z displacements re-sort the atoms (Atom_Data::sort_by_z, stable)
    if (frozen() && pn_z) std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return dz[a] < dz[b]; });
    x.resize(order.size()); y.resize(order.size()); z.resize(order.size());
    for (size_t i = 0; i < order.size(); ++i) { x[i] = dx[order[i]]; y[i] = dy[order[i]]; z[i] = dz[order[i]]; }
    sl = slices;
    if (subslicing && frozen() && pn_z)
        for (SliceInfo &s : sl) atoms_in(z, s.z_int_0, s.z_int_e, true, s.iatom_0, s.iatom_e);
}

template <class R>
void ModelT<R>::subslice_rule(std::vector<double> &x, std::vector<double> &w) const
{
    // nw_TanhSinh on [-1 + eps, 1 - eps], eps = Epsilon<float>::eps, limits in R
    const int n = 128;
    const double kHalfPi = 1.570796326794896619231;
    const float eps = 10.0f * 1.19209290e-07f;
    const double x_min = (double)(R)(-1.0 + eps), x_max = (double)(R)(1.0 - eps);
    auto aprox = [&](double v) { return std::asinh(std::atanh(v) / kHalfPi); };
    const double t_min = aprox(x_min), t_max = aprox(x_max);
    const double h = (t_max - t_min) / (n - 1);
    x.resize(n); w.resize(n);
    for (int i = 0; i < n; ++i) {
        const double ti = t_min + i * h, tti = kHalfPi * std::sinh(ti);
        x[i] = std::tanh(tti);
        w[i] = h * kHalfPi * std::cosh(ti) / std::pow(std::cosh(tti), 2);
    }
}

template <class R>
void ModelT<R>::output_dims(mt_output_multislice *out) const
{
    out->n_thk = (int)thk.size();
    if (is_stem()) {
        out->output_type = pn_coh_contrib ? MT_OT_IMAGE_TOT_COH : MT_OT_IMAGE_TOT;
        out->nx = sc_nx; out->ny = sc_ny; out->n_det = (int)det_g2_min.size();
        out->dx = sc_dRx; out->dy = sc_dRy;
        return;
    }
    out->nx = oa_ix_e - oa_ix_0; out->ny = oa_iy_e - oa_iy_0; out->n_det = 0;
    out->dx = is_grid_fs() ? dgx : dRx;
    out->dy = is_grid_fs() ? dgy : dRy;
    // Output_Multislice::set_output_type (src/output_multislice.hpp:1249-1280)
    if (!(is_ewfs() || is_ewrs())) out->output_type = pn_coh_contrib ? MT_OT_M2PSI_TOT_COH : MT_OT_M2PSI_TOT;
    else out->output_type = (!frozen() || pn_single_conf) ? MT_OT_PSI_COH : MT_OT_M2PSI_TOT_PSI_COH;
}

template <class R>
void ModelT<R>::fill_axes(mt_output_multislice *out) const
{
    if (out->thick) for (size_t i = 0; i < thk.size(); ++i) out->thick[i] = thk[i].z;
    if (is_stem()) {
        if (out->x) for (int i = 0; i < sc_nx; ++i) out->x[i] = sc_x0 + i * sc_dRx;
        if (out->y) for (int i = 0; i < sc_ny; ++i) out->y[i] = sc_y0 + i * sc_dRy;
        return;
    }
    bool rs = !is_grid_fs();
    if (out->x) for (int ix = oa_ix_0; ix < oa_ix_e; ++ix) out->x[ix - oa_ix_0] = rs ? ix * dRx : (ix - nxh) * dgx;
    if (out->y) for (int iy = oa_iy_0; iy < oa_iy_e; ++iy) out->y[iy - oa_iy_0] = rs ? iy * dRy : (iy - nyh) * dgy;
}

template <class R>
void ModelT<R>::cond_quadratures(Quad1 &qt, Quad2 &qs) const
{
    const LensModel &L = cond;
    qt = Quad1{};
    qs = Quad2{};
    // defocus spread: Gaussian (weight ti_a, sigma ti_sigma) + exponential
    // (ti_beta), integrated with a Gauss-Hermite rule of the pooled width
    if (L.ti_npts > 1 && !near_zero(L.ti_sigma)) {
        const R s2 = L.ti_sigma * L.ti_sigma, b2 = L.ti_beta * L.ti_beta;
        const double a = (double)L.ti_a / (std::sqrt(kTwoPi) * (double)L.ti_sigma);
        const double b = 0.5 / (double)s2;
        const double c = near_zero(L.ti_beta) ? 0.0 : (1.0 - (double)L.ti_a) / (2.0 * (double)L.ti_beta);
        const double d = near_zero(L.ti_beta) ? 0.0 : 1.0 / (double)L.ti_beta;
        const R st2 = L.ti_a * s2 + R(2) * (R(1) - L.ti_a) * b2;
        const double sigma_t = std::sqrt((double)st2);
        const double b_q = 0.5 / (sigma_t * sigma_t);
        const GaussRule r = gauss_gen_hermite(L.ti_npts, 0.0, 0.0, b_q);
        for (int i = 0; i < L.ti_npts; ++i) {
            const double x = std::fabs(r.x[i]);
            qt.x.push_back(r.x[i]);
            qt.w.push_back(r.w[i] * (a * std::exp(-b * x * x + b_q * x * x) + c * std::exp(-d * x + b_q * x * x)));
        }
    } else {
        qt.x = {0.0};
        qt.w = {1.0};
    }
    // source spread: radial generalized-Laguerre rule x uniform azimuths
    if ((L.si_rad_npts > 1 || L.si_azm_npts > 1) && (!near_zero(L.si_sigma) || !near_zero(L.si_beta))) {
        const R s2 = L.si_sigma * L.si_sigma, b2 = L.si_beta * L.si_beta;
        const double a = near_zero(L.si_sigma) ? 0.0 : (double)L.si_a / (kTwoPi * (double)s2);
        const double b = near_zero(L.si_sigma) ? 0.0 : 0.5 / (double)s2;
        const double c = near_zero(L.si_beta) ? 0.0 : (1.0 - (double)L.si_a) / (kTwoPi * (double)b2);
        const double d = near_zero(L.si_beta) ? 0.0 : 1.0 / (double)L.si_beta;
        const R st2 = (R(2) * L.si_a * s2 + R(6) * (R(1) - L.si_a) * b2) / R(6);
        const double b_q = 1.0 / std::sqrt((double)st2);
        // a one-point rule is (0, 1) whatever the weight (src/quadrature.hpp:37-43)
        GaussRule qr = L.si_rad_npts > 1 ? gauss_gen_laguerre(L.si_rad_npts, 1.0, 0.0, b_q) : GaussRule{{0.0}, {1.0}};
        for (size_t i = 0; i < qr.x.size(); ++i) {
            const double r = std::fabs(qr.x[i]);
            qr.w[i] *= a * std::exp(-b * r * r + b_q * r) + c * std::exp(-d * r + b_q * r);
        }
        const double h = kTwoPi / L.si_azm_npts;
        for (size_t ir = 0; ir < qr.x.size(); ++ir)
            for (int ia = 0; ia < L.si_azm_npts; ++ia) {
                const double th = ia * h;
                qs.x.push_back(qr.x[ir] * std::cos(th));
                qs.y.push_back(qr.x[ir] * std::sin(th));
                qs.w.push_back(qr.w[ir] * h);
            }
    } else {
        qs.x = {0.0};
        qs.y = {0.0};
        qs.w = {1.0};
    }
}

template <class R>
void ModelT<R>::obj_temporal_quadrature(Quad1 &qt) const
{
    // Gauss-Hermite rule in double, stored in R, weights divided by sqrt(pi)
    // in double and stored in R
    const GaussRule r = obj.ti_npts > 1 ? gauss_gen_hermite(obj.ti_npts, 0.0, 0.0, 1.0) : GaussRule{{0.0}, {1.0}};
    qt = Quad1{};
    for (size_t i = 0; i < r.x.size(); ++i) {
        qt.x.push_back((double)(R)r.x[i]);
        qt.w.push_back((double)(R)((double)(R)r.w[i] / kSqrtPi));
    }
}

template class ModelT<float>;
template class ModelT<double>;

} // namespace mtg
