// HIP kernels of the multislice engine (gfx950): projected potential and
// transmission, propagator and probe tables, detector reductions, output
// accumulation, and the launchers.  The slice-pass kernels are in
// slice_passes.hpp; this unit instantiates them for the lengths 2^k and
// hands the lengths 3 x 2^k and 5 x 2^k to slice_passes_m3.hip and
// slice_passes_m5.hip.
#include "slice_passes.hpp"

namespace mtg {

// ---------------------------------------------------------------------------
// Projected potential -> transmission function, gathered per pixel tile.
// V(r) = sum over atoms of the slice of occ * spline(|r - r_a|^2)
// (Projected_Potential, src/projected_potential.cuh:89-182; eval_cubic_poly,
// src/cgpu_fcns.cuh:1496-1566), restricted to each atom's window
// [floor((x-R_max)/dR), ceil((x+R_max)/dR)] (get_bn, src/cpu_fcns.hpp:1669-1684)
// with periodic images; T = exp(i w V) written fft-shifted.
constexpr int kTfTy = 64, kTfTx = 16, kTfThreads = 256;
constexpr int kTfPx = kTfTx * kTfTy / kTfThreads;   // pixels per thread, kTfThreads / kTfTy apart in x
constexpr int kKnots = 128;

// Knot interval of R2 in a table of kKnots increasing knots: the largest
// i in [0, kKnots - 2] with tb[i] <= R2, or 0 -- the result of the
// reference's 7-step binary search (eval_cubic_poly, src/cgpu_fcns.cuh:1496-1566).
// The knots are log-spaced in R (R_min e^(i dlnr)), so log2(R2) is linear in
// i: the guess from the hardware log2 and the table's end knots (lg.x =
// log2 tb[0], lg.y = (kKnots - 1) / (log2 tb[kKnots - 1] - lg.x)) lands
// within one interval of the answer, and one comparison on each side
// settles it.
template <class R>
__device__ __forceinline__ int knot_interval(const R *tb, R R2, float2 lg)
{
    float f = (__builtin_amdgcn_logf((float)R2) - lg.x) * lg.y;
    f = fminf(fmaxf(f, 0.0f), (float)(kKnots - 2));   // also maps log2(0) = -inf to 0
    const int g = (int)f;
    const R lo = tb[g], hi = tb[g + 1];   // both loads issued at once, no branches
    const int dn = (int)(R2 < lo) & (int)(g > 0);
    const int up = (int)(R2 >= hi) & (int)(g < kKnots - 2) & (dn ^ 1);
    return g - dn + up;
}

// SUB: the spline coefficients are per (slice, atom) pair in global memory
// (sub-slicing) instead of the element's, held in LDS.
template <class R, bool SUB>
__global__ __launch_bounds__(kTfThreads) void k_tf_build(TfArgsT<R> a)
{
    // Every product and sum rounds separately, as on the reference's host
    // path: the pixel values then depend only on the summation order.
#pragma clang fp contract(off)
    using R2v = cplx_t<R>;
    using A4 = vec4_t<R>;
    // spline tables of the run's atom types only (dynamic LDS, n_types x 5 x
    // 128 values): the LDS per workgroup, not registers, would otherwise cap
    // the resident workgroups (16 types: three per CU in FP32, one in FP64)
    extern __shared__ __align__(16) unsigned char tf_lds[];
    R *tab = reinterpret_cast<R *>(tf_lds);
    __shared__ R2v rmx[kMaxTypes];
    __shared__ float2 klog[kMaxTypes];    // knot_interval's constants per type
    __shared__ int list[kTfThreads];
    __shared__ int4 lwin[kTfThreads];     // the listed atoms' pixel windows [ix0, ixe] x [iy0, iye]
    __shared__ int2 lbase[kTfThreads];    // and image offsets: the tile's last column / row + offset
                                          // is the smallest image not before the window start
    __shared__ A4 latom[kTfThreads];      // and their records
    __shared__ int wave_cnt[kTfThreads / 64];

    const int tid = threadIdx.x;
    const int ntx = a.nx / kTfTx;
    const int tile = blockIdx.x;
    const int px0 = (tile % ntx) * kTfTx;
    const int py0 = (tile / ntx) * kTfTy;
    const int is = a.slice0 + blockIdx.y;

    for (int i = tid; i < a.n_types * 5 * 128; i += kTfThreads) tab[i] = a.tables[i];
    if (tid < a.n_types) {
        rmx[tid] = a.rmax[tid];
        const float l0 = __builtin_amdgcn_logf((float)a.tables[tid * 5 * kKnots]);
        const float l1 = __builtin_amdgcn_logf((float)a.tables[tid * 5 * kKnots + kKnots - 1]);
        klog[tid] = make_float2(l0, (float)(kKnots - 1) / (l1 - l0));
    }

    const int ty = tid % kTfTy, tx = tid / kTfTy;   // pixels (px0 + tx + 4 h, py), h < kTfPx
    const int py = py0 + ty;
    const bool py_ok = py < a.ny;
    R vacc[kTfPx];
#pragma unroll
    for (int h = 0; h < kTfPx; ++h) vacc[h] = R(0);

    const int2 rng = a.slice_range[is];
    __syncthreads();
    for (int c0 = rng.x; c0 <= rng.y; c0 += kTfThreads) {
        // cull this chunk of atoms against the tile (all periodic images)
        const int ia = c0 + tid;
        bool hit = false;
        A4 at{};
        int ix0 = 0, ixe = 0, iy0 = 0, iye = 0, kx = 0, ky = 0;
        if (ia <= rng.y) {
            at = a.atoms[ia];
            const R R_max = rmx[atom_type(at)].x;
            ix0 = (int)floor((at.x - R_max) / a.dRx); ixe = (int)ceil((at.x + R_max) / a.dRx);
            iy0 = (int)floor((at.y - R_max) / a.dRy); iye = (int)ceil((at.y + R_max) / a.dRy);
            if (ix0 != ixe && iy0 != iye) {
                // smallest image of the tile's last column/row not left of the window start
                auto fdiv = [](int p, int q) { return p >= 0 ? p / q : -((-p + q - 1) / q); };
                kx = -fdiv(-(ix0 - (px0 + kTfTx - 1)), a.nx);   // ceil((ix0 - px_last)/nx)
                ky = -fdiv(-(iy0 - (py0 + kTfTy - 1)), a.ny);
                hit = (px0 + kx * a.nx <= ixe) && (py0 + ky * a.ny <= iye);
            }
        }
        const unsigned long long mask = __ballot(hit);
        const int lane = tid & 63, wave = tid >> 6;
        if (lane == 0) wave_cnt[wave] = __popcll(mask);
        __syncthreads();
        int base = 0, total = 0;
        for (int w = 0; w < kTfThreads / 64; ++w) {
            if (w < wave) base += wave_cnt[w];
            total += wave_cnt[w];
        }
        if (hit) {
            const int pos = base + __popcll(mask & ((1ull << lane) - 1ull));
            list[pos] = ia;
            lwin[pos] = make_int4(ix0, ixe, iy0, iye);
            lbase[pos] = make_int2(kx * a.nx, ky * a.ny);
            latom[pos] = at;
        }
        __syncthreads();
        for (int li = 0; li < total; ++li) {
            const A4 at = latom[li];
            const int4 wn = lwin[li];
            const int2 lb = lbase[li];
            const int ty_ = atom_type(at);
            const R *tb = tab + ty_ * 5 * kKnots;
            const float2 lg = klog[ty_];
            // the knots are the element's; the coefficients c0..c3 the
            // element's, or those of this atom's sub-slice
            const R *cf = SUB ? a.sub_coef + (size_t)(a.pair_base[is] + (list[li] - rng.x)) * 512 : tb + kKnots;
            const R R2_max = rmx[ty_].y;
            // images iy = py + k*ny inside [iy0, iye]: the first is the smallest
            // not below iy0 (py is at most 63 rows above the tile's last row,
            // and ny >= 64, so it is lb.y or lb.y + ny above py)
            int iy_first = py + lb.y;
            if (iy_first < wn.z) iy_first += a.ny;
#pragma unroll
            for (int h = 0; h < kTfPx; ++h) {
                const int px = px0 + tx + (kTfThreads / kTfTy) * h;
                // images ix = px + k*nx inside [ix0, ixe], in increasing order,
                // from the smallest not below ix0 (likewise lb.x or lb.x + nx)
                int ix = px + lb.x;
                if (ix < wn.x) ix += a.nx;
                for (; ix <= wn.y; ix += a.nx) {
                    const R rx = (R)ix * a.dRx - at.x;
                    for (int iy = iy_first; iy <= wn.w; iy += a.ny) {
                        const R ry = (R)iy * a.dRy - at.y;
                        const R R2 = rx * rx + ry * ry;
                        if (R2 < R2_max) {
                            const int i0 = knot_interval(tb, R2, lg);
                            const R dx = R2 - tb[i0];
                            const R p = ((cf[3 * kKnots + i0] * dx + cf[2 * kKnots + i0]) * dx + cf[1 * kKnots + i0]) * dx
                                            + cf[i0];
                            vacc[h] = vacc[h] + at.z * p;
                        }
                    }
                }
            }
        }
        __syncthreads();
    }
    if (!py_ok) return;
    const long long nxy = (long long)a.nx * a.ny;
#pragma unroll
    for (int h = 0; h < kTfPx; ++h) {
        const int px = px0 + tx + (kTfThreads / kTfTy) * h;
        const long long sl = is - a.slice0;
        if (a.V) a.V[sl * nxy + (long long)px * a.ny + py] = vacc[h];
        if (a.trans) {
            // exp(i w V), or 1 + i w V for the weak-phase object (transmission_function,
            // src/cgpu_fcns.cuh:3071-3086)
            R s, c;
            if (a.weak_phase) {
                c = R(1);
                s = a.w * vacc[h];
            } else {
                rsincos(a.w * vacc[h], &s, &c);
            }
            a.trans[sl * nxy + (long long)shift_idx(px, a.nxh) * a.ny + shift_idx(py, a.nyh)] = cmake(c, s);
        }
    }
}

// Lobato 3D potential and its radial derivative over r, times f
// (Vr_dVrir_Lobato_neutral_0_12 with Vr_dVrir's factor, src/cgpu_fcns.cuh:1460-1475,
// 3026-3064): Vr = sum a_i e^(-k_i r) (2/(k_i r) + 1),
// dVrir = -(Vr + sum a_i e^(-k_i r) (k_i r + 1)) / r^2.
__device__ __forceinline__ void vr_dvrir(double z, double R2, const double *P, double f, double &Vr, double &dVrir)
{
    const double r = sqrt(z * z + R2);
    double e[5], kr[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        kr[i] = P[5 + i] * r;
        e[i] = P[i] * exp(-kr[i]);
    }
    Vr = e[0] * (2.0 / kr[0] + 1.0) + e[1] * (2.0 / kr[1] + 1.0) + e[2] * (2.0 / kr[2] + 1.0) +
         e[3] * (2.0 / kr[3] + 1.0) + e[4] * (2.0 / kr[4] + 1.0);
    dVrir = -(Vr + e[0] * (kr[0] + 1.0) + e[1] * (kr[1] + 1.0) + e[2] * (kr[2] + 1.0) + e[3] * (kr[3] + 1.0) +
              e[4] * (kr[4] + 1.0)) / (r * r);
    Vr *= f;
    dVrir *= f;
}

// One workgroup per (slice, atom) pair, one thread per knot R2_k of the
// atom's element (linear_Vz + cubic_poly_coef, src/gpu_fcns.cuh:175-285):
//   V(R2_k)  = integral over z in [z_0, z_e] - z_atom of Vr(sqrt(z^2 + R2_k)),
//   dV(R2_k) = d V / d R^2,
// on the 128-node rule mapped to the range (split at the atom when the
// slice contains it), nodes summed in order; then the cosine taper beyond
// R2_tap and the cubic Hermite coefficients in R^2 between knots.  All in
// FP64, the coefficients cast to R at the end (the reference's FP32 path
// evaluates these in float, and its Hermite coefficients, differences of
// neighbouring float values, then lose three to four digits).
template <class R>
__global__ __launch_bounds__(128) void k_subslice_coef(SubArgsT<R> a)
{
    __shared__ double c0s[128], c1s[128];
    const int p = blockIdx.x, k = threadIdx.x;
    const int2 pr = a.pairs[p];
    const int ty = atom_type(a.atoms[pr.y]);
    const double za = (double)a.z[pr.y];
    const double z0h = 0.5 * ((double)a.slice_z[2 * pr.x] - za), zeh = 0.5 * ((double)a.slice_z[2 * pr.x + 1] - za);
    const bool split = (z0h < 0.0) && (0.0 < zeh);
    const double *P = a.par + ty * 12;
    const double R2 = (double)a.tables[ty * 640 + k];
    const double a1 = split ? -z0h : zeh - z0h, b1 = split ? z0h : zeh + z0h;
    double V = 0.0, dV = 0.0;
    for (int i = 0; i < 128; ++i) {
        const double x = a.qx[i], w = a.qw[i];
        double v, d;
        vr_dvrir(a1 * x + b1, R2, P, a1 * w, v, d);
        if (split) {
            double v2, d2;
            vr_dvrir(zeh * x + zeh, R2, P, zeh * w, v2, d2);
            v += v2;
            d += d2;
        }
        V += v;
        dV += d;
    }
    dV = 0.5 * dV;
    // apply_tapering (src/cgpu_fcns.cuh:1528-1539)
    const double R2_tap = P[10], tap_cf = P[11];
    if (R2_tap < R2) {
        double sn, cs;
        sincos(tap_cf * (R2 - R2_tap), &sn, &cs);
        dV = dV * cs - tap_cf * V * sn;
        V *= cs;
    }
    c0s[k] = V;
    c1s[k] = dV;
    __syncthreads();
    R *out = a.coef + (size_t)p * 512;
    out[k] = (R)V;
    out[128 + k] = (R)dV;
    double c2 = 0.0, c3 = 0.0;
    if (k < 127) {
        // cubic_poly_coef (src/cgpu_fcns.cuh:1544-1555)
        const double idR = 1.0 / ((double)a.tables[ty * 640 + k + 1] - R2);
        const double m = (c0s[k + 1] - V) * idR, n = dV + c1s[k + 1];
        c2 = (3.0 * m - n - dV) * idR;
        c3 = (n - 2.0 * m) * idR * idR;
    }
    out[256 + k] = (R)c2;
    out[384 + k] = (R)c3;
}

// Transmission of stored potentials: T = exp(i w V) (or 1 + i w V), V in the
// natural layout [slice][nx*ny], T written fft-shifted -- the same operations
// as the end of k_tf_build, so T is bit-identical to the one built from the
// atoms (Transmission_Function::trans with a stored potential,
// src/transmission_function.cuh:72-90).
template <class R>
__global__ __launch_bounds__(256) void k_trans_from_pot(const R *V, int nx, int ny, int nxh, int nyh, R w,
                                                        int weak_phase, long long n, cplx_t<R> *trans)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long nxy = (long long)nx * ny;
    const long long sl = i / nxy, p = i - sl * nxy;
    const int px = (int)(p / ny), py = (int)(p - (long long)px * ny);
    const R v = V[i];
    R sn, c;
    if (weak_phase) {
        c = R(1);
        sn = w * v;
    } else {
        rsincos(w * v, &sn, &c);
    }
    trans[sl * nxy + (long long)shift_idx(px, nxh) * ny + shift_idx(py, nyh)] = cmake(c, sn);
}

// ---------------------------------------------------------------------------
// Detector partials -> image (FP64), in fixed order.
__global__ void k_det_reduce(const double *part, int n_members, int n_thk, int n_det, int ntile,
                             const Member *members, double w, double *image, int nscan)
{
    const int m = blockIdx.x;
    const int iscan = members[m].iscan;
    if (iscan < 0) return;
    const double wm = w * members[m].wq;
    for (int k = threadIdx.x; k < n_thk * n_det; k += blockDim.x) {
        const double *p = part + ((long long)m * n_thk * n_det + k) * ntile;
        double s = 0.0;
        for (int i = 0; i < ntile; ++i) s += p[i];
        image[(long long)k * nscan + iscan] += wm * s;
    }
}

// Fresnel propagator factors, separable in gx and gy, with the FP64 phase
// reduced mod 2 pi before the FP32 sincos:
// polar(1/(nx ny), w |g - g0|^2), w = -mul_sign pi lambda z
// (propagate, src/cgpu_fcns.cuh:1835-1845; g0 the beam tilt).
template <class Z>
__global__ void k_prop_tables(int nx, int ny, const double *z, int nz, double lambda, int mul_sign,
                              double lx, double ly, double gx0, double gy0, Z *px, Z *py)
{
    using R = real_t<Z>;
    const int n = nx + ny;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)n * nz) return;
    const int zi = (int)(i / n), j = (int)(i % n);
    const double w = -mul_sign * 3.141592653589793238463 * lambda * z[zi];
    double g;
    if (j < nx) g = signed_freq(j, nx) / lx - gx0;
    else g = signed_freq(j - nx, ny) / ly - gy0;
    double arg = w * g * g;
    arg -= kTwoPiD * rint(arg * (1.0 / kTwoPiD));
    R s, c;
    rsincos((R)arg, &s, &c);
    if (j < nx) {
        const R inv = R(1) / (R)(nx * ny);
        px[(long long)zi * nx + j] = cmake(c * inv, s * inv);
    } else {
        py[(long long)zi * ny + (j - nx)] = cmake(c, s);
    }
}

// Tilt phase ramp exp(-2 pi i (gx0 (Rx - Rx_c) + gy0 (Ry - Ry_c))) of the
// output (Wave_Function::phase_multiplication with -g0, src/wave_function.cuh:77-105;
// exp_r_factor_2d, src/cgpu_fcns.cuh:1867-1877) as separable tables over
// the shifted layout, rx with the extra factor `scale`; FP64 argument
// reduced mod 2 pi before the FP32 sincos.
template <class Z>
__global__ void k_ramp_tables(int nx, int ny, double gx0, double gy0, double dRx, double dRy, double Rx_c,
                              double Ry_c, real_t<Z> scale, Z *rx, Z *ry)
{
    using R = real_t<Z>;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nx + ny) return;
    double arg;
    if (j < nx) arg = -kTwoPiD * gx0 * (shift_idx(j, nx >> 1) * dRx - Rx_c);
    else arg = -kTwoPiD * gy0 * (shift_idx(j - nx, ny >> 1) * dRy - Ry_c);
    arg -= kTwoPiD * rint(arg * (1.0 / kTwoPiD));
    R sn, cs;
    rsincos((R)arg, &sn, &cs);
    if (j < nx) rx[j] = cmake(cs * scale, sn * scale);
    else ry[j - nx] = cmake(cs, sn);
}

template <class R>
__global__ void k_count_aperture(int nx, int ny, R dgx, R dgy, R ux, R uy, R g2_min, R g2_max, unsigned int *count)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    bool in = false;
    if (i < (long long)nx * ny) {
        const int ix = (int)(i / ny), iy = (int)(i % ny);
        const R g2 = g2_grid_shift(signed_freq(ix, nx), signed_freq(iy, ny), dgx, dgy, ux, uy);
        in = (g2_min <= g2 && g2 < g2_max);
    }
    const unsigned long long m = __ballot(in);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(count, (unsigned int)__popcll(m));
}

// Output accumulation with fftshift + crop (add_scale_*_crop_shift_2d,
// src/cgpu_fcns.cuh:1664-1778); members summed in order.
template <class Z>
__global__ void k_out_accum(OutArgsT<Z> a)
{
    const long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= (long long)a.nx_o * a.ny_o) return;
    const int ixo = (int)(o / a.ny_o), iyo = (int)(o % a.ny_o);
    const int sx = shift_idx(ixo + a.ix_0, a.nx >> 1), sy = shift_idx(iyo + a.iy_0, a.ny >> 1);
    const long long si = (long long)sx * a.ny + sy;
    for (int m = 0; m < a.n_members; ++m) {
        const Z v = a.src[(long long)m * a.member_stride + si];
        const double w = a.members ? a.w * a.members[m].wq : a.w;
        if (a.mode & OUT_SET_PSI) {
            a.psi[2 * o] = v.x;
            a.psi[2 * o + 1] = v.y;
        }
        if (a.mode & OUT_ADD_M2) a.m2[o] += w * (double)(v.x * v.x + v.y * v.y);
        if (a.mode & OUT_ADD_PSI) {
            a.psi[2 * o] += w * (double)v.x;
            a.psi[2 * o + 1] += w * (double)v.y;
        }
    }
    if (a.mode & OUT_ADD_PSI_FULL) {
        // natural index o maps to the same pixel of the uncropped wave
        for (long long k = o; k < (long long)a.nx * a.ny; k += (long long)a.nx_o * a.ny_o) {
            for (int m = 0; m < a.n_members; ++m) {
                const Z v = a.src[(long long)m * a.member_stride + k];
                const double w = a.members ? a.w * a.members[m].wq : a.w;
                a.psi_full[2 * k] += w * (double)v.x;
                a.psi_full[2 * k + 1] += w * (double)v.y;
            }
        }
    }
}

template <class R>
__global__ void k_ctf_table(int nx, int ny, CtfParamsT<R> P, cplx_t<R> *ctf)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)nx * ny) return;
    const int ix = (int)(i / ny), iy = (int)(i % ny);
    const int igx = signed_freq(ix, nx), igy = signed_freq(iy, ny);
    const R g2f = g2_grid(igx, igy, P.dgx_f, P.dgy_f);
    cplx_t<R> v = cmake(R(0), R(0));
    if (P.g2_min <= g2f && g2f < P.g2_max) {
        const double gx = igx * P.dgx, gy = igy * P.dgy, g2 = gx * gx + gy * gy;
        double chi, amp = 1.0;
        if (P.mode == 1) {
            chi = (P.cc[0] + (P.cc[1] + P.cc[2] * g2) * g2) * g2;
            if (P.azim) {
                constexpr int pw[11] = {2, 3, 3, 4, 4, 5, 5, 5, 6, 6, 6};
                constexpr int mu[11] = {2, 1, 3, 2, 4, 1, 3, 5, 2, 4, 6};
                const double g1 = sqrt(g2), ph = atan2(gy, gx);
                double gp[7];
                gp[0] = 1.0;
                for (int k = 1; k < 7; ++k) gp[k] = gp[k - 1] * g1;
                if (ph != 0.0) chi += P.az_m * ph;
                for (int k = 0; k < 11; ++k)
                    if (P.az_c[k] != 0.0) chi += P.az_c[k] * gp[pw[k]] * sin(mu[k] * (ph - P.az_phi[k]));
            }
        } else {
            const double pi = 3.141592653589793238463;
            chi = g2 * (P.cc[1] * g2 + P.cc[0]);
            double c = pi * P.sth * P.ti;
            const double u = 1.0 + c * c * g2;
            c = pi * P.ti * P.lambda * g2;
            const double temp_inc = 0.25 * c * c;
            c = pi * P.sth * (P.cc[4] * P.lambda * P.lambda * g2 + P.cc[3]);
            const double spa_inc = c * c * g2;
            amp = exp(-(spa_inc + temp_inc) / u) / sqrt(u);
        }
        chi -= kTwoPiD * rint(chi * (1.0 / kTwoPiD));
        R sn, cs;
        rsincos((R)chi, &sn, &cs);
        v = cmake((R)amp * cs, (R)amp * sn);
    }
    ctf[i] = v;
}

// Coherent STEM image (Wave_Function::set_m2psi_coh, src/wave_function.cuh:188-205):
// one workgroup per (scan point, thickness), FP64 sums reduced in fixed order.
template <class R>
struct CohDet {
    R g2_min[kMaxDet], g2_max[kMaxDet];
};

template <class R>
__global__ __launch_bounds__(256) void k_coh_image(const double2 *coh, int nscan, int n_thk, int n_det, int nx, int ny,
                                                   R dgx, R dgy, CohDet<R> det, const R *fR, int extra_bwl,
                                                   R bwl_alpha, R gl2_max, double *image_coh)
{
    __shared__ double red[kMaxDet][256];
    const int iscan = blockIdx.x % nscan, ithk = blockIdx.x / nscan;
    const double2 *c = coh + ((long long)iscan * n_thk + ithk) * ((long long)nx * ny);
    double acc[kMaxDet];
    for (int d = 0; d < kMaxDet; ++d) acc[d] = 0.0;
    for (long long i = threadIdx.x; i < (long long)nx * ny; i += 256) {
        const int ix = (int)(i / ny), iy = (int)(i % ny);
        const R g2 = g2_grid(signed_freq(ix, nx), signed_freq(iy, ny), dgx, dgy);
        const double2 v = c[i];
        double m2 = v.x * v.x + v.y * v.y;
        if (extra_bwl) {
            const double f = (double)bwl_factor(g2, bwl_alpha, gl2_max);
            m2 *= f * f;
        }
        if (fR) {
            for (int d = 0; d < n_det; ++d) acc[d] += (double)fR[(long long)d * nx * ny + i] * m2;
        } else {
            for (int d = 0; d < n_det; ++d)
                if (det.g2_min[d] <= g2 && g2 < det.g2_max[d]) acc[d] += m2;
        }
    }
    for (int d = 0; d < n_det; ++d) red[d][threadIdx.x] = acc[d];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int d = 0; d < n_det; ++d) red[d][threadIdx.x] += red[d][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < n_det) image_coh[((long long)ithk * n_det + threadIdx.x) * nscan + iscan] = red[threadIdx.x][0];
}

template <class R>
__global__ void k_slice_sum(const R *V, int n_slices, long long nxy, double *out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nxy) return;
    double s = 0.0;
    for (int k = 0; k < n_slices; ++k) s += (double)V[(long long)k * nxy + i];
    out[i] = s;
}

// Streaming copy, the HBM bandwidth reference: one 16-byte element per
// thread and one workgroup per 4 KiB, nontemporal loads and stores (the
// fastest of the copy forms measured on MI355X: plain one-shot, several
// elements per thread, grid-stride loops).
typedef float stream_f4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_stream_copy(const stream_f4 *__restrict__ a, stream_f4 *__restrict__ b,
                                                     long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
}

__global__ void k_m2_from_psi(const double *psi_full, int nx, int ny, int ix_0, int iy_0, int nx_o, int ny_o, double *m2)
{
    const long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= (long long)nx_o * ny_o) return;
    const int ixo = (int)(o / ny_o), iyo = (int)(o % ny_o);
    const long long si = (long long)shift_idx(ixo + ix_0, nx >> 1) * ny + shift_idx(iyo + iy_0, ny >> 1);
    const double re = psi_full[2 * si], im = psi_full[2 * si + 1];
    m2[o] = re * re + im * im;
}

template <class Z>
__global__ void k_shift_complex(const Z *src, Z *dst, int nx, int ny)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)nx * ny) return;
    const int ix = (int)(i / ny), iy = (int)(i % ny);
    dst[i] = src[(long long)shift_idx(ix, nx >> 1) * ny + shift_idx(iy, ny >> 1)];
}

// Natural-layout FP64 complex (ny x nx, interleaved) into the shifted layout
// of the wave type, and back (the staging steps of a propagation whose input
// and output are FP64 host arrays).
template <class Z>
__global__ void k_load_shifted(const double2 *src, Z *dst, int nx, int ny)
{
    using R = real_t<Z>;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)nx * ny) return;
    const int ix = (int)(i / ny), iy = (int)(i % ny);
    const double2 v = src[(long long)shift_idx(ix, nx >> 1) * ny + shift_idx(iy, ny >> 1)];
    dst[i] = cmake((R)v.x, (R)v.y);
}

template <class Z>
__global__ void k_store_unshifted(const Z *src, double2 *dst, int nx, int ny)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)nx * ny) return;
    const int ix = (int)(i / ny), iy = (int)(i % ny);
    const Z v = src[(long long)shift_idx(ix, nx >> 1) * ny + shift_idx(iy, ny >> 1)];
    dst[i] = make_double2((double)v.x, (double)v.y);
}

// ---------------------------------------------------------------------------
// launchers

int col_tile_width(int nx) { return col_variant(nx, 0).tile; }
int row_group(int ny) { return row_variant(ny, 0).tile; }
int col_tile_width_fp64(int nx) { return col_tile0_fp64(nx); }

bool row_variants_apply(int flags) { return flags == kRowSlice; }
bool col_variants_apply(int flags, bool det) { return det ? flags == kColLast : flags == kColSlice; }

template <class Z>
hipError_t launch_row_pass(int n, const RowArgsT<Z> &a, hipStream_t s)
{
    switch (n) {
    case 64: return row_launch<64>(a, s);
    case 128: return row_launch<128>(a, s);
    case 256: return row_launch<256>(a, s);
    case 512: return row_launch<512>(a, s);
    case 1024: return row_launch<1024>(a, s);
    case 2048: return row_launch<2048>(a, s);
    }
    if (n % 3 == 0) return launch_row_pass_m3(n, a, s);
    if (n % 5 == 0) return launch_row_pass_m5(n, a, s);
    return hipErrorInvalidValue;
}

template <class Z>
hipError_t launch_col_pass(int n, const ColArgsT<Z> &a, bool probe, bool det, hipStream_t s)
{
    if (probe && det) return hipErrorInvalidValue;
    switch (n) {
    case 64: return col_launch<64>(a, probe, det, s);
    case 128: return col_launch<128>(a, probe, det, s);
    case 256: return col_launch<256>(a, probe, det, s);
    case 512: return col_launch<512>(a, probe, det, s);
    case 1024: return col_launch<1024>(a, probe, det, s);
    case 2048: return col_launch<2048>(a, probe, det, s);
    }
    if (n % 3 == 0) return launch_col_pass_m3(n, a, probe, det, s);
    if (n % 5 == 0) return launch_col_pass_m5(n, a, probe, det, s);
    return hipErrorInvalidValue;
}

template <class R, bool SUB>
hipError_t tf_build_go(const TfArgsT<R> &a, int tiles, size_t lds, hipStream_t s)
{
    if (lds > 48 * 1024) {
        // allow the dynamic LDS of the largest table set (kMaxTypes types) once
        static const hipError_t allowed = hipFuncSetAttribute(
            reinterpret_cast<const void *>(&k_tf_build<R, SUB>), hipFuncAttributeMaxDynamicSharedMemorySize,
            (int)(kMaxTypes * 5 * kKnots * sizeof(R)));
        if (allowed != hipSuccess) return allowed;
    }
    hipLaunchKernelGGL((k_tf_build<R, SUB>), dim3(tiles, a.n_slices), dim3(kTfThreads), lds, s, a);
    return hipGetLastError();
}

template <class R>
hipError_t launch_tf_build(const TfArgsT<R> &a, hipStream_t s)
{
    // the pixel tile's rows and columns lie within one grid period (the
    // image offsets of k_tf_build rely on it)
    if (a.nx % kTfTx || a.ny < kTfTy || a.n_types > kMaxTypes || a.n_slices <= 0) return hipErrorInvalidValue;
    const int tiles = (a.nx / kTfTx) * ((a.ny + kTfTy - 1) / kTfTy);
    const size_t lds = (size_t)a.n_types * 5 * kKnots * sizeof(R);
    return a.sub_coef ? tf_build_go<R, true>(a, tiles, lds, s) : tf_build_go<R, false>(a, tiles, lds, s);
}

template <class R>
hipError_t launch_subslice_coef(const SubArgsT<R> &a, hipStream_t s)
{
    if (a.n_pairs <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_subslice_coef<R>), dim3(a.n_pairs), dim3(128), 0, s, a);
    return hipGetLastError();
}

template <class R>
hipError_t launch_trans_from_pot(const R *V, int n_slices, int nx, int ny, R w, int weak_phase, cplx_t<R> *trans,
                                 hipStream_t s)
{
    if (n_slices <= 0 || (nx & 1) || (ny & 1)) return hipErrorInvalidValue;
    const long long n = (long long)n_slices * nx * ny;
    hipLaunchKernelGGL((k_trans_from_pot<R>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, V, nx, ny, nx / 2,
                       ny / 2, w, weak_phase, n, trans);
    return hipGetLastError();
}

hipError_t launch_det_reduce(const double *part, int n_members, int n_thk, int n_det, int ntile,
                             const Member *members, double w, double *image, int nscan, hipStream_t s)
{
    hipLaunchKernelGGL(k_det_reduce, dim3(n_members), dim3(64), 0, s, part, n_members, n_thk, n_det, ntile,
                       members, w, image, nscan);
    return hipGetLastError();
}

template <class Z>
hipError_t launch_prop_tables(int nx, int ny, const double *z, int nz, double lambda, int mul_sign,
                              double lx, double ly, double gx0, double gy0, Z *px, Z *py, hipStream_t s)
{
    const long long n = (long long)(nx + ny) * nz;
    hipLaunchKernelGGL((k_prop_tables<Z>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nx, ny, z, nz, lambda,
                       mul_sign, lx, ly, gx0, gy0, px, py);
    return hipGetLastError();
}

template <class Z>
hipError_t launch_ramp_tables(int nx, int ny, double gx0, double gy0, double dRx, double dRy, double Rx_c,
                              double Ry_c, real_t<Z> scale, Z *rx, Z *ry, hipStream_t s)
{
    hipLaunchKernelGGL((k_ramp_tables<Z>), dim3((unsigned)((nx + ny + 255) / 256)), dim3(256), 0, s, nx, ny, gx0, gy0,
                       dRx, dRy, Rx_c, Ry_c, scale, rx, ry);
    return hipGetLastError();
}

template <class R>
hipError_t launch_count_aperture(int nx, int ny, R dgx, R dgy, R ux, R uy, R g2_min, R g2_max, unsigned int *count,
                                 hipStream_t s)
{
    const long long n = (long long)nx * ny;
    hipLaunchKernelGGL((k_count_aperture<R>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nx, ny, dgx, dgy, ux,
                       uy, g2_min, g2_max, count);
    return hipGetLastError();
}

template <class Z>
hipError_t launch_out_accum(const OutArgsT<Z> &a, hipStream_t s)
{
    const long long n = (long long)a.nx_o * a.ny_o;
    hipLaunchKernelGGL((k_out_accum<Z>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_m2_from_psi(const double *psi_full, int nx, int ny, int ix_0, int iy_0, int nx_o, int ny_o,
                              double *m2, hipStream_t s)
{
    const long long n = (long long)nx_o * ny_o;
    hipLaunchKernelGGL(k_m2_from_psi, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, psi_full, nx, ny, ix_0,
                       iy_0, nx_o, ny_o, m2);
    return hipGetLastError();
}

template <class R>
hipError_t launch_coh_image(const double2 *coh, int nscan, int n_thk, int n_det, int nx, int ny, R dgx, R dgy,
                            const R *det_g2_min, const R *det_g2_max, const R *det_fR, int extra_bwl, R bwl_alpha,
                            R gl2_max, double *image_coh, hipStream_t s)
{
    if (n_det > kMaxDet || n_det < 0) return hipErrorInvalidValue;
    CohDet<R> d{};
    for (int k = 0; k < n_det; ++k) { d.g2_min[k] = det_g2_min[k]; d.g2_max[k] = det_g2_max[k]; }
    hipLaunchKernelGGL((k_coh_image<R>), dim3((unsigned)(nscan * n_thk)), dim3(256), 0, s, coh, nscan, n_thk, n_det,
                       nx, ny, dgx, dgy, d, det_fR, extra_bwl, bwl_alpha, gl2_max, image_coh);
    return hipGetLastError();
}

template <class R>
hipError_t launch_ctf_table(int nx, int ny, const CtfParamsT<R> &p, cplx_t<R> *ctf, hipStream_t s)
{
    const long long n = (long long)nx * ny;
    hipLaunchKernelGGL((k_ctf_table<R>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nx, ny, p, ctf);
    return hipGetLastError();
}

template <class R>
hipError_t launch_slice_sum(const R *V, int n_slices, long long nxy, double *out, hipStream_t s)
{
    hipLaunchKernelGGL((k_slice_sum<R>), dim3((unsigned)((nxy + 255) / 256)), dim3(256), 0, s, V, n_slices, nxy, out);
    return hipGetLastError();
}

hipError_t launch_stream_copy(const void *a, void *b, size_t bytes, hipStream_t s)
{
    if (bytes % 16) return hipErrorInvalidValue;
    const long long n = (long long)(bytes / 16);
    const long long blocks = (n + 255) / 256;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)blocks), dim3(256), 0, s, (const stream_f4 *)a, (stream_f4 *)b, n);
    return hipGetLastError();
}

template <class Z>
hipError_t launch_shift_complex(const Z *src, Z *dst, int nx, int ny, hipStream_t s)
{
    const long long n = (long long)nx * ny;
    hipLaunchKernelGGL((k_shift_complex<Z>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, nx, ny);
    return hipGetLastError();
}

template <class Z>
hipError_t launch_load_shifted(const double2 *src, Z *dst, int nx, int ny, hipStream_t s)
{
    if ((nx & 1) || (ny & 1)) return hipErrorInvalidValue;
    const long long n = (long long)nx * ny;
    hipLaunchKernelGGL((k_load_shifted<Z>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, nx, ny);
    return hipGetLastError();
}

template <class Z>
hipError_t launch_store_unshifted(const Z *src, double2 *dst, int nx, int ny, hipStream_t s)
{
    if ((nx & 1) || (ny & 1)) return hipErrorInvalidValue;
    const long long n = (long long)nx * ny;
    hipLaunchKernelGGL((k_store_unshifted<Z>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, nx, ny);
    return hipGetLastError();
}

// both precisions
#define MTG_INSTANTIATE(R, Z)                                                                                          \
    template hipError_t launch_row_pass<Z>(int, const RowArgsT<Z> &, hipStream_t);                                     \
    template hipError_t launch_col_pass<Z>(int, const ColArgsT<Z> &, bool, bool, hipStream_t);                         \
    template hipError_t launch_tf_build<R>(const TfArgsT<R> &, hipStream_t);                                           \
    template hipError_t launch_subslice_coef<R>(const SubArgsT<R> &, hipStream_t);                                     \
    template hipError_t launch_trans_from_pot<R>(const R *, int, int, int, R, int, Z *, hipStream_t);                  \
    template hipError_t launch_prop_tables<Z>(int, int, const double *, int, double, int, double, double, double,      \
                                              double, Z *, Z *, hipStream_t);                                          \
    template hipError_t launch_ramp_tables<Z>(int, int, double, double, double, double, double, double, R, Z *, Z *,   \
                                              hipStream_t);                                                            \
    template hipError_t launch_count_aperture<R>(int, int, R, R, R, R, R, R, unsigned int *, hipStream_t);            \
    template hipError_t launch_out_accum<Z>(const OutArgsT<Z> &, hipStream_t);                                         \
    template hipError_t launch_coh_image<R>(const double2 *, int, int, int, int, int, R, R, const R *, const R *,      \
                                            const R *, int, R, R, double *, hipStream_t);                              \
    template hipError_t launch_ctf_table<R>(int, int, const CtfParamsT<R> &, Z *, hipStream_t);                        \
    template hipError_t launch_slice_sum<R>(const R *, int, long long, double *, hipStream_t);                         \
    template hipError_t launch_shift_complex<Z>(const Z *, Z *, int, int, hipStream_t);                               \
    template hipError_t launch_load_shifted<Z>(const double2 *, Z *, int, int, hipStream_t);                          \
    template hipError_t launch_store_unshifted<Z>(const Z *, double2 *, int, int, hipStream_t);

MTG_INSTANTIATE(float, float2)
MTG_INSTANTIATE(double, double2)
#undef MTG_INSTANTIATE

} // namespace mtg
