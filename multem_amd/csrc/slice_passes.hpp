// Slice-pass kernels (row and column passes of the slice step) and their
// launch templates.  Shared by the translation units that instantiate them
// for different transform lengths: kernels.hip (2^k), slice_passes_m3.hip
// (3 x 2^k) and slice_passes_m5.hip (5 x 2^k), compiled in parallel.  All
// of it has internal linkage, so each unit holds its own instances.
//
// Layout of a wave in HBM: column-major, element (ix, iy) at ix*ny + iy, real
// space fft-shifted (box centre at index 0), reciprocal space in natural FFT
// order -- the conventions of the reference (src/types.cuh:2528-2618).
//
// One slice step is two passes over the wave (see row/col pass below):
//   row pass  (along y, contiguous):  IFFT_y -> x T_s(r) -> FFT_y
//   col pass  (along x, strided):     FFT_x -> x P_s(g) [-> detector sums] -> IFFT_x
// so each probe.slice reads and writes the wave twice and reads T once.
#pragma once

#include "kernels.hpp"
#include "fft_core.hpp"

#include <type_traits>
#include <algorithm>
#include <map>
#include <mutex>
#include <utility>

namespace mtg {

namespace {

constexpr double kTwoPiD = 6.283185307179586476925;

__device__ __forceinline__ int shift_idx(int i, int nh) { return i < nh ? i + nh : i - nh; }
__device__ __forceinline__ int signed_freq(int i, int n) { return i < (n >> 1) ? i : i - n; }

// exp / sincos of the precision in use
__device__ __forceinline__ float rexp(float x) { return expf(x); }
__device__ __forceinline__ double rexp(double x) { return exp(x); }
__device__ __forceinline__ void rsincos(float x, float *s, float *c) { sincosf(x, s, c); }
__device__ __forceinline__ void rsincos(double x, double *s, double *c) { sincos(x, s, c); }

// g^2 on the grid exactly as the reference rounds it in precision R (no
// contraction): Grid_2d::g2_shift (src/types.cuh:2549-2552)
template <class R>
__device__ __forceinline__ R g2_grid(int igx, int igy, R dgx, R dgy)
{
#pragma clang fp contract(off)
    const R gx = (R)igx * dgx, gy = (R)igy * dgy;
    return gx * gx + gy * gy;
}

// |g + u|^2 with the rounding of exp_i_chi (src/cgpu_fcns.cuh:1949-1956):
// (igx dgx + ux)^2 + (igy dgy + uy)^2
template <class R>
__device__ __forceinline__ R g2_grid_shift(int igx, int igy, R dgx, R dgy, R ux, R uy)
{
#pragma clang fp contract(off)
    const R gx = (R)igx * dgx + ux, gy = (R)igy * dgy + uy;
    return gx * gx + gy * gy;
}

// bandwidth-limit factor 1/(1+exp(alpha (g^2 - gl2_max))) (Grid_2d::bwl_factor_shift)
template <class R>
__device__ __forceinline__ R bwl_factor(R g2, R alpha, R gl2_max)
{
    return R(1) / (R(1) + rexp(alpha * (g2 - gl2_max)));
}

// atom type stored in the fourth component of an atom record
__device__ __forceinline__ int atom_type(float4 a) { return __float_as_int(a.w); }
__device__ __forceinline__ int atom_type(double4 a) { return (int)a.w; }

// blockIdx -> (row group, member) such that one row group of every member
// runs on the same XCD (blockIdx % 8): the transmission rows it multiplies
// by stay in that XCD's L2 across the batch.
__device__ __forceinline__ void map_rows(int wg, int G, int M, int &rg, int &member)
{
    if ((G & 7) == 0) {
        const int xcd = wg & 7, k = wg >> 3;
        rg = (k / M) * 8 + xcd;
        member = k % M;
    } else {
        rg = wg % G;
        member = wg / G;
    }
}


// ---------------------------------------------------------------------------
// Row pass: RW rows of length N per workgroup, T = N/E threads per row.
// S: radix schedule; F: compile-time RowFlags of the hot paths, or -1 for
// a.flags at run time; TWT: twiddles from the table (1) or in registers (0).
template <class S, int RW, int F, int TWT, class Z>
__global__ __launch_bounds__(RW *S::T) void k_row_pass(RowArgsT<Z> a)
{
    constexpr int N = S::N, E = S::E, T = S::T;
    using TW = std::conditional_t<TWT != 0, TwTable<S, Z>, TwReg<S, Z>>;
    using R = real_t<Z>;
    // one pad element per 2^PSH (bank conflicts of the stage exchanges):
    // every 8 up to 512, every 32 from 1024 (4 % faster there, measured)
    constexpr int PSH = N >= 1024 ? 5 : 3;
    constexpr int PADN = N + (N >> PSH);
    __shared__ Z lds[RW * PADN];
    const int flags = (F < 0) ? a.flags : F;

    const int tid = threadIdx.x;
    const int t = tid % T, r = tid / T;
    int rg, member;
    map_rows(blockIdx.x, a.nrows / RW, a.n_members, rg, member);
    const int row = rg * RW + r;
    Z *lp = lds + r * PADN;
    auto idx = [](int i) { return i + (i >> PSH); };
    auto sync = [] { __syncthreads(); };

    const long long off = (long long)member * a.member_stride + (long long)row * (a.pitch ? a.pitch : N);
    Z v[E], tv[E];
    // issue the wave and transmission loads together, then the twiddles
    if (flags & ROW_LOAD) {
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = a.src[off + t + T * e];
    } else {
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = cmake(R(1), R(0));
    }
    if (flags & ROW_MUL_T) {
        const int slot = a.members[member].tslot;
        const Z *T_s = a.trans + (long long)slot * a.tslot_stride + (long long)a.slice * a.nrows * N
                            + (long long)row * N;
#pragma unroll
        for (int e = 0; e < E; ++e) tv[e] = T_s[t + T * e];
    }
    TW tw;
    tw.load(t, a.tw);
    if (flags & ROW_IFFT) fft_run<S, true>(v, t, tw, lp, idx, sync);
    if (flags & ROW_MUL_T) {
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = cmul(tv[e], v[e]);
    }
    if (flags & ROW_MUL_RAMP) {
        const Z *rx = a.ramp_x, *ry = a.ramp_y;
        if (a.rot_x) {
            const int ir = a.members[member].irot;
            rx += ir * a.rot_x;
            ry += ir * a.rot_y;
        }
        const Z rxv = rx[row];
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = cmul(cmul(rxv, ry[t + T * e]), v[e]);
    }
    if (flags & ROW_FFT) fft_run<S, false>(v, t, tw, lp, idx, sync);
    if (flags & ROW_STORE) {
#pragma unroll
        for (int e = 0; e < E; ++e) a.dst[off + t + T * e] = v[e];
    }
}

// ---------------------------------------------------------------------------
// Column pass: C adjacent columns of length N per workgroup.  Thread (c, t)
// holds elements ix = t + T*e of column tile*C + c.
// S, F, TWT as for the row pass.
template <class S, int C, int F, bool PROBE, bool DET, int TWT, class Z>
__global__ __launch_bounds__(C *S::T) void k_col_pass(ColArgsT<Z> a)
{
    constexpr int N = S::N, E = S::E, T = S::T;
    using TW = std::conditional_t<TWT != 0, TwTable<S, Z>, TwReg<S, Z>>;
    using R = real_t<Z>;
    constexpr int NWAVE = (C * T + 63) / 64;
    __shared__ Z lds[N * C];
    __shared__ double red[DET ? NWAVE * kMaxDet : 1];
    const int flags = (F < 0) ? a.flags : F;

    const int tid = threadIdx.x;
    const int c = tid % C, t = tid / C;
    const int ntile = a.ncols / C;
    const int tile = blockIdx.x % ntile, member = blockIdx.x / ntile;
    const int col = tile * C + c;
    auto idx = [c](int i) { return i * C + c; };
    auto sync = [] { __syncthreads(); };
    const long long base = (long long)member * a.member_stride + col;
    const long long pitch = a.pitch ? a.pitch : a.ncols;   // x stride of the wave buffers
    const int igy = signed_freq(col, a.ny);

    Z v[E];
    if constexpr (!PROBE) {
        if (flags & COL_LOAD) {
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = a.src[base + (long long)(t + T * e) * pitch];
        }
    }
    TW tw;
    tw.load(t, a.tw);
    if constexpr (PROBE) {
      if (flags & COL_USER_WAVE) {
        // user-defined incident wave moved to the first beam's position:
        // fpsi0(g) exp(i (x gx + y gy)) / (nx ny) (exp_g_factor_2d,
        // src/cgpu_fcns.cuh:1920-1927), the phase in FP64 reduced mod 2 pi
        const Member m = a.members[member];
        const ProbeParamsT<R> &P = a.probe;
        const Beam bm = a.beams[m.beam0];
        const double gy = igy * P.dgy;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int ix = t + T * e;
            const double gx = signed_freq(ix, a.nx) * P.dgx;
            double ph = bm.xph * gx + bm.yph * gy;
            ph -= kTwoPiD * rint(ph * (1.0 / kTwoPiD));
            R s, co;
            rsincos((R)ph, &s, &co);
            v[e] = cscale(cmul(cmake(co, s), a.fpsi0[(long long)ix * a.ncols + col]), a.inv_nxy);
        }
      } else {
        // Convergent probe in Fourier space (exp_i_chi, src/cgpu_fcns.cuh:1949-1979):
        // aperture mask in R, FP64 phase reduced mod 2 pi, sincos in R.
        const Member m = a.members[member];
        const ProbeParamsT<R> &P = a.probe;
        const double gy = igy * P.dgy + P.gyu;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int ix = t + T * e;
            const int igx = signed_freq(ix, a.nx);
            const R g2f = g2_grid_shift(igx, igy, P.dgx_f, P.dgy_f, P.gxu_f, P.gyu_f);
            Z acc = cmake(R(0), R(0));
            if (P.g2_min <= g2f && g2f < P.g2_max) {
                const double gx = igx * P.dgx + P.gxu;
                const double g2 = gx * gx + gy * gy;
                const double c35 = (P.c30 + P.c50 * g2) * g2;
                double aber = 0.0;   // azimuthal terms; the defocus term is per beam
                if (P.azim) {
                    // powers of |g| and multiplicities of the 11 azimuthal terms
                    constexpr int pw[11] = {2, 3, 3, 4, 4, 5, 5, 5, 6, 6, 6};
                    constexpr int mu[11] = {2, 1, 3, 2, 4, 1, 3, 5, 2, 4, 6};
                    const double g1 = sqrt(g2), ph = atan2(gy, gx);
                    double gp[7];
                    gp[0] = 1.0;
                    for (int k = 1; k < 7; ++k) gp[k] = gp[k - 1] * g1;
                    if (ph != 0.0) aber += P.az_m * ph;
                    for (int k = 0; k < 11; ++k)
                        if (P.az_c[k] != 0.0) aber += P.az_c[k] * gp[pw[k]] * sin(mu[k] * (ph - P.az_phi[k]));
                }
                for (int b = 0; b < m.nbeam; ++b) {
                    const Beam bm = a.beams[m.beam0 + b];
                    double chi = bm.xph * gx + bm.yph * gy + (bm.c10 + c35) * g2 + aber;
                    chi -= kTwoPiD * rint(chi * (1.0 / kTwoPiD));
                    R s, co;
                    rsincos((R)chi, &s, &co);
                    acc.x += co * P.norm;
                    acc.y += s * P.norm;
                }
            }
            v[e] = acc;
        }
      }
    }
    if (flags & COL_FFT) fft_run<S, false>(v, t, tw, lds, idx, sync);
    if (flags & (COL_MUL_P | COL_BWL)) {
        // the propagator factors are tiny tables, cache-resident across the batch
        const Z *px = a.px, *py = a.py;
        if (flags & COL_ROT) {   // never set in the specialised slice steps
            const int ir = a.members[member].irot;
            px += ir * a.rot_x;
            py += ir * a.rot_y;
        }
        const Z pyv = (flags & COL_MUL_P) ? cscale(py[col], a.p_scale) : cmake(a.inv_nxy, R(0));
#pragma unroll
        for (int e = 0; e < E; ++e) {
            Z p = (flags & COL_MUL_P) ? cmul(px[t + T * e], pyv) : pyv;
            if (flags & COL_BWL) {
                const R f = bwl_factor(g2_grid(signed_freq(t + T * e, a.nx), igy, a.dgx_f, a.dgy_f),
                                           a.bwl_alpha, a.gl2_max);
                p = cscale(p, f);
            }
            v[e] = cmul(p, v[e]);
        }
    }
    if (flags & COL_MUL_CTF) {
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = cmul(a.ctf[(long long)(t + T * e) * a.ncols + col], v[e]);
    }
    if constexpr (DET) {
        if (flags & COL_COH) {
            // coherent sums over configurations (Output_Multislice::add_scale_psi_coh)
            const int iscan = a.members[member].iscan;
            if (iscan >= 0) {
                double2 *cp = a.coh + ((long long)iscan * a.n_thk + a.ithk) * ((long long)a.nx * a.ny) + col;
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    double2 q = cp[(long long)(t + T * e) * a.ncols];
                    q.x += a.coh_w * (double)v[e].x;
                    q.y += a.coh_w * (double)v[e].y;
                    cp[(long long)(t + T * e) * a.ncols] = q;
                }
            }
        }
        // Annular detector sums (sum_square_over_Det, src/cgpu_fcns.cuh:1794-1805):
        // FP64 per-thread sums, wave and workgroup reduction in fixed order.
        double acc[kMaxDet];
#pragma unroll
        for (int d = 0; d < kMaxDet; ++d) acc[d] = 0.0;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int ix = t + T * e;
            const R g2f = g2_grid(signed_freq(ix, a.nx), igy, a.dgx_f, a.dgy_f);
            R m2 = cnorm2(v[e]);
            if (flags & COL_DET_BWL) {
                const R f = bwl_factor(g2f, a.bwl_alpha, a.gl2_max);
                m2 *= f * f;
            }
            if (flags & COL_DET_FR) {   // never set in the specialised detector passes
                // matrix detectors: sum fR |psi|^2 (sum_square_over_Det with fR,
                // src/wave_function.cuh:125-129)
                const long long px = (long long)ix * a.ncols + col;
                for (int d = 0; d < a.n_det; ++d) acc[d] += (double)(a.det_fR[(long long)d * a.nx * a.ny + px] * m2);
            } else {
                for (int d = 0; d < a.n_det; ++d)
                    if (a.det_g2_min[d] <= g2f && g2f < a.det_g2_max[d]) acc[d] += (double)m2;
            }
        }
        const int lane = tid & 63, wave = tid >> 6;
        for (int d = 0; d < a.n_det; ++d) {
            double s = acc[d];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
            if (lane == 0) red[wave * kMaxDet + d] = s;
        }
        __syncthreads();
        if (tid < a.n_det) {
            double s = 0.0;
            for (int w = 0; w < NWAVE; ++w) s += red[w * kMaxDet + tid];
            a.det_part[(((long long)member * a.n_thk + a.ithk) * a.n_det + tid) * ntile + tile] = s;
        }
    }
    if (flags & COL_IFFT) fft_run<S, true>(v, t, tw, lds, idx, sync);
    if (flags & COL_STORE) {
#pragma unroll
        for (int e = 0; e < E; ++e) a.dst[base + (long long)(t + T * e) * pitch] = v[e];
    }
}

// ---------------------------------------------------------------------------
// Persistent slice steps: one workgroup per resident slot loops over the
// tiles w = blockIdx.x, blockIdx.x + gridDim.x, ...  and issues the loads of
// its next tile before transforming the current one, so HBM traffic of one
// tile overlaps the FFT work of the previous one.  The arithmetic per tile is
// that of k_row_pass / k_col_pass with the full slice-step flags, operation
// for operation, so the waves are bit-identical to the variant with the same
// schedule.  gridDim.x is a multiple of 8 (or the tile count), so tile w
// runs on XCD w % 8 as in the one-tile-per-workgroup launch.
template <class S, int RW, int TWT, class Z>
__global__ __launch_bounds__(RW *S::T) void k_row_slice_pp(RowArgsT<Z> a)
{
    constexpr int N = S::N, E = S::E, T = S::T;
    using TW = std::conditional_t<TWT != 0, TwTable<S, Z>, TwReg<S, Z>>;
    // one pad element per 2^PSH (bank conflicts of the stage exchanges):
    // every 8 up to 512, every 32 from 1024 (4 % faster there, measured)
    constexpr int PSH = N >= 1024 ? 5 : 3;
    constexpr int PADN = N + (N >> PSH);
    __shared__ Z lds[RW * PADN];
    const int tid = threadIdx.x;
    const int t = tid % T, r = tid / T;
    const int G = a.nrows / RW, n_items = G * a.n_members;
    Z *lp = lds + r * PADN;
    auto idx = [](int i) { return i + (i >> PSH); };
    auto sync = [] { __syncthreads(); };
    Z nv[E], ntv[E];
    long long noff = 0;
    // loads of tile `item` into nv (wave) and ntv (transmission rows)
    auto prefetch = [&](int item) {
        int rg, member;
        map_rows(item, G, a.n_members, rg, member);
        const int row = rg * RW + r;
        noff = (long long)member * a.member_stride + (long long)row * (a.pitch ? a.pitch : N);
        const Z *T_s = a.trans + (long long)a.members[member].tslot * a.tslot_stride +
                       (long long)a.slice * a.nrows * N + (long long)row * N;
#pragma unroll
        for (int e = 0; e < E; ++e) nv[e] = a.src[noff + t + T * e];
#pragma unroll
        for (int e = 0; e < E; ++e) ntv[e] = T_s[t + T * e];
    };

    TW tw;
    tw.load(t, a.tw);
    int item = blockIdx.x;
    if (item < n_items) prefetch(item);
    // everything loaded so far has arrived: inside the loop the only loads in
    // flight are the next tile's, and waits for them stay at the loop top
    __builtin_amdgcn_s_waitcnt(0);
    for (; item < n_items; item += gridDim.x) {
        Z v[E], tv[E];
#pragma unroll
        for (int e = 0; e < E; ++e) { v[e] = nv[e]; tv[e] = ntv[e]; }
        const long long off = noff;
        if (item + (int)gridDim.x < n_items) prefetch(item + gridDim.x);
        // keep the next tile's loads ahead of this tile's transforms
        __builtin_amdgcn_sched_barrier(0);
        fft_run<S, true>(v, t, tw, lp, idx, sync);
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = cmul(tv[e], v[e]);
        fft_run<S, false>(v, t, tw, lp, idx, sync);
#pragma unroll
        for (int e = 0; e < E; ++e) a.dst[off + t + T * e] = v[e];
    }
}

template <class S, int C, int TWT, class Z>
__global__ __launch_bounds__(C *S::T) void k_col_slice_pp(ColArgsT<Z> a)
{
    constexpr int N = S::N, E = S::E, T = S::T;
    using TW = std::conditional_t<TWT != 0, TwTable<S, Z>, TwReg<S, Z>>;
    __shared__ Z lds[N * C];
    const int tid = threadIdx.x;
    const int c = tid % C, t = tid / C;
    const int ntile = a.ncols / C, n_items = ntile * a.n_members;
    const long long pitch = a.pitch ? a.pitch : a.ncols;   // x stride of the wave buffers
    auto idx = [c](int i) { return i * C + c; };
    auto sync = [] { __syncthreads(); };
    Z nv[E], npy;
    long long nbase = 0;
    // loads of tile `item` into nv (wave columns) and npy (propagator factor of the column)
    auto prefetch = [&](int item) {
        const int col = (item % ntile) * C + c;
        nbase = (long long)(item / ntile) * a.member_stride + col;
#pragma unroll
        for (int e = 0; e < E; ++e) nv[e] = a.src[nbase + (long long)(t + T * e) * pitch];
        npy = a.py[col];
    };

    // the x factors of the propagator are the same for every tile of this thread
    Z pxv[E];
#pragma unroll
    for (int e = 0; e < E; ++e) pxv[e] = a.px[t + T * e];
    TW tw;
    tw.load(t, a.tw);
    int item = blockIdx.x;
    if (item < n_items) prefetch(item);
    // everything loaded so far has arrived: inside the loop the only loads in
    // flight are the next tile's, and waits for them stay at the loop top
    __builtin_amdgcn_s_waitcnt(0);
    for (; item < n_items; item += gridDim.x) {
        Z v[E];
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = nv[e];
        const Z pyv = cscale(npy, a.p_scale);
        const long long base = nbase;
        if (item + (int)gridDim.x < n_items) prefetch(item + gridDim.x);
        // keep the next tile's loads ahead of this tile's transforms
        __builtin_amdgcn_sched_barrier(0);
        fft_run<S, false>(v, t, tw, lds, idx, sync);
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = cmul(cmul(pxv[e], pyv), v[e]);
        fft_run<S, true>(v, t, tw, lds, idx, sync);
#pragma unroll
        for (int e = 0; e < E; ++e) a.dst[base + (long long)(t + T * e) * pitch] = v[e];
    }
}

constexpr int kRowSlice = ROW_LOAD | ROW_IFFT | ROW_MUL_T | ROW_FFT | ROW_STORE;
constexpr int kRowStart = ROW_MUL_T | ROW_FFT | ROW_STORE;
constexpr int kColSlice = COL_LOAD | COL_FFT | COL_MUL_P | COL_IFFT | COL_STORE;
constexpr int kColLast = COL_LOAD | COL_FFT | COL_MUL_P;
constexpr int kColProbe = COL_IFFT | COL_STORE;

template <class Z> constexpr bool kFp64 = std::is_same<Z, double2>::value;

// Workgroups of a persistent launch: the resident slots of the current
// device (occupancy of `kernel` at `threads` per workgroup times the CU
// count), rounded down to a multiple of 8 so that a workgroup stays on one
// XCD's tiles, and at most the tile count.  Cached per kernel and device.
int persistent_grid(const void *kernel, int threads, int tiles)
{
    static std::mutex mu;
    static std::map<std::pair<const void *, int>, int> slots;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return tiles;
    int n = 0;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = slots.find({kernel, dev});
        if (it != slots.end()) {
            n = it->second;
        } else {
            int per_cu = 0, cus = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess ||
                hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
                (void)hipGetLastError();
                return tiles;
            }
            n = std::max(per_cu, 1) * std::max(cus, 1);
            n = n >= 8 ? n / 8 * 8 : n;
            slots[{kernel, dev}] = n;
        }
    }
    return std::min(n, tiles);
}

// FP64 launch variants.  The FP64 passes always read twiddles from the
// table (the double registers leave no room for register twiddles), so the
// variants that differ from another only in where twiddles live are left
// out; column tiles are the FP32 variant's scaled to the FP64 default width,
// and every double LDS tile stays within a CU's 160 KiB.
constexpr bool fp64_twiddle_twin(int v) { return v == 1 || v == 4 || v == 6; }
constexpr int col_tile_fp64(int n, int v) { return col_tile0_fp64(n) * col_variant(n, v).tile / col_tile0(n); }
constexpr bool fp64_row_ok(int n, int v)
{
    return !fp64_twiddle_twin(v) && row_variant(n, v).tile * (n + n / 8) * 16 <= kLdsBytes;
}
constexpr bool fp64_col_ok(int n, int v)
{
    const int c = col_tile_fp64(n, v);
    return !fp64_twiddle_twin(v) && c >= 1 && n * c * 16 <= kLdsBytes &&
           c * (n / sched_elems(n, col_variant(n, v).sched)) <= 1024;
}

// variant V of the row pass at length N with compile-time flags F
template <int N, int V, int F, class Z>
hipError_t row_go(const RowArgsT<Z> &a, hipStream_t s)
{
    if constexpr (kFp64<Z> && !fp64_row_ok(N, V)) {
        return hipErrorInvalidValue;
    } else {
        constexpr PassVariant pv = row_variant(N, V);
        using S = Sched<N, pv.sched>;
        static_assert(S::E == sched_elems(N, pv.sched) && pv.tile * S::T <= 1024, "row variant");
        if (a.nrows % pv.tile) return hipErrorInvalidValue;
        const int grid = (a.nrows / pv.tile) * a.n_members;
        constexpr int twt = kFp64<Z> ? 1 : pv.tw_table;
        if constexpr (pv.persist != 0) {
            if constexpr (F != kRowSlice) {
                return hipErrorInvalidValue;
            } else {
                auto k = k_row_slice_pp<S, pv.tile, twt, Z>;
                hipLaunchKernelGGL(k, dim3(persistent_grid((const void *)k, pv.tile * S::T, grid)),
                                   dim3(pv.tile * S::T), 0, s, a);
            }
        } else {
            hipLaunchKernelGGL((k_row_pass<S, pv.tile, F, twt, Z>), dim3(grid), dim3(pv.tile * S::T), 0, s, a);
        }
        return hipGetLastError();
    }
}

template <int N, int F, class Z, int... V>
hipError_t row_pick(const RowArgsT<Z> &a, hipStream_t s, std::integer_sequence<int, V...>)
{
    hipError_t e = hipErrorInvalidValue;
    ((a.variant == V ? (e = row_go<N, V, F>(a, s), 0) : 0), ...);
    return e;
}

// Only the full slice step has launch variants; any other flag set runs
// variant 0 whatever a.variant holds.
template <int N, class Z>
hipError_t row_launch(const RowArgsT<Z> &a, hipStream_t s)
{
    if (a.flags == kRowSlice)
        return row_pick<N, kRowSlice>(a, s, std::make_integer_sequence<int, pass_variants(N)>{});
    if (a.flags == kRowStart) return row_go<N, 0, kRowStart>(a, s);
    return row_go<N, 0, -1>(a, s);
}

template <int N, int V, int F, bool PROBE, bool DET, class Z>
hipError_t col_go(const ColArgsT<Z> &a, hipStream_t s)
{
    if constexpr (kFp64<Z>) {
        // FP64: table twiddles, columns narrowed so the double LDS tile fits
        // (col_tile_fp64); the detector pass has variant 0 only
        constexpr PassVariant pv = col_variant(N, V);
        using S = Sched<N, pv.sched>;
        constexpr int C = col_tile_fp64(N, V);
        if constexpr (!fp64_col_ok(N, V) || (DET && V != 0) || (pv.persist != 0 && (F != kColSlice || PROBE))) {
            return hipErrorInvalidValue;
        } else {
            static_assert(C * S::T <= 1024 && N * C * 16 <= kLdsBytes, "FP64 column tile");
            if (a.ncols % C) return hipErrorInvalidValue;
            const int grid = (a.ncols / C) * a.n_members;
            if constexpr (pv.persist != 0) {
                auto k = k_col_slice_pp<S, C, 1, Z>;
                hipLaunchKernelGGL(k, dim3(persistent_grid((const void *)k, C * S::T, grid)), dim3(C * S::T), 0, s,
                                   a);
            } else {
                hipLaunchKernelGGL((k_col_pass<S, C, F, PROBE, DET, 1, Z>), dim3(grid), dim3(C * S::T), 0, s, a);
            }
            return hipGetLastError();
        }
    } else {
        constexpr PassVariant pv = col_variant(N, V);
        using S = Sched<N, pv.sched>;
        static_assert(S::E == sched_elems(N, pv.sched) && pv.tile * S::T <= 1024, "column variant");
        if constexpr (DET && (pv.tile != col_tile0(N) || pv.persist)) {
            return hipErrorInvalidValue;   // the detector partials assume the default width
        } else if constexpr (pv.persist != 0 && (F != kColSlice || PROBE)) {
            return hipErrorInvalidValue;   // persistent workgroups run the full slice step only
        } else {
            if (a.ncols % pv.tile) return hipErrorInvalidValue;
            const int grid = (a.ncols / pv.tile) * a.n_members;
            if constexpr (pv.persist != 0) {
                auto k = k_col_slice_pp<S, pv.tile, pv.tw_table, Z>;
                hipLaunchKernelGGL(k, dim3(persistent_grid((const void *)k, pv.tile * S::T, grid)),
                                   dim3(pv.tile * S::T), 0, s, a);
            } else {
                hipLaunchKernelGGL((k_col_pass<S, pv.tile, F, PROBE, DET, pv.tw_table, Z>), dim3(grid),
                                   dim3(pv.tile * S::T), 0, s, a);
            }
            return hipGetLastError();
        }
    }
}

template <int N, int F, bool DET, class Z, int... V>
hipError_t col_pick(const ColArgsT<Z> &a, hipStream_t s, std::integer_sequence<int, V...>)
{
    hipError_t e = hipErrorInvalidValue;
    ((a.variant == V ? (e = col_go<N, V, F, false, DET>(a, s), 0) : 0), ...);
    return e;
}

template <int N, class Z>
hipError_t col_launch(const ColArgsT<Z> &a, bool probe, bool det, hipStream_t s)
{
    constexpr auto all = std::make_integer_sequence<int, pass_variants(N)>{};
    if (probe) {
        if (a.flags == kColProbe) return col_go<N, 0, kColProbe, true, false>(a, s);
        return col_go<N, 0, -1, true, false>(a, s);
    }
    if (det) {
        if (a.flags == kColLast) return col_pick<N, kColLast, true>(a, s, all);
        if (a.flags == kColSlice) return col_go<N, 0, kColSlice, false, true>(a, s);
        return col_go<N, 0, -1, false, true>(a, s);
    }
    if (a.flags == kColSlice) return col_pick<N, kColSlice, false>(a, s, all);
    return col_go<N, 0, -1, false, false>(a, s);
}


} // namespace

// slice passes at lengths 3 x 2^k and 5 x 2^k (hipErrorInvalidValue for any
// other length)
template <class Z>
hipError_t launch_row_pass_m3(int n, const RowArgsT<Z> &a, hipStream_t s);
template <class Z>
hipError_t launch_col_pass_m3(int n, const ColArgsT<Z> &a, bool probe, bool det, hipStream_t s);
template <class Z>
hipError_t launch_row_pass_m5(int n, const RowArgsT<Z> &a, hipStream_t s);
template <class Z>
hipError_t launch_col_pass_m5(int n, const ColArgsT<Z> &a, bool probe, bool det, hipStream_t s);

} // namespace mtg
