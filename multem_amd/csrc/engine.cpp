// Multislice engine: plans device memory, builds transmission stacks, and
// runs batches of waves through the two-pass slice step on one GPU.
//
// Replaces Multislice::STEM_ISTEM / EWFS_EWRS / CBED_CBEI
// (src/tem_simulation.cuh:100-327) and Wave_Function::psi
// (src/wave_function.cuh:229-252): instead of one probe at a time with a host
// synchronisation per detector, a batch of B waves (probe positions or
// phonon configurations) advances slice by slice, each slice one row pass and
// one column pass over the whole batch, and the slice loop of a batch is
// captured once in a hipGraph and replayed.
#include "engine.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <tuple>

namespace mtg {

namespace {

constexpr double kPi = 3.141592653589793238463;

// isZero of the reference: |x| <= epsilon_abs<T>, 1e-5 for float and 1e-13
// for double (src/types.cuh:383-391)
template <class R>
bool is_zero(R x)
{
    return std::fabs(x) <= (sizeof(R) == 4 ? R(1e-5) : R(1e-13));
}

template <class Z>
Z cz(double x, double y)
{
    Z z;
    z.x = (real_t<Z>)x;
    z.y = (real_t<Z>)y;
    return z;
}

// columns per column-pass workgroup of the precision (the detector partial layout)
template <class R>
int col_tile_w(int nx)
{
    return sizeof(R) == 8 ? col_tile_width_fp64(nx) : col_tile_width(nx);
}

thread_local const char *t_last_ok = "none";

void check(hipError_t e, const char *what)
{
    if (e == hipSuccess) t_last_ok = what;
    if (e != hipSuccess) {
        std::string m = std::string(what) + ": " + hipGetErrorString(e);
        throw Error{e == hipErrorOutOfMemory ? MT_ERR_OUT_OF_MEMORY : MT_ERR_DEVICE, m};
    }
}

// Device blocks kept between runs.  A simulation allocates the same buffers
// each time it is called (the transmission stack of a 1024^2 CBED with 20
// configurations x 325 slices is 52 GB), and returning them to the runtime
// and asking again costs host time that varies from run to run.  A released
// block is kept in a per-device list keyed by its byte size; an allocation
// takes a block of exactly its size from that list, else calls hipMalloc,
// and when hipMalloc runs out of memory every kept block of the device is
// freed and the call retried.  A release first waits for the device, as
// hipFree does, so that a block is never handed on while kernels still use
// it.  The lists are never destroyed (their blocks live until the process
// ends).
struct BlockCache {
    std::mutex mu;
    std::multimap<std::pair<int, size_t>, void *> blocks;   // (device, bytes) -> block
    std::map<int, size_t> kept;                             // kept bytes per device
};

BlockCache &block_cache()
{
    static BlockCache *c = new BlockCache;
    return *c;
}

void trim_blocks(int dev)
{
    BlockCache &c = block_cache();
    std::lock_guard<std::mutex> lk(c.mu);
    for (auto it = c.blocks.begin(); it != c.blocks.end();) {
        if (it->first.first == dev) {
            (void)hipFree(it->second);
            it = c.blocks.erase(it);
        } else {
            ++it;
        }
    }
    c.kept[dev] = 0;
}

size_t kept_bytes(int dev)
{
    BlockCache &c = block_cache();
    std::lock_guard<std::mutex> lk(c.mu);
    auto it = c.kept.find(dev);
    return it == c.kept.end() ? 0 : it->second;
}

void *dev_alloc(size_t bytes, int &dev)
{
    check(hipGetDevice(&dev), "hipGetDevice");
    BlockCache &c = block_cache();
    {
        std::lock_guard<std::mutex> lk(c.mu);
        auto it = c.blocks.find({dev, bytes});
        if (it != c.blocks.end()) {
            void *p = it->second;
            c.blocks.erase(it);
            c.kept[dev] -= bytes;
            return p;
        }
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        trim_blocks(dev);
        e = hipMalloc(&p, bytes);
    }
    check(e, "hipMalloc");
    return p;
}

// MULTEM_HIP_KEEP_BLOCKS=0 turns the kept blocks off (every release frees)
bool keep_blocks()
{
    static const bool on = [] {
        const char *e = std::getenv("MULTEM_HIP_KEEP_BLOCKS");
        return !(e && e[0] == '0');
    }();
    return on;
}

void dev_release(void *p, size_t bytes, int dev)
{
    if (!keep_blocks()) {
        (void)hipFree(p);
        return;
    }
    (void)hipDeviceSynchronize();
    BlockCache &c = block_cache();
    std::lock_guard<std::mutex> lk(c.mu);
    c.blocks.insert({{dev, bytes}, p});
    c.kept[dev] += bytes;
}

template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    int dev = 0;   // the device the block belongs to
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    void alloc(size_t count)
    {
        release();
        if (count) p = (T *)dev_alloc(count * sizeof(T), dev);
        n = count;
    }
    void release()
    {
        if (p) dev_release(p, n * sizeof(T), dev);
        p = nullptr;
        n = 0;
    }
    ~DBuf() { release(); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    // Host buffers are pageable and short-lived: the copy is complete on return.
    void upload(const T *h, size_t count, hipStream_t s)
    {
        check(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
        check(hipStreamSynchronize(s), "hipStreamSynchronize");
    }
};

struct Stream {
    hipStream_t s = nullptr;
    Stream() { check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate"); }
    ~Stream() { if (s) (void)hipStreamDestroy(s); }
};

struct Event {
    hipEvent_t e = nullptr;
    Event() { check(hipEventCreate(&e), "hipEventCreate"); }
    ~Event() { if (e) (void)hipEventDestroy(e); }
};

// HIP-event timing of the engine's own stream.  MULTEM_HIP_EVENTS=0 turns it
// off (stream synchronisation only), for runs under an external tracer.
bool events_enabled()
{
    static const bool on = [] {
        const char *v = std::getenv("MULTEM_HIP_EVENTS");
        return !(v && v[0] == '0');
    }();
    return on;
}

void record(const Event &ev, hipStream_t s)
{
    if (events_enabled()) check(hipEventRecord(ev.e, s), "hipEventRecord");
}

void wait(const Event &ev, hipStream_t s)
{
    if (events_enabled()) check(hipEventSynchronize(ev.e), "hipEventSynchronize");
    else check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

float elapsed(const Event &a, const Event &b)
{
    if (!events_enabled()) return 0.f;
    float ms = 0;
    check(hipEventElapsedTime(&ms, a.e, b.e), "hipEventElapsedTime");
    return ms;
}

template <class Z>
std::vector<Z> twiddles(int n)
{
    std::vector<Z> w(n);
    for (int m = 0; m < n; ++m) {
        double a = -2.0 * kPi * (double)m / (double)n;
        w[m] = cz<Z>(std::cos(a), std::sin(a));
    }
    return w;
}

// One configuration's atoms on the device: x, y, occ, type in the
// configuration's z order, and with sub-slicing the configuration's atom
// range per slice and the spline coefficients of every (slice, atom) pair.
// Captured graphs hold these buffers' addresses: once `fixed` (Setup::reserve,
// sized for every configuration of the run) they are never reallocated.
template <class R>
struct DevAtoms {
    DBuf<vec4_t<R>> a;
    DBuf<R> z;
    DBuf<int2> ranges, pairs;
    DBuf<int> pair_base;
    DBuf<R> coef;
    bool fixed = false;
    size_t size() const { return a.n; }
};

template <class T>
void upload_grow(DBuf<T> &b, const std::vector<T> &h, hipStream_t s, bool fixed = false)
{
    if (b.n < h.size()) {
        if (fixed) throw Error{MT_ERR_DEVICE, "reserved atom buffers too small for a configuration"};
        b.alloc(h.size());
    }
    if (!h.empty()) b.upload(h.data(), h.size(), s);
}

// Device-resident data shared by all entries of one model; R is the real
// type of the wave path (float or double).
template <class R>
struct Setup {
    using Z = cplx_t<R>;
    using V4 = vec4_t<R>;
    const ModelT<R> &m;
    hipStream_t s;
    int nx, ny;
    size_t nxy;
    DBuf<Z> tw_x, tw_y;
    DBuf<R> tables;
    DBuf<Z> rmax;
    DBuf<int2> ranges;
    std::vector<int2> h_ranges;
    DBuf<R> sub_slice_z;                             // sub-slicing: slice bounds
    DBuf<double> sub_par, sub_qx, sub_qw;            // sub-slicing: element terms, rule

    Setup(const ModelT<R> &mm, hipStream_t st) : m(mm), s(st), nx(mm.nx), ny(mm.ny), nxy((size_t)mm.nx * mm.ny)
    {
        auto wx = twiddles<Z>(nx), wy = twiddles<Z>(ny);
        tw_x.alloc(nx); tw_x.upload(wx.data(), nx, s);
        tw_y.alloc(ny); tw_y.upload(wy.data(), ny, s);
        const int nt = (int)m.tables.size();
        if (nt > kMaxTypes) throw Error{MT_ERR_UNSUPPORTED, "more than 16 distinct elements"};
        std::vector<R> tb((size_t)nt * 5 * 128);
        std::vector<Z> rm(nt);
        for (int i = 0; i < nt; ++i) {
            const AtomTableT<R> &t = m.tables[i];
            for (int k = 0; k < 128; ++k) {
                tb[(size_t)i * 640 + k] = t.R2[k];
                tb[(size_t)i * 640 + 128 + k] = t.c0[k];
                tb[(size_t)i * 640 + 256 + k] = t.c1[k];
                tb[(size_t)i * 640 + 384 + k] = t.c2[k];
                tb[(size_t)i * 640 + 512 + k] = t.c3[k];
            }
            rm[i] = cz<Z>(t.R_max_f, t.R2_max_f);
        }
        tables.alloc(tb.size()); tables.upload(tb.data(), tb.size(), s);
        rmax.alloc(nt); rmax.upload(rm.data(), nt, s);
        for (const auto &sl : m.slices) h_ranges.push_back(make_int2(sl.iatom_0, sl.iatom_e));
        ranges.alloc(h_ranges.size()); ranges.upload(h_ranges.data(), h_ranges.size(), s);
        if (m.subslicing) {
            std::vector<R> sz;
            std::vector<double> par((size_t)nt * 12), qx, qw;
            for (const auto &sl : m.slices) { sz.push_back(sl.z_0); sz.push_back(sl.z_e); }
            for (int i = 0; i < nt; ++i) {
                // the element's terms and taper in double, as its projected
                // table is built (get_atom_coef, src/atomic_data.hpp:1610-1665)
                const AtomTableT<R> &t = m.tables[i];
                for (int j = 0; j < 5; ++j) {
                    par[(size_t)i * 12 + j] = t.vr_a[j];
                    par[(size_t)i * 12 + 5 + j] = t.vr_k[j];
                }
                par[(size_t)i * 12 + 10] = t.R_tap * t.R_tap;
                par[(size_t)i * 12 + 11] = t.tap_cf;
            }
            m.subslice_rule(qx, qw);
            upload_grow(sub_slice_z, sz, s);
            upload_grow(sub_par, par, s);
            upload_grow(sub_qx, qx, s);
            upload_grow(sub_qw, qw, s);
        }
    }

    // Allocates d for every configuration of the run, before any graph is
    // captured on it: the atoms, and with sub-slicing the ranges and the
    // pairs of the configuration that has the most.
    void reserve(DevAtoms<R> &d) const
    {
        const size_t na = m.atoms_u.x.size(), ns = (size_t)m.n_slice();
        d.a.alloc(na);
        if (m.subslicing) {
            size_t np = 0;
            auto count = [&](const std::vector<SliceInfoT<R>> &sl) {
                size_t c = 0;
                for (const auto &x : sl) c += (size_t)std::max(0, x.iatom_e - x.iatom_0 + 1);
                np = std::max(np, c);
            };
            if (m.frozen() && m.pn_z) {
                std::vector<int> order;
                std::vector<R> x, y, z;
                std::vector<SliceInfoT<R>> sl;
                for (int ic = m.fp_iconf_0; ic <= m.pn_nconf; ++ic) {
                    m.config_atoms(ic, order, x, y, z, sl);
                    count(sl);
                }
            } else {
                count(m.slices);
            }
            d.z.alloc(na);
            d.ranges.alloc(ns);
            d.pair_base.alloc(ns);
            d.pairs.alloc(std::max<size_t>(np, 1));
            d.coef.alloc(std::max<size_t>(np, 1) * 512);
        }
        d.fixed = true;
    }

    // Configuration iconf's atoms onto the device (with sub-slicing: in the
    // configuration's z order, with its atom ranges and the coefficients of
    // every (slice, atom) pair, computed on stream st).
    void load(int iconf, DevAtoms<R> &d, hipStream_t st = nullptr)
    {
        if (!st) st = s;
        if (!m.subslicing) {
            upload_grow(d.a, atoms_of(iconf), st, d.fixed);
            return;
        }
        std::vector<int> order;
        std::vector<R> x, y, z;
        std::vector<SliceInfoT<R>> sl;
        m.config_atoms(iconf, order, x, y, z, sl);
        std::vector<V4> a(order.size());
        for (size_t i = 0; i < order.size(); ++i) {
            const int ty = m.type_of_atom[order[i]];
            R tyr;
            if constexpr (sizeof(R) == 4) std::memcpy(&tyr, &ty, sizeof tyr);
            else tyr = (R)ty;
            a[i].x = x[i]; a[i].y = y[i]; a[i].z = (R)m.atoms_u.occ[order[i]]; a[i].w = tyr;
        }
        std::vector<int2> rg, pairs;
        std::vector<int> base;
        for (size_t is = 0; is < sl.size(); ++is) {
            rg.push_back(make_int2(sl[is].iatom_0, sl[is].iatom_e));
            base.push_back((int)pairs.size());
            for (int ia = sl[is].iatom_0; ia <= sl[is].iatom_e; ++ia) pairs.push_back(make_int2((int)is, ia));
        }
        upload_grow(d.a, a, st, d.fixed);
        upload_grow(d.z, z, st, d.fixed);
        upload_grow(d.ranges, rg, st, d.fixed);
        upload_grow(d.pair_base, base, st, d.fixed);
        upload_grow(d.pairs, pairs, st, d.fixed);
        if (d.coef.n < pairs.size() * 512) {
            if (d.fixed) throw Error{MT_ERR_DEVICE, "reserved sub-slice coefficients too small for a configuration"};
            d.coef.alloc(pairs.size() * 512);
        }
        SubArgsT<R> sa{};
        sa.n_pairs = (int)pairs.size();
        sa.pairs = d.pairs.p; sa.atoms = d.a.p; sa.z = d.z.p;
        sa.slice_z = sub_slice_z.p; sa.tables = tables.p; sa.par = sub_par.p;
        sa.qx = sub_qx.p; sa.qw = sub_qw.p; sa.coef = d.coef.p;
        check(launch_subslice_coef(sa, st), "sub-slice coefficients");
    }

    // atoms (x, y, occ, type) of configuration iconf: the type is the bit
    // pattern of an int in .w for float, its value for double
    std::vector<V4> atoms_of(int iconf) const
    {
        std::vector<R> x, y;
        m.displaced(iconf, x, y);
        std::vector<V4> a(x.size());
        for (size_t i = 0; i < x.size(); ++i) {
            const int ty = m.type_of_atom[i];
            R tyr;
            if constexpr (sizeof(R) == 4) std::memcpy(&tyr, &ty, sizeof tyr);
            else tyr = (R)ty;
            a[i].x = x[i]; a[i].y = y[i]; a[i].z = (R)m.atoms_u.occ[i]; a[i].w = tyr;
        }
        return a;
    }

    // Transmission stack of one configuration into trans (n_slice * nxy),
    // optionally the potential V (natural layout) of the same slices; on
    // stream st (default: the setup stream).
    void build_tf(const DevAtoms<R> &atoms, int slice0, int n_slices, Z *trans, R *V, hipStream_t st = nullptr)
    {
        TfArgsT<R> a{};
        a.nx = nx; a.ny = ny; a.nxh = m.nxh; a.nyh = m.nyh;
        a.dRx = m.dRx; a.dRy = m.dRy;
        a.atoms = atoms.a.p;
        a.slice_range = m.subslicing ? atoms.ranges.p : ranges.p;
        a.sub_coef = m.subslicing ? atoms.coef.p : nullptr;
        a.pair_base = m.subslicing ? atoms.pair_base.p : nullptr;
        a.slice0 = slice0; a.n_slices = n_slices;
        a.tables = tables.p; a.rmax = rmax.p; a.n_types = (int)m.tables.size();
        a.w = m.Vr_factor;
        a.weak_phase = m.interaction_model == 3 ? 1 : 0;
        a.trans = trans; a.V = V;
        check(launch_tf_build(a, st ? st : s), "tf_build");
        if (m.bwl && trans) bandwidth_limit(trans, n_slices, st);
    }

    // Transmission of n_slices stored potentials V into trans (+ bandwidth
    // limit): the same values build_tf writes for those slices.
    void tf_from_pot(const R *V, int n_slices, Z *trans, hipStream_t st = nullptr)
    {
        check(launch_trans_from_pot(V, n_slices, nx, ny, m.Vr_factor, m.interaction_model == 3 ? 1 : 0, trans,
                                    st ? st : s),
              "trans_from_pot");
        if (m.bwl) bandwidth_limit(trans, n_slices, st);
    }

    // FFT -> x bwl/(nx ny) -> IFFT on n waves in place (Transmission_Function::trans,
    // src/transmission_function.cuh:72-82)
    void bandwidth_limit(Z *w, int n, hipStream_t st = nullptr)
    {
        if (!st) st = s;
        RowArgsT<Z> r{};
        r.src = w; r.dst = w; r.member_stride = (long long)nxy; r.nrows = nx; r.n_members = n;
        r.tw = tw_y.p;
        r.flags = ROW_LOAD | ROW_FFT | ROW_STORE;
        check(launch_row_pass(ny, r, st), "row pass");
        ColArgsT<Z> c = col_base(w, w, n);
        c.flags = COL_LOAD | COL_FFT | COL_BWL | COL_IFFT | COL_STORE;
        check(launch_col_pass(nx, c, false, false, st), "col pass");
        r.flags = ROW_LOAD | ROW_IFFT | ROW_STORE;
        check(launch_row_pass(ny, r, st), "row pass");
    }

    ColArgsT<Z> col_base(const Z *src, Z *dst, int n_members) const
    {
        ColArgsT<Z> c{};
        c.src = src; c.dst = dst; c.member_stride = (long long)nxy;
        c.ncols = ny; c.n_members = n_members;
        c.tw = tw_x.p;
        c.dgx_f = m.dgx; c.dgy_f = m.dgy;
        c.nx = nx; c.ny = ny;
        c.bwl_alpha = m.bwl_alpha; c.gl2_max = m.gl2_max;
        c.inv_nxy = R(1) / (R)nxy;
        c.p_scale = R(1);
        return c;
    }

    ProbeParamsT<R> probe_params(unsigned int n_aperture) const
    {
        ProbeParamsT<R> P{};
        const double lam = m.lambda_d;
        P.c10 = probe_c10(m.cond.c_10);
        P.c30 = is_zero(m.cond.c_30) ? 0.0 : -kPi * (double)m.cond.c_30 * std::pow(lam, 3) / 2.0;
        P.c50 = is_zero(m.cond.c_50) ? 0.0 : -kPi * (double)m.cond.c_50 * std::pow(lam, 5) / 3.0;
        P.dgx = 1.0 / m.lx_d; P.dgy = 1.0 / m.ly_d;
        // azimuthal coefficients c_c_* of Lens::set_input_data (src/types.cuh:2924-2976)
        const LensModelT<R> &L = m.cond;
        if (L.phi_required) {
            P.azim = 1;
            P.az_m = (double)L.m;
            const double l1 = lam, l2 = lam * lam, l3 = l2 * lam, l4 = l3 * lam, l5 = l4 * lam;
            const double cf[11] = {-kPi * l1, -2.0 * kPi * l2 / 3.0, -2.0 * kPi * l2 / 3.0, -kPi * l3 / 2.0,
                                   -kPi * l3 / 2.0, -2.0 * kPi * l4 / 5.0, -2.0 * kPi * l4 / 5.0,
                                   -2.0 * kPi * l4 / 5.0, -kPi * l5 / 3.0, -kPi * l5 / 3.0, -kPi * l5 / 3.0};
            const R cv[11] = {L.c_12, L.c_21, L.c_23, L.c_32, L.c_34, L.c_41, L.c_43, L.c_45, L.c_52, L.c_54, L.c_56};
            const R pv[11] = {L.phi_12, L.phi_21, L.phi_23, L.phi_32, L.phi_34, L.phi_41, L.phi_43, L.phi_45,
                              L.phi_52, L.phi_54, L.phi_56};
            for (int k = 0; k < 11; ++k) {
                P.az_c[k] = is_zero(cv[k]) ? 0.0 : cf[k] * (double)cv[k];
                P.az_phi[k] = (double)pv[k];
            }
        }
        P.dgx_f = m.dgx; P.dgy_f = m.dgy;
        P.g2_min = m.cond.g2_min; P.g2_max = m.cond.g2_max;
        P.norm = n_aperture ? (R)std::sqrt(1.0 / (double)n_aperture) : R(0);
        return P;
    }

    // defocus coefficient of chi for condenser defocus c_10, referred to the
    // first slice (Incident_Wave::operator(), src/incident_wave.cuh:72-91)
    double probe_c10(R c_10) const
    {
        const R f_s = c_10 - (m.cond.zero_defocus_plane - m.z_init());
        return is_zero(f_s) ? 0.0 : -kPi * (double)f_s * m.lambda_d;
    }

    Beam beam(R xb, R yb, double c10) const
    {
        // x = -2 pi (x_b - lx/2)  (Grid_2d::exp_factor_Rx, src/types.cuh:2599-2611)
        return Beam{-2.0 * kPi * ((double)xb - 0.5 * m.lx_d), -2.0 * kPi * ((double)yb - 0.5 * m.ly_d), c10};
    }

    // Defocus points of the condenser lens: the probe's defocus coefficient
    // and the weight of each point.  Fully integrated illumination follows
    // the defocus loop of STEM_ISTEM / CBED_CBEI (src/tem_simulation.cuh:138-178,
    // 207-261): c_10 = c_10_0 + x_t, weight w_t; otherwise one point.
    void cond_defocus_points(const Quad1 &qt, std::vector<double> &c10, std::vector<double> &wq) const
    {
        c10.clear();
        wq.clear();
        const double c_10_0 = (double)m.cond.c_10;
        for (size_t i = 0; i < qt.x.size(); ++i) {
            c10.push_back(probe_c10((R)(c_10_0 + qt.x[i])));
            wq.push_back(qt.w[i]);
        }
    }

    // pixels inside the condenser aperture, |g + u| evaluated as the probe does
    unsigned int count_aperture(R ux = 0, R uy = 0)
    {
        DBuf<unsigned int> cnt(1);
        unsigned int zero = 0;
        cnt.upload(&zero, 1, s);
        check(launch_count_aperture(nx, ny, m.dgx, m.dgy, ux, uy, m.cond.g2_min, m.cond.g2_max, cnt.p, s),
              "count aperture");
        unsigned int h = 0;
        check(hipMemcpyAsync(&h, cnt.p, sizeof h, hipMemcpyDeviceToHost, s), "D2H");
        check(hipStreamSynchronize(s), "sync");
        return h;
    }

    // Objective-lens transfer functions of HRTEM (MicroscopeEffects,
    // src/microscope_effects.cuh:57-202): one table for the coherent and the
    // partially coherent models; for the fully integrated model one coherent
    // table per defocus point c_10 = c_10_0 + ti_iehwgd x_t of num_int_TEM,
    // with its weight in wq.
    void ctf_tables(DBuf<Z> &ctf, std::vector<double> &wq)
    {
        const LensModelT<R> &L = m.obj;
        const double lam = m.lambda_d;
        CtfParamsT<R> P{};
        P.mode = m.illumination_model == 2 ? 2 : 1;
        P.dgx = 1.0 / m.lx_d; P.dgy = 1.0 / m.ly_d;
        P.dgx_f = m.dgx; P.dgy_f = m.dgy;
        P.g2_min = L.g2_min; P.g2_max = L.g2_max;
        auto nz = [](R v) { return !is_zero(v); };
        P.cc[0] = nz(L.c_10) ? -kPi * (double)L.c_10 * lam : 0.0;
        P.cc[1] = nz(L.c_30) ? -kPi * (double)L.c_30 * std::pow(lam, 3) / 2.0 : 0.0;
        P.cc[2] = nz(L.c_50) ? -kPi * (double)L.c_50 * std::pow(lam, 5) / 3.0 : 0.0;
        P.cc[3] = (double)L.c_10;
        P.cc[4] = (double)L.c_30;
        if (L.phi_required && P.mode == 1) {
            P.azim = 1;
            P.az_m = (double)L.m;
            const double l1 = lam, l2 = lam * lam, l3 = l2 * lam, l4 = l3 * lam, l5 = l4 * lam;
            const double cf[11] = {-kPi * l1, -2.0 * kPi * l2 / 3.0, -2.0 * kPi * l2 / 3.0, -kPi * l3 / 2.0,
                                   -kPi * l3 / 2.0, -2.0 * kPi * l4 / 5.0, -2.0 * kPi * l4 / 5.0,
                                   -2.0 * kPi * l4 / 5.0, -kPi * l5 / 3.0, -kPi * l5 / 3.0, -kPi * l5 / 3.0};
            const R cv[11] = {L.c_12, L.c_21, L.c_23, L.c_32, L.c_34, L.c_41, L.c_43, L.c_45, L.c_52, L.c_54, L.c_56};
            const R pv[11] = {L.phi_12, L.phi_21, L.phi_23, L.phi_32, L.phi_34, L.phi_41, L.phi_43, L.phi_45,
                              L.phi_52, L.phi_54, L.phi_56};
            for (int k = 0; k < 11; ++k) {
                P.az_c[k] = nz(cv[k]) ? cf[k] * (double)cv[k] : 0.0;
                P.az_phi[k] = (double)pv[k];
            }
        }
        // temporal / spatial selection of PCTF_LI_WPO_TEM
        P.ti = m.temporal_spatial_incoh == 3 ? 0.0 : (double)L.ti_iehwgd;
        P.sth = m.temporal_spatial_incoh == 2 ? 0.0 : (double)L.si_theta_c;
        P.lambda = lam;
        wq.assign(1, 1.0);
        if (m.illumination_model == 4) {
            Quad1 qt;
            m.obj_temporal_quadrature(qt);
            wq = qt.w;
            ctf.alloc(qt.x.size() * nxy);
            for (size_t j = 0; j < qt.x.size(); ++j) {
                const R c10 = L.ti_iehwgd * (R)qt.x[j] + L.c_10;
                P.cc[0] = nz(c10) ? -kPi * (double)c10 * lam : 0.0;
                P.cc[3] = (double)c10;
                check(launch_ctf_table(nx, ny, P, ctf.p + j * nxy, s), "ctf table");
            }
        } else {
            ctf.alloc(nxy);
            check(launch_ctf_table(nx, ny, P, ctf.p, s), "ctf table");
        }
        check(hipStreamSynchronize(s), "sync");
    }

    // tilt ramp of the outputs, exp(-2 pi i g0.(r - r_c)), with 1/ny folded
    // into rx (it follows an unnormalised inverse transform along y)
    // (one pair per tilt azimuth: rx [nrot][nx], ry [nrot][ny])
    void ramp_tables(DBuf<Z> &rx, DBuf<Z> &ry, int nrot = 1)
    {
        rx.alloc((size_t)nrot * nx);
        ry.alloc((size_t)nrot * ny);
        // Grid_2d::Rx_c() is an int (src/types.cuh:2284-2287)
        const double Rx_c = (double)(int)(R(0.5) * m.lx), Ry_c = (double)(int)(R(0.5) * m.ly);
        for (int ir = 0; ir < nrot; ++ir)
            check(launch_ramp_tables(nx, ny, (double)m.gx0_rot(ir), (double)m.gy0_rot(ir), m.lx_d / nx, m.ly_d / ny,
                                     Rx_c, Ry_c, R(1) / (R)ny, rx.p + (size_t)ir * nx, ry.p + (size_t)ir * ny, s),
                  "ramp tables");
        check(hipStreamSynchronize(s), "sync");
    }

    // Spectrum of the user-defined incident wave: fft2_shift then the
    // unnormalised forward FFT (Incident_Wave::set_input_data,
    // src/incident_wave.cuh:50-55), natural k-space order
    void user_wave_spectrum(DBuf<Z> &f)
    {
        std::vector<Z> h(nxy);
        for (size_t i = 0; i < nxy; ++i) h[i] = cz<Z>(m.iw_psi[2 * i], m.iw_psi[2 * i + 1]);
        DBuf<Z> a(nxy);
        a.upload(h.data(), nxy, s);
        f.alloc(nxy);
        check(launch_shift_complex(a.p, f.p, nx, ny, s), "shift");
        Member hm{0, -1, 0, 0, 1.0};
        DBuf<Member> mem(1);
        mem.upload(&hm, 1, s);
        RowArgsT<Z> r{};
        r.src = f.p; r.dst = f.p; r.member_stride = (long long)nxy; r.nrows = nx; r.n_members = 1;
        r.tw = tw_y.p; r.members = mem.p;
        r.flags = ROW_LOAD | ROW_FFT | ROW_STORE;
        check(launch_row_pass(ny, r, s), "row pass");
        ColArgsT<Z> c = col_base(f.p, f.p, 1);
        c.flags = COL_LOAD | COL_FFT | COL_STORE;
        check(launch_col_pass(nx, c, false, false, s), "col pass");
        check(hipStreamSynchronize(s), "sync");
    }

    // propagator tables for the given distances: px [nz][nx], py [nz][ny],
    // one set per tilt azimuth ([nrot][nz][nx], [nrot][nz][ny])
    void prop_tables(const std::vector<double> &z, DBuf<Z> &px, DBuf<Z> &py, int nrot = 1)
    {
        DBuf<double> dz(z.size());
        dz.upload(z.data(), z.size(), s);
        px.alloc((size_t)nrot * z.size() * nx);
        py.alloc((size_t)nrot * z.size() * ny);
        for (int ir = 0; ir < nrot; ++ir)
            check(launch_prop_tables(nx, ny, dz.p, (int)z.size(), m.lambda_d, m.mul_sign, m.lx_d, m.ly_d,
                                     (double)m.gx0_rot(ir), (double)m.gy0_rot(ir), px.p + (size_t)ir * z.size() * nx,
                                     py.p + (size_t)ir * z.size() * ny, s),
                  "prop tables");
        check(hipStreamSynchronize(s), "sync");
    }
};

// Slice storage of one configuration (Transmission_Function::Memory_Slice,
// src/transmission_function.cuh:146-215, used at :84-103 and :118-134).
// When the transmission stack of every slice fits the memory given to it,
// the stack is resident and the row pass reads slice s from it.  Otherwise,
// as in the reference, the potentials of as many slices as fit are stored
// (half the bytes of a transmission), and each slice's transmission is
// rebuilt just before the row pass that multiplies by it, into one slot:
// from its stored potential, or, beyond those, from the atoms.  Either way
// the values are those of the resident stack, bit for bit.
template <class R>
struct SliceStore {
    using Z = cplx_t<R>;
    using V4 = vec4_t<R>;
    Setup<R> &S;
    int n_slice = 0;
    bool resident = true;
    int n_pot = 0;        // rebuilt mode: slices whose potential is stored
    int K = 1;            // resident mode: configurations held at once
    DBuf<Z> trans;        // resident: [K][n_slice][nxy]; rebuilt: the slot [nxy]
    DBuf<R> V;            // rebuilt: [n_pot][nxy], natural layout

    explicit SliceStore(Setup<R> &s) : S(s), n_slice(s.m.n_slice()) {}

    size_t stack_bytes() const { return (size_t)n_slice * S.nxy * sizeof(Z); }
    size_t slot_bytes() const { return S.nxy * sizeof(Z); }
    size_t pot_bytes() const { return S.nxy * sizeof(R); }

    // Plan for `avail` bytes and up to k_max configurations: the stacks of
    // as many configurations as fit (at least one), else the rebuilt mode
    // with as many stored potentials as fit beside the slot.  Returns the
    // bytes taken; throws if not even the slot fits.
    size_t plan(size_t avail, int k_max)
    {
        if (stack_bytes() <= avail) {
            resident = true;
            K = (int)std::max<size_t>(1, std::min<size_t>((size_t)k_max, avail / stack_bytes()));
            trans.alloc((size_t)K * n_slice * S.nxy);
            return (size_t)K * stack_bytes();
        }
        if (slot_bytes() > avail) throw Error{MT_ERR_OUT_OF_MEMORY, "no room for one transmission slice"};
        resident = false;
        K = 1;
        n_pot = (int)std::min<size_t>((size_t)n_slice, (avail - slot_bytes()) / pot_bytes());
        trans.alloc(S.nxy);
        check(hipMemsetAsync(trans.p, 0, slot_bytes(), S.s), "hipMemsetAsync");
        if (n_pot > 0) V.alloc((size_t)n_pot * S.nxy);
        return slot_bytes() + (size_t)n_pot * pot_bytes();
    }

    // what a configuration needs before its waves run: its transmission stack
    // (into configuration slot j), or its stored potentials
    void build(const DevAtoms<R> &atoms, int j, hipStream_t st = nullptr)
    {
        if (resident) S.build_tf(atoms, 0, n_slice, trans.p + (size_t)j * n_slice * S.nxy, nullptr, st);
        else if (n_pot > 0) S.build_tf(atoms, 0, n_pot, nullptr, V.p, st);
    }

    // Row-pass operands for slice is: r.trans, r.tslot_stride and r.slice
    template <class RA>
    void bind(int is, RA &r) const
    {
        r.trans = trans.p;
        r.tslot_stride = resident ? (long long)n_slice * S.nxy : 0;
        r.slice = resident ? is : 0;
    }

    // bind(), after enqueueing (rebuilt mode) the slice's transmission into
    // the slot on st, ordered before the row pass on the same stream
    template <class RA>
    void for_slice(const DevAtoms<R> &atoms, int is, RA &r, hipStream_t st)
    {
        if (!resident) {
            if (is < n_pot) S.tf_from_pot(V.p + (size_t)is * S.nxy, 1, trans.p, st);
            else S.build_tf(atoms, is, 1, trans.p, nullptr, st);
        }
        bind(is, r);
    }

    std::string describe() const
    {
        if (resident) return "slices=resident";
        return "slices=rebuilt stored_potentials=" + std::to_string(n_pot) + "/" + std::to_string(n_slice);
    }
};

size_t free_device_memory()
{
    size_t fr = 0, tot = 0;
    check(hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    return fr + kept_bytes(dev);   // blocks kept from earlier runs are free for this one
}

// Row pitch of a STEM wave batch: ny plus a pad (mt_hip_set_wave_pad;
// default 0).  The column pass reads 128-byte row segments at this stride.
// A bare copy with the column pattern at 512^2 runs 6 % faster at pitch 528
// than at 512, but the slice step does not: at batch 118 the column pass
// takes 99.6 us at pitch 528 and 98.5 us at 512, so the default keeps ny.
// The pad changes no arithmetic: the images are bit-identical.
int wave_pad()
{
    const int p = run_control().wave_pad;
    return p < 0 ? 0 : p;
}

size_t budget_bytes()
{
    size_t fr = free_device_memory();
    size_t b = (size_t)(0.85 * (double)fr);
    if (run_control().memory_budget) b = std::min(b, run_control().memory_budget);
    return b;
}

void select_device(int device)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) throw Error{MT_ERR_NO_DEVICE, "no HIP device visible"};
    if (device < 0 || device >= n) throw Error{MT_ERR_INVALID_ARGUMENT, "gpu_device out of range"};
    check(hipSetDevice(device), "hipSetDevice");
}

void progress_add(int k) { run_control().iter.fetch_add(k); }
void stop_check()
{
    if (run_control().stop.load()) throw Error{MT_ERR_STOPPED, "stopped by mt_hip_request_stop"};
}

// ---------------------------------------------------------------------------
// Autotuner of the slice-step launch variants.  For each pass it launches
// every candidate variant on the run's own buffers (one warm-up, then kReps
// timed launches) and keeps the lowest-numbered variant within 2 % of the
// fastest, so near-ties resolve the same way from run to run.  Choices are
// cached per shape (nx, ny, batch, detector epilogue) for the process.
//
// Tuning launches overwrite the wave buffer and the detector partials; the
// caller runs it before the first batch, which rewrites both.
// The key holds each pass's flags: a run whose passes go to the generic
// kernels (bandwidth limit, per-azimuth tables, matrix detectors) must not
// share choices with one that runs the specialised kernels.
struct TuneKey {
    int nx, ny, B, row_flags, col_flags, det_flags, real_bytes;
    bool operator<(const TuneKey &o) const
    {
        return std::tie(nx, ny, B, row_flags, col_flags, det_flags, real_bytes) <
               std::tie(o.nx, o.ny, o.B, o.row_flags, o.col_flags, o.det_flags, o.real_bytes);
    }
};

std::map<TuneKey, PassChoice> &tune_cache()
{
    static std::map<TuneKey, PassChoice> c;
    return c;
}

// mean duration (ms) of kReps launches of `launch`, or -1 if the variant is
// not valid for this shape
template <class F>
double time_variant(F &&launch, hipStream_t s)
{
    constexpr int kReps = 4;
    hipError_t e = launch();
    if (e == hipErrorInvalidValue) {
        (void)hipGetLastError();
        return -1.0;
    }
    check(e, "tuning launch");
    check(hipStreamSynchronize(s), "hipStreamSynchronize");
    double ms;
    if (events_enabled()) {
        Event a, b;
        record(a, s);
        for (int i = 0; i < kReps; ++i) check(launch(), "tuning launch");
        record(b, s);
        wait(b, s);
        ms = elapsed(a, b) / kReps;
    } else {
        const auto t0 = std::chrono::steady_clock::now();
        for (int i = 0; i < kReps; ++i) check(launch(), "tuning launch");
        check(hipStreamSynchronize(s), "hipStreamSynchronize");
        ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / kReps;
    }
    return ms;
}

// forced >= 0: use that variant (checked to be valid for the shape);
// tune == false: variant 0 unless forced; variants == false: the pass runs
// the generic kernel, which has one variant (0)
template <class F>
int pick_variant(const char *name, int forced, bool tune, bool variants, F &&launch, hipStream_t s,
                 std::ostringstream &log)
{
    if (!variants) {
        if (forced > 0) {
            throw Error{MT_ERR_INVALID_ARGUMENT, std::string("launch variant ") + std::to_string(forced) + " of the " +
                                                     name + " pass is not valid for this run (generic kernel)"};
        }
        check(launch(0), "variant launch");
        log << name << "=v0(generic) ";
        return 0;
    }
    if (forced >= 0 || !tune) {
        const int v = std::max(forced, 0);
        hipError_t e = v < kPassVariants ? launch(v) : hipErrorInvalidValue;
        if (e == hipErrorInvalidValue) {
            (void)hipGetLastError();
            throw Error{MT_ERR_INVALID_ARGUMENT, std::string("launch variant ") + std::to_string(v) + " of the " +
                                                     name + " pass is not valid for this grid"};
        }
        check(e, "variant launch");
        log << name << "=v" << v << (forced >= 0 ? "(forced) " : "(fixed) ");
        return v;
    }
    std::vector<double> t(kPassVariants, -1.0);
    for (int v = 0; v < kPassVariants; ++v) t[v] = time_variant([&] { return launch(v); }, s);
    double best = -1.0;
    for (double x : t)
        if (x >= 0 && (best < 0 || x < best)) best = x;
    int pick = 0;
    for (int v = 0; v < kPassVariants; ++v)
        if (t[v] >= 0 && t[v] <= best * 1.02) { pick = v; break; }
    log << name << "=v" << pick << "[";
    for (int v = 0; v < kPassVariants; ++v) {
        if (v) log << ",";
        if (t[v] < 0) log << "-";
        else log << (int)std::lround(t[v] * 1000.0);
    }
    log << "us] ";
    return pick;
}

// r: a full row slice step over the batch; c: a full column slice step;
// c_last: the last-slice column pass with detector sums (nullptr if the run
// has none).
template <class Z>
PassChoice choose_variants(int nx, int ny, int B, RowArgsT<Z> r, ColArgsT<Z> c, const ColArgsT<Z> *c_last,
                           hipStream_t s)
{
    const RunControl &rc = run_control();
    const bool forced = rc.force_row >= 0 || rc.force_col >= 0 || rc.force_col_det >= 0;
    const bool tune = rc.autotune != 0;
    const TuneKey key{nx, ny, B, r.flags, c.flags, c_last ? c_last->flags : -1, (int)sizeof(real_t<Z>)};
    if (tune && rc.autotune == 1 && !forced) {
        auto it = tune_cache().find(key);
        if (it != tune_cache().end()) {
            PassChoice cached = it->second;
            cached.log = "cached: " + cached.log;
            return cached;
        }
    }
    PassChoice ch;
    std::ostringstream log;
    log << "nx=" << nx << " ny=" << ny << " batch=" << B << ": ";
    ch.row = pick_variant("row", rc.force_row, tune, row_variants_apply(r.flags),
                          [&](int v) { r.variant = v; return launch_row_pass(ny, r, s); }, s, log);
    ch.col = pick_variant("col", rc.force_col, tune, col_variants_apply(c.flags, false),
                          [&](int v) { c.variant = v; return launch_col_pass(nx, c, false, false, s); }, s, log);
    if (c_last) {
        ColArgsT<Z> cl = *c_last;
        ch.col_det = pick_variant("col_det", rc.force_col_det, tune, col_variants_apply(cl.flags, true),
                                  [&](int v) { cl.variant = v; return launch_col_pass(nx, cl, false, true, s); }, s,
                                  log);
    }
    check(hipStreamSynchronize(s), "hipStreamSynchronize");
    ch.log = log.str();
    if (tune && !forced) tune_cache()[key] = ch;
    return ch;
}

// ---------------------------------------------------------------------------
// STEM: batches of probe positions of one configuration.
//
// A StemJob holds one specimen's device state.  Its work is a sequence of
// items (configuration, probe batch); prepare(k) makes item k's inputs
// resident (transmission stack of the configuration, the batch's member and
// beam tables) and pipeline(st) enqueues the item's whole device work on
// stream st: probe generation, the slice loop, detector partials -> image.
// The member and beam tables of every item are built once and kept on the
// device; prepare(k) enqueues a device copy of item k's rows into the fixed
// buffers the pipeline reads, so the host never waits for a batch to end
// before enqueueing the next one.
template <class R>
class StemJob {
public:
    using Z = cplx_t<R>;
    const ModelT<R> &m;
    Stream own;
    hipStream_t s;            // uploads and transmission builds
    Setup<R> S;
    SliceStore<R> store;      // transmission stack, or stored potentials rebuilt per slice
    int ns_run, nscan, n_thk, n_det, B, ntile, nbatch;
    size_t nxy;
    long long pitch;          // row pitch of the wave batch (ny + pad)
    size_t wave;              // complex elements per wave: nx * pitch
    DBuf<Z> psi;
    DBuf<Member> members;
    DBuf<Beam> beams;
    DBuf<Member> all_members;   // [n_t][nbatch][B]: every item's member table
    DBuf<Beam> all_beams;       // [n_t][nbatch][B]
    DBuf<double> part, image;
    DBuf<Z> px, py;           // [n_slice + n_thk]: slice propagators, then back-propagations
    DBuf<Z> rx, ry, scratch;   // tilted illumination: output ramp and the output copy of the batch
    DBuf<double2> coh;              // pn_coh_contrib: sum over configurations of w psi(g), [scan][n_thk][nxy]
    DBuf<Z> fpsi0;                  // iw_type 3: spectrum of the user's incident wave
    DBuf<R> det_fR;                 // matrix detectors
    DevAtoms<R> atoms;
    ProbeParamsT<R> P{};
    unsigned int nap = 0;
    R w = 0;
    int n_t = 1;                     // defocus points (fully integrated illumination)
    std::vector<double> tc10, twq;   // their probe defocus coefficients and weights
    RowArgsT<Z> r{};
    ColArgsT<Z> c{};
    PassChoice choice;
    std::vector<Member> hm;
    int cur_conf = -1;
    double ms_tf = 0;
    std::string batch_log;       // batch-width selection log
    int n_chain = 1;             // concurrent chains the batch is split into
    std::vector<std::unique_ptr<Stream>> cs;
    std::vector<std::unique_ptr<Event>> joins;
    Event fork;

    // budget: bytes this job may use for its transmission stack and waves
    StemJob(const ModelT<R> &mm, size_t budget, int batch_override)
        : m(mm), s(own.s), S(mm, own.s), store(S), nxy((size_t)mm.nx * mm.ny),
          pitch((long long)mm.ny + wave_pad()), wave((size_t)mm.nx * pitch)
    {
        ns_run = 0;
        for (const auto &t : m.thk) ns_run = std::max(ns_run, t.islice + 1);   // deeper slices add nothing
        nscan = (int)m.sc_x.size();
        n_thk = (int)m.thk.size();
        n_det = (int)m.det_g2_min.size();
        if (n_det > kMaxDet) throw Error{MT_ERR_UNSUPPORTED, "more than 8 detectors"};

        // memory plan: the coherent accumulators of every scan point (all
        // resident), the slice store -- one transmission stack, or stored
        // potentials when the stack does not fit beside two waves -- and B waves
        const size_t coh_bytes = m.pn_coh_contrib ? (size_t)nscan * n_thk * nxy * sizeof(double2) : 0;
        const size_t min_waves = 2 * nxy * sizeof(Z);
        if (coh_bytes + min_waves > budget)
            throw Error{MT_ERR_OUT_OF_MEMORY, "coherent accumulators exceed the memory budget"};
        const size_t stack_bytes = coh_bytes + store.plan(budget - coh_bytes - min_waves, 1);
        const size_t per_wave = (m.tilted() ? 2 : 1) * wave * sizeof(Z);
        const int B_cap = (int)std::min<size_t>((size_t)std::max(1, nscan), (budget - stack_bytes) / per_wave);
        if (B_cap < 1) throw Error{MT_ERR_OUT_OF_MEMORY, "no room for a wave batch"};
        ntile = m.ny / col_tile_w<R>(m.nx);
        std::vector<double> zl;
        for (int i = 0; i < m.n_slice(); ++i) zl.push_back((double)m.dz_slice(i));
        for (const auto &t : m.thk) zl.push_back((double)t.z_back_prop);
        S.prop_tables(zl, px, py);

        B = batch_override;
        if (B <= 0) {
            // enough waves to fill 256 CUs in the column pass, and a working set
            // that stays near the 256 MB last-level cache; the tuner may double
            // or quadruple it
            const int nt = m.ny / col_tile_w<R>(m.nx);
            B = std::max(1, (int)std::min<size_t>(256, (size_t)(160u << 20) / (nxy * sizeof(Z))));
            B = std::max(B, (2048 + nt - 1) / nt);
            B = std::min(B, B_cap);
            if (run_control().autotune != 0) B = choose_batch(B, std::min(B_cap, std::max(B, 256)));   // B0..2 B0
            B = balanced(std::min(B, B_cap));
        }
        B = std::min(B, B_cap);
        nbatch = (nscan + B - 1) / B;

        psi.alloc((size_t)B * wave);
        members.alloc(B);
        beams.alloc(B);
        part.alloc((size_t)B * n_thk * n_det * ntile);
        image.alloc((size_t)n_thk * n_det * nscan);
        {
            std::vector<double> zero(image.n, 0.0);
            image.upload(zero.data(), zero.size(), s);
        }
        if (m.tilted()) {
            S.ramp_tables(rx, ry);
            scratch.alloc((size_t)B * wave);
        }
        if (m.pn_coh_contrib) {
            coh.alloc((size_t)nscan * n_thk * nxy);
            check(hipMemsetAsync(coh.p, 0, coh.n * sizeof(double2), s), "hipMemsetAsync");
        }
        nap = S.count_aperture();
        P = S.probe_params(nap);
        w = m.phonon_weight();
        if (m.iw_type == 3) S.user_wave_spectrum(fpsi0);
        {
            // STEM integrates the defocus spread only (src/tem_simulation.cuh:138-178)
            Quad1 qt;
            Quad2 qs;
            if (m.full_integration()) m.cond_quadratures(qt, qs);
            else qt = Quad1{{0.0}, {1.0}};
            S.cond_defocus_points(qt, tc10, twq);
            n_t = (int)tc10.size();
        }
        // the captured pipeline reads the atom buffers (rebuilt slices): their
        // addresses are fixed before any capture
        S.reserve(atoms);
        hm.resize(B);
        {
            std::vector<Member> am((size_t)n_t * nbatch * B);
            std::vector<Beam> ab(am.size());
            for (int it = 0; it < n_t; ++it)
                for (int is = 0; is < nbatch * B; ++is) {
                    const bool valid = is < nscan;
                    const int iq = valid ? is : nscan - 1;
                    const size_t i = (size_t)it * nbatch * B + is;
                    am[i] = Member{0, valid ? is : -1, is % B, 1, twq[it]};
                    ab[i] = S.beam(m.sc_x[iq], m.sc_y[iq], tc10[it]);
                }
            all_members.alloc(am.size());
            all_beams.alloc(ab.size());
            all_members.upload(am.data(), am.size(), s);
            all_beams.upload(ab.data(), ab.size(), s);
        }

        // slice-step arguments shared by the tuner and the pipeline
        r.src = psi.p; r.dst = psi.p; r.member_stride = (long long)wave; r.pitch = pitch;
        r.nrows = m.nx; r.n_members = B;
        r.tw = S.tw_y.p; r.trans = store.trans.p; r.tslot_stride = 0; r.members = members.p;
        r.flags = ROW_LOAD | ROW_IFFT | ROW_MUL_T | ROW_FFT | ROW_STORE;
        c = S.col_base(psi.p, psi.p, B);
        c.member_stride = (long long)wave; c.pitch = pitch;
        c.n_det = n_det;
        for (int d = 0; d < n_det; ++d) { c.det_g2_min[d] = m.det_g2_min[d]; c.det_g2_max[d] = m.det_g2_max[d]; }
        if (!m.det_fR.empty()) {
            det_fR.alloc(m.det_fR.size());
            det_fR.upload(m.det_fR.data(), m.det_fR.size(), s);
        }
        c.det_fR = det_fR.p;
        c.det_part = part.p; c.n_thk = n_thk;
        c.members = members.p;
        c.coh = coh.p; c.coh_w = (double)w;

        // launch variants, picked before any pipeline is captured
        RowArgsT<Z> rt = r;
        rt.slice = 0;
        ColArgsT<Z> ct = c;
        ct.flags = col_flags(false);
        ct.px = px.p; ct.py = py.p; ct.ithk = -1;
        ColArgsT<Z> cl = c;
        cl.flags = col_flags(true) | (det_fR.p ? COL_DET_FR : 0);
        cl.px = px.p; cl.py = py.p; cl.ithk = m.slices[ns_run - 1].ithk;
        for (int j = 0; j < B; ++j) hm[j] = Member{0, -1, j, 1, 1.0};   // valid entries while tuning
        members.upload(hm.data(), B, s);
        check(hipMemsetAsync(psi.p, 0, psi.n * sizeof(Z), s), "hipMemsetAsync");
        choice = choose_variants(m.nx, m.ny, B, rt, ct, cl.ithk >= 0 ? &cl : nullptr, s);
        r.variant = choice.row;

        // concurrent chains: two for grids of 512^2 and more with a batch of
        // at least 128 MiB of waves (measured: C2, 512^2 x 118 waves, 5.8 %
        // faster; 256^2 batches were slower split), when the transmission
        // stack is resident (a rebuilt slice is one buffer that every chain
        // reads) and the scan fills more than one chain
        const int rq = run_control().chains;
        const bool auto2 = nxy >= 512u * 512u && (size_t)B * nxy >= 64u * 512u * 512u && B >= 4 && nscan > B / 2;
        n_chain = rq > 0 ? rq : (auto2 ? 2 : 1);
        if (!store.resident) n_chain = 1;
        n_chain = std::max(1, std::min({n_chain, 4, B}));
        for (int k = 0; k < n_chain && n_chain > 1; ++k) {
            cs.emplace_back(new Stream());
            joins.emplace_back(new Event());
        }
    }

    // The narrowest batch that covers the scan in as many batches as width b
    // does: the last batch then pads at most nbatch - 1 waves, instead of up
    // to b - 1 (padding waves cost a full wave's work).
    int balanced(int b) const
    {
        if (b <= 0 || nscan <= 0) return b;
        const int nb = (nscan + b - 1) / b;
        return (nscan + nb - 1) / nb;
    }

    // Batch width: B0, 1.5 B0 and 2 B0 (up to B_max, each balanced over the
    // scan) timed on the full slice step (launch variants tuned for the
    // first); the cost of a width is its time per wave times the waves it
    // runs, padding included.  The smallest width within 2 % of the best
    // wins.  Cached per (grid, B0, B_max, scan length).
    int choose_batch(int B0, int B_max)
    {
        std::vector<int> cand;
        for (int b : {B0, B0 + B0 / 2, 2 * B0}) {
            if (b > B_max) continue;
            b = balanced(b);
            if (cand.empty() || b > cand.back()) cand.push_back(b);
        }
        if (cand.size() < 2) return B0;
        static std::map<std::tuple<int, int, int, int, int>, int> cache;   // one per precision
        const auto key = std::make_tuple(m.nx, m.ny, B0, B_max, nscan);
        const bool reuse = run_control().autotune == 1;
        if (reuse) {
            auto it = cache.find(key);
            if (it != cache.end()) {
                batch_log = "batch=" + std::to_string(it->second) + "(cached)";
                return it->second;
            }
        }
        const int bm = cand.back();
        DBuf<Z> w((size_t)bm * wave);
        DBuf<Member> mem(bm);
        std::vector<Member> hmv(bm, Member{0, -1, 0, 1, 1.0});
        mem.upload(hmv.data(), bm, s);
        check(hipMemsetAsync(w.p, 0, w.n * sizeof(Z), s), "hipMemsetAsync");
        std::ostringstream log;
        log << "batch[";
        double best = -1;
        std::vector<double> per(cand.size());
        auto args = [&](int b, RowArgsT<Z> &rr, ColArgsT<Z> &cc) {
            rr = RowArgsT<Z>{};
            rr.src = w.p; rr.dst = w.p; rr.member_stride = (long long)wave; rr.pitch = pitch;
            rr.nrows = m.nx; rr.n_members = b;
            rr.tw = S.tw_y.p; rr.trans = store.trans.p; rr.members = mem.p; rr.slice = 0;
            rr.flags = ROW_LOAD | ROW_IFFT | ROW_MUL_T | ROW_FFT | ROW_STORE;
            cc = S.col_base(w.p, w.p, b);
            cc.member_stride = (long long)wave; cc.pitch = pitch;
            cc.flags = col_flags(false);
            cc.px = px.p; cc.py = py.p; cc.ithk = -1;
        };
        // the candidates are compared with the launch variants the tuner picks
        // for the smallest one, so that the comparison sees tuned passes
        PassChoice pc0;
        {
            RowArgsT<Z> rr;
            ColArgsT<Z> cc;
            args(cand[0], rr, cc);
            pc0 = choose_variants(m.nx, m.ny, cand[0], rr, cc, (const ColArgsT<Z> *)nullptr, s);
        }
        for (size_t i = 0; i < cand.size(); ++i) {
            RowArgsT<Z> rr;
            ColArgsT<Z> cc;
            args(cand[i], rr, cc);
            rr.variant = pc0.row;
            cc.variant = pc0.col;
            auto step = [&] {
                hipError_t e = launch_row_pass(m.ny, rr, s);
                return e != hipSuccess ? e : launch_col_pass(m.nx, cc, false, false, s);
            };
            const double t1 = time_variant(step, s), t2 = time_variant(step, s);
            const double t = (t1 > 0 && t2 > 0) ? std::min(t1, t2) : std::max(t1, t2);
            const int nb = (std::max(nscan, 1) + cand[i] - 1) / cand[i];
            per[i] = t * nb;   // the whole scan at this width
            if (per[i] > 0 && (best < 0 || per[i] < best)) best = per[i];
            log << (i ? "," : "") << cand[i] << ":" << (int)std::lround(t * 1000.0) << "us";
        }
        int pick = B0;
        for (size_t i = 0; i < cand.size(); ++i)
            if (per[i] > 0 && per[i] <= best * 1.02) { pick = cand[i]; break; }
        log << "]=" << pick;
        batch_log = log.str();
        cache[key] = pick;
        return pick;
    }

    int col_flags(bool last) const
    {
        // the (weak) phase object is not propagated inside the specimen, so its
        // slice step has no bandwidth limit (Wave_Function::psi_slice,
        // src/wave_function.cuh:229-238); the back-propagation's still applies
        return COL_LOAD | COL_FFT | COL_MUL_P | (m.bwl && m.multislice() ? COL_BWL : 0) | (m.bwl ? COL_DET_BWL : 0) |
               (last ? 0 : (COL_IFFT | COL_STORE));
    }

    // work item k: configuration k / (n_t nbatch), defocus point (k / nbatch) % n_t,
    // probe batch k % nbatch
    int n_work() const { return m.n_conf_run() * n_t * nbatch; }
    int batch_size(int k) const { return std::min(B, nscan - (k % nbatch) * B); }

    std::string plan() const
    {
        std::ostringstream o;
        o << "stem " << (sizeof(R) == 8 ? "fp64" : "fp32") << " nx=" << m.nx << " ny=" << m.ny << " batch=" << B << " slices=" << ns_run
          << " row_variant=" << choice.row << " col_variant=" << choice.col << " col_det_variant=" << choice.col_det
          << " graph=" << (run_control().use_graph ? 1 : 0) << " aperture_px=" << nap;
        if (n_chain > 1) o << " chains=" << n_chain;
        if (n_t > 1) o << " defocus_points=" << n_t;
        if (!store.resident) o << " " << store.describe();
        return o.str();
    }

    // Inputs of work item k.  A new configuration is loaded and its
    // transmission built on the job's stream after everything enqueued there
    // (gate, if given, is an event that stream waits for first: the end of
    // the work that still reads the previous configuration); the host
    // returns once the build is done.  The batch's member and beam tables
    // are copied on stream st, stream-ordered before the pipeline that reads
    // them: no host wait.
    void prepare(int k, hipStream_t st, const Event *gate)
    {
        const int iconf = m.fp_iconf_0 + k / (n_t * nbatch), it = (k / nbatch) % n_t, ib = k % nbatch;
        if (iconf != cur_conf) {
            if (gate) check(hipStreamWaitEvent(s, gate->e, 0), "hipStreamWaitEvent");
            S.load(iconf, atoms, s);
            Event t0, t1;
            record(t0, s);
            store.build(atoms, 0);
            record(t1, s);
            wait(t1, s);
            ms_tf += elapsed(t0, t1);
            cur_conf = iconf;
        }
        const size_t row = ((size_t)it * nbatch + ib) * B;
        check(hipMemcpyAsync(members.p, all_members.p + row, B * sizeof(Member), hipMemcpyDeviceToDevice, st),
              "hipMemcpyAsync D2D");
        check(hipMemcpyAsync(beams.p, all_beams.p + row, B * sizeof(Beam), hipMemcpyDeviceToDevice, st),
              "hipMemcpyAsync D2D");
    }

    // the batch pipeline on stream st; ev (optional): 3 events per slice
    // bracketing the row and column passes
    // The batch's device work on stream st.  With n_chain > 1 the members
    // are split into that many chains, each on its own stream (forked from
    // st and joined back before the detector reduction): while one chain's
    // pass drains its last workgroups, the other's fills the CUs.  Every
    // wave's arithmetic is the same either way.  ev (per-pass events) runs
    // one chain.
    void pipeline(hipStream_t st, std::vector<Event> *ev)
    {
        const int K = ev ? 1 : n_chain;
        if (K <= 1) {
            chain(st, ev, 0, B);
        } else {
            check(hipEventRecord(fork.e, st), "hipEventRecord");
            const int per = (B + K - 1) / K;
            for (int k = 0; k < K && k * per < B; ++k) {
                check(hipStreamWaitEvent(cs[k]->s, fork.e, 0), "hipStreamWaitEvent");
                chain(cs[k]->s, nullptr, k * per, std::min(per, B - k * per));
                check(hipEventRecord(joins[k]->e, cs[k]->s), "hipEventRecord");
                check(hipStreamWaitEvent(st, joins[k]->e, 0), "hipStreamWaitEvent");
            }
        }
        check(launch_det_reduce(part.p, B, n_thk, n_det, ntile, members.p, (double)w, image.p, nscan, st),
              "det reduce");
    }

    // members m0 .. m0 + nm - 1 of the batch: probe generation, the slice
    // loop, detector partials (the member tables are indexed from m0; beam
    // indices in them are batch-wide)
    void chain(hipStream_t st, std::vector<Event> *ev, int m0, int nm)
    {
        Z *wv = psi.p + (size_t)m0 * wave;
        RowArgsT<Z> r = this->r;
        ColArgsT<Z> c = this->c;
        r.src = r.dst = wv; r.n_members = nm; r.members = members.p + m0;
        c.src = c.dst = wv; c.n_members = nm; c.members = members.p + m0;
        c.det_part = part.p + (size_t)m0 * n_thk * n_det * ntile;
        ColArgsT<Z> c0 = S.col_base(nullptr, wv, nm);
        c0.member_stride = (long long)wave; c0.pitch = pitch;
        c0.flags = COL_IFFT | COL_STORE | (m.iw_type == 3 ? COL_USER_WAVE : 0);
        c0.fpsi0 = fpsi0.p;
        c0.members = members.p + m0; c0.beams = beams.p; c0.probe = P;
        check(launch_col_pass(m.nx, c0, true, false, st), "probe pass");
        int k = 0;
        for (int is = 0; is < ns_run; ++is) {
            store.for_slice(atoms, is, r, st);
            if (ev) record((*ev)[k++], st);
            check(launch_row_pass(m.ny, r, st), "row pass");
            if (ev) record((*ev)[k++], st);
            const int ithk = m.slices[is].ithk;
            c.px = px.p + (size_t)is * m.nx;
            c.py = py.p + (size_t)is * m.ny;
            if (m.tilted() && ithk >= 0) {
                // complete the slice step, then detect a copy with the tilt ramp
                // removed: O = FFT[psi exp(-2 pi i g0.r)] x P_back (get_psi,
                // src/wave_function.cuh:96-105)
                c.flags = col_flags(false);
                c.ithk = -1;
                c.variant = choice.col;
                check(launch_col_pass(m.nx, c, false, false, st), "col pass");
                RowArgsT<Z> ro = r;
                ro.dst = scratch.p + (size_t)m0 * wave;
                ro.flags = ROW_LOAD | ROW_IFFT | ROW_MUL_RAMP | ROW_FFT | ROW_STORE;
                ro.ramp_x = rx.p; ro.ramp_y = ry.p;
                check(launch_row_pass(m.ny, ro, st), "output ramp pass");
                ColArgsT<Z> cd = c;
                cd.src = scratch.p + (size_t)m0 * wave;
                cd.flags = COL_LOAD | COL_FFT | COL_MUL_P | (m.bwl ? COL_BWL : 0) | (m.pn_coh_contrib ? COL_COH : 0) |
                           (det_fR.p ? COL_DET_FR : 0);
                cd.px = px.p + (size_t)(m.n_slice() + ithk) * m.nx;
                cd.py = py.p + (size_t)(m.n_slice() + ithk) * m.ny;
                cd.ithk = ithk;
                cd.variant = 0;
                // the stored wave lacks the 1/ny that the next row pass would supply
                cd.p_scale = (R)m.ny;
                check(launch_col_pass(m.nx, cd, false, true, st), "output detector pass");
            } else {
                c.flags = col_flags(is == ns_run - 1) | ((ithk >= 0 && m.pn_coh_contrib) ? COL_COH : 0) |
                          ((ithk >= 0 && det_fR.p) ? COL_DET_FR : 0);
                c.ithk = ithk;
                c.variant = ithk >= 0 ? choice.col_det : choice.col;
                check(launch_col_pass(m.nx, c, false, ithk >= 0, st), "col pass");
            }
            if (ev) record((*ev)[k++], st);
        }
    }

    void finish(mt_output_multislice *out)
    {
        if (out->image_tot) {
            check(hipMemcpyAsync(out->image_tot, image.p, image.n * sizeof(double), hipMemcpyDeviceToHost, s), "D2H");
            check(hipStreamSynchronize(s), "sync");
        }
        if (out->image_coh && m.pn_coh_contrib) {
            DBuf<double> ic((size_t)n_thk * n_det * nscan);
            check(launch_coh_image(coh.p, nscan, n_thk, n_det, m.nx, m.ny, m.dgx, m.dgy, m.det_g2_min.data(),
                                   m.det_g2_max.data(), (const R *)det_fR.p, (m.bwl && !m.tilted()) ? 1 : 0,
                                   m.bwl_alpha, m.gl2_max, ic.p,
                                   s),
                  "coherent image");
            check(hipMemcpyAsync(out->image_coh, ic.p, ic.n * sizeof(double), hipMemcpyDeviceToHost, s), "D2H");
            check(hipStreamSynchronize(s), "sync");
        }
    }
};

// Begin/end events of the last two batches enqueued on one stream.  begin()
// first waits for the batch that used the slot two batches ago and returns
// its time (ms); drain() waits for both and returns their times.
struct BatchRing {
    Event b[2], e[2];
    bool pending[2] = {false, false};
    int k = 0;
    double collect(int i)
    {
        if (!pending[i]) return 0.0;
        pending[i] = false;
        check(hipEventSynchronize(e[i].e), "hipEventSynchronize");
        return events_enabled() ? (double)elapsed(b[i], e[i]) : 0.0;
    }
    double begin(hipStream_t s)
    {
        const double t = collect(k & 1);
        record(b[k & 1], s);
        return t;
    }
    void end(hipStream_t s)
    {
        check(hipEventRecord(e[k & 1].e, s), "hipEventRecord");
        pending[k & 1] = true;
        ++k;
    }
    // end event of the last batch enqueued, or null
    const Event *last() const { return k > 0 && pending[(k - 1) & 1] ? &e[(k - 1) & 1] : nullptr; }
    double drain() { return collect(k & 1) + collect((k + 1) & 1); }
};

struct GraphExec {
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
    GraphExec() = default;
    GraphExec(const GraphExec &) = delete;
    GraphExec &operator=(const GraphExec &) = delete;
    ~GraphExec()
    {
        if (e) (void)hipGraphExecDestroy(e);
        if (g) (void)hipGraphDestroy(g);
    }
};

template <class R>
void run_stem_one(const ModelT<R> &m, mt_output_multislice *out, size_t budget, bool reset_progress)
{
    Timing &tm = last_timing();
    tm = Timing{};
    StemJob<R> J(m, budget, run_control().batch_override);
    hipStream_t s = J.s;
    tm.plan = J.plan();
    tm.tune_log = (J.batch_log.empty() ? "" : J.batch_log + " ") + J.choice.log;
    if (reset_progress) {
        run_control().niter = J.n_work() / J.nbatch * J.nscan;
        run_control().iter = 0;
    }

    GraphExec G;
    if (run_control().use_graph) {
        check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
        J.pipeline(s, nullptr);
        check(hipStreamEndCapture(s, &G.g), "end capture");
        check(hipGraphInstantiate(&G.e, G.g, nullptr, nullptr, 0), "graph instantiate");
    }

    std::vector<Event> kev(run_control().time_kernels ? 3 * J.ns_run : 0);
    Event t0, t1;
    // Two batches in flight: the host enqueues batch k while batch k-1 runs,
    // and collects batch k-2's time (its events are reused for batch k).
    BatchRing ring;
    double ms_slice = 0, ps = 0;
    record(t0, s);
    int last_conf = -1;
    for (int k = 0; k < J.n_work(); ++k) {
        stop_check();
        const int iconf = k / (J.n_t * J.nbatch);
        if (iconf != last_conf) ms_slice += ring.drain();   // the configuration's build waits for the stream anyway
        J.prepare(k, s, nullptr);
        // per-kernel events on the first batch of each configuration
        const bool timed = run_control().time_kernels && iconf != last_conf;
        last_conf = iconf;
        ms_slice += ring.begin(s);
        if (timed || !G.e) J.pipeline(s, timed ? &kev : nullptr);
        else check(hipGraphLaunch(G.e, s), "graph launch");
        ring.end(s);
        ps += (double)J.batch_size(k) * J.ns_run;
        if (timed) {
            ms_slice += ring.drain();
            for (int is = 0; is < J.ns_run; ++is) {
                tm.ms_row += elapsed(kev[3 * is], kev[3 * is + 1]);
                tm.ms_col += elapsed(kev[3 * is + 1], kev[3 * is + 2]);
            }
            tm.n_row += J.ns_run;
            tm.n_col += J.ns_run;
        }
        progress_add(J.batch_size(k));
    }
    ms_slice += ring.drain();
    record(t1, s);
    wait(t1, s);
    tm.ms_total = elapsed(t0, t1);
    tm.ms_slice = ms_slice;
    tm.ms_tf = J.ms_tf;
    tm.probe_slices = ps;
    J.finish(out);
}

// STEM.  The coherent image keeps FP64 sums of every scan point's wave
// resident; when they would take more than half of the memory left beside
// the transmission stack, the scan runs in blocks of points, each with all
// configurations (scan points are independent: Multislice::STEM_ISTEM,
// src/tem_simulation.cuh:115-133, loops them outermost).
template <class R>
void run_stem(const ModelT<R> &m, mt_output_multislice *out)
{
    const size_t budget = budget_bytes();
    const size_t nxy = (size_t)m.nx * m.ny, nscan = m.sc_x.size(), n_thk = m.thk.size();
    const size_t stack = (size_t)m.n_slice() * nxy * sizeof(cplx_t<R>), wave = 2 * nxy * sizeof(cplx_t<R>);
    const size_t per_scan = n_thk * nxy * sizeof(double2);
    if (!m.pn_coh_contrib || stack + wave >= budget || nscan * per_scan <= (budget - stack) / 2) {
        run_stem_one(m, out, budget, true);
        return;
    }
    size_t nblk = ((budget - stack) / 2) / per_scan;
    if (nblk < 1) nblk = (budget - stack - wave) / per_scan;
    if (nblk < 1) throw Error{MT_ERR_OUT_OF_MEMORY, "coherent accumulators of one scan point exceed the memory budget"};
    const int n_det = (int)m.det_g2_min.size();
    run_control().niter = (int)(nscan * (size_t)m.n_conf_run());   // the coherent loop: every configuration per point
    run_control().iter = 0;
    Timing total;
    int nb = 0;
    for (size_t s0 = 0; s0 < nscan; s0 += nblk, ++nb) {
        const size_t n = std::min(nblk, nscan - s0);
        ModelT<R> mb = m;
        mb.sc_x.assign(m.sc_x.begin() + s0, m.sc_x.begin() + s0 + n);
        mb.sc_y.assign(m.sc_y.begin() + s0, m.sc_y.begin() + s0 + n);
        std::vector<double> it(out->image_tot ? n_thk * n_det * n : 0), ic(out->image_coh ? n_thk * n_det * n : 0);
        mt_output_multislice ob = *out;
        ob.image_tot = out->image_tot ? it.data() : nullptr;
        ob.image_coh = out->image_coh ? ic.data() : nullptr;
        run_stem_one(mb, &ob, budget, false);
        // images [n_thk][n_det][scan]: the block's points are a contiguous run of the scan
        for (size_t k = 0; k < n_thk * n_det; ++k) {
            if (out->image_tot) std::copy(it.begin() + k * n, it.begin() + (k + 1) * n, out->image_tot + k * nscan + s0);
            if (out->image_coh) std::copy(ic.begin() + k * n, ic.begin() + (k + 1) * n, out->image_coh + k * nscan + s0);
        }
        const Timing &t = last_timing();
        total.ms_total += t.ms_total; total.ms_slice += t.ms_slice; total.probe_slices += t.probe_slices;
        total.ms_row += t.ms_row; total.ms_col += t.ms_col; total.n_row += t.n_row; total.n_col += t.n_col;
        total.ms_tf += t.ms_tf;
        if (s0 == 0) { total.plan = t.plan; total.tune_log = t.tune_log; }
    }
    total.plan += " scan_blocks=" + std::to_string(nb) + " scan_points_per_block=" + std::to_string(nblk);
    last_timing() = total;
}

// Several STEM specimens, possibly of different grid sizes, advance
// together: each round gives every specimen with work left its next work
// item, and the round's device work -- one branch per specimen, each with its
// own tuned kernels -- is one hipGraph launch (branches forked onto
// per-specimen streams during capture, joined at the end).  Graphs are kept
// per set of active specimens.
template <class R>
void run_stem_group(const std::vector<const ModelT<R> *> &ms, const std::vector<mt_output_multislice *> &outs)
{
    Timing &tm = last_timing();
    tm = Timing{};
    const size_t n = ms.size();
    if (n == 0 || n > 16) throw Error{MT_ERR_INVALID_ARGUMENT, "group of 1..16 specimens"};
    // memory: each specimen gets a share of the budget proportional to its
    // transmission stack plus a full batch
    std::vector<double> want(n);
    double tot = 0;
    for (size_t i = 0; i < n; ++i) {
        const double nxy = (double)ms[i]->nx * ms[i]->ny;
        want[i] = nxy * (double)sizeof(cplx_t<R>) * (ms[i]->n_slice() + 256.0);
        tot += want[i];
    }
    const size_t budget = budget_bytes();
    std::vector<std::unique_ptr<StemJob<R>>> jobs;
    std::ostringstream plan, tune;
    int rounds = 0, niter = 0;
    for (size_t i = 0; i < n; ++i) {
        const size_t share = (size_t)((double)budget * want[i] / tot);
        jobs.emplace_back(new StemJob<R>(*ms[i], share, run_control().batch_override));
        plan << (i ? " | " : "") << jobs.back()->plan();
        tune << (i ? " | " : "") << jobs.back()->batch_log << (jobs.back()->batch_log.empty() ? "" : " ")
             << jobs.back()->choice.log;
        rounds = std::max(rounds, jobs.back()->n_work());
        niter += jobs.back()->n_work() / jobs.back()->nbatch * jobs.back()->nscan;
    }
    tm.plan = "group " + plan.str();
    tm.tune_log = tune.str();
    run_control().niter = niter;
    run_control().iter = 0;

    Stream main;
    hipStream_t s = main.s;
    std::vector<std::unique_ptr<Stream>> bs;
    std::vector<std::unique_ptr<Event>> join;
    for (size_t i = 0; i < n; ++i) {
        bs.emplace_back(new Stream());
        join.emplace_back(new Event());
    }
    Event fork;
    std::map<unsigned, std::unique_ptr<GraphExec>> graphs;
    auto enqueue = [&](unsigned mask) {
        check(hipEventRecord(fork.e, s), "hipEventRecord");
        for (size_t i = 0; i < n; ++i) {
            if (!(mask >> i & 1u)) continue;
            check(hipStreamWaitEvent(bs[i]->s, fork.e, 0), "hipStreamWaitEvent");
            jobs[i]->pipeline(bs[i]->s, nullptr);
            check(hipEventRecord(join[i]->e, bs[i]->s), "hipEventRecord");
            check(hipStreamWaitEvent(s, join[i]->e, 0), "hipStreamWaitEvent");
        }
    };

    Event t0, t1;
    BatchRing ring;
    double ms_slice = 0, ps = 0;
    record(t0, s);
    for (int k = 0; k < rounds; ++k) {
        stop_check();
        unsigned mask = 0;
        for (size_t i = 0; i < n; ++i)
            if (k < jobs[i]->n_work()) {
                // a new configuration's build waits for the last round enqueued
                jobs[i]->prepare(k, s, ring.last());
                mask |= 1u << i;
            }
        ms_slice += ring.begin(s);
        if (run_control().use_graph) {
            auto it = graphs.find(mask);
            if (it == graphs.end()) {
                std::unique_ptr<GraphExec> G(new GraphExec());
                check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
                enqueue(mask);
                check(hipStreamEndCapture(s, &G->g), "end capture");
                check(hipGraphInstantiate(&G->e, G->g, nullptr, nullptr, 0), "graph instantiate");
                it = graphs.emplace(mask, std::move(G)).first;
            }
            check(hipGraphLaunch(it->second->e, s), "graph launch");
        } else {
            enqueue(mask);
        }
        ring.end(s);
        for (size_t i = 0; i < n; ++i)
            if (mask >> i & 1u) {
                ps += (double)jobs[i]->batch_size(k) * jobs[i]->ns_run;
                progress_add(jobs[i]->batch_size(k));
            }
    }
    ms_slice += ring.drain();
    record(t1, s);
    wait(t1, s);
    tm.ms_total = elapsed(t0, t1);
    tm.ms_slice = ms_slice;
    tm.probe_slices = ps;
    for (size_t i = 0; i < n; ++i) {
        tm.ms_tf += jobs[i]->ms_tf;
        jobs[i]->finish(outs[i]);
    }
}

// ---------------------------------------------------------------------------
// Exit wave, diffraction and image simulations (EWFS, EWRS, CBED, CBEI, ED,
// HRTEM, ISTEM, PED, HCTEM): the members of a launch are the waves of one or
// more phonon configurations.
//
// One wave in the reference's loop order: its beams, integration weight and
// tilt azimuth.
struct EwWave {
    int beam0, nbeam;
    double wq;
    int irot;
};

template <class R>
void run_ew(const ModelT<R> &m, mt_output_multislice *out)
{
    using Z = cplx_t<R>;
    Stream stream;
    hipStream_t s = stream.s;
    Setup<R> S(m, s);
    Timing &tm = last_timing();
    tm = Timing{};
    const size_t nxy = S.nxy;
    const int n_thk = (int)m.thk.size();
    mt_output_multislice dims{};
    m.output_dims(&dims);
    const int ot = dims.output_type;
    const int nxo = m.oa_ix_e - m.oa_ix_0, nyo = m.oa_iy_e - m.oa_iy_0;
    const size_t no = (size_t)nxo * nyo;
    const bool fs = m.is_grid_fs();     // Fourier-space output (EWFS, CBED, ED, PED)
    const bool img = m.is_image();      // objective-lens image (HRTEM, ISTEM, CBEI, HCTEM)
    const bool ramp = m.out_ramp();
    const int ns_run = [&] {
        int last = 0;
        for (const auto &t : m.thk) last = std::max(last, t.islice);
        return last + 1;
    }();

    const int nconf = m.n_conf_run();
    // Waves of one configuration, all sharing its transmission stack, in the
    // loop order of the reference:
    //  - CBED / CBEI fully integrated: (source point, defocus point), beams
    //    shifted by the source offset (CBED_CBEI, src/tem_simulation.cuh:207-261);
    //  - ISTEM: (defocus point, scan position) (STEM_ISTEM, :100-200; the
    //    defocus spread only when fully integrated);
    //  - PED / HCTEM: the tilt azimuths (PED_HCTEM, :274-301);
    //  - otherwise one wave.
    Quad1 qt{{0.0}, {1.0}};
    Quad2 qs{{0.0}, {0.0}, {1.0}};
    if (m.full_integration() && (m.is_cbed_cbei() || m.is_istem())) m.cond_quadratures(qt, qs);
    if (m.is_istem()) qs = Quad2{{0.0}, {0.0}, {1.0}};
    std::vector<double> tc10, twq;
    S.cond_defocus_points(qt, tc10, twq);
    const int n_s = (int)qs.w.size(), n_t = (int)tc10.size();
    const int n_pos = m.is_istem() ? (int)m.sc_x.size() : 1;
    const int n_rot = m.nrot;
    const int nq = n_s * n_t * n_pos * n_rot;
    // beams of a wave: every iw_x, iw_y of a convergent wave (the scan
    // position in ISTEM); the first for a user wave
    const int nbeam = m.iw_type == 2 ? (m.is_istem() ? 1 : (int)m.iw_x.size()) : (m.iw_type == 3 ? 1 : 0);
    std::vector<EwWave> wv(nq);
    std::vector<Beam> hb;
    for (int i_s = 0; i_s < n_s; ++i_s)
        for (int i_t = 0; i_t < n_t; ++i_t)
            for (int ip = 0; ip < n_pos; ++ip)
                for (int ir = 0; ir < n_rot; ++ir) {
                    const int q = ((i_s * n_t + i_t) * n_pos + ip) * n_rot + ir;
                    wv[q] = EwWave{q * nbeam, nbeam, qs.w[i_s] * twq[i_t], ir};
                    for (int ib = 0; ib < nbeam; ++ib) {
                        if (m.is_istem()) {
                            hb.push_back(S.beam(m.sc_x[ip], m.sc_y[ip], tc10[i_t]));
                        } else {
                            // beam_x = iw_x + source offset, stored in R (src/tem_simulation.cuh:233-237)
                            const R xb = (R)((double)m.iw_x[ib] + qs.x[i_s]);
                            const R yb = (R)((double)m.iw_y[ib] + qs.y[i_s]);
                            hb.push_back(S.beam(xb, yb, tc10[i_t]));
                        }
                    }
                }

    // Configurations per group (their stacks resident) and waves per launch:
    // all waves of the group while they fit the wave cap and the memory, else
    // one configuration at a time in chunks of waves; when even one stack and
    // one wave do not fit, the configuration's stored potentials are rebuilt
    // slice by slice (SliceStore).
    SliceStore<R> store(S);
    const size_t stack_bytes = store.stack_bytes();
    const size_t wave_bytes = (size_t)(img && ramp ? 3 : 2) * nxy * sizeof(Z);
    const size_t fixed = (size_t)n_thk * (no * 3 + 2 * nxy) * sizeof(double);
    const size_t budget = budget_bytes();
    if (fixed + store.slot_bytes() + wave_bytes > budget)
        throw Error{MT_ERR_OUT_OF_MEMORY, "one wave and one transmission slice exceed the memory budget"};
    const size_t avail = budget - fixed;
    const int wave_cap = run_control().batch_override > 0 ? run_control().batch_override : 256;
    int K = 0, W;
    if (nq <= wave_cap)
        K = (int)std::min<size_t>({(size_t)nconf, avail / (stack_bytes + (size_t)nq * wave_bytes),
                                   (size_t)std::max(1, wave_cap / nq)});
    if (K >= 1) {
        store.plan((size_t)K * stack_bytes, K);
        W = K * nq;
    } else {
        K = 1;
        const size_t taken = store.plan(avail - wave_bytes, 1);
        W = (int)std::min<size_t>((size_t)std::min(nq, wave_cap), (avail - taken) / wave_bytes);
        W = std::max(W, 1);
    }

    DBuf<Z> psi((size_t)W * nxy), scratch((size_t)W * nxy);
    DBuf<Member> members((size_t)W);
    DBuf<Beam> beams(std::max<size_t>(1, hb.size()));
    if (!hb.empty()) beams.upload(hb.data(), hb.size(), s);
    DBuf<double> m2(std::max<size_t>(1, (size_t)n_thk * no)), pc(std::max<size_t>(1, (size_t)n_thk * no * 2));
    const bool need_full = (ot == MT_OT_M2PSI_TOT_COH);
    DBuf<double> pfull(need_full ? (size_t)n_thk * nxy * 2 : 1);
    {
        std::vector<double> zero(std::max(m2.n, std::max(pc.n, pfull.n)), 0.0);
        m2.upload(zero.data(), m2.n, s);
        pc.upload(zero.data(), pc.n, s);
        pfull.upload(zero.data(), pfull.n, s);
    }
    std::vector<double> zl;
    for (int i = 0; i < m.n_slice(); ++i) zl.push_back((double)m.dz_slice(i));
    for (const auto &t : m.thk) zl.push_back((double)t.z_back_prop);
    // propagator (and output ramp) tables, one set per tilt azimuth, picked
    // per wave by Member::irot
    DBuf<Z> px, py;
    S.prop_tables(zl, px, py, n_rot);
    const long long rot_px = n_rot > 1 ? (long long)zl.size() * m.nx : 0;
    const long long rot_py = n_rot > 1 ? (long long)zl.size() * m.ny : 0;
    const int col_rot = n_rot > 1 ? COL_ROT : 0;   // keeps the specialised slice steps for one azimuth
    DBuf<Z> rx, ry;
    if (ramp) S.ramp_tables(rx, ry, n_rot);
    DBuf<Z> ctf;
    std::vector<double> cwq{1.0};   // weights of the objective-lens tables
    if (img) S.ctf_tables(ctf, cwq);
    const int n_ctf = (int)cwq.size();
    // an image run with the output ramp that reads the ramp output more than
    // once (several lens tables, or the coherent pass) keeps it apart from
    // the output passes
    DBuf<Z> scratch2(img && ramp && (n_ctf > 1 || need_full) ? (size_t)W * nxy : 0);
    ProbeParamsT<R> P{};
    DBuf<Z> fpsi0;
    if (m.iw_type == 3) S.user_wave_spectrum(fpsi0);
    if (m.iw_type >= 2) P = S.probe_params(S.count_aperture());
    const char *kind = m.is_cbed() ? "cbed" : m.is_cbei() ? "cbei" : m.is_ed() ? "ed" : m.is_hrtem() ? "hrtem"
                     : m.is_istem() ? "istem" : m.is_ped() ? "ped" : m.is_hctem() ? "hctem"
                     : m.is_ewrs() ? "ewrs" : "ewfs";
    std::ostringstream plan;
    plan << kind << (sizeof(R) == 8 ? " fp64" : " fp32") << " nx=" << m.nx << " ny=" << m.ny
         << " configs_per_batch=" << K << " waves_per_launch=" << W << " slices=" << ns_run;
    if (nq > 1) plan << " waves_per_config=" << nq;
    if (n_ctf > 1) plan << " ctf_defocus_points=" << n_ctf;
    if (!store.resident) plan << " " << store.describe();
    tm.plan = plan.str();

    // column-pass arguments with the wave's tilt-azimuth tables
    auto with_rot = [&](ColArgsT<Z> c, int zi) {
        c.px = px.p + (size_t)zi * m.nx;
        c.py = py.p + (size_t)zi * m.ny;
        c.rot_x = rot_px;
        c.rot_y = rot_py;
        c.members = members.p;
        return c;
    };

    const double w = (double)m.phonon_weight();
    run_control().niter = nconf * nq;
    run_control().iter = 0;
    Event t0, t1;
    record(t0, s);
    DevAtoms<R> atoms;
    S.reserve(atoms);
    PassChoice choice;
    bool tuned = false;
    for (int c0 = 0; c0 < nconf; c0 += K) {
        stop_check();
        const int nb = std::min(K, nconf - c0);   // configurations of this group
        for (int j = 0; j < nb; ++j) {
            S.load(m.fp_iconf_0 + c0 + j, atoms, s);
            store.build(atoms, j, s);
            // the atoms buffer is reused by the next configuration (in the rebuilt
            // mode only after this one's waves: K = 1)
            check(hipStreamSynchronize(s), "sync");
        }
        for (int w0 = 0; w0 < nb * nq; w0 += W) {
            stop_check();
            const int nm = std::min(W, nb * nq - w0);   // waves of this launch
            std::vector<Member> hm(nm);
            for (int i = 0; i < nm; ++i) {
                const int g = w0 + i, j = g / nq;
                const EwWave &e = wv[g % nq];
                hm[i] = Member{j, -1, e.beam0, e.nbeam, e.wq, e.irot};
            }
            members.upload(hm.data(), nm, s);

            RowArgsT<Z> r{};
            r.src = psi.p; r.dst = psi.p; r.member_stride = (long long)nxy; r.nrows = m.nx; r.n_members = nm;
            r.tw = S.tw_y.p; r.members = members.p;
            store.bind(0, r);   // the tuner's launches read slice 0's operands
            if (!tuned) {   // pick the slice-step launch variants on this launch's buffers
                RowArgsT<Z> rt = r;
                rt.slice = 0;
                rt.flags = ROW_LOAD | ROW_IFFT | ROW_MUL_T | ROW_FFT | ROW_STORE;
                ColArgsT<Z> ct = with_rot(S.col_base(psi.p, psi.p, nm), 0);
                ct.flags = COL_LOAD | COL_FFT | COL_MUL_P | (m.bwl && m.multislice() ? COL_BWL : 0) | COL_IFFT |
                           COL_STORE | col_rot;
                check(hipMemsetAsync(psi.p, 0, (size_t)nm * nxy * sizeof(Z), s), "hipMemsetAsync");
                choice = choose_variants(m.nx, m.ny, nm, rt, ct, (const ColArgsT<Z> *)nullptr, s);
                tm.tune_log = choice.log;
                tm.plan += " row_variant=" + std::to_string(choice.row) + " col_variant=" + std::to_string(choice.col);
                tuned = true;
            }
            r.variant = choice.row;
            if (m.iw_type >= 2) {
                ColArgsT<Z> cp = S.col_base(nullptr, psi.p, nm);
                cp.flags = COL_IFFT | COL_STORE | (m.iw_type == 3 ? COL_USER_WAVE : 0);
                cp.fpsi0 = fpsi0.p;
                cp.members = members.p; cp.beams = beams.p; cp.probe = P;
                check(launch_col_pass(m.nx, cp, true, false, s), "probe pass");
            }
            for (int is = 0; is < ns_run; ++is) {
                store.for_slice(atoms, is, r, s);
                r.flags = ((is == 0 && m.iw_type == 1) ? 0 : (ROW_LOAD | ROW_IFFT)) | ROW_MUL_T | ROW_FFT | ROW_STORE;
                check(launch_row_pass(m.ny, r, s), "row pass");
                ColArgsT<Z> c = with_rot(S.col_base(psi.p, psi.p, nm), is);
                c.flags = COL_LOAD | COL_FFT | COL_MUL_P | (m.bwl && m.multislice() ? COL_BWL : 0) | COL_IFFT |
                          COL_STORE | col_rot;
                c.variant = choice.col;
                check(launch_col_pass(m.nx, c, false, false, s), "col pass");
                const int ithk = m.slices[is].ithk;
                if (ithk < 0) continue;
                // O = F_x(L)/nx * bwl * exp(i phi_back)  (get_psi, src/wave_function.cuh:96-105)
                const R zb = m.thk[ithk].z_back_prop;
                const Z *osrc0 = psi.p;
                if (ramp) {
                    // psi exp(-2 pi i g0.r) in real space first (get_psi, src/wave_function.cuh:96-105)
                    RowArgsT<Z> ro = r;
                    ro.src = psi.p; ro.dst = scratch.p;
                    ro.flags = ROW_LOAD | ROW_IFFT | ROW_MUL_RAMP | ROW_FFT | ROW_STORE;
                    ro.ramp_x = rx.p; ro.ramp_y = ry.p;
                    ro.rot_x = n_rot > 1 ? m.nx : 0;
                    ro.rot_y = n_rot > 1 ? m.ny : 0;
                    check(launch_row_pass(m.ny, ro, s), "output ramp pass");
                    osrc0 = scratch.p;
                }
                Z *odst = scratch2.n ? scratch2.p : scratch.p;
                for (int jc = 0; jc < n_ctf; ++jc) {
                    const Z *osrc = osrc0;
                    if (fs || zb != R(0) || m.bwl || m.tilted() || img) {
                        // images: the objective lens acts on O(g) before the inverse
                        // transforms (microscope_effects, src/wave_function.cuh:172-180)
                        ColArgsT<Z> co = with_rot(S.col_base(osrc, odst, nm), m.n_slice() + ithk);
                        co.flags = COL_LOAD | COL_FFT | COL_MUL_P | (m.bwl ? COL_BWL : 0) | (fs ? 0 : COL_IFFT) |
                                   COL_STORE | (img ? COL_MUL_CTF : 0) | col_rot;
                        co.ctf = ctf.p + (size_t)jc * nxy;
                        co.p_scale = (R)m.ny;
                        check(launch_col_pass(m.nx, co, false, false, s), "output col pass");
                        osrc = odst;
                    }
                    if (!fs) {
                        RowArgsT<Z> ro = r;
                        ro.src = osrc; ro.dst = odst;
                        ro.flags = ROW_LOAD | ROW_IFFT | ROW_STORE;
                        check(launch_row_pass(m.ny, ro, s), "output row pass");
                        osrc = odst;
                    }
                    OutArgsT<Z> oa{};
                    oa.src = osrc; oa.member_stride = (long long)nxy; oa.n_members = nm;
                    oa.nx = m.nx; oa.ny = m.ny; oa.ix_0 = m.oa_ix_0; oa.iy_0 = m.oa_iy_0; oa.nx_o = nxo; oa.ny_o = nyo;
                    oa.w = w * cwq[jc];
                    oa.members = members.p;
                    oa.m2 = m2.p + (size_t)ithk * no;
                    oa.psi = pc.p + (size_t)ithk * no * 2;
                    oa.psi_full = pfull.p + (need_full ? (size_t)ithk * nxy * 2 : 0);
                    if (ot == MT_OT_PSI_COH) oa.mode = OUT_SET_PSI;
                    else if (ot == MT_OT_M2PSI_TOT_PSI_COH) oa.mode = OUT_ADD_M2 | OUT_ADD_PSI;
                    else oa.mode = OUT_ADD_M2 | (need_full && !img ? OUT_ADD_PSI_FULL : 0);
                    check(launch_out_accum(oa, s), "output accumulate");
                }
                if (need_full && img) {
                    // psi_coh += w O(g), the back-propagated wave before the objective
                    // lens (add_scale_psi_coh, src/wave_function.cuh:166-172)
                    ColArgsT<Z> co = with_rot(S.col_base(osrc0, odst, nm), m.n_slice() + ithk);
                    co.flags = COL_LOAD | COL_FFT | COL_MUL_P | (m.bwl ? COL_BWL : 0) | COL_STORE | col_rot;
                    co.p_scale = (R)m.ny;
                    check(launch_col_pass(m.nx, co, false, false, s), "coherent output col pass");
                    OutArgsT<Z> oa{};
                    oa.src = odst; oa.member_stride = (long long)nxy; oa.n_members = nm;
                    oa.nx = m.nx; oa.ny = m.ny; oa.ix_0 = m.oa_ix_0; oa.iy_0 = m.oa_iy_0; oa.nx_o = nxo; oa.ny_o = nyo;
                    oa.w = w;
                    oa.members = members.p;
                    oa.psi_full = pfull.p + (size_t)ithk * nxy * 2;
                    oa.mode = OUT_ADD_PSI_FULL;
                    check(launch_out_accum(oa, s), "coherent output accumulate");
                }
            }
            progress_add(nm);
        }
    }
    record(t1, s);
    wait(t1, s);
    tm.ms_total = elapsed(t0, t1);
    tm.probe_slices = (double)nconf * nq * ns_run;

    if (out->m2psi_tot && (ot == MT_OT_M2PSI_TOT || ot == MT_OT_M2PSI_TOT_COH || ot == MT_OT_M2PSI_TOT_PSI_COH))
        check(hipMemcpyAsync(out->m2psi_tot, m2.p, (size_t)n_thk * no * sizeof(double), hipMemcpyDeviceToHost, s), "D2H");
    if (out->psi_coh && (ot == MT_OT_PSI_COH || ot == MT_OT_M2PSI_TOT_PSI_COH))
        check(hipMemcpyAsync(out->psi_coh, pc.p, (size_t)n_thk * no * 2 * sizeof(double), hipMemcpyDeviceToHost, s), "D2H");
    if (need_full && out->m2psi_coh && img) {
        // set_m2psi_coh for HRTEM (src/wave_function.cuh:206-215): the objective
        // lens applied to psi_coh, |IFFT|^2 (summed over the defocus points of a
        // fully integrated lens), cropped and shifted
        DBuf<double> m2c(no);
        std::vector<double> zero(no, 0.0), hp(nxy * 2);
        m2c.upload(zero.data(), no, s);
        std::vector<Z> hz(nxy);
        DBuf<Z> pz(nxy), wz(nxy);
        for (int k = 0; k < n_thk; ++k) {
            check(hipMemcpyAsync(hp.data(), pfull.p + (size_t)k * nxy * 2, nxy * 2 * sizeof(double),
                                 hipMemcpyDeviceToHost, s), "D2H");
            check(hipStreamSynchronize(s), "sync");
            for (size_t i = 0; i < nxy; ++i) hz[i] = cz<Z>(hp[2 * i], hp[2 * i + 1]);
            pz.upload(hz.data(), nxy, s);
            for (int jc = 0; jc < n_ctf; ++jc) {
                ColArgsT<Z> cc = S.col_base(pz.p, wz.p, 1);
                cc.flags = COL_LOAD | COL_MUL_CTF | COL_IFFT | COL_STORE;
                cc.ctf = ctf.p + (size_t)jc * nxy;
                check(launch_col_pass(m.nx, cc, false, false, s), "coherent image col pass");
                RowArgsT<Z> rr{};
                rr.src = wz.p; rr.dst = wz.p; rr.member_stride = (long long)nxy; rr.nrows = m.nx; rr.n_members = 1;
                rr.tw = S.tw_y.p; rr.members = members.p;
                rr.flags = ROW_LOAD | ROW_IFFT | ROW_STORE;
                check(launch_row_pass(m.ny, rr, s), "coherent image row pass");
                OutArgsT<Z> oa{};
                oa.src = wz.p; oa.member_stride = (long long)nxy; oa.n_members = 1;
                oa.nx = m.nx; oa.ny = m.ny; oa.ix_0 = m.oa_ix_0; oa.iy_0 = m.oa_iy_0; oa.nx_o = nxo; oa.ny_o = nyo;
                oa.w = cwq[jc];
                oa.m2 = m2c.p;
                oa.mode = OUT_ADD_M2;
                check(launch_out_accum(oa, s), "coherent image accumulate");
            }
            check(hipMemcpyAsync(out->m2psi_coh + (size_t)k * no, m2c.p, no * sizeof(double), hipMemcpyDeviceToHost, s),
                  "D2H");
            check(hipStreamSynchronize(s), "sync");
            m2c.upload(zero.data(), no, s);
        }
    } else if (need_full && out->m2psi_coh) {
        DBuf<double> tmp(no);
        for (int k = 0; k < n_thk; ++k) {
            check(launch_m2_from_psi(pfull.p + (size_t)k * nxy * 2, m.nx, m.ny, m.oa_ix_0, m.oa_iy_0, nxo, nyo, tmp.p, s),
                  "m2 from psi");
            check(hipMemcpyAsync(out->m2psi_coh + (size_t)k * no, tmp.p, no * sizeof(double), hipMemcpyDeviceToHost, s), "D2H");
            check(hipStreamSynchronize(s), "sync");
        }
    }
    check(hipStreamSynchronize(s), "sync");
}

} // namespace

const char *last_successful_call() { return t_last_ok; }

RunControl &run_control()
{
    static RunControl rc;
    return rc;
}

Timing &last_timing()
{
    static Timing t;
    return t;
}

template <class R>
void run_multislice_group(const std::vector<const ModelT<R> *> &ms, int device,
                          const std::vector<mt_output_multislice *> &outs)
{
    select_device(device);
    for (size_t i = 0; i < ms.size(); ++i) {
        const ModelT<R> &m = *ms[i];
        if (!m.is_stem()) throw Error{MT_ERR_UNSUPPORTED, "group runs take STEM inputs only"};
        m.output_dims(outs[i]);
        m.fill_axes(outs[i]);
    }
    run_stem_group(ms, outs);
}

template <class R>
void run_multislice(const ModelT<R> &m, int device, mt_output_multislice *out)
{
    select_device(device);
    m.output_dims(out);
    m.fill_axes(out);
    if (m.is_stem()) run_stem(m, out);
    else run_ew(m, out);
}

// ---------------------------------------------------------------------------
// per-piece entries

template <class R>
void projected_potential(const ModelT<R> &m, int device, int iconf, int islice, double *V)
{
    select_device(device);
    if (islice < 0 || islice >= m.n_slice()) throw Error{MT_ERR_INVALID_ARGUMENT, "islice out of range"};
    Stream st;
    Setup<R> S(m, st.s);
    DevAtoms<R> atoms;
    S.load(iconf, atoms, st.s);
    DBuf<R> dv(S.nxy);
    S.build_tf(atoms, islice, 1, nullptr, dv.p);
    std::vector<R> h(S.nxy);
    check(hipMemcpyAsync(h.data(), dv.p, S.nxy * sizeof(R), hipMemcpyDeviceToHost, st.s), "D2H");
    check(hipStreamSynchronize(st.s), "sync");
    for (size_t i = 0; i < S.nxy; ++i) V[i] = h[i];
}

double stream_copy_gbs(int device, size_t bytes, int reps)
{
    select_device(device);
    Stream st;
    DBuf<char> a(bytes), b(bytes);
    check(hipMemsetAsync(a.p, 1, bytes, st.s), "hipMemsetAsync");
    check(launch_stream_copy(a.p, b.p, bytes, st.s), "stream copy");
    Event t0, t1;
    check(hipEventRecord(t0.e, st.s), "hipEventRecord");
    for (int i = 0; i < reps; ++i) check(launch_stream_copy(a.p, b.p, bytes, st.s), "stream copy");
    check(hipEventRecord(t1.e, st.s), "hipEventRecord");
    check(hipEventSynchronize(t1.e), "hipEventSynchronize");
    float ms = 0;
    check(hipEventElapsedTime(&ms, t0.e, t1.e), "hipEventElapsedTime");
    return 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
}

// Sum over all slices of the projected potential of configuration iconf
// (natural layout), accumulated in FP64 in slice order; the device time of
// the build and the sum is reported as the run's ms_total.
template <class R>
void projected_potential_sum(const ModelT<R> &m, int device, int iconf, double *V)
{
    select_device(device);
    Timing &tm = last_timing();
    tm = Timing{};
    Stream st;
    Setup<R> S(m, st.s);
    DevAtoms<R> atoms;
    S.load(iconf, atoms, st.s);
    DBuf<R> dv((size_t)m.n_slice() * S.nxy);
    DBuf<double> sum(S.nxy);
    Event t0, t1;
    record(t0, st.s);
    S.build_tf(atoms, 0, m.n_slice(), nullptr, dv.p);
    check(launch_slice_sum(dv.p, m.n_slice(), (long long)S.nxy, sum.p, st.s), "slice sum");
    record(t1, st.s);
    check(hipMemcpyAsync(V, sum.p, S.nxy * sizeof(double), hipMemcpyDeviceToHost, st.s), "D2H");
    check(hipStreamSynchronize(st.s), "sync");
    tm.ms_total = elapsed(t0, t1);
}

template <class R>
static void download_unshifted(Setup<R> &S, const cplx_t<R> *d, double *out)
{
    using Z = cplx_t<R>;
    DBuf<Z> tmp(S.nxy);
    check(launch_shift_complex(d, tmp.p, S.nx, S.ny, S.s), "shift");
    std::vector<Z> h(S.nxy);
    check(hipMemcpyAsync(h.data(), tmp.p, S.nxy * sizeof(Z), hipMemcpyDeviceToHost, S.s), "D2H");
    check(hipStreamSynchronize(S.s), "sync");
    for (size_t i = 0; i < S.nxy; ++i) { out[2 * i] = h[i].x; out[2 * i + 1] = h[i].y; }
}

template <class R>
void transmission_function(const ModelT<R> &m, int device, int iconf, int islice, double *trans)
{
    select_device(device);
    if (islice < 0 || islice >= m.n_slice()) throw Error{MT_ERR_INVALID_ARGUMENT, "islice out of range"};
    Stream st;
    Setup<R> S(m, st.s);
    DevAtoms<R> atoms;
    S.load(iconf, atoms, st.s);
    DBuf<cplx_t<R>> t(S.nxy);
    S.build_tf(atoms, islice, 1, t.p, nullptr);
    download_unshifted(S, t.p, trans);
}

template <class R>
void incident_wave(const ModelT<R> &m, int device, double *psi0)
{
    using Z = cplx_t<R>;
    select_device(device);
    Stream st;
    Setup<R> S(m, st.s);
    DBuf<Z> w(S.nxy);
    if (m.iw_type == 1) {
        std::vector<Z> one(S.nxy, cz<Z>(1.0, 0.0));
        w.upload(one.data(), S.nxy, st.s);
    } else {
        // the stand-alone incident wave evaluates chi at g + g0 (Incident_Wave::operator()
        // with gx_0, gy_0, src/incident_wave.cuh:116-131); inside a simulation the probe is untilted
        ProbeParamsT<R> P = S.probe_params(S.count_aperture(m.gx0, m.gy0));
        P.gxu = (double)m.gx0; P.gyu = (double)m.gy0;
        P.gxu_f = m.gx0; P.gyu_f = m.gy0;
        std::vector<Beam> hb;
        for (size_t i = 0; i < m.iw_x.size(); ++i) hb.push_back(S.beam(m.iw_x[i], m.iw_y[i], P.c10));
        if (hb.empty()) throw Error{MT_ERR_INVALID_ARGUMENT, "convergent wave needs iw_x, iw_y"};
        if (m.iw_type == 3) hb.resize(1);   // a user wave is placed at the first position
        DBuf<Beam> beams(hb.size());
        beams.upload(hb.data(), hb.size(), st.s);
        Member hm{0, -1, 0, (int)hb.size(), 1.0};
        DBuf<Member> mem(1);
        mem.upload(&hm, 1, st.s);
        DBuf<Z> fpsi0;
        if (m.iw_type == 3) S.user_wave_spectrum(fpsi0);
        ColArgsT<Z> c = S.col_base(nullptr, w.p, 1);
        c.flags = COL_IFFT | COL_STORE | (m.iw_type == 3 ? COL_USER_WAVE : 0);
        c.fpsi0 = fpsi0.p;
        c.members = mem.p; c.beams = beams.p; c.probe = P;
        check(launch_col_pass(m.nx, c, true, false, st.s), "probe pass");
        RowArgsT<Z> r{};
        r.src = w.p; r.dst = w.p; r.member_stride = (long long)S.nxy; r.nrows = m.nx; r.n_members = 1;
        r.tw = S.tw_y.p; r.members = mem.p;
        r.flags = ROW_LOAD | ROW_IFFT | ROW_STORE;
        check(launch_row_pass(m.ny, r, st.s), "row pass");
        check(hipStreamSynchronize(st.s), "sync");
    }
    download_unshifted(S, w.p, psi0);
}

// Propagator::operator()(space, ..., output) (src/propagator.cuh:95-104)
template <class R>
void propagate(const ModelT<R> &m, int device, const double *psi_in, double z, int space, double *psi_out)
{
    using Z = cplx_t<R>;
    select_device(device);
    if (space != 1 && space != 2) throw Error{MT_ERR_INVALID_ARGUMENT, "space must be 1 (real) or 2 (reciprocal)"};
    Stream st;
    Setup<R> S(m, st.s);
    std::vector<Z> h(S.nxy);
    for (size_t i = 0; i < S.nxy; ++i) h[i] = cz<Z>((R)psi_in[2 * i], (R)psi_in[2 * i + 1]);
    DBuf<Z> a(S.nxy), w(S.nxy);
    a.upload(h.data(), S.nxy, st.s);
    check(launch_shift_complex(a.p, w.p, S.nx, S.ny, st.s), "shift");
    DBuf<Z> px, py;
    std::vector<double> zl{z};
    S.prop_tables(zl, px, py);
    const bool zero = is_zero((R)z);
    Member hm{0, -1, 0, 0, 1.0};
    DBuf<Member> mem(1);
    mem.upload(&hm, 1, st.s);
    RowArgsT<Z> r{};
    r.src = w.p; r.dst = w.p; r.member_stride = (long long)S.nxy; r.nrows = m.nx; r.n_members = 1;
    r.tw = S.tw_y.p; r.members = mem.p;
    bool transformed = false;
    if (!zero || m.bwl || space == 2) {
        r.flags = ROW_LOAD | ROW_FFT | ROW_STORE;
        check(launch_row_pass(m.ny, r, st.s), "row pass");
        ColArgsT<Z> c = S.col_base(w.p, w.p, 1);
        if (!zero) c.flags = COL_LOAD | COL_FFT | COL_MUL_P | (m.bwl ? COL_BWL : 0) | COL_STORE;
        else c.flags = COL_LOAD | COL_FFT | (m.bwl ? COL_BWL : 0) | COL_STORE;
        if (zero && !m.bwl) {
            c.flags = COL_LOAD | COL_FFT | COL_MUL_P | COL_STORE;   // scale 1/(nx ny): z = 0 tables
        }
        c.px = px.p; c.py = py.p;
        if (space == 1) c.flags |= COL_IFFT;
        check(launch_col_pass(m.nx, c, false, false, st.s), "col pass");
        transformed = true;
    }
    if (transformed && space == 1) {
        r.flags = ROW_LOAD | ROW_IFFT | ROW_STORE;
        check(launch_row_pass(m.ny, r, st.s), "row pass");
    }
    check(hipStreamSynchronize(st.s), "sync");
    download_unshifted(S, w.p, psi_out);
}

// ---------------------------------------------------------------------------
// Propagation without device allocation.  All working memory of the call --
// the input/output staging area, the wave, the twiddle and propagator
// tables -- lives in one static device array per GPU, `scratch`, whose
// device address is looked up once per device (hipGetSymbolAddress) on the
// first call there and kept; host code uses only that address.  The array
// keeps the last call's contents; every region is fully written before it is
// read, so nothing needs clearing.  Calls on one GPU share the array: the C
// ABI serialises them behind its lock, and calls that overlap on one GPU
// from outside that lock would race.
constexpr size_t kNoAllocMaxNxy = size_t(1) << 20;   // nx * ny at most 1024^2 (FP32 and FP64)
constexpr size_t kNoAllocScratchBytes = kNoAllocMaxNxy * 16 * 2 + (size_t(1) << 18);
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

unsigned char *scratch_on(int device)
{
    static std::mutex mu;
    static std::map<int, unsigned char *> ptr;
    std::lock_guard<std::mutex> lk(mu);
    auto it = ptr.find(device);
    if (it != ptr.end()) return it->second;
    void *p = nullptr;
    check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    ptr[device] = (unsigned char *)p;
    return (unsigned char *)p;
}

} // namespace

size_t propagate_noalloc_max_nxy() { return kNoAllocMaxNxy; }

size_t kept_device_bytes(int device) { return kept_bytes(device); }

void release_kept_blocks(int device)
{
    check(hipSetDevice(device), "hipSetDevice");
    check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    trim_blocks(device);
}

template <class R>
void propagate_noalloc(const ModelT<R> &m, int device, const double *psi_in, double z, int space, double *psi_out)
{
    using Z = cplx_t<R>;
    select_device(device);
    if (space != 1 && space != 2) throw Error{MT_ERR_INVALID_ARGUMENT, "space must be 1 (real) or 2 (reciprocal)"};
    const int nx = m.nx, ny = m.ny;
    const size_t nxy = (size_t)nx * ny;
    if (nxy > kNoAllocMaxNxy)
        throw Error{MT_ERR_INVALID_ARGUMENT, "nx * ny = " + std::to_string(nxy) + " exceeds the " +
                                                 std::to_string(kNoAllocMaxNxy) +
                                                 " points of the allocation-free propagation"};
    unsigned char *base = scratch_on(device);
    auto carve = [&](size_t bytes) {
        unsigned char *q = base;
        base += (bytes + 255) / 256 * 256;
        return q;
    };
    double2 *stage = (double2 *)carve(nxy * sizeof(double2));
    Z *w = (Z *)carve(nxy * sizeof(Z));
    Z *tw_x = (Z *)carve(nx * sizeof(Z)), *tw_y = (Z *)carve(ny * sizeof(Z));
    Z *px = (Z *)carve(nx * sizeof(Z)), *py = (Z *)carve(ny * sizeof(Z));
    double *zd = (double *)carve(sizeof(double));
    Member *mem = (Member *)carve(sizeof(Member));
    if ((size_t)(base - scratch_on(device)) > kNoAllocScratchBytes)
        throw Error{MT_ERR_INVALID_ARGUMENT, "grid exceeds the allocation-free scratch buffer"};

    const hipStream_t s = nullptr;   // the null stream: creating a stream may allocate
    const auto wx = twiddles<Z>(nx), wy = twiddles<Z>(ny);
    const Member hm{0, -1, 0, 0, 1.0};
    check(hipMemcpyAsync(stage, psi_in, nxy * sizeof(double2), hipMemcpyHostToDevice, s), "H2D psi");
    check(hipMemcpyAsync(tw_x, wx.data(), nx * sizeof(Z), hipMemcpyHostToDevice, s), "H2D twiddles");
    check(hipMemcpyAsync(tw_y, wy.data(), ny * sizeof(Z), hipMemcpyHostToDevice, s), "H2D twiddles");
    check(hipMemcpyAsync(zd, &z, sizeof(double), hipMemcpyHostToDevice, s), "H2D z");
    check(hipMemcpyAsync(mem, &hm, sizeof(Member), hipMemcpyHostToDevice, s), "H2D member");
    check(launch_load_shifted(stage, w, nx, ny, s), "load shifted");
    check(launch_prop_tables(nx, ny, zd, 1, m.lambda_d, m.mul_sign, m.lx_d, m.ly_d, (double)m.gx0_rot(0),
                             (double)m.gy0_rot(0), px, py, s),
          "prop tables");

    // the passes of propagate(), on the scratch buffers
    const bool zero = is_zero((R)z);
    RowArgsT<Z> r{};
    r.src = w; r.dst = w; r.member_stride = (long long)nxy; r.nrows = nx; r.n_members = 1;
    r.tw = tw_y; r.members = mem;
    bool transformed = false;
    if (!zero || m.bwl || space == 2) {
        r.flags = ROW_LOAD | ROW_FFT | ROW_STORE;
        check(launch_row_pass(ny, r, s), "row pass");
        ColArgsT<Z> c{};
        c.src = w; c.dst = w; c.member_stride = (long long)nxy;
        c.ncols = ny; c.n_members = 1;
        c.tw = tw_x;
        c.dgx_f = m.dgx; c.dgy_f = m.dgy;
        c.nx = nx; c.ny = ny;
        c.bwl_alpha = m.bwl_alpha; c.gl2_max = m.gl2_max;
        c.inv_nxy = R(1) / (R)nxy;
        c.p_scale = R(1);
        if (!zero) c.flags = COL_LOAD | COL_FFT | COL_MUL_P | (m.bwl ? COL_BWL : 0) | COL_STORE;
        else c.flags = COL_LOAD | COL_FFT | (m.bwl ? COL_BWL : 0) | COL_STORE;
        if (zero && !m.bwl) c.flags = COL_LOAD | COL_FFT | COL_MUL_P | COL_STORE;
        c.px = px; c.py = py;
        if (space == 1) c.flags |= COL_IFFT;
        check(launch_col_pass(nx, c, false, false, s), "col pass");
        transformed = true;
    }
    if (transformed && space == 1) {
        r.flags = ROW_LOAD | ROW_IFFT | ROW_STORE;
        check(launch_row_pass(ny, r, s), "row pass");
    }
    check(launch_store_unshifted(w, stage, nx, ny, s), "store unshifted");
    check(hipMemcpyAsync(psi_out, stage, nxy * sizeof(double2), hipMemcpyDeviceToHost, s), "D2H psi");
    check(hipStreamSynchronize(s), "sync");
}

#define MTG_ENGINE_INSTANTIATE(R)                                                                                      \
    template void run_multislice<R>(const ModelT<R> &, int, mt_output_multislice *);                                   \
    template void run_multislice_group<R>(const std::vector<const ModelT<R> *> &, int,                                 \
                                          const std::vector<mt_output_multislice *> &);                               \
    template void projected_potential<R>(const ModelT<R> &, int, int, int, double *);                                  \
    template void projected_potential_sum<R>(const ModelT<R> &, int, int, double *);                                   \
    template void transmission_function<R>(const ModelT<R> &, int, int, int, double *);                                \
    template void incident_wave<R>(const ModelT<R> &, int, double *);                                                  \
    template void propagate<R>(const ModelT<R> &, int, const double *, double, int, double *);                       \
    template void propagate_noalloc<R>(const ModelT<R> &, int, const double *, double, int, double *);

MTG_ENGINE_INSTANTIATE(float)
MTG_ENGINE_INSTANTIATE(double)
#undef MTG_ENGINE_INSTANTIATE

} // namespace mtg
