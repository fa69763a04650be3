// C ABI of libmafm: context, device tables, launches.  Host code only; the
// kernels live in scan_plain.hip, scan_fast.hpp (instantiated by scan_fast_mid.hip and
// scan_fast_gen.hip), fft.hip, adjoint.hip, deltaf.hip and util.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <array>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mafm.h"
#include "ctx_ref.hpp"
#include "device_common.hpp"
#include "flex.hpp"

namespace mafm {
hipError_t launch_scan_plain(const ScanArgs& a, hipStream_t st);
hipError_t launch_eval_plain(const ScanArgs& a, int mol, long n, const double* pos, double* out, hipStream_t st);
hipError_t launch_grid_interp(const ScanArgs& a, int grid, long n, const double* pos, double* out, hipStream_t st);
size_t valu_calibration_doubles();
hipError_t launch_valu_calibration(int kind, int n, double* out, int64_t* wave_insts, hipStream_t st);
hipError_t launch_grid_build(const ScanArgs& a, int mol, const int* n, const double* s, const double* o,
                             double4* out, hipStream_t st);
hipError_t launch_scan_fast(const ScanArgs& a, bool f32, int comp, int G, int n_blocks, int W, int lds_sites,
                            bool scalar_sites, bool mid_family, hipStream_t st);
hipError_t launch_to_f32(const double4* in, float4* out, long n, hipStream_t st);
hipError_t launch_record_totals(const DevRecord* rec, long n, int64_t* out, hipStream_t st);
hipError_t noalloc_scratch_address(void** out);
bool deltaf_weights(const mafm_deltaf_desc& d, int nz, double zhigh, double dz_input, mafm_deltaf_weights_out& o,
                    std::vector<double>* z_low, char* err, size_t err_len);
hipError_t launch_delta_f(const DevRecord* rec, const mafm_deltaf_weights_out& w, int nx, int ny, int nz, bool text,
                          double* out, hipStream_t st);
bool fft_length_supported(int n);
void fft_twiddles(int n, std::vector<double>& out);
hipError_t fft3d_device(double2* data, const int* n, const double2* const* tw, bool inv, hipStream_t st);
hipError_t launch_real_to_complex(const double* in, double2* out, long n, hipStream_t st);
hipError_t launch_zero2(double2* out, long n, hipStream_t st);
hipError_t launch_gauss_tip(double2* rho, const int* n, const int* cut, bool ortho, const double* b, double qn,
                            double w2, hipStream_t st);
hipError_t launch_mul_vol(double2* pot, const double2* rho, double vol, long n, hipStream_t st);
hipError_t launch_spectral_grad(const double2* ek, double2* out, const int* n, int c, const double* kb,
                                hipStream_t st);
hipError_t launch_extract(const double2* in, double* fe, int slot, double inv_n, long n, int* flag, hipStream_t st);
hipError_t launch_scale2(double2* x, double s, long n, hipStream_t st);
hipError_t launch_adjoint(const ScanArgs& a, int mol, int n_sites, long n, const double* tip, const double* dummy,
                          const double* g, double* v, double* out6, int* singular, bool f32, hipStream_t st);
}  // namespace mafm

using namespace mafm;

namespace {

struct HostGrid {
    bool live = false;
    int n[3] = {0, 0, 0};
    int periodic = 0;
    int orthogonal = 1;
    double basis[9] = {0}, inv[9] = {0}, offset[3] = {0};
    double4* fe64 = nullptr;
    float4* fe32 = nullptr;
};

void inv3(const double* m, double* o) {
    double det = m[0] * (m[4] * m[8] - m[5] * m[7]);
    det += m[1] * (m[5] * m[6] - m[3] * m[8]);
    det += m[2] * (m[3] * m[7] - m[4] * m[6]);
    double inv_det = 1 / det;
    double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6], h = m[7], i = m[8];
    o[0] = inv_det * (e * i - f * h);
    o[1] = -inv_det * (b * i - c * h);
    o[2] = inv_det * (b * f - c * e);
    o[3] = -inv_det * (d * i - f * g);
    o[4] = inv_det * (a * i - c * g);
    o[5] = -inv_det * (a * f - c * d);
    o[6] = inv_det * (d * h - e * g);
    o[7] = -inv_det * (a * h - b * g);
    o[8] = inv_det * (a * e - b * d);
}

// Matrix3::isDiagonal (matrices.hpp:72-86)
bool is_diag(const double* m) {
    double diag = 0, nondiag = 0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) (i == j ? diag : nondiag) += std::fabs(m[3 * i + j]);
    return nondiag < 1.0e-10 * diag;
}

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }   // also on early error returns
    hipError_t reserve(size_t n) {
        if (n <= cap && p) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T));
        if (e == hipSuccess) cap = std::max<size_t>(n, 1);
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Pinned, coherent host memory mapped into the device's address space: the
// kernels store into it through `dp` over PCIe while the host reads it
// through `p`.  Grows only.
template <typename T>
struct HostBuf {
    T* p = nullptr;
    T* dp = nullptr;
    size_t cap = 0;
    HostBuf() = default;
    HostBuf(const HostBuf&) = delete;
    HostBuf& operator=(const HostBuf&) = delete;
    ~HostBuf() { release(); }
    hipError_t reserve(size_t n) {
        if (n <= cap && p) return hipSuccess;
        release();
        void* h = nullptr;
        hipError_t e = hipHostMalloc(&h, std::max<size_t>(n, 1) * sizeof(T), hipHostMallocMapped | hipHostMallocCoherent);
        if (e != hipSuccess) return e;
        void* d = nullptr;
        e = hipHostGetDevicePointer(&d, h, 0);
        if (e != hipSuccess) {
            (void)hipHostFree(h);
            return e;
        }
        p = static_cast<T*>(h);
        dp = static_cast<T*>(d);
        cap = std::max<size_t>(n, 1);
        return hipSuccess;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = dp = nullptr;
        cap = 0;
    }
};

// What finishing a launched scan needs (mafm_scan_host_wait).
struct PendingScan {
    mafm_scan_desc d{};
    bool fast = false;
    int G = 1, wps = 0, n_blocks = 0, W = 1;
    size_t n_rec = 0;
};

}  // namespace

struct mafm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    int cus = 0;
    size_t lds_per_cu = 0;   // LDS bytes per compute unit
    size_t lds_per_block = 0;   // LDS bytes one workgroup may allocate
    std::vector<HostGrid> grids;
    DevBuf<DevGrid> d_grids;
    std::vector<DevMol> mols;
    DevBuf<DevMol> d_mols;
    DevBuf<Site64> d_ref;
    DevBuf<SiteF64> d_seg64;
    DevBuf<float4> d_seg32a, d_seg32b;
    DevBuf<DevRecord> d_rec;
    DevBuf<int32_t> d_next;
    DevBuf<float> d_cum;
    DevBuf<int32_t> d_order;
    DevBuf<int64_t> d_counters;
    DevBuf<double> d_tmp;
    DevBuf<float4> d_margins;      // per-position decision margins (MAFM_SCAN_MARGINS)
    // streamed scans (mafm_scan_host): records and per-column flags in pinned host memory
    HostBuf<DevRecord> h_rec;
    HostBuf<int32_t> h_done;
    HostBuf<int64_t> h_counters;   // the streamed launch's counters, copied right behind it
    bool host_pending = false;     // launched, not yet ended by mafm_scan_host_wait
    PendingScan pending;
    // flexible scan: term list, index tables, initial positions, kept states, per-block work
    DevBuf<mafm_flex_term> d_fx_terms;
    DevBuf<int32_t> d_fx_ints;
    DevBuf<double> d_fx_pos, d_fx_state, d_fx_work;
    DevBuf<double> d_sys;          // System operators: positions, velocities, forces, energies, steps
    bool last_flex = false;        // the records come from mafm_scan_flexible
    int fx_state_atoms = 0, fx_state_heights = 0;   // kept state of the last flexible scan (0: none)
    bool have_margins = false;
    uint64_t scan_generation = 0;  // bumped by every scan (the records are replaced)
    void* scratch = nullptr;       // device address of the fixed workspace (mafm_scan_noalloc)
    std::unique_ptr<ScanScratch> scratch_img{new ScanScratch()};   // its host image
    std::vector<int64_t> mol_cost;
    size_t n_records = 0;
    bool have_systems = false;
    mafm_scan_desc last_scan{};   // geometry of the records in d_rec
    bool have_scan = false;
    std::map<int, DevBuf<double2>> twiddles;   // FFT twiddle tables by length
    DevBuf<int> d_flag;
    int sims = 0;                  // open simulations on this context (mafm_sim_open)
    bool destroy_pending = false;  // mafm_destroy called while sims > 0
};

static mafm_status fail(mafm_ctx* c, mafm_status s, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return s;
}

#define HIPCHK(ctx, expr)                                                                          \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(ctx, MAFM_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));        \
    } while (0)

static mafm_status sync_grids(mafm_ctx* c) {
    std::vector<DevGrid> dg(std::max<size_t>(c->grids.size(), 1));
    for (size_t k = 0; k < c->grids.size(); ++k) {
        const HostGrid& h = c->grids[k];
        DevGrid& g = dg[k];
        memset(&g, 0, sizeof g);
        for (int d = 0; d < 3; ++d) g.n[d] = h.n[d];
        g.periodic = h.periodic;
        g.orthogonal = h.orthogonal;
        memcpy(g.basis, h.basis, sizeof g.basis);
        memcpy(g.inv, h.inv, sizeof g.inv);
        memcpy(g.offset, h.offset, sizeof g.offset);
        g.fe64 = h.fe64;
        g.fe32 = h.fe32;
    }
    HIPCHK(c, c->d_grids.reserve(dg.size()));
    HIPCHK(c, hipMemcpyAsync(c->d_grids.p, dg.data(), dg.size() * sizeof(DevGrid), hipMemcpyHostToDevice,
                             c->stream));
    return MAFM_OK;
}

static mafm_status new_grid(mafm_ctx* c, const mafm_grid_desc* g, HostGrid& h) {
    if (!g) return fail(c, MAFM_ERR_INVALID, "grid descriptor is NULL");
    for (int d = 0; d < 3; ++d) {
        if (g->n[d] <= 0) return fail(c, MAFM_ERR_INVALID, "grid dimension %d is %d", d, g->n[d]);
        h.n[d] = g->n[d];
        h.offset[d] = g->offset[d];
    }
    h.periodic = g->periodic ? 1 : 0;
    memcpy(h.basis, g->basis, sizeof h.basis);
    h.orthogonal = is_diag(h.basis);
    inv3(h.basis, h.inv);
    if (h.orthogonal)   // exact reciprocals of the spacings (the fast lookups multiply by them)
        for (int d = 0; d < 3; ++d) h.inv[4 * d] = 1.0 / h.basis[4 * d];
    size_t np = (size_t)h.n[0] * h.n[1] * h.n[2];
    HIPCHK(c, hipMalloc(&h.fe64, np * sizeof(double4)));
    HIPCHK(c, hipMalloc(&h.fe32, np * sizeof(float4)));
    h.live = true;
    return MAFM_OK;
}

// A new grid's device buffers, freed on every early return unless the grid
// was stored in the context.
struct GridGuard {
    HostGrid& h;
    bool kept = false;
    explicit GridGuard(HostGrid& g) : h(g) {}
    GridGuard(const GridGuard&) = delete;
    GridGuard& operator=(const GridGuard&) = delete;
    ~GridGuard() {
        if (kept) return;
        if (h.fe64) (void)hipFree(h.fe64);
        if (h.fe32) (void)hipFree(h.fe32);
        h.fe64 = nullptr;
        h.fe32 = nullptr;
    }
};

static int store_grid(mafm_ctx* c, const HostGrid& h) {
    for (size_t k = 0; k < c->grids.size(); ++k)
        if (!c->grids[k].live) {
            c->grids[k] = h;
            return (int)k;
        }
    c->grids.push_back(h);
    return (int)c->grids.size() - 1;
}

static ScanArgs base_args(mafm_ctx* c) {
    ScanArgs a;
    memset(&a, 0, sizeof a);
    a.mols = c->d_mols.p;
    a.sites_ref = c->d_ref.p;
    a.segf64 = c->d_seg64.p;
    a.seg32a = c->d_seg32a.p;
    a.seg32b = c->d_seg32b.p;
    a.grids = c->d_grids.p;
    a.counters = c->d_counters.p;
    a.n_mols = (int)c->mols.size();
    return a;
}

// Lower mafm_system_desc tables to the device layout (device_common.hpp).
static mafm_status build_tables(mafm_ctx* c, int n, const mafm_system_desc* sys, std::vector<DevMol>& mols,
                                std::vector<Site64>& ref, std::vector<SiteF64>& seg64,
                                std::vector<Site32>& seg32, std::vector<int64_t>& cost) {
    const double LOG2E = 1.4426950408889634;
    for (int m = 0; m < n; ++m) {
        const mafm_system_desc& s = sys[m];
        if (s.n_sites < 0 || (s.n_sites > 0 && !s.sites))
            return fail(c, MAFM_ERR_INVALID, "system %d: bad site table", m);
        for (int g : {s.force_grid, s.potential_grid})
            if (g >= 0 && (g >= (int)c->grids.size() || !c->grids[g].live))
                return fail(c, MAFM_ERR_INVALID, "system %d: grid handle %d is not live", m, g);
        DevMol d;
        memset(&d, 0, sizeof d);
        d.ref_off = (int)ref.size();
        d.n_sites = s.n_sites;
        d.seg_off = (int)seg64.size();
        d.force_grid = s.force_grid;
        d.potential_grid = s.potential_grid;
        d.tip_pbc = s.tip_pbc;
        d.td_vdw = s.tip_dummy.vdw;
        d.td_coul = s.tip_dummy.has_coulomb;
        d.td_p0 = s.tip_dummy.p0;
        d.td_p1 = s.tip_dummy.p1;
        d.td_p2 = s.tip_dummy.p2;
        d.td_qq = s.tip_dummy.qq;
        d.td_k = s.tip_dummy.k;
        d.td_r0 = s.tip_dummy.r0;
        d.tip_dummy_d = s.tip_dummy_d;
        memcpy(d.cell, s.cell, sizeof d.cell);
        memcpy(d.offset, s.offset, sizeof d.offset);
        if (s.tip_pbc) inv3(d.cell, d.cell_inv);
        if (s.n_images < 0 || s.n_images > 27 || (s.n_images > 0 && !s.image_shifts))
            return fail(c, MAFM_ERR_INVALID, "system %d: n_images must be 0..27 with shifts", m);
        d.n_images = s.n_images > 0 ? s.n_images : 1;
        d.planar = 1;
        for (int k = 0; k < d.n_images; ++k) {
            for (int q = 0; q < 3; ++q) {
                d.shift[k][q] = s.n_images > 0 ? s.image_shifts[3 * k + q] : 0.0;
                d.shift32[k][q] = (float)d.shift[k][q];
            }
            d.shift32[k][3] = 0.f;
            if (d.shift32[k][2] != 0.f) d.planar = 0;
        }
        // FP32 coordinate origin: centre of the sites' bounding box
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int k = 0; k < s.n_sites; ++k) {
            const double p[3] = {s.sites[k].x, s.sites[k].y, s.sites[k].z};
            for (int q = 0; q < 3; ++q) {
                lo[q] = std::min(lo[q], p[q]);
                hi[q] = std::max(hi[q], p[q]);
            }
        }
        for (int q = 0; q < 3; ++q) d.origin[q] = s.n_sites ? 0.5 * (lo[q] + hi[q]) : 0.0;
        // FP32 coordinates stay relative to the base image; image shifts are
        // applied to the tip in FP64 before the conversion
        for (int k = 0; k < s.n_sites; ++k) {
            const mafm_site& t = s.sites[k];
            if (t.vdw != MAFM_VDW_LJ && t.vdw != MAFM_VDW_MORSE)
                return fail(c, MAFM_ERR_INVALID, "system %d site %d: unknown vdw kind %d", m, k, t.vdw);
            Site64 r = {t.x, t.y, t.z, t.p0, t.p1, t.p2, t.has_coulomb ? t.qq : 0.0, t.vdw, t.has_coulomb ? 1 : 0};
            ref.push_back(r);
        }
        for (int segk = 0; segk < 4; ++segk) {
            const int want_vdw = segk < 2 ? MAFM_VDW_LJ : MAFM_VDW_MORSE;
            const int want_coul = (segk % 2 == 0) ? 1 : 0;
            int cnt = 0;
            for (int k = 0; k < s.n_sites; ++k) {
                const mafm_site& t = s.sites[k];
                const int coul = (t.has_coulomb && t.qq != 0.0) ? 1 : 0;
                if (t.vdw != want_vdw || coul != want_coul) continue;
                // pre-scaled constants (scan_fast.hpp, pair32 / pair64)
                const double qq = coul ? t.qq : 0.0;
                SiteF64 g;
                g.x = t.x; g.y = t.y; g.z = t.z; g.qq = qq;
                Site32 f;
                f.a = make_float4((float)(t.x - d.origin[0]), (float)(t.y - d.origin[1]),
                                  (float)(t.z - d.origin[2]), 0.f);
                if (t.vdw == MAFM_VDW_LJ) {
                    g.p0 = 6 * t.p0; g.p1 = 12 * t.p1; g.p2 = 0; g.p3 = 0;
                    f.a.w = (float)(6 * t.p0);
                    f.b = make_float4((float)(12 * t.p1), (float)qq, 0.f, 0.f);
                } else {
                    g.p0 = 12 * t.p0; g.p1 = t.p1; g.p2 = t.p2; g.p3 = 2 * t.p0 * t.p1;
                    f.a.w = (float)(12 * t.p0);
                    f.b = make_float4((float)(t.p1 * LOG2E), (float)qq, (float)(t.p1 * LOG2E * t.p2), (float)g.p3);
                }
                seg64.push_back(g);
                seg32.push_back(f);
                ++cnt;
            }
            d.n_seg[segk] = cnt;
        }
        mols.push_back(d);
        cost.push_back(s.force_grid >= 0 ? 8 : std::max(1, s.n_sites * d.n_images));
    }
    return MAFM_OK;
}

// Launch shape of a FAST scan: cols[m] pixel columns of molecule m; lanes /
// waves_per_simd / waves_per_group 0 = automatic.  Uploads the work-stealing
// state into a.
struct FastPlan {
    int G = 1, wps = 0, W = 1, n_blocks = 0, lds_sites = 0;
    bool scalar_sites = false, mid_family = false;
};

// LDS bytes the fast scan may stage one molecule's sites in
static size_t scan_lds_limit(const mafm_ctx* c) {
    return c->lds_per_block > 1024 ? c->lds_per_block - 1024 : 0;
}

static mafm_status plan_fast(mafm_ctx* c, const mafm_scan_desc* d, bool f32, const std::vector<long>& cols, int lanes,
                             int waves_per_simd, int waves_per_group, ScanArgs& a, FastPlan& p,
                             ScanScratch* host_ws = nullptr, ScanScratch* dev_ws = nullptr) {
    const int n_mols = (int)c->mols.size();
    long total_cols = 0;
    for (long v : cols) total_cols += v;
    // kernel family: FIRE + midpoint on non-periodic systems runs the
    // specialised kernels (built for three waves per SIMD); periodic
    // systems, whose image loop needs more registers, and every other
    // minimiser run the generic ones (two waves per SIMD)
    bool periodic = false;
    for (const DevMol& m : c->mols) periodic |= m.n_images > 1;
    p.mid_family = d->minimiser == MAFM_MIN_FIRE && d->integrator == MAFM_INT_MIDPOINT && !periodic;
    p.wps = waves_per_simd > 0 ? waves_per_simd : (p.mid_family ? 3 : 2);
    const long lanes_total = (long)c->cus * 4 * p.wps * 64;
    int64_t max_sites = 1;
    for (int m = 0; m < n_mols; ++m)
        if (cols[m] > 0) max_sites = std::max<int64_t>(max_sites, c->mols[m].force_grid >= 0 ? 1 : c->mols[m].n_sites);
    if (lanes > 0) {
        p.G = lanes;
    } else {
        p.G = 1;
        // split pixels over lanes until the columns give every resident lane
        // two pixels' share of work, and a lane keeps >= 8 sites
        while (p.G < 64 && (double)total_cols * p.G < 2.0 * lanes_total && p.G * 8 <= max_sites) p.G <<= 1;
    }
    // a workgroup stages its molecule's sites in LDS (up to what one
    // workgroup may allocate, less 1 KiB for the kernel's static LDS: 159 KiB
    // on gfx950); larger molecules are streamed with scalar loads, one lane
    // per pixel
    const size_t site_bytes = f32 ? sizeof(Site32) : sizeof(SiteF64);
    const size_t lds_limit = scan_lds_limit(c);
    p.scalar_sites = (size_t)max_sites * site_bytes > lds_limit;
    if (p.scalar_sites && p.G > 1) {
        if (lanes > 0)
            return fail(c, MAFM_ERR_INVALID, "lanes_per_pixel > 1 needs the sites of one molecule to fit %zu bytes of LDS",
                        lds_limit);
        p.G = 1;
    }
    p.lds_sites = (int)max_sites;
    // waves per workgroup share one LDS copy of the sites: grow W until the
    // LDS a CU holds admits wps waves per SIMD
    p.W = 1;
    if (waves_per_group > 0) {
        p.W = waves_per_group;
    } else if (!p.scalar_sites) {
        const size_t lds_bytes = std::max<size_t>((size_t)max_sites * site_bytes, 1);
        const long blocks_per_cu = (long)(c->lds_per_cu / lds_bytes);
        while (p.W < 8 && blocks_per_cu * p.W < 4L * p.wps) ++p.W;
        // FIRE + midpoint kernels over site tables: at least four waves per
        // workgroup (one per SIMD), which measured faster than two on the
        // batched workload (fast64 531-537 -> 516-518 ms, fast32 -0.8 %); any
        // W that holds the LDS constraint above also holds it at 4.  Force-
        // grid scans stage nothing and keep the smallest W (four measured
        // 6 % slower there).
        bool stages_sites = false;
        for (int m = 0; m < n_mols; ++m) stages_sites |= cols[m] > 0 && c->mols[m].force_grid < 0;
        if (p.mid_family && stages_sites) p.W = std::max(p.W, 4);
    }
    const long groups_per_wave = 64 / p.G;
    const long need_waves = std::max<long>((total_cols + groups_per_wave - 1) / groups_per_wave, 1);
    int n_waves = (int)std::min<long>((long)c->cus * 4 * p.wps, need_waves);
    n_waves = std::max(n_waves, 1);
    p.n_blocks = (n_waves + p.W - 1) / p.W;
    // Molecules by rank: decreasing work per column (sites x images), so the
    // longest columns start first.  Initial workgroup shares: when the batch
    // needs more workgroups than the grid holds, fill ranks in order with
    // enough workgroups to take every column of the molecule at once
    // (longest-processing-time first); otherwise share in proportion to each
    // molecule's total work.
    std::vector<int32_t> zero(n_mols, 0), order(n_mols);
    for (int m = 0; m < n_mols; ++m) order[m] = m;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        const bool wx = cols[x] > 0, wy = cols[y] > 0;
        if (wx != wy) return wx;
        return c->mol_cost[x] > c->mol_cost[y];
    });
    const long cols_per_block = groups_per_wave * (long)p.W;
    long need_blocks_total = 0;
    for (int m = 0; m < n_mols; ++m) need_blocks_total += (cols[m] + cols_per_block - 1) / cols_per_block;
    std::vector<double> share(n_mols, 0.0);
    if (need_blocks_total > p.n_blocks) {
        long left = p.n_blocks;
        for (int k = 0; k < n_mols && left > 0; ++k) {
            const long b = std::min((cols[order[k]] + cols_per_block - 1) / cols_per_block, left);
            share[k] = (double)b;
            left -= b;
        }
    } else {
        for (int k = 0; k < n_mols; ++k) share[k] = (double)c->mol_cost[order[k]] * (double)cols[order[k]];
    }
    double tot = 0;
    for (double v : share) tot += v;
    if (!(tot > 0)) {
        tot = 1;
        share[0] = 1;
    }
    std::vector<float> cum(n_mols + 1, 0.f);
    double run = 0;
    for (int k = 0; k < n_mols; ++k) {
        run += share[k];
        cum[k + 1] = (float)(run / tot);
    }
    cum[n_mols] = 1.0f;
    if (dev_ws) {
        // fixed workspace: the caller copies the host image over in one go
        std::copy(zero.begin(), zero.end(), host_ws->next_col);
        std::copy(cum.begin(), cum.end(), host_ws->cum);
        std::copy(order.begin(), order.end(), host_ws->order);
        a.next_col = dev_ws->next_col;
        a.mol_cum = dev_ws->cum;
        a.mol_order = dev_ws->order;
        return MAFM_OK;
    }
    HIPCHK(c, c->d_next.reserve(n_mols));
    HIPCHK(c, c->d_cum.reserve(n_mols + 1));
    HIPCHK(c, c->d_order.reserve(n_mols));
    HIPCHK(c, hipMemcpyAsync(c->d_next.p, zero.data(), n_mols * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_cum.p, cum.data(), (n_mols + 1) * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_order.p, order.data(), n_mols * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    // the host vectors die with this call: the copies must be done first
    HIPCHK(c, hipStreamSynchronize(c->stream));
    a.next_col = c->d_next.p;
    a.mol_cum = c->d_cum.p;
    a.mol_order = c->d_order.p;
    return MAFM_OK;
}

static void destroy_now(mafm_ctx* c);

extern "C" {

mafm_status mafm_create(int device, mafm_ctx** out) {
    if (!out) return MAFM_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return MAFM_ERR_NO_DEVICE;
    if (device < 0 || device >= count) return MAFM_ERR_INVALID;
    mafm_ctx* c = new mafm_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        c->d_counters.reserve(8) != hipSuccess) {
        delete c;
        return MAFM_ERR_HIP;
    }
    // the allocation-free scan's workspace: its address on this device, taken
    // once with the device current
    if (noalloc_scratch_address(&c->scratch) != hipSuccess || !c->scratch) {
        destroy_now(c);
        return MAFM_ERR_HIP;
    }
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) == hipSuccess) {
        c->cus = p.multiProcessorCount;
        c->lds_per_cu = p.maxSharedMemoryPerMultiProcessor;
        c->lds_per_block = p.sharedMemPerBlock;
    }
    if (c->cus <= 0) c->cus = 256;
    if (c->lds_per_cu < 64 * 1024) c->lds_per_cu = 64 * 1024;
    if (c->lds_per_block < 64 * 1024) c->lds_per_block = 64 * 1024;
    *out = c;
    return MAFM_OK;
}

static void destroy_now(mafm_ctx* c) {
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto& g : c->grids) {
        if (g.fe64) (void)hipFree(g.fe64);
        if (g.fe32) (void)hipFree(g.fe32);
    }
    c->d_grids.release();
    c->d_mols.release();
    c->d_ref.release();
    c->d_seg64.release();
    c->d_seg32a.release();
    c->d_seg32b.release();
    c->d_rec.release();
    c->d_next.release();
    c->d_cum.release();
    c->d_order.release();
    c->d_counters.release();
    c->d_tmp.release();
    c->d_margins.release();
    c->h_rec.release();
    c->h_done.release();
    c->h_counters.release();
    c->d_fx_terms.release();
    c->d_fx_ints.release();
    c->d_fx_pos.release();
    c->d_fx_state.release();
    c->d_fx_work.release();
    c->d_sys.release();
    for (auto& kv : c->twiddles) kv.second.release();
    c->d_flag.release();
    (void)hipEventDestroy(c->ev0);
    (void)hipEventDestroy(c->ev1);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

void mafm_destroy(mafm_ctx* c) {
    if (!c) return;
    if (c->sims > 0) {              // finished by the last mafm_sim_close on this context
        c->destroy_pending = true;
        return;
    }
    destroy_now(c);
}

const char* mafm_last_error(const mafm_ctx* c) { return c ? c->err.c_str() : "no context"; }

mafm_status mafm_device_info(mafm_ctx* c, mafm_device_props* props) {
    if (!c || !props) return MAFM_ERR_INVALID;
    hipDeviceProp_t p;
    HIPCHK(c, hipGetDeviceProperties(&p, c->device));
    memset(props, 0, sizeof *props);
    snprintf(props->name, sizeof props->name, "%s", p.gcnArchName);
    props->compute_units = p.multiProcessorCount;
    props->clock_mhz = p.clockRate / 1000;
    props->total_mem = (int64_t)p.totalGlobalMem;
    props->lds_per_block = (int32_t)p.sharedMemPerBlock;
    props->warp_size = p.warpSize;
    return MAFM_OK;
}

mafm_status mafm_grid_upload(mafm_ctx* c, const mafm_grid_desc* g, int* handle) {
    if (!c || !handle) return MAFM_ERR_INVALID;
    if (!g || !g->fe) return fail(c, MAFM_ERR_INVALID, "grid upload needs sample values");
    HIPCHK(c, hipSetDevice(c->device));
    HostGrid h;
    GridGuard guard(h);
    mafm_status s = new_grid(c, g, h);
    if (s != MAFM_OK) return s;
    size_t np = (size_t)h.n[0] * h.n[1] * h.n[2];
    HIPCHK(c, hipMemcpyAsync(h.fe64, g->fe, np * sizeof(double4), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_to_f32(h.fe64, h.fe32, (long)np, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *handle = store_grid(c, h);
    guard.kept = true;
    return sync_grids(c);
}

mafm_status mafm_grid_download(mafm_ctx* c, int handle, double* fe) {
    if (!c || !fe) return MAFM_ERR_INVALID;
    if (handle < 0 || handle >= (int)c->grids.size() || !c->grids[handle].live)
        return fail(c, MAFM_ERR_INVALID, "grid handle %d is not live", handle);
    const HostGrid& h = c->grids[handle];
    size_t np = (size_t)h.n[0] * h.n[1] * h.n[2];
    HIPCHK(c, hipMemcpyAsync(fe, h.fe64, np * sizeof(double4), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MAFM_OK;
}

mafm_status mafm_grid_interpolate(mafm_ctx* c, int handle, size_t n, const double* pos, double* fe) {
    if (!c || (n && (!pos || !fe))) return MAFM_ERR_INVALID;
    if (handle < 0 || handle >= (int)c->grids.size() || !c->grids[handle].live)
        return fail(c, MAFM_ERR_INVALID, "grid handle %d is not live", handle);
    if (n == 0) return MAFM_OK;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, c->d_tmp.reserve(7 * n));
    double* dpos = c->d_tmp.p;
    double* dout = c->d_tmp.p + 3 * n;
    HIPCHK(c, hipMemcpyAsync(dpos, pos, 3 * n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(c->d_counters.p, 0, 8 * sizeof(int64_t), c->stream));
    ScanArgs a = base_args(c);
    HIPCHK(c, launch_grid_interp(a, handle, (long)n, dpos, dout, c->stream));
    int64_t counters[8];
    HIPCHK(c, hipMemcpyAsync(fe, dout, 4 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(counters, c->d_counters.p, sizeof counters, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (counters[3] != 0)
        return fail(c, MAFM_ERR_OUT_OF_GRID, "%lld positions needed a grid point past the grid",
                    (long long)counters[3]);
    return MAFM_OK;
}

mafm_status mafm_valu_calibration(mafm_ctx* c, int kind, int iterations, double* ms, int64_t* wave_instructions) {
    if (!c || !ms || !wave_instructions) return MAFM_ERR_INVALID;
    if (kind < 0 || kind > 2) return fail(c, MAFM_ERR_INVALID, "calibration kind %d (0 FP32 FMA, 1 FP64 FMA, 2 FP64 rsq)", kind);
    if (iterations < 1 || iterations > (1 << 20)) return fail(c, MAFM_ERR_INVALID, "iterations is %d", iterations);
    if (c->host_pending) return fail(c, MAFM_ERR_STATE, "a streamed scan is still running (mafm_scan_host_wait ends it)");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, c->d_tmp.reserve(valu_calibration_doubles()));
    int64_t insts = 0;
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK(c, launch_valu_calibration(kind, iterations, c->d_tmp.p, &insts, c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float t = 0;
    HIPCHK(c, hipEventElapsedTime(&t, c->ev0, c->ev1));
    *ms = t;
    *wave_instructions = insts;
    return MAFM_OK;
}

mafm_status mafm_grid_free(mafm_ctx* c, int handle) {
    if (!c) return MAFM_ERR_INVALID;
    if (handle < 0 || handle >= (int)c->grids.size() || !c->grids[handle].live)
        return fail(c, MAFM_ERR_INVALID, "grid handle %d is not live", handle);
    for (const DevMol& m : c->mols)
        if (m.force_grid == handle || m.potential_grid == handle)
            return fail(c, MAFM_ERR_STATE, "grid %d is used by a loaded system", handle);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HostGrid& h = c->grids[handle];
    (void)hipFree(h.fe64);
    (void)hipFree(h.fe32);
    h = HostGrid();
    return sync_grids(c);
}

static mafm_status upload_tables(mafm_ctx* c, const std::vector<DevMol>& mols, const std::vector<Site64>& ref,
                                 const std::vector<SiteF64>& seg64, const std::vector<Site32>& seg32) {
    HIPCHK(c, c->d_mols.reserve(mols.size()));
    HIPCHK(c, c->d_ref.reserve(ref.size()));
    HIPCHK(c, c->d_seg64.reserve(seg64.size()));
    HIPCHK(c, c->d_seg32a.reserve(seg32.size()));
    HIPCHK(c, c->d_seg32b.reserve(seg32.size()));
    if (!mols.empty())
        HIPCHK(c, hipMemcpyAsync(c->d_mols.p, mols.data(), mols.size() * sizeof(DevMol), hipMemcpyHostToDevice, c->stream));
    if (!ref.empty())
        HIPCHK(c, hipMemcpyAsync(c->d_ref.p, ref.data(), ref.size() * sizeof(Site64), hipMemcpyHostToDevice, c->stream));
    if (!seg64.empty())
        HIPCHK(c, hipMemcpyAsync(c->d_seg64.p, seg64.data(), seg64.size() * sizeof(SiteF64), hipMemcpyHostToDevice, c->stream));
    if (!seg32.empty()) {
        std::vector<float4> sa(seg32.size()), sb(seg32.size());
        for (size_t k = 0; k < seg32.size(); ++k) {
            sa[k] = seg32[k].a;
            sb[k] = seg32[k].b;
        }
        HIPCHK(c, hipMemcpyAsync(c->d_seg32a.p, sa.data(), sa.size() * sizeof(float4), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->d_seg32b.p, sb.data(), sb.size() * sizeof(float4), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MAFM_OK;
}

mafm_status mafm_grid_build(mafm_ctx* c, const mafm_system_desc* sys, const mafm_grid_desc* g, int* handle) {
    if (!c || !sys || !handle) return MAFM_ERR_INVALID;
    if (sys->force_grid >= 0) return fail(c, MAFM_ERR_INVALID, "grid build needs the tip-surface sites, not a grid");
    HIPCHK(c, hipSetDevice(c->device));
    // sample with a one-molecule table, then restore the loaded systems
    std::vector<DevMol> mols;
    std::vector<Site64> ref;
    std::vector<SiteF64> seg64;
    std::vector<Site32> seg32;
    std::vector<int64_t> cost;
    mafm_status s = build_tables(c, 1, sys, mols, ref, seg64, seg32, cost);
    if (s != MAFM_OK) return s;
    HostGrid h;
    GridGuard guard(h);
    s = new_grid(c, g, h);
    if (s != MAFM_OK) return s;
    if (!h.orthogonal) return fail(c, MAFM_ERR_INVALID, "grid build expects a diagonal spacing basis");
    DevBuf<DevMol> dm;
    DevBuf<Site64> dr;
    HIPCHK(c, dm.reserve(1));
    HIPCHK(c, dr.reserve(ref.size()));
    HIPCHK(c, hipMemcpyAsync(dm.p, mols.data(), sizeof(DevMol), hipMemcpyHostToDevice, c->stream));
    if (!ref.empty())
        HIPCHK(c, hipMemcpyAsync(dr.p, ref.data(), ref.size() * sizeof(Site64), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(c->d_counters.p, 0, 8 * sizeof(int64_t), c->stream));
    ScanArgs a = base_args(c);
    a.mols = dm.p;
    a.sites_ref = dr.p;
    a.n_mols = 1;
    const double sp[3] = {h.basis[0], h.basis[4], h.basis[8]};
    HIPCHK(c, launch_grid_build(a, 0, h.n, sp, h.offset, h.fe64, c->stream));
    size_t np = (size_t)h.n[0] * h.n[1] * h.n[2];
    HIPCHK(c, launch_to_f32(h.fe64, h.fe32, (long)np, c->stream));
    int64_t counters[8];
    HIPCHK(c, hipMemcpyAsync(counters, c->d_counters.p, sizeof counters, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dm.release();
    dr.release();
    if (counters[3] != 0) return fail(c, MAFM_ERR_OUT_OF_GRID, "potential-grid lookup left the grid during grid build");
    *handle = store_grid(c, h);
    guard.kept = true;
    return sync_grids(c);
}

mafm_status mafm_set_systems(mafm_ctx* c, int n, const mafm_system_desc* sys) {
    if (!c || n <= 0 || !sys) return fail(c, MAFM_ERR_INVALID, "need at least one system");
    if (c->host_pending)
        return fail(c, MAFM_ERR_STATE, "a streamed scan of these systems is still running (mafm_scan_host_wait ends it)");
    HIPCHK(c, hipSetDevice(c->device));
    std::vector<DevMol> mols;
    std::vector<Site64> ref;
    std::vector<SiteF64> seg64;
    std::vector<Site32> seg32;
    std::vector<int64_t> cost;
    mafm_status s = build_tables(c, n, sys, mols, ref, seg64, seg32, cost);
    if (s != MAFM_OK) return s;
    s = upload_tables(c, mols, ref, seg64, seg32);
    if (s != MAFM_OK) return s;
    c->mols = mols;
    c->mol_cost = cost;
    c->have_systems = true;
    return MAFM_OK;
}

// The scan behind mafm_scan (records and work-stealing state in buffers the
// context allocates: out = nullptr, ws = nullptr) and mafm_scan_noalloc
// (records into the caller's device buffer out of `cap` records, state in
// the fixed workspace ws).
// The scan geometry and minimiser settings every scan entry point checks.
static mafm_status check_scan_desc(mafm_ctx* c, const mafm_scan_desc* d) {
    if (d->nx <= 0 || d->ny <= 0 || d->nz <= 0) return fail(c, MAFM_ERR_INVALID, "empty scan grid");
    if (d->minimiser != MAFM_MIN_SD && d->minimiser != MAFM_MIN_FIRE) return fail(c, MAFM_ERR_INVALID, "minimiser");
    if (d->integrator < 0 || d->integrator > 2) return fail(c, MAFM_ERR_INVALID, "integrator");
    if (d->minterm < 0 || d->minterm > 2) return fail(c, MAFM_ERR_INVALID, "minterm");
    if (d->precision < 0 || d->precision > 2) return fail(c, MAFM_ERR_INVALID, "precision");
    if ((int64_t)d->nx * d->ny > INT32_MAX) return fail(c, MAFM_ERR_INVALID, "more than 2^31 pixel columns");
    if ((int64_t)d->nx * d->ny * d->nz > ((int64_t)1 << 40)) return fail(c, MAFM_ERR_INVALID, "scan grid too large");
    if (!std::isfinite(d->dz) || d->dz == 0 || !std::isfinite(d->dt))
        return fail(c, MAFM_ERR_INVALID, "dz must be finite and non-zero, dt finite");
    return MAFM_OK;
}

// Totals of a streamed scan's records on the host (what k_record_totals
// computes on the device for the other scans): [0] minimiser iterations,
// [1] unconverged positions, [2] grid-clamp warnings, [3] out-of-grid lookups.
static void host_record_totals(const DevRecord* r, size_t n, int64_t* out) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < ((size_t)1 << 16)) nt = 1;
    std::vector<std::array<int64_t, 4>> part(nt);
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; ++t)
        pool.emplace_back([&, t]() {
            int64_t s = 0, u = 0, w = 0, o = 0;
            for (size_t i = n * t / nt, e = n * (t + 1) / nt; i < e; ++i) {
                s += (uint32_t)r[i].n_steps;
                u += (r[i].flags & REC_UNCONVERGED) != 0;
                w += (r[i].flags & REC_GRID_CLAMP) != 0;
                o += (r[i].flags & REC_OUT_OF_GRID) != 0;
            }
            part[t] = {s, u, w, o};
        });
    for (auto& th : pool) th.join();
    for (auto& p : part)
        for (int k = 0; k < 4; ++k) out[k] += p[k];
}

// The statistics of a finished scan from its counters and event times.
static mafm_status scan_stats(mafm_ctx* c, const PendingScan& ps, const int64_t* counters, mafm_scan_stats* stats) {
    const mafm_scan_desc* d = &ps.d;
    const int n_mols = (int)c->mols.size();
    const size_t per_mol = (size_t)d->nx * d->ny * d->nz;
    const long cols_per_mol = (long)d->nx * d->ny;
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (stats) {
        memset(stats, 0, sizeof *stats);
        stats->kernel_ms = ms;
        stats->positions = (int64_t)ps.n_rec;
        stats->total_steps = counters[0];
        stats->unconverged = counters[1];
        stats->grid_warnings = counters[2];
        int64_t evals_per_step = d->minimiser == MAFM_MIN_SD ? 1 : (d->integrator == 1 ? 2 : (d->integrator == 2 ? 4 : 1));
        int64_t pe = 0;
        for (int m = 0; m < n_mols; ++m) {
            int64_t sites = c->mols[m].force_grid >= 0 ? 0 : (int64_t)c->mols[m].n_sites * c->mols[m].n_images;
            pe += sites;
        }
        // algorithmic pair evaluations: (steps * evals + one getOutput eval) per position
        stats->pair_evals = n_mols ? (counters[0] * evals_per_step + (int64_t)per_mol * n_mols) * (pe / n_mols) : 0;
        stats->lanes_per_pixel = ps.fast ? ps.G : 1;
        stats->group_trips = counters[4];
        stats->active_group_trips = counters[5];
        stats->wide_trips = counters[6];
        stats->waves_per_simd = ps.fast ? ps.wps : 0;
        stats->grid_blocks = ps.fast ? ps.n_blocks : (int)((cols_per_mol * n_mols + 255) / 256);
        stats->block_threads = ps.fast ? 64 * ps.W : 256;
    }
    if (counters[3] != 0)
        return fail(c, MAFM_ERR_OUT_OF_GRID, "%lld positions needed a force-grid point past the grid",
                    (long long)counters[3]);
    return MAFM_OK;
}

static mafm_status scan_impl(mafm_ctx* c, const mafm_scan_desc* d, mafm_scan_stats* stats, DevRecord* out, size_t cap,
                             ScanScratch* ws, bool to_host = false) {
    if (!c || !d) return MAFM_ERR_INVALID;
    if (c->host_pending) return fail(c, MAFM_ERR_STATE, "a streamed scan is still running (mafm_scan_host_wait ends it)");
    if (!c->have_systems) return fail(c, MAFM_ERR_STATE, "mafm_scan before mafm_set_systems");
    mafm_status sd = check_scan_desc(c, d);
    if (sd != MAFM_OK) return sd;
    if (d->precision != MAFM_PREC_PLAIN64 && d->lanes_per_pixel > 0) {
        const int G = d->lanes_per_pixel;
        if (G != 1 && G != 2 && G != 4 && G != 8 && G != 16 && G != 32 && G != 64)
            return fail(c, MAFM_ERR_INVALID, "lanes_per_pixel must be a power of two <= 64");
    }
    if (d->waves_per_group != 0 && (d->waves_per_group < 1 || d->waves_per_group > 8))
        return fail(c, MAFM_ERR_INVALID, "waves_per_group must be 1 .. 8");
    const int n_mols = (int)c->mols.size();
    const size_t per_mol = (size_t)d->nx * d->ny * d->nz;
    const size_t n_rec = per_mol * n_mols;
    const long cols_per_mol = (long)d->nx * d->ny;
    if (ws) {
        if (n_mols > NOALLOC_MAX_SYSTEMS)
            return fail(c, MAFM_ERR_INVALID, "%d systems: the allocation-free scan takes at most %d", n_mols,
                        NOALLOC_MAX_SYSTEMS);
        if (!out) return fail(c, MAFM_ERR_INVALID, "the allocation-free scan needs a device record buffer");
        if (cap < n_rec)
            return fail(c, MAFM_ERR_INVALID, "record buffer holds %zu records, the scan writes %zu", cap, n_rec);
        if (d->options & MAFM_SCAN_MARGINS)
            return fail(c, MAFM_ERR_INVALID, "the allocation-free scan records no decision margins");
    }
    if (to_host && (d->options & MAFM_SCAN_MARGINS))
        return fail(c, MAFM_ERR_INVALID, "a streamed scan records no decision margins");
    HIPCHK(c, hipSetDevice(c->device));
    for (int m = 0; m < n_mols; ++m) c->mols[m].rec_off = (int64_t)(per_mol * m);
    HIPCHK(c, hipMemcpyAsync(c->d_mols.p, c->mols.data(), n_mols * sizeof(DevMol), hipMemcpyHostToDevice, c->stream));
    // the records of an earlier scan are no longer described by the context
    ++c->scan_generation;
    c->have_margins = false;
    c->last_flex = false;
    c->fx_state_atoms = c->fx_state_heights = 0;
    if (to_host) {
        // the records of this scan live in pinned host memory only
        HIPCHK(c, c->h_rec.reserve(n_rec));
        HIPCHK(c, c->h_done.reserve((size_t)cols_per_mol * n_mols));
        HIPCHK(c, c->h_counters.reserve(8));
        HIPCHK(c, hipStreamSynchronize(c->stream));   // no earlier work still reads the flags
        memset(c->h_done.p, 0, (size_t)cols_per_mol * n_mols * sizeof(int32_t));
        c->have_scan = false;
        c->n_records = 0;
        out = c->h_rec.dp;
    } else if (out) {
        c->have_scan = false;
        c->n_records = 0;
    } else {
        HIPCHK(c, c->d_rec.reserve(n_rec));
        c->n_records = n_rec;
        c->last_scan = *d;
        c->have_scan = true;
        out = c->d_rec.p;
    }
    ScanArgs a = base_args(c);
    if (ws) a.counters = ws->counters;
    a.nx = d->nx; a.ny = d->ny; a.nz = d->nz;
    a.dx = d->dx; a.dy = d->dy; a.dz = d->dz; a.zhigh = d->zhigh;
    a.minimiser = d->minimiser; a.integrator = d->integrator; a.minterm = d->minterm;
    a.maxsteps = d->maxsteps; a.etol = d->etol; a.ftol = d->ftol; a.dt = d->dt;
    a.rec = out;
    if (to_host) a.col_done = c->h_done.dp;
    // last groups of a wave: regroup over 64 lanes once that halves the sites per lane
    a.wide_tail = (d->options & MAFM_SCAN_REPRODUCIBLE) ? 0 : 2;
    a.inv_2ftol = d->ftol > 0 ? 1.0 / (2.0 * d->ftol) : 0.0;
    const bool fast = d->precision != MAFM_PREC_PLAIN64;
    if (fast && (d->options & MAFM_SCAN_MARGINS)) {
        HIPCHK(c, c->d_margins.reserve(n_rec));
        a.margins = c->d_margins.p;
        c->have_margins = true;
    }
    // host image of the fixed workspace: zero counters plus the launch plan
    ScanScratch* img = ws ? c->scratch_img.get() : nullptr;
    if (ws) memset(img->counters, 0, sizeof img->counters);
    else HIPCHK(c, hipMemsetAsync(c->d_counters.p, 0, 8 * sizeof(int64_t), c->stream));

    FastPlan plan;
    if (fast) {
        std::vector<long> cols(n_mols, cols_per_mol);
        mafm_status s = plan_fast(c, d, d->precision == MAFM_PREC_FAST32, cols, d->lanes_per_pixel,
                                  d->waves_per_simd, d->waves_per_group, a, plan, img, ws);
        if (s != MAFM_OK) return s;
    }
    if (ws) {
        HIPCHK(c, hipMemcpyAsync(ws, img, sizeof(ScanScratch), hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    if (!fast) {
        HIPCHK(c, launch_scan_plain(a, c->stream));
    } else {
        // 0 / 3: block-compensated (default), 1: Kahan, 2: plain FP32
        const int comp = d->compensation == 1 ? 1 : (d->compensation == 2 ? 0 : 2);
        HIPCHK(c, launch_scan_fast(a, d->precision == MAFM_PREC_FAST32, comp, plan.G, plan.n_blocks, plan.W,
                                   plan.lds_sites, plan.scalar_sites, plan.mid_family, c->stream));
    }
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    PendingScan ps;
    ps.d = *d;
    ps.fast = fast;
    ps.G = plan.G;
    ps.wps = plan.wps;
    ps.n_blocks = plan.n_blocks;
    ps.W = plan.W;
    ps.n_rec = n_rec;
    if (to_host) {
        // mafm_scan_host_wait finishes it; the counters are copied out in
        // stream order now, so later work on the context may reuse them
        HIPCHK(c, hipMemcpyAsync(c->h_counters.p, a.counters, 8 * sizeof(int64_t), hipMemcpyDeviceToHost,
                                 c->stream));
        c->pending = ps;
        c->host_pending = true;
        return MAFM_OK;
    }
    // the FAST kernels leave the step / convergence / grid totals to the records
    if (fast) HIPCHK(c, launch_record_totals(out, (long)n_rec, a.counters, c->stream));
    int64_t counters[8];
    HIPCHK(c, hipMemcpyAsync(counters, a.counters, sizeof counters, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return scan_stats(c, ps, counters, stats);
}

mafm_status mafm_scan(mafm_ctx* c, const mafm_scan_desc* d, mafm_scan_stats* stats) {
    return scan_impl(c, d, stats, nullptr, 0, nullptr);
}

mafm_status mafm_scan_host(mafm_ctx* c, const mafm_scan_desc* d, const mafm_record** records,
                           const int32_t** column_done) {
    if (!c || !d || !records || !column_done) return MAFM_ERR_INVALID;
    mafm_status s = scan_impl(c, d, nullptr, nullptr, 0, nullptr, true);
    if (s != MAFM_OK) return s;
    static_assert(sizeof(mafm_record) == sizeof(DevRecord), "record layout");
    *records = reinterpret_cast<const mafm_record*>(c->h_rec.p);
    *column_done = c->h_done.p;
    return MAFM_OK;
}

mafm_status mafm_scan_host_query(mafm_ctx* c, int* finished) {
    if (!c || !finished) return MAFM_ERR_INVALID;
    if (!c->host_pending) return fail(c, MAFM_ERR_STATE, "no streamed scan is running");
    hipError_t e = hipEventQuery(c->ev1);
    if (e == hipErrorNotReady) {
        *finished = 0;
        return MAFM_OK;
    }
    HIPCHK(c, e);
    *finished = 1;
    return MAFM_OK;
}

mafm_status mafm_scan_host_wait(mafm_ctx* c, mafm_scan_stats* stats) {
    if (!c) return MAFM_ERR_INVALID;
    if (!c->host_pending) return fail(c, MAFM_ERR_STATE, "no streamed scan is running");
    c->host_pending = false;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int64_t counters[8];
    memcpy(counters, c->h_counters.p, sizeof counters);
    if (c->pending.fast) host_record_totals(c->h_rec.p, c->pending.n_rec, counters);
    return scan_stats(c, c->pending, counters, stats);
}

mafm_status mafm_scan_noalloc(mafm_ctx* c, const mafm_scan_desc* d, mafm_record* records, size_t capacity,
                              mafm_scan_stats* stats) {
    if (!c || !d) return MAFM_ERR_INVALID;
    if (!c->scratch) return fail(c, MAFM_ERR_STATE, "no device workspace on this context");
    return scan_impl(c, d, stats, reinterpret_cast<DevRecord*>(records), capacity,
                     static_cast<ScanScratch*>(c->scratch));
}

// Number of atoms a flexible term acts on, and the order in which the
// reference's eval updates them (interactions.cpp:15-51, 252-362): an angle
// updates atom_i1, atom_i2, then the shared atom.
static int flex_arity(int kind) {
    switch (kind) {
        case MAFM_FX_LJ: case MAFM_FX_MORSE: case MAFM_FX_COULOMB: case MAFM_FX_TIP_HARMONIC: case MAFM_FX_BOND:
            return 2;
        case MAFM_FX_ANGLE: return 3;
        case MAFM_FX_DIHEDRAL: return 4;
        case MAFM_FX_SUBSTRATE: case MAFM_FX_XY_HARMONIC: return 1;
        default: return 0;
    }
}

static const int* flex_update_order(int kind) {
    static const int pair[4] = {0, 1, 2, 3}, angle[3] = {1, 2, 0};
    return kind == MAFM_FX_ANGLE ? angle : pair;
}

static bool flex_tip_surface(const mafm_flex_term& t) {   // isTipSurface (interactions.hpp:79-137)
    return (t.kind == MAFM_FX_LJ || t.kind == MAFM_FX_MORSE || t.kind == MAFM_FX_COULOMB) && t.atom[0] == 1;
}

// Relative length of one term evaluation's dependent chain by kind (square
// roots, IEEE divisions and libm calls set it: a dihedral has three roots,
// seven divisions and an atan, an angle an acos, a Morse term an exp).
static double flex_cost(int kind) {
    switch (kind) {
        case MAFM_FX_DIHEDRAL: return 16;
        case MAFM_FX_ANGLE: return 8;
        case MAFM_FX_MORSE: return 5;
        case MAFM_FX_SUBSTRATE: return 4;
        case MAFM_FX_XY_HARMONIC: return 1;
        default: return 3;
    }
}

// Which terms each thread of a `waves`-wavefront block evaluates (the
// FlexArgs::sched table): the terms of each kind, in list order, are cut
// into wave-rounds of 64, so the 64 lanes of a round run one code path, and
// the rounds are dealt longest first, each to the wave with the least work
// so far.  Returns the number of rounds.
static int flex_schedule(const std::vector<mafm_flex_term>& eval, int waves, std::vector<int32_t>& sched) {
    std::map<int, std::vector<int32_t>> by_kind;
    for (int k = 0; k < (int)eval.size(); ++k) by_kind[eval[k].kind].push_back(k);
    struct Round {
        double cost;
        const int32_t* terms;
        int n;
    };
    std::vector<Round> rounds;
    for (const auto& kv : by_kind)
        for (size_t q = 0; q < kv.second.size(); q += 64)
            rounds.push_back({flex_cost(kv.first), kv.second.data() + q, (int)std::min<size_t>(64, kv.second.size() - q)});
    std::stable_sort(rounds.begin(), rounds.end(), [](const Round& x, const Round& y) { return x.cost > y.cost; });
    std::vector<double> load(waves, 0.0);
    std::vector<std::vector<int>> mine(waves);
    for (int r = 0; r < (int)rounds.size(); ++r) {
        const int w = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[w] += rounds[r].cost;
        mine[w].push_back(r);
    }
    int n_rounds = 0;
    for (const auto& m : mine) n_rounds = std::max(n_rounds, (int)m.size());
    const int T = 64 * waves;
    sched.assign((size_t)n_rounds * T, -1);
    for (int w = 0; w < waves; ++w)
        for (int r = 0; r < (int)mine[w].size(); ++r) {
            const Round& R = rounds[mine[w][r]];
            for (int l = 0; l < R.n; ++l) sched[(size_t)r * T + 64 * w + l] = R.terms[l];
        }
    return n_rounds;
}

mafm_status mafm_scan_flexible(mafm_ctx* c, const mafm_flex_desc* s, const mafm_scan_desc* d,
                               mafm_scan_stats* stats) {
    if (!c || !s || !d) return MAFM_ERR_INVALID;
    if (c->host_pending) return fail(c, MAFM_ERR_STATE, "a streamed scan is still running (mafm_scan_host_wait ends it)");
    mafm_status sd = check_scan_desc(c, d);
    if (sd != MAFM_OK) return sd;
    const int n = s->n_atoms;
    if (n < 2 || n > (1 << 20)) return fail(c, MAFM_ERR_INVALID, "n_atoms is %d", n);
    if (!s->positions || !s->fixed || s->n_terms < 0 || (s->n_terms > 0 && !s->terms))
        return fail(c, MAFM_ERR_INVALID, "flexible system: missing positions, fixed flags or terms");
    if (s->fixed[0] != 1 || s->fixed[1] == 1)
        return fail(c, MAFM_ERR_INVALID, "the dummy (atom 0) must be fixed and the tip (atom 1) free");
    for (int i = 0; i < 3 * n; ++i)
        if (i >= 6 && !std::isfinite(s->positions[i]))
            return fail(c, MAFM_ERR_INVALID, "atom %d has a non-finite position", i / 3);
    const long n_cols = (long)d->nx * d->ny;
    if (s->state_column < -1 || s->state_column >= n_cols)
        return fail(c, MAFM_ERR_INVALID, "state column %d outside the %ld columns", s->state_column, n_cols);
    // tip-surface prefix, arity, indices
    int n_ts = 0;
    while (n_ts < s->n_terms && flex_tip_surface(s->terms[n_ts])) ++n_ts;
    for (int t = 0; t < s->n_terms; ++t) {
        const mafm_flex_term& T = s->terms[t];
        const int ar = flex_arity(T.kind);
        if (ar == 0) return fail(c, MAFM_ERR_INVALID, "term %d: unknown kind %d", t, T.kind);
        for (int r = 0; r < ar; ++r)
            if (T.atom[r] < 0 || T.atom[r] >= n)
                return fail(c, MAFM_ERR_INVALID, "term %d: atom %d outside 0..%d", t, T.atom[r], n - 1);
        if (t >= n_ts && flex_tip_surface(T))
            return fail(c, MAFM_ERR_INVALID, "term %d is a tip-surface term after the first non-tip-surface term", t);
    }
    // terms acting on a free atom, their result slots, and each free atom's slots in list order
    // the terms acting on a free atom (in list order, each with its first
    // result slot), each free atom's slots in list order, the tip's terms
    std::vector<mafm_flex_term> eval;
    std::vector<int32_t> csr_off(n + 1, 0), csr, tip_eval;
    int n_slots = 0, ts_slots = 0, n_ts_eval = 0;
    std::vector<std::vector<int32_t>> per_atom(n);
    for (int t = 0; t < s->n_terms; ++t) {
        const mafm_flex_term& T = s->terms[t];
        const int ar = flex_arity(T.kind);
        bool any_free = false;
        for (int r = 0; r < ar; ++r) any_free |= s->fixed[T.atom[r]] != 1;
        if (!any_free) continue;
        const int k = (int)eval.size();
        eval.push_back(T);
        eval.back().reserved = n_slots;
        const int* order = flex_update_order(T.kind);
        for (int q = 0; q < ar; ++q) {
            const int r = order[q];
            if (s->fixed[T.atom[r]] != 1) per_atom[T.atom[r]].push_back(n_slots + r);
            if (T.atom[r] == 1) tip_eval.push_back(k);   // energies_[1] += e
        }
        n_slots += ar;
        if (t < n_ts) {
            ts_slots = n_slots;
            n_ts_eval = k + 1;
        }
    }
    for (int i = 0; i < n; ++i) {
        csr_off[i + 1] = csr_off[i] + (int32_t)per_atom[i].size();
        csr.insert(csr.end(), per_atom[i].begin(), per_atom[i].end());
    }
    const int n_eval = (int)eval.size();
    int tip_ts_slots = 0, tip_ts_terms = 0;
    for (int32_t sl : per_atom[1]) tip_ts_slots += sl < ts_slots;
    for (int32_t k : tip_eval) tip_ts_terms += k < n_ts_eval;
    const size_t n_rec = (size_t)n_cols * d->nz;

    HIPCHK(c, hipSetDevice(c->device));
    ++c->scan_generation;
    c->have_margins = false;
    c->have_scan = false;
    c->n_records = 0;
    c->last_flex = false;
    c->fx_state_atoms = c->fx_state_heights = 0;
    std::vector<int32_t> ints;
    ints.reserve((n + 1) + csr.size() + tip_eval.size() + n);
    ints.insert(ints.end(), csr_off.begin(), csr_off.end());
    ints.insert(ints.end(), csr.begin(), csr.end());
    ints.insert(ints.end(), tip_eval.begin(), tip_eval.end());
    ints.insert(ints.end(), s->fixed, s->fixed + n);

    FlexArgs a;
    memset(&a, 0, sizeof a);
    a.n_atoms = n; a.n_eval = n_eval; a.n_ts = n_ts_eval;
    a.n_slots = n_slots; a.ts_slots = ts_slots;
    a.n_csr = (int32_t)csr.size();
    a.n_tip = (int32_t)tip_eval.size();
    a.tip_ts_slots = tip_ts_slots;
    a.tip_ts_terms = tip_ts_terms;
    a.nx = d->nx; a.ny = d->ny; a.nz = d->nz;
    a.dx = d->dx; a.dy = d->dy; a.dz = d->dz; a.zhigh = d->zhigh;
    a.minimiser = d->minimiser; a.integrator = d->integrator; a.minterm = d->minterm;
    a.maxsteps = d->maxsteps; a.etol = d->etol; a.ftol = d->ftol; a.dt = d->dt;
    a.tip_pbc = s->tip_pbc ? 1 : 0;
    a.state_col = s->state_column;
    memcpy(a.cell, s->cell, sizeof a.cell);
    if (a.tip_pbc) inv3(a.cell, a.cell_inv);
    memcpy(a.offset, s->offset, sizeof a.offset);
    a.tip_dummy_d = s->tip_dummy_d;
    const bool rk4 = d->minimiser == MAFM_MIN_FIRE && d->integrator == MAFM_INT_RK4;
    // the column state in LDS when one workgroup may hold it (gfx950: up to
    // 160 KiB), as many columns per CU as its LDS holds (up to 8)
    const size_t block_bytes = flex_block_lds(a) + 64;
    const bool in_lds = !(d->options & MAFM_SCAN_STATE_IN_HBM) && block_bytes <= c->lds_per_block;
    // up to 4 waves share a column (terms and atoms spread over 64-256 threads);
    // two resident waves per SIMD, and as many columns per CU as its LDS holds
    const int waves = std::max(n_eval, n) > 128 ? 4 : (std::max(n_eval, n) > 64 ? 2 : 1);
    int per_cu = 8 / waves;
    if (in_lds) per_cu = (int)std::max<size_t>(1, std::min<size_t>(per_cu, c->lds_per_cu / block_bytes));
    const int n_blocks = (int)std::min<long>(n_cols, (long)c->cus * per_cu);
    // the term schedule goes behind the index tables (read once per thread,
    // not copied to LDS)
    std::vector<int32_t> sched;
    const int sched_rounds = flex_schedule(eval, waves, sched);
    const size_t sched_at = ints.size();
    ints.insert(ints.end(), sched.begin(), sched.end());

    HIPCHK(c, c->d_fx_terms.reserve(std::max(n_eval, 1)));
    HIPCHK(c, c->d_fx_ints.reserve(ints.size()));
    HIPCHK(c, c->d_fx_pos.reserve(3 * (size_t)n));
    HIPCHK(c, c->d_rec.reserve(n_rec));
    HIPCHK(c, c->d_next.reserve(1));
    if (s->state_column >= 0) HIPCHK(c, c->d_fx_state.reserve((size_t)d->nz * n * 3));
    if (!in_lds) HIPCHK(c, c->d_fx_work.reserve((size_t)n_blocks * flex_block_doubles(n, n_slots, n_eval, rk4)));
    if (n_eval > 0)
        HIPCHK(c, hipMemcpyAsync(c->d_fx_terms.p, eval.data(), n_eval * sizeof(mafm_flex_term),
                                 hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_fx_ints.p, ints.data(), ints.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_fx_pos.p, s->positions, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipMemsetAsync(c->d_next.p, 0, sizeof(int32_t), c->stream));
    HIPCHK(c, hipMemsetAsync(c->d_counters.p, 0, 8 * sizeof(int64_t), c->stream));

    a.terms = c->d_fx_terms.p;
    a.idx = c->d_fx_ints.p;
    a.sched = sched_rounds > 0 ? c->d_fx_ints.p + sched_at : nullptr;
    a.sched_rounds = sched_rounds;
    a.pos0 = c->d_fx_pos.p;
    a.rec = c->d_rec.p;
    a.next_col = c->d_next.p;
    a.state = s->state_column >= 0 ? c->d_fx_state.p : nullptr;
    a.work = in_lds ? nullptr : c->d_fx_work.p;

    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK(c, launch_scan_flex(a, n_blocks, waves, in_lds, c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, launch_record_totals(c->d_rec.p, (long)n_rec, c->d_counters.p, c->stream));
    int64_t counters[8];
    HIPCHK(c, hipMemcpyAsync(counters, c->d_counters.p, sizeof counters, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->n_records = n_rec;
    c->last_scan = *d;
    c->have_scan = true;
    c->last_flex = true;
    if (s->state_column >= 0) {
        c->fx_state_atoms = n;
        c->fx_state_heights = d->nz;
    }
    if (stats) {
        memset(stats, 0, sizeof *stats);
        stats->kernel_ms = ms;
        stats->positions = (int64_t)n_rec;
        stats->total_steps = counters[0];
        stats->unconverged = counters[1];
        const int64_t evals = d->minimiser == MAFM_MIN_SD ? 1 : (d->integrator == 1 ? 2 : (d->integrator == 2 ? 4 : 1));
        // term evaluations: every evaluated term per integrator stage, plus the tip-surface terms of each output
        stats->pair_evals = counters[0] * evals * n_eval + (int64_t)n_rec * a.n_ts;
        stats->lanes_per_pixel = 64 * waves;
        stats->grid_blocks = n_blocks;
        stats->block_threads = 64 * waves;
    }
    return MAFM_OK;
}

// ---------------------------------------------------------------- System operators
// The tables of one interaction list for k_system_op: every term with its
// first result slot, each atom's slots in the order Interaction::eval adds
// to forces_[i], each atom's terms in the order it adds to energies_[i].
static mafm_status system_op(mafm_ctx* c, const mafm_interaction_list* L, int op, const mafm_minimiser_opts* o,
                             int integrator, double dt, int64_t n_states, double* pos, double* vel, double* frc,
                             double* ener, int32_t* steps) {
    if (!c || !L) return MAFM_ERR_INVALID;
    if (n_states < 0) return fail(c, MAFM_ERR_INVALID, "n_states is %lld", (long long)n_states);
    if (n_states == 0) return MAFM_OK;
    const int n = L->n_atoms;
    if (n < 2 || n > (1 << 20)) return fail(c, MAFM_ERR_INVALID, "n_atoms is %d (the tip is atom 1)", n);
    if (!L->fixed || L->n_terms < 0 || (L->n_terms > 0 && !L->terms))
        return fail(c, MAFM_ERR_INVALID, "interaction list: missing fixed flags or terms");
    if (!pos || !frc || !ener || (op != SYS_EVAL && !vel) || (op == SYS_MINIMISE && !steps))
        return fail(c, MAFM_ERR_INVALID, "missing state array");
    if (op == SYS_STEP && (integrator < 0 || integrator > 2)) return fail(c, MAFM_ERR_INVALID, "integrator");
    if (op == SYS_MINIMISE) {
        if (!o) return fail(c, MAFM_ERR_INVALID, "minimiser options are NULL");
        if (o->minimiser != MAFM_MIN_SD && o->minimiser != MAFM_MIN_FIRE)
            return fail(c, MAFM_ERR_INVALID, "minimiser");
        if (o->integrator < 0 || o->integrator > 2) return fail(c, MAFM_ERR_INVALID, "integrator");
        if (o->minterm < 0 || o->minterm > 2) return fail(c, MAFM_ERR_INVALID, "minterm");
    }
    std::vector<mafm_flex_term> terms(L->terms, L->terms + L->n_terms);
    std::vector<std::vector<int32_t>> fslots(n), eterms(n);
    int n_slots = 0;
    for (int k = 0; k < L->n_terms; ++k) {
        mafm_flex_term& T = terms[k];
        const int ar = flex_arity(T.kind);
        if (ar == 0) return fail(c, MAFM_ERR_INVALID, "term %d: unknown kind %d", k, T.kind);
        for (int r = 0; r < ar; ++r)
            if (T.atom[r] < 0 || T.atom[r] >= n)
                return fail(c, MAFM_ERR_INVALID, "term %d: atom %d outside 0..%d", k, T.atom[r], n - 1);
        T.reserved = n_slots;
        const int* order = flex_update_order(T.kind);
        for (int q = 0; q < ar; ++q) {
            const int r = order[q];
            fslots[T.atom[r]].push_back(n_slots + r);
            eterms[T.atom[r]].push_back(k);
        }
        n_slots += ar;
    }
    std::vector<int32_t> ints;
    auto csr = [&](const std::vector<std::vector<int32_t>>& lists) {
        size_t start = ints.size(), total = 0;
        ints.resize(start + n + 1);
        for (int i = 0; i < n; ++i) {
            ints[start + i] = (int32_t)total;
            total += lists[i].size();
        }
        ints[start + n] = (int32_t)total;
        for (const auto& l : lists) ints.insert(ints.end(), l.begin(), l.end());
        return (int32_t)total;
    };
    SysArgs a;
    memset(&a, 0, sizeof a);
    a.n_fcsr = csr(fslots);
    a.n_ecsr = csr(eterms);
    ints.insert(ints.end(), L->fixed, L->fixed + n);
    a.n_atoms = n;
    a.n_terms = L->n_terms;
    a.n_slots = n_slots;
    a.op = op;
    a.integrator = integrator;
    a.dt = dt;
    if (o) {
        a.minimiser = o->minimiser;
        a.integrator = op == SYS_MINIMISE ? o->integrator : integrator;
        a.minterm = o->minterm;
        a.maxsteps = o->maxsteps;
        a.etol = o->etol;
        a.ftol = o->ftol;
        a.dt = op == SYS_MINIMISE ? o->dt : dt;
    }
    a.n_states = n_states;
    const size_t lds = sys_block_bytes(a) + 64;
    const bool in_lds = lds <= c->lds_per_block;
    const int waves = std::max(L->n_terms, n) > 128 ? 4 : (std::max(L->n_terms, n) > 64 ? 2 : 1);
    const size_t n3 = (size_t)n_states * n * 3, n1 = (size_t)n_states * n;
    const int64_t blocks = std::min<int64_t>(n_states, 1 << 20);
    const size_t nd = (sys_block_bytes(a) - ints.size() * sizeof(int32_t)) / sizeof(double);

    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, c->d_fx_terms.reserve(std::max(L->n_terms, 1)));
    HIPCHK(c, c->d_fx_ints.reserve(ints.size()));
    HIPCHK(c, c->d_sys.reserve(3 * n3 + n1 + (n_states + 1) / 2 + 1));
    if (!in_lds) HIPCHK(c, c->d_fx_work.reserve((size_t)blocks * nd));
    double* d_pos = c->d_sys.p;
    double* d_vel = d_pos + n3;
    double* d_frc = d_vel + n3;
    double* d_ener = d_frc + n3;
    int32_t* d_steps = reinterpret_cast<int32_t*>(d_ener + n1);
    if (L->n_terms > 0)
        HIPCHK(c, hipMemcpyAsync(c->d_fx_terms.p, terms.data(), terms.size() * sizeof(mafm_flex_term),
                                 hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_fx_ints.p, ints.data(), ints.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipMemcpyAsync(d_pos, pos, n3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(d_frc, frc, n3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(d_ener, ener, n1 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (op != SYS_EVAL) HIPCHK(c, hipMemcpyAsync(d_vel, vel, n3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    a.terms = c->d_fx_terms.p;
    a.idx = c->d_fx_ints.p;
    a.pos = d_pos;
    a.vel = d_vel;
    a.frc = d_frc;
    a.ener = d_ener;
    a.steps = d_steps;
    a.work = in_lds ? nullptr : c->d_fx_work.p;
    HIPCHK(c, launch_system_op(a, waves, in_lds, c->stream));
    HIPCHK(c, hipMemcpyAsync(pos, d_pos, n3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(frc, d_frc, n3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(ener, d_ener, n1 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (op != SYS_EVAL) HIPCHK(c, hipMemcpyAsync(vel, d_vel, n3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (op == SYS_MINIMISE)
        HIPCHK(c, hipMemcpyAsync(steps, d_steps, n_states * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MAFM_OK;
}

mafm_status mafm_system_eval(mafm_ctx* c, const mafm_interaction_list* list, int64_t n_states, const double* positions,
                             double* forces, double* energies) {
    return system_op(c, list, SYS_EVAL, nullptr, 0, 0.0, n_states, const_cast<double*>(positions), nullptr, forces,
                     energies, nullptr);
}

mafm_status mafm_system_step(mafm_ctx* c, const mafm_interaction_list* list, int32_t integrator, double dt,
                             int64_t n_states, double* positions, double* velocities, double* forces,
                             double* energies) {
    return system_op(c, list, SYS_STEP, nullptr, integrator, dt, n_states, positions, velocities, forces, energies,
                     nullptr);
}

mafm_status mafm_system_minimise(mafm_ctx* c, const mafm_interaction_list* list, const mafm_minimiser_opts* opts,
                                 int64_t n_states, double* positions, double* velocities, double* forces,
                                 double* energies, int32_t* steps) {
    return system_op(c, list, SYS_MINIMISE, opts, 0, 0.0, n_states, positions, velocities, forces, energies, steps);
}

mafm_status mafm_flex_state(mafm_ctx* c, int32_t first, int32_t count, double* out) {
    if (!c || (!out && count)) return MAFM_ERR_INVALID;
    if (c->fx_state_heights == 0) return fail(c, MAFM_ERR_STATE, "the last scan kept no flexible state");
    if (first < 0 || count < 0 || first + count > c->fx_state_heights)
        return fail(c, MAFM_ERR_INVALID, "heights [%d, %d) beyond the %d of the last scan", first, first + count,
                    c->fx_state_heights);
    if (count == 0) return MAFM_OK;
    const size_t per = 3 * (size_t)c->fx_state_atoms;
    HIPCHK(c, hipMemcpyAsync(out, c->d_fx_state.p + per * first, per * count * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MAFM_OK;
}

mafm_status mafm_margins(mafm_ctx* c, size_t first, size_t count, float* out) {
    if (!c || (!out && count)) return MAFM_ERR_INVALID;
    if (!c->have_margins) return fail(c, MAFM_ERR_STATE, "the last scan did not record decision margins");
    if (first + count > c->n_records)
        return fail(c, MAFM_ERR_INVALID, "margins [%zu, %zu) beyond the %zu of the last scan", first, first + count,
                    c->n_records);
    if (count == 0) return MAFM_OK;
    HIPCHK(c, hipMemcpyAsync(out, c->d_margins.p + first, count * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MAFM_OK;
}

mafm_status mafm_deltaf_weights(mafm_ctx* c, const mafm_deltaf_desc* d, int32_t nz, double zhigh, double dz,
                                mafm_deltaf_weights_out* out) {
    if (!d || !out) return fail(c, MAFM_ERR_INVALID, "null argument");
    char err[256];
    if (!deltaf_weights(*d, nz, zhigh, dz, *out, nullptr, err, sizeof err)) return fail(c, MAFM_ERR_INVALID, "%s", err);
    return MAFM_OK;
}

mafm_status mafm_delta_f(mafm_ctx* c, int system, const mafm_deltaf_desc* d, double* df, double* z_low) {
    if (!c || !d || !df) return MAFM_ERR_INVALID;
    if (!c->have_scan) return fail(c, MAFM_ERR_STATE, "frequency shift before a scan");
    const int n_systems = c->last_flex ? 1 : (int)c->mols.size();
    if (system < 0 || system >= n_systems) return fail(c, MAFM_ERR_INVALID, "system %d", system);
    const mafm_scan_desc& s = c->last_scan;
    mafm_deltaf_weights_out w;
    std::vector<double> zl;
    char err[256];
    if (!deltaf_weights(*d, s.nz, s.zhigh, s.dz, w, &zl, err, sizeof err)) return fail(c, MAFM_ERR_INVALID, "%s", err);
    if (w.n_images == 0) return fail(c, MAFM_ERR_INVALID, "%d heights do not fill a %d-height window", s.nz, w.n_weights);
    HIPCHK(c, hipSetDevice(c->device));
    const size_t n = (size_t)w.n_images * s.nx * s.ny;
    DevBuf<double> out;
    HIPCHK(c, out.reserve(n));
    const DevRecord* rec = c->d_rec.p + (c->last_flex ? 0 : c->mols[system].rec_off);
    HIPCHK(c, launch_delta_f(rec, w, s.nx, s.ny, s.nz, d->fz_from_text != 0, out.p, c->stream));
    HIPCHK(c, hipMemcpyAsync(df, out.p, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (z_low) memcpy(z_low, zl.data(), zl.size() * sizeof(double));
    return MAFM_OK;
}

mafm_status mafm_records(mafm_ctx* c, size_t first, size_t count, mafm_record* out) {
    if (!c || (!out && count)) return MAFM_ERR_INVALID;
    if (first + count > c->n_records)
        return fail(c, MAFM_ERR_INVALID, "records [%zu, %zu) beyond the %zu of the last scan", first, first + count,
                    c->n_records);
    static_assert(sizeof(mafm_record) == sizeof(DevRecord), "record layout");
    if (count == 0) return MAFM_OK;
    HIPCHK(c, hipMemcpyAsync(out, c->d_rec.p + first, count * sizeof(DevRecord), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MAFM_OK;
}

mafm_status mafm_eval_tip_surface(mafm_ctx* c, int system, size_t n, const double* pos, double* fe) {
    if (!c || (n && (!pos || !fe))) return MAFM_ERR_INVALID;
    if (!c->have_systems) return fail(c, MAFM_ERR_STATE, "eval before mafm_set_systems");
    if (system < 0 || system >= (int)c->mols.size()) return fail(c, MAFM_ERR_INVALID, "system %d", system);
    if (n == 0) return MAFM_OK;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, c->d_tmp.reserve(7 * n));
    double* dpos = c->d_tmp.p;
    double* dout = c->d_tmp.p + 3 * n;
    HIPCHK(c, hipMemcpyAsync(dpos, pos, 3 * n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(c->d_counters.p, 0, 8 * sizeof(int64_t), c->stream));
    ScanArgs a = base_args(c);
    HIPCHK(c, launch_eval_plain(a, system, (long)n, dpos, dout, c->stream));
    int64_t counters[8];
    HIPCHK(c, hipMemcpyAsync(fe, dout, 4 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(counters, c->d_counters.p, sizeof counters, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (counters[3] != 0) return fail(c, MAFM_ERR_OUT_OF_GRID, "tip position past the force grid");
    return MAFM_OK;
}

static mafm_status twiddles_for(mafm_ctx* c, int n, const double2** out) {
    auto it = c->twiddles.find(n);
    if (it == c->twiddles.end()) {
        std::vector<double> h;
        fft_twiddles(n, h);
        DevBuf<double2>& b = c->twiddles[n];
        hipError_t e = b.reserve(n);
        if (e == hipSuccess)
            e = hipMemcpyAsync(b.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) {
            c->twiddles.erase(n);
            return fail(c, MAFM_ERR_HIP, "twiddle table of length %d: %s", n, hipGetErrorString(e));
        }
        it = c->twiddles.find(n);
    }
    *out = it->second.p;
    return MAFM_OK;
}

static mafm_status check_fft_dims(mafm_ctx* c, const int32_t* n) {
    for (int a = 0; a < 3; ++a)
        if (!fft_length_supported(n[a]))
            return fail(c, MAFM_ERR_INVALID, "FFT length %d unsupported (1..2048, prime factors <= 13)", n[a]);
    return MAFM_OK;
}

mafm_status mafm_fft3d(mafm_ctx* c, const int32_t n[3], int inverse, const double* in, double* out) {
    if (!c || !n || !in || !out) return MAFM_ERR_INVALID;
    mafm_status s = check_fft_dims(c, n);
    if (s != MAFM_OK) return s;
    HIPCHK(c, hipSetDevice(c->device));
    const long N = (long)n[0] * n[1] * n[2];
    const double2* tw[3];
    for (int a = 0; a < 3; ++a)
        if ((s = twiddles_for(c, n[a], &tw[a])) != MAFM_OK) return s;
    DevBuf<double2> d;
    HIPCHK(c, d.reserve(N));
    HIPCHK(c, hipMemcpyAsync(d.p, in, N * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    int nn[3] = {n[0], n[1], n[2]};
    HIPCHK(c, fft3d_device(d.p, nn, tw, inverse != 0, c->stream));
    if (inverse) HIPCHK(c, launch_scale2(d.p, 1.0 / (double)N, N, c->stream));
    HIPCHK(c, hipMemcpyAsync(out, d.p, N * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    d.release();
    return MAFM_OK;
}

mafm_status mafm_potential_grid(mafm_ctx* c, const mafm_potential_desc* pd, int* handle) {
    if (!c || !pd || !handle || !pd->potential) return MAFM_ERR_INVALID;
    mafm_status s = check_fft_dims(c, pd->n);
    if (s != MAFM_OK) return s;
    if (!(pd->gaussian_width > 0)) return fail(c, MAFM_ERR_INVALID, "gaussian width must be positive");
    HIPCHK(c, hipSetDevice(c->device));
    const int n[3] = {pd->n[0], pd->n[1], pd->n[2]};
    const long N = (long)n[0] * n[1] * n[2];
    const double* B = pd->basis;   // row-major, columns = voxel vectors
    const bool ortho = is_diag(B);
    // cutoff of the Gaussian per axis (interactions.cpp:71-95)
    const double w = pd->gaussian_width, w2 = std::pow(w, 2);
    const double norm = 1.0 / std::pow(w * std::sqrt(2.0 * 3.14159265358979323846), 3);
    const int n_min = std::min(std::min(n[0], n[1]), n[2]);
    int cut[3];
    for (int a = 0; a < 3; ++a) {
        cut[a] = n_min;
        for (int i = 0; i < n_min - 1; ++i) {
            const double x = i * B[a], y = i * B[3 + a], z = i * B[6 + a];
            const double r2 = std::pow(x, 2) + std::pow(y, 2) + std::pow(z, 2);
            if (std::exp(-0.5 * r2 / w2) < 1.0e-10) {
                cut[a] = i;
                break;
            }
        }
    }
    // k-space basis (fft.cpp:41-47): (basis * diag(n))^-1, transposed
    double bn[9], inv[9], kb[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) bn[3 * i + j] = B[3 * i + j] * n[j];
    inv3(bn, inv);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) kb[3 * i + j] = inv[3 * j + i];
    const double vol = B[0] * (B[4] * B[8] - B[5] * B[7]) + B[1] * (B[5] * B[6] - B[3] * B[8]) +
                       B[2] * (B[3] * B[7] - B[4] * B[6]);
    const double2* tw[3];
    for (int a = 0; a < 3; ++a)
        if ((s = twiddles_for(c, n[a], &tw[a])) != MAFM_OK) return s;
    DevBuf<double2> pot, rho, tmp;
    DevBuf<double> real;
    HIPCHK(c, pot.reserve(N));
    HIPCHK(c, rho.reserve(N));
    HIPCHK(c, tmp.reserve(N));
    HIPCHK(c, real.reserve(N));
    HIPCHK(c, c->d_flag.reserve(1));
    HIPCHK(c, hipMemsetAsync(c->d_flag.p, 0, sizeof(int), c->stream));
    HIPCHK(c, hipMemcpyAsync(real.p, pd->potential, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK(c, launch_real_to_complex(real.p, pot.p, N, c->stream));
    HIPCHK(c, launch_zero2(rho.p, N, c->stream));
    HIPCHK(c, launch_gauss_tip(rho.p, n, cut, ortho, B, pd->tip_charge * norm, w2, c->stream));
    HIPCHK(c, fft3d_device(pot.p, n, tw, false, c->stream));
    HIPCHK(c, fft3d_device(rho.p, n, tw, false, c->stream));
    HIPCHK(c, launch_mul_vol(pot.p, rho.p, vol, N, c->stream));   // pot = E_k
    HostGrid h;
    mafm_grid_desc gd{};
    for (int a = 0; a < 3; ++a) {
        gd.n[a] = n[a];
        gd.offset[a] = pd->origin[a];
    }
    gd.periodic = 1;
    memcpy(gd.basis, B, sizeof gd.basis);
    GridGuard guard(h);
    s = new_grid(c, &gd, h);
    if (s != MAFM_OK) return s;
    double* fe = reinterpret_cast<double*>(h.fe64);
    const double inv_n = 1.0 / (double)N;
    HIPCHK(c, hipMemcpyAsync(tmp.p, pot.p, N * sizeof(double2), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, fft3d_device(tmp.p, n, tw, true, c->stream));
    HIPCHK(c, launch_extract(tmp.p, fe, 3, inv_n, N, c->d_flag.p, c->stream));
    for (int comp = 0; comp < 3; ++comp) {
        HIPCHK(c, launch_spectral_grad(pot.p, tmp.p, n, comp, kb, c->stream));
        HIPCHK(c, fft3d_device(tmp.p, n, tw, true, c->stream));
        HIPCHK(c, launch_extract(tmp.p, fe, comp, inv_n, N, c->d_flag.p, c->stream));
    }
    HIPCHK(c, launch_to_f32(h.fe64, h.fe32, N, c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    int flag = 0;
    HIPCHK(c, hipMemcpyAsync(&flag, c->d_flag.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    pot.release();
    rho.release();
    tmp.release();
    real.release();
    if (flag) {
        return fail(c, MAFM_ERR_FFT_IMAG, "Imaginary part of inverse FFT exceeded threshold 1.0e-8!");
    }
    *handle = store_grid(c, h);
    guard.kept = true;
    return sync_grids(c);
}

mafm_status mafm_adjoint(mafm_ctx* c, int system, const mafm_adjoint_desc* d, mafm_adjoint_out* o) {
    if (!c || !d || !o) return MAFM_ERR_INVALID;
    if (!c->have_systems) return fail(c, MAFM_ERR_STATE, "adjoint before mafm_set_systems");
    if (system < 0 || system >= (int)c->mols.size()) return fail(c, MAFM_ERR_INVALID, "system %d", system);
    const DevMol& m = c->mols[system];
    if (m.force_grid >= 0 || m.potential_grid >= 0)
        return fail(c, MAFM_ERR_INVALID, "the adjoint covers pair sites only (no force / potential grids)");
    if (m.n_seg[2] + m.n_seg[3] > 0) return fail(c, MAFM_ERR_INVALID, "the adjoint covers LJ (+ Coulomb) sites only");
    const long n = (long)d->n_positions;
    if (n < 0 || (n > 0 && (!d->tip || !d->dummy || !d->g))) return fail(c, MAFM_ERR_INVALID, "adjoint inputs");
    const int ns = m.n_sites;
    if (ns > 0 && (!o->d_es6 || !o->d_es12 || !o->d_qq || !o->d_pos)) return fail(c, MAFM_ERR_INVALID, "adjoint outputs");
    HIPCHK(c, hipSetDevice(c->device));
    DevBuf<double> buf;
    HIPCHK(c, buf.reserve((size_t)10 * n + 6 * (size_t)ns + 1));
    double* tip = buf.p;
    double* dummy = tip + 3 * n;
    double* g = dummy + 3 * n;
    double* v = g + n;
    double* out6 = v + 3 * n;
    HIPCHK(c, c->d_flag.reserve(1));
    HIPCHK(c, hipMemsetAsync(c->d_flag.p, 0, sizeof(int), c->stream));
    if (n) {
        HIPCHK(c, hipMemcpyAsync(tip, d->tip, 3 * n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(dummy, d->dummy, 3 * n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(g, d->g, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(c, hipMemsetAsync(out6, 0, 6 * (size_t)ns * sizeof(double), c->stream));
    ScanArgs a = base_args(c);
    if (n) HIPCHK(c, launch_adjoint(a, system, ns, n, tip, dummy, g, v, out6, c->d_flag.p,
                                    d->precision == MAFM_PREC_FAST32, c->stream));
    std::vector<double> h(6 * (size_t)ns);
    int singular = 0;
    if (ns) HIPCHK(c, hipMemcpyAsync(h.data(), out6, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(&singular, c->d_flag.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    buf.release();
    for (int i = 0; i < ns; ++i) {
        o->d_es6[i] = h[6 * i];
        o->d_es12[i] = h[6 * i + 1];
        o->d_qq[i] = h[6 * i + 2];
        o->d_pos[3 * i] = h[6 * i + 3];
        o->d_pos[3 * i + 1] = h[6 * i + 4];
        o->d_pos[3 * i + 2] = h[6 * i + 5];
    }
    if (singular) return fail(c, MAFM_ERR_INVALID, "singular tip-force Jacobian at a relaxed position");
    return MAFM_OK;
}

}  // extern "C"

namespace mafm {

void ctx_retain(mafm_ctx* c) {
    if (c) ++c->sims;
}

void ctx_release(mafm_ctx* c) {
    if (c && --c->sims == 0 && c->destroy_pending) destroy_now(c);
}

uint64_t ctx_scan_generation(const mafm_ctx* c) { return c ? c->scan_generation : 0; }

}  // namespace mafm
