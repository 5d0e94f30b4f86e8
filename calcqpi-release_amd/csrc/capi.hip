// C ABI of the continuum-LDOS library: context, GPUQPI drop-in entry points
// and the whole-chain device path.  See include/cqpi_amd.h for the contract.
#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "cqpi_common.h"
#include "mixed.h"

using namespace cqpi;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(CQPI_E_HIP, std::string(#call " failed: ") + hipGetErrorString(e_)); \
    } while (0)

#define RETCHK(call)                                                         \
    do {                                                                     \
        int r_ = (call);                                                     \
        if (r_) {                                                            \
            if (g_err.empty()) g_err = std::string(#call " failed");         \
            return r_;                                                       \
        }                                                                    \
    } while (0)

template <class T>
int dev_grow(T *&ptr, size_t &cap, size_t count, size_t &gpumem)
{
    if (count <= cap && ptr) return 0;
    if (ptr) {
        (void)hipFree(ptr);
        gpumem -= cap * sizeof(T);
        ptr = nullptr;
        cap = 0;
    }
    hipError_t e = hipMalloc((void **)&ptr, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) {
        ptr = nullptr;
        return fail(CQPI_E_HIP, std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) +
                                    " bytes failed: " + hipGetErrorString(e));
    }
    cap = std::max<size_t>(count, 1);
    gpumem += cap * sizeof(T);
    return 0;
}

bool is_supported_bands(size_t b)
{
    return (b >= 1 && b <= 12) || b == 14 || b == 16;
}

}  // namespace

namespace cqpi {
int set_error(int code, const std::string &msg) { return fail(code, msg); }
}  // namespace cqpi

// Fixed device scratch of cqpi_wannierldos_device, the entry point that
// allocates nothing: the periodic-term table, and per precision the fp16
// planes and their scales (mixed) or the FP32 copies of the windows and T
// (FP32).  Every device holds its own copy until the process ends.  Host code
// never passes the array's name to a kernel (on the host it is only a
// stand-in): it uses the device address hipGetSymbolAddress returned for the
// current device at context creation.
static __device__ __align__(256) unsigned char scratch[CQPI_NOALLOC_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;
std::map<int, void *> g_scratch_ptr;  // device -> address of its copy of scratch

// resolves (once per device, that device current) the address of scratch
int scratch_setup(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr.count(device)) return 0;
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(CQPI_E_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_ptr[device] = ptr;
    return 0;
}

void *scratch_of(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    return it == g_scratch_ptr.end() ? nullptr : it->second;
}

}  // namespace

// streams an eager run's energy chunk is spread over (3 and 4 measured
// slower than 2 on the 3-orbital workload: 0.77 and 0.81 ms against 0.755)
constexpr int kSplit = 2;

struct cqpi_ctx {
    Geom g{};
    size_t wanniern = 0, kpoints = 0;
    int precision = CQPI_PREC_FP64;
    int device = 0;
    hipStream_t stream = nullptr;
    // extra streams (created on first use): the parts of an energy chunk run
    // concurrently in eager runs; batch instances keep to one stream (the
    // batch graph has its own parallel branches)
    hipStream_t xstream[kSplit - 1] = {};
    hipEvent_t ev_fork = nullptr, ev_join[kSplit - 1] = {};
    bool single_stream = false;
    size_t gpumem = 0;

    // tiles
    double *d_tiles = nullptr;
    size_t c_tiles = 0;
    std::vector<double> h_tiles;

    // model
    bool has_model = false;
    DevRGroup *d_groups = nullptr;
    size_t c_groups = 0;
    double2 *d_tmat = nullptr;  // [group][o1][o2] summed amplitudes
    size_t c_tmat = 0;
    int ngroups = 0;
    double efermi = 0.0;
    int scmax = 0;
    // H(k)^T = H(-k) for the model (cqpi_set_model), and whether the
    // transform passes may use it (cqpi_set_pair_symmetry)
    bool tsym_model = false;
    bool tsym_on = true;
    bool tsym_used = false;  // the last prepared run used it
    double *d_eval = nullptr;
    size_t c_eval = 0;
    double2 *d_evec = nullptr;
    size_t c_evec = 0;
    bool eigen_valid = false;

    // impurity
    double2 *d_V = nullptr;
    size_t c_V = 0;

    // transform
    double2 *d_tw = nullptr;
    size_t c_tw = 0;
    float2 *d_tw32 = nullptr;  // the same table rounded to FP32 (reduced-precision transform)
    size_t c_tw32 = 0;

    // per-chunk buffers
    double2 *d_F1 = nullptr;
    size_t c_F1 = 0;
    double2 *d_gkw = nullptr;  // G(k) workspace of the many-band resolvent path
    size_t c_gkw = 0;
    double2 *d_G0w = nullptr;
    size_t c_G0w = 0;
    double *d_bulk = nullptr;
    size_t c_bulk = 0;
    // per-run buffers
    double2 *d_T = nullptr;
    size_t c_T = 0;
    double2 *d_tips = nullptr;  // tip factor of each energy of a Josephson map
    size_t c_tips = 0;
    // spectral-function context (cqpi_create_spectral): positions and
    // prefactors of the maxband orbitals, G0(k) of the drop-in calls, the
    // map cqpi_run makes (0: none, 1: SPF, 2: USPF), DOS buffers
    bool spectral = false;
    int spec_mode = 0;
    double *d_spos = nullptr;
    double2 *d_g0k = nullptr, *h_g0k = nullptr;
    size_t c_g0k = 0, c_hg0k = 0, c_spos = 0;
    double *d_dos = nullptr, *d_dosp = nullptr, *d_domega = nullptr;
    size_t c_dos = 0, c_dosp = 0, c_domega = 0;
    // threshold list (cqpi_create_list): per-pair factor table and pair index
    bool list_mode = false;
    double *d_lcoef = nullptr, *d_lccoef = nullptr;
    int *d_lr1off = nullptr, *d_lpair = nullptr, *d_lcoff = nullptr, *d_lco2 = nullptr;
    int lnact = 0;
    // surface / bulk Green's function of a layered model (cqpi_set_surface):
    // 0 off (G0(k) from the eigen data), 1 surface layer, 2 bulk.  Hopping
    // groups of the bulk in-plane, surface in-plane and coupling blocks;
    // d_sstat: decimation steps per energy (surf_maxit + 1: not converged)
    int surf_mode = 0;
    double surf_eps = 1e-5;
    int surf_maxit = 500;
    std::vector<cqpi_hopping> h_shops;
    DevRGroup *d_sgroups = nullptr;
    double2 *d_stmat = nullptr;
    size_t c_sgroups = 0, c_stmat = 0;
    int sg_n[3] = {0, 0, 0};
    size_t sg_goff[3] = {0, 0, 0}, sg_toff[3] = {0, 0, 0};
    int *d_sstat = nullptr;
    size_t c_sstat = 0;
    int *d_status = nullptr;
    size_t c_status = 0;
    double *d_omega = nullptr;
    size_t c_omega = 0;
    double *d_ldos = nullptr;
    size_t c_ldos = 0;
    size_t nlayers = 0;
    size_t last_chunk_e0 = 0, last_chunk_ec = 0;

    // FP32 contraction: tiles, narrowed windows and T
    float *d_tilesf = nullptr;
    size_t c_tilesf = 0;
    float2 *d_G0f = nullptr, *d_Tf = nullptr;
    size_t c_G0f = 0, c_Tf = 0;

    // mixed precision: tile fragments, per-energy scales, fp16 planes
    MixGeom mg{};
    bool mix_ready = false;
    double sA = 1.0;
    _Float16 *d_afrag = nullptr;
    size_t c_afrag = 0;
    double *d_mscale = nullptr, *d_cscale = nullptr;
    size_t c_mscale = 0, c_cscale = 0;
    _Float16 *d_planes = nullptr;
    size_t c_planes = 0;

    // gradient: kept forward state and adjoint buffers
    std::vector<cqpi_hopping> h_hops;  // model as given to cqpi_set_model
    bool keep_windows = false, windows_kept = false;
    // Josephson maps from cqpi_run (tip gap and broadening; etat < 0: the run's eta)
    bool josephson = false, maps_josephson = false;
    double jos_deltat = 0.001, jos_etat = -1.0;
    std::vector<double2> up_tips;
    bool window_whole = true;  // the last run wrote the whole G0 window
    std::vector<double> last_omegas;
    double last_eta = 0.0;
    double2 *d_twn = nullptr;  // exp(-2 pi i m / K)
    size_t c_twn = 0;
    double2 *d_abar = nullptr, *d_bbar = nullptr, *d_Tpart = nullptr, *d_Tbar = nullptr, *d_Gbar = nullptr;
    double2 *d_Vbar = nullptr, *d_Hacc = nullptr, *d_Rlist = nullptr, *d_SR = nullptr;
    size_t c_abar = 0, c_bbar = 0, c_Tpart = 0, c_Tbar = 0, c_Gbar = 0, c_Vbar = 0, c_Hacc = 0, c_Rlist = 0, c_SR = 0;
    double *d_rhobar = nullptr, *d_rhosum = nullptr, *d_ry = nullptr, *d_Q = nullptr;
    size_t c_rhobar = 0, c_rhosum = 0, c_ry = 0, c_Q = 0;
    double2 *d_U = nullptr;
    size_t c_U = 0;
    // last uploaded hopping vectors and y offsets, and the pinned read-back
    // buffer of the adjoint's sums
    std::vector<double2> up_R;
    std::vector<double> up_ry;
    double2 *h_grad = nullptr;
    size_t c_hgrad = 0;

    // uploads of the last prepared run (skipped when unchanged) and the pinned
    // status read-back buffer
    std::vector<double> up_omegas, up_scale;
    int *h_status = nullptr;
    size_t c_hstatus = 0;

    // host staging for the drop-in path (pinned)
    double2 *h_win = nullptr;
    size_t c_hwin = 0;
    double2 *h_T = nullptr;
    hipEvent_t copy_done = nullptr;

    // timing
    bool timing = false;
    bool eigen_timed = false;  // an explicit cqpi_compute_eigen opened the current timing window
    bool in_run = false;       // cqpi_compute_eigen called from inside cqpi_run
    double stage_ms[CQPI_NSTAGES] = {0};
    int64_t stage_n[CQPI_NSTAGES] = {0};
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
};

namespace {

void release(cqpi_ctx *c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    void *bufs[] = {c->d_tiles, c->d_groups, c->d_tmat, c->d_eval, c->d_evec, c->d_V, c->d_tw, c->d_tw32,
                    c->d_F1, c->d_gkw, c->d_G0w, c->d_bulk, c->d_T, c->d_tips, c->d_lcoef, c->d_lccoef,
                    c->d_lcoff, c->d_lco2, c->d_lr1off,
                    c->d_lpair, c->d_spos, c->d_g0k, c->d_dos, c->d_dosp, c->d_domega,
                    c->d_sgroups, c->d_stmat, c->d_sstat, c->d_status, c->d_omega, c->d_ldos,
                    c->d_tilesf, c->d_G0f, c->d_Tf, c->d_afrag, c->d_mscale, c->d_cscale, c->d_planes, c->d_twn, c->d_abar, c->d_bbar, c->d_Tpart,
                    c->d_Tbar, c->d_Gbar, c->d_Vbar, c->d_Hacc, c->d_Rlist, c->d_SR, c->d_rhobar,
                    c->d_rhosum, c->d_ry, c->d_U, c->d_Q};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    if (c->h_win) (void)hipHostFree(c->h_win);
    if (c->h_T) (void)hipHostFree(c->h_T);
    if (c->h_g0k) (void)hipHostFree(c->h_g0k);
    if (c->h_status) (void)hipHostFree(c->h_status);
    if (c->h_grad) (void)hipHostFree(c->h_grad);
    for (auto &p : c->pending) {
        (void)hipEventDestroy(p.second.first);
        (void)hipEventDestroy(p.second.second);
    }
    if (c->copy_done) (void)hipEventDestroy(c->copy_done);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    for (int i = 0; i < kSplit - 1; i++) {
        if (c->xstream[i]) (void)hipStreamDestroy(c->xstream[i]);
        if (c->ev_join[i]) (void)hipEventDestroy(c->ev_join[i]);
    }
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    delete c;
}

// stage timing: begin returns an index into pending, end records the stop event
int stage_begin(cqpi_ctx *c, int stage)
{
    if (!c->timing) return -1;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return -1;
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return -1;
    }
    (void)hipEventRecord(a, c->stream);
    c->pending.push_back({stage, {a, b}});
    return (int)c->pending.size() - 1;
}

void stage_end(cqpi_ctx *c, int idx)
{
    if (idx < 0) return;
    (void)hipEventRecord(c->pending[idx].second.second, c->stream);
}

void stage_collect(cqpi_ctx *c)
{
    for (auto &p : c->pending) {
        float ms = 0.f;
        if (hipEventSynchronize(p.second.second) == hipSuccess &&
            hipEventElapsedTime(&ms, p.second.first, p.second.second) == hipSuccess) {
            c->stage_ms[p.first] += ms;
            c->stage_n[p.first] += 1;
        }
        (void)hipEventDestroy(p.second.first);
        (void)hipEventDestroy(p.second.second);
    }
    c->pending.clear();
}

void reset_stages(cqpi_ctx *c)
{
    for (int s = 0; s < CQPI_NSTAGES; s++) {
        c->stage_ms[s] = 0.0;
        c->stage_n[s] = 0;
    }
}

int upload_identity_V(cqpi_ctx *c)
{
    int B = c->g.B;
    std::vector<double2> V((size_t)B * B, make_double2(0.0, 0.0));
    for (int i = 0; i < B; i++) V[(size_t)i * B + i] = make_double2(1.0, 0.0);
    RETCHK(dev_grow(c->d_V, c->c_V, V.size(), c->gpumem));
    HIPCHK(hipMemcpy(c->d_V, V.data(), V.size() * sizeof(double2), hipMemcpyHostToDevice));
    return 0;
}

// fp16 planes and per-energy scales for the mixed contraction of E energies
int mix_buffers(cqpi_ctx *c, int E)
{
    RETCHK(dev_grow(c->d_mscale, c->c_mscale, (size_t)E, c->gpumem));
    RETCHK(dev_grow(c->d_cscale, c->c_cscale, (size_t)E, c->gpumem));
    RETCHK(dev_grow(c->d_planes, c->c_planes, (size_t)E * mix_plane_halves(c->g, c->mg), c->gpumem));
    return 0;
}

// bulk table + contraction for E energies whose G0 window, T are on the
// device.  planes_ready: the transform already wrote the mixed-precision
// planes and scales (whole-chain path); otherwise they are made from G0w.
int contract(cqpi_ctx *c, const double2 *G0w, const double2 *T, int E, double *ldos, bool planes_ready,
             size_t eo = 0, hipStream_t s = nullptr, const double2 *tips = nullptr, bool bulk_ready = false)
{
    // eo: offset of these energies in the chunk buffers (periodic term, planes,
    // scales, FP32 copies); s: stream (default: the context stream); tips:
    // device tip factors of the E energies for the Josephson map (NULL: LDOS)
    const Geom &g = c->g;
    if (!s) s = c->stream;
    const int NP = g.B * g.B;
    if (tips && 2 * g.maxb > g.B) return fail(CQPI_E_ARG, "the Josephson map needs 2*maxband <= bands");
    const size_t per = tips ? 2 : 1;  // doubles per periodic-table entry
    RETCHK(dev_grow(c->d_bulk, c->c_bulk, per * (eo + E) * g.os * g.os, c->gpumem));
    double *bulk = c->d_bulk + per * eo * g.os * g.os;
    if (c->list_mode) {
        // the periodic part (FP64) and the map from the threshold list, in
        // FP64 or on windows and T narrowed to FP32
        float2 *G0f = nullptr, *Tf = nullptr;
        if (c->precision == CQPI_PREC_FP32) {
            RETCHK(dev_grow(c->d_G0f, c->c_G0f, (eo + E) * g.nw * g.nw * NP, c->gpumem));
            RETCHK(dev_grow(c->d_Tf, c->c_Tf, (eo + E) * NP, c->gpumem));
            G0f = c->d_G0f + eo * g.nw * g.nw * NP;
            Tf = c->d_Tf + eo * NP;
        }
        int st = stage_begin(c, 4);
        int r = launch_contract_list(g, G0w, T, c->d_lcoef, c->d_lcoff, c->d_lco2, c->d_lccoef, c->d_lr1off,
                                     c->d_lpair, c->lnact, E, bulk, ldos, s,
                                     G0f, Tf);
        stage_end(c, st);
        if (r) return fail(r, "threshold-list contraction launch failed");
        return 0;
    }
    int st = 0;
    if (!bulk_ready) {
        st = stage_begin(c, 3);
        if (tips)
            RETCHK(launch_bulk_josephson(g, G0w, c->d_tiles, E, reinterpret_cast<double2 *>(bulk), s));
        else
            RETCHK(launch_bulk(g, G0w, c->d_tiles, E, bulk, s));
        stage_end(c, st);
    }
    st = stage_begin(c, 4);
    int r = 0;
    switch (c->precision) {
    case CQPI_PREC_MIXED: {
        const bool tf = !tips && mix_tfold(g, c->mg);
        if (!planes_ready) {
            RETCHK(mix_buffers(c, (int)(eo + E)));
            r = launch_mix_prepare(g, c->mg, G0w, E, c->d_mscale + eo,
                                   c->d_planes + eo * mix_plane_halves(g, c->mg), s, tf);
            c->up_scale.clear();  // the device scales now differ from the last upload
        }
        if (!r)
            r = launch_contract_mixed(g, c->mg, c->d_planes + eo * mix_plane_halves(g, c->mg), c->d_afrag,
                                      c->d_mscale + eo, c->sA, T, bulk, E, ldos, s, tips, tf, c->d_cscale + eo);
        break;
    }
    case CQPI_PREC_FP32:
        RETCHK(dev_grow(c->d_G0f, c->c_G0f, (eo + E) * g.nw * g.nw * NP, c->gpumem));
        RETCHK(dev_grow(c->d_Tf, c->c_Tf, (eo + E) * NP, c->gpumem));
        r = launch_contract_fp32(g, G0w, T, bulk, c->d_tilesf, E, ldos, c->d_G0f + eo * g.nw * g.nw * NP,
                                 c->d_Tf + eo * NP, s, tips);
        break;
    default:
        r = launch_contract_fp64(g, G0w, T, bulk, c->d_tiles, E, ldos, s, tips);
        break;
    }
    stage_end(c, st);
    if (r) return fail(r, "contraction launch failed");
    return 0;
}

// energies per chunk and mixed-precision plane scale of one cqpi_run
struct RunPlan {
    size_t E = 0, Ec = 0;
    double eta = 0.0, sG = 1.0;
    bool mixed = false;
    size_t work_bytes = 0;  // G(k) workspace of the resolvent pass (0: not used)
    int rwin = 0;           // half-width of the G0 window part that is written
    bool tsym = false;      // the transform passes use G0(R)_cr = G0(-R)_rc
};

// allocations and uploads of a run (synchronous); no kernel launch
int prepare_run(cqpi_ctx *c, const double *omegas, size_t E, double eta, RunPlan &plan)
{
    const Geom &g = c->g;
    const int NP = g.B * g.B;
    size_t WW = (size_t)g.W * g.W;
    RETCHK(dev_grow(c->d_ldos, c->c_ldos, E * WW, c->gpumem));
    RETCHK(dev_grow(c->d_T, c->c_T, E * NP, c->gpumem));
    RETCHK(dev_grow(c->d_status, c->c_status, E, c->gpumem));
    const bool grew = c->c_omega < E || !c->d_omega;
    RETCHK(dev_grow(c->d_omega, c->c_omega, E, c->gpumem));
    // energies: uploaded from a member copy (it outlives the asynchronous
    // copy), and only when they changed
    if (grew || c->up_omegas.size() != E || !std::equal(omegas, omegas + E, c->up_omegas.begin())) {
        c->up_omegas.assign(omegas, omegas + E);
        HIPCHK(hipMemcpyAsync(c->d_omega, c->up_omegas.data(), E * sizeof(double), hipMemcpyHostToDevice,
                              c->stream));
    }
    // energies per chunk: intermediate of the two-pass transform kept under ~2 GiB
    size_t per_e = (size_t)g.K * g.nw * NP * sizeof(double2);
    size_t Ec = std::max<size_t>(1, std::min<size_t>(E, ((size_t)2 << 30) / std::max<size_t>(per_e, 1)));
    if (c->keep_windows) {
        // one chunk: the window and T of every energy stay for cqpi_gradient
        if (per_e * E > ((size_t)64 << 30))
            return fail(CQPI_E_ARG, "keeping the windows of this many energies needs more than 64 GiB");
        Ec = E;
    }
    c->windows_kept = false;
    RETCHK(dev_grow(c->d_F1, c->c_F1, Ec * (size_t)g.K * g.nw * NP, c->gpumem));
    RETCHK(dev_grow(c->d_G0w, c->c_G0w, Ec * (size_t)g.nw * g.nw * NP, c->gpumem));
    RETCHK(dev_grow(c->d_bulk, c->c_bulk, (c->josephson ? 2 : 1) * Ec * g.os * g.os, c->gpumem));
    if (c->josephson) {
        // tip factor of each energy (src/calcqpi.cc:705-709):
        // i (-sign(w) pi deltat) / sqrt((w + i etat)^2 - deltat^2), principal root
        const double dt = c->jos_deltat, et = c->jos_etat < 0.0 ? eta : c->jos_etat;
        std::vector<double2> tips(E);
        for (size_t e = 0; e < E; e++) {
            const std::complex<double> oc(omegas[e], et);
            const std::complex<double> t =
                std::complex<double>(0.0, -std::copysign(1.0, omegas[e]) * M_PI * dt) / std::sqrt(oc * oc - dt * dt);
            tips[e] = make_double2(t.real(), t.imag());
        }
        const bool tgrew = c->c_tips < E || !c->d_tips;
        RETCHK(dev_grow(c->d_tips, c->c_tips, E, c->gpumem));
        if (tgrew || c->up_tips.size() != E ||
            !std::equal(tips.begin(), tips.end(), c->up_tips.begin(),
                        [](double2 a, double2 b) { return a.x == b.x && a.y == b.y; })) {
            c->up_tips = tips;
            HIPCHK(hipMemcpyAsync(c->d_tips, c->up_tips.data(), E * sizeof(double2), hipMemcpyHostToDevice,
                                  c->stream));
        }
    }
    c->maps_josephson = c->josephson;
    // mixed precision needs the FP64 window only for |R| <= 2w (T-matrix and
    // bulk term) unless it is kept for the gradient
    plan.rwin = (c->precision == CQPI_PREC_MIXED && !c->keep_windows) ? std::min(g.h, 2 * g.w) : g.h;
    c->window_whole = plan.rwin >= g.h;
    plan.work_bytes = resolvent_work_bytes(g, (int)Ec, c->precision != CQPI_PREC_FP64 && transform32_supported(g));
    if (c->surf_mode) {
        // G(k) of the decimation: a block of energies at a time (under ~2 GiB)
        const size_t cs =
            c->precision != CQPI_PREC_FP64 && transform32_supported(g) ? sizeof(float2) : sizeof(double2);
        const size_t per_gk = (size_t)NP * g.K * g.K * cs;
        plan.work_bytes = std::max<size_t>(1, std::min<size_t>(Ec, ((size_t)2 << 30) / per_gk)) * per_gk;
        RETCHK(dev_grow(c->d_sstat, c->c_sstat, E, c->gpumem));
    }
    if (plan.work_bytes)
        RETCHK(dev_grow(c->d_gkw, c->c_gkw, (plan.work_bytes + sizeof(double2) - 1) / sizeof(double2), c->gpumem));
    // mixed precision: the k->x transform writes the fp16 planes directly.  Their
    // scale comes from |G0(R)_ab| <= ||G0(k)|| <= 1/eta, a bound for every R.
    plan.mixed = c->precision == CQPI_PREC_MIXED;
    plan.sG = 1.0;
    if (plan.mixed) {
        const bool sgrew = c->c_mscale < Ec || !c->d_mscale;
        RETCHK(mix_buffers(c, (int)Ec));
        int ex;
        std::frexp(1.0 / eta, &ex);
        plan.sG = std::ldexp(1.0, 13 - ex);
        if (sgrew || c->up_scale.size() != Ec || c->up_scale[0] != plan.sG) {
            c->up_scale.assign(Ec, plan.sG);
            HIPCHK(hipMemcpyAsync(c->d_mscale, c->up_scale.data(), Ec * sizeof(double), hipMemcpyHostToDevice,
                                  c->stream));
        }
    }
    // the transform passes carry only the pairs r <= c where the model has
    // H(k)^T = H(-k): on the Stockham path, with G0(k) from the eigen data (not
    // the decimation workspace) and no fp16 planes to write (their writer in
    // the kx pass has no mirrored copies; packing the planes from a whole
    // FP64 window instead cost more than the symmetry saved: mixed Josephson
    // map of 6 bands 5.3 -> 7.6 ms)
    const bool planes_written = plan.mixed && (c->josephson || !mix_tfold(g, c->mg));
    plan.tsym = c->tsym_model && c->tsym_on && !c->surf_mode && g.B > 1 && transform32_supported(g) &&
                !planes_written;
    c->tsym_used = plan.tsym;
    if (c->precision == CQPI_PREC_FP32) {
        RETCHK(dev_grow(c->d_G0f, c->c_G0f, Ec * g.nw * g.nw * NP, c->gpumem));
        RETCHK(dev_grow(c->d_Tf, c->c_Tf, Ec * NP, c->gpumem));
    }
    plan.E = E;
    plan.Ec = Ec;
    plan.eta = eta;
    c->last_omegas.assign(omegas, omegas + E);
    c->last_eta = eta;
    return 0;
}

// device work of a prepared run on c->stream: resolvent, transforms, T-matrix
// and contraction of every energy chunk.  Launches only (no allocation, copy
// or synchronisation), so it can be captured into a graph.
// energies [e0, e0 + ec) of the current chunk, at offset eo in the chunk
// buffers, on stream s
int enqueue_range(cqpi_ctx *c, const RunPlan &plan, size_t e0, int ec, size_t eo, hipStream_t s)
{
    const Geom &g = c->g;
    // the transform passes' geometry: pairs r <= c only where the model allows
    // it (prepare_run)
    Geom gt = g;
    gt.ts = plan.tsym ? 1 : 0;
    const int NP = g.B * g.B;
    const size_t WW = (size_t)g.W * g.W;
    const size_t f1e = (size_t)g.K * g.nw * NP, we = (size_t)g.nw * g.nw * NP;
    double2 *G0w = c->d_G0w + eo * we;
    _Float16 *planes = plan.mixed ? c->d_planes + eo * mix_plane_halves(g, c->mg) : nullptr;
    // reduced-precision modes run the resolvent and both transform passes in
    // FP32 where the Stockham path applies (F1 then holds float2)
    const bool t32 = c->precision != CQPI_PREC_FP64 && transform32_supported(g);
    // the T-folded contraction reads the FP32 window, which the FP32 kx pass
    // writes into the plane buffer; after an FP64 kx pass the contraction
    // packs it from the FP64 window itself
    const bool tf = plan.mixed && !c->josephson && mix_tfold(g, c->mg);
    const bool planes_ready = plan.mixed && (!tf || t32);
    float2 *F1f = reinterpret_cast<float2 *>(c->d_F1) + eo * f1e;
    double2 *F1d = c->d_F1 + eo * f1e;
    int st = stage_begin(c, 1);
    if (c->surf_mode) {
        // G(k) by decimation into the workspace, then the ky pass from it
        const size_t per_gk = (size_t)NP * g.K * g.K * (t32 ? sizeof(float2) : sizeof(double2));
        const int EBa = (int)std::max<size_t>(1, std::min<size_t>((size_t)ec, plan.work_bytes / per_gk));
        const DevRGroup *sg[3];
        const double2 *stm[3];
        for (int q = 0; q < 3; q++) {
            sg[q] = c->d_sgroups + c->sg_goff[q];
            stm[q] = c->d_stmat + c->sg_toff[q];
        }
        for (int ea = 0; ea < ec; ea += EBa) {
            const int eb = std::min(EBa, ec - ea);
            const double *om = c->d_omega + e0 + ea;
            RETCHK(launch_surface_gk(g.B, g.K, sg, stm, c->sg_n, c->d_tw, c->efermi, c->scmax, om, eb, plan.eta,
                                     c->surf_eps, c->surf_maxit, c->surf_mode == CQPI_GREEN_SURFACE, c->d_gkw, t32,
                                     c->d_sstat + e0 + ea, s));
            if (t32)
                RETCHK(launch_resolvent_ky_gk(g, c->d_gkw, om, eb, c->d_tw32, F1f + (size_t)ea * f1e, true, s));
            else
                RETCHK(launch_resolvent_ky_gk(g, c->d_gkw, om, eb, c->d_tw, F1d + (size_t)ea * f1e, false, s));
        }
    } else if (t32)
        RETCHK(launch_resolvent_ky32(gt, c->d_eval, c->d_evec, c->d_omega + e0, ec, plan.eta, c->d_tw32, F1f,
                                     c->d_gkw, plan.work_bytes, s));
    else
        RETCHK(launch_resolvent_ky(gt, c->d_eval, c->d_evec, c->d_omega + e0, ec, plan.eta, c->d_tw, F1d, c->d_gkw,
                                   plan.work_bytes, s));
    stage_end(c, st);
    st = stage_begin(c, 2);
    if (t32)
        RETCHK(launch_transform_kx32(gt, plan.mixed ? &c->mg : nullptr, F1f, ec, c->d_tw32, plan.sG, G0w, planes,
                                     plan.rwin, s, tf ? reinterpret_cast<float2 *>(planes) : nullptr));
    else
        RETCHK(launch_transform_kx(gt, planes_ready ? &c->mg : nullptr, F1d, ec, c->d_tw, plan.sG, G0w, planes, s));
    stage_end(c, st);
    st = stage_begin(c, 3);
    // the T-matrix and (but for the threshold list, whose kernel forms it) the
    // periodic term in one launch
    const bool tb = !c->list_mode;
    if (tb)
        RETCHK(launch_tmatrix_bulk(g, G0w, c->d_V, ec, c->d_T + e0 * NP, c->d_status + e0, c->d_tiles,
                                   c->d_bulk + (c->josephson ? 2 : 1) * eo * g.os * g.os, c->josephson, s));
    else
        RETCHK(launch_tmatrix(g, G0w, c->d_V, ec, c->d_T + e0 * NP, c->d_status + e0, s));
    stage_end(c, st);
    RETCHK(contract(c, G0w, c->d_T + e0 * NP, ec, c->d_ldos + e0 * WW, planes_ready, eo, s,
                    c->josephson ? c->d_tips + e0 : nullptr, tb));
    return 0;
}

// device work of a prepared run: resolvent, transforms, T-matrix and
// contraction of every energy chunk.  In eager runs a chunk of 8 or more
// energies is split in two halves that run concurrently on the context's two
// streams (a latency-bound transform pass of one half overlaps the contraction
// of the other); with stage timing on, and for batch instances, everything
// runs in order on the context stream.  Launches only (no
// allocation, copy or synchronisation, apart from creating the extra streams
// on first use outside a capture), so it can be captured into a graph.
int enqueue_chunks(cqpi_ctx *c, const RunPlan &plan)
{
    const size_t E = plan.E, Ec = plan.Ec;
    for (size_t e0 = 0; e0 < E; e0 += Ec) {
        const int ec = (int)std::min(Ec, E - e0);
        // parts of at least 4 energies (the G(k) workspace path has one
        // workspace: no split)
        const int parts = std::min(kSplit, ec / 4);
        const bool split = parts >= 2 && !c->timing && !c->single_stream && !plan.work_bytes && c->g.streams != 1;
        if (split && !c->ev_fork) {
            bool ok = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) == hipSuccess;
            for (int i = 0; ok && i < kSplit - 1; i++)
                ok = hipStreamCreate(&c->xstream[i]) == hipSuccess &&
                     hipEventCreateWithFlags(&c->ev_join[i], hipEventDisableTiming) == hipSuccess;
            if (!ok) return fail(CQPI_E_HIP, "extra streams/events could not be created");
        }
        if (split) {
            HIPCHK(hipEventRecord(c->ev_fork, c->stream));
            int off = 0;
            for (int p = 0; p < parts; p++) {
                const int cnt = ec / parts + (p < ec % parts ? 1 : 0);
                hipStream_t sp = p == 0 ? c->stream : c->xstream[p - 1];
                if (p > 0) HIPCHK(hipStreamWaitEvent(sp, c->ev_fork, 0));
                RETCHK(enqueue_range(c, plan, e0 + off, cnt, (size_t)off, sp));
                off += cnt;
            }
            for (int p = 1; p < parts; p++) {
                HIPCHK(hipEventRecord(c->ev_join[p - 1], c->xstream[p - 1]));
                HIPCHK(hipStreamWaitEvent(c->stream, c->ev_join[p - 1], 0));
            }
        } else {
            RETCHK(enqueue_range(c, plan, e0, ec, 0, c->stream));
        }
        c->last_chunk_e0 = e0;
        c->last_chunk_ec = ec;
    }
    return 0;
}

}  // namespace

static int check_surface_steps(cqpi_ctx *c, size_t E);

namespace {

// waits for the run and reports a singular 1 - V G0(0)
int finish_run(cqpi_ctx *c, const RunPlan &plan)
{
    const size_t E = plan.E;
    c->nlayers = E;
    c->windows_kept = c->keep_windows;
    if (c->c_hstatus < E) {
        if (c->h_status) (void)hipHostFree(c->h_status);
        c->h_status = nullptr;
        c->c_hstatus = 0;
        HIPCHK(hipHostMalloc((void **)&c->h_status, E * sizeof(int), hipHostMallocDefault));
        c->c_hstatus = E;
    }
    HIPCHK(hipMemcpyAsync(c->h_status, c->d_status, E * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (size_t e = 0; e < E; e++)
        if (c->h_status[e])
            return fail(CQPI_E_SINGULAR, "1 - V G0(0) is singular at energy index " + std::to_string(e));
    if (c->surf_mode) RETCHK(check_surface_steps(c, E));
    return 0;
}

}  // namespace

// One energy of the drop-in path: the G0(R) window and T handed over by the
// caller, then the LDOS map (tip == NULL) or the Josephson map with the tip
// factor tip[0] + i tip[1].
static int dropin_map(cqpi_ctx *c, const double *scat, const double *const *g0, size_t tda, const double *tip)
{
    g_err.clear();
    if (!c || !scat || !g0) return fail(CQPI_E_ARG, "NULL argument");
    const Geom &g = c->g;
    const int B = g.B, NP = B * B, K = g.K;
    if (tda < (size_t)B) return fail(CQPI_E_ARG, "tda < bands");
    size_t nwin = (size_t)g.nw * g.nw * NP;
    // previous work must be done before the pinned staging and window buffer are reused
    HIPCHK(hipStreamSynchronize(c->stream));
    reset_stages(c);
    if (nwin > c->c_hwin) {
        if (c->h_win) (void)hipHostFree(c->h_win);
        c->h_win = nullptr;
        c->c_hwin = 0;
        HIPCHK(hipHostMalloc((void **)&c->h_win, nwin * sizeof(double2)));
        c->c_hwin = nwin;
    }
    // T (up to 16 x 16) and, after it, the tip factor
    if (!c->h_T) HIPCHK(hipHostMalloc((void **)&c->h_T, (16 * 16 + 1) * sizeof(double2)));
    // window of G0(R), R = slot - h wrapped onto the k-grid (src/qpihipkernel.cc:383-393)
    for (int sx = 0; sx < g.nw; sx++) {
        int ix = ((sx - g.h) % K + K) % K;
        for (int sy = 0; sy < g.nw; sy++) {
            int iy = ((sy - g.h) % K + K) % K;
            const double *m = g0[(size_t)ix * K + iy];
            if (!m) return fail(CQPI_E_ARG, "NULL matrix in g0 table");
            double2 *dst = c->h_win + ((size_t)sx * g.nw + sy) * NP;
            for (int r = 0; r < B; r++)
                for (int cc = 0; cc < B; cc++)
                    dst[r * B + cc] = make_double2(m[2 * (r * tda + cc)], m[2 * (r * tda + cc) + 1]);
        }
    }
    for (int i = 0; i < NP; i++) c->h_T[i] = make_double2(scat[2 * i], scat[2 * i + 1]);
    if (tip) {
        c->h_T[16 * 16] = make_double2(tip[0], tip[1]);
        RETCHK(dev_grow(c->d_tips, c->c_tips, 1, c->gpumem));
    }
    RETCHK(dev_grow(c->d_G0w, c->c_G0w, nwin, c->gpumem));
    RETCHK(dev_grow(c->d_T, c->c_T, (size_t)NP, c->gpumem));
    RETCHK(dev_grow(c->d_ldos, c->c_ldos, (size_t)g.W * g.W, c->gpumem));
    HIPCHK(hipMemcpyAsync(c->d_G0w, c->h_win, nwin * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_T, c->h_T, NP * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    if (tip)
        HIPCHK(hipMemcpyAsync(c->d_tips, c->h_T + 16 * 16, sizeof(double2), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipEventRecord(c->copy_done, c->stream));
    RETCHK(contract(c, c->d_G0w, c->d_T, 1, c->d_ldos, false, 0, nullptr, tip ? c->d_tips : nullptr));
    c->nlayers = 1;
    c->windows_kept = false;
    c->window_whole = true;
    c->last_chunk_e0 = 0;
    c->last_chunk_ec = 1;
    // the caller may overwrite g0 and scat once the copies have landed
    HIPCHK(hipEventSynchronize(c->copy_done));
    return 0;
}

extern "C" {

const char *cqpi_last_error(void) { return g_err.c_str(); }

const char *cqpi_version(void) { return "cqpi_amd 0.1 (gfx950)"; }

int cqpi_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int cqpi_create(size_t wanniern, size_t kpoints, size_t n, size_t window, size_t oversamp, size_t bands,
                size_t maxband, int spin, const double *wf, cqpi_ctx **out)
{
    g_err.clear();
    if (!out) return fail(CQPI_E_ARG, "out is NULL");
    *out = nullptr;
    if (!n || !kpoints || !oversamp || !bands) return fail(CQPI_E_ARG, "n, kpoints, oversamp and bands must be > 0");
    if (wanniern != oversamp * n) return fail(CQPI_E_ARG, "wanniern must equal oversamp*n");
    if (!is_supported_bands(bands)) return fail(CQPI_E_ARG, "unsupported band count (1-12, 14, 16)");
    if (maxband == 0) maxband = bands;
    if (maxband > bands) return fail(CQPI_E_ARG, "maxband > bands");
    if (spin && maxband < 2) return fail(CQPI_E_ARG, "spin needs maxband >= 2");
    if (!wf) return fail(CQPI_E_ARG, "wf is NULL");
    if (kpoints * sizeof(double2) > 80 * 1024) return fail(CQPI_E_ARG, "kpoints > 5120 not supported");
    if (cqpi_device_count() <= 0) return fail(CQPI_E_NODEV, "no HIP device visible");
    cqpi_ctx *c = new cqpi_ctx();
    Geom &g = c->g;
    g.K = (int)kpoints;
    g.n = (int)n;
    g.w = (int)window;
    g.os = (int)oversamp;
    g.B = (int)bands;
    g.maxb = (int)maxband;
    g.spin = spin ? 1 : 0;
    int n2 = g.n >> 1;
    g.h = std::max(n2 + g.w, 2 * g.w);
    g.nw = 2 * g.h + 1;
    g.T = (2 * g.w + 1) * g.os;
    g.W = g.os * g.n;
    c->wanniern = wanniern;
    c->kpoints = kpoints;
    c->scmax = g.B;
    (void)hipGetDevice(&c->device);
    if (scratch_setup(c->device)) {
        std::string msg = g_err;
        release(c);
        return fail(CQPI_E_HIP, msg);
    }
    if (hipStreamCreate(&c->stream) != hipSuccess) {
        release(c);
        return fail(CQPI_E_HIP, "hipStreamCreate failed");
    }
    if (hipEventCreateWithFlags(&c->copy_done, hipEventDisableTiming) != hipSuccess) {
        release(c);
        return fail(CQPI_E_HIP, "hipEventCreate failed");
    }
    size_t ntile = (size_t)maxband * g.T * g.T;
    c->h_tiles.assign(wf, wf + ntile);
    int r = dev_grow(c->d_tiles, c->c_tiles, ntile, c->gpumem);
    if (!r && hipMemcpy(c->d_tiles, wf, ntile * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        r = fail(CQPI_E_HIP, "tile upload failed");
    // twiddles exp(+2 pi i m / K), m < K
    if (!r) {
        std::vector<double2> tw(kpoints);
        for (size_t m = 0; m < kpoints; m++) {
            long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)kpoints;
            tw[m] = make_double2((double)cosl(a), (double)sinl(a));
        }
        r = dev_grow(c->d_tw, c->c_tw, kpoints, c->gpumem);
        if (!r && hipMemcpy(c->d_tw, tw.data(), kpoints * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess)
            r = fail(CQPI_E_HIP, "twiddle upload failed");
        std::vector<float2> tw32(kpoints);
        for (size_t m = 0; m < kpoints; m++) tw32[m] = make_float2((float)tw[m].x, (float)tw[m].y);
        if (!r) r = dev_grow(c->d_tw32, c->c_tw32, kpoints, c->gpumem);
        if (!r && hipMemcpy(c->d_tw32, tw32.data(), kpoints * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess)
            r = fail(CQPI_E_HIP, "twiddle upload failed");
    }
    if (!r) r = upload_identity_V(c);
    if (!r) r = dev_grow(c->d_ldos, c->c_ldos, (size_t)g.W * g.W, c->gpumem);
    if (r) {
        release(c);
        return r;
    }
    *out = c;
    return 0;
}

void cqpi_destroy(cqpi_ctx *ctx) { release(ctx); }

int cqpi_create_spectral(size_t kpoints, size_t n, size_t bands, size_t maxbands, int spin, const double *pos,
                         size_t npos, const double *prearr, size_t nprearr, cqpi_ctx **out)
{
    g_err.clear();
    if (!out) return fail(CQPI_E_ARG, "NULL argument");
    *out = nullptr;
    if (!maxbands) maxbands = bands;
    if (n > kpoints) return fail(CQPI_E_ARG, "the map side n must not exceed kpoints");
    if (maxbands > bands) return fail(CQPI_E_ARG, "maxbands > bands");
    if (npos && npos < maxbands) return fail(CQPI_E_ARG, "pos must hold maxbands positions (or none)");
    if ((npos && !pos) || (nprearr && !prearr)) return fail(CQPI_E_ARG, "NULL pos or prearr");
    // orbital x, y and prefactor (the reference's d_mpos, src/qpihipkernel.cc:511-521)
    std::vector<double> sp(3 * maxbands, 0.0);
    for (size_t i = 0; i < maxbands; i++) {
        if (npos) {
            sp[3 * i] = pos[2 * i];
            sp[3 * i + 1] = pos[2 * i + 1];
        }
        sp[3 * i + 2] = nprearr ? prearr[i % nprearr] : 1.0;
    }
    std::vector<double> zero_tiles(maxbands, 0.0);  // window 0, oversampling 1: map n x n
    RETCHK(cqpi_create(n, kpoints, n, 0, 1, bands, maxbands, spin, zero_tiles.data(), out));
    cqpi_ctx *c = *out;
    c->spectral = true;
    int r = dev_grow(c->d_spos, c->c_spos, sp.size(), c->gpumem);
    if (!r && hipMemcpy(c->d_spos, sp.data(), sp.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        r = fail(CQPI_E_HIP, "position upload failed");
    if (r) {
        release(c);
        *out = nullptr;
    }
    return r;
}

// one spectral map of the drop-in path from the caller's G0(k) on the whole grid
static int dropin_spectral(cqpi_ctx *c, const double *const *g0k, size_t tda, int mode)
{
    g_err.clear();
    if (!c || !g0k) return fail(CQPI_E_ARG, "NULL argument");
    if (!c->spectral) return fail(CQPI_E_STATE, "this context was not created by cqpi_create_spectral");
    const Geom &g = c->g;
    const size_t B = g.B, NP = B * B, KK = (size_t)g.K * g.K, n = KK * NP;
    if (tda < B) return fail(CQPI_E_ARG, "tda < bands");
    HIPCHK(hipStreamSynchronize(c->stream));
    reset_stages(c);
    if (n > c->c_hg0k) {
        if (c->h_g0k) (void)hipHostFree(c->h_g0k);
        c->h_g0k = nullptr;
        c->c_hg0k = 0;
        HIPCHK(hipHostMalloc((void **)&c->h_g0k, n * sizeof(double2)));
        c->c_hg0k = n;
    }
    for (size_t kk = 0; kk < KK; kk++) {
        const double *m = g0k[kk];
        if (!m) return fail(CQPI_E_ARG, "NULL matrix in g0 table");
        for (size_t r = 0; r < B; r++)
            for (size_t cc = 0; cc < B; cc++)
                c->h_g0k[kk * NP + r * B + cc] = make_double2(m[2 * (r * tda + cc)], m[2 * (r * tda + cc) + 1]);
    }
    RETCHK(dev_grow(c->d_g0k, c->c_g0k, n, c->gpumem));
    RETCHK(dev_grow(c->d_ldos, c->c_ldos, (size_t)g.W * g.W, c->gpumem));
    HIPCHK(hipMemcpyAsync(c->d_g0k, c->h_g0k, n * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipEventRecord(c->copy_done, c->stream));
    int st = stage_begin(c, 4);
    int r = launch_spectral(g.K, g.n, g.B, g.maxb, g.spin, mode, c->d_g0k, nullptr, nullptr, nullptr, 1, 0.0,
                            c->d_spos, c->d_ldos, c->stream);
    stage_end(c, st);
    if (r) return fail(r, "spectral-function launch failed");
    c->nlayers = 1;
    c->windows_kept = false;
    HIPCHK(hipEventSynchronize(c->copy_done));  // the caller may overwrite g0 now
    return 0;
}

int cqpi_spf(cqpi_ctx *c, const double *const *g0k, size_t tda) { return dropin_spectral(c, g0k, tda, 1); }

int cqpi_uspf(cqpi_ctx *c, const double *const *g0k, size_t tda) { return dropin_spectral(c, g0k, tda, 2); }

int cqpi_set_spectral(cqpi_ctx *c, int mode)
{
    if (!c) return fail(CQPI_E_ARG, "ctx is NULL");
    if (mode < 0 || mode > 2) return fail(CQPI_E_ARG, "mode must be 0 (none), 1 (SPF) or 2 (USPF)");
    if (mode && !c->spectral) return fail(CQPI_E_STATE, "this context was not created by cqpi_create_spectral");
    c->spec_mode = mode;
    return 0;
}

// energies per G(k) block of the surface Green's function (FP64 workspace
// kept under ~2 GiB); grows the workspace and the step counters
static int surface_blocks(cqpi_ctx *c, size_t E, int &EB)
{
    const Geom &g = c->g;
    const size_t per_gk = (size_t)g.B * g.B * g.K * g.K * sizeof(double2);
    EB = (int)std::max<size_t>(1, std::min<size_t>(E, ((size_t)2 << 30) / per_gk));
    RETCHK(dev_grow(c->d_gkw, c->c_gkw, (size_t)EB * per_gk / sizeof(double2), c->gpumem));
    RETCHK(dev_grow(c->d_sstat, c->c_sstat, E, c->gpumem));
    return 0;
}

// FP64 G(k) of the surface Green's function of energies d_om[0, E) into the workspace
static int surface_gk(cqpi_ctx *c, const double *d_om, int E, double eta, int *stat)
{
    const DevRGroup *sg[3];
    const double2 *stm[3];
    for (int q = 0; q < 3; q++) {
        sg[q] = c->d_sgroups + c->sg_goff[q];
        stm[q] = c->d_stmat + c->sg_toff[q];
    }
    return launch_surface_gk(c->g.B, c->g.K, sg, stm, c->sg_n, c->d_tw, c->efermi, c->scmax, d_om, E, eta,
                             c->surf_eps, c->surf_maxit, c->surf_mode == CQPI_GREEN_SURFACE, c->d_gkw, false, stat,
                             c->stream);
}

// after the stream is done: CQPI_E_CONVERGE if a k-point of energy e < E ran out of steps
static int check_surface_steps(cqpi_ctx *c, size_t E)
{
    std::vector<int> st(E);
    HIPCHK(hipMemcpy(st.data(), c->d_sstat, E * sizeof(int), hipMemcpyDeviceToHost));
    for (size_t e = 0; e < E; e++)
        if (st[e] > c->surf_maxit)
            return fail(CQPI_E_CONVERGE, "the surface Green's function did not converge within " +
                                             std::to_string(c->surf_maxit) + " steps at energy index " +
                                             std::to_string(e));
    return 0;
}

int cqpi_dos(cqpi_ctx *c, const double *omegas, size_t E, double eta, double *dos)
{
    g_err.clear();
    if (!c || !omegas || !dos) return fail(CQPI_E_ARG, "NULL argument");
    if (!E) return fail(CQPI_E_ARG, "no energies");
    if (!c->has_model) return fail(CQPI_E_STATE, "cqpi_set_model has not been called");
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!c->eigen_valid && !c->surf_mode) {
        c->in_run = true;
        int r = cqpi_compute_eigen(c);
        c->in_run = false;
        if (r) return r;
    }
    const Geom &g = c->g;
    const int nc = dos_chunks(g.K);
    RETCHK(dev_grow(c->d_domega, c->c_domega, E, c->gpumem));
    RETCHK(dev_grow(c->d_dos, c->c_dos, E, c->gpumem));
    RETCHK(dev_grow(c->d_dosp, c->c_dosp, E * (size_t)nc, c->gpumem));
    HIPCHK(hipMemcpy(c->d_domega, omegas, E * sizeof(double), hipMemcpyHostToDevice));
    if (c->surf_mode) {
        int EB = 1;
        RETCHK(surface_blocks(c, E, EB));
        for (size_t ea = 0; ea < E; ea += EB) {
            const int eb = (int)std::min<size_t>(EB, E - ea);
            int r = surface_gk(c, c->d_domega + ea, eb, eta, c->d_sstat + ea);
            if (!r) r = launch_dos_gk(g.K, g.B, g.maxb, c->d_gkw, eb, c->d_dosp + ea * nc, c->d_dos + ea, c->stream);
            if (r) return fail(r, "density-of-states launch failed");
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        RETCHK(check_surface_steps(c, E));
    } else {
        int r = launch_dos(g.K, g.B, g.maxb, c->d_eval, c->d_evec, c->d_domega, (int)E, eta, c->d_dosp, c->d_dos,
                           c->stream);
        if (r) return fail(r, "density-of-states launch failed");
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    HIPCHK(hipMemcpy(dos, c->d_dos, E * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int cqpi_create_list(size_t wanniern, size_t kpoints, size_t n, size_t window, size_t oversamp, size_t bands,
                     const cqpi_flagentry *flags, const size_t *counts, cqpi_ctx **out)
{
    g_err.clear();
    if (!out || !counts) return fail(CQPI_E_ARG, "NULL argument");
    *out = nullptr;
    if (bands > 16) return fail(CQPI_E_ARG, "the threshold-list contraction supports up to 16 bands");
    const size_t side = 2 * window + 1, NR = side * side, npair = NR * NR;
    const size_t NP = bands * bands, npix = oversamp * oversamp;
    size_t total = 0;
    for (size_t p = 0; p < npair; p++) total += counts[p];
    if (total && !flags) return fail(CQPI_E_ARG, "flags is NULL");
    // dense factor table of the active pairs, [k][o2][o1][pix], duplicates summed
    std::vector<int> r1off(NR + 1, 0), lpair;
    std::vector<double> coef;
    size_t pos = 0;
    for (size_t p = 0; p < npair; p++) {
        const size_t cnt = counts[p];
        if (!cnt) continue;
        const size_t k = lpair.size();
        lpair.push_back((int)p);  // p = r1 * NR + r2 in IDX4CC order
        r1off[p / NR + 1]++;
        coef.resize((k + 1) * NP * npix, 0.0);
        for (size_t m = 0; m < cnt; m++, pos++) {
            const cqpi_flagentry &f = flags[pos];
            if (f.i >= oversamp || f.j >= oversamp || f.o1 >= bands || f.o2 >= bands)
                return fail(CQPI_E_ARG, "threshold-list entry out of range (i, j < oversamp; o1, o2 < bands)");
            coef[(k * NP + (size_t)f.o2 * bands + f.o1) * npix + (size_t)f.i * oversamp + f.j] += f.factor;
        }
    }
    for (size_t r = 0; r < NR; r++) r1off[r + 1] += r1off[r];
    // the orbital pairs present per active pair: combos ordered by (k, o1, o2),
    // coff[k B + o1] the first of (k, o1); ccoef their factors per sub-pixel
    const size_t nact = lpair.size();
    std::vector<int> coff(nact * bands + 1, 0), co2;
    std::vector<double> ccoef;
    for (size_t k = 0; k < nact; k++)
        for (size_t o1 = 0; o1 < bands; o1++) {
            coff[k * bands + o1] = (int)co2.size();
            for (size_t o2 = 0; o2 < bands; o2++) {
                const double *cp = &coef[(k * NP + o2 * bands + o1) * npix];
                if (std::none_of(cp, cp + npix, [](double v) { return v != 0.0; })) continue;
                co2.push_back((int)o2);
                ccoef.insert(ccoef.end(), cp, cp + npix);
            }
        }
    coff[nact * bands] = (int)co2.size();
    const size_t T = side * oversamp;
    std::vector<double> zero_tiles(bands * T * T, 0.0);  // the list carries the factors
    RETCHK(cqpi_create(wanniern, kpoints, n, window, oversamp, bands, bands, 0, zero_tiles.data(), out));
    cqpi_ctx *c = *out;
    c->list_mode = true;
    c->lnact = (int)lpair.size();
    size_t cap_coef = 0, cap_off = 0, cap_pair = 0, cap_cc = 0, cap_co = 0, cap_c2 = 0;
    int r = dev_grow(c->d_lcoef, cap_coef, std::max<size_t>(coef.size(), 1), c->gpumem);
    if (!r) r = dev_grow(c->d_lccoef, cap_cc, std::max<size_t>(ccoef.size(), 1), c->gpumem);
    if (!r) r = dev_grow(c->d_lcoff, cap_co, coff.size(), c->gpumem);
    if (!r) r = dev_grow(c->d_lco2, cap_c2, std::max<size_t>(co2.size(), 1), c->gpumem);
    if (!r && !ccoef.empty() &&
        hipMemcpy(c->d_lccoef, ccoef.data(), ccoef.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        r = fail(CQPI_E_HIP, "threshold-list upload failed");
    if (!r && hipMemcpy(c->d_lcoff, coff.data(), coff.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
        r = fail(CQPI_E_HIP, "threshold-list upload failed");
    if (!r && !co2.empty() &&
        hipMemcpy(c->d_lco2, co2.data(), co2.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
        r = fail(CQPI_E_HIP, "threshold-list upload failed");
    if (!r) r = dev_grow(c->d_lr1off, cap_off, NR + 1, c->gpumem);
    if (!r) r = dev_grow(c->d_lpair, cap_pair, std::max<size_t>(lpair.size(), 1), c->gpumem);
    if (!r && !coef.empty() &&
        hipMemcpy(c->d_lcoef, coef.data(), coef.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        r = fail(CQPI_E_HIP, "threshold-list upload failed");
    if (!r && hipMemcpy(c->d_lr1off, r1off.data(), (NR + 1) * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
        r = fail(CQPI_E_HIP, "threshold-list upload failed");
    if (!r && !lpair.empty() &&
        hipMemcpy(c->d_lpair, lpair.data(), lpair.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
        r = fail(CQPI_E_HIP, "threshold-list upload failed");
    if (r) {
        release(c);
        *out = nullptr;
        return r;
    }
    return 0;
}

int cqpi_make_threshold_list(const double *wf, size_t norb, size_t window, size_t oversamp, double threshold,
                             size_t maxband, int spin, cqpi_flagentry *out, size_t cap, size_t *counts,
                             size_t *total)
{
    g_err.clear();
    if (!wf || !counts || !total) return fail(CQPI_E_ARG, "NULL argument");
    if (!maxband) maxband = norb;
    if (maxband > norb) return fail(CQPI_E_ARG, "maxband > number of tiles");
    const size_t T = (2 * window + 1) * oversamp, wins = 2 * window + 1, sb = maxband >> 1;
    auto wv = [&](size_t o, size_t x, size_t y) { return wf[(o * T + x) * T + y]; };
    double maxval = 0.0;
    for (size_t l = 0; l < norb; l++)
        for (size_t i = 0; i < T; i++)
            for (size_t j = 0; j < T; j++) maxval = std::max(maxval, wv(l, i, j) * wv(l, i, j));
    const double cut = threshold * maxval;
    const size_t np = spin ? sb : maxband;
    size_t cnt = 0;
    for (size_t xn1 = 0; xn1 < wins; xn1++)
        for (size_t yn1 = 0; yn1 < wins; yn1++)
            for (size_t xn2 = 0; xn2 < wins; xn2++)
                for (size_t yn2 = 0; yn2 < wins; yn2++) {
                    const size_t first = cnt;
                    for (size_t o1 = 0; o1 < np; o1++)
                        for (size_t o2 = 0; o2 < np; o2++)
                            for (size_t i = 0; i < oversamp; i++)
                                for (size_t j = 0; j < oversamp; j++) {
                                    const size_t x = i + (wins - 1) * oversamp, y = j + (wins - 1) * oversamp;
                                    for (int part = 0; part < (spin ? 2 : 1); part++) {
                                        const size_t a = part ? sb + o1 : o1, b = part ? sb + o2 : o2;
                                        const double f = wv(a, x - xn1 * oversamp, y - yn1 * oversamp) *
                                                         wv(b, x - xn2 * oversamp, y - yn2 * oversamp);
                                        if (std::fabs(f) > cut) {
                                            if (out && cnt < cap)
                                                out[cnt] = cqpi_flagentry{(uint32_t)i, (uint32_t)j, (uint32_t)a,
                                                                          (uint32_t)b, f};
                                            cnt++;
                                        }
                                    }
                                }
                    counts[((xn1 * wins + yn1) * wins + xn2) * wins + yn2] = cnt - first;
                }
    *total = cnt;
    if (out && cnt > cap) return fail(CQPI_E_ARG, "cap is smaller than the list; *total holds its length");
    return 0;
}

int cqpi_make_threshold_list_spinarr(const double *wf, size_t norb, size_t window, size_t oversamp,
                                     double threshold, const int *spinarr, size_t nspinarr, cqpi_flagentry *out,
                                     size_t cap, size_t *counts, size_t *total)
{
    g_err.clear();
    if (!wf || !counts || !total || (!spinarr && nspinarr)) return fail(CQPI_E_ARG, "NULL argument");
    // the spin label of every tile: spinarr resized to norb, missing labels 0
    // (no entries), as src/qpi.cc:217
    std::vector<int> spin(norb, 0);
    for (size_t o = 0; o < std::min(norb, nspinarr); o++) spin[o] = spinarr[o];
    const size_t T = (2 * window + 1) * oversamp, wins = 2 * window + 1;
    auto wv = [&](size_t o, size_t x, size_t y) { return wf[(o * T + x) * T + y]; };
    double maxval = 0.0;
    for (size_t l = 0; l < norb; l++)
        for (size_t i = 0; i < T; i++)
            for (size_t j = 0; j < T; j++) maxval = std::max(maxval, wv(l, i, j) * wv(l, i, j));
    const double cut = threshold * maxval;
    size_t cnt = 0;
    for (size_t xn1 = 0; xn1 < wins; xn1++)
        for (size_t yn1 = 0; yn1 < wins; yn1++)
            for (size_t xn2 = 0; xn2 < wins; xn2++)
                for (size_t yn2 = 0; yn2 < wins; yn2++) {
                    const size_t first = cnt;
                    for (size_t o1 = 0; o1 < norb; o1++) {
                        if (!spin[o1]) continue;
                        for (size_t o2 = 0; o2 < norb; o2++) {
                            if (spin[o2] != spin[o1]) continue;
                            for (size_t i = 0; i < oversamp; i++)
                                for (size_t j = 0; j < oversamp; j++) {
                                    const size_t x = i + (wins - 1) * oversamp, y = j + (wins - 1) * oversamp;
                                    const double f = wv(o1, x - xn1 * oversamp, y - yn1 * oversamp) *
                                                     wv(o2, x - xn2 * oversamp, y - yn2 * oversamp);
                                    if (std::fabs(f) > cut) {
                                        if (out && cnt < cap)
                                            out[cnt] = cqpi_flagentry{(uint32_t)i, (uint32_t)j, (uint32_t)o1,
                                                                      (uint32_t)o2, f};
                                        cnt++;
                                    }
                                }
                        }
                    }
                    counts[((xn1 * wins + yn1) * wins + xn2) * wins + yn2] = cnt - first;
                }
    *total = cnt;
    if (out && cnt > cap) return fail(CQPI_E_ARG, "cap is smaller than the list; *total holds its length");
    return 0;
}

int cqpi_set_precision(cqpi_ctx *c, int precision)
{
    if (!c) return fail(CQPI_E_ARG, "ctx is NULL");
    if (precision != CQPI_PREC_FP64 && precision != CQPI_PREC_FP32 && precision != CQPI_PREC_MIXED)
        return fail(CQPI_E_ARG, "precision mode not available");
    if (precision == CQPI_PREC_MIXED && c->list_mode)
        return fail(CQPI_E_ARG, "the threshold-list contraction is computed in FP64 or FP32 (it is no matrix product)");
    if (precision == CQPI_PREC_FP32 && !c->d_tilesf) {
        std::vector<float> tf(c->h_tiles.begin(), c->h_tiles.end());
        RETCHK(dev_grow(c->d_tilesf, c->c_tilesf, tf.size(), c->gpumem));
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipMemcpy(c->d_tilesf, tf.data(), tf.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    if (precision == CQPI_PREC_MIXED && !c->mix_ready) {
        if (c->g.os > 16) return fail(CQPI_E_ARG, "mixed precision needs oversamp <= 16");
        c->mg = mix_geometry(c->g);
        std::vector<_Float16> F = mix_tile_fragments(c->g, c->mg, c->h_tiles.data(), c->sA);
        RETCHK(dev_grow(c->d_afrag, c->c_afrag, F.size(), c->gpumem));
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipMemcpy(c->d_afrag, F.data(), F.size() * sizeof(_Float16), hipMemcpyHostToDevice));
        c->mix_ready = true;
    }
    c->precision = precision;
    return 0;
}

int cqpi_set_timing(cqpi_ctx *c, int on)
{
    if (!c) return fail(CQPI_E_ARG, "ctx is NULL");
    c->timing = on != 0;
    return 0;
}

int cqpi_stage_times(cqpi_ctx *c, double *ms, int64_t *launches)
{
    if (!c) return fail(CQPI_E_ARG, "ctx is NULL");
    for (int s = 0; s < CQPI_NSTAGES; s++) {
        if (ms) ms[s] = c->stage_ms[s];
        if (launches) launches[s] = c->stage_n[s];
    }
    return 0;
}

int cqpi_wannierldos(cqpi_ctx *c, const double *scat, const double *const *g0, size_t tda)
{
    if (c && c->list_mode)
        return fail(CQPI_E_STATE, "this context holds a threshold list: use cqpi_wannierldoslist");
    if (c && c->spectral) return fail(CQPI_E_STATE, "this context makes spectral-function maps (cqpi_spf, cqpi_uspf)");
    return dropin_map(c, scat, g0, tda, nullptr);
}

int cqpi_wannierldoslist(cqpi_ctx *c, const double *scat, const double *const *g0, size_t tda)
{
    if (c && !c->list_mode) return fail(CQPI_E_STATE, "this context was not created by cqpi_create_list");
    return dropin_map(c, scat, g0, tda, nullptr);
}

int cqpi_wannierjosephson(cqpi_ctx *c, const double *scat, const double *const *g0, size_t tda, double tip_re,
                          double tip_im)
{
    const double tip[2] = {tip_re, tip_im};
    if (c && (c->list_mode || c->spectral))
        return fail(CQPI_E_STATE, "the Josephson map needs a context made by cqpi_create");
    return dropin_map(c, scat, g0, tda, tip);
}

int cqpi_wannierldos_device(cqpi_ctx *c, const double *d_T, const double *d_g0win, size_t E, double *d_ldos)
{
    g_err.clear();
    if (!c || !d_T || !d_g0win || !d_ldos) return fail(CQPI_E_ARG, "NULL argument");
    if (c->list_mode || c->spectral) return fail(CQPI_E_STATE, "the LDOS map needs a context made by cqpi_create");
    if (!E || E > 65535) return fail(CQPI_E_ARG, "energies per call must be 1 to 65535");
    void *base = scratch_of(c->device);
    if (!base) return fail(CQPI_E_STATE, "no scratch address for the context's device");
    const Geom &g = c->g;
    const size_t NP = (size_t)g.B * g.B;
    // scratch layout: periodic table, then (mixed) the plane scales, the scales
    // of the T-folded b planes and the planes or (FP32)
    // narrowed windows and T, each part 256-byte aligned
    size_t off = 0;
    auto carve = [&off](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) / 256 * 256;
        return o;
    };
    const size_t o_bulk = carve(E * g.os * g.os * sizeof(double));
    size_t o_a = 0, o_b = 0;
    size_t o_c = 0;
    if (c->precision == CQPI_PREC_MIXED) {
        o_a = carve(E * sizeof(double));
        o_c = carve(E * sizeof(double));
        o_b = carve(E * mix_plane_halves(g, c->mg) * sizeof(_Float16));
    } else if (c->precision == CQPI_PREC_FP32) {
        o_a = carve(E * g.nw * g.nw * NP * sizeof(float2));
        o_b = carve(E * NP * sizeof(float2));
    }
    if (off > (size_t)CQPI_NOALLOC_SCRATCH_BYTES)
        return fail(CQPI_E_ARG, "this call needs " + std::to_string(off) + " bytes of scratch, more than the fixed " +
                                    std::to_string((size_t)CQPI_NOALLOC_SCRATCH_BYTES) +
                                    ": pass fewer energies per call");
    unsigned char *sp = static_cast<unsigned char *>(base);
    const double2 *G0w = reinterpret_cast<const double2 *>(d_g0win);
    const double2 *T = reinterpret_cast<const double2 *>(d_T);
    double *bulk = reinterpret_cast<double *>(sp + o_bulk);
    const int En = (int)E;
    hipStream_t s = c->stream;
    RETCHK(launch_bulk(g, G0w, c->d_tiles, En, bulk, s));
    int r = 0;
    switch (c->precision) {
    case CQPI_PREC_MIXED: {
        double *scale = reinterpret_cast<double *>(sp + o_a);
        _Float16 *planes = reinterpret_cast<_Float16 *>(sp + o_b);
        const bool tf = mix_tfold(g, c->mg);
        r = launch_mix_prepare(g, c->mg, G0w, En, scale, planes, s, tf);
        if (!r)
            r = launch_contract_mixed(g, c->mg, planes, c->d_afrag, scale, c->sA, T, bulk, En, d_ldos, s, nullptr, tf,
                                      reinterpret_cast<double *>(sp + o_c));
        break;
    }
    case CQPI_PREC_FP32:
        r = launch_contract_fp32(g, G0w, T, bulk, c->d_tilesf, En, d_ldos, reinterpret_cast<float2 *>(sp + o_a),
                                 reinterpret_cast<float2 *>(sp + o_b), s);
        break;
    default:
        r = launch_contract_fp64(g, G0w, T, bulk, c->d_tiles, En, d_ldos, s);
        break;
    }
    if (r) return fail(r, "contraction launch failed");
    return 0;
}

int cqpi_retrieve(cqpi_ctx *c, double *ldos) { return cqpi_retrieve_layers(c, 0, 1, ldos); }

int cqpi_retrieve_layers(cqpi_ctx *c, size_t first, size_t count, double *ldos)
{
    g_err.clear();
    if (!c || !ldos) return fail(CQPI_E_ARG, "NULL argument");
    if (first + count > c->nlayers) return fail(CQPI_E_STATE, "layer range beyond the last run");
    HIPCHK(hipStreamSynchronize(c->stream));
    stage_collect(c);
    size_t WW = (size_t)c->g.W * c->g.W;
    HIPCHK(hipMemcpy(ldos, c->d_ldos + first * WW, count * WW * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int cqpi_device_ldos(cqpi_ctx *c, const double **dptr, size_t *count)
{
    if (!c || !dptr) return fail(CQPI_E_ARG, "NULL argument");
    *dptr = c->d_ldos;
    if (count) *count = c->nlayers * (size_t)c->g.W * c->g.W;
    return 0;
}

int cqpi_printinfo(cqpi_ctx *c, char *buf, size_t len)
{
    if (!c || !buf || !len) return fail(CQPI_E_ARG, "NULL argument");
    hipDeviceProp_t props;
    std::string name = "unknown";
    if (hipGetDeviceProperties(&props, c->device) == hipSuccess) name = props.name;
    char tmp[1024];
    snprintf(tmp, sizeof tmp,
             "Running continuum QPI on HIP GPU %d: %s (%s)\n"
             "Additional memory requirements:\n"
             "GPU memory:       %.2fGB\n"
             "Map:              %dx%d pixels, %d bands, window %d, oversamp %d, k-grid %d\n"
             "G0 window:        %dx%d (R in [-%d,%d])\n",
             c->device, name.c_str(), cqpi_version(), (double)c->gpumem / 1073741824.0, c->g.W, c->g.W,
             c->g.B, c->g.w, c->g.os, c->g.K, c->g.nw, c->g.nw, c->g.h, c->g.h);
    strncpy(buf, tmp, len - 1);
    buf[len - 1] = 0;
    return 0;
}

// Hopping terms grouped by lattice vector (rx, ry), 'anti' terms folded onto
// -R; each group sums its amplitudes into a dense B x B block.  layer < 0:
// every term (kz = 0, so rz drops out of the phase); 0: the in-plane terms
// (rz == 0); 1: the coupling to the next layer (rz < 0), as
// tightbind::setmatrix with kzdepinplane / kzdepoutofplane
// (src/tightbinding.cc:184-208).  Returns the index of a term whose orbital
// is out of range, else -1.
static long group_hoppings(const cqpi_hopping *hops, size_t nhops, int B, int K, int layer, std::vector<DevRGroup> &G,
                           std::vector<double2> &T)
{
    std::map<std::pair<double, double>, size_t> index;
    G.clear();
    T.clear();
    for (size_t i = 0; i < nhops; i++) {
        const cqpi_hopping &h = hops[i];
        if (h.o1 < 0 || h.o2 < 0 || h.o1 >= B || h.o2 >= B) return (long)i;
        if (layer == 0 && !(h.rz == 0.0)) continue;
        if (layer == 1 && !(h.rz < 0.0)) continue;
        const double rx = (h.anti ? -h.rx : h.rx) + 0.0, ry = (h.anti ? -h.ry : h.ry) + 0.0;
        auto key = std::make_pair(rx, ry);
        auto it = index.find(key);
        if (it == index.end()) {
            DevRGroup r;
            r.rx = rx;
            r.ry = ry;
            const bool integral = rx == std::floor(rx) && ry == std::floor(ry) && std::fabs(rx) < 1e9 &&
                                  std::fabs(ry) < 1e9 && K <= 32768;
            r.integral = integral ? 1 : 0;
            r.irx = integral ? (int)(((long long)rx % K + K) % K) : 0;
            r.iry = integral ? (int)(((long long)ry % K + K) % K) : 0;
            r.pad = 0;
            it = index.emplace(key, G.size()).first;
            G.push_back(r);
            T.resize(T.size() + (size_t)B * B, make_double2(0.0, 0.0));
        }
        double2 &t = T[it->second * B * B + (size_t)h.o1 * B + h.o2];
        t.x += h.tre;
        t.y += h.tim;
    }
    return -1;
}

// uploads the three hopping-group sets of the decimation (bulk in-plane,
// surface in-plane, coupling) from the context's models
static int upload_surface_groups(cqpi_ctx *c)
{
    const int B = c->g.B, K = c->g.K;
    std::vector<DevRGroup> Gall, G;
    std::vector<double2> Tall, T;
    for (int q = 0; q < 3; q++) {
        const std::vector<cqpi_hopping> &src = (q == 1 && !c->h_shops.empty()) ? c->h_shops : c->h_hops;
        const long bad = group_hoppings(src.data(), src.size(), B, K, q == 2 ? 1 : 0, G, T);
        if (bad >= 0)
            return fail(CQPI_E_ARG, std::string(q == 1 && !c->h_shops.empty() ? "surface " : "") +
                                        "hopping orbital index out of range at term " + std::to_string(bad));
        c->sg_goff[q] = Gall.size();
        c->sg_toff[q] = Tall.size();
        c->sg_n[q] = (int)G.size();
        Gall.insert(Gall.end(), G.begin(), G.end());
        Tall.insert(Tall.end(), T.begin(), T.end());
    }
    RETCHK(dev_grow(c->d_sgroups, c->c_sgroups, std::max<size_t>(Gall.size(), 1), c->gpumem));
    RETCHK(dev_grow(c->d_stmat, c->c_stmat, std::max<size_t>(Tall.size(), 1), c->gpumem));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!Gall.empty())
        HIPCHK(hipMemcpy(c->d_sgroups, Gall.data(), Gall.size() * sizeof(DevRGroup), hipMemcpyHostToDevice));
    if (!Tall.empty()) HIPCHK(hipMemcpy(c->d_stmat, Tall.data(), Tall.size() * sizeof(double2), hipMemcpyHostToDevice));
    return 0;
}

// H(k)^T = H(-k) on the k-grid: every block of H(k) = sum_R T_R e^{2 pi i k.R}
// (the groups k_eigen sums, hole-block terms at their conjugate-phase R) sits
// at an integral R and has its exact transpose at -R (real hoppings of a
// Hermitian model, and a particle-hole model built from them with real
// pairing); the diagonal -/+ efermi is real.  Then G0(k)^T = G0(-k) at every
// energy, and the transform of the pair (r, c) at R is the pair (c, r) at -R.
static bool transpose_symmetric(const std::vector<DevRGroup> &G, const std::vector<double2> &T, int B)
{
    std::map<std::pair<double, double>, size_t> at;
    for (size_t i = 0; i < G.size(); i++) {
        if (!G[i].integral) return false;
        at.emplace(std::make_pair(G[i].rx, G[i].ry), i);
    }
    for (size_t i = 0; i < G.size(); i++) {
        auto it = at.find(std::make_pair(-G[i].rx + 0.0, -G[i].ry + 0.0));
        if (it == at.end()) return false;
        const double2 *a = &T[i * B * B], *m = &T[it->second * B * B];
        for (int r = 0; r < B; r++)
            for (int q = 0; q < B; q++)
                if (!(a[r * B + q].x == m[q * B + r].x && a[r * B + q].y == m[q * B + r].y)) return false;
    }
    return true;
}

int cqpi_set_model(cqpi_ctx *c, const cqpi_hopping *hops, size_t nhops, double efermi, int scmodel)
{
    g_err.clear();
    if (!c) return fail(CQPI_E_ARG, "ctx is NULL");
    if (nhops && !hops) return fail(CQPI_E_ARG, "hops is NULL");
    const int B = c->g.B;
    std::vector<DevRGroup> G;
    std::vector<double2> T;
    const long bad = group_hoppings(hops, nhops, B, c->g.K, -1, G, T);
    if (bad >= 0) return fail(CQPI_E_ARG, "hopping orbital index out of range at term " + std::to_string(bad));
    RETCHK(dev_grow(c->d_groups, c->c_groups, std::max<size_t>(G.size(), 1), c->gpumem));
    RETCHK(dev_grow(c->d_tmat, c->c_tmat, std::max<size_t>(T.size(), 1), c->gpumem));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!G.empty()) HIPCHK(hipMemcpy(c->d_groups, G.data(), G.size() * sizeof(DevRGroup), hipMemcpyHostToDevice));
    if (!T.empty()) HIPCHK(hipMemcpy(c->d_tmat, T.data(), T.size() * sizeof(double2), hipMemcpyHostToDevice));
    c->ngroups = (int)G.size();
    c->efermi = efermi;
    c->scmax = scmodel ? (B >> 1) : B;
    c->tsym_model = transpose_symmetric(G, T, B);
    c->has_model = true;
    c->h_hops.assign(hops, hops + nhops);
    c->windows_kept = false;
    c->eigen_valid = false;
    if (c->surf_mode) RETCHK(upload_surface_groups(c));
    return 0;
}

int cqpi_bands_histogram(cqpi_ctx *c, size_t pix, size_t oversamp, double le, double ue, size_t layers,
                         double *hist)
{
    g_err.clear();
    if (!c || !hist) return fail(CQPI_E_ARG, "NULL argument");
    if (!c->has_model) return fail(CQPI_E_STATE, "cqpi_set_model has not been called");
    if (!pix || !oversamp || pix * oversamp < 2 || pix * oversamp > 65536 || layers < 2 || !(ue > le))
        return fail(CQPI_E_ARG, "histogram needs pix*oversamp in [2, 65536], layers >= 2 and ue > le");
    const size_t n = pix * pix * layers;
    double *d = nullptr;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMalloc((void **)&d, n * sizeof(double)));
    int r = hipMemset(d, 0, n * sizeof(double)) == hipSuccess ? 0 : CQPI_E_HIP;
    if (!r)
        r = launch_bands_hist(c->g.B, c->d_groups, c->ngroups, c->d_tmat, c->efermi, c->scmax, (int)pix,
                              (int)oversamp, le, ue, (int)layers, d, c->stream);
    if (!r) r = hipStreamSynchronize(c->stream) == hipSuccess ? 0 : CQPI_E_HIP;
    if (!r) r = hipMemcpy(hist, d, n * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess ? 0 : CQPI_E_HIP;
    (void)hipFree(d);
    if (r) return fail(r, "band-structure histogram failed");
    return 0;
}

int cqpi_set_pair_symmetry(cqpi_ctx *c, int enable)
{
    g_err.clear();
    if (!c) return fail(CQPI_E_ARG, "ctx is NULL");
    c->tsym_on = enable != 0;
    return 0;
}

int cqpi_get_pair_symmetry(const cqpi_ctx *c, int *model, int *used)
{
    g_err.clear();
    if (!c) return fail(CQPI_E_ARG, "ctx is NULL");
    if (model) *model = c->tsym_model ? 1 : 0;
    if (used) *used = c->tsym_used ? 1 : 0;
    return 0;
}

int cqpi_set_surface_steps(cqpi_ctx *c, int maxit)
{
    g_err.clear();
    if (!c) return fail(CQPI_E_ARG, "ctx is NULL");
    if (maxit < 1 || maxit > 100000) return fail(CQPI_E_ARG, "the step limit must lie in [1, 100000]");
    c->surf_maxit = maxit;
    return 0;
}

int cqpi_set_surface(cqpi_ctx *c, int mode, double epserr, const cqpi_hopping *shops, size_t nshops)
{
    g_err.clear();
    if (!c) return fail(CQPI_E_ARG, "ctx is NULL");
    if (mode < CQPI_GREEN_NORMAL || mode > CQPI_GREEN_BULK) return fail(CQPI_E_ARG, "unknown Green's function mode");
    if (mode == CQPI_GREEN_NORMAL) {
        c->surf_mode = 0;
        c->h_shops.clear();
        return 0;
    }
    if (!(epserr > 0.0)) return fail(CQPI_E_ARG, "epserr must be positive");
    if (nshops && !shops) return fail(CQPI_E_ARG, "shops is NULL");
    if (c->g.B > 12 && c->g.B != 14 && c->g.B != 16)
        return fail(CQPI_E_ARG, "the surface Green's function supports 1-12, 14 and 16 bands");
    if (!c->has_model) return fail(CQPI_E_STATE, "cqpi_set_model has not been called");
    c->h_shops.assign(shops, shops + nshops);
    c->surf_mode = mode;
    c->surf_eps = epserr;
    const int r = upload_surface_groups(c);
    if (r) {
        c->surf_mode = 0;
        c->h_shops.clear();
        return r;
    }
    c->windows_kept = false;
    return 0;
}

int cqpi_set_impurity(cqpi_ctx *c, const double *V)
{
    g_err.clear();
    if (!c || !V) return fail(CQPI_E_ARG, "NULL argument");
    int B = c->g.B;
    RETCHK(dev_grow(c->d_V, c->c_V, (size_t)B * B, c->gpumem));
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(c->d_V, V, (size_t)B * B * sizeof(double2), hipMemcpyHostToDevice));
    return 0;
}

int cqpi_compute_eigen(cqpi_ctx *c)
{
    g_err.clear();
    if (!c) return fail(CQPI_E_ARG, "ctx is NULL");
    if (!c->has_model) return fail(CQPI_E_STATE, "cqpi_set_model has not been called");
    const Geom &g = c->g;
    if (!c->in_run) {
        HIPCHK(hipStreamSynchronize(c->stream));
        stage_collect(c);
        reset_stages(c);
        c->eigen_timed = true;
    }
    size_t KK = (size_t)g.K * g.K;
    RETCHK(dev_grow(c->d_eval, c->c_eval, KK * g.B, c->gpumem));
    RETCHK(dev_grow(c->d_evec, c->c_evec, KK * g.B * g.B, c->gpumem));
    int st = stage_begin(c, 0);
    RETCHK(launch_eigen(g.B, c->d_groups, c->ngroups, c->d_tmat, c->d_tw, c->efermi, c->scmax, g.K, c->d_eval,
                        c->d_evec, c->stream));
    stage_end(c, st);
    c->eigen_valid = true;
    return 0;
}

int cqpi_run(cqpi_ctx *c, const double *omegas, size_t E, double eta, double *ldos_host)
{
    g_err.clear();
    if (!c || (!omegas && E)) return fail(CQPI_E_ARG, "NULL argument");
    if (!c->has_model) return fail(CQPI_E_STATE, "cqpi_set_model has not been called");
    if (E == 0) return fail(CQPI_E_ARG, "no energies");
    // no stream synchronisation here: an eigensolve enqueued just before
    // keeps running while this run is prepared and enqueued (the previous
    // run ended synchronised in finish_run; stage events wait for themselves)
    stage_collect(c);
    // the timing window covers this run, plus a cqpi_compute_eigen called just before it
    if (!c->eigen_timed) reset_stages(c);
    c->eigen_timed = false;
    // the surface Green's function (every output, the spectral maps too) does
    // not use the eigen data
    if (!c->eigen_valid && !c->surf_mode) {
        c->in_run = true;
        int r = cqpi_compute_eigen(c);
        c->in_run = false;
        if (r) return r;
    }
    if (c->spectral && !c->spec_mode)
        return fail(CQPI_E_STATE, "a spectral-function context: select SPF or USPF with cqpi_set_spectral");
    if (c->spec_mode) {
        // SPF / USPF maps from the eigen data, n x n per energy
        const Geom &g = c->g;
        RETCHK(dev_grow(c->d_ldos, c->c_ldos, E * (size_t)g.W * g.W, c->gpumem));
        RETCHK(dev_grow(c->d_omega, c->c_omega, E, c->gpumem));
        c->up_omegas.assign(omegas, omegas + E);
        HIPCHK(hipMemcpyAsync(c->d_omega, c->up_omegas.data(), E * sizeof(double), hipMemcpyHostToDevice, c->stream));
        int st = stage_begin(c, 4);
        int r = 0;
        if (c->surf_mode) {
            int EB = 1;
            RETCHK(surface_blocks(c, E, EB));
            for (size_t ea = 0; ea < E && !r; ea += EB) {
                const int eb = (int)std::min<size_t>(EB, E - ea);
                r = surface_gk(c, c->d_omega + ea, eb, eta, c->d_sstat + ea);
                if (!r)
                    r = launch_spectral_gk(g.K, g.n, g.B, g.maxb, g.spin, c->spec_mode, c->d_gkw, eb, c->d_spos,
                                           c->d_ldos + ea * (size_t)g.W * g.W, c->stream);
            }
        } else {
            r = launch_spectral(g.K, g.n, g.B, g.maxb, g.spin, c->spec_mode, nullptr, c->d_eval, c->d_evec,
                                c->d_omega, (int)E, eta, c->d_spos, c->d_ldos, c->stream);
        }
        stage_end(c, st);
        if (r) return fail(r, "spectral-function launch failed");
        if (c->surf_mode) {
            HIPCHK(hipStreamSynchronize(c->stream));
            RETCHK(check_surface_steps(c, E));
        }
        c->nlayers = E;
        c->windows_kept = false;
        c->last_omegas.assign(omegas, omegas + E);
        c->last_eta = eta;
        stage_collect(c);
        if (ldos_host)
            HIPCHK(hipMemcpy(ldos_host, c->d_ldos, E * (size_t)g.W * g.W * sizeof(double), hipMemcpyDeviceToHost));
        return 0;
    }
    RunPlan plan;
    RETCHK(prepare_run(c, omegas, E, eta, plan));
    RETCHK(enqueue_chunks(c, plan));
    RETCHK(finish_run(c, plan));
    stage_collect(c);
    const size_t WW = (size_t)c->g.W * c->g.W;
    if (ldos_host) HIPCHK(hipMemcpy(ldos_host, c->d_ldos, E * WW * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int cqpi_set_josephson(cqpi_ctx *c, int on, double deltat, double etat)
{
    if (!c) return fail(CQPI_E_ARG, "ctx is NULL");
    if (on && (c->list_mode || c->spectral))
        return fail(CQPI_E_ARG, "the Josephson map needs a context made by cqpi_create");
    if (on && 2 * c->g.maxb > c->g.B) return fail(CQPI_E_ARG, "the Josephson map needs 2*maxband <= bands");
    c->josephson = on != 0;
    c->jos_deltat = deltat;
    c->jos_etat = std::isnan(etat) ? -1.0 : etat;
    return 0;
}

int cqpi_set_keep_windows(cqpi_ctx *c, int on)
{
    if (!c) return fail(CQPI_E_ARG, "ctx is NULL");
    c->keep_windows = on != 0;
    return 0;
}

static int gradient_impl(cqpi_ctx *c, const double *rhobar, bool on_device, double *g_hops, double *g_mu,
                         double *g_V);

int cqpi_gradient(cqpi_ctx *c, const double *rhobar, double *g_hops, double *g_mu, double *g_V)
{
    return gradient_impl(c, rhobar, false, g_hops, g_mu, g_V);
}

int cqpi_gradient_device(cqpi_ctx *c, const double *d_rhobar, double *g_hops, double *g_mu, double *g_V)
{
    return gradient_impl(c, d_rhobar, true, g_hops, g_mu, g_V);
}

static int gradient_impl(cqpi_ctx *c, const double *rhobar, bool on_device, double *g_hops, double *g_mu,
                         double *g_V)
{
    g_err.clear();
    if (!c || !rhobar) return fail(CQPI_E_ARG, "NULL argument");
    if (!c->windows_kept || c->nlayers == 0)
        return fail(CQPI_E_STATE, "the last cqpi_run did not keep its windows (cqpi_set_keep_windows)");
    if (c->maps_josephson) return fail(CQPI_E_ARG, "the gradient is of the LDOS map, not the Josephson map");
    if (c->surf_mode) return fail(CQPI_E_ARG, "the gradient needs G0(k) from the eigen data, not the surface Green's function");
    if (c->list_mode) return fail(CQPI_E_ARG, "the gradient is of the dense LDOS map, not a threshold-list map");
    const Geom &g = c->g;
    if (!adjoint_supported(g))
        return fail(CQPI_E_ARG, "the adjoint needs a power-of-two k-grid (<= 1024), at most 5 bands and window <= 4");
    const int B = g.B, NP = B * B;
    const size_t E = c->nlayers, WW = (size_t)g.W * g.W;
    const int nsets = g.spin ? 2 : 1;
    // distinct lattice vectors of the hopping phases (sign folded for 'anti'
    // terms), as (rx, index of ry among the distinct ry values)
    std::vector<double> ryv;
    std::vector<double2> R;
    std::vector<int> hop_r(c->h_hops.size());
    for (size_t i = 0; i < c->h_hops.size(); i++) {
        const cqpi_hopping &h = c->h_hops[i];
        double rx = (h.anti ? -h.rx : h.rx) + 0.0, ry = (h.anti ? -h.ry : h.ry) + 0.0;
        int iy = -1;
        for (size_t j = 0; j < ryv.size(); j++)
            if (ryv[j] == ry) iy = (int)j;
        if (iy < 0) {
            iy = (int)ryv.size();
            ryv.push_back(ry);
        }
        int idx = -1;
        for (size_t j = 0; j < R.size(); j++)
            if (R[j].x == rx && R[j].y == (double)iy) idx = (int)j;
        if (idx < 0) {
            idx = (int)R.size();
            R.push_back(make_double2(rx, (double)iy));
        }
        hop_r[i] = idx;
    }
    if (ryv.size() > 16) return fail(CQPI_E_ARG, "the adjoint handles at most 16 distinct hopping y offsets");
    const int nR = (int)R.size(), nry = (int)ryv.size();
    R.push_back(make_double2(0.0, (double)nry));  // plain sums, for the trace (mu)
    if (ryv.empty()) ryv.push_back(0.0);
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!c->d_twn) {
        std::vector<double2> tw(g.K);
        for (int m = 0; m < g.K; m++) {
            long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)g.K;
            tw[m] = make_double2((double)cosl(a), (double)sinl(a));
        }
        RETCHK(dev_grow(c->d_twn, c->c_twn, (size_t)g.K, c->gpumem));
        HIPCHK(hipMemcpy(c->d_twn, tw.data(), g.K * sizeof(double2), hipMemcpyHostToDevice));
    }
    const size_t nblk = adjoint_pix_blocks(g);
    const size_t NG = (size_t)adjoint_energy_groups(g, (int)E);
    if (!on_device) RETCHK(dev_grow(c->d_rhobar, c->c_rhobar, E * WW, c->gpumem));
    // the hopping vectors, y offsets and energies are uploaded only when they
    // differ from the last upload (the energies normally are the forward run's)
    auto same2 = [](const std::vector<double2> &a, const std::vector<double2> &b) {
        if (a.size() != b.size()) return false;
        for (size_t i = 0; i < a.size(); i++)
            if (a[i].x != b[i].x || a[i].y != b[i].y) return false;
        return true;
    };
    const bool newR = c->c_Rlist < R.size() || !same2(R, c->up_R);
    const bool newry = c->c_ry < ryv.size() || ryv != c->up_ry;
    const bool newom = c->c_omega < E || c->up_omegas != c->last_omegas;
    RETCHK(dev_grow(c->d_ry, c->c_ry, ryv.size(), c->gpumem));
    RETCHK(dev_grow(c->d_U, c->c_U, (size_t)g.K * (nry + 1) * NP + (size_t)nry * g.K, c->gpumem));
    RETCHK(dev_grow(c->d_abar, c->c_abar, E * nsets * WW * B, c->gpumem));
    RETCHK(dev_grow(c->d_bbar, c->c_bbar, E * nsets * WW * B, c->gpumem));
    RETCHK(dev_grow(c->d_Tpart, c->c_Tpart, E * nblk * NP, c->gpumem));
    RETCHK(dev_grow(c->d_Tbar, c->c_Tbar, E * NP, c->gpumem));
    RETCHK(dev_grow(c->d_rhosum, c->c_rhosum, E * g.os * g.os, c->gpumem));
    // Gbar in two halves: the a and the b gather of the window adjoint
    RETCHK(dev_grow(c->d_Gbar, c->c_Gbar, 2 * E * (size_t)g.nw * g.nw * NP, c->gpumem));
    RETCHK(dev_grow(c->d_Vbar, c->c_Vbar, E * NP, c->gpumem));
    RETCHK(dev_grow(c->d_F1, c->c_F1, E * (size_t)g.K * g.nw * NP, c->gpumem));
    RETCHK(dev_grow(c->d_Hacc, c->c_Hacc, NG * g.K * g.K * NP, c->gpumem));
    RETCHK(dev_grow(c->d_Rlist, c->c_Rlist, R.size(), c->gpumem));
    RETCHK(dev_grow(c->d_SR, c->c_SR, R.size() * NP, c->gpumem));
    RETCHK(dev_grow(c->d_Q, c->c_Q, adjoint_q_size(g), c->gpumem));
    // the Gbar(k) workspace of the many-line ky adjoint shares the G(k) workspace
    const size_t gkb = adjoint_workspace_bytes(g, (int)E);
    if (gkb) RETCHK(dev_grow(c->d_gkw, c->c_gkw, gkb / sizeof(double2), c->gpumem));
    const double *drb = rhobar;
    if (!on_device) {
        HIPCHK(hipMemcpyAsync(c->d_rhobar, rhobar, E * WW * sizeof(double), hipMemcpyHostToDevice, c->stream));
        drb = c->d_rhobar;
    }
    if (newR) {
        c->up_R = R;
        HIPCHK(hipMemcpyAsync(c->d_Rlist, c->up_R.data(), R.size() * sizeof(double2), hipMemcpyHostToDevice,
                              c->stream));
    }
    if (newry) {
        c->up_ry = ryv;
        HIPCHK(hipMemcpyAsync(c->d_ry, c->up_ry.data(), ryv.size() * sizeof(double), hipMemcpyHostToDevice,
                              c->stream));
    }
    if (newom) {
        RETCHK(dev_grow(c->d_omega, c->c_omega, E, c->gpumem));
        c->up_omegas = c->last_omegas;
        HIPCHK(hipMemcpyAsync(c->d_omega, c->up_omegas.data(), E * sizeof(double), hipMemcpyHostToDevice,
                              c->stream));
    }
    RETCHK(launch_adjoint(g, c->d_G0w, c->d_T, c->d_V, c->d_tiles, drb, c->d_eval, c->d_evec, c->d_omega, (int)E,
                          c->last_eta, c->d_twn, c->d_ry, nry, c->d_Rlist, (int)R.size(), c->d_abar, c->d_bbar,
                          c->d_Tpart, c->d_Tbar, c->d_rhosum, c->d_Gbar, c->d_Vbar, c->d_F1, c->d_Hacc, c->d_U,
                          c->d_SR, c->d_Q, c->d_gkw, c->c_gkw * sizeof(double2), c->stream));
    const size_t nsr = R.size() * NP, ngrad = nsr + E * NP;
    if (c->c_hgrad < ngrad) {
        if (c->h_grad) (void)hipHostFree(c->h_grad);
        c->h_grad = nullptr;
        c->c_hgrad = 0;
        HIPCHK(hipHostMalloc((void **)&c->h_grad, ngrad * sizeof(double2), hipHostMallocDefault));
        c->c_hgrad = ngrad;
    }
    HIPCHK(hipMemcpyAsync(c->h_grad, c->d_SR, nsr * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(c->h_grad + nsr, c->d_Vbar, E * NP * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const double2 *SR = c->h_grad, *Vb = c->h_grad + nsr;
    // (the adjoint reused the F1 buffer; the kept windows and T stay valid)
    if (g_hops)
        for (size_t i = 0; i < c->h_hops.size(); i++) {
            const double2 s = SR[(size_t)hop_r[i] * NP + c->h_hops[i].o1 * B + c->h_hops[i].o2];
            g_hops[2 * i] = s.x;
            g_hops[2 * i + 1] = s.y;
        }
    if (g_mu) {  // dH/dmu = -1 on the particle block, +1 on the hole block
        double s = 0.0;
        for (int d = 0; d < B; d++) s += (d < c->scmax ? -1.0 : 1.0) * SR[(size_t)nR * NP + d * B + d].x;
        *g_mu = s;
    }
    if (g_V)
        for (int p = 0; p < NP; p++) {
            double re = 0.0, im = 0.0;
            for (size_t e = 0; e < E; e++) {
                re += Vb[e * NP + p].x;
                im += Vb[e * NP + p].y;
            }
            g_V[2 * p] = re;
            g_V[2 * p + 1] = im;
        }
    return 0;
}

int cqpi_window_half(cqpi_ctx *c, size_t *h)
{
    if (!c || !h) return fail(CQPI_E_ARG, "NULL argument");
    *h = (size_t)c->g.h;
    return 0;
}

int cqpi_retrieve_g0_window(cqpi_ctx *c, size_t e, double *out)
{
    g_err.clear();
    if (!c || !out) return fail(CQPI_E_ARG, "NULL argument");
    if (e < c->last_chunk_e0 || e >= c->last_chunk_e0 + c->last_chunk_ec)
        return fail(CQPI_E_STATE, "G0 window is kept for the last energy chunk only");
    if (!c->window_whole)
        return fail(CQPI_E_STATE, "a mixed-precision run writes only |R| <= 2w of the G0 window; "
                                  "call cqpi_set_keep_windows(ctx, 1) before the run to keep it whole");
    size_t nwin = (size_t)c->g.nw * c->g.nw * c->g.B * c->g.B;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(out, c->d_G0w + (e - c->last_chunk_e0) * nwin, nwin * sizeof(double2), hipMemcpyDeviceToHost));
    return 0;
}

int cqpi_retrieve_tmatrix(cqpi_ctx *c, size_t e, double *out)
{
    g_err.clear();
    if (!c || !out) return fail(CQPI_E_ARG, "NULL argument");
    if (e >= c->nlayers) return fail(CQPI_E_STATE, "energy index beyond the last run");
    size_t NP = (size_t)c->g.B * c->g.B;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(out, c->d_T + e * NP, NP * sizeof(double2), hipMemcpyDeviceToHost));
    return 0;
}

void cqpi_free(void *p) { free(p); }

}  // extern "C"

/* ------------------------------------------------------------------------ */
/* Autotuning                                                                */
/* ------------------------------------------------------------------------ */

namespace {

// best launch tuning per problem shape, shared by every context of the process
std::mutex g_tune_mu;
std::map<std::string, std::array<int, 4>> g_tune;  // tpw, eb, js, streams

std::string tune_key(const cqpi_ctx *c, size_t E)
{
    const Geom &g = c->g;
    char k[160];
    snprintf(k, sizeof k, "K%d n%d w%d os%d b%d m%d s%d p%d E%zu", g.K, g.n, g.w, g.os, g.B, g.maxb, g.spin,
             c->precision, E);
    return k;
}

// device time of one prepared run (eigen data cached) with the context's tuning
int time_run(cqpi_ctx *c, const RunPlan &plan, int reps, float &ms)
{
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    int r = enqueue_chunks(c, plan);  // warm-up
    if (!r && hipEventRecord(a, c->stream) != hipSuccess) r = CQPI_E_HIP;
    for (int i = 0; i < reps && !r; i++) r = enqueue_chunks(c, plan);
    if (!r && hipEventRecord(b, c->stream) != hipSuccess) r = CQPI_E_HIP;
    if (!r && hipEventSynchronize(b) != hipSuccess) r = CQPI_E_HIP;
    if (!r && hipEventElapsedTime(&ms, a, b) != hipSuccess) r = CQPI_E_HIP;
    ms /= std::max(reps, 1);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return r ? fail(r, "autotune timing failed") : 0;
}

}  // namespace

extern "C" {

int cqpi_set_tuning(cqpi_ctx *c, int tpw, int eb)
{
    if (!c) return fail(CQPI_E_ARG, "ctx is NULL");
    if (tpw != 0 && tpw != 2 && tpw != 4 && tpw != 8) return fail(CQPI_E_ARG, "tpw must be 0, 2, 4 or 8");
    if (eb < 0) return fail(CQPI_E_ARG, "eb must be >= 0");
    c->g.tpw = tpw;
    c->g.eb = eb;
    return 0;
}

int cqpi_set_tuning_ext(cqpi_ctx *c, int tpw, int eb, int js, int streams)
{
    RETCHK(cqpi_set_tuning(c, tpw, eb));
    if (js < 0 || js > 2) return fail(CQPI_E_ARG, "js must be 0, 1 or 2");
    if (streams < 0 || streams > 2) return fail(CQPI_E_ARG, "streams must be 0, 1 or 2");
    c->g.js = js;
    c->g.streams = streams;
    return 0;
}

int cqpi_get_tuning_ext(cqpi_ctx *c, int *tpw, int *eb, int *js, int *streams)
{
    if (!c) return fail(CQPI_E_ARG, "ctx is NULL");
    if (tpw) *tpw = c->g.tpw;
    if (eb) *eb = c->g.eb;
    if (js) *js = c->g.js;
    if (streams) *streams = c->g.streams;
    return 0;
}

int cqpi_get_tuning(cqpi_ctx *c, int *tpw, int *eb)
{
    if (!c) return fail(CQPI_E_ARG, "ctx is NULL");
    if (tpw) *tpw = c->g.tpw;
    if (eb) *eb = c->g.eb;
    return 0;
}

int cqpi_autotune(cqpi_ctx *c, const double *omegas, size_t E, double eta, int reps, int *tpw, int *eb)
{
    g_err.clear();
    if (!c || !omegas || !E) return fail(CQPI_E_ARG, "NULL argument or no energies");
    if (!c->has_model) return fail(CQPI_E_STATE, "cqpi_set_model has not been called");
    const std::string key = tune_key(c, E);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) {
            c->g.tpw = it->second[0];
            c->g.eb = it->second[1];
            c->g.js = it->second[2];
            c->g.streams = it->second[3];
            if (tpw) *tpw = c->g.tpw;
            if (eb) *eb = c->g.eb;
            return 0;
        }
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    stage_collect(c);
    const bool timing = c->timing;
    c->timing = false;
    int r = 0;
    if (!c->eigen_valid) {
        c->in_run = true;
        r = cqpi_compute_eigen(c);
        c->in_run = false;
    }
    RunPlan plan;
    if (!r) r = prepare_run(c, omegas, E, eta, plan);
    // energies per workgroup of the first transform pass, then (mixed) cell
    // rows per wave of the contraction; each knob is timed with the other fixed
    int best_eb = 0, best_tpw = 0;
    float best = 1e30f;
    const int ebs[] = {0, (int)plan.Ec, (int)((plan.Ec + 1) / 2), (int)((plan.Ec + 3) / 4), (int)((plan.Ec + 7) / 8)};
    for (int cand : ebs) {
        if (r) break;
        c->g.eb = cand;
        c->g.tpw = 0;
        c->g.js = 0;
        c->g.streams = 0;
        float ms = 0.f;
        r = time_run(c, plan, reps, ms);
        if (!r && ms < best) {
            best = ms;
            best_eb = cand;
        }
    }
    if (!r && plan.mixed) {
        const int tpws[] = {4, 2};
        for (int cand : tpws) {
            c->g.eb = best_eb;
            c->g.tpw = cand;
            float ms = 0.f;
            r = time_run(c, plan, reps, ms);
            if (r) break;
            if (ms < best) {
                best = ms;
                best_tpw = cand;
            }
        }
    }
    // waves per cell row of the T-folded contraction (its default is per shape)
    int best_js = 0;
    if (!r && plan.mixed && !c->josephson && mix_tfold(c->g, c->mg)) {
        const int jss[] = {1, 2};
        for (int cand : jss) {
            c->g.eb = best_eb;
            c->g.tpw = best_tpw;
            c->g.js = cand;
            float ms = 0.f;
            r = time_run(c, plan, reps, ms);
            if (r) break;
            if (ms < best) {
                best = ms;
                best_js = cand;
            }
        }
    }
    // one stream instead of the split over two (where a run would split)
    int best_streams = 0;
    if (!r && plan.Ec >= 8 && !plan.work_bytes && !c->single_stream) {
        c->g.eb = best_eb;
        c->g.tpw = best_tpw;
        c->g.js = best_js;
        c->g.streams = 1;
        float ms = 0.f;
        r = time_run(c, plan, reps, ms);
        if (!r && ms < best) {
            best = ms;
            best_streams = 1;
        }
    }
    c->timing = timing;
    c->g.eb = best_eb;
    c->g.tpw = best_tpw;
    c->g.js = best_js;
    c->g.streams = best_streams;
    if (!r) r = finish_run(c, plan);
    if (r) return r;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune[key] = {best_tpw, best_eb, best_js, best_streams};
    }
    if (tpw) *tpw = best_tpw;
    if (eb) *eb = best_eb;
    return 0;
}

}  // extern "C"

/* ------------------------------------------------------------------------ */
/* Batches: many independent instances in one captured graph               */
/* ------------------------------------------------------------------------ */

struct cqpi_batch {
    std::vector<cqpi_ctx *> inst;
    std::vector<std::vector<double>> omegas;
    std::vector<double> eta;
    std::vector<RunPlan> plans;
    std::vector<int> branch_of;
    hipStream_t main = nullptr;
    std::vector<hipStream_t> branches;
    hipEvent_t fork = nullptr, t0 = nullptr, t1 = nullptr;
    std::vector<hipEvent_t> joins;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    size_t nodes = 0;
    float last_ms = 0.f;
};

namespace {

constexpr int kBatchBranches = 4;

void batch_release(cqpi_batch *b)
{
    if (!b) return;
    if (b->main) (void)hipStreamSynchronize(b->main);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (hipStream_t s : b->branches)
        if (s) (void)hipStreamDestroy(s);
    for (hipEvent_t e : b->joins)
        if (e) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->t0) (void)hipEventDestroy(b->t0);
    if (b->t1) (void)hipEventDestroy(b->t1);
    if (b->main) (void)hipStreamDestroy(b->main);
    for (cqpi_ctx *c : b->inst) release(c);
    delete b;
}

// relative device work of an instance: transforms ~ E K^2 b^2 log K, contraction ~ E W^2 (2w+1)^2 b^2
double batch_cost(const cqpi_ctx *c, size_t E)
{
    const Geom &g = c->g;
    double side = 2.0 * g.w + 1.0;
    return (double)E * ((double)g.K * g.K * g.B * g.B * std::log2((double)g.K + 1.0) +
                        (double)g.W * g.W * side * side * g.B * g.B);
}

// eager run of every instance (allocates its buffers, validates it), then
// capture of all their launches into one graph: instances spread over
// kBatchBranches parallel branches, heaviest first onto the least loaded one
int batch_capture(cqpi_batch *b)
{
    const size_t n = b->inst.size();
    b->plans.assign(n, RunPlan());
    b->branch_of.assign(n, 0);
    for (size_t i = 0; i < n; i++) {
        cqpi_ctx *c = b->inst[i];
        RETCHK(cqpi_run(c, b->omegas[i].data(), b->omegas[i].size(), b->eta[i], nullptr));
        RETCHK(prepare_run(c, b->omegas[i].data(), b->omegas[i].size(), b->eta[i], b->plans[i]));
    }
    std::vector<size_t> order(n);
    for (size_t i = 0; i < n; i++) order[i] = i;
    std::vector<double> cost(n);
    for (size_t i = 0; i < n; i++) cost[i] = batch_cost(b->inst[i], b->omegas[i].size());
    std::sort(order.begin(), order.end(), [&](size_t x, size_t y) { return cost[x] > cost[y]; });
    const int nb = (int)b->branches.size();
    std::vector<double> load(nb, 0.0);
    HIPCHK(hipStreamBeginCapture(b->main, hipStreamCaptureModeThreadLocal));
    int r = 0;
    if (hipEventRecord(b->fork, b->main) != hipSuccess) r = CQPI_E_HIP;
    for (int k = 0; k < nb && !r; k++)
        if (hipStreamWaitEvent(b->branches[k], b->fork, 0) != hipSuccess) r = CQPI_E_HIP;
    for (size_t oi = 0; oi < n && !r; oi++) {
        const size_t i = order[oi];
        const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[k] += cost[i];
        b->branch_of[i] = k;
        cqpi_ctx *c = b->inst[i];
        hipStream_t own = c->stream;
        const bool timing = c->timing;
        c->timing = false;
        c->stream = b->branches[k];
        const Geom &g = c->g;
        r = launch_eigen(g.B, c->d_groups, c->ngroups, c->d_tmat, c->d_tw, c->efermi, c->scmax, g.K, c->d_eval, c->d_evec,
                         c->stream);
        if (!r) r = enqueue_chunks(c, b->plans[i]);
        c->stream = own;
        c->timing = timing;
    }
    for (int k = 0; k < nb && !r; k++) {
        if (hipEventRecord(b->joins[k], b->branches[k]) != hipSuccess ||
            hipStreamWaitEvent(b->main, b->joins[k], 0) != hipSuccess)
            r = CQPI_E_HIP;
    }
    hipGraph_t graph = nullptr;
    hipError_t ec = hipStreamEndCapture(b->main, &graph);
    if (r || ec != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return fail(r ? r : CQPI_E_HIP, "batch graph capture failed");
    }
    b->graph = graph;
    HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    size_t nodes = 0;
    HIPCHK(hipGraphGetNodes(b->graph, nullptr, &nodes));
    b->nodes = nodes;
    return 0;
}

}  // namespace

extern "C" {

int cqpi_batch_create(const cqpi_instance *in, size_t count, int precision, cqpi_batch **out)
{
    g_err.clear();
    if (!out || (!in && count)) return fail(CQPI_E_ARG, "NULL argument");
    *out = nullptr;
    if (!count) return fail(CQPI_E_ARG, "empty batch");
    cqpi_batch *b = new cqpi_batch();
    int r = 0;
    for (size_t i = 0; i < count && !r; i++) {
        const cqpi_instance &d = in[i];
        if (!d.omegas || !d.E || !d.V) {
            r = fail(CQPI_E_ARG, "instance " + std::to_string(i) + ": omegas, E and V are required");
            break;
        }
        cqpi_ctx *c = nullptr;
        r = cqpi_create(d.oversamp * d.n, d.kpoints, d.n, d.window, d.oversamp, d.bands, d.maxband, d.spin, d.wf,
                        &c);
        if (r) break;
        b->inst.push_back(c);
        c->single_stream = true;
        r = cqpi_set_precision(c, precision);
        if (!r) r = cqpi_set_model(c, d.hops, d.nhops, d.efermi, d.scmodel);
        if (!r) r = cqpi_set_impurity(c, d.V);
        if (!r) {
            b->omegas.emplace_back(d.omegas, d.omegas + d.E);
            b->eta.push_back(d.eta);
        }
    }
    if (!r && hipStreamCreateWithFlags(&b->main, hipStreamNonBlocking) != hipSuccess) r = fail(CQPI_E_HIP, "stream");
    for (int k = 0; k < kBatchBranches && !r; k++) {
        hipStream_t s = nullptr;
        hipEvent_t e = nullptr;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            r = fail(CQPI_E_HIP, "stream/event");
        b->branches.push_back(s);
        b->joins.push_back(e);
    }
    if (!r && (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
               hipEventCreate(&b->t0) != hipSuccess || hipEventCreate(&b->t1) != hipSuccess))
        r = fail(CQPI_E_HIP, "event");
    if (r) {
        std::string msg = g_err;
        batch_release(b);
        g_err = msg;
        return r;
    }
    *out = b;
    return 0;
}

int cqpi_batch_autotune(cqpi_batch *b, int reps)
{
    g_err.clear();
    if (!b) return fail(CQPI_E_ARG, "batch is NULL");
    if (b->exec) return fail(CQPI_E_STATE, "the batch graph is already captured");
    for (size_t i = 0; i < b->inst.size(); i++) {
        int tpw = 0, eb = 0;
        RETCHK(cqpi_autotune(b->inst[i], b->omegas[i].data(), b->omegas[i].size(), b->eta[i], reps, &tpw, &eb));
    }
    return 0;
}

int cqpi_batch_run(cqpi_batch *b)
{
    g_err.clear();
    if (!b) return fail(CQPI_E_ARG, "batch is NULL");
    if (!b->exec) RETCHK(batch_capture(b));
    HIPCHK(hipEventRecord(b->t0, b->main));
    HIPCHK(hipGraphLaunch(b->exec, b->main));
    HIPCHK(hipEventRecord(b->t1, b->main));
    HIPCHK(hipStreamSynchronize(b->main));
    HIPCHK(hipEventElapsedTime(&b->last_ms, b->t0, b->t1));
    for (size_t i = 0; i < b->inst.size(); i++) {
        cqpi_ctx *c = b->inst[i];
        const size_t E = b->omegas[i].size();
        std::vector<int> status(E);
        HIPCHK(hipMemcpy(status.data(), c->d_status, E * sizeof(int), hipMemcpyDeviceToHost));
        for (size_t e = 0; e < E; e++)
            if (status[e])
                return fail(CQPI_E_SINGULAR, "instance " + std::to_string(i) + ": 1 - V G0(0) is singular at energy index " +
                                                 std::to_string(e));
        c->nlayers = E;
    }
    return 0;
}

int cqpi_batch_size(cqpi_batch *b, size_t *count)
{
    if (!b || !count) return fail(CQPI_E_ARG, "NULL argument");
    *count = b->inst.size();
    return 0;
}

int cqpi_batch_context(cqpi_batch *b, size_t i, cqpi_ctx **ctx)
{
    if (!b || !ctx || i >= b->inst.size()) return fail(CQPI_E_ARG, "bad batch index");
    *ctx = b->inst[i];
    return 0;
}

int cqpi_batch_info(cqpi_batch *b, double *last_ms, size_t *graph_nodes, size_t *gpumem)
{
    if (!b) return fail(CQPI_E_ARG, "batch is NULL");
    if (last_ms) *last_ms = b->last_ms;
    if (graph_nodes) *graph_nodes = b->nodes;
    if (gpumem) {
        size_t m = 0;
        for (cqpi_ctx *c : b->inst) m += c->gpumem;
        *gpumem = m;
    }
    return 0;
}

void cqpi_batch_destroy(cqpi_batch *b) { batch_release(b); }

}  // extern "C"
