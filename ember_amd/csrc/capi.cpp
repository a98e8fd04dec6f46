// C-ABI of the ember_amd library (declared in include/ember_amd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ember_amd.h"
#include "emx_common.h"
#include "emx_math.h"
#include "kernels.h"
#include "mech_host.h"

using namespace emx;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

// A failed HIP call is reported here; the runtime's sticky last-error is cleared so that
// the launch checks of later calls (hipGetLastError after a launch) report their own
// launches only.
#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) {                                                         \
            (void)hipGetLastError();                                                    \
            return fail(-10, std::string(#expr) + ": " + hipGetErrorString(e_));        \
        }                                                                               \
    } while (0)

// growable device buffer
struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= n) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc(&p, bytes);
        if (e == hipSuccess) n = bytes;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

}  // namespace

struct emx_context {
    int device = 0;
    int num_cu = 256;
    hipStream_t stream = nullptr;
    double P = 101325.0;
    emx_options opts{};
    DevOpts dopts{};
    HostMech hm;
    DevLayout lay;
    DevMech dm{};
    void* blob = nullptr;
    double* perturb = nullptr;  // device, slot order
    unsigned int* queue = nullptr;
    int waves_per_block = 4;
    bool waves_set = false;  // waves_per_block chosen by the caller (else the default, which
                             // the BDF lowers to what fits in LDS)
    int blocks_per_cu = 0;  // 0: as many as LDS allows
    DevBuf b_in, b_out, b_sc, b_ql, b_status, b_s1, b_s2;
    DevBuf d_B, d_Brow, d_D, d_sc, d_x, d_y, d_h, d_ns, d_work, d_hmin;
    DevBuf t_a, t_b, t_c, t_k;
    DevBuf jac, plain_ws;
    int variant = EMX_VARIANT_WAVE;
    // cost-ordered scheduling: per-point cost of the last stage and the order built from it
    int schedule = EMX_SCHED_COST;
    DevBuf cost, order;
    size_t cost_n = 0;  // number of points the recorded costs belong to (0: none)
    // adjoint: checkpoints of the last adjoint forward stage, per-wave records
    DevBuf ck_y, ck_s, nckpt, record;
    int adj_cap = 64;       // checkpoints per point
    size_t adj_n = 0;       // points of the last adjoint forward stage (0: none)
    double adj_dt = 0.0;
    // rate multiplier function: [T0, T1, c_0 .. c_{n-1}] on the device, n = ratefn_n (0: none);
    // while set, the kernels take unit per-slot multipliers (`ones`) times a(t)
    DevBuf ratefn, ones;
    int ratefn_n = 0;
    // igniter: point positions (host) and the per-point pulse amplitudes (device), rebuilt
    // when the positions or the ignition options change
    std::vector<double> xpos;
    DevBuf qign;
    bool qign_valid = false;
    // InterpKinetics table grid (Ember's defaults, grown by emx_set_interp_range or by the
    // host-array stage entry point)
    // convection stage work arrays
    DevBuf cv_work, cv_rec_t, cv_rec_v, cv_ints, cv_aux, cv_yleft;
    std::vector<double> cv_yleft_host;  // the inflow composition cv_yleft holds
    int interp_n = kInterpN;
    double interp_tmin = kInterpTmin, interp_tmax = kInterpTmax;
};

// ---------------------------------------------------------------------------
// Scratch of the allocation-free production stage (emx_source_integrate_device_noalloc):
// bytes [0, 256) hold the work-queue counter, then kNoAllocMaxPoints per-point costs and
// kNoAllocMaxPoints work-queue positions (int32 each): 2 MiB + 256 B per GPU, held until
// the process ends.  Host code never uses the array's name as an address: each device's
// copy is reached through the pointer hipGetSymbolAddress returned for it.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocMaxPoints = 262144;
constexpr size_t kScratchBytes = 256 + 2 * kNoAllocMaxPoints * sizeof(int);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

struct DeviceScratch {
    unsigned char* ptr = nullptr;  // this device's copy of `scratch`
    size_t last_n = 0;             // points of the last allocation-free stage (0: none)
};
std::mutex g_scratch_mu;
std::map<int, DeviceScratch> g_scratch;

// Looks up the device copy of `scratch` once per device; the device must be current.
int scratch_setup(int device, std::string& err) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device].ptr) return 0;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return -10;
    }
    g_scratch[device].ptr = static_cast<unsigned char*>(p);
    return 0;
}

DevOpts to_dev(const emx_options& o) {
    DevOpts d;
    d.qss_epsmin = o.qss_epsmin;
    d.qss_epsmax = o.qss_epsmax;
    d.qss_dtmin = o.qss_dtmin;
    d.qss_dtmax = o.qss_dtmax;
    d.qss_itermax = o.qss_iteration_count;
    d.qss_abstol = o.qss_abstol;
    d.qss_minval = o.qss_minval;
    d.qss_stability = o.qss_stability_check;
    d.rtol = o.integrator_rel_tol;
    d.atol_momentum = o.integrator_momentum_abs_tol;
    d.atol_energy = o.integrator_energy_abs_tol;
    d.atol_species = o.integrator_species_abs_tol;
    d.min_step = o.integrator_min_timestep;
    d.max_steps = o.integrator_max_steps;
    d.ratefn = nullptr;
    d.ratefn_n = 0;
    d.t_start = 0.0;
    d.qign = nullptr;
    d.ign_t0 = o.ignition_t_start;
    d.ign_t1 = o.ignition_t_start + o.ignition_duration;  // as getQdotIgniter forms the bound
    return d;
}

// pack host arrays into one device allocation (256-byte aligned pieces)
struct Packer {
    std::vector<char> bytes;
    std::vector<std::pair<size_t, const void**>> fix;
    template <class T>
    void add(const std::vector<T>& v, const T** dst) {
        size_t off = (bytes.size() + 255) & ~size_t(255);
        bytes.resize(off + std::max<size_t>(v.size(), 1) * sizeof(T), 0);
        if (!v.empty()) memcpy(bytes.data() + off, v.data(), v.size() * sizeof(T));
        fix.emplace_back(off, reinterpret_cast<const void**>(dst));
    }
};

int upload_mech(emx_context* c) {
    const HostMech& m = c->hm;
    const DevLayout& L = c->lay;
    DevMech& d = c->dm;
    d.K = m.K;
    d.I = m.I;
    d.nVars = m.K + 2;
    d.nRounds = L.nRounds;
    d.nFallRounds = L.nFallRounds;
    d.nGroups = L.nGroups;
    d.nChunks = L.nChunks;
    d.nChunkRounds = L.nChunkRounds;
    d.nEntries = (int)L.entry.size();
    d.nGrpEntries = (int)L.grp_sp.size();
    d.grpMaxLen = 0;
    for (int g = 0; g < L.nGroups; g++) d.grpMaxLen = std::max(d.grpMaxLen, L.grp_off[g + 1] - L.grp_off[g]);
    d.nGrpSlots = (int)L.grp_slot.size();
    d.nSpGrp = (int)L.sp_grp.size();
    d.grpSlotMax = d.spGrpMax = 0;
    for (int g = 0; g < L.nGroups; g++) d.grpSlotMax = std::max(d.grpSlotMax, L.grp_slot_off[g + 1] - L.grp_slot_off[g]);
    for (int k = 0; k < m.K; k++) d.spGrpMax = std::max(d.spGrpMax, L.sp_grp_off[k + 1] - L.sp_grp_off[k]);
    // slots with a third-body group come first in device order (falloff, three-body)
    int last_tb = 0;
    for (int s = 0; s < m.I; s++)  // padding slots (kNone) read as group 0: not counted
        if (info_group(L.info[s]) < L.nGroups) last_tb = s;
    d.nTbRounds = last_tb / 64 + 1;
    d.interp = c->opts.kinetics_model == EMX_KIN_INTERP ? 1 : 0;
    d.P = c->P;
    d.Pref = m.Pref;
    d.lnRPref = std::log(kGasConstant / m.Pref);
    std::vector<double> kf_c, kf_s, kc_c, kc_s;
    d.interpN = c->interp_n;
    d.interpTmin = c->interp_tmin;
    d.interpTmax = c->interp_tmax;
    if (d.interp)
        build_interp_tables(m, L, c->P, c->interp_n, c->interp_tmin, c->interp_tmax, kf_c, kf_s, kc_c, kc_s);
    std::vector<int32_t> aux;
    std::vector<double> auxd;
    build_pressure_tables(m, L, c->P, aux, auxd);
    d.has_pdep = 0;
    for (const auto& r : m.rxn) d.has_pdep |= (r.type == 4 || r.type == 5);
    d.has_orders = 0;
    for (const auto& r : m.rxn) d.has_orders |= r.hasOrd;
    Packer pk;
    pk.add(m.W, &d.W);
    pk.add(L.rW, &d.rW);
    pk.add(m.Tmid, &d.Tmid);
    pk.add(m.nasa, &d.nasa);
    pk.add(std::vector<double>(kExp2Tab, kExp2Tab + 16), &d.exp2t);
    pk.add(L.A, &d.A);
    pk.add(L.b, &d.b);
    pk.add(L.EaR, &d.EaR);
    pk.add(L.dn, &d.dn);
    pk.add(L.A0, &d.A0);
    pk.add(L.b0, &d.b0);
    pk.add(L.EaR0, &d.EaR0);
    pk.add(L.troeA, &d.troeA);
    pk.add(L.rT3, &d.rT3);
    pk.add(L.rT1, &d.rT1);
    pk.add(L.T2, &d.T2);
    pk.add(L.reac, &d.reac);
    pk.add(L.prod, &d.prod);
    pk.add(L.info, &d.info);
    pk.add(L.ord, &d.ord);
    pk.add(L.orig, &d.orig);
    pk.add(L.grp_off, &d.grp_off);
    pk.add(L.grp_sp, &d.grp_sp);
    pk.add(L.grp_eff, &d.grp_eff);
    pk.add(L.grp_def, &d.grp_def);
    pk.add(L.entry, &d.entry);
    pk.add(L.sp_chunk0, &d.sp_chunk0);
    pk.add(L.sp_nchunk, &d.sp_nchunk);
    pk.add(L.grp_slot_off, &d.grp_slot_off);
    pk.add(L.grp_slot, &d.grp_slot);
    pk.add(L.sp_grp_off, &d.sp_grp_off);
    pk.add(L.sp_grp, &d.sp_grp);
    pk.add(L.sp_grp_eff, &d.sp_grp_eff);
    pk.add(L.chunk_sp, &d.chunk_sp);
    pk.add(aux, &d.aux);
    pk.add(auxd, &d.auxd);
    d.tr_visc = d.tr_cond = d.tr_diff = d.tr_wilke = nullptr;
    std::vector<double> wilke;
    if (m.has_transport()) {
        pk.add(m.visc, &d.tr_visc);
        pk.add(m.cond, &d.tr_cond);
        pk.add(m.diff, &d.tr_diff);
        // Wilke's weight factor of the pair (k, j), sqrt(W_j) / sqrt(8 (W_j + W_k)), at
        // [j * K + k] (the transport kernel's Phi_kj = (1 + a_k / a_j)^2 times this)
        wilke.resize((size_t)m.K * m.K);
        for (int j = 0; j < m.K; j++)
            for (int k = 0; k < m.K; k++)
                wilke[(size_t)j * m.K + k] = std::sqrt(m.W[j]) / std::sqrt(8.0 * (m.W[j] + m.W[k]));
        pk.add(wilke, &d.tr_wilke);
    }
    pk.add(kf_c, &d.kf_c);
    pk.add(kf_s, &d.kf_s);
    pk.add(kc_c, &d.kc_c);
    pk.add(kc_s, &d.kc_s);
    if (c->blob) {
        // stages queued on the context stream still read the old tables
        HIP_TRY(hipStreamSynchronize(c->stream));
        (void)hipFree(c->blob);
    }
    c->blob = nullptr;
    HIP_TRY(hipMalloc(&c->blob, pk.bytes.size()));
    HIP_TRY(hipMemcpy(c->blob, pk.bytes.data(), pk.bytes.size(), hipMemcpyHostToDevice));
    for (auto& f : pk.fix) *f.second = static_cast<char*>(c->blob) + f.first;
    return 0;
}

int check_ctx(emx_context* c) {
    if (!c) return fail(-1, "null context");
    HIP_TRY(hipSetDevice(c->device));
    return 0;
}

// PLOG, Chebyshev or SRI reactions, or explicit reaction orders (no adjoint, no plain
// baseline kernel for them)
bool has_kinds_beyond_troe(const HostMech& m) {
    for (const auto& r : m.rxn)
        if (r.type >= 4 || r.hasOrd) return true;
    return false;
}

// per-slot multipliers for the production kernels: ones while the function is set
const double* stage_perturb(const emx_context* c) {
    return c->ratefn_n > 0 ? c->ones.as<double>() : c->perturb;
}

// kernel options of a stage starting at absolute time t_start (igniter amplitudes as
// prepared by igniter_ready)
DevOpts stage_opts(const emx_context* c, double t_start) {
    DevOpts d = c->dopts;
    if (c->ratefn_n > 0) {
        d.ratefn = c->ratefn.as<double>();
        d.ratefn_n = c->ratefn_n;
    }
    if (c->opts.ignition_energy != 0.0) d.qign = c->qign.as<double>();
    d.t_start = t_start;
    return d;
}

// The igniter's per-point amplitudes for a stage over nPoints points: getQdotIgniter's
// spatial factor at each position (src/sourceSystem.cpp:38-41), evaluated on the host.
int igniter_ready(emx_context* c, size_t nPoints) {
    if (c->opts.ignition_energy == 0.0) return 0;
    if (c->xpos.size() != nPoints)
        return fail(-1, "the igniter needs emx_set_point_positions for the stage's points");
    if (c->qign_valid) return 0;
    const emx_options& o = c->opts;
    std::vector<double> a(nPoints);
    for (size_t j = 0; j < nPoints; j++)
        a[j] = o.ignition_energy / (o.ignition_stddev * std::sqrt(2 * M_PI) * o.ignition_duration) *
               std::exp(-std::pow(c->xpos[j] - o.ignition_center, 2) / (2 * std::pow(o.ignition_stddev, 2)));
    HIP_TRY(c->qign.ensure(nPoints * sizeof(double)));
    HIP_TRY(hipStreamSynchronize(c->stream));  // stages in flight read the previous amplitudes
    HIP_TRY(hipMemcpy(c->qign.p, a.data(), nPoints * sizeof(double), hipMemcpyHostToDevice));
    c->qign_valid = true;
    return 0;
}

int grid_blocks(emx_context* c, int nPoints, bool bdf, int wpb) {
    size_t lds_per_block = bdf ? bdf_lds_bytes(c->dm, wpb) : qss_lds_bytes(c->dm, wpb);
    int per_cu = c->blocks_per_cu;
    if (per_cu <= 0) {
        per_cu = (int)std::max<size_t>(1, (160 * 1024) / std::max<size_t>(lds_per_block, 1));
        per_cu = std::min(per_cu, 32 / wpb);
    }
    long blocks = (long)per_cu * c->num_cu;
    long needed = (nPoints + wpb - 1) / wpb;
    return (int)std::max<long>(1, std::min(blocks, needed));
}

}  // namespace

extern "C" {

void emx_default_options(emx_options* o) {
    o->chemistry_integrator = EMX_QSS;
    o->kinetics_model = EMX_KIN_STANDARD;
    o->qss_epsmin = 2e-2;
    o->qss_epsmax = 10.0;
    o->qss_dtmin = 1e-16;
    o->qss_dtmax = 1e-6;
    o->qss_iteration_count = 1;
    o->qss_abstol = 1e-11;
    o->qss_minval = 1e-60;
    o->qss_stability_check = 0;
    o->integrator_rel_tol = 1e-6;
    o->integrator_momentum_abs_tol = 1e-7;
    o->integrator_energy_abs_tol = 1e-8;
    o->integrator_species_abs_tol = 1e-13;
    o->integrator_min_timestep = 1e-18;
    o->integrator_max_steps = 1000000;
    o->diffusion_timestep_multiplier = 10.0;
    o->global_timestep = 2e-5;
    o->ignition_t_start = 0.0;  // python/ember/input.py class Ignition
    o->ignition_duration = 1e-3;
    o->ignition_energy = 0.0;
    o->ignition_center = 0.0;
    o->ignition_stddev = 1e-4;
    o->transport_threshold = 1e-5;  // python/ember/input.py class Chemistry, threshold
}

const char* emx_last_error(void) { return g_err.c_str(); }

int emx_create(const char* path, double pressure, const emx_options* opts, emx_context** out) {
    if (!out) return fail(-1, "out is NULL");
    *out = nullptr;
    auto* c = new emx_context();
    if (opts) c->opts = *opts; else emx_default_options(&c->opts);
    c->dopts = to_dev(c->opts);
    c->P = pressure;
    std::string err = read_table(path ? path : "", c->hm);
    if (err.empty()) err = build_layout(c->hm, c->lay);
    if (!err.empty()) {
        delete c;
        return fail(-2, err);
    }
    hipError_t e = hipGetDevice(&c->device);
    if (e != hipSuccess) {
        delete c;
        return fail(-10, std::string("hipGetDevice: ") + hipGetErrorString(e));
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) c->num_cu = prop.multiProcessorCount;
    {
        std::string serr;
        if (int src = scratch_setup(c->device, serr)) {
            delete c;
            return fail(src, serr);
        }
    }
    int rc = upload_mech(c);
    if (rc) { delete c; return rc; }
    std::vector<double> ones(c->lay.nRounds * 64, 1.0);
    if (hipMalloc(&c->perturb, ones.size() * sizeof(double)) != hipSuccess ||
        hipMemcpy(c->perturb, ones.data(), ones.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc(&c->queue, 256) != hipSuccess) {
        emx_destroy(c);
        return fail(-10, "device allocation failed");
    }
    *out = c;
    return 0;
}

void emx_destroy(emx_context* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->blob) (void)hipFree(c->blob);
    if (c->perturb) (void)hipFree(c->perturb);
    if (c->queue) (void)hipFree(c->queue);
    for (DevBuf* b : {&c->b_in, &c->b_out, &c->b_sc, &c->b_ql, &c->b_status, &c->b_s1, &c->b_s2,
                      &c->d_B, &c->d_Brow, &c->d_D, &c->d_sc, &c->d_x, &c->d_y, &c->d_h, &c->d_ns,
                      &c->d_work, &c->d_hmin, &c->t_a, &c->t_b, &c->t_c, &c->t_k, &c->jac, &c->plain_ws,
                      &c->cost, &c->order, &c->ck_y, &c->ck_s, &c->nckpt, &c->record, &c->ratefn,
                      &c->ones, &c->qign, &c->cv_work, &c->cv_rec_t, &c->cv_rec_v, &c->cv_ints,
                      &c->cv_aux, &c->cv_yleft})
        b->release();
    delete c;
}

int emx_num_species(const emx_context* c) { return c ? c->hm.K : -1; }
int emx_num_reactions(const emx_context* c) { return c ? c->hm.I : -1; }

int emx_set_options(emx_context* c, const emx_options* o) {
    if (int rc = check_ctx(c)) return rc;
    if (!o) return fail(-1, "options is NULL");
    bool interp_changed = o->kinetics_model != c->opts.kinetics_model;
    c->opts = *o;
    c->dopts = to_dev(*o);
    c->qign_valid = false;
    if (interp_changed) return upload_mech(c);
    return 0;
}

int emx_set_interp_range(emx_context* c, int n, double Tmin, double Tmax) {
    if (int rc = check_ctx(c)) return rc;
    if (n < 2 || n > 100000 || !(Tmin > 0.0) || !(Tmax > Tmin))
        return fail(-1, "interpolation grid needs n >= 2 and 0 < Tmin < Tmax");
    c->interp_n = n;
    c->interp_tmin = Tmin;
    c->interp_tmax = Tmax;
    return c->opts.kinetics_model == EMX_KIN_INTERP ? upload_mech(c) : 0;
}

int emx_get_interp_range(const emx_context* c, int* n, double* Tmin, double* Tmax) {
    if (!c) return fail(-1, "null context");
    if (n) *n = c->interp_n;
    if (Tmin) *Tmin = c->interp_tmin;
    if (Tmax) *Tmax = c->interp_tmax;
    return 0;
}

int emx_set_point_positions(emx_context* c, size_t nPoints, const double* x) {
    if (int rc = check_ctx(c)) return rc;
    if (x)
        c->xpos.assign(x, x + nPoints);
    else
        c->xpos.clear();
    c->qign_valid = false;
    return 0;
}

int emx_set_rate_multipliers(emx_context* c, const double* m) {
    if (int rc = check_ctx(c)) return rc;
    std::vector<double> p(c->lay.nRounds * 64, 1.0);
    if (m)
        for (int i = 0; i < c->hm.I; i++) p[c->lay.slot_of_rxn[i]] = m[i];
    // stages already queued on the context stream read the previous multipliers
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(c->perturb, p.data(), p.size() * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

int emx_set_rate_multiplier_function(emx_context* c, int n, const double* data) {
    if (int rc = check_ctx(c)) return rc;
    if (n == 0 || !data) {
        c->ratefn_n = 0;
        return 0;
    }
    if (n < 2) return fail(-1, "the rate multiplier function needs n >= 2 (T0, T1, coefficients)");
    if (n - 2 > kMaxRateFnCoeffs) return fail(-1, "too many Chebyshev coefficients");
    if (!(data[1] != data[0])) return fail(-1, "the rate multiplier function needs T1 != T0");
    // ChebyshevFunction::setCoefficients: at least two coefficients, zero-filled
    std::vector<double> f = {data[0], data[1]};
    if (n >= 4)
        f.insert(f.end(), data + 2, data + n);
    else if (n == 3)
        f.insert(f.end(), {data[2], 0.0});
    else
        f.insert(f.end(), {0.0, 0.0});
    HIP_TRY(c->ratefn.ensure((2 + kMaxRateFnCoeffs) * sizeof(double)));
    if (!c->ones.p) {
        const std::vector<double> one(c->lay.nRounds * 64, 1.0);
        HIP_TRY(c->ones.ensure(one.size() * sizeof(double)));
        HIP_TRY(hipMemcpy(c->ones.p, one.data(), one.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    // stages already queued on the context stream read the previous coefficients
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(c->ratefn.p, f.data(), f.size() * sizeof(double), hipMemcpyHostToDevice));
    c->ratefn_n = (int)f.size() - 2;
    return 0;
}

int emx_set_stream(emx_context* c, void* stream) {
    if (!c) return fail(-1, "null context");
    c->stream = static_cast<hipStream_t>(stream);
    return 0;
}

int emx_set_launch(emx_context* c, int waves_per_block, int blocks_per_cu) {
    if (!c) return fail(-1, "null context");
    if (waves_per_block < 0 || waves_per_block > 16 || blocks_per_cu < 0)
        return fail(-1, "invalid launch shape");
    c->waves_per_block = waves_per_block ? waves_per_block : 4;
    c->waves_set = waves_per_block != 0;
    c->blocks_per_cu = blocks_per_cu;
    return 0;
}

int emx_set_kernel_variant(emx_context* c, int variant) {
    if (!c) return fail(-1, "null context");
    if (variant != EMX_VARIANT_WAVE && variant != EMX_VARIANT_PLAIN && variant != EMX_VARIANT_WAVE_PER_POINT)
        return fail(-1, "unknown variant");
    c->variant = variant;
    return 0;
}

int emx_set_schedule(emx_context* c, int mode) {
    if (!c) return fail(-1, "null context");
    if (mode != EMX_SCHED_FIFO && mode != EMX_SCHED_COST) return fail(-1, "unknown schedule");
    c->schedule = mode;
    return 0;
}

int emx_source_integrate_device(emx_context* c, size_t nPoints, size_t nVars, const double* state_in,
                                const double* split_const, const double* qloss, double t_stage_start,
                                double dt, double* state_out, int32_t* status, int32_t* stat1,
                                int32_t* stat2, double* qint) {
    if (int rc = check_ctx(c)) return rc;
    if ((int)nVars != c->hm.K + 2) return fail(-1, "nVars must equal nSpecies + 2");
    if (nPoints == 0) return 0;
    if (nPoints > 0x7fffffffULL) return fail(-1, "too many points");
    if (!state_in || !state_out) return fail(-1, "state pointers are NULL");
    if (c->variant == EMX_VARIANT_PLAIN) {
        if (c->opts.chemistry_integrator != EMX_QSS || c->opts.qss_stability_check || c->ratefn_n ||
            c->opts.ignition_energy != 0.0 || has_kinds_beyond_troe(c->hm))
            return fail(-1, "the plain baseline kernel implements QSS without stability check, "
                            "rate multiplier function, igniter, PLOG, Chebyshev, SRI rates or "
                            "explicit reaction orders only");
        HIP_TRY(c->plain_ws.ensure(plain_workspace_doubles(c->dm, (int)nPoints) * sizeof(double)));
        int rc = launch_qss_plain(c->dm, c->dopts, state_in, split_const, qloss, c->perturb, dt,
                                  (int)nPoints, state_out, status, stat1, stat2,
                                  c->plain_ws.as<double>(), c->stream);
        if (rc) return fail(-10, std::string("plain launch: ") + hipGetErrorString((hipError_t)rc));
        return 0;
    }
    if (int rc = igniter_ready(c, nPoints)) return rc;
    // work queue order: descending cost of the previous stage over the same points
    HIP_TRY(c->cost.ensure(nPoints * sizeof(int)));
    HIP_TRY(c->order.ensure(nPoints * sizeof(int)));
    const int* order = nullptr;
    if (c->schedule == EMX_SCHED_COST && c->cost_n == nPoints) {
        int rc = launch_cost_order(c->cost.as<int>(), (int)nPoints, c->order.as<int>(), c->stream);
        if (rc) return fail(-10, std::string("order launch: ") + hipGetErrorString((hipError_t)rc));
        order = c->order.as<int>();
    }
    if (c->opts.chemistry_integrator == EMX_BDF) {
        if (c->variant == EMX_VARIANT_WAVE_PER_POINT)
            return fail(-1, "the per-point wavefront variant implements QSS only");
        // a wave's LU pair matrix grows with the state (32 KB at 64 variables): from about 55
        // species on, four waves of a CU no longer fit in its 160 KB, and the default shape
        // runs fewer waves per workgroup (one workgroup per CU)
        int wpb = c->waves_per_block;
        if (!c->waves_set)
            while (wpb > 1 && bdf_lds_bytes(c->dm, wpb) > 160 * 1024) wpb--;
        if (bdf_lds_bytes(c->dm, wpb) > 160 * 1024)
            return fail(-1, "BDF workgroup needs more than 160 KB of LDS; use fewer waves per block");
        const int blocks = grid_blocks(c, (int)nPoints, true, wpb);
        HIP_TRY(c->jac.ensure((size_t)blocks * wpb * 64 * 64 * sizeof(double)));
        int rc = launch_bdf_stage(c->dm, stage_opts(c, t_stage_start), state_in, split_const, qloss, stage_perturb(c), dt,
                                  (int)nPoints, order, c->queue, state_out, status, stat1, stat2,
                                  c->cost.as<int>(), qint, blocks, wpb,
                                  c->jac.as<double>(), c->stream);
        if (rc) return fail(-10, std::string("bdf launch: ") + hipGetErrorString((hipError_t)rc));
        c->cost_n = nPoints;
        return 0;
    }
    const bool per_point = c->variant == EMX_VARIANT_WAVE_PER_POINT;
    // with the rate multiplier function or the igniter the kernels run at up to 8 waves
    // per workgroup
    const bool ext = c->ratefn_n > 0 || c->opts.ignition_energy != 0.0;
    const int wpb = ext ? std::min(c->waves_per_block, 8) : c->waves_per_block;
    const int qblocks = per_point ? (int)((nPoints + wpb - 1) / wpb)
                                  : grid_blocks(c, (int)nPoints, false, wpb);
    int rc = launch_qss_stage(c->dm, stage_opts(c, t_stage_start), state_in, split_const, qloss,
                              stage_perturb(c), dt, (int)nPoints, per_point ? nullptr : order,
                              per_point ? nullptr : c->queue,
                              state_out, status, stat1, stat2, c->cost.as<int>(), qint, qblocks,
                              wpb, c->stream);
    if (rc) return fail(-10, std::string("qss launch: ") + hipGetErrorString((hipError_t)rc));
    c->cost_n = nPoints;
    return 0;
}

int emx_source_integrate_device_noalloc(emx_context* c, size_t nPoints, size_t nVars,
                                        const double* state_in, const double* split_const,
                                        const double* qloss, double t_stage_start, double dt,
                                        double* state_out, int32_t* status, int32_t* stat1,
                                        int32_t* stat2, double* qint) {
    if (int rc = check_ctx(c)) return rc;
    if ((int)nVars != c->hm.K + 2) return fail(-1, "nVars must equal nSpecies + 2");
    if (nPoints == 0) return 0;
    if (nPoints > kNoAllocMaxPoints)
        return fail(-1, "emx_source_integrate_device_noalloc takes at most 262144 points per call");
    if (!state_in || !state_out) return fail(-1, "state pointers are NULL");
    if (c->opts.chemistry_integrator != EMX_QSS)
        return fail(-1, "emx_source_integrate_device_noalloc runs the QSS integrator only (the BDF "
                        "needs a Jacobian workspace per wave)");
    if (c->variant == EMX_VARIANT_PLAIN)
        return fail(-1, "emx_source_integrate_device_noalloc does not run the plain baseline kernel");
    if (c->opts.ignition_energy != 0.0 && !(c->qign_valid && c->xpos.size() == nPoints))
        return fail(-1, "with the igniter on, run one emx_source_integrate_device stage over these "
                        "points first (it prepares the igniter amplitudes)");
    unsigned char* base;
    bool cost_valid;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        DeviceScratch& d = g_scratch[c->device];
        if (!d.ptr) return fail(-10, "scratch of this device was not set up");
        base = d.ptr;
        cost_valid = d.last_n == nPoints;
        d.last_n = nPoints;
    }
    unsigned int* queue = reinterpret_cast<unsigned int*>(base);
    int* cost = reinterpret_cast<int*>(base + 256);
    int* order_buf = cost + kNoAllocMaxPoints;
    // the costs left by the previous allocation-free stage on this device over the same
    // number of points give the work-queue order (results do not depend on it)
    const int* order = nullptr;
    if (c->schedule == EMX_SCHED_COST && cost_valid && c->variant == EMX_VARIANT_WAVE) {
        int rc = launch_cost_order(cost, (int)nPoints, order_buf, c->stream);
        if (rc) return fail(-10, std::string("order launch: ") + hipGetErrorString((hipError_t)rc));
        order = order_buf;
    }
    const bool per_point = c->variant == EMX_VARIANT_WAVE_PER_POINT;
    const bool ext = c->ratefn_n > 0 || c->opts.ignition_energy != 0.0;
    const int wpb = ext ? std::min(c->waves_per_block, 8) : c->waves_per_block;
    const int qblocks = per_point ? (int)((nPoints + wpb - 1) / wpb)
                                  : grid_blocks(c, (int)nPoints, false, wpb);
    int rc = launch_qss_stage(c->dm, stage_opts(c, t_stage_start), state_in, split_const, qloss,
                              stage_perturb(c), dt, (int)nPoints, per_point ? nullptr : order,
                              per_point ? nullptr : queue, state_out, status, stat1, stat2, cost, qint,
                              qblocks, wpb, c->stream);
    if (rc) return fail(-10, std::string("qss launch: ") + hipGetErrorString((hipError_t)rc));
    return 0;
}

// ---------------------------------------------------------------------------
// discrete adjoint of the QSS stage
// ---------------------------------------------------------------------------
static int adjoint_check(emx_context* c, size_t nPoints, size_t nVars) {
    if (int rc = check_ctx(c)) return rc;
    if ((int)nVars != c->hm.K + 2) return fail(-1, "nVars must equal nSpecies + 2");
    if (nPoints > 0x7fffffffULL) return fail(-1, "too many points");
    if (c->opts.chemistry_integrator != EMX_QSS || c->opts.qss_iteration_count != 1 ||
        c->opts.qss_stability_check)
        return fail(-1, "the adjoint supports the QSS integrator with qss_iteration_count = 1 and "
                        "no stability check");
    if (c->ratefn_n)
        return fail(-1, "the adjoint's ln-multiplier parameters need per-reaction multipliers; "
                        "remove the rate multiplier function");
    if (c->opts.ignition_energy != 0.0) return fail(-1, "the adjoint does not support the igniter");
    if (has_kinds_beyond_troe(c->hm))
        return fail(-1, "the adjoint does not support PLOG, Chebyshev, SRI reactions or explicit reaction orders");
    return 0;
}

int emx_set_adjoint_capacity(emx_context* c, int max_checkpoints) {
    if (!c) return fail(-1, "null context");
    if (max_checkpoints < 1) return fail(-1, "max_checkpoints must be positive");
    c->adj_cap = max_checkpoints;
    c->adj_n = 0;
    return 0;
}

int emx_source_adjoint_forward(emx_context* c, size_t nPoints, size_t nVars, const double* state_in,
                               const double* split_const, const double* qloss, double dt,
                               double* state_out, int32_t* status, double* heat_release) {
    if (int rc = adjoint_check(c, nPoints, nVars)) return rc;
    c->adj_n = 0;
    if (nPoints == 0) return 0;
    if (!state_in || !state_out) return fail(-1, "state pointers are NULL");
    const size_t cap = (size_t)c->adj_cap;
    HIP_TRY(c->ck_y.ensure(nPoints * cap * nVars * sizeof(double)));
    HIP_TRY(c->ck_s.ensure(nPoints * cap * 2 * sizeof(double)));
    HIP_TRY(c->nckpt.ensure(nPoints * sizeof(int)));
    HIP_TRY(c->cost.ensure(nPoints * sizeof(int)));
    HIP_TRY(c->order.ensure(nPoints * sizeof(int)));
    const int* order = nullptr;
    if (c->schedule == EMX_SCHED_COST && c->cost_n == nPoints) {
        int rc = launch_cost_order(c->cost.as<int>(), (int)nPoints, c->order.as<int>(), c->stream);
        if (rc) return fail(-10, std::string("order launch: ") + hipGetErrorString((hipError_t)rc));
        order = c->order.as<int>();
    }
    const int wpb = std::min(c->waves_per_block, 8);
    const size_t lds = adjoint_lds_bytes(c->dm, wpb, false);
    int per_cu = std::max<int>(1, (int)((160 * 1024) / lds));
    per_cu = std::min(per_cu, 32 / wpb);
    const int blocks = (int)std::max<long>(1, std::min<long>((long)per_cu * c->num_cu, ((long)nPoints + wpb - 1) / wpb));
    int rc = launch_qss_adjoint_forward(c->dm, c->dopts, state_in, split_const, qloss, c->perturb, dt,
                                        (int)nPoints, order, c->queue, state_out, status, heat_release,
                                        c->cost.as<int>(), c->ck_y.as<double>(), c->ck_s.as<double>(),
                                        c->nckpt.as<int>(), c->adj_cap, blocks, wpb, c->stream);
    if (rc) return fail(-10, std::string("adjoint forward launch: ") + hipGetErrorString((hipError_t)rc));
    c->cost_n = nPoints;
    c->adj_n = nPoints;
    c->adj_dt = dt;
    return 0;
}

int emx_source_adjoint_backward(emx_context* c, size_t nPoints, size_t nVars, const double* state_in,
                                const double* split_const, const double* qloss, double dt,
                                const double* Qbar, const double* ybar_final, double* grad_state,
                                double* grad_lnk, int32_t* status) {
    if (int rc = adjoint_check(c, nPoints, nVars)) return rc;
    if (nPoints == 0) return 0;
    if (c->adj_n != nPoints || c->adj_dt != dt)
        return fail(-1, "emx_source_adjoint_backward needs a preceding emx_source_adjoint_forward "
                        "over the same points and dt");
    if (!state_in || !grad_state) return fail(-1, "state pointers are NULL");
    const int wpb = 8;
    const size_t lds = adjoint_lds_bytes(c->dm, wpb, true);
    if (lds > 160 * 1024) return fail(-1, "adjoint workgroup needs more than 160 KB of LDS");
    const int blocks = (int)std::max<long>(1, std::min<long>((long)c->num_cu, ((long)nPoints + wpb - 1) / wpb));
    HIP_TRY(c->record.ensure((size_t)blocks * wpb * adjoint_record_doubles_per_wave() * sizeof(double)));
    // most expensive points first (costs of the forward stage)
    const int* order = nullptr;
    if (c->schedule == EMX_SCHED_COST && c->cost_n == nPoints) {
        int rc = launch_cost_order(c->cost.as<int>(), (int)nPoints, c->order.as<int>(), c->stream);
        if (rc) return fail(-10, std::string("order launch: ") + hipGetErrorString((hipError_t)rc));
        order = c->order.as<int>();
    }
    int rc = launch_qss_adjoint_backward(c->dm, c->dopts, state_in, split_const, qloss, c->perturb, dt,
                                         (int)nPoints, order, c->queue, c->ck_y.as<double>(),
                                         c->ck_s.as<double>(), c->nckpt.as<int>(), c->adj_cap, Qbar,
                                         ybar_final, c->record.as<double>(), grad_state, grad_lnk,
                                         status, blocks, wpb, c->stream);
    if (rc) return fail(-10, std::string("adjoint backward launch: ") + hipGetErrorString((hipError_t)rc));
    return 0;
}

// InterpKinetics::updateROP grows its tables when T leaves them (src/chemistry0d.cpp:
// 279-291): 10 more nodes and 100 K more above, or 2 more nodes and a lower bound of
// max(T - 20, 10) K below.  The host-array stage applies the same growth to the points
// that failed with -2 (T outside the table) and integrates them again from their input
// states, up to 40 growth steps.  The grown tables stay for later stages, as Ember's do.
static int interp_range_retry(emx_context* c, size_t nPoints, size_t nVars, const double* state_in,
                              const double* split_const, const double* qloss, double t0, double dt,
                              double* state_out, std::vector<int32_t>& stat,
                              std::vector<int32_t>& g1, std::vector<int32_t>& g2) {
    hipStream_t st = c->stream;
    for (int attempt = 0; attempt < 40; attempt++) {
        std::vector<size_t> idx;
        double Tlow = c->interp_tmin;
        for (size_t j = 0; j < nPoints; j++)
            if (stat[j] == -2) {
                idx.push_back(j);
                Tlow = std::min(Tlow, state_in[j * nVars + 1]);
            }
        if (idx.empty()) return 0;
        if (Tlow < c->interp_tmin) {
            c->interp_n += 2;
            c->interp_tmin = std::max(Tlow - 20.0, 10.0);
        } else {
            c->interp_n += 10;
            c->interp_tmax += 100.0;
        }
        if (int rc = upload_mech(c)) return rc;
        const size_t n = idx.size();
        std::vector<double> sin(n * nVars), sc(split_const ? n * nVars : 0), ql(qloss ? n : 0);
        for (size_t i = 0; i < n; i++) {
            memcpy(&sin[i * nVars], state_in + idx[i] * nVars, nVars * 8);
            if (split_const) memcpy(&sc[i * nVars], split_const + idx[i] * nVars, nVars * 8);
            if (qloss) ql[i] = qloss[idx[i]];
        }
        HIP_TRY(hipMemcpyAsync(c->b_in.p, sin.data(), n * nVars * 8, hipMemcpyHostToDevice, st));
        if (split_const) HIP_TRY(hipMemcpyAsync(c->b_sc.p, sc.data(), n * nVars * 8, hipMemcpyHostToDevice, st));
        if (qloss) HIP_TRY(hipMemcpyAsync(c->b_ql.p, ql.data(), n * 8, hipMemcpyHostToDevice, st));
        c->cost_n = 0;  // the recorded costs are for the full point set
        int rc = emx_source_integrate_device(c, n, nVars, c->b_in.as<double>(),
                                             split_const ? c->b_sc.as<double>() : nullptr,
                                             qloss ? c->b_ql.as<double>() : nullptr, t0, dt,
                                             c->b_out.as<double>(), c->b_status.as<int32_t>(),
                                             c->b_s1.as<int32_t>(), c->b_s2.as<int32_t>(), nullptr);
        if (rc) return rc;
        std::vector<double> sout(n * nVars);
        std::vector<int32_t> s0(n), s1(n), s2(n);
        HIP_TRY(hipMemcpyAsync(sout.data(), c->b_out.p, n * nVars * 8, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(s0.data(), c->b_status.p, n * 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(s1.data(), c->b_s1.p, n * 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(s2.data(), c->b_s2.p, n * 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        c->cost_n = 0;
        for (size_t i = 0; i < n; i++) {
            memcpy(state_out + idx[i] * nVars, &sout[i * nVars], nVars * 8);
            stat[idx[i]] = s0[i];
            g1[idx[i]] = s1[i];
            g2[idx[i]] = s2[i];
        }
    }
    return 0;
}

int emx_source_integrate(emx_context* c, size_t nPoints, size_t nVars, const double* state_in,
                         const double* split_const, const double* qloss, double t_stage_start,
                         double dt, double* state_out, int32_t* status, int32_t* stat1,
                         int32_t* stat2) {
    if (int rc = check_ctx(c)) return rc;
    if ((int)nVars != c->hm.K + 2) return fail(-1, "nVars must equal nSpecies + 2");
    if (nPoints == 0) return 0;
    if (!state_in || !state_out) return fail(-1, "state pointers are NULL");
    const size_t bytes = nPoints * nVars * sizeof(double);
    HIP_TRY(c->b_in.ensure(bytes));
    HIP_TRY(c->b_out.ensure(bytes));
    HIP_TRY(c->b_status.ensure(nPoints * sizeof(int32_t)));
    HIP_TRY(c->b_s1.ensure(nPoints * sizeof(int32_t)));
    HIP_TRY(c->b_s2.ensure(nPoints * sizeof(int32_t)));
    hipStream_t st = c->stream;
    HIP_TRY(hipMemcpyAsync(c->b_in.p, state_in, bytes, hipMemcpyHostToDevice, st));
    const double* dsc = nullptr;
    const double* dql = nullptr;
    if (split_const) {
        HIP_TRY(c->b_sc.ensure(bytes));
        HIP_TRY(hipMemcpyAsync(c->b_sc.p, split_const, bytes, hipMemcpyHostToDevice, st));
        dsc = c->b_sc.as<double>();
    }
    if (qloss) {
        HIP_TRY(c->b_ql.ensure(nPoints * sizeof(double)));
        HIP_TRY(hipMemcpyAsync(c->b_ql.p, qloss, nPoints * sizeof(double), hipMemcpyHostToDevice, st));
        dql = c->b_ql.as<double>();
    }
    int rc = emx_source_integrate_device(c, nPoints, nVars, c->b_in.as<double>(), dsc, dql,
                                         t_stage_start, dt, c->b_out.as<double>(),
                                         c->b_status.as<int32_t>(), c->b_s1.as<int32_t>(),
                                         c->b_s2.as<int32_t>(), nullptr);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(state_out, c->b_out.p, bytes, hipMemcpyDeviceToHost, st));
    std::vector<int32_t> stat(nPoints), g1(nPoints), g2(nPoints);
    HIP_TRY(hipMemcpyAsync(stat.data(), c->b_status.p, nPoints * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(g1.data(), c->b_s1.p, nPoints * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(g2.data(), c->b_s2.p, nPoints * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (c->opts.kinetics_model == EMX_KIN_INTERP)
        if (int rc2 = interp_range_retry(c, nPoints, nVars, state_in, split_const, qloss, t_stage_start,
                                         dt, state_out, stat, g1, g2))
            return rc2;
    if (status) memcpy(status, stat.data(), nPoints * 4);
    if (stat1) memcpy(stat1, g1.data(), nPoints * 4);
    if (stat2) memcpy(stat2, g2.data(), nPoints * 4);
    return 0;
}

int emx_rates(emx_context* c, size_t n, const double* T, const double* Y, int nonorm, double* wQ,
              double* wD, double* wnet, int32_t* status) {
    if (int rc = check_ctx(c)) return rc;
    if (n == 0) return 0;
    if (!T || !Y) return fail(-1, "T or Y is NULL");
    const int K = c->hm.K;
    DevBuf dT, dY, dQ, dD, dN, dS;
    auto cleanup = [&]() { dT.release(); dY.release(); dQ.release(); dD.release(); dN.release(); dS.release(); };
    size_t by = n * K * sizeof(double);
    if (dT.ensure(n * 8) || dY.ensure(by) || dQ.ensure(by) || dD.ensure(by) || dN.ensure(by) ||
        dS.ensure(n * 4)) {
        cleanup();
        return fail(-10, "device allocation failed");
    }
    hipStream_t st = c->stream;
    int rc = 0;
    if (hipMemcpyAsync(dT.p, T, n * 8, hipMemcpyHostToDevice, st) ||
        hipMemcpyAsync(dY.p, Y, by, hipMemcpyHostToDevice, st)) rc = fail(-10, "copy failed");
    if (!rc) {
        int e = launch_rates(c->dm, dT.as<double>(), dY.as<double>(), nonorm, c->perturb, (int)n,
                             dQ.as<double>(), dD.as<double>(), dN.as<double>(), dS.as<int>(), st);
        if (e) rc = fail(-10, std::string("rates launch: ") + hipGetErrorString((hipError_t)e));
    }
    if (!rc && ((wQ && hipMemcpyAsync(wQ, dQ.p, by, hipMemcpyDeviceToHost, st)) ||
                (wD && hipMemcpyAsync(wD, dD.p, by, hipMemcpyDeviceToHost, st)) ||
                (wnet && hipMemcpyAsync(wnet, dN.p, by, hipMemcpyDeviceToHost, st)) ||
                (status && hipMemcpyAsync(status, dS.p, n * 4, hipMemcpyDeviceToHost, st)) ||
                hipStreamSynchronize(st)))
        rc = fail(-10, "copy back failed");
    cleanup();
    return rc;
}

// ---------------------------------------------------------------------------
// diffusion
// ---------------------------------------------------------------------------
// Per variable: Integrator::initialize(t0, h) and integrateToTime(t1)
// (src/integrator.cpp:37-47) with h from the caller or, when h_host is NULL, from
// FlameSolver::setDiffusionSolverState's rule (src/flameSolver.cpp:709-711), which needs
// one device-to-host read of the grid minimum.  Otherwise the call is asynchronous.
static int diffusion_run(emx_context* c, const DiffArgs& a, const double* h_host, double t0,
                         double t1, double* y_dev) {
    hipStream_t st = c->stream;
    const int nV = a.nV, N = a.N;
    HIP_TRY(c->d_work.ensure(diffusion_work_doubles(nV, N) * sizeof(double)));
    DiffWork w;
    diffusion_carve(c->d_work.as<double>(), nV, N, w);
    DiffSteps steps{};
    if (h_host) {
        for (int v = 0; v < nV; v++) steps.h[v] = h_host[v];
    } else {
        HIP_TRY(c->d_hmin.ensure(nV * sizeof(double)));
        int e = launch_diffusion_hmin(a, w, c->d_hmin.as<double>(), st);
        if (e) return fail(-10, "hmin launch failed");
        std::vector<double> hm(nV);
        HIP_TRY(hipMemcpyAsync(hm.data(), c->d_hmin.p, nV * 8, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        for (int v = 0; v < nV; v++)
            steps.h[v] = std::min(c->opts.diffusion_timestep_multiplier * hm[v], c->opts.global_timestep);
    }
    // Integrator::integrateToTime: equal steps (src/integrator.cpp:41-43)
    int maxs = 1;
    for (int v = 0; v < nV; v++) {
        if (!(steps.h[v] > 0.0)) return fail(-1, "diffusion step sizes must be positive");
        int n = static_cast<int>((t1 - t0) / steps.h[v] + 1e-5);
        n = std::max(n, 1);
        steps.h[v] = (t1 - t0) / n;
        steps.ns[v] = n;
        maxs = std::max(maxs, n);
    }
    HIP_TRY(c->d_h.ensure(nV * sizeof(double)));
    HIP_TRY(c->d_ns.ensure(nV * sizeof(int)));
    const int nsolve = 8 + (maxs - 1);
    int e = launch_diffusion_stage(a, y_dev, steps, c->d_h.as<double>(), c->d_ns.as<int>(), nsolve, w, st);
    if (e) return fail(-10, std::string("diffusion launch: ") + hipGetErrorString((hipError_t)e));
    return 0;
}

static int check_diff(emx_context* c, size_t nVars, size_t nPoints, const emx_grid* g) {
    if (int rc = check_ctx(c)) return rc;
    if (!g || !g->x) return fail(-1, "grid is NULL");
    if (nPoints < 4) return fail(-1, "diffusion needs at least 4 points");
    if (nVars < 1 || nVars > 64) return fail(-1, "nVars must be in [1, 64]");
    if (g->alpha < 0 || g->alpha > 1) return fail(-1, "alpha must be 0 or 1");
    if (g->left_bc < 0 || g->left_bc > 3 || g->right_bc < 0 || g->right_bc > 1)
        return fail(-1, "invalid boundary condition");
    return 0;
}

int emx_diffusion_integrate_device(emx_context* c, size_t nVars, size_t nPoints, const double* B,
                                   size_t nB, const int32_t* B_row, const double* D,
                                   const double* split_const, const emx_grid* g,
                                   const double* h_per_var, double t0, double t1, double* state) {
    if (int rc = check_diff(c, nVars, nPoints, g)) return rc;
    if (!B_row && nB != nVars) return fail(-1, "B_row is NULL but nB != nVars");
    if (!B || !D || !state) return fail(-1, "B, D or state is NULL");
    DiffArgs a{B, B_row, D, split_const, g->x, (int)nPoints, (int)nVars, g->alpha, g->left_bc,
               g->right_bc, g->y_inf, g->wall_const, g->wall_var};
    return diffusion_run(c, a, h_per_var, t0, t1, state);
}

int emx_diffusion_integrate(emx_context* c, size_t nVars, size_t nPoints, const double* B, size_t nB,
                            const int32_t* B_row, const double* D, const double* split_const,
                            const emx_grid* g, const double* h_per_var, double t0, double t1,
                            double* state) {
    if (int rc = check_diff(c, nVars, nPoints, g)) return rc;
    if (!B_row && nB != nVars) return fail(-1, "B_row is NULL but nB != nVars");
    if (!B || !D || !state) return fail(-1, "B, D or state is NULL");
    if (B_row)
        for (size_t v = 0; v < nVars; v++)
            if (B_row[v] < 0 || (size_t)B_row[v] >= nB) return fail(-1, "B_row entry out of range");
    hipStream_t st = c->stream;
    const size_t NV = nVars * nPoints * sizeof(double);
    HIP_TRY(c->d_B.ensure(nB * nPoints * sizeof(double)));
    HIP_TRY(c->d_D.ensure(NV));
    HIP_TRY(c->d_y.ensure(NV));
    HIP_TRY(c->d_x.ensure(nPoints * sizeof(double)));
    HIP_TRY(hipMemcpyAsync(c->d_B.p, B, nB * nPoints * 8, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->d_D.p, D, NV, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->d_y.p, state, NV, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->d_x.p, g->x, nPoints * 8, hipMemcpyHostToDevice, st));
    const int32_t* brow = nullptr;
    if (B_row) {
        HIP_TRY(c->d_Brow.ensure(nVars * 4));
        HIP_TRY(hipMemcpyAsync(c->d_Brow.p, B_row, nVars * 4, hipMemcpyHostToDevice, st));
        brow = c->d_Brow.as<int32_t>();
    }
    const double* sc = nullptr;
    if (split_const) {
        HIP_TRY(c->d_sc.ensure(NV));
        HIP_TRY(hipMemcpyAsync(c->d_sc.p, split_const, NV, hipMemcpyHostToDevice, st));
        sc = c->d_sc.as<double>();
    }
    DiffArgs a{c->d_B.as<double>(), brow, c->d_D.as<double>(), sc, c->d_x.as<double>(),
               (int)nPoints, (int)nVars, g->alpha, g->left_bc, g->right_bc, g->y_inf, g->wall_const, g->wall_var};
    int rc = diffusion_run(c, a, h_per_var, t0, t1, c->d_y.as<double>());
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(state, c->d_y.p, NV, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

// ---------------------------------------------------------------------------
// device-resident split step
// ---------------------------------------------------------------------------
// Ember's diffusion step rule (src/flameSolver.cpp:709-711) for every variable, from one
// read of the per-variable grid minimum.
static int diffusion_steps(emx_context* c, const DiffArgs& a, std::vector<double>& h) {
    hipStream_t st = c->stream;
    const int nV = a.nV, N = a.N;
    HIP_TRY(c->d_work.ensure(diffusion_work_doubles(nV, N) * sizeof(double)));
    DiffWork w;
    diffusion_carve(c->d_work.as<double>(), nV, N, w);
    HIP_TRY(c->d_hmin.ensure(nV * sizeof(double)));
    if (launch_diffusion_hmin(a, w, c->d_hmin.as<double>(), st)) return fail(-10, "hmin launch failed");
    std::vector<double> hm(nV);
    HIP_TRY(hipMemcpyAsync(hm.data(), c->d_hmin.p, nV * 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    h.resize(nV);
    for (int v = 0; v < nV; v++)
        h[v] = std::min(c->opts.diffusion_timestep_multiplier * hm[v], c->opts.global_timestep);
    return 0;
}

// The convection stage's point range, continuity condition, geometry and boundary
// conditions (emx_convection_integrate_device, and emx_split_step_conv_device before its
// first stage).
static int check_conv_params(size_t nPoints, const emx_convection_params* cp) {
    if (nPoints < 3 || nPoints > (size_t)kConvMaxPoints)
        return fail(-1, "the convection stage takes 3 to 4096 grid points");
    if (cp->continuity_bc != 0)
        return fail(-1, "only the fixedLeft continuity boundary condition is supported");
    if (cp->alpha < 0 || cp->alpha > 1 || (cp->beta != 1 && cp->beta != 2) || cp->left_bc < 0 ||
        cp->left_bc > 3 || cp->right_bc < 0 || cp->right_bc > 1)
        return fail(-1, "invalid grid geometry or boundary condition");
    return 0;
}

// SplitSolver::step on the device (emx_split_step_device, emx_split_step_conv_device);
// conv NULL: empty convection stages.
static int split_step(emx_context* c, size_t nVars, size_t nPoints, const emx_split_arrays* s,
                      const emx_split_convection* conv, const double* B, size_t nB, const int32_t* B_row,
                      const double* D, const emx_grid* g, const double* qloss, const double* h_per_var,
                      double t, double dt, int balanced, int32_t* status) {
    if (int rc = check_diff(c, nVars, nPoints, g)) return rc;
    if ((int)nVars != c->hm.K + 2) return fail(-1, "nVars must equal nSpecies + 2");
    if (!s || !s->state || !s->ddt_conv || !s->ddt_diff || !s->ddt_prod || !s->delta_diff ||
        !s->delta_prod || !s->work)
        return fail(-1, "split arrays: only ddt_cross may be NULL");
    if (!B_row && nB != nVars) return fail(-1, "B_row is NULL but nB != nVars");
    if (!(dt > 0.0)) return fail(-1, "dt must be positive");
    if (conv) {
        if (!conv->params || !conv->delta_conv) return fail(-1, "convection: params and delta_conv are required");
        if ((conv->rho == nullptr) != (conv->Wmx == nullptr))
            return fail(-1, "convection: rho and Wmx are given together or not at all");
        // refused before any stage runs, so that a refusal leaves the state untouched
        if (int rc = check_conv_params(nPoints, conv->params)) return rc;
    }
    hipStream_t st = c->stream;
    const size_t n = nVars * nPoints;
    double* state = s->state;
    double* start = s->work;
    double* sc_diff = s->work + n;
    double* sc_prod = s->work + 2 * n;
    double* sc_conv = conv ? s->work + 3 * n : nullptr;
    double* drhodt = conv && conv->rho ? s->work + 4 * n : nullptr;
    DiffArgs a{B, B_row, D, sc_diff, g->x, (int)nPoints, (int)nVars, g->alpha, g->left_bc,
               g->right_bc, g->y_inf, g->wall_const, g->wall_var};
    std::vector<double> h;
    if (h_per_var) {
        h.assign(h_per_var, h_per_var + nVars);
    } else if (int rc = diffusion_steps(c, a, h)) {
        return rc;
    }
    HIP_TRY(c->b_status.ensure(nPoints * sizeof(int32_t)));
    HIP_TRY(c->b_s1.ensure(nPoints * sizeof(int32_t)));
    HIP_TRY(c->b_s2.ensure(nPoints * sizeof(int32_t)));
    auto launched = [&](int e, const char* what) {
        return e ? fail(-10, std::string(what) + ": " + hipGetErrorString((hipError_t)e)) : 0;
    };
    // SplitSolver::step: split constants from the previous derivatives; with convection
    // stages, FlameSolver::prepareIntegrators' density derivative from the same derivatives
    if (conv) {
        if (int rc = launched(launch_split_constants_conv(n, balanced, s->ddt_conv, s->ddt_diff, s->ddt_prod,
                                                          s->ddt_cross, sc_diff, sc_prod, sc_conv, st),
                              "split constants"))
            return rc;
        if (drhodt)
            if (int rc = launched(launch_density_derivative(c->dm, (int)nPoints, balanced, state, s->ddt_conv,
                                                            s->ddt_diff, s->ddt_prod, s->ddt_cross, conv->rho,
                                                            conv->Wmx, drhodt, st),
                                  "density derivative"))
                return rc;
    } else if (int rc = launched(launch_split_constants(n, balanced, s->ddt_conv, s->ddt_diff, s->ddt_prod,
                                                        s->ddt_cross, sc_diff, sc_prod, st),
                                 "split constants")) {
        return rc;
    }
    // integrateSplitTerms (src/splitSolver.cpp:60-75)
    if (int rc = launched(launch_fill(n, 0.0, s->delta_diff, s->delta_prod, st), "fill")) return rc;
    if (conv)
        if (int rc = launched(launch_fill(n, 0.0, conv->delta_conv, nullptr, st), "fill")) return rc;
    auto diffusion_stage = [&](double t0, double t1) -> int {
        if (int rc = launched(launch_copy(n, state, start, st), "copy")) return rc;
        if (int rc = diffusion_run(c, a, h.data(), t0, t1, state)) return rc;
        return launched(launch_accumulate(n, state, start, s->delta_diff, st), "accumulate");
    };
    // FlameSolver::integrateConvectionTerms (flameSolver.cpp:736-751); empty without conv
    auto convection_stage = [&](double t0, double t1, int stage) -> int {
        if (!conv) return 0;
        if (int rc = launched(launch_copy(n, state, start, st), "copy")) return rc;
        int32_t* steps = conv->steps ? conv->steps + (size_t)stage * (c->hm.K + 1) : nullptr;
        if (int rc = emx_convection_integrate_device(c, nVars, nPoints, conv->params, g->x, start, sc_conv,
                                                     drhodt, t0, t1, state, steps))
            return rc;
        return launched(launch_accumulate(n, state, start, conv->delta_conv, st), "accumulate");
    };
    if (int rc = diffusion_stage(t, t + 0.25 * dt)) return rc;
    if (int rc = convection_stage(t, t + 0.5 * dt, 0)) return rc;
    if (int rc = diffusion_stage(t + 0.25 * dt, t + 0.5 * dt)) return rc;
    if (int rc = launched(launch_copy(n, state, start, st), "copy")) return rc;
    // the stage length as SplitSolver forms it: tStageEnd - tStageStart
    const double dt_prod = (t + dt) - t;
    if (int rc = emx_source_integrate_device(c, nPoints, nVars, start, sc_prod, qloss, t, dt_prod, state,
                                             status ? status : c->b_status.as<int32_t>(),
                                             c->b_s1.as<int32_t>(), c->b_s2.as<int32_t>(), nullptr))
        return rc;
    if (int rc = launched(launch_accumulate(n, state, start, s->delta_prod, st), "accumulate")) return rc;
    if (int rc = diffusion_stage(t + 0.5 * dt, t + 0.75 * dt)) return rc;
    if (int rc = convection_stage(t + 0.5 * dt, t + dt, 1)) return rc;
    if (int rc = diffusion_stage(t + 0.75 * dt, t + dt)) return rc;
    // calculateTimeDerivatives
    if (conv)
        if (int rc = launched(launch_conv_time_derivative(n, dt, conv->delta_conv, sc_conv, s->ddt_conv, st),
                              "time derivatives"))
            return rc;
    return launched(launch_time_derivatives(n, dt, s->delta_diff, s->delta_prod, sc_diff, sc_prod,
                                            s->ddt_diff, s->ddt_prod, st), "time derivatives");
}

int emx_split_step_device(emx_context* c, size_t nVars, size_t nPoints, const emx_split_arrays* s,
                          const double* B, size_t nB, const int32_t* B_row, const double* D,
                          const emx_grid* g, const double* qloss, const double* h_per_var, double t,
                          double dt, int balanced, int32_t* status) {
    return split_step(c, nVars, nPoints, s, nullptr, B, nB, B_row, D, g, qloss, h_per_var, t, dt, balanced,
                      status);
}

int emx_split_step_conv_device(emx_context* c, size_t nVars, size_t nPoints, const emx_split_arrays* s,
                               const emx_split_convection* conv, const double* B, size_t nB,
                               const int32_t* B_row, const double* D, const emx_grid* g, const double* qloss,
                               const double* h_per_var, double t, double dt, int balanced, int32_t* status) {
    if (!conv) return fail(-1, "convection description is NULL");
    return split_step(c, nVars, nPoints, s, conv, B, nB, B_row, D, g, qloss, h_per_var, t, dt, balanced,
                      status);
}

int emx_density_derivative_device(emx_context* c, size_t nVars, size_t nPoints, const double* state,
                                  const double* ddt_conv, const double* ddt_diff, const double* ddt_prod,
                                  const double* ddt_cross, int balanced, const double* rho, const double* Wmx,
                                  double* drhodt) {
    if (int rc = check_ctx(c)) return rc;
    if ((int)nVars != c->hm.K + 2) return fail(-1, "nVars must equal nSpecies + 2");
    if (nPoints > 0x7fffffffULL) return fail(-1, "too many points");
    if (nPoints == 0) return 0;
    if (!state || !ddt_conv || !ddt_diff || !ddt_prod || !rho || !Wmx || !drhodt)
        return fail(-1, "only ddt_cross may be NULL");
    int e = launch_density_derivative(c->dm, (int)nPoints, balanced, state, ddt_conv, ddt_diff, ddt_prod,
                                      ddt_cross, rho, Wmx, drhodt, c->stream);
    return e ? fail(-10, std::string("density derivative launch: ") + hipGetErrorString((hipError_t)e)) : 0;
}

int emx_correct_mass_fractions_device(emx_context* c, size_t nVars, size_t nPoints, double* state) {
    if (int rc = check_ctx(c)) return rc;
    if ((int)nVars != c->hm.K + 2) return fail(-1, "nVars must equal nSpecies + 2");
    if (nPoints > 0x7fffffffULL) return fail(-1, "too many points");
    if (nPoints == 0) return 0;
    if (!state) return fail(-1, "state is NULL");
    int e = launch_correct_mass_fractions((int)nVars, (int)nPoints, state, c->stream);
    return e ? fail(-10, "mass fraction launch failed") : 0;
}

int emx_heat_release_device(emx_context* c, size_t nVars, size_t nPoints, const double* state,
                            double t_now, double* qdot, int32_t* status) {
    if (int rc = check_ctx(c)) return rc;
    if ((int)nVars != c->hm.K + 2) return fail(-1, "nVars must equal nSpecies + 2");
    if (nPoints > 0x7fffffffULL) return fail(-1, "too many points");
    if (nPoints == 0) return 0;
    if (!state || !qdot) return fail(-1, "state or qdot is NULL");
    int e = launch_heat_release(c->dm, stage_opts(c, t_now), state, stage_perturb(c), (int)nPoints, qdot,
                                status, c->stream);
    return e ? fail(-10, std::string("heat release launch: ") + hipGetErrorString((hipError_t)e)) : 0;
}

int emx_thermo_properties(emx_context* c, size_t nVars, size_t nPoints, const double* state,
                          double* rho, double* Wmx, double* cp, double* cpSpec, double* hk) {
    if (int rc = check_ctx(c)) return rc;
    if ((int)nVars != c->hm.K + 2) return fail(-1, "nVars must equal nSpecies + 2");
    if (nPoints > 0x7fffffffULL) return fail(-1, "too many points");
    if (nPoints == 0) return 0;
    if (!state || !rho || !cp) return fail(-1, "state, rho or cp is NULL");
    const size_t N = nPoints, K = nVars - 2;
    hipStream_t st = c->stream;
    HIP_TRY(c->b_in.ensure(N * nVars * sizeof(double)));
    HIP_TRY(c->b_out.ensure((3 * N + 2 * N * K) * sizeof(double)));
    // staging: rho, Wmx, cp (N each), then cpSpec and hk (N x K each)
    double* o = c->b_out.as<double>();
    double *d_rho = o, *d_wmx = o + N, *d_cp = o + 2 * N, *d_cps = o + 3 * N, *d_hk = d_cps + N * K;
    HIP_TRY(hipMemcpyAsync(c->b_in.p, state, N * nVars * sizeof(double), hipMemcpyHostToDevice, st));
    int e = launch_thermo_props(c->dm, c->b_in.as<double>(), (int)N, d_rho, Wmx ? d_wmx : nullptr, d_cp,
                                cpSpec ? d_cps : nullptr, hk ? d_hk : nullptr, st);
    if (e) return fail(-10, std::string("thermo launch: ") + hipGetErrorString((hipError_t)e));
    HIP_TRY(hipMemcpyAsync(rho, d_rho, N * sizeof(double), hipMemcpyDeviceToHost, st));
    if (Wmx) HIP_TRY(hipMemcpyAsync(Wmx, d_wmx, N * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(cp, d_cp, N * sizeof(double), hipMemcpyDeviceToHost, st));
    if (cpSpec) HIP_TRY(hipMemcpyAsync(cpSpec, d_cps, N * K * sizeof(double), hipMemcpyDeviceToHost, st));
    if (hk) HIP_TRY(hipMemcpyAsync(hk, d_hk, N * K * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

int emx_thermo_properties_device(emx_context* c, size_t nVars, size_t nPoints, const double* state,
                                 double* rho, double* Wmx, double* cp, double* cpSpec, double* hk) {
    if (int rc = check_ctx(c)) return rc;
    if ((int)nVars != c->hm.K + 2) return fail(-1, "nVars must equal nSpecies + 2");
    if (nPoints > 0x7fffffffULL) return fail(-1, "too many points");
    if (nPoints == 0) return 0;
    if (!state || !rho || !cp) return fail(-1, "state, rho or cp is NULL");
    int e = launch_thermo_props(c->dm, state, (int)nPoints, rho, Wmx, cp, cpSpec, hk, c->stream);
    return e ? fail(-10, std::string("thermo launch: ") + hipGetErrorString((hipError_t)e)) : 0;
}

static int check_transport(emx_context* c, size_t nVars, size_t nPoints) {
    if (int rc = check_ctx(c)) return rc;
    if ((int)nVars != c->hm.K + 2) return fail(-1, "nVars must equal nSpecies + 2");
    if (nPoints > 0x7fffffffULL) return fail(-1, "too many points");
    if (!c->hm.has_transport())
        return fail(-1, "the mechanism table has no transport data (species `transport` blocks)");
    if (!(c->opts.transport_threshold >= 0.0)) return fail(-1, "transport_threshold must be >= 0");
    return 0;
}

int emx_transport_properties_device(emx_context* c, size_t nVars, size_t nPoints, const double* state,
                                    double* mu, double* lambda, double* rhoD) {
    if (int rc = check_transport(c, nVars, nPoints)) return rc;
    if (nPoints == 0) return 0;
    if (!state || !mu || !lambda || !rhoD) return fail(-1, "state, mu, lambda or rhoD is NULL");
    int e = launch_transport_props(c->dm, state, (int)nPoints, c->opts.transport_threshold, mu, lambda, 1,
                                   rhoD, (int)(nVars - 2), c->stream);
    return e ? fail(-10, std::string("transport launch: ") + hipGetErrorString((hipError_t)e)) : 0;
}

int emx_transport_properties(emx_context* c, size_t nVars, size_t nPoints, const double* state, double* mu,
                             double* lambda, double* rhoD) {
    if (int rc = check_transport(c, nVars, nPoints)) return rc;
    if (nPoints == 0) return 0;
    if (!state || !mu || !lambda || !rhoD) return fail(-1, "state, mu, lambda or rhoD is NULL");
    const size_t N = nPoints, K = nVars - 2;
    hipStream_t st = c->stream;
    HIP_TRY(c->b_in.ensure(N * nVars * sizeof(double)));
    HIP_TRY(c->b_out.ensure((2 * N + N * K) * sizeof(double)));
    double* o = c->b_out.as<double>();
    HIP_TRY(hipMemcpyAsync(c->b_in.p, state, N * nVars * sizeof(double), hipMemcpyHostToDevice, st));
    int e = launch_transport_props(c->dm, c->b_in.as<double>(), (int)N, c->opts.transport_threshold, o, o + N,
                                   1, o + 2 * N, (int)K, st);
    if (e) return fail(-10, std::string("transport launch: ") + hipGetErrorString((hipError_t)e));
    HIP_TRY(hipMemcpyAsync(mu, o, N * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(lambda, o + N, N * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(rhoD, o + 2 * N, N * K * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

int emx_diffusion_coefficients_device(emx_context* c, size_t nVars, size_t nPoints, const double* state,
                                      double* B, double* D) {
    if (int rc = check_transport(c, nVars, nPoints)) return rc;
    if (nPoints == 0) return 0;
    if (!state || !B || !D) return fail(-1, "state, B or D is NULL");
    const int n = (int)nPoints, nv = (int)nVars;
    hipStream_t st = c->stream;
    // rho into B row 0, cp into B row 1, then B = (1/rho, 1/(rho cp)) in place
    int e = launch_thermo_props(c->dm, state, n, B, nullptr, B + nPoints, nullptr, nullptr, st);
    if (!e) e = launch_diffusion_B(n, B, st);
    // D rows per point: mu, lambda, rho D_k (FlameSolver::prepareIntegrators)
    if (!e) e = launch_transport_props(c->dm, state, n, c->opts.transport_threshold, D, D + 1, nv, D + 2, nv, st);
    return e ? fail(-10, std::string("diffusion coefficient launch: ") + hipGetErrorString((hipError_t)e)) : 0;
}

// ---------------------------------------------------------------------------
// convection stage
// ---------------------------------------------------------------------------
static constexpr int kConvMaxRecords = 2048;  // v records: UTW steps + 2

int emx_convection_integrate_device(emx_context* c, size_t nVars, size_t nPoints,
                                    const emx_convection_params* cp, const double* x, const double* state_in,
                                    const double* split_const, const double* drhodt, double t0, double t1,
                                    double* state_out, int32_t* steps) {
    if (int rc = check_ctx(c)) return rc;
    if ((int)nVars != c->hm.K + 2) return fail(-1, "nVars must equal nSpecies + 2");
    if (!cp || !x || !state_in || !state_out) return fail(-1, "params, x or a state pointer is NULL");
    if (int rc = check_conv_params(nPoints, cp)) return rc;
    if (!(t1 > t0)) return fail(-1, "t1 must exceed t0");
    const int N = (int)nPoints, K = c->hm.K;
    hipStream_t st = c->stream;
    // setLeftBC: Wleft = the mean molecular weight of the (normalised) left composition
    double Wleft = 0.0;
    if (cp->Yleft) {
        double norm = 0.0, sum = 0.0;
        for (int k = 0; k < K; k++) norm += std::max(cp->Yleft[k], 0.0);
        if (!(norm > 0.0) || !std::isfinite(norm)) return fail(-1, "Yleft must hold a positive, finite composition");
        for (int k = 0; k < K; k++) sum += std::max(cp->Yleft[k], 0.0) * (1.0 / norm) * (1.0 / c->hm.W[k]);
        Wleft = 1.0 / sum;
        // uploaded when it changes, so that a call captured in a graph after a first call
        // records no host copy
        if (c->cv_yleft_host.size() != (size_t)K ||
            !std::equal(c->cv_yleft_host.begin(), c->cv_yleft_host.end(), cp->Yleft)) {
            HIP_TRY(c->cv_yleft.ensure(K * sizeof(double)));
            HIP_TRY(hipMemcpyAsync(c->cv_yleft.p, cp->Yleft, K * sizeof(double), hipMemcpyHostToDevice, st));
            c->cv_yleft_host.assign(cp->Yleft, cp->Yleft + K);
        }
    }
    HIP_TRY(c->cv_work.ensure(convection_work_doubles(N, K) * sizeof(double)));
    HIP_TRY(c->cv_rec_t.ensure(kConvMaxRecords * sizeof(double)));
    HIP_TRY(c->cv_rec_v.ensure((size_t)kConvMaxRecords * N * sizeof(double)));
    HIP_TRY(c->cv_ints.ensure((2 + K) * sizeof(int)));
    // aux: Wmx, rho, cp, scU, scT, scW, drhodt (zeros when NULL)
    HIP_TRY(c->cv_aux.ensure(7 * (size_t)N * sizeof(double)));
    double* aux = c->cv_aux.as<double>();
    double *Wmx = aux, *rho = aux + N, *cpm = aux + 2 * N, *scU = aux + 3 * N, *scT = aux + 4 * N,
           *scW = aux + 5 * N, *dr0 = aux + 6 * N;
    // the register path (N <= kConvRegMax) forms Wmx and the UTW split constants itself
    int e = 0;
    if (N > kConvRegMax) {
        e = launch_thermo_props(c->dm, state_in, N, rho, Wmx, cpm, nullptr, nullptr, st);
        if (!e) e = launch_conv_split_constants(c->dm, N, split_const, scU, scT, scW, st);
        if (!e && !drhodt) e = launch_fill((size_t)N, 0.0, dr0, nullptr, st);
    }
    if (e) return fail(-10, std::string("convection set-up launch: ") + hipGetErrorString((hipError_t)e));
    ConvParams p{};
    p.N = N;
    p.K = K;
    p.nv = (int)nVars;
    p.alpha = cp->alpha;
    p.beta = cp->beta;
    p.leftBC = cp->left_bc;
    p.rightBC = cp->right_bc;
    p.P = c->P;
    p.rVzero = cp->rVzero;
    p.Tleft = cp->Tleft;
    p.Wleft = Wleft;
    p.rhou = cp->rhou;
    p.aInitial = cp->strain_initial;
    p.aFinal = cp->strain_final;
    p.T0 = cp->strain_t0;
    p.Dt = cp->strain_dt;
    // ConvectionSystemSplit::setTolerances (src/convectionSystem.cpp:448-455)
    p.rtol = c->opts.integrator_rel_tol;
    p.atolU = c->opts.integrator_momentum_abs_tol;
    p.atolT = c->opts.integrator_energy_abs_tol;
    p.atolW = c->opts.integrator_species_abs_tol * 20;
    p.atolY = c->opts.integrator_species_abs_tol;
    p.t0 = t0;
    p.t1 = t1;
    p.maxSteps = cp->max_steps > 0 ? cp->max_steps : 1000000;
    p.maxRec = kConvMaxRecords;
    ConvArrays a{};
    a.state = state_in;
    a.sc = split_const;
    a.x = x;
    a.drhodt = drhodt ? drhodt : (N > kConvRegMax ? dr0 : nullptr);
    a.rW = c->dm.rW;
    a.Wmx = Wmx;
    a.scU = scU;
    a.scT = scT;
    a.scW = scW;
    a.Yleft = cp->Yleft ? c->cv_yleft.as<double>() : nullptr;
    a.state_out = state_out;
    a.rec_t = c->cv_rec_t.as<double>();
    a.rec_v = c->cv_rec_v.as<double>();
    a.nrec = c->cv_ints.as<int>();
    a.steps = steps ? steps : c->cv_ints.as<int>() + 1;
    a.work = c->cv_work.as<double>();
    e = launch_convection(p, a, st);
    return e ? fail(-10, std::string("convection launch: ") + hipGetErrorString((hipError_t)e)) : 0;
}

int emx_convection_integrate(emx_context* c, size_t nVars, size_t nPoints, const emx_convection_params* cp,
                             const double* x, const double* split_const, const double* drhodt, double t0,
                             double t1, double* state, int32_t* steps) {
    if (int rc = check_ctx(c)) return rc;
    if (!state || !x) return fail(-1, "state or x is NULL");
    if (nPoints < 3 || nPoints > (size_t)kConvMaxPoints)
        return fail(-1, "the convection stage takes 3 to 4096 grid points");
    const size_t NV = nVars * nPoints * sizeof(double);
    const int K = c->hm.K;
    hipStream_t st = c->stream;
    HIP_TRY(c->d_y.ensure(NV));
    HIP_TRY(c->d_x.ensure(nPoints * sizeof(double)));
    HIP_TRY(hipMemcpyAsync(c->d_y.p, state, NV, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->d_x.p, x, nPoints * sizeof(double), hipMemcpyHostToDevice, st));
    const double* sc = nullptr;
    if (split_const) {
        HIP_TRY(c->d_sc.ensure(NV));
        HIP_TRY(hipMemcpyAsync(c->d_sc.p, split_const, NV, hipMemcpyHostToDevice, st));
        sc = c->d_sc.as<double>();
    }
    const double* dr = nullptr;
    if (drhodt) {
        HIP_TRY(c->d_hmin.ensure(nPoints * sizeof(double)));
        HIP_TRY(hipMemcpyAsync(c->d_hmin.p, drhodt, nPoints * sizeof(double), hipMemcpyHostToDevice, st));
        dr = c->d_hmin.as<double>();
    }
    HIP_TRY(c->b_status.ensure((1 + K) * sizeof(int32_t)));
    int rc = emx_convection_integrate_device(c, nVars, nPoints, cp, c->d_x.as<double>(), c->d_y.as<double>(), sc,
                                             dr, t0, t1, c->d_y.as<double>(), c->b_status.as<int32_t>());
    if (rc) return rc;
    std::vector<int32_t> s(1 + K);
    HIP_TRY(hipMemcpyAsync(s.data(), c->b_status.p, (1 + K) * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(state, c->d_y.p, NV, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (steps) memcpy(steps, s.data(), (1 + K) * sizeof(int32_t));
    for (int i = 0; i <= K; i++)
        if (s[i] < 0)
            return fail(-3, std::string("convection integration failed (") + (i == 0 ? "UTW system" : "species") +
                                ", code " + std::to_string(s[i]) + ")");
    return 0;
}

int emx_tridiagonal_integrate(emx_context* c, size_t nSys, size_t N, const double* a, const double* b,
                              const double* cc, const double* k, const double* h_in, double t0,
                              double t1, double* y) {
    if (int rc = check_ctx(c)) return rc;
    if (N < 2 || nSys < 1 || nSys > 64) return fail(-1, "need N >= 2 and 1 <= nSys <= 64");
    if (!h_in || !a || !b || !cc || !k || !y) return fail(-1, "an array argument is NULL");
    hipStream_t st = c->stream;
    const size_t bytes = nSys * N * sizeof(double);
    HIP_TRY(c->t_a.ensure(bytes));
    HIP_TRY(c->t_b.ensure(bytes));
    HIP_TRY(c->t_c.ensure(bytes));
    HIP_TRY(c->t_k.ensure(bytes));
    HIP_TRY(c->d_y.ensure(bytes));
    HIP_TRY(hipMemcpyAsync(c->t_a.p, a, bytes, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->t_b.p, b, bytes, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->t_c.p, cc, bytes, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->t_k.p, k, bytes, hipMemcpyHostToDevice, st));
    // y: system-major on the host, point-major on the device
    std::vector<double> yt(nSys * N);
    for (size_t s = 0; s < nSys; s++)
        for (size_t j = 0; j < N; j++) yt[j * nSys + s] = y[s * N + j];
    HIP_TRY(hipMemcpyAsync(c->d_y.p, yt.data(), bytes, hipMemcpyHostToDevice, st));
    HIP_TRY(c->d_work.ensure(diffusion_work_doubles((int)nSys, (int)N) * sizeof(double)));
    DiffWork w;
    diffusion_carve(c->d_work.as<double>(), (int)nSys, (int)N, w);
    std::vector<double> h(nSys);
    std::vector<int> ns(nSys);
    int maxs = 1;
    for (size_t s = 0; s < nSys; s++) {
        int n = std::max(static_cast<int>((t1 - t0) / h_in[s] + 1e-5), 1);
        h[s] = (t1 - t0) / n;
        ns[s] = n;
        maxs = std::max(maxs, n);
    }
    HIP_TRY(c->d_h.ensure(nSys * 8));
    HIP_TRY(c->d_ns.ensure(nSys * 4));
    HIP_TRY(hipMemcpyAsync(c->d_h.p, h.data(), nSys * 8, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->d_ns.p, ns.data(), nSys * 4, hipMemcpyHostToDevice, st));
    for (int q = 0; q < 8 + maxs - 1; q++) {
        int e = launch_explicit_solve(c->t_a.as<double>(), c->t_b.as<double>(), c->t_c.as<double>(),
                                      c->t_k.as<double>(), (int)nSys, (int)N, c->d_y.as<double>(),
                                      c->d_h.as<double>(), c->d_ns.as<int>(), q, w, st);
        if (e) return fail(-10, "tridiagonal launch failed");
    }
    HIP_TRY(hipMemcpyAsync(yt.data(), c->d_y.p, bytes, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (size_t s = 0; s < nSys; s++)
        for (size_t j = 0; j < N; j++) y[s * N + j] = yt[j * nSys + s];
    return 0;
}

}  // extern "C"
